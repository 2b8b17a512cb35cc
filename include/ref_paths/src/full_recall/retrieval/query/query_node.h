// Forwarding header at the reference's include path (engine/retrieval/query/query_node.h).
#pragma once
#include "../../../../../../hybridsearchsystem_amd/csrc/engine/query_node.h"
