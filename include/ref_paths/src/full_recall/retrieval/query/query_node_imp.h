// Forwarding header at the include path the reference's search driver uses
// (harness/run/run_context.cpp:13, harness/query/filter_setup.h:13): the QueryNode hierarchy of this build.
#pragma once
#include "../../../../../../hybridsearchsystem_amd/csrc/engine/query_node.h"
