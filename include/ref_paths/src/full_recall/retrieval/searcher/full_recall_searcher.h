// Forwarding header at the include path the reference's search driver uses
// (harness/run/run_context.cpp:14, harness/run/schedule_serial.cpp:18): FullRecallSearcher of this build.
#pragma once
#include "../../../../../../hybridsearchsystem_amd/csrc/engine/full_recall_searcher.h"
