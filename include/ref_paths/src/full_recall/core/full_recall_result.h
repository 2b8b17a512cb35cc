// Forwarding header at the include path the reference's search driver uses (harness/run/run_context.h:23):
// FullRecallResult, ScoreWithIndex, GmBlock and GmPoolName of this build.
#pragma once
#include "../../../../../hybridsearchsystem_amd/csrc/engine/full_recall_result.h"
