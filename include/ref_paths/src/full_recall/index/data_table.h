// Forwarding header at the include path the reference's search driver uses (harness/run/run_context.h:24):
// DataTable of this build.
#pragma once
#include "../../../../../hybridsearchsystem_amd/csrc/engine/data_table.h"
