// Forwarding header at the include path the reference's search driver uses (harness/query/filter_setup.cpp:8):
// NpuRetrieval::Hash64 (MurmurHash3_x64_128, seed 0x20160713, out[0] ^ out[1], core/hash/hash.cpp:10-16).
#pragma once
#include "../../../../../../hybridsearchsystem_amd/csrc/engine/format.h"
