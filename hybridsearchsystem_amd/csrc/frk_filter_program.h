// frk_filter_program.h -- the structural check of a postfix filter program, shared by the filter kernel
// (which runs it once per workgroup before evaluating the program without further checks) and the
// host (frk_check_program), so the same rule can be exercised on the CPU.
//
// The rule is that of the reference's host-side check of a compiled program (TextFilter::CheckPostExpr,
// engine/retrieval/searcher/stage/text_filter.cpp:1220-1352): every op reads only words that exist,
// consumes postings in order without running past the operand list, pops only entries that are on the
// stack, NOT takes exactly one operand, AND/OR at least one, CONJ at least one part of at least one
// posting, and the program ends with every posting consumed and exactly one entry left. Every
// comparison is written so that no sum of program words can wrap (the reference adds 32-bit counts,
// text_filter.cpp:1230,1260,1291, so a count of 0xFFFFFFFF could pass there): a huge word is rejected
// like any other count that exceeds what is left. The stack limit is this kernel's own.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/frk_capi.h"

namespace frk {

// Returns true when prog[0 .. len) is a well-formed program over n_postings operands whose stack stays
// within max_stack entries; *depth (optional) receives the deepest stack it builds.
__host__ __device__ inline bool ValidateProgram(const uint32_t* prog, uint32_t len, uint32_t n_postings,
                                                int* depth_out, int max_stack = FRK_FILTER_MAX_STACK) {
    int sp = 0, depth = 0;
    uint32_t pc = 0;  // postings consumed so far; pc <= n_postings holds throughout
    for (uint32_t i = 0; i < len; i++) {
        const uint32_t op = prog[i];
        if (op == 0) {  // CONJ: [0, parts, n_1 .. n_parts], parts >= 1, every n_x >= 1
            if (len - i < 2) return false;
            const uint32_t parts = prog[i + 1];
            if (parts == 0 || parts > len - i - 2) return false;  // the part counts lie inside the program
            for (uint32_t x = 0; x < parts; x++) {
                const uint32_t n = prog[i + 2 + x];
                if (n == 0 || n > n_postings - pc) return false;
                pc += n;
            }
            i += 1 + parts;
        } else if (op == 1 || op == 2 || op == 3) {  // AND / OR / NOT: [op, postingNum, stackNum]
            if (len - i < 3) return false;
            const uint32_t pn = prog[i + 1], sn = prog[i + 2];
            if (pn > n_postings - pc || sn > static_cast<uint32_t>(sp)) return false;
            if (op == 3 ? !((pn == 1 && sn == 0) || (pn == 0 && sn == 1)) : (pn == 0 && sn == 0)) return false;
            pc += pn;
            sp -= static_cast<int>(sn);
            i += 2;
        } else {
            return false;
        }
        if (sp >= max_stack) return false;
        sp++;
        depth = sp > depth ? sp : depth;
    }
    if (sp != 1 || pc != n_postings) return false;
    if (depth_out) *depth_out = depth;
    return true;
}

}  // namespace frk
