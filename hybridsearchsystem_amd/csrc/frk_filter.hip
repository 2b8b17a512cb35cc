// frk_filter.hip -- boolean tag filter: one postfix program per query evaluated over posting operands
// in all five on-disk layouts, written as per-segment u64 bit words.
//
// Semantics follow the reference's filter kernel (engine/retrieval/device/filter/
// kernel_bitmap_text_filter.cpp:52-119 with filter_and_op.h, filter_or_op.h, filter_not_op.h,
// filter_conjunction_op.h): AND starts from all-ones and folds the postings, then the stack entries
// popped from the top; OR starts from zero; NOT takes one operand; CONJ is an AND over parts of the
// OR of each part's postings; the last op's value is the result.
//
// Work split: one workgroup of 256 threads per (query, segment, 256-word chunk); each thread owns one
// u64 word (64 docs). Phase 1 stages every non-dense operand of the chunk into an LDS slot:
// SPARSE_BITMAP (1) and SPARSE_PACKED (4) units of the chunk are located by a two-round cooperative
// search (their units are ascending, posting_serializer.cpp:142-186) and scattered with LDS atomic ORs, SPARSE_IDLIST (3)
// ids are filtered and scattered, DENSE_BYTEMAP (2) bytes are packed to bits. Phase 2 runs the
// program per word with the stack in registers; DENSE_BITMAP (0) operands are read straight from
// HBM with coalesced 8-byte loads.
#include <stdarg.h>
#include <stdio.h>

#include "frk_common.h"

namespace frk {
namespace {

constexpr int kChunkWords = 256;
constexpr int kSlots = 32;
constexpr uint32_t kStatusBadProgram = 1u;
constexpr uint32_t kStatusTooManyPostings = 2u;

enum OperandKind : uint32_t { kZero = 0, kDense = 1, kSlot = 2, kSlow = 3 };

struct PostingHead {
    const uint8_t* payload;
    uint32_t layout;
    uint32_t hits;
    uint32_t len;
};

__device__ inline PostingHead ReadHead(uint64_t addr) {
    PostingHead h;
    const uint32_t* p = reinterpret_cast<const uint32_t*>(addr);
    uint32_t w0 = p[0];
    h.layout = w0 >> 28;
    h.hits = w0 & 0x0FFFFFFFu;
    h.len = p[1];
    h.payload = reinterpret_cast<const uint8_t*>(addr) + 8;
    return h;
}

// First index i in [0, k) whose unit is >= unit_lo, for the two sorted sparse layouts.
__device__ inline uint32_t LowerBoundUnit(const PostingHead& h, uint32_t k, uint32_t unit_lo) {
    uint32_t lo = 0, hi = k;
    const uint32_t* words = reinterpret_cast<const uint32_t*>(h.payload);
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        uint32_t unit = h.layout == 4 ? (words[mid] >> 16) : (words[mid] >> 1);
        if (unit < unit_lo) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ inline void UnitAt(const PostingHead& h, uint32_t k, uint32_t i, uint32_t& unit, uint32_t& mask) {
    const uint32_t* words = reinterpret_cast<const uint32_t*>(h.payload);
    if (h.layout == 4) {
        uint32_t w = words[i];
        unit = w >> 16;
        mask = w & 0xFFFFu;
    } else {
        unit = words[i] >> 1;  // byte offset of the u16 unit in the dense bitmap
        mask = reinterpret_cast<const uint16_t*>(h.payload + 4ull * k)[i];
    }
}

// Both lower bounds (first unit >= a, first unit >= b) of a sorted unit list, found by the whole
// workgroup in two rounds of parallel probes instead of a dependent binary search per thread: round 1
// probes every s-th entry (s = ceil(k / 256)) and counts the probes below each target, round 2 probes
// the s entries of the bracket that count selects. Called by every thread of the workgroup.
__device__ inline void CoopLowerBounds(const PostingHead& h, uint32_t k, uint32_t a, uint32_t b, uint32_t& lo,
                                       uint32_t& hi) {
    const uint32_t t = threadIdx.x;
    const uint32_t s = (k + kChunkWords - 1) / kChunkWords;  // <= 32 for a segment of 8192 units
    if (s > kChunkWords) {  // more units than two rounds cover (segments beyond 1M docs): per-thread search
        lo = LowerBoundUnit(h, k, a);
        hi = LowerBoundUnit(h, k, b);
        return;
    }
    const uint32_t* words = reinterpret_cast<const uint32_t*>(h.payload);
    auto unit_at = [&](uint32_t i) { return h.layout == 4 ? (words[i] >> 16) : (words[i] >> 1); };
    const uint32_t p1 = t * s;
    uint32_t u1 = 0xFFFFFFFFu;
    if (s && p1 < k) u1 = unit_at(p1);
    const uint32_t ca = __syncthreads_count(p1 < k && u1 < a);
    const uint32_t cb = __syncthreads_count(p1 < k && u1 < b);
    // the bound lies in (p_{c-1}, p_c]: probe the up to s - 1 entries strictly between
    auto round2 = [&](uint32_t c, uint32_t target) {
        if (c == 0) return __syncthreads_count(false) * 0u;
        const uint32_t base = (c - 1) * s + 1;
        const uint32_t p2 = base + t;
        const bool below = t + 1 < s && p2 < k && unit_at(p2) < target;
        return base + __syncthreads_count(below);
    };
    lo = round2(ca, a);
    hi = round2(cb, b);
}

// One operand's 64 bits at segment word w, read directly (dense and the per-thread fallback).
__device__ inline uint64_t OperandWordDirect(const PostingHead& h, uint32_t w) {
    if (h.hits == 0) return 0;
    switch (h.layout) {
        case 0: {
            if (8ull * w >= h.len) return 0;
            return reinterpret_cast<const uint64_t*>(h.payload)[w];
        }
        case 1:
        case 4: {
            uint32_t k = h.len / (h.layout == 4 ? 4u : 6u);
            uint32_t i = LowerBoundUnit(h, k, 4 * w);
            uint64_t v = 0;
            for (; i < k; i++) {
                uint32_t unit, mask;
                UnitAt(h, k, i, unit, mask);
                if (unit >= 4 * w + 4) break;
                v |= static_cast<uint64_t>(mask) << (16 * (unit & 3));
            }
            return v;
        }
        case 2: {
            uint64_t v = 0;
            uint32_t base = 64 * w;
            for (uint32_t b = 0; b < 64 && base + b < h.len; b++) v |= static_cast<uint64_t>(h.payload[base + b] != 0) << b;
            return v;
        }
        case 3: {
            uint64_t v = 0;
            const uint32_t* ids = reinterpret_cast<const uint32_t*>(h.payload);
            uint32_t n = h.len / 4;
            for (uint32_t i = 0; i < n; i++)
                if ((ids[i] >> 6) == w) v |= 1ull << (ids[i] & 63);
            return v;
        }
    }
    return 0;
}

__global__ __launch_bounds__(kChunkWords) void FilterKernel(const frk_filter_query* __restrict__ queries,
                                                            const uint32_t* __restrict__ expr,
                                                            const uint64_t* __restrict__ table, uint32_t seg_docs,
                                                            uint64_t n_docs, uint64_t* __restrict__ out,
                                                            uint64_t out_stride, uint32_t* __restrict__ counts,
                                                            uint32_t* __restrict__ status) {
    __shared__ uint64_t slots[kSlots][kChunkWords];
    __shared__ uint32_t kinds[FRK_FILTER_MAX_POSTINGS];
    __shared__ uint32_t wave_pop[kChunkWords / 64];
    __shared__ uint32_t slot_posting[kSlots];
    __shared__ uint32_t n_slots_used;

    const uint32_t q = blockIdx.z;
    const uint32_t seg = blockIdx.y;
    const uint32_t seg_words = seg_docs / 64;
    const uint32_t w0 = blockIdx.x * kChunkWords;
    const uint32_t t = threadIdx.x;
    const uint32_t w = w0 + t;
    const frk_filter_query Q = queries[q];

    // docs of this segment; bits at or past it stay 0
    const uint64_t seg_first = static_cast<uint64_t>(seg) * seg_docs;
    const uint64_t n_seg = n_docs > seg_first ? (n_docs - seg_first < seg_docs ? n_docs - seg_first : seg_docs) : 0;
    uint64_t valid = 0;
    if (64ull * (w + 1) <= n_seg) valid = ~0ull;
    else if (64ull * w < n_seg) valid = (1ull << (n_seg - 64ull * w)) - 1;

    uint64_t result = 0;
    if (Q.flags & FRK_FILTER_MATCH_ALL) {
        result = valid;
    } else if (Q.n_postings > FRK_FILTER_MAX_POSTINGS) {
        if (t == 0) atomicOr(status, kStatusTooManyPostings);
    } else {
        const uint64_t* tab = table + Q.table_off + static_cast<uint64_t>(seg) * Q.n_postings;
        // ---- classify operands, give the first kSlots non-dense ones an LDS slot, stage them
        for (uint32_t p = t; p < Q.n_postings; p += kChunkWords) {
            uint64_t a = tab[p];
            uint32_t kind = kZero;
            if (a != 0) {
                PostingHead h = ReadHead(a);
                if (h.hits != 0) kind = h.layout == 0 ? kDense : kSlot;
            }
            kinds[p] = kind;
        }
        __syncthreads();
        if (t == 0) {
            uint32_t n = 0;
            for (uint32_t p = 0; p < Q.n_postings; p++) {
                if (kinds[p] != kSlot) continue;
                if (n < kSlots) {
                    kinds[p] = kSlot | (n << 2);
                    slot_posting[n++] = p;
                } else {
                    kinds[p] = kSlow;
                }
            }
            n_slots_used = n;
        }
        __syncthreads();
        const uint32_t n_slots = n_slots_used;
        for (uint32_t sl = 0; sl < n_slots; sl++) {
            const uint32_t p = slot_posting[sl];
            PostingHead h = ReadHead(tab[p]);
            if (h.layout == 2) {
                // DENSE_BYTEMAP: 64 bytes per word, no sharing between threads
                slots[sl][t] = (w < seg_words) ? OperandWordDirect(h, w) : 0;
                continue;
            }
            slots[sl][t] = 0;
            __syncthreads();
            if (h.layout == 1 || h.layout == 4) {
                const uint32_t k = h.len / (h.layout == 4 ? 4u : 6u);
                uint32_t lo, hi;
                CoopLowerBounds(h, k, 4 * w0, 4 * (w0 + kChunkWords), lo, hi);
                for (uint32_t i = lo + t; i < hi; i += kChunkWords) {
                    uint32_t unit, mask;
                    UnitAt(h, k, i, unit, mask);
                    atomicOr(reinterpret_cast<unsigned long long*>(&slots[sl][(unit >> 2) - w0]),
                             static_cast<unsigned long long>(mask) << (16 * (unit & 3)));
                }
            } else if (h.layout == 3) {
                const uint32_t* ids = reinterpret_cast<const uint32_t*>(h.payload);
                const uint32_t n = h.len / 4;
                for (uint32_t i = t; i < n; i += kChunkWords) {
                    uint32_t id = ids[i];
                    uint32_t wd = id >> 6;
                    if (wd >= w0 && wd < w0 + kChunkWords)
                        atomicOr(reinterpret_cast<unsigned long long*>(&slots[sl][wd - w0]), 1ull << (id & 63));
                }
            }
        }
        __syncthreads();

        // ---- evaluate the program for this thread's word
        uint64_t stack[FRK_FILTER_MAX_STACK];
        int sp = 0;
        uint32_t pc = 0;
        bool bad = false;
        auto operand = [&](uint32_t p) -> uint64_t {
            uint32_t kind = kinds[p];
            switch (kind & 3u) {
                case kDense: return OperandWordDirect(ReadHead(tab[p]), w);
                case kSlot: return slots[kind >> 2][t];
                case kSlow: return OperandWordDirect(ReadHead(tab[p]), w);
                default: return 0;
            }
        };
        const uint32_t* prog = expr + Q.expr_off;
        for (uint32_t i = 0; i < Q.expr_len && !bad; i++) {
            const uint32_t op = prog[i];
            uint64_t v;
            if (op == 0) {  // CONJ
                if (i + 1 >= Q.expr_len) { bad = true; break; }
                const uint32_t parts = prog[i + 1];
                if (i + 1 + parts >= Q.expr_len) { bad = true; break; }
                v = ~0ull;
                for (uint32_t j = 0; j < parts; j++) {
                    const uint32_t n = prog[i + 2 + j];
                    if (pc + n > Q.n_postings) { bad = true; break; }
                    uint64_t acc = 0;
                    for (uint32_t x = 0; x < n; x++) acc |= operand(pc + x);
                    pc += n;
                    v &= acc;
                }
                i += parts + 1;
            } else if (op == 1 || op == 2 || op == 3) {
                if (i + 2 >= Q.expr_len) { bad = true; break; }
                const uint32_t pn = prog[i + 1], sn = prog[i + 2];
                if (pc + pn > Q.n_postings || sn > static_cast<uint32_t>(sp)) { bad = true; break; }
                if (op == 3) {
                    if (pn + sn != 1) { bad = true; break; }
                    v = ~(pn ? operand(pc) : stack[sp - 1]);
                } else {
                    const bool is_and = op == 1;
                    v = is_and ? ~0ull : 0ull;
                    for (uint32_t x = 0; x < pn; x++) v = is_and ? (v & operand(pc + x)) : (v | operand(pc + x));
                    for (uint32_t x = 0; x < sn; x++) v = is_and ? (v & stack[sp - 1 - x]) : (v | stack[sp - 1 - x]);
                }
                pc += pn;
                sp -= static_cast<int>(sn);
                i += 2;
            } else {
                bad = true;
                break;
            }
            if (sp >= FRK_FILTER_MAX_STACK) { bad = true; break; }
            stack[sp++] = v;
        }
        if (bad || sp != 1 || pc != Q.n_postings) {
            if (t == 0) atomicOr(status, kStatusBadProgram);
            result = 0;
        } else {
            result = stack[0] & valid;
        }
    }
    if (w < seg_words) out[static_cast<uint64_t>(q) * out_stride + static_cast<uint64_t>(seg) * seg_words + w] = result;

    // passing-doc count for the query
    uint32_t pop = static_cast<uint32_t>(__popcll(result));
    for (int off = 32; off > 0; off >>= 1) pop += __shfl_xor(pop, off, 64);
    if ((t & 63) == 0) wave_pop[t >> 6] = pop;
    __syncthreads();
    if (t == 0) {
        uint32_t s = 0;
        for (int i = 0; i < kChunkWords / 64; i++) s += wave_pop[i];
        if (s) atomicAdd(&counts[q], s);
    }
}

}  // namespace
}  // namespace frk

extern "C" int frk_filter_eval(const frk_filter_query* d_queries, uint32_t n_queries, const uint32_t* d_expr,
                               const uint64_t* d_table, uint32_t n_segments, uint32_t seg_docs, uint64_t n_docs,
                               uint64_t* d_out, uint64_t out_stride, uint32_t* d_counts, uint32_t* d_status,
                               frk_stream_t stream) {
    if (n_queries == 0 || n_segments == 0) return FRK_OK;
    if (seg_docs == 0 || seg_docs % 64 != 0) {
        frk::SetError("frk_filter_eval: seg_docs %u must be a positive multiple of 64", seg_docs);
        return FRK_ERR_UNSUPPORTED;
    }
    if (!d_queries || !d_expr || !d_table || !d_out || !d_counts || !d_status) {
        frk::SetError("frk_filter_eval: null device pointer");
        return FRK_ERR_INVALID;
    }
    const uint32_t seg_words = seg_docs / 64;
    if (out_stride < static_cast<uint64_t>(n_segments) * seg_words) {
        frk::SetError("frk_filter_eval: out_stride %llu < %llu words", (unsigned long long)out_stride,
                      (unsigned long long)n_segments * seg_words);
        return FRK_ERR_CAPACITY;
    }
    if (static_cast<uint64_t>(n_segments) * seg_docs < n_docs) {
        frk::SetError("frk_filter_eval: %u segments of %u docs do not cover %llu docs", n_segments, seg_docs,
                      (unsigned long long)n_docs);
        return FRK_ERR_INVALID;
    }
    if (n_segments > 65535 || n_queries > 65535) {
        frk::SetError("frk_filter_eval: grid limit (segments %u, queries %u)", n_segments, n_queries);
        return FRK_ERR_UNSUPPORTED;
    }
    dim3 grid((seg_words + frk::kChunkWords - 1) / frk::kChunkWords, n_segments, n_queries);
    hipLaunchKernelGGL(frk::FilterKernel, grid, dim3(frk::kChunkWords), 0, static_cast<hipStream_t>(stream), d_queries,
                       d_expr, d_table, seg_docs, n_docs, d_out, out_stride, d_counts, d_status);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}
