// frk_filter.hip -- boolean tag filter: one postfix program per query evaluated over posting operands
// in all five on-disk layouts, written as per-segment u64 bit words.
//
// Semantics follow the reference's filter kernel (engine/retrieval/device/filter/
// kernel_bitmap_text_filter.cpp:52-119 with filter_and_op.h, filter_or_op.h, filter_not_op.h,
// filter_conjunction_op.h): AND starts from all-ones and folds the postings, then the stack entries
// popped from the top; OR starts from zero; NOT takes one operand; CONJ is an AND over parts of the
// OR of each part's postings; the last op's value is the result.
//
// Work split: one workgroup of 512 threads per (query, segment, chunk of 512*W words); each thread
// owns W words (64 docs each), at w0 + t + 512 j. The kernel runs in three phases:
//   1. the addresses of the first S operands go to LDS (S slots, at most 16 / W: as many as the
//      batch's largest program needs);
//   2. each of those operands is staged whole into an LDS slot holding the chunk's words:
//      DENSE_BITMAP (0) words are copied with coalesced 8-byte loads, DENSE_BYTEMAP (2) bytes are
//      packed to bits, SPARSE_BITMAP (1) and SPARSE_PACKED (4) units are scattered with LDS atomic
//      ORs (a chunk smaller than its segment first finds its unit range with a two-round cooperative
//      search; units are ascending, posting_serializer.cpp:142-186), SPARSE_IDLIST (3) ids are
//      range-checked and scattered;
//   3. the program runs per word with its stack in registers and its operands read from LDS.
// Operands past the S slots (programs with more postings than the launch was sized for) are read
// per word straight from HBM: correct, slower.
// LDS: S slots x 512 W words x 8 B, at most 64 KiB.
#include <stdarg.h>
#include <stdio.h>

#include <utility>

#include "frk_common.h"
#include "frk_filter_program.h"

namespace frk {
namespace {

constexpr int kThreads = 512;
constexpr int kMaxSlots = 16;  // S * 512 W <= 8192 words (64 KiB)
constexpr uint32_t kLayoutEmpty = 15;  // a staged operand with no hits (absent posting or hitCount 0)
constexpr uint32_t kStatusBadProgram = 1u;
constexpr uint32_t kStatusTooManyPostings = 2u;


struct PostingHead {
    const uint8_t* payload;
    uint32_t layout;
    uint32_t hits;
    uint32_t len;
};

__device__ __forceinline__ PostingHead ReadHead(uint64_t addr) {
    PostingHead h;
    if (addr == 0) {
        h.payload = nullptr;
        h.layout = 0;
        h.hits = 0;
        h.len = 0;
        return h;
    }
    const uint2 w = *reinterpret_cast<const uint2*>(addr);
    h.layout = w.x >> 28;
    h.hits = w.x & 0x0FFFFFFFu;
    h.len = w.y;
    h.payload = reinterpret_cast<const uint8_t*>(addr) + 8;
    return h;
}

// Units in a sorted sparse payload: SPARSE_PACKED holds K u32 (unit << 16 | mask); SPARSE_BITMAP
// holds K u32 byte offsets (2 * unit) then K u16 masks.
__device__ __forceinline__ uint32_t UnitCount(const PostingHead& h) { return h.len / (h.layout == 4 ? 4u : 6u); }

__device__ __forceinline__ uint32_t UnitIndex(const PostingHead& h, uint32_t i) {
    const uint32_t v = reinterpret_cast<const uint32_t*>(h.payload)[i];
    return h.layout == 4 ? (v >> 16) : (v >> 1);
}

__device__ __forceinline__ void UnitAt(const PostingHead& h, uint32_t k, uint32_t i, uint32_t& unit, uint32_t& mask) {
    const uint32_t v = reinterpret_cast<const uint32_t*>(h.payload)[i];
    if (h.layout == 4) {
        unit = v >> 16;
        mask = v & 0xFFFFu;
    } else {
        unit = v >> 1;
        mask = reinterpret_cast<const uint16_t*>(h.payload + 4ull * k)[i];
    }
}

// First index i in [0, k) whose unit is >= unit_lo (per-thread binary search).
__device__ __forceinline__ uint32_t LowerBoundUnit(const PostingHead& h, uint32_t k, uint32_t unit_lo) {
    uint32_t lo = 0, hi = k;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (UnitIndex(h, mid) < unit_lo) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Both lower bounds (first unit >= a, first unit >= b) of a sorted unit list, found by the whole
// workgroup in two rounds of parallel probes instead of a dependent binary search per thread: round 1
// probes every s-th entry (s = ceil(k / threads)) and counts the probes below each target, round 2
// probes the s - 1 entries of the bracket that count selects. Called by every thread of the workgroup.
__device__ __forceinline__ void CoopLowerBounds(const PostingHead& h, uint32_t k, uint32_t a, uint32_t b,
                                                uint32_t& lo, uint32_t& hi) {
    const uint32_t t = threadIdx.x;
    const uint32_t s = (k + kThreads - 1) / kThreads;
    if (s > kThreads) {  // more units than two rounds cover: per-thread search
        lo = LowerBoundUnit(h, k, a);
        hi = LowerBoundUnit(h, k, b);
        return;
    }
    const uint32_t p1 = t * s;
    uint32_t u1 = 0xFFFFFFFFu;
    if (s && p1 < k) u1 = UnitIndex(h, p1);
    const uint32_t ca = __syncthreads_count(p1 < k && u1 < a);
    const uint32_t cb = __syncthreads_count(p1 < k && u1 < b);
    // the bound lies in (p_{c-1}, p_c]: probe the up to s - 1 entries strictly between
    auto round2 = [&](uint32_t c, uint32_t target) {
        if (c == 0) return __syncthreads_count(false) * 0u;
        const uint32_t base = (c - 1) * s + 1;
        const uint32_t p2 = base + t;
        const bool below = t + 1 < s && p2 < k && UnitIndex(h, p2) < target;
        return base + __syncthreads_count(below);
    };
    lo = round2(ca, a);
    hi = round2(cb, b);
}

// 8 bytes of a DENSE_BYTEMAP -> 8 bits (bit i = byte i != 0): fold each byte onto its low bit,
// then gather the eight low bits into the top byte with one multiply.
__device__ __forceinline__ uint32_t PackBytes(uint64_t x) {
    x |= x >> 4;
    x |= x >> 2;
    x |= x >> 1;
    return static_cast<uint32_t>(((x & 0x0101010101010101ull) * 0x0102040810204080ull) >> 56);
}

// One DENSE_BYTEMAP word (64 docs from byte 64 w): bytes at or past len read as 0.
__device__ __forceinline__ uint64_t BytemapWord(const PostingHead& h, uint32_t w) {
    const uint64_t* p = reinterpret_cast<const uint64_t*>(h.payload);
    uint64_t v = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; i++) {
        const uint32_t byte0 = 64 * w + 8 * i;
        if (byte0 >= h.len) break;
        uint64_t x = p[8 * w + i];
        if (h.len - byte0 < 8) x &= (1ull << (8 * (h.len - byte0))) - 1;
        v |= static_cast<uint64_t>(PackBytes(x)) << (8 * i);
    }
    return v;
}

// One operand's 64 bits at segment word w, read directly from HBM (operands without an LDS slot).
__device__ uint64_t OperandWordDirect(const PostingHead& h, uint32_t w) {
    if (h.hits == 0) return 0;
    switch (h.layout) {
        case 0:
            return 8ull * w < h.len ? reinterpret_cast<const uint64_t*>(h.payload)[w] : 0;
        case 1:
        case 4: {
            const uint32_t k = UnitCount(h);
            uint64_t v = 0;
            for (uint32_t i = LowerBoundUnit(h, k, 4 * w); i < k; i++) {
                uint32_t unit, mask;
                UnitAt(h, k, i, unit, mask);
                if (unit >= 4 * w + 4) break;
                v |= static_cast<uint64_t>(mask) << (16 * (unit & 3));
            }
            return v;
        }
        case 2:
            return BytemapWord(h, w);
        case 3: {
            const uint32_t* ids = reinterpret_cast<const uint32_t*>(h.payload);
            const uint32_t n = h.len / 4;
            uint64_t v = 0;
            for (uint32_t i = 0; i < n; i++)
                if ((ids[i] >> 6) == w) v |= 1ull << (ids[i] & 63);
            return v;
        }
    }
    return 0;
}

template <int W>
__global__ __launch_bounds__(kThreads, 4) void FilterKernel(const frk_filter_query* __restrict__ queries,
                                                         const uint32_t* __restrict__ expr,
                                                         const uint64_t* __restrict__ table, uint32_t seg_docs,
                                                         uint64_t n_docs, uint64_t* __restrict__ out,
                                                         uint64_t out_stride, uint32_t* __restrict__ counts,
                                                         uint32_t* __restrict__ status, uint32_t S) {
    constexpr uint32_t C = kThreads * W;  // words per chunk
    extern __shared__ uint64_t slot_words[];  // S slots of C words
    __shared__ uint64_t head_addr[kMaxSlots];
    __shared__ uint32_t head_layout[kMaxSlots], head_len[kMaxSlots], head_hits[kMaxSlots];
    __shared__ uint32_t wave_pop[kThreads / 64];

    const uint32_t q = blockIdx.z;
    const uint32_t seg = blockIdx.y;
    const uint32_t seg_words = seg_docs / 64;
    const uint32_t w0 = blockIdx.x * C;
    const uint32_t t = threadIdx.x;
    const frk_filter_query Q = queries[q];

    // docs of this segment; bits at or past its doc count stay 0
    const uint64_t seg_first = static_cast<uint64_t>(seg) * seg_docs;
    const uint64_t n_seg = n_docs > seg_first ? (n_docs - seg_first < seg_docs ? n_docs - seg_first : seg_docs) : 0;
    auto valid_bits = [&](uint32_t w) -> uint64_t {
        if (64ull * (w + 1) <= n_seg) return ~0ull;
        if (64ull * w < n_seg) return (1ull << (n_seg - 64ull * w)) - 1;
        return 0;
    };

    uint64_t result[W];
    if (Q.flags & FRK_FILTER_MATCH_ALL) {
#pragma unroll
        for (int j = 0; j < W; j++) result[j] = ~0ull;
    } else if (Q.n_postings > FRK_FILTER_MAX_POSTINGS) {
        if (t == 0) atomicOr(status, kStatusTooManyPostings);
#pragma unroll
        for (int j = 0; j < W; j++) result[j] = 0;
    } else {
        const uint64_t* tab = table + Q.table_off + static_cast<uint64_t>(seg) * Q.n_postings;
        const uint32_t n_staged = Q.n_postings < S ? Q.n_postings : S;
        // ---- phase 1: the staged operands' heads to LDS, read in parallel
        if (t < n_staged) {
            head_addr[t] = tab[t];
            const PostingHead h = ReadHead(tab[t]);
            head_layout[t] = h.hits == 0 ? kLayoutEmpty : h.layout;
            head_len[t] = h.len;
            head_hits[t] = h.hits;
        }
        __syncthreads();

        // ---- phase 2: stage every slotted operand's chunk words in rounds over all operands, each
        // round's loads independent of the others', so their latencies overlap:
        //   a. dense operands' W words per thread go through registers into their slots; the slots of
        //      sparse and empty operands are cleared;
        //   b. (barrier) sparse operands scatter into their slots;
        //   c. (barrier) the program reads the slots.
        const bool whole = C >= seg_words;  // the chunk covers the segment: sparse lists need no search
        {
            constexpr int kP = kMaxSlots / W;  // most slots at this W
            uint64_t v[kP][W];
#pragma unroll
            for (int p = 0; p < kP; p++) {
#pragma unroll
                for (int j = 0; j < W; j++) v[p][j] = 0;
                if (static_cast<uint32_t>(p) >= n_staged) continue;
                const uint32_t lay = head_layout[p];
                if (lay != 0 && lay != 2) continue;
                PostingHead h;
                h.payload = reinterpret_cast<const uint8_t*>(head_addr[p]) + 8;
                h.layout = lay;
                h.hits = head_hits[p];
                h.len = head_len[p];
#pragma unroll
                for (int j = 0; j < W; j++) {
                    const uint32_t w = w0 + t + kThreads * j;
                    if (w < seg_words) {
                        if (lay == 0) {
                            if (8ull * w < h.len) v[p][j] = reinterpret_cast<const uint64_t*>(h.payload)[w];
                        } else {
                            v[p][j] = BytemapWord(h, w);
                        }
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < kP; p++) {
                if (static_cast<uint32_t>(p) >= n_staged) continue;
                uint64_t* slot = slot_words + static_cast<size_t>(p) * C;
#pragma unroll
                for (int j = 0; j < W; j++) slot[t + kThreads * j] = v[p][j];
            }
        }
        __syncthreads();
        for (uint32_t p = 0; p < n_staged; p++) {
            const uint32_t lay = head_layout[p];
            if (lay != 1 && lay != 3 && lay != 4) continue;
            PostingHead h;
            h.payload = reinterpret_cast<const uint8_t*>(head_addr[p]) + 8;
            h.layout = lay;
            h.hits = head_hits[p];
            h.len = head_len[p];
            uint64_t* slot = slot_words + static_cast<size_t>(p) * C;
            if (lay == 1 || lay == 4) {
                const uint32_t k = UnitCount(h);
                uint32_t lo = 0, hi = k;
                if (!whole) CoopLowerBounds(h, k, 4 * w0, 4 * (w0 + C), lo, hi);
                for (uint32_t i = lo + t; i < hi; i += kThreads) {
                    uint32_t unit, mask;
                    UnitAt(h, k, i, unit, mask);
                    const uint32_t wd = unit >> 2;
                    if (wd >= w0 && wd < w0 + C)
                        atomicOr(reinterpret_cast<unsigned long long*>(&slot[wd - w0]),
                                 static_cast<unsigned long long>(mask) << (16 * (unit & 3)));
                }
            } else {
                const uint32_t* ids = reinterpret_cast<const uint32_t*>(h.payload);
                const uint32_t n = h.len / 4;
                for (uint32_t i = t; i < n; i += kThreads) {
                    const uint32_t id = ids[i];
                    const uint32_t wd = id >> 6;
                    if (wd >= w0 && wd < w0 + C)
                        atomicOr(reinterpret_cast<unsigned long long*>(&slot[wd - w0]), 1ull << (id & 63));
                }
            }
        }
        __syncthreads();

        // ---- phase 3: the program. It is uniform over the workgroup, so its decoding runs on the scalar
        // unit, and that decoding (repeated per word) bounds this phase. One scalar pass first
        // validates the program and finds its deepest stack (a malformed one yields all-zero words
        // and the status flag, the same verdict in every thread); the evaluation then needs no
        // checks, and a program whose stack stays within kPairDepth entries evaluates two words per
        // pass, with one 2-word vector per stack entry (a 1-D array indexed by the uniform stack
        // pointer, so it stays in registers).
        const uint32_t* prog = expr + Q.expr_off;
        int depth = 0;
        bool bad = !ValidateProgram(prog, Q.expr_len, Q.n_postings, &depth);
        // the evaluation of a validated program: v(e) = op over operands and popped entries, kE words
        auto evaluate = [&](auto ew_c, auto depth_c, int j0) __attribute__((always_inline)) {
            constexpr int kE = decltype(ew_c)::value;
            constexpr int kDepth = decltype(depth_c)::value;
            typedef uint64_t vec __attribute__((ext_vector_type(kE)));
            uint32_t wl[kE];
#pragma unroll
            for (int e = 0; e < kE; e++) wl[e] = t + kThreads * (j0 + e);
            auto operand = [&](uint32_t p) -> vec {
                vec o;
#pragma unroll
                for (int e = 0; e < kE; e++) {
                    if (p < S) {
                        o[e] = slot_words[static_cast<size_t>(p) * C + wl[e]];
                    } else {
                        const uint32_t w = w0 + wl[e];
                        o[e] = w < seg_words ? OperandWordDirect(ReadHead(tab[p]), w) : 0;
                    }
                }
                return o;
            };
            vec stack[kDepth];
            int sp = 0;
            uint32_t pc = 0;
            for (uint32_t i = 0; i < Q.expr_len; i++) {
                const uint32_t op = prog[i];
                vec v;
                if (op == 0) {  // CONJ: AND over parts of the OR of each part's postings
                    const uint32_t parts = prog[i + 1];
                    v = ~vec(0);
                    for (uint32_t x = 0; x < parts; x++) {
                        const uint32_t n = prog[i + 2 + x];
                        vec acc = vec(0);
                        for (uint32_t y = 0; y < n; y++) acc |= operand(pc + y);
                        pc += n;
                        v &= acc;
                    }
                    i += parts + 1;
                } else {
                    const uint32_t pn = prog[i + 1], sn = prog[i + 2];
                    if (op == 3) {
                        v = ~(pn ? operand(pc) : stack[sp - 1]);
                    } else if (op == 1) {
                        v = ~vec(0);
                        for (uint32_t x = 0; x < pn; x++) v &= operand(pc + x);
                        for (uint32_t x = 0; x < sn; x++) v &= stack[sp - 1 - x];
                    } else {
                        v = vec(0);
                        for (uint32_t x = 0; x < pn; x++) v |= operand(pc + x);
                        for (uint32_t x = 0; x < sn; x++) v |= stack[sp - 1 - x];
                    }
                    pc += pn;
                    sp -= static_cast<int>(sn);
                    i += 2;
                }
                stack[sp++] = v;
            }
#pragma unroll
            for (int e = 0; e < kE; e++) result[j0 + e] = stack[0][e];
        };
        constexpr int kPairDepth = FRK_FILTER_MAX_STACK / 2;
        constexpr int kQuadDepth = FRK_FILTER_MAX_STACK / 4;
        if (!bad) {
            if (W == 4 && depth <= kQuadDepth) {
                evaluate(std::integral_constant<int, W == 4 ? 4 : 1>{}, std::integral_constant<int, kQuadDepth>{}, 0);
            } else if (W >= 2 && depth <= kPairDepth) {
#pragma unroll
                for (int j0 = 0; j0 < W; j0 += 2)
                    evaluate(std::integral_constant<int, W >= 2 ? 2 : 1>{}, std::integral_constant<int, kPairDepth>{}, j0);
            } else {
#pragma unroll
                for (int j0 = 0; j0 < W; j0++)
                    evaluate(std::integral_constant<int, 1>{}, std::integral_constant<int, FRK_FILTER_MAX_STACK>{}, j0);
            }
        }
        if (bad) {
            if (t == 0) atomicOr(status, kStatusBadProgram);
#pragma unroll
            for (int j = 0; j < W; j++) result[j] = 0;
        }
    }

    uint32_t pop = 0;
    uint64_t* dst = out + static_cast<uint64_t>(q) * out_stride + static_cast<uint64_t>(seg) * seg_words;
#pragma unroll
    for (int j = 0; j < W; j++) {
        const uint32_t w = w0 + t + kThreads * j;
        if (w < seg_words) {
            const uint64_t r = result[j] & valid_bits(w);
            dst[w] = r;
            pop += static_cast<uint32_t>(__popcll(r));
        }
    }
    // passing-doc count for the query
    for (int off = 32; off > 0; off >>= 1) pop += __shfl_xor(pop, off, 64);
    if ((t & 63) == 0) wave_pop[t >> 6] = pop;
    __syncthreads();
    if (t == 0) {
        uint32_t s = 0;
        for (int i = 0; i < kThreads / 64; i++) s += wave_pop[i];
        if (s) atomicAdd(&counts[q], s);
    }
}

}  // namespace
}  // namespace frk

extern "C" int frk_filter_eval(const frk_filter_query* d_queries, uint32_t n_queries, const uint32_t* d_expr,
                               const uint64_t* d_table, uint32_t n_segments, uint32_t seg_docs, uint64_t n_docs,
                               uint64_t* d_out, uint64_t out_stride, uint32_t* d_counts, uint32_t* d_status,
                               uint32_t max_postings, frk_stream_t stream) {
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
    // W words per thread: the most that still give every operand of the batch's largest program an
    // LDS slot (16 / W slots; all 16 when unknown), and no more than the segment needs. Whole-segment
    // chunks (W = 4 at 131,072 docs) spare the sparse operands their range search. S slots: one per
    // operand, so the LDS shrinks with the programs.
    const uint32_t need = max_postings == 0 || max_postings > 16 ? 16u : max_postings;
    uint32_t w = need <= 4 ? 4u : need <= 8 ? 2u : 1u;
    while (w > 1 && static_cast<uint32_t>(frk::kThreads) * (w / 2) >= seg_words) w /= 2;
    const uint32_t slots = need < 16u / w ? need : 16u / w;
    const uint32_t chunk = frk::kThreads * w;
    const size_t lds = static_cast<size_t>(slots) * chunk * sizeof(uint64_t);
    dim3 grid((seg_words + chunk - 1) / chunk, n_segments, n_queries);
    hipStream_t st = static_cast<hipStream_t>(stream);
    // allow 64 KiB of dynamic LDS for every instantiation, once per process (a function-local static
    // is initialised exactly once even when several host threads launch at the same time)
    static const hipError_t lds_set = [] {
        const int max_lds = frk::kMaxSlots * frk::kThreads * static_cast<int>(sizeof(uint64_t));
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&frk::FilterKernel<4>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&frk::FilterKernel<2>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(&frk::FilterKernel<1>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, max_lds);
        return e;
    }();
    FRK_HIP_CHECK(lds_set);
    switch (w) {
        case 4:
            hipLaunchKernelGGL(frk::FilterKernel<4>, grid, dim3(frk::kThreads), lds, st, d_queries, d_expr, d_table,
                               seg_docs, n_docs, d_out, out_stride, d_counts, d_status, slots);
            break;
        case 2:
            hipLaunchKernelGGL(frk::FilterKernel<2>, grid, dim3(frk::kThreads), lds, st, d_queries, d_expr, d_table,
                               seg_docs, n_docs, d_out, out_stride, d_counts, d_status, slots);
            break;
        default:
            hipLaunchKernelGGL(frk::FilterKernel<1>, grid, dim3(frk::kThreads), lds, st, d_queries, d_expr, d_table,
                               seg_docs, n_docs, d_out, out_stride, d_counts, d_status, slots);
            break;
    }
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

extern "C" int frk_check_program(const uint32_t* expr, uint32_t expr_len, uint32_t n_postings, uint32_t* max_depth) {
    if (!expr && expr_len) {
        frk::SetError("frk_check_program: null program");
        return FRK_ERR_INVALID;
    }
    int depth = 0;
    if (n_postings > FRK_FILTER_MAX_POSTINGS || !frk::ValidateProgram(expr, expr_len, n_postings, &depth)) {
        frk::SetError("frk_check_program: malformed program (%u words, %u postings)", expr_len, n_postings);
        return FRK_ERR_INVALID;
    }
    if (max_depth) *max_depth = static_cast<uint32_t>(depth);
    return FRK_OK;
}
