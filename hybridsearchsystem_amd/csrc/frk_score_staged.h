// frk_score_staged.h -- the fused scorer for wide docs and large batches (part of frk_score_impl.h,
// which includes it after the shared pieces: filter words, accumulator init, epilogues, survivor queue).
//
// ScoreKernel keeps the query tile in LDS and streams doc tiles into each wave's registers; a 768-dim
// BF16 tile of 96 queries already takes 144 KiB, so a larger batch runs as several query groups that
// each read the doc range again. Here the roles swap: every wave holds one 32-query row block's
// fragments in registers for the whole launch (4 registers per 16-byte chunk, so up to 48 chunks), and
// the workgroup stages each 32-doc block's tile in LDS once, with direct global->LDS loads, for all of
// its waves. A workgroup of W waves thus scores 32 W queries (256 with 8 waves) from one read of the doc
// range. Per block every wave reads the whole tile from LDS (W x CH KiB), twice the HBM bytes at 8
// waves and 768 BF16 dims, which the LDS sustains; the MFMA chain of a wave's row block and the
// epilogues are those of ScoreKernel, so scores, survivors and block maxima are bit-identical.
//
// Pipeline per workgroup (contiguous block range, all waves on the same block): two LDS tile buffers;
// at the top of iteration i one barrier (after each wave's loads have landed) makes tile i visible and
// frees buffer (i + 1) % 2, whose tile i + 1 is then requested together with the filter words (and
// FP8 doc scales) of block i + 1, so one block's loads are in flight under each block's MFMAs.
//
// Measured on 10M docs x 768 dims, no survivors (ScoreKernel -> this): BF16 256 queries 5.12 -> 3.76
// ms, 192 3.58 -> 3.37, 128 2.74 -> 2.69, 96 2.54 -> 2.66 (so BF16 starts above 96); FP8 256 2.31 ->
// 2.06, 192 1.96 -> 1.85, 128 1.35 -> 1.31, 96 1.34 -> 1.29. At 256 BF16 queries the kernel is bound
// by the MFMAs, not by HBM: 2 x 256 x 10M x 768 flops in 3.76 ms is 1.05 PFLOP/s, the MFMA pipes are
// busy 63% of the cycles, and the clock under this load is ~1.5 GHz. The same loop without epilogue
// runs 2.66 ms on constant tiles (2.1 GHz) but 3.87 ms on random ones, and 2.50 ms (HBM-bound) with
// half its MFMAs; halving its LDS reads instead gains 5%. With 192 queries six row blocks share four
// SIMDs, two of which run two, so a block costs nearly what it does at 256. A third tile buffer and
// two row blocks per wave (4 waves, 512 registers each) both measured slower; running the epilogue of
// one of the two waves on each SIMD at the other end of the iteration made no difference.
//
// At 256 and 512 dims ScoreKernel holds 128 queries in one LDS tile, so the doc range is read once
// either way; the staged loop is still as fast or faster from 65 queries (10M docs, no survivors,
// ScoreKernel -> this): BF16 256 dims 96 queries 0.904 -> 0.882 ms, 128 0.906 -> 0.895, 192 1.297 ->
// 1.134, 256 1.363 -> 1.287; BF16 512 dims 96 2.155 -> 1.684, 128 1.729 -> 1.711, 192 and 256 equal;
// FP8 256 dims 96 0.487 -> 0.473, 128 0.495 -> 0.489, 192 0.727 -> 0.648, 256 0.731 -> 0.741; BF16 384
// dims 72 1.456 -> 1.352 ms, 80 1.284 -> 1.267, 96 1.305 -> 1.264 (so 65 there too). In the
// BF16 + learned-floor pipeline at 128 queries (~5,000 survivors per query) the BF16 256-dim pass
// takes 0.978 -> 0.930 ms.
//
// Tried for the 8 and 16 KiB tiles of 256-dim docs and left out: three or four tile buffers, tiles
// requested two or three blocks ahead with the wait at the top of each block leaving the newest tile
// in flight (a bare s_barrier, since __syncthreads' workgroup fence waits for every load): FP8 at 128
// queries 0.497 -> 0.509 / 0.519 ms with no survivors and 0.519 -> 0.534 / 0.565 ms with ~3,100 per
// query, BF16 0.901 -> 0.944 / 0.933 ms. Fewer workgroups fit a CU, and the compiler makes every LDS
// write of the survivor queue wait for all of the wave's loads (vmcnt(0): it cannot tell the queue
// from the tile buffers the global->LDS loads are filling), which then includes the deeper prefetch.
// Queuing a block's survivors at the top of the next iteration instead (after the wait, before the next
// request; the scores kept in 16 more registers) measured slower too: FP8 at 128 queries with ~3,100
// survivors per query 0.539 -> 0.557 ms, with ~15,600 0.575 -> 0.635 ms.
#pragma once

namespace frk {

// Shapes the staged scorer is built for: BF16 stores with 16-48 chunks per doc (256-768 dims) and FP8
// ones with 8 (256 dims) or 24-32 (768-1024 dims). Their query fragments (4 CH registers) leave room
// for the accumulators and the LDS read ring under the 256 registers of two waves per SIMD (FP8 at 48
// chunks spills there). FP8 at 16 chunks (512 dims) measured no faster than ScoreKernel from 96
// queries up (0.908 -> 0.903 ms at 96, 1.25 -> 1.28 ms at 192, 1.28 -> 1.41 ms at 256).
template <int DT, int CH>
constexpr bool kStagedShape = CH % 8 == 0 && ((DT == FRK_DTYPE_BF16 && CH >= 16 && CH <= 48) ||
                                              (DT == FRK_DTYPE_FP8E4M3 && (CH == 8 || (CH >= 24 && CH <= 32))));

// Smallest batches the staged scorer takes (the measurements above and below): BF16 at 48 chunks from
// 97 queries, the narrower BF16 shapes and FP8 from 65. At 64 queries ScoreKernel is as fast or faster
// in every shape (10M docs, no survivors: BF16 256 dims 0.866 vs 0.873 ms, 512 dims 1.662 vs 1.676).
template <int DT, int CH>
constexpr uint32_t kStagedMinQueries = DT == FRK_DTYPE_BF16 && CH > 32 ? 97u : 65u;

// This wave's query fragments for row block q_row0 .. q_row0 + 31, converted exactly as StageQueries
// converts them: lane l holds row q_row0 + l % 32, chunk c's elements of lane half l / 32. FP8: the
// row's scale is Fp8Scale(max |q|) over the padded row (each lane half sees half of the row's dims;
// the halves combine with one exchange), returned with its E8M0 exponent for the scaled MFMA.
// F32ToBf16 without a branch (both results computed, one selected): the same bits for every input.
__device__ __forceinline__ uint32_t Bf16Select(float f) {
    const uint32_t u = __float_as_uint(f);
    const uint32_t special = (u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u);  // infinity or NaN (quieted)
    const uint32_t rounded = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;    // round to nearest even
    return (u & 0x7F800000u) == 0x7F800000u ? special : rounded;
}

template <int DT, int CH>
__device__ __forceinline__ void StageRowBlock(const ScoreArgs& A, uint32_t q_row0, u32x4 (&a)[CH], float& q_scale,
                                              int& q_exp) {
    constexpr int E = Traits<DT>::kElemsPerLane;
    // chunks converted per round: 32 loaded values in registers at a time beside the fragments (the last
    // rounds of a 48-chunk row block run with 176 fragment registers live)
    constexpr int kBatch = 32 / E;
    static_assert(CH % kBatch == 0, "whole rounds of chunks");
    const int lane = threadIdx.x & 63;
    const uint32_t q = q_row0 + (lane & 31);
    const int h = lane >> 5;
    const bool live = q < A.n_queries;
    const floatx4* row = reinterpret_cast<const floatx4*>(A.queries + static_cast<uint64_t>(live ? q : 0) * A.q_dim);
    // rows past the batch read row 0 (the loads stay unconditional, so each round is straight-line
    // code) and take zeros
    auto elems = [&](const floatx4* r, int c, float (&v)[E]) __attribute__((always_inline)) {
        const int d0 = c * 2 * E + h * E;
#pragma unroll
        for (int e4 = 0; e4 < E / 4; e4++) {
            const floatx4 x = r[d0 / 4 + e4];
#pragma unroll
            for (int e = 0; e < 4; e++) v[4 * e4 + e] = live ? x[e] : 0.0f;
        }
    };
    q_scale = 1.0f;
    q_exp = 127;
    float inv = 1.0f;
    if constexpr (DT == FRK_DTYPE_FP8E4M3) {
        float mx = 0.0f;
#pragma unroll 4
        for (int c = 0; c < CH; c++) {
            float v[E];
            elems(row, c, v);
#pragma unroll
            for (int e = 0; e < E; e++) mx = fmaxf(mx, fabsf(v[e]));
        }
        mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
        q_scale = Fp8Scale(mx);
        q_exp = static_cast<int>((__float_as_uint(q_scale) >> 23) & 0xFFu);
        inv = 1.0f / q_scale;
    }
    // the row is read again below: an opaque copy of its address keeps the compiler from holding every
    // value of the first pass in registers for the second
    const floatx4* row2 = row;
    asm volatile("" : "+v"(row2));
#pragma unroll
    for (int c0 = 0; c0 < CH; c0 += kBatch) {
        float v[kBatch][E];
#pragma unroll
        for (int k = 0; k < kBatch; k++) elems(row2, c0 + k, v[k]);
#pragma unroll
        for (int k = 0; k < kBatch; k++) {
            u32x4 packed;
            if constexpr (DT == FRK_DTYPE_BF16) {
#pragma unroll
                for (int e = 0; e < 4; e++)
                    packed[e] = Bf16Select(v[k][2 * e]) | (Bf16Select(v[k][2 * e + 1]) << 16);
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    int w = __builtin_amdgcn_cvt_pk_fp8_f32(v[k][4 * e] * inv, v[k][4 * e + 1] * inv, 0, false);
                    w = __builtin_amdgcn_cvt_pk_fp8_f32(v[k][4 * e + 2] * inv, v[k][4 * e + 3] * inv, w, true);
                    packed[e] = static_cast<uint32_t>(w);
                }
            }
            // converted here (not where the fragment is first used), so the round's raw values die here
            asm volatile("" : "+v"(packed));
            a[c0 + k] = packed;
        }
        // one round's loads at a time: the next round's address is made to depend on this round's last
        // fragment (these empty statements keep their order), so no load of it is issued before this
        // round's values are converted
        asm volatile("" : "+v"(row2) : "v"(a[c0 + kBatch - 1]));
    }
}

// s_waitcnt with vmcnt(0) and the LDS/export counters left alone (gfx9 encoding: vmcnt bits 3:0 and
// 15:14, expcnt 6:4, lgkmcnt 11:8)
constexpr int kWaitVmcnt0 = 0x0F70;

template <int DT, int CH, int W, int kMode>
__global__ __launch_bounds__(W * 64, 1) void ScoreStagedKernel(ScoreArgs A) {
    static_assert(kStagedShape<DT, CH>, "not a staged shape");
    static_assert(kMode != kModeDense, "the score matrix is written by ScoreKernel");
    static_assert(CH % W == 0, "every wave requests the same number of chunks per tile");
    extern __shared__ u32x4 smem[];
    u32x4* tile_lds = smem;                                                   // [2][CH][64]
    float* q_scale_lds = reinterpret_cast<float*>(smem + 2 * CH * 64);        // [32 W]
    float* floor_lds = q_scale_lds + 32 * W;                                  // [32 W]
    uint8_t* queue_lds = reinterpret_cast<uint8_t*>(floor_lds + 32 * W);      // [W][kQueueBytes]
    // workgroup -> (doc range, query group) as in ScoreKernel: the groups of one range on one XCD
    const uint32_t groups = A.n_groups;
    const uint32_t n_ranges = gridDim.x / groups;
    uint32_t range_x, group;
    if (groups == 1) {
        range_x = blockIdx.x;
        group = 0;
    } else {
        const uint32_t xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
        group = slot % groups;
        range_x = (slot / groups) * kXcds + xcd;
    }
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t q_row0 = group * (32 * W) + 32 * wave;  // this wave's row block
    const bool active = q_row0 < A.n_queries;               // uniform over the wave
    u32x4 a[CH];
    float q_scale;
    int q_exp;
    StageRowBlock<DT, CH>(A, q_row0, a, q_scale, q_exp);
    if (lane < 32) {
        const uint32_t q = q_row0 + lane;
        q_scale_lds[32 * wave + lane] = q_scale;
        floor_lds[32 * wave + lane] = q < A.n_queries ? (A.floor ? A.floor[q] : -INFINITY) : INFINITY;
    }
    WaveQueue Qu;
    Qu.ent = reinterpret_cast<uint2*>(queue_lds + wave * kQueueBytes);
    Qu.row = queue_lds + wave * kQueueBytes + kQueueKeys * sizeof(uint2);
    Qu.row_at = reinterpret_cast<uint32_t*>(Qu.row + kQueueKeys);
    Qu.n = 0;
    const uint64_t per_wg = (A.block_count + n_ranges - 1) / n_ranges;
    const uint64_t i_lo = static_cast<uint64_t>(range_x) * per_wg;
    const uint64_t i_hi = i_lo + per_wg < A.block_count ? i_lo + per_wg : A.block_count;
    if (i_lo >= i_hi) return;  // the whole workgroup: no barrier below is reached by part of it
    const uint64_t n = i_hi - i_lo;
    auto blk_of = [&](uint64_t i) { return A.block_begin + (i_lo + i) * A.block_step; };
    // this wave's share of block i's tile (chunks wave, wave + W, ...): one 1 KiB global->LDS load per
    // chunk, lane l's 16 bytes landing at lane offset l of the chunk's LDS slot (the tile layout as stored)
    auto request = [&](uint64_t i, uint32_t buf) __attribute__((always_inline)) {
        const u32x4* src = A.tiles + blk_of(i) * CH * 64 + lane;
        u32x4* dst = tile_lds + buf * (CH * 64);
#pragma unroll
        for (int c = 0; c < CH / W; c++) {
            const int ch = wave + c * W;
            __builtin_amdgcn_global_load_lds(static_cast<const void*>(src + ch * 64),
                                             (__attribute__((address_space(3))) void*)(dst + ch * 64), 16, 0, 0);
        }
    };
    // block i's filter word (bit x: query q_row0 + x passes doc l % 32) and FP8 doc scale, taken as
    // loaded: a select on the value would make the compiler wait for it (and so for every tile load
    // issued before it) right after the load. The word of a wave past the batch is never used; it reads
    // word 0 of the block's row instead of one past it.
    const uint32_t m_off = A.bmask_stride && active ? q_row0 + (lane & 31) : 0u;
    auto load_words = [&](uint64_t i, uint32_t& m, float& sc) __attribute__((always_inline)) {
        const uint64_t blk = blk_of(i);
        m = A.bmask[blk * A.bmask_stride + m_off];
        if constexpr (DT == FRK_DTYPE_FP8E4M3) {
            const uint64_t doc = blk * 32 + (lane & 31);
            sc = A.doc_scale[doc < A.n_docs ? doc : A.n_docs - 1];
        }
    };
    floatx16 acc[1];
    auto epilogue = [&](uint64_t i) __attribute__((always_inline)) {
        if constexpr (kMode == kModeBlockMax)
            EpilogueBlockMax<DT, 1>(A, q_scale_lds + 32 * wave, q_row0, i_lo + i, 1.0f, acc);
        else
            Epilogue<DT, 1>(A, floor_lds + 32 * wave, q_scale_lds + 32 * wave, q_row0, blk_of(i), 1.0f, acc, Qu);
    };
    uint32_t m_next = 0;
    float s_next = 1.0f;
    request(0, 0);
    load_words(0, m_next, s_next);
    for (uint64_t i = 0; i < n; i++) {
        // every wave's loads of tile i have landed (this wave's by the wait, the others' by the barrier
        // after theirs), and every wave is done with tile i - 1, whose buffer is requested next
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        __syncthreads();
        const uint32_t m_cur[1] = {m_next};
        const float s_cur = s_next;
        if (i + 1 < n) {
            request(i + 1, static_cast<uint32_t>((i + 1) & 1));
            load_words(i + 1, m_next, s_next);
        }
        if (!active) continue;
        const uint64_t blk = blk_of(i);
        InitBias<1>(m_cur, blk * 32 + (lane & 31) < A.n_docs, acc);
        const int d_exp = DT == FRK_DTYPE_FP8E4M3 ? static_cast<int>((__float_as_uint(s_cur) >> 23) & 0xFFu) : 127;
        const u32x4* t = tile_lds + (i & 1) * (CH * 64) + lane;
        // the tile's chunks through a 4-deep register ring: the LDS read of chunk c + 3 is issued before
        // chunk c's MFMA (FP8: chunk pairs, a 3-deep ring of pairs)
        if constexpr (DT == FRK_DTYPE_BF16) {
            u32x4 b[4];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                b[c] = t[c * 64];
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int c = 0; c < CH; c++) {
                if (c + 3 < CH) b[(c + 3) % 4] = t[(c + 3) * 64];
                __builtin_amdgcn_sched_barrier(0);
                MfmaChunk<DT>(a[c], b[c % 4], acc[0]);
                __builtin_amdgcn_sched_barrier(0);
            }
        } else {
            constexpr int P = CH / 2;
            u32x4 b[3][2];
#pragma unroll
            for (int p = 0; p < 2; p++) {
                b[p][0] = t[(2 * p) * 64];
                b[p][1] = t[(2 * p + 1) * 64];
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int p = 0; p < P; p++) {
                if (p + 2 < P) {
                    b[(p + 2) % 3][0] = t[(2 * p + 4) * 64];
                    b[(p + 2) % 3][1] = t[(2 * p + 5) * 64];
                }
                __builtin_amdgcn_sched_barrier(0);
                MfmaChunkPairFp8(a[2 * p], a[2 * p + 1], b[p % 3][0], b[p % 3][1], acc[0], q_exp, d_exp);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        epilogue(i);
    }
    if constexpr (kMode == kModeSelect) {
        if (active) FlushQueue(A, q_row0, Qu, nullptr);
    }
}

template <int DT, int CH, int W, int kMode>
int LaunchStaged(const ScoreArgs& A, hipStream_t stream, uint32_t grid_x) {
    constexpr int kThreads = W * 64;
    const size_t lds = static_cast<size_t>(2) * CH * 64 * sizeof(u32x4) + 2 * 32 * W * sizeof(float) + W * kQueueBytes;
    const void* fn = reinterpret_cast<const void*>(&ScoreStagedKernel<DT, CH, W, kMode>);
    struct Residency {
        hipError_t err;
        int per_cu;
    };
    // once per process (thread-safe initialisation of a function-local static)
    static const Residency res = [&] {
        Residency r{hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)), 1};
        int occ = 0;
        if (r.err == hipSuccess) r.err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kThreads, lds);
        r.per_cu = occ > 0 ? occ : 1;
        return r;
    }();
    FRK_HIP_CHECK(res.err);
    int cus = 0;
    if (int rc = CurrentDeviceCUs(&cus)) return rc;
    const uint32_t groups = (A.n_queries + 32 * W - 1) / (32 * W);
    const uint32_t resident = static_cast<uint32_t>(cus > 0 ? cus : 1) * static_cast<uint32_t>(res.per_cu);
    const uint32_t per_group = resident / groups > 0 ? resident / groups : 1;
    if (grid_x > per_group) grid_x = per_group;
    if (groups > 1) grid_x = grid_x >= kXcds ? grid_x / kXcds * kXcds : kXcds;
    ScoreArgs a = A;
    a.n_groups = groups;
    hipLaunchKernelGGL((ScoreStagedKernel<DT, CH, W, kMode>), dim3(grid_x * groups), dim3(kThreads), lds, stream, a);
    FRK_HIP_CHECK(hipGetLastError());
    return FRK_OK;
}

// Launches the staged scorer where it serves the batch (*done set): 4-wave workgroups up to 128 queries,
// 8-wave ones (256 queries per group) above.
template <int DT, int CH, int kMode>
int TryStaged(const ScoreArgs& A, hipStream_t stream, uint32_t grid_x, bool* done) {
    *done = false;
    if constexpr (kStagedShape<DT, CH> && kMode != kModeDense) {
        if (A.n_queries >= kStagedMinQueries<DT, CH>) {
            *done = true;
            if (A.n_queries <= 128) return LaunchStaged<DT, CH, 4, kMode>(A, stream, grid_x);
            return LaunchStaged<DT, CH, 8, kMode>(A, stream, grid_x);
        }
    }
    return FRK_OK;
}

}  // namespace frk
