"""Scorer micro-benchmark: ``frk_score_select`` alone over a synthetic store, per dtype, batch size and floor.

    python -m hybridsearchsystem_amd.scorebench --docs 10000000 --dim 768 --dtypes bf16,fp8 --queries 64,128,256

Docs and queries are random unit vectors tiled on the device (``frk_tile_rows_device``); every doc passes
a random half of the queries (per-doc filter words drawn at random). Each line reports the kernel time from
HIP events over ``--reps`` launches on the current stream (each after a counter reset), the doc-store bytes
read per launch over that time, and the mean survivors per query under the floor (``inf``: none survive, the epilogue's fast path
only). The launch picks the scorer variant as the search pipeline does (``ScoreKernel`` or, for wide BF16
and FP8 docs and large batches, ``ScoreStagedKernel``).
"""
import argparse
import sys

from . import FRK_DTYPE_BF16, FRK_DTYPE_F16X2, FRK_DTYPE_F32, FRK_DTYPE_FP8E4M3, check, lib, stream_ptr
from . import kernels as K

DTYPES = {"f32": FRK_DTYPE_F32, "bf16": FRK_DTYPE_BF16, "fp8": FRK_DTYPE_FP8E4M3, "f16x2": FRK_DTYPE_F16X2}
ELEM_BYTES = {FRK_DTYPE_F32: 4, FRK_DTYPE_BF16: 2, FRK_DTYPE_FP8E4M3: 1, FRK_DTYPE_F16X2: 4}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--docs", type=int, default=10_000_000)
    ap.add_argument("--dim", type=int, default=768)
    ap.add_argument("--dtypes", default="bf16,fp8")
    ap.add_argument("--queries", default="64,128,256", help="batch sizes")
    ap.add_argument("--floors", default="inf", help="per-query floors to time (inf: no survivors)")
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args(argv)
    import torch
    L = lib()
    torch.manual_seed(0)
    n, dim = args.docs, args.dim
    nb = (n + 31) // 32
    for name in args.dtypes.split(","):
        dt = DTYPES[name.strip()]
        dp = K.padded_dim(dim, dt)
        tiles = torch.empty(int(L.frk_tiles_bytes(n, dp, dt)), dtype=torch.uint8, device="cuda")
        scales = torch.empty(n, dtype=torch.float32, device="cuda") if dt in (FRK_DTYPE_FP8E4M3, FRK_DTYPE_F16X2) else None
        chunk = 1 << 20
        for at in range(0, n, chunk):
            m = min(chunk, n - at)
            v = torch.randn(m, dim, device="cuda")
            v /= v.norm(dim=1, keepdim=True)
            check(L.frk_tile_rows_device(v.data_ptr(), m, dim, at, dt, dp, tiles.data_ptr(),
                                         scales.data_ptr() if scales is not None else None, stream_ptr()), "tile")
        torch.cuda.synchronize()
        store = K.doc_store(tiles, scales, dt, dp, n)
        doc_bytes = nb * 32 * dp * ELEM_BYTES[dt] + (4 * n if scales is not None else 0)
        for q in [int(x) for x in args.queries.split(",")]:
            qv = torch.randn(q, dp, device="cuda")
            qv[:, dim:] = 0
            qv /= qv.norm(dim=1, keepdim=True)
            q_stride = (q + 31) // 32 * 32
            words = torch.randint(0, 2 ** 31, (nb, q_stride), dtype=torch.int64, device="cuda").to(torch.int32)
            words ^= torch.randint(0, 2, (nb, q_stride), dtype=torch.int32, device="cuda") << 31
            cap = 1 << 17
            cand = torch.empty(q * cap, dtype=torch.int64, device="cuda")
            counts = torch.zeros(q * 32, dtype=torch.int32, device="cuda")
            for fl in [float(x) for x in args.floors.split(",")]:
                floor = torch.full((q,), fl, dtype=torch.float32, device="cuda")
                ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                for i in range(args.reps + 3):
                    if i == 3:
                        ev0.record()
                    counts.zero_()
                    K.score_select(store, qv, q, words, q_stride, floor, 0, 1, nb, cand, cap, counts, counts_stride=32)
                ev1.record()
                torch.cuda.synchronize()
                ms = ev0.elapsed_time(ev1) / args.reps
                surv = counts.view(q, 32)[:, 0].float().mean().item()
                print("%s %dx%d queries %4d floor %5s: %.4f ms  %.2f TB/s doc bytes  survivors/query %.0f"
                      % (name, n, dim, q, fl, ms, doc_bytes / ms / 1e9, surv), flush=True)
        del tiles, scales
    return 0


if __name__ == "__main__":
    sys.exit(main())
