"""bench.py's host-side accounting, without a GPU: the algorithmic bytes its roofline divides by the score
kernel's time, and the committed counter summary it reads the kernel's HBM traffic from."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

FP8, BF16 = 2, 1


def test_kernel_roofline_counts_each_input_once():
    # 10M x 256 FP8 at 128 queries: every doc tile once (1 byte per element), the per-doc filter words of
    # 128 queries (4 bytes per doc and 32-query group), one 4-byte scale per doc, 8 bytes per candidate key;
    # the time is the mean over the batches' score-kernel events
    stats = [{"candidates": 1000, "ms_score": 0.5}, {"candidates": 3000, "ms_score": 0.7}]
    alg, ms, gbs = bench.kernel_roofline(stats, 10_000_000, 256, 128, FP8)
    nb = 312_500
    want = nb * 32 * 256 + nb * 32 * 4 + nb * 128 * 4 + 2000 * 8
    assert alg == want
    assert abs(ms - 0.6) < 1e-12
    assert abs(gbs - want / 0.6e-3 / 1e9) < 1e-6
    # a ragged last block counts whole (the store is tiled in 32-doc blocks), 100 queries pad the filter
    # words to 4 groups of 32, and a BF16 store has 2-byte elements and no scales
    alg, _, _ = bench.kernel_roofline([{"candidates": 0, "ms_score": 1.0}], 1000, 256, 100, BF16)
    assert alg == 32 * 32 * 256 * 2 + 32 * 128 * 4


def test_committed_fp8_traffic_profile_matches_the_algorithmic_bytes():
    # the profile's calibrated FETCH_SIZE + WRITE_SIZE per launch of the headline score kernel: each doc
    # byte, filter word and scale read from HBM once, plus the candidate writes (under 1.5% over)
    t = bench.traffic_from_profile(os.path.join(ROOT, "profiles", "score_fp8_10Mx256_q128_traffic.md"))
    assert t is not None
    alg, _, _ = bench.kernel_roofline([{"candidates": 0, "ms_score": 1.0}], 10_000_000, 256, 128, FP8)
    assert 1.0 <= t / alg <= 1.015, t / alg
    assert bench.traffic_from_profile(os.path.join(ROOT, "profiles", "no_such_summary.md")) is None
