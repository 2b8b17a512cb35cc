"""frk_mlp_forward_noalloc -- the floor model's inference with its workspace in a fixed device array (for
callers that cannot allocate) -- against frk_mlp_forward with a caller workspace: the same kernels, so
the outputs must be bit-identical, on two calls in a row (the second with fewer rows, so a stale
activation of the first call would show), and sizes beyond the fixed array are refused."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_noalloc_forward_matches_workspace_forward():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hybridsearchsystem_amd import learned
    from hybridsearchsystem_amd import lib
    L = lib()
    rng = np.random.default_rng(5)
    model = learned.FloorModel(256, hidden=256, seed=3)
    for k in ("b1", "b2", "b3"):
        model.p[k] = torch.from_numpy(rng.standard_normal(model.p[k].numel()).astype(np.float32) * 0.05).cuda()
    p = model.params_struct()
    for rows in (1000, 133):
        print("noalloc forward: %d rows" % rows, flush=True)
        x = torch.from_numpy(rng.standard_normal((rows, model.in_dim)).astype(np.float32)).cuda()
        want = model.forward(x).clone()
        got = torch.full((rows,), float("nan"), device="cuda")
        rc = L.frk_mlp_forward_noalloc(ctypes.byref(p), rows, x.data_ptr(), model.in_dim, got.data_ptr(), None)
        assert rc == 0, L.frk_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, want), rows
    # beyond the fixed array: refused with a capacity error, nothing launched
    x = torch.zeros((1025, model.in_dim), device="cuda")
    y = torch.zeros(1025, device="cuda")
    assert L.frk_mlp_forward_noalloc(ctypes.byref(p), 1025, x.data_ptr(), model.in_dim, y.data_ptr(), None) == -4
    big = learned.FloorModel(256, hidden=512, seed=3)
    pb = big.params_struct()
    assert L.frk_mlp_forward_noalloc(ctypes.byref(pb), 8, x.data_ptr(), big.in_dim, y.data_ptr(), None) == -4
