"""The learned score floor: the MLP oracle pinned by finite differences, the HIP BF16 training and
inference kernels against it, and the trained floor inside the searcher (exact results, fewer
survivors than without it)."""
import numpy as np
import pytest

from oracle import mlp as omlp

torch = pytest.importorskip("torch")


# ----------------------------------------------------------------------------- oracle (CPU)
def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = {k: v.astype(np.float64) for k, v in omlp.init_params(6, 5, seed=1).items()}
    p["b1"] = rng.standard_normal(5) * 0.1
    p["b2"] = rng.standard_normal(5) * 0.1
    p["b3"] = np.array([0.05])
    x = rng.standard_normal((7, 6))
    lab = rng.standard_normal(7)
    _, g, _ = omlp.loss_and_grads(p, x, lab, 0.1, round_bf16=False)
    for k in p:
        for i in np.ndindex(p[k].shape):
            q = {kk: vv.copy() for kk, vv in p.items()}
            q[k][i] += 1e-6
            lp = omlp.loss_and_grads(q, x, lab, 0.1, False)[0]
            q[k][i] -= 2e-6
            lm = omlp.loss_and_grads(q, x, lab, 0.1, False)[0]
            assert abs((lp - lm) / 2e-6 - g[k][i]) < 1e-7, (k, i)


def test_oracle_pinball_minimiser_is_the_tau_quantile():
    """What the mathematics fixes about the pinball loss, independent of how the oracle writes it: over
    a sample of n labels, the constant prediction that minimises the mean loss at level tau is the
    tau-quantile of the sample, the order statistic s[ceil(tau n) - 1] (the subgradient
    (#(l < y) - tau n) / n changes sign there). A tau <-> 1 - tau swap, a sign error in the residual or
    a dropped (tau - 1) factor each move the minimiser to another order statistic."""
    rng = np.random.default_rng(3)
    n = 2001
    lab = rng.standard_normal(n) * 2.0 + 0.5
    s = np.sort(lab)
    for tau in (0.02, 0.3, 0.9):
        j = int(np.ceil(tau * n)) - 1
        assert s[j] == np.quantile(lab, tau, method="inverted_cdf")
        losses = np.array([omlp.pinball(np.full(n, y), lab, tau)[0] for y in s])
        assert int(np.argmin(losses)) == j, tau
        # the loss is piecewise linear with these slopes on either side of the minimiser
        for a, b, slope in ((s[j - 1], s[j], j / n - tau), (s[j], s[j + 1], (j + 1) / n - tau)):
            y, eps = 0.5 * (a + b), 0.4 * (b - a)  # the loss is linear between neighbouring labels
            fd = (omlp.pinball(np.full(n, y + eps), lab, tau)[0] - omlp.pinball(np.full(n, y - eps), lab, tau)[0]) / (2 * eps)
            assert abs(fd - slope) < 1e-5, (tau, fd, slope)
            # the returned gradient (per row, mean over rows) sums to the same slope
            assert abs(omlp.pinball(np.full(n, y), lab, tau)[1].sum() - slope) < 1e-12
    # a single residual: tau * r above the label, (1 - tau) * |r| below
    assert np.isclose(omlp.pinball([0.0], [2.0], 0.1)[0], 0.2)
    assert np.isclose(omlp.pinball([2.0], [0.0], 0.1)[0], 1.8)


def test_oracle_bf16_rounding_matches_torch():
    x = np.concatenate([np.random.default_rng(1).standard_normal(100000).astype(np.float32) * 10.0 ** np.arange(-5, 5).repeat(10000),
                        np.array([1.00390625, 1.005859375, -1.00390625, 65504.0, np.inf, -np.inf], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(omlp.bf16(x), want)


def _tree_of(expr):
    """Nested (op, children|tag) of a grammar string via the oracle's parser (pre-order pairs)."""
    import oracle
    pairs = oracle.parse_filter(expr)

    def walk(i):
        typ, arg = int(pairs[i][0]), int(pairs[i][1])
        if typ == 0:
            return i + 1, ("t", arg)
        kids, j = [], i + 1
        for _ in range(arg):
            j, c = walk(j)
            kids.append(c)
        return j, ({1: "and", 2: "or", 3: "not"}[typ], kids)
    return walk(0)[1]


def _expect(node, dens):
    """(selectivity, depth, [and, or, not, terms], term densities) by the independence rule."""
    op = node[0]
    if op == "t":
        return dens[node[1]], 1, np.array([0, 0, 0, 1]), [dens[node[1]]]
    parts = [_expect(c, dens) for c in node[1]]
    cnt = sum(p[2] for p in parts) + np.array([op == "and", op == "or", op == "not", 0])
    terms = [d for p in parts for d in p[3]]
    sels = [p[0] for p in parts]
    sel = 1 - sels[0] if op == "not" else (np.prod(sels) if op == "and" else 1 - np.prod([1 - x for x in sels]))
    return sel, 1 + max(p[1] for p in parts), cnt, terms


@pytest.mark.gpu
def test_floor_features_follow_independence_estimates():
    """Features the library computes for the searcher, against the independence rule on the table's
    actual per-tag doc counts (the estimate itself is a model input, not a claimed truth)."""
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hybridsearchsystem_amd import search, synth
    n, n_tags = 20000, 16
    rng = np.random.default_rng(9)
    dens = synth.loguniform_densities(rng, n_tags, lo=0.001, hi=0.5)
    off, ids = synth.gen_postings(rng, n, dens)
    actual = np.diff(off) / n
    docs = rng.standard_normal((n, 32)).astype(np.float32)
    t = search.Table.build(docs, off, ids, docs_per_segment=8192)
    filters, _ = synth.random_filters(10, dens, 60, depth=3, sel=(1e-4, 0.9))
    filters = ["", "3"] + filters
    ks = np.full(len(filters), 40, np.uint32)
    qs = rng.standard_normal((len(filters), 32)).astype(np.float32)
    f = t.floor_features(qs, filters, ks)
    for i, expr in enumerate(filters):
        if expr == "":
            assert f[i, 14] == 1.0 and f[i, 0] == 0.0
            continue
        sel, depth, cnt, terms = _expect(_tree_of(expr), actual)
        sel = min(max(sel, 1e-9), 1.0)
        assert np.isclose(f[i, 0], np.log10(sel), atol=1e-5), expr
        assert np.isclose(f[i, 1], np.log10(40.0), atol=1e-6)
        assert np.isclose(f[i, 2], np.log10(max(sel * n, 1.0)), atol=1e-5)
        assert f[i, 4] == depth and list(f[i, 5:9]) == list(cnt), expr
        rare = sorted(np.log10(max(d, 1e-9)) for d in terms)[:4]
        assert np.allclose(f[i, 9:9 + len(rare)], rare, atol=1e-5)
        assert np.isclose(f[i, 13], np.linalg.norm(qs[i].astype(np.float64)), rtol=1e-5)


# ------------------------------------------------------------------------------ kernels (GPU)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hybridsearchsystem_amd import learned
    return learned


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


@pytest.mark.gpu
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("epi", [0, 1, 2])
def test_gemm_matches_bf16_reference(ta, tb, epi):
    L = _gpu()
    rng = np.random.default_rng(10 * ta + 3 * tb + epi)
    M, N, K = 70, 97, 45
    a = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
    b = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    z = rng.standard_normal((M, N)).astype(np.float32)
    print("gemm ta=%d tb=%d epi=%d" % (ta, tb, epi), flush=True)
    da, db, dbias, dz = _dev(a), _dev(b), _dev(bias), _dev(z)
    c = torch.zeros((M, N), device="cuda")
    h = torch.zeros((M, N), device="cuda")
    L.gemm(M, N, K, da, M if ta else K, ta, db, K if tb else N, tb, c, N, epi, bias=dbias if epi == 1 else None,
           z=dz if epi == 2 else None, ldz=N, h=h if epi == 1 else None, ldh=N)
    torch.cuda.synchronize()
    A = omlp.bf16(a.T if ta else a).astype(np.float64)
    B = omlp.bf16(b.T if tb else b).astype(np.float64)
    want = A @ B
    scale = np.abs(A) @ np.abs(B)
    if epi == 1:
        want = want + bias
    if epi == 2:
        want = np.where(z > 0, want, 0.0)
    got = c.cpu().numpy()
    assert np.all(np.abs(got - want) <= 1e-6 * scale + 1e-6)
    if epi == 1:
        assert np.array_equal(h.cpu().numpy(), np.maximum(got, 0))


@pytest.mark.gpu
def test_small_kernels_match_reference():
    L = _gpu()
    rng = np.random.default_rng(3)
    B, D = 300, 257
    hmat = rng.standard_normal((B, D)).astype(np.float32)
    w = rng.standard_normal(D).astype(np.float32)
    b3 = np.array([0.25], np.float32)
    y = torch.zeros(B, device="cuda")
    L.head_forward(B, D, _dev(hmat), D, _dev(w), _dev(b3), y)
    want = omlp.bf16(hmat).astype(np.float64) @ omlp.bf16(w).astype(np.float64) + 0.25
    assert np.allclose(y.cpu().numpy(), want, rtol=0, atol=1e-5 * np.abs(hmat).sum(1).max())
    dy = rng.standard_normal(B).astype(np.float32)
    dz = torch.zeros((B, D), device="cuda")
    L.head_backward(B, D, _dev(dy), _dev(w), _dev(hmat), D, dz, D)
    assert np.allclose(dz.cpu().numpy(), np.where(hmat > 0, dy[:, None] * omlp.bf16(w)[None, :], 0), rtol=1e-6, atol=0)
    out = torch.zeros(D, device="cuda")
    L.column_sum(B, D, _dev(hmat), D, out, weights=_dev(dy), round_bf16=True)
    want = dy.astype(np.float64) @ omlp.bf16(hmat).astype(np.float64)
    assert np.allclose(out.cpu().numpy(), want, atol=1e-4)
    lab = rng.standard_normal(B).astype(np.float32)
    g = torch.zeros(B, device="cuda")
    lo = torch.zeros(B, device="cuda")
    ynp = y.cpu().numpy()
    L.pinball(B, _dev(ynp), _dev(lab), 0.1, g, lo)
    l_ref, g_ref = omlp.pinball(ynp, lab, 0.1)
    assert np.allclose(g.cpu().numpy(), g_ref, rtol=1e-6) and np.isclose(lo.cpu().numpy().mean(), l_ref, rtol=1e-5)
    # Adam against the textbook update
    p0 = rng.standard_normal(1000).astype(np.float32)
    gr = rng.standard_normal(1000).astype(np.float32)
    p, m, v = _dev(p0), torch.zeros(1000, device="cuda"), torch.zeros(1000, device="cuda")
    L.adam(p, _dev(gr), m, v, 0.01, 1)
    mh = (0.1 * gr) / 0.1
    vh = (0.001 * gr * gr) / 0.001
    assert np.allclose(p.cpu().numpy(), p0 - 0.01 * mh / (np.sqrt(vh) + 1e-8), rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_pinball_kernel_minimiser_is_the_tau_quantile():
    """frk_mlp_pinball against what fixes the pinball loss independently of the oracle: for a constant
    prediction y, the summed gradient is (#(labels < y) - tau n) / n, so it changes sign at the
    tau-quantile of the labels, and the mean loss there is below its value one order statistic away."""
    L = _gpu()
    rng = np.random.default_rng(12)
    n = 1001
    lab = rng.standard_normal(n).astype(np.float32)
    s = np.sort(lab)
    g = torch.zeros(n, device="cuda")
    lo = torch.zeros(n, device="cuda")
    for tau in (0.05, 0.7):
        j = int(np.ceil(tau * n)) - 1
        means = []
        for y in (s[j - 1], s[j], s[j + 1]):
            mid_lo = float(y) - 1e-4 * float(s[j + 1] - s[j - 1])
            L.pinball(n, _dev(np.full(n, mid_lo, np.float32)), _dev(lab), tau, g, lo)
            below = int((lab < mid_lo).sum())
            assert np.isclose(float(g.sum().item()), (below - tau * n) / n, atol=1e-5), (tau, y)
            L.pinball(n, _dev(np.full(n, y, np.float32)), _dev(lab), tau, g, lo)
            means.append(float(lo.mean().item()))
        assert means[1] < means[0] and means[1] < means[2], (tau, means)


@pytest.mark.gpu
def test_training_step_gradients_match_oracle():
    L = _gpu()
    rng = np.random.default_rng(4)
    dim = 256
    model = L.FloorModel(dim, hidden=256, seed=5, tau=0.05)
    # non-zero biases so every path of the backward is exercised
    for k in ("b1", "b2", "b3"):
        model.p[k] = _dev(rng.standard_normal(model.p[k].numel()).astype(np.float32) * 0.05)
    B = 512
    x = rng.standard_normal((B, model.in_dim)).astype(np.float32)
    lab = rng.standard_normal(B).astype(np.float32)
    loss, cache = model.loss_and_grads(_dev(x), _dev(lab))
    torch.cuda.synchronize()
    host = {k: v.cpu().numpy() for k, v in model.p.items()}
    l_ref, g_ref, c_ref = omlp.loss_and_grads(host, x, lab, 0.05, round_bf16=True)
    y = cache["y"].cpu().numpy()
    assert np.allclose(y, c_ref["y"], rtol=0, atol=2e-3 * np.abs(c_ref["y"]).max())
    # the oracle rounds at the same points, so only FP32 accumulation order separates them (a BF16
    # rounding that lands one ulp apart would show as ~4e-3; a dropped or transposed term as O(1))
    for k in model.p:
        got = model.g[k].cpu().numpy().reshape(g_ref[k].shape)
        ref = g_ref[k]
        err = np.abs(got - ref)
        scale = np.abs(ref).max() + 1e-12
        print(k, "max rel err %.2e mean %.2e" % (err.max() / scale, err.mean() / scale), flush=True)
        assert err.max() <= 1e-4 * scale, k
        assert err.mean() <= 1e-5 * scale, k


@pytest.mark.gpu
def test_training_learns_and_calibrates_a_quantile():
    L = _gpu()
    rng = np.random.default_rng(6)
    dim, n = 64, 49152
    q = rng.standard_normal((n, dim)).astype(np.float32) / 8
    feats = rng.standard_normal((n, L.N_FEAT)).astype(np.float32)
    target = 0.5 * q[:, 0] + 0.3 * np.tanh(feats[:, 1]) + 0.05 * rng.standard_normal(n).astype(np.float32)
    tr, va, te = slice(0, 40000), slice(40000, 44000), slice(44000, n)
    model = L.FloorModel(dim, hidden=256, seed=7, tau=0.1)
    hist = model.fit(q[tr], feats[tr], target[tr], epochs=6, batch=512, lr=2e-3)
    assert hist[-1] < 0.75 * hist[0]  # hist[0] is already the end of the first epoch
    # it learned the signal: held-out pinball loss well below the best constant predictor's
    pred = model.predict_raw(q[te], feats[te])
    const = np.quantile(target[tr], 0.1)
    l_model = omlp.pinball(pred, target[te], 0.1)[0]
    l_const = omlp.pinball(np.full(len(pred), const), target[te], 0.1)[0]
    # and the calibrated margin holds the floor below the label on ~90% of new rows
    model.calibrate(q[va], feats[va], target[va], fallback=0.1)
    below = np.mean(pred - model.margin <= target[te])
    print("held-out pinball %.4f vs constant %.4f; calibrated margin %.4f, below %.3f" % (
        l_model, l_const, model.margin, below), flush=True)
    assert l_model < 0.7 * l_const
    assert 0.86 <= below <= 0.94


@pytest.mark.gpu
def test_learned_floor_is_exact_and_prunes():
    L = _gpu()
    from hybridsearchsystem_amd import search, synth
    n, dim, n_tags = 300_000, 64, 64
    docs = synth.mixture_vectors(41, n, dim, 64)
    rng = np.random.default_rng(42)
    dens = synth.loguniform_densities(rng, n_tags, lo=0.002, hi=0.6)
    off, ids = synth.gen_postings(rng, n, dens)
    t = search.Table.build(docs, off, ids, docs_per_segment=65536)
    n_train, n_val, n_test = 20000, 2048, 512
    nq = n_train + n_val + n_test
    qs = synth.mixture_queries(43, nq, dim, 64, corpus_seed=41)
    filters, _ = synth.random_filters(44, dens, nq, depth=3, sel=(0.001, 0.6))
    ks = synth.draw_ks(45, nq, [10, 100]).astype(np.uint32)
    labels, passing = L.kth_labels(t, qs, filters, ks)
    feats = L.features(t, qs, filters, ks)
    model = L.FloorModel(dim, hidden=256, seed=8, tau=0.05)
    va = slice(n_train, n_train + n_val)
    te = slice(n_train + n_val, nq)
    model.fit(qs[:n_train], feats[:n_train], labels[:n_train], epochs=10, batch=512, lr=2e-3)
    model.calibrate(qs[va], feats[va], labels[va], fallback=0.02)
    fl = model.floors(qs[te], feats[te], ks[te])
    ft = filters[te]
    base = t.search(qs[te], ft, ks[te], probe=False)
    host = t.search(qs[te], ft, ks[te], probe=False, floors=fl)
    # the same model run by the searcher itself on the device, batch by batch
    dev = t.search(qs[te], ft, ks[te], probe=False, floor_model=model)
    for name, got in (("host floors", host), ("device floors", dev)):
        for i, what in ((0, "ids"), (1, "scores"), (2, "counts")):
            bad = np.nonzero(np.any(np.atleast_2d(base[i] != got[i]).reshape(len(base[2]), -1), axis=1))[0]
            assert len(bad) == 0, "%s: %s differ for queries %s" % (name, what, bad[:10])
    # the learned floor inside a captured graph: replays equal the unpruned results too
    replayed = 0
    for rep in range(3):
        g = t.search(qs[te], ft, ks[te], probe=False, floor_model=model, graph=True)
        for i, what in ((0, "ids"), (1, "scores"), (2, "counts")):
            assert np.array_equal(base[i], g[i]), "graph run %d: %s differ" % (rep, what)
        replayed += g[4]["graph_replayed"]
    assert replayed >= 1
    gain = base[4]["candidates"] / max(dev[4]["candidates"], 1)
    print("candidates without floor %d, learned floor %d on the host / %d on the device (x%.1f), reruns %d / %d "
          "of %d, margin %.4f" % (base[4]["candidates"], host[4]["candidates"], dev[4]["candidates"], gain,
                                  host[4]["reruns"], dev[4]["reruns"], n_test, model.margin), flush=True)
    assert gain >= 3.0
    assert dev[4]["reruns"] <= 0.1 * n_test
    # the device-side prediction equals the host-side one up to summation order
    assert abs(int(host[4]["candidates"]) - int(dev[4]["candidates"])) <= 0.01 * host[4]["candidates"]
