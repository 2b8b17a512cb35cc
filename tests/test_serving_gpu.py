"""Graph replay and the continuous-batching server: a replayed pipeline graph returns what the eager
sequence returns (every replay, with eager work and new data in between), and each request's result is
the same alone and mixed into batches with others (bit-identical ids and scores: every query's result
depends only on that query), under HBM-budgeted admission too."""
import threading

import numpy as np
import pytest

from hybridsearchsystem_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def corpus():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from hybridsearchsystem_amd import search
    n, dim, n_tags = 300_000, 128, 48
    docs = synth.mixture_vectors(61, n, dim, 64)
    rng = np.random.default_rng(62)
    dens = synth.loguniform_densities(rng, n_tags, lo=0.001, hi=0.6)
    off, ids = synth.gen_postings(rng, n, dens)
    t = search.Table.build(docs, off, ids, docs_per_segment=65536,
                           dtypes=(search.FRK_DTYPE_BF16, search.FRK_DTYPE_FP8E4M3))
    qs = synth.mixture_queries(63, 2048, dim, 64, corpus_seed=61)
    fs, _ = synth.random_filters(64, dens, 2048, depth=3, sel=(0.0005, 0.6))
    yield t, qs, fs
    t.free()


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.parametrize("dtype", [0, 1, 2])
def test_graph_replay_matches_eager(corpus, dtype):
    t, qs, fs = corpus
    ks = np.random.default_rng(65).choice([10, 100], size=96).astype(np.uint32)
    fa = fs[:96]
    qa, qb = qs[:96], qs[96:192]
    print("graph replay dtype %d" % dtype, flush=True)
    eager_a = t.search(qa, fa, ks, dtype=dtype)
    eager_b = t.search(qb, fa, ks, dtype=dtype)
    replays = 0
    for i in range(4):
        g = t.search(qa, fa, ks, dtype=dtype, graph=True)
        assert _same(g, eager_a), "replay %d of batch A" % i
        replays += g[4]["graph_replayed"]
        t.search(qs[192:256], fs[192:256], 10, dtype=dtype)  # eager work of another shape in between
        gb = t.search(qb, fa, ks, dtype=dtype, graph=True)  # same shape, new query data
        assert _same(gb, eager_b), "replay %d of batch B" % i
        replays += gb[4]["graph_replayed"]
    assert replays >= 6  # the first batch of the shape runs eagerly, later ones replay


def _requests(rng, qs, fs, n_req, max_size):
    out, at = [], 0
    for _ in range(n_req):
        s = int(rng.integers(1, max_size + 1))
        idx = np.arange(at, at + s) % len(qs)
        at += s
        out.append((idx, rng.choice([10, 100, 1000], size=s).astype(np.uint32)))
    return out


@pytest.mark.parametrize("budget_queries", [0, 48])
def test_server_result_same_alone_and_in_mix(corpus, budget_queries):
    t, qs, fs = corpus
    from hybridsearchsystem_amd import serve
    srv = serve.Server(t, dtype=1, rescore_factor=4, probe=True, graph=True, workers=2, max_batch_queries=256)
    if budget_queries:
        srv.close()
        probe_srv = serve.Server(t, dtype=1, graph=False)
        budget = probe_srv.batch_bytes(budget_queries)
        probe_srv.close()
        srv = serve.Server(t, dtype=1, rescore_factor=4, probe=True, graph=True, workers=2, max_batch_queries=256,
                           hbm_budget_bytes=budget)
    reqs = _requests(np.random.default_rng(66 + budget_queries), qs, fs, 40, 24 if budget_queries else 64)
    print("server mix: %d requests, budget %d queries" % (len(reqs), budget_queries), flush=True)
    # in the mix: every request submitted at once from another thread, collected as they finish
    tickets = [None] * len(reqs)

    def producer():
        for i, (idx, k) in enumerate(reqs):
            tickets[i] = srv.submit(qs[idx], [fs[j] for j in idx], k)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    mixed = [srv.wait(tk) for tk in tickets]
    st = srv.stats()
    assert st["batches"] < len(reqs)  # requests shared batches
    assert any(m[5] > len(r[0]) for m, r in zip(mixed, reqs))
    if budget_queries:
        assert all(m[5] <= max(budget_queries, len(r[0])) for m, r in zip(mixed, reqs))
    # alone: one request at a time on an idle server, so each batch holds only that request
    for (idx, k), m in zip(reqs, mixed):
        a = srv.wait(srv.submit(qs[idx], [fs[j] for j in idx], k))
        assert a[5] == len(idx)
        assert np.array_equal(a[2], m[2]) and np.array_equal(a[0], m[0]) and np.array_equal(a[1], m[1])
        assert np.array_equal(a[3], m[3])
    if budget_queries:
        # idle workspaces are not kept beyond the budget: after the run (every batch done, every
        # workspace released) the pool holds at most the budget
        pool = serve.pool_stats()
        print("pool after the budgeted run:", pool, "budget", budget, flush=True)
        assert pool["retain_limit"] == budget
        assert pool["live_bytes"] <= budget
    srv.close()
    if budget_queries:
        assert serve.pool_stats()["retain_limit"] == 0


def test_server_splits_requests_over_batches(corpus):
    """With split_requests the scheduler cuts requests at batch boundaries (here 64-query batches and
    requests of up to 150 queries, so most span two or three batches): every query's ids, scores,
    counts and passing counts equal those of the same request served whole and alone, no batch
    exceeds its 64 queries, and every query is served once."""
    t, qs, fs = corpus
    from hybridsearchsystem_amd import serve
    reqs = _requests(np.random.default_rng(71), qs, fs, 20, 150)
    total = sum(len(r[0]) for r in reqs)
    print("server split: %d requests, %d queries" % (len(reqs), total), flush=True)
    srv = serve.Server(t, dtype=1, rescore_factor=4, probe=True, graph=True, workers=2, max_batch_queries=64,
                       split_requests=True)
    tickets = [None] * len(reqs)

    def producer():
        for i, (idx, k) in enumerate(reqs):
            tickets[i] = srv.submit(qs[idx], [fs[j] for j in idx], k)

    th = threading.Thread(target=producer)
    th.start()
    th.join()
    split = [srv.wait(tk) for tk in tickets]
    st = srv.stats()
    srv.close()
    assert st["queries"] == total and st["requests"] == len(reqs)
    assert st["batches"] >= (total + 63) // 64
    assert all(m[5] <= 64 for m in split)
    assert any(len(r[0]) > 64 for r in reqs)
    whole = serve.Server(t, dtype=1, rescore_factor=4, probe=True, graph=True, workers=2, max_batch_queries=256)
    for (idx, k), m in zip(reqs, split):
        a = whole.wait(whole.submit(qs[idx], [fs[j] for j in idx], k))
        assert a[5] == len(idx)
        assert np.array_equal(a[2], m[2]) and np.array_equal(a[0], m[0]) and np.array_equal(a[1], m[1])
        assert np.array_equal(a[3], m[3])
        assert m[4][0] <= m[4][1] <= m[4][2]  # submitted <= started <= finished
    whole.close()
