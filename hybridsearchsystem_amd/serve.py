"""Serving: the continuous-batching request server of the C ABI (frk_server_*) and a Poisson load
generator that measures it.

``Server`` wraps one server over a resident ``search.Table``. ``poisson_load`` submits requests of
U{1..max_request} queries with k drawn from {10, 100, 1000} at exponential inter-arrival times (a
target rate in queries/s) from one thread while another collects them, and reports per-request
latency (finish - submit on the server's clock) percentiles, achieved queries/s, batches formed,
graph replays and the doc-stream rate against the HBM peak.

``python -m hybridsearchsystem_amd.serve`` builds the synthetic serving corpus (20M x 768 with BF16
and FP8 stores beside the FP32 originals, 4,096 Zipf tags), trains the learned floor, and sweeps
arrival rates for each low-precision pipeline (store + learned floor + FP32 rescoring, graph replay)
and the plain FP32 path, printing one JSON line per run.
"""
import argparse
import ctypes
import json
import sys
import threading
import time

import numpy as np

from . import check, lib
from ._capi_sigs import SearchOptions

HBM_PEAK = 8.0e12
# published dense MFMA peaks of an MI355X per store dtype (FP32, BF16, FP8; F16X2 runs FP16 MFMAs)
MFMA_PEAK = {0: 157.3e12, 1: 2.5e15, 2: 5.0e15, 3: 2.5e15}


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_batch_queries", ctypes.c_uint32), ("workers", ctypes.c_uint32),
                ("hbm_budget_bytes", ctypes.c_uint64), ("pad_queries", ctypes.c_uint32),
                ("split_requests", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("queries", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("graph_replays", ctypes.c_uint64), ("reruns", ctypes.c_uint64), ("overflows", ctypes.c_uint64),
                ("device_ms", ctypes.c_double),
                ("doc_bytes_streamed", ctypes.c_uint64)]


def _bind(L):
    if getattr(L, "_frk_server_bound", False):
        return L
    vp, u32, u64, cp = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_char_p
    L.frk_server_create.restype = ctypes.c_int
    L.frk_server_create.argtypes = [vp, cp, cp, ctypes.POINTER(SearchOptions), ctypes.POINTER(ServerConfig),
                                    ctypes.POINTER(vp)]
    L.frk_server_free.restype = None
    L.frk_server_free.argtypes = [vp]
    L.frk_server_submit.restype = ctypes.c_int
    L.frk_server_submit.argtypes = [vp, u32, vp, u32, vp, vp, ctypes.POINTER(u64)]
    L.frk_server_wait.restype = ctypes.c_int
    L.frk_server_wait.argtypes = [vp, u64, u32, vp, vp, vp, vp, vp, ctypes.POINTER(u32)]
    L.frk_server_stats_get.restype = ctypes.c_int
    L.frk_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStats)]
    L.frk_server_batch_bytes.restype = u64
    L.frk_server_batch_bytes.argtypes = [vp, u32]
    L.frk_server_now.restype = ctypes.c_double
    L.frk_server_now.argtypes = [vp]
    L.frk_workspace_pool_stats.restype = ctypes.c_int
    L.frk_workspace_pool_stats.argtypes = [vp]
    L._frk_server_bound = True
    return L


class Server:
    def __init__(self, table, dtype=0, rescore_factor=4, probe=True, floor_model=None, graph=True,
                 max_batch_queries=256, workers=2, hbm_budget_bytes=0, pad_queries=64, split_requests=False):
        L = _bind(lib())
        self.table = table
        self._model = floor_model.c_model() if floor_model is not None else None  # must outlive the server
        opts = SearchOptions(dtype, rescore_factor, int(bool(probe)), 0, None, 0, None, int(bool(graph)))
        if self._model is not None:
            opts.floor_model = ctypes.cast(ctypes.byref(self._model), ctypes.c_void_p)
        cfg = ServerConfig(max_batch_queries, workers, hbm_budget_bytes, pad_queries, int(bool(split_requests)))
        h = ctypes.c_void_p()
        check(L.frk_server_create(table.h, table.vector_field.encode(), table.posting_field.encode(),
                                  ctypes.byref(opts), ctypes.byref(cfg), ctypes.byref(h)), "frk_server_create")
        self.h = h
        self._k = {}

    def submit(self, queries, filters, ks):
        q = np.ascontiguousarray(queries, dtype=np.float32)
        n, dim = q.shape
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(ks, dtype=np.uint32), (n,)))
        enc = (ctypes.c_char_p * n)(*[(f or "").encode() for f in filters])
        t = ctypes.c_uint64()
        check(lib().frk_server_submit(self.h, n, q.ctypes.data, dim, enc, ks.ctypes.data, ctypes.byref(t)),
              "frk_server_submit")
        self._k[t.value] = (n, int(ks.max()))
        return t.value

    def wait(self, ticket):
        """(ids [n, k], scores, counts, passing, (submitted, started, finished), batch queries)."""
        n, k = self._k.pop(ticket)
        ids = np.full((n, k), -1, np.int64)
        sc = np.zeros((n, k), np.float32)
        cnt = np.zeros(n, np.uint32)
        pas = np.zeros(n, np.uint32)
        times = np.zeros(3, np.float64)
        bq = ctypes.c_uint32()
        check(lib().frk_server_wait(self.h, ticket, k, ids.ctypes.data, sc.ctypes.data, cnt.ctypes.data,
                                    pas.ctypes.data, times.ctypes.data, ctypes.byref(bq)), "frk_server_wait")
        return ids, sc, cnt, pas, tuple(times), int(bq.value)

    def stats(self):
        st = ServerStats()
        check(lib().frk_server_stats_get(self.h, ctypes.byref(st)), "frk_server_stats_get")
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def batch_bytes(self, n_queries):
        return int(lib().frk_server_batch_bytes(self.h, n_queries))

    def now(self):
        return lib().frk_server_now(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().frk_server_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


class PoolStats(ctypes.Structure):
    """frk_pool_stats of include/frk_capi.h."""
    _fields_ = [("live_bytes", ctypes.c_uint64), ("peak_bytes", ctypes.c_uint64), ("mallocs", ctypes.c_uint64),
                ("frees", ctypes.c_uint64), ("retain_limit", ctypes.c_uint64), ("workspaces", ctypes.c_uint32),
                ("pooled", ctypes.c_uint32)]


def pool_stats():
    """The searchers' workspace pool: device bytes held now and at the peak, workspaces alive/idle."""
    st = PoolStats()
    check(_bind(lib()).frk_workspace_pool_stats(ctypes.byref(st)), "frk_workspace_pool_stats")
    return {f: getattr(st, f) for f, _ in PoolStats._fields_}


def poisson_load(server, queries, filters, rate_qps, n_requests, max_request=256, ks=(10, 100, 1000), seed=0):
    """Submits n_requests requests (sizes U{1..max_request}, per-query k from ks, queries and filters
    taken round-robin from the pools) at Poisson arrivals with mean rate_qps queries/s; returns the
    latency/throughput summary and the per-request results."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(1, max_request + 1, n_requests)
    mean = float(sizes.mean())
    gaps = rng.exponential(mean / rate_qps, n_requests)
    pool = len(queries)
    reqs = []
    at = 0
    for s in sizes:
        idx = (np.arange(at, at + s) % pool)
        at += s
        reqs.append((idx, rng.choice(np.asarray(ks, np.uint32), size=s)))
    tickets = [None] * n_requests
    done = [None] * n_requests
    ready = threading.Semaphore(0)

    def producer():
        t0 = time.perf_counter()
        due = 0.0
        for i, (idx, k) in enumerate(reqs):
            due += gaps[i]
            delay = t0 + due - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = server.submit(queries[idx], [filters[j] for j in idx], k)
            ready.release()

    def consumer():
        for i in range(n_requests):
            ready.acquire()
            done[i] = server.wait(tickets[i])

    s0 = server.stats()
    w0 = time.perf_counter()
    th = [threading.Thread(target=producer), threading.Thread(target=consumer)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - w0
    s1 = server.stats()
    lat = np.array([(d[4][2] - d[4][0]) * 1e3 for d in done])
    q = int(sizes.sum())
    doc_bytes = s1["doc_bytes_streamed"] - s0["doc_bytes_streamed"]
    out = {"offered_qps": rate_qps, "requests": n_requests, "queries": q, "mean_request": round(mean, 1),
           "achieved_qps": round(q / wall, 1), "wall_s": round(wall, 3),
           "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
           "batches": s1["batches"] - s0["batches"], "graph_replays": s1["graph_replays"] - s0["graph_replays"],
           "mean_batch_queries": round(q / max(1, s1["batches"] - s0["batches"]), 1),
           "reruns": s1["reruns"] - s0["reruns"], "overflows": s1["overflows"] - s0["overflows"],
           "doc_stream_gbs": round(doc_bytes / wall / 1e9, 1),
           "hbm_roofline_frac": round(doc_bytes / wall / HBM_PEAK, 4)}
    return out, reqs, done


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--docs", type=int, default=20_000_000)
    ap.add_argument("--dim", type=int, default=768)
    ap.add_argument("--tags", type=int, default=4096)
    ap.add_argument("--rates", default="500,2000,5000,10000,20000,30000,40000", help="offered queries/s")
    ap.add_argument("--plain-rates", default="500,2000,5000", help="offered queries/s for the plain FP32 path")
    ap.add_argument("--seconds", type=float, default=8.0, help="offered load per rate")
    ap.add_argument("--train", type=int, default=150_000, help="learned-floor training queries (measured at "
                    "saturation, BF16: 50k -> margin 0.164, 24.7k q/s; 150k -> 0.144, 27.2k q/s; 300k -> 0.126, "
                    "no further gain)")
    ap.add_argument("--workers", type=int, default=4, help="batches in flight (measured at saturation, 20M x 768: "
                    "BF16 18.0k / 19.5k / 20.1k q/s and FP8 27.8k / 30.5k / 33.0k with 2 / 3 / 4)")
    ap.add_argument("--whole-requests", action="store_true", help="never split a request over batches (by "
                    "default the scheduler fills every batch to 256 queries, cutting the last request)")
    ap.add_argument("--dtypes", default="bf16,fp8", help="low-precision stores to serve from, each with the "
                    "learned floor and FP32 rescoring")
    ap.add_argument("--plain", action="store_true", help="also run the plain FP32 path at each rate")
    ap.add_argument("--diag", action="store_true", help="compare host- and device-side learned floors first")
    args = ap.parse_args()
    import torch
    from . import learned, search, synth
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    seed, cents = synth.SEEDS["corpus"], 1024
    t0 = time.time()
    dens = synth.zipf_densities(args.tags)
    off, ids = synth.gen_postings(np.random.default_rng(synth.SEEDS["tags"]), args.docs, dens)
    vec = torch.empty((args.docs, args.dim), dtype=torch.float32, device="cuda")
    for at in range(0, args.docs, 1 << 20):
        m = min(1 << 20, args.docs - at)
        vec[at:at + m] = synth.mixture_vectors_torch(seed, m, args.dim, cents, first=at)
    names = {"bf16": search.FRK_DTYPE_BF16, "fp8": search.FRK_DTYPE_FP8E4M3}
    dtypes = [d.strip() for d in args.dtypes.split(",") if d.strip()]
    table = search.Table.build(vec, off, ids, n_docs=args.docs, dtypes=tuple(names[d] for d in dtypes),
                               docs_per_segment=131072)
    del vec
    torch.cuda.empty_cache()
    log("serving corpus %d x %d (%s) built in %.1fs" % (args.docs, args.dim, "+".join(dtypes), time.time() - t0))
    S = synth.SEEDS
    tq = synth.mixture_queries(S["train_queries"], args.train, args.dim, cents, corpus_seed=seed)
    tf, _ = synth.random_filters(S["train_filters"], dens, args.train, depth=4, sel=(1e-5, 0.6))
    tk = synth.draw_ks(S["train_kdraw"], args.train, [10, 100, 1000])
    # a rerun costs a whole extra pass over the corpus, so the serving floor targets few of them
    model, info = learned.train_for_table(table, args.dim, tq, tf, tk, val_fraction=0.2, fallback=0.001, log=log)
    log("learned floor trained: %s" % info)
    pool = 8192
    qs = synth.mixture_queries(S["eval_queries"], pool, args.dim, cents, corpus_seed=seed)
    fs, _ = synth.random_filters(S["eval_filters"], dens, pool, depth=4, sel=(1e-5, 0.6))
    # how often the floor would sit above the true k'-th score on the serving pool, per k' (k * 4)
    diag = {}
    for k in (10, 100, 1000):
        kp = np.full(pool, 4 * k, np.uint32)
        lab, pas = learned.kth_labels(table, qs, fs, kp)
        fl = model.floors(qs, learned.features(table, qs, fs, kp), kp)
        live = np.isfinite(fl)
        over = live & (fl > lab)
        diag[4 * k] = {"floored": int(live.sum()), "above_kth": int(over.sum()),
                       "above_kth_with_fewer_passing": int((over & (pas < kp)).sum())}
    log("floor check on the serving pool: %s" % diag)
    if args.diag:
        ks = synth.draw_ks(5, 1024, [10, 100, 1000]).astype(np.uint32)
        fl = model.floors(qs[:1024], learned.features(table, qs[:1024], fs[:1024], ks * 4), ks * 4)
        for name, kw in (("host floors", dict(floors=None)), ("device floors", dict(floor_model=model)),
                         ("device floors + graph", dict(floor_model=model, graph=True))):
            rr = ov = 0
            for lo in range(0, 1024, 128):
                sl = slice(lo, lo + 128)
                extra = dict(kw)
                if name == "host floors":
                    extra["floors"] = fl[sl]
                r = table.search(qs[sl], fs[sl], ks[sl], dtype=search.FRK_DTYPE_BF16, rescore_factor=4, probe=False,
                                 **extra)
                rr += r[4]["reruns"]
                ov += r[4]["overflowed"]
            log("diag %s: reruns %d overflowed %d of 1024" % (name, rr, ov))
    runs = [("%s + learned floor + fp32 rescore, graph replay" % d, args.rates,
             dict(dtype=names[d], rescore_factor=4, probe=False, floor_model=model, graph=True)) for d in dtypes]
    if args.plain:
        runs.append(("plain fp32, graph replay", args.plain_rates,
                     dict(dtype=search.FRK_DTYPE_F32, probe=True, graph=True)))
    for name, rates, kw in runs:
        srv = Server(table, workers=args.workers, split_requests=not args.whole_requests, **kw)
        poisson_load(srv, qs, fs, 2000.0, 20, seed=99)  # warm: workspaces grown, graphs captured
        for rate in [float(r) for r in rates.split(",")]:
            n_req = int(min(2000, max(40, rate * args.seconds / 128.5)))
            log("%s: offered %.0f queries/s, %d requests" % (name, rate, n_req))
            res, _, _ = poisson_load(srv, qs, fs, rate, n_req, seed=int(rate))
            # the main pass's multiply-adds against the store dtype's MFMA peak (the same queries'
            # doc bytes against HBM are hbm_roofline_frac)
            res["mfma_frac"] = round(2.0 * res["queries"] * args.docs * args.dim / res["wall_s"] / MFMA_PEAK[kw["dtype"]], 4)
            res.update(pipeline=name, config={"workload": "serving %d x %d, %d tags" % (args.docs, args.dim, args.tags),
                                              "workers": args.workers, "max_batch_queries": 256,
                                              "split_requests": not args.whole_requests})
            print(json.dumps(res), flush=True)
        srv.close()


if __name__ == "__main__":
    main()
