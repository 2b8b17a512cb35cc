"""Learned score floor: a small MLP that predicts, per query, a lower bound on the k-th best filtered
score, used by FullRecallSearcher as a per-query floor so the fused scorer drops docs below it before
compaction.

Input: the query vector and 32 filter features (estimated log selectivity under tag independence,
expected passing docs, k relative to them, AST depth and operator counts, the rarest tags' densities,
query norm), computed by the library (frk_floor_features) exactly as the searcher computes them when
it runs the model on the device for a batch (``Table.search(floor_model=...)``). Network:
(dim + 32) -> 256 -> 256 -> 1 with ReLU, trained with a pinball (quantile) loss on
k-th scores the plain FP32 method produces, then shifted down by a margin calibrated on held-out
queries. Every forward, backward and optimizer step runs in the library's HIP kernels
(``frk_mlp_*`` in include/frk_capi.h: BF16 MFMA GEMMs with FP32 accumulation); torch only provides
device memory.

Exactness does not depend on the model: the searcher reruns unpruned any query whose floor left
fewer than min(k, passing) survivors, so results always equal the plain method's.
"""
import ctypes

import numpy as np

from . import check, lib

N_FEAT = 32
EPI_NONE, EPI_BIAS_RELU, EPI_RELU_GRAD = 0, 1, 2


class MlpParams(ctypes.Structure):
    _fields_ = [("w1", ctypes.c_void_p), ("b1", ctypes.c_void_p), ("w2", ctypes.c_void_p), ("b2", ctypes.c_void_p),
                ("w3", ctypes.c_void_p), ("b3", ctypes.c_void_p), ("in_dim", ctypes.c_int32),
                ("hidden", ctypes.c_int32)]


def _bind(L):
    if getattr(L, "_frk_mlp_bound", False):
        return L
    vp, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    L.frk_mlp_gemm.restype = i32
    L.frk_mlp_gemm.argtypes = [i32, i32, i32, vp, i32, i32, vp, i32, i32, vp, i32, i32, vp, vp, i32, vp, i32, vp]
    L.frk_mlp_head_forward.restype = i32
    L.frk_mlp_head_forward.argtypes = [i32, i32, vp, i32, vp, vp, vp, vp]
    L.frk_mlp_head_backward.restype = i32
    L.frk_mlp_head_backward.argtypes = [i32, i32, vp, vp, vp, i32, vp, i32, vp]
    L.frk_mlp_column_sum.restype = i32
    L.frk_mlp_column_sum.argtypes = [i32, i32, vp, i32, vp, i32, vp, vp]
    L.frk_mlp_pinball.restype = i32
    L.frk_mlp_pinball.argtypes = [i32, vp, vp, f32, vp, vp, vp]
    L.frk_mlp_adam.restype = i32
    L.frk_mlp_adam.argtypes = [i64, vp, vp, vp, vp, f32, f32, f32, f32, i32, vp]
    L.frk_mlp_forward.restype = i32
    L.frk_mlp_forward.argtypes = [ctypes.POINTER(MlpParams), i32, vp, i32, vp, vp, vp]
    L._frk_mlp_bound = True
    return L


def _L():
    return _bind(lib())


def _p(t):
    return None if t is None else t.data_ptr()


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# ------------------------------------------------------------------------------------------ kernels
def gemm(M, N, K, a, lda, trans_a, b, ldb, trans_b, c, ldc, epilogue=EPI_NONE, bias=None, z=None, ldz=0, h=None,
         ldh=0):
    check(_L().frk_mlp_gemm(M, N, K, _p(a), lda, int(trans_a), _p(b), ldb, int(trans_b), _p(c), ldc, epilogue,
                            _p(bias), _p(z), ldz, _p(h), ldh, _stream()), "frk_mlp_gemm")


def head_forward(rows, dim, h, ldh, w, bias, y):
    check(_L().frk_mlp_head_forward(rows, dim, _p(h), ldh, _p(w), _p(bias), _p(y), _stream()), "frk_mlp_head_forward")


def head_backward(rows, dim, dy, w, z, ldz, dz, lddz):
    check(_L().frk_mlp_head_backward(rows, dim, _p(dy), _p(w), _p(z), ldz, _p(dz), lddz, _stream()),
          "frk_mlp_head_backward")


def column_sum(rows, cols, x, ldx, out, weights=None, round_bf16=False):
    check(_L().frk_mlp_column_sum(rows, cols, _p(x), ldx, _p(weights), int(round_bf16), _p(out), _stream()),
          "frk_mlp_column_sum")


def pinball(rows, y, label, tau, dy, loss=None):
    check(_L().frk_mlp_pinball(rows, _p(y), _p(label), float(tau), _p(dy), _p(loss), _stream()), "frk_mlp_pinball")


def adam(param, grad, m, v, lr, step, beta1=0.9, beta2=0.999, eps=1e-8):
    check(_L().frk_mlp_adam(param.numel(), _p(param), _p(grad), _p(m), _p(v), float(lr), beta1, beta2, eps, int(step),
                            _stream()), "frk_mlp_adam")


# ----------------------------------------------------------------------------------------- features
def features(table, queries, filters, ks):
    """The 32 floor features per (query, filter, k) as the searcher computes them for a batch
    (frk_floor_features): log10 selectivity under tag independence from the table's per-token doc
    counts, log10 k, log10 expected passing docs, k relative to them, AST depth and AND/OR/NOT/term
    counts, the rarest terms' log10 densities, query norm, match-all flag."""
    return table.floor_features(queries, filters, ks)


def expected_passing(feats):
    return np.power(10.0, np.asarray(feats, np.float64)[:, 2])


class FloorModelC(ctypes.Structure):
    """frk_floor_model of include/frk_capi.h."""
    _fields_ = [("mlp", MlpParams), ("feat_mean", ctypes.c_void_p), ("feat_std", ctypes.c_void_p),
                ("margin", ctypes.c_float), ("min_passing_ratio", ctypes.c_float), ("margin_by_k", ctypes.c_void_p)]


def k_decade(feats):
    """Decade bucket of k (0: k < 10, 1: < 100, 2: < 1000, 3: >= 1000) from the log10 k feature."""
    return np.clip(np.floor(np.asarray(feats, np.float32)[:, 1] + 1e-4), 0, 3).astype(np.int64)


# ------------------------------------------------------------------------------------------- model
class FloorModel:
    """The floor MLP on the device: FP32 master weights, BF16 arithmetic in the kernels."""

    def __init__(self, dim, hidden=256, seed=0, tau=0.02):
        import torch
        self.dim, self.hidden, self.tau = dim, hidden, tau
        self.in_dim = dim + N_FEAT
        rng = np.random.default_rng(seed)

        def lin(fan_out, fan_in):
            lim = np.sqrt(6.0 / fan_in)
            return rng.uniform(-lim, lim, size=(fan_out, fan_in)).astype(np.float32)

        host = {"w1": lin(hidden, self.in_dim), "b1": np.zeros(hidden, np.float32), "w2": lin(hidden, hidden),
                "b2": np.zeros(hidden, np.float32), "w3": lin(1, hidden).reshape(hidden),
                "b3": np.zeros(1, np.float32)}
        self.p = {k: torch.from_numpy(v).cuda() for k, v in host.items()}
        self.g = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.step = 0
        self.feat_mean = np.zeros(N_FEAT, np.float32)
        self.feat_std = np.ones(N_FEAT, np.float32)
        self.margin = 0.0
        self.margin_by_k = None  # optional float32 [4]: per decade of k
        self.min_passing_ratio = 2.0
        self._ws = {}

    # -- buffers
    def _buf(self, name, shape):
        import torch
        t = self._ws.get(name)
        if t is None or t.numel() < int(np.prod(shape)):
            t = torch.empty(int(np.prod(shape)), dtype=torch.float32, device="cuda")
            self._ws[name] = t
        return t[:int(np.prod(shape))].view(*shape)

    def params_struct(self):
        p = self.p
        return MlpParams(p["w1"].data_ptr(), p["b1"].data_ptr(), p["w2"].data_ptr(), p["b2"].data_ptr(),
                         p["w3"].data_ptr(), p["b3"].data_ptr(), self.in_dim, self.hidden)

    def c_model(self):
        """The model as the searcher takes it (frk_search_options.floor_model); the returned struct
        holds references to the host arrays it points at."""
        mean = np.ascontiguousarray(self.feat_mean, np.float32)
        std = np.ascontiguousarray(self.feat_std, np.float32)
        mk = None if self.margin_by_k is None else np.ascontiguousarray(self.margin_by_k, np.float32)
        m = FloorModelC(self.params_struct(), mean.ctypes.data, std.ctypes.data, float(self.margin),
                        float(self.min_passing_ratio), None if mk is None else mk.ctypes.data)
        m._keep = (mean, std, mk)
        return m

    def inputs(self, queries, feats):
        """Host rows [query | standardized features] (float32)."""
        q = np.asarray(queries, np.float32)
        f = (np.asarray(feats, np.float32) - self.feat_mean) / self.feat_std
        return np.ascontiguousarray(np.concatenate([q, f], axis=1), dtype=np.float32)

    # -- passes
    def forward(self, x):
        """x: device [rows, in_dim] -> device y [rows] through frk_mlp_forward."""
        rows = x.shape[0]
        work = self._buf("work", (4 * rows * self.hidden,))
        y = self._buf("y_inf", (rows,))
        p = self.params_struct()
        check(_L().frk_mlp_forward(ctypes.byref(p), rows, x.data_ptr(), self.in_dim, work.data_ptr(), y.data_ptr(),
                                   _stream()), "frk_mlp_forward")
        return y

    def loss_and_grads(self, x, label):
        """Forward + pinball loss + backward on the device; gradients land in self.g. Returns the
        device per-row loss and the cached activations."""
        B, H, D = x.shape[0], self.hidden, self.in_dim
        p, g = self.p, self.g
        z1, h1 = self._buf("z1", (B, H)), self._buf("h1", (B, H))
        z2, h2 = self._buf("z2", (B, H)), self._buf("h2", (B, H))
        y, dy, loss = self._buf("y", (B,)), self._buf("dy", (B,)), self._buf("loss", (B,))
        dz2, dz1 = self._buf("dz2", (B, H)), self._buf("dz1", (B, H))
        gemm(B, H, D, x, D, 0, p["w1"], D, 1, z1, H, EPI_BIAS_RELU, bias=p["b1"], h=h1, ldh=H)
        gemm(B, H, H, h1, H, 0, p["w2"], H, 1, z2, H, EPI_BIAS_RELU, bias=p["b2"], h=h2, ldh=H)
        head_forward(B, H, h2, H, p["w3"], p["b3"], y)
        pinball(B, y, label, self.tau, dy, loss)
        column_sum(B, H, h2, H, g["w3"], weights=dy, round_bf16=True)
        column_sum(B, 1, dy, 1, g["b3"])
        head_backward(B, H, dy, p["w3"], z2, H, dz2, H)
        gemm(H, H, B, dz2, H, 1, h1, H, 0, g["w2"], H)
        column_sum(B, H, dz2, H, g["b2"])
        gemm(B, H, H, dz2, H, 0, p["w2"], H, 0, dz1, H, EPI_RELU_GRAD, z=z1, ldz=H)
        gemm(H, D, B, dz1, H, 1, x, D, 0, g["w1"], D)
        column_sum(B, H, dz1, H, g["b1"])
        return loss, {"z1": z1, "h1": h1, "z2": z2, "h2": h2, "y": y, "dy": dy}

    def apply_grads(self, lr):
        self.step += 1
        for k in self.p:
            adam(self.p[k], self.g[k], self.m[k], self.v[k], lr, self.step)

    # -- training
    def fit(self, queries, feats, labels, epochs=30, batch=1024, lr=2e-3, seed=0, log=None):
        """Trains on rows with a finite label (the k-th score exists). Feature standardization is
        fitted here."""
        import torch
        ok = np.isfinite(labels)
        feats = np.asarray(feats, np.float32)[ok]
        self.feat_mean = feats.mean(axis=0).astype(np.float32)
        self.feat_std = np.maximum(feats.std(axis=0), 1e-3).astype(np.float32)
        x = torch.from_numpy(self.inputs(np.asarray(queries)[ok], feats)).cuda()
        y = torch.from_numpy(np.asarray(labels, np.float32)[ok]).cuda()
        n = x.shape[0]
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        history = []
        for ep in range(epochs):
            lr_ep = lr * 0.5 * (1.0 + np.cos(np.pi * ep / epochs))
            perm = torch.randperm(n, device="cuda", generator=gen)
            tot = 0.0
            for lo in range(0, n - batch + 1, batch):
                idx = perm[lo:lo + batch]
                loss, _ = self.loss_and_grads(x[idx].contiguous(), y[idx].contiguous())
                self.apply_grads(lr_ep)
                if lo == 0 or lo + 2 * batch > n:
                    tot = float(loss.mean())
            history.append(tot)
            if log:
                log("epoch %d loss %.5f" % (ep, tot))
        return history

    def predict_raw(self, queries, feats):
        import torch
        x = torch.from_numpy(self.inputs(queries, feats)).cuda()
        out = np.empty(x.shape[0], np.float32)
        for lo in range(0, x.shape[0], 8192):
            out[lo:lo + 8192] = self.forward(x[lo:lo + 8192].contiguous()).cpu().numpy()
        return out

    def calibrate(self, queries, feats, labels, fallback=0.01, per_k=True):
        """Margin so that at most `fallback` of the held-out queries get a floor above their label;
        with per_k, one margin per decade of k (the score density near the k-th best, and with it
        the cost of a loose floor, differs by orders of magnitude between k = 10 and k = 4000),
        falling back to the global margin for a decade with too few held-out queries."""
        ok = np.isfinite(labels)
        feats = np.asarray(feats)[ok]
        pred = self.predict_raw(np.asarray(queries)[ok], feats)
        over = pred - np.asarray(labels, np.float32)[ok]
        self.margin = float(max(np.quantile(over, 1.0 - fallback), 0.0))
        if per_k:
            dec = k_decade(feats)
            mk = np.full(4, self.margin, np.float32)
            for b in range(4):
                sel = over[dec == b]
                if len(sel) >= max(50, int(2.0 / fallback)):
                    mk[b] = max(float(np.quantile(sel, 1.0 - fallback)), 0.0)
            self.margin_by_k = mk
        return self.margin

    def floors(self, queries, feats, ks):
        """Per-query floors on the host, as the searcher computes them on the device: prediction minus
        the margin; -inf where fewer than min_passing_ratio * k docs are expected to pass (too little
        to prune, too easy to overshoot)."""
        margin = self.margin if self.margin_by_k is None else np.asarray(self.margin_by_k)[k_decade(feats)]
        fl = self.predict_raw(queries, feats) - margin
        fl[expected_passing(feats) < self.min_passing_ratio * np.asarray(ks, np.float64)] = -np.inf
        return fl.astype(np.float32)

    # -- persistence (plain arrays, no pickling)
    def state(self):
        st = {k: v.cpu().numpy() for k, v in self.p.items()}
        st.update(feat_mean=self.feat_mean, feat_std=self.feat_std, margin=np.float32(self.margin),
                  tau=np.float32(self.tau),
                  margin_by_k=np.full(4, np.nan, np.float32) if self.margin_by_k is None else self.margin_by_k)
        return st

    def load_state(self, st):
        import torch
        for k in self.p:
            self.p[k] = torch.from_numpy(np.asarray(st[k], np.float32)).cuda()
        self.feat_mean = np.asarray(st["feat_mean"], np.float32)
        self.feat_std = np.asarray(st["feat_std"], np.float32)
        self.margin = float(st["margin"])
        mk = np.asarray(st.get("margin_by_k", np.full(4, np.nan)), np.float32)
        self.margin_by_k = None if np.isnan(mk).any() else mk


def kth_labels(table, queries, filters, ks, batch=128, **search_kw):
    """The k-th best score per query from the plain FP32 method (-inf when fewer than k pass), and
    the passing-doc counts."""
    from .search import FRK_DTYPE_F32
    n = len(queries)
    labels = np.full(n, -np.inf, np.float32)
    passing = np.zeros(n, np.int64)
    for lo in range(0, n, batch):
        sl = slice(lo, lo + batch)
        ids, sc, cnt, pas, _ = table.search(queries[sl], filters[sl], ks[sl], dtype=FRK_DTYPE_F32, **search_kw)
        for i in range(len(cnt)):
            q = lo + i
            passing[q] = pas[i]
            if cnt[i] >= ks[q] and cnt[i] > 0:
                labels[q] = sc[i, cnt[i] - 1]
    return labels, passing


def train_for_table(table, dim, queries, filters, ks, rescore_factors=(1, 4), seed=0, epochs=8, batch=1024, lr=2e-3,
                    tau=0.02, val_fraction=0.05, fallback=0.005, log=None):
    """Trains a floor model for `table` on the given seeded queries: each query gets one main-pass key
    count k' = k * m (m drawn from rescore_factors, so one model serves the plain FP32 path, m = 1,
    and the low-precision paths), labelled with the k'-th best score of the plain FP32 method. The
    margin is calibrated on the last val_fraction of the queries for the given fallback rate.
    Returns (model, info)."""
    import time
    rng = np.random.default_rng(seed)
    m = rng.choice(np.asarray(rescore_factors, np.uint32), size=len(ks))
    kp = np.minimum(np.asarray(ks, np.uint32) * m, 4096).astype(np.uint32)
    t0 = time.time()
    labels, passing = kth_labels(table, queries, filters, kp)
    t1 = time.time()
    feats = features(table, queries, filters, kp)
    n_val = max(1, int(len(kp) * val_fraction))
    tr, va = slice(0, len(kp) - n_val), slice(len(kp) - n_val, len(kp))
    model = FloorModel(dim, seed=seed, tau=tau)
    hist = model.fit(queries[tr], feats[tr], labels[tr], epochs=epochs, batch=batch, lr=lr, seed=seed, log=log)
    model.calibrate(queries[va], feats[va], labels[va], fallback=fallback)
    t2 = time.time()
    info = {"train_queries": int(np.isfinite(labels[tr]).sum()), "val_queries": int(np.isfinite(labels[va]).sum()),
            "label_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2), "loss": [round(x, 5) for x in hist],
            "margin": round(model.margin, 5), "target_fallback": fallback,
            "margin_by_k_decade": None if model.margin_by_k is None else [round(float(x), 5) for x in model.margin_by_k]}
    return model, info
