"""Seeded synthetic workloads of the benchmark configurations.

Corpus: L2-normalised draws from a Gaussian mixture (centroids N(0, I), sigma 0.3). Tags: independent
Bernoulli per (doc, tag), densities log-uniform (small config) or a Zipf(1.1) rank curve rescaled
log-linearly onto [1e-5, 0.5]. Queries: mixture draws plus N(0, 0.1^2) noise. Filters: random
AND/OR/NOT trees over distinct tags, rejection-sampled to a selectivity range computed exactly
under tag independence.
"""
import math

import numpy as np

CONFIGS = {
    # name: docs, dim, centroids, tags, tag densities, queries, batch, ks, filter depth, tags per
    # filter, selectivity range
    "c1": dict(n_docs=131072, dim=128, centroids=64, n_tags=64, densities="loguniform", n_queries=256, batch=256,
               ks=(100,), depth=3, tags_per_filter=(1, 4), sel=(0.001, 0.5), docs_per_segment=131072),
    "c2": dict(n_docs=5_000_000, dim=256, centroids=1024, n_tags=1024, densities="zipf", n_queries=4096, batch=64,
               ks=(10, 100, 1000), depth=3, tags_per_filter=(1, 4), sel=(0.0001, 0.6), docs_per_segment=131072),
    # the precision configuration takes the small configuration's tag and filter generator
    "c3": dict(n_docs=10_000_000, dim=256, centroids=1024, n_tags=64, densities="loguniform", n_queries=1024,
               batch=128, ks=(100,), depth=3, tags_per_filter=(1, 4), sel=(0.001, 0.5), docs_per_segment=131072),
}

SEEDS = dict(corpus=20261203, queries=1, filters=2, kdraw=3, tags=4, train_queries=11, train_filters=12,
             train_kdraw=13, eval_queries=21, eval_filters=22, eval_kdraw=23)


def zipf_densities(n_tags, hi=0.5, lo=1e-5):
    r = np.arange(1, n_tags + 1, dtype=np.float64)
    if n_tags == 1:
        return np.array([hi])
    alpha = math.log(hi / lo) / math.log(n_tags)
    return hi * r ** (-alpha)


def loguniform_densities(rng, n_tags, lo=0.001, hi=0.5):
    return np.exp(rng.uniform(math.log(lo), math.log(hi), n_tags))


def densities_for(cfg, seed=SEEDS["tags"]):
    if cfg["densities"] == "zipf":
        return zipf_densities(cfg["n_tags"])
    return loguniform_densities(np.random.default_rng(seed), cfg["n_tags"])


def gen_postings(rng, n_docs, densities):
    """Bernoulli(p_t) membership for every (doc, tag), as CSR: tag t's ascending doc ids are
    ids[offsets[t]:offsets[t+1]]. Sampled as geometric gaps, so the cost follows the hit count."""
    parts = []
    offsets = [0]
    for p in densities:
        if p >= 1.0:
            pos = np.arange(n_docs, dtype=np.int64)
        else:
            want = int(n_docs * p * 1.05 + 10 * math.sqrt(n_docs * p + 1) + 16)
            pos = np.empty(0, dtype=np.int64)
            last = -1
            chunks = []
            while last < n_docs:
                gaps = rng.geometric(p, size=want)
                c = last + np.cumsum(gaps, dtype=np.int64)
                chunks.append(c)
                last = int(c[-1])
            pos = np.concatenate(chunks)
            pos = pos[pos < n_docs]
        parts.append(pos.astype(np.uint32))
        offsets.append(offsets[-1] + len(pos))
    ids = np.concatenate(parts) if parts else np.zeros(0, np.uint32)
    return np.asarray(offsets, dtype=np.uint64), ids


def doc_tag_words(offsets, ids, n_docs, n_tags, lo=0, hi=None):
    """Per-doc tag bitmaps (u64 words, bit t%64 of word t/64) for docs [lo, hi)."""
    hi = n_docs if hi is None else hi
    words = (n_tags + 63) // 64
    out = np.zeros((hi - lo, words), dtype=np.uint64)
    for t in range(n_tags):
        d = ids[int(offsets[t]):int(offsets[t + 1])].astype(np.int64)
        d = d[(d >= lo) & (d < hi)] - lo
        out[d, t // 64] |= np.uint64(1) << np.uint64(t % 64)
    return out


def mixture_centroids(seed, n_centroids, dim):
    return np.random.default_rng(seed).standard_normal((n_centroids, dim)).astype(np.float32)


def mixture_vectors(seed, n, dim, n_centroids, sigma=0.3, first=0):
    """Docs [first, first + n) of the seeded corpus (each chunk seeded by its position)."""
    cent = mixture_centroids(seed, n_centroids, dim)
    rng = np.random.default_rng([seed, first])
    z = rng.integers(0, n_centroids, n)
    v = cent[z] + sigma * rng.standard_normal((n, dim), dtype=np.float32)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    return v.astype(np.float32)


def mixture_vectors_torch(seed, n, dim, n_centroids, sigma=0.3, first=0, device="cuda"):
    """The same corpus model generated on the GPU (a different stream of draws than the numpy path)."""
    import torch
    cent = torch.from_numpy(mixture_centroids(seed, n_centroids, dim)).to(device)
    g = torch.Generator(device=device)
    g.manual_seed(seed * 1000003 + first)
    z = torch.randint(0, n_centroids, (n,), generator=g, device=device)
    v = cent[z] + sigma * torch.randn((n, dim), generator=g, device=device)
    v /= v.norm(dim=1, keepdim=True)
    return v


def mixture_queries(seed, n, dim, n_centroids, corpus_seed=SEEDS["corpus"], sigma=0.3, noise=0.1):
    cent = mixture_centroids(corpus_seed, n_centroids, dim)
    rng = np.random.default_rng(seed)
    z = rng.integers(0, n_centroids, n)
    q = cent[z] + sigma * rng.standard_normal((n, dim)).astype(np.float32)
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q += noise * rng.standard_normal((n, dim)).astype(np.float32)
    return q.astype(np.float32)


def _selectivity(node, dens):
    kind = node[0]
    if kind == "t":
        return float(dens[node[1]])
    if kind == "not":
        return 1.0 - _selectivity(node[1], dens)
    vals = [_selectivity(c, dens) for c in node[1]]
    if kind == "and":
        return float(np.prod(vals))
    return 1.0 - float(np.prod([1.0 - v for v in vals]))


def _to_str(node):
    kind = node[0]
    if kind == "t":
        return str(node[1])
    if kind == "not":
        return "NOT " + _to_str(node[1])
    inner = (" AND " if kind == "and" else " OR ").join(_to_str(c) for c in node[1])
    return "(" + inner + ")"


def _tree(rng, tags, depth):
    if len(tags) == 1 or depth <= 1:
        if len(tags) == 1:
            leaf = ("t", tags[0])
            return ("not", leaf) if rng.random() < 0.2 else leaf
        kind = "and" if rng.random() < 0.5 else "or"
        return (kind, [("t", t) for t in tags])
    cut = int(rng.integers(1, len(tags)))
    groups = [tags[:cut], tags[cut:]]
    kind = "and" if rng.random() < 0.55 else "or"
    kids = [_tree(rng, g, depth - 1) for g in groups]
    node = (kind, kids)
    return ("not", node) if rng.random() < 0.1 else node


def random_filters(seed, densities, n, tags_per_filter=(1, 4), depth=3, sel=(0.001, 0.5), max_tries=None):
    """n filter strings with their exact selectivities under tag independence."""
    max_tries = max_tries or max(100000, 200 * n)
    rng = np.random.default_rng(seed)
    n_tags = len(densities)
    out, sels = [], []
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not draw filters in the selectivity range")
        k = int(rng.integers(tags_per_filter[0], tags_per_filter[1] + 1))
        # tag ranks log-uniform over 1..T: dense tags are drawn often, rare ones still appear
        ranks = np.unique(np.exp(rng.uniform(0, math.log(n_tags), 4 * k)).astype(np.int64) - 1)
        ranks = rng.permutation(ranks[ranks < n_tags])[:k]
        if len(ranks) < k:
            continue
        node = _tree(rng, [int(t) for t in ranks], depth)
        s = _selectivity(node, densities)
        if sel[0] <= s <= sel[1]:
            out.append(_to_str(node))
            sels.append(s)
    return out, np.asarray(sels)


def shift_tags(filters, offset):
    """Filter strings with every tag id moved up by offset (two tag families in one posting field)."""
    import re
    return [None if f is None else re.sub(r"\d+", lambda m: str(int(m.group(0)) + offset), f) for f in filters]


def draw_ks(seed, n, ks):
    return np.random.default_rng(seed).choice(np.asarray(ks, dtype=np.uint32), size=n)
