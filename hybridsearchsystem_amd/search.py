"""Index tables and batch search through the C ABI handle layer (include/frk_capi.h, layer 2).

``Table.build`` stages an index in device memory from arrays (vectors host- or device-resident,
per-tag posting lists as CSR); ``write_index`` writes the same arrays as an index directory on disk
and ``Table.load`` reads one. ``Table.search`` runs one batch through
``NpuRetrieval::FullRecallSearcher``.
"""
import ctypes

import numpy as np

from . import check, lib
from ._capi_sigs import IndexPostingField, IndexVectorField, SearchOptions, SearchStats

FRK_DTYPE_F32, FRK_DTYPE_BF16, FRK_DTYPE_FP8E4M3, FRK_DTYPE_F16X2 = 0, 1, 2, 3
VECTOR_FILE_FP16_ZN, VECTOR_FILE_FP32_ROWS, VECTOR_FILE_BF16_ROWS, VECTOR_FILE_FP8_ROWS = 0, 1, 2, 3


def _mask(dtypes):
    m = 0
    for d in dtypes:
        m |= 1 << int(d)
    return m


class PreparedFilter:
    """A filter string parsed and built into its query tree once (frk_filter_prepare)."""

    def __init__(self, expr, posting_field="tag"):
        self.expr = expr
        h = ctypes.c_void_p()
        check(lib().frk_filter_prepare(posting_field.encode(), expr.encode() if expr is not None else None,
                                       ctypes.byref(h)), "frk_filter_prepare")
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().frk_filter_free(self.h)
            except Exception:  # interpreter shutdown: the library may already be gone
                pass
            self.h = None


class Table:
    def __init__(self, handle, vector_field="vec", posting_field="tag", dim=None):
        self.h = handle
        self.vector_field = vector_field
        self.posting_field = posting_field
        self.dim = dim
        n = ctypes.c_uint64()
        s = ctypes.c_uint32()
        d = ctypes.c_uint32()
        check(lib().frk_table_info(self.h, ctypes.byref(n), ctypes.byref(s), ctypes.byref(d)), "frk_table_info")
        self.n_docs, self.n_segments, self.docs_per_segment = int(n.value), int(s.value), int(d.value)

    def __del__(self):
        try:
            self.free()
        except Exception:  # interpreter shutdown: the library may already be gone
            pass

    def free(self):
        if getattr(self, "h", None):
            lib().frk_table_free(self.h)
            self.h = None

    @staticmethod
    def load(path, device=0, dtypes=(), vector_field="vec", posting_field="tag", dim=None):
        h = ctypes.c_void_p()
        check(lib().frk_table_load(device, path.encode(), _mask(dtypes), ctypes.byref(h)), "frk_table_load")
        return Table(h, vector_field, posting_field, dim)

    @staticmethod
    def build(vectors, tag_offsets, tag_doc_ids, n_docs=None, tags=None, docs_per_segment=131072, dtypes=(),
              vector_field="vec", posting_field="tag", is_matrix=False, threshold=0.05, packed=True, gdocids=None,
              device=0, chunk_rows=1 << 20):
        """vectors: numpy [n, dim] float32 or a torch CUDA tensor; tag postings as CSR over tags
        0..T-1 (or the ids in `tags`): docs of tag t are tag_doc_ids[tag_offsets[t]:tag_offsets[t+1]]."""
        L = lib()
        n = int(n_docs if n_docs is not None else vectors.shape[0])
        dim = int(vectors.shape[1])
        b = ctypes.c_void_p()
        check(L.frk_builder_create(device, n, docs_per_segment, ctypes.byref(b)), "frk_builder_create")
        try:
            check(L.frk_builder_add_vector_field(b, vector_field.encode(), dim, _mask(dtypes)), "add_vector_field")
            on_dev = hasattr(vectors, "is_cuda") and vectors.is_cuda
            if on_dev:
                vt = vectors.contiguous()
                for at in range(0, n, chunk_rows):
                    m = min(chunk_rows, n - at)
                    check(L.frk_builder_set_vectors(b, vector_field.encode(), vt[at:at + m].data_ptr(), at, m, 1),
                          "set_vectors")
            else:
                v = np.ascontiguousarray(vectors, dtype=np.float32)
                check(L.frk_builder_set_vectors(b, vector_field.encode(), v.ctypes.data, 0, n, 0), "set_vectors")
            check(L.frk_builder_add_posting_field(b, posting_field.encode(), int(is_matrix), float(threshold),
                                                  int(packed)), "add_posting_field")
            off = np.ascontiguousarray(tag_offsets, dtype=np.uint64)
            ids = np.ascontiguousarray(tag_doc_ids, dtype=np.uint32)
            t = np.ascontiguousarray(tags if tags is not None else np.arange(len(off) - 1), dtype=np.uint64)
            check(L.frk_builder_add_tag_postings(b, posting_field.encode(), len(t), t.ctypes.data, off.ctypes.data,
                                                 ids.ctypes.data if len(ids) else None), "add_tag_postings")
            if gdocids is not None:
                g = np.ascontiguousarray(gdocids, dtype=np.uint64)
                check(L.frk_builder_set_gdocids(b, g.ctypes.data, len(g)), "set_gdocids")
            h = ctypes.c_void_p()
            check(L.frk_builder_finish(b, ctypes.byref(h)), "frk_builder_finish")
        finally:
            L.frk_builder_free(b)
        return Table(h, vector_field, posting_field, dim)

    def prepare_filters(self, filters):
        """PreparedFilter handles for a list of filter strings (None = match all)."""
        return [PreparedFilter(f, self.posting_field) for f in filters]

    def floor_features(self, queries, filters, ks):
        """The floor model's 32 features per (query, filter, k), as the searcher computes them."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        n, dim = q.shape
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(ks, dtype=np.uint32), (n,)))
        enc = [f.encode() if f is not None else None for f in filters]
        fl = (ctypes.c_char_p * n)(*enc)
        out = np.zeros((n, 32), dtype=np.float32)
        check(lib().frk_floor_features(self.h, self.posting_field.encode(), n, fl, ks.ctypes.data, q.ctypes.data, dim,
                                       out.ctypes.data), "frk_floor_features")
        return out

    def compile_filter(self, expr):
        cap = 4 * len(expr) + 16
        buf = np.zeros(cap, dtype=np.uint32)
        npost, ops, ma = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_int32()
        n = lib().frk_compile_filter(self.h, self.posting_field.encode(), expr.encode(), buf.ctypes.data, cap,
                                     ctypes.byref(npost), ctypes.byref(ops), ctypes.byref(ma))
        check(0 if n >= 0 else n, "frk_compile_filter")
        return buf[:n].copy(), int(npost.value), int(ops.value), bool(ma.value)

    def search(self, queries, filters, ks, dtype=FRK_DTYPE_F32, rescore_factor=4, probe=True, probe_step=0,
               floors=None, cand_cap=0, k_stride=None, floor_model=None, graph=False):
        """Returns (ids int64 [Q, k_stride], scores float32, counts uint32, passing uint32, stats dict).
        filters: strings (None = match all), or PreparedFilter handles (parsed once, see
        prepare_filters), or None for no filter at all.
        floor_model: a learned.FloorModel whose floors the searcher predicts on the device.
        graph: replay the device sequence from a hipGraph captured per batch shape."""
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq, dim = q.shape
        ks = np.ascontiguousarray(np.broadcast_to(np.asarray(ks, dtype=np.uint32), (nq,)))
        k_stride = int(k_stride or max(1, int(ks.max()) if nq else 1))
        fl = None
        prepared = filters is not None and len(filters) > 0 and all(isinstance(f, PreparedFilter) for f in filters)
        if prepared:
            fl = (ctypes.c_void_p * nq)(*[f.h.value for f in filters])
        elif filters is not None:
            enc = [f.encode() if f is not None else None for f in filters]
            fl = (ctypes.c_char_p * nq)(*enc)
        opts = SearchOptions(dtype, rescore_factor, int(bool(probe)), probe_step, None, cand_cap, None, int(bool(graph)))
        fm = None
        if floor_model is not None:
            fm = floor_model.c_model()  # keeps the host arrays it points at alive for the call
            opts.floor_model = ctypes.cast(ctypes.byref(fm), ctypes.c_void_p)
        if floors is not None:
            fa = np.ascontiguousarray(floors, dtype=np.float32)
            opts.floors = fa.ctypes.data
        ids = np.full((nq, k_stride), -1, dtype=np.int64)
        sc = np.zeros((nq, k_stride), dtype=np.float32)
        cnt = np.zeros(nq, dtype=np.uint32)
        passing = np.zeros(nq, dtype=np.uint32)
        st = SearchStats()
        if prepared:
            check(lib().frk_search_prepared(self.h, self.vector_field.encode(), nq, q.ctypes.data, dim,
                                            ctypes.cast(fl, ctypes.c_void_p), ks.ctypes.data, ctypes.byref(opts),
                                            k_stride, ids.ctypes.data, sc.ctypes.data, cnt.ctypes.data,
                                            passing.ctypes.data, ctypes.byref(st)), "frk_search_prepared")
        else:
            check(lib().frk_search(self.h, self.vector_field.encode(), self.posting_field.encode(), nq, q.ctypes.data,
                                   dim, fl, ks.ctypes.data, ctypes.byref(opts), k_stride, ids.ctypes.data,
                                   sc.ctypes.data, cnt.ctypes.data, passing.ctypes.data, ctypes.byref(st)), "frk_search")
        stats = {f: getattr(st, f) for f, _ in SearchStats._fields_}
        return ids, sc, cnt, passing, stats


def write_index(path, vectors, tag_offsets, tag_doc_ids, tags=None, terms=None, docs_per_segment=131072,
                version=0, split_doc_num_zn=32, vector_file=VECTOR_FILE_FP16_ZN, vector_field="vec",
                posting_field="tag", is_matrix=False, threshold=0.05, packed=True, gdocids=None):
    """Writes an index directory (.meta, id/*.idm, <vector_field>/*.vector, <posting_field>/*.posting)
    from host arrays: vectors [n, dim] float32; postings as CSR over tokens (tags[t] as decimal text,
    or the strings in `terms`)."""
    v = np.ascontiguousarray(vectors, dtype=np.float32)
    n, dim = v.shape
    off = np.ascontiguousarray(tag_offsets, dtype=np.uint64)
    ids = np.ascontiguousarray(tag_doc_ids, dtype=np.uint32)
    nt = len(off) - 1
    t = np.ascontiguousarray(tags if tags is not None else np.arange(nt), dtype=np.uint64)
    term_arr = None
    if terms is not None:
        term_arr = (ctypes.c_char_p * nt)(*[x.encode() for x in terms])
    vf = IndexVectorField(vector_field.encode(), dim, vector_file, v.ctypes.data)
    pf = IndexPostingField(posting_field.encode(), int(is_matrix), float(threshold), int(packed), nt, t.ctypes.data,
                           ctypes.cast(term_arr, ctypes.c_void_p) if term_arr is not None else None, off.ctypes.data,
                           ids.ctypes.data if len(ids) else None)
    g = None
    if gdocids is not None:
        g = np.ascontiguousarray(gdocids, dtype=np.uint64)
        if len(g) != n:
            raise ValueError("gdocids must hold one id per doc")
    check(lib().frk_write_index(str(path).encode(), version, n, docs_per_segment, split_doc_num_zn, 1,
                                ctypes.byref(vf), 1, ctypes.byref(pf), g.ctypes.data if g is not None else None),
          "frk_write_index")


def f32_to_f16(x):
    """The index writer's FP32 -> FP16 rounding, elementwise (uint16 bit patterns)."""
    a = np.ascontiguousarray(x, dtype=np.float32).ravel()
    f = lib().frk_f32_to_f16
    return np.array([f(float(e)) for e in a], dtype=np.uint16).reshape(np.shape(x))


def parse_filter(expr):
    """The product's filter grammar as (n, 2) int64 pre-order pairs; None if it does not parse."""
    cap = 4 * len(expr) + 8
    buf = np.zeros(2 * cap, dtype=np.int64)
    n = lib().frk_parse_filter(expr.encode(), buf.ctypes.data, cap)
    if n < 0:
        return None
    return buf[:2 * n].reshape(n, 2)
