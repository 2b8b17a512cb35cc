"""ctypes signatures of the C ABI beyond the core kernel entry points (include/frk_capi.h)."""
import ctypes

vp, u32, u64, i32, f32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float
cp = ctypes.c_char_p


class SearchOptions(ctypes.Structure):
    _fields_ = [("dtype", u32), ("rescore_factor", u32), ("probe", ctypes.c_int32), ("probe_step", u32),
                ("floors", vp), ("cand_cap", u32), ("floor_model", vp), ("graph", u32)]


class SearchStats(ctypes.Structure):
    _fields_ = [("ms_filter", f32), ("ms_probe", f32), ("ms_score", f32), ("ms_select", f32), ("ms_rescore", f32),
                ("ms_total", f32), ("candidates", u64), ("reruns", u32), ("overflowed", u32), ("probe_step", u32),
                ("cand_cap", u32), ("graph_replayed", u32), ("ms_host_prepare", f32), ("ms_host_finish", f32),
                ("ms_call", f32)]


class IndexVectorField(ctypes.Structure):
    _fields_ = [("name", cp), ("dim", u32), ("file_type", u32), ("rows", vp)]


class IndexPostingField(ctypes.Structure):
    _fields_ = [("name", cp), ("is_matrix", ctypes.c_int32), ("density_threshold", f32), ("packed", ctypes.c_int32),
                ("n_tokens", u64), ("tags", vp), ("terms", vp), ("offsets", vp), ("doc_ids", vp)]


def bind(L):
    L.frk_read_probe.restype = i32
    L.frk_read_probe.argtypes = [vp, u64, vp, u32, vp]
    L.frk_score_matrix.restype = i32
    L.frk_score_matrix.argtypes = [vp, vp, u32, vp, u64, vp]
    L.frk_aggregate.restype = i32
    L.frk_aggregate.argtypes = [vp, vp, u64, vp, u32, vp, vp]
    L.frk_check_program.restype = i32
    L.frk_mlp_forward_noalloc.restype = i32
    L.frk_mlp_forward_noalloc.argtypes = [vp, i32, vp, i32, vp, vp]
    L.frk_check_program.argtypes = [vp, u32, u32, vp]
    L.frk_tiles_bytes.restype = u64
    L.frk_tiles_bytes.argtypes = [u64, u32, u32]
    L.frk_tile_docs.restype = i32
    L.frk_tile_docs.argtypes = [vp, u64, u32, u32, u32, vp, vp, u32]
    L.frk_tile_rows_device.restype = i32
    L.frk_tile_rows_device.argtypes = [vp, u64, u32, u64, u32, u32, vp, vp, vp]
    L.frk_score_blockmax.restype = i32
    L.frk_score_blockmax.argtypes = [vp, vp, u32, vp, u32, u64, u64, u64, vp, u32, vp]
    L.frk_masks_to_blocks.restype = i32
    L.frk_masks_to_blocks.argtypes = [vp, u64, u32, u64, vp, u32, vp]
    L.frk_kth_largest.restype = i32
    L.frk_kth_largest.argtypes = [vp, u32, u32, u32, vp, u32, vp, vp]
    L.frk_rescore.restype = i32
    L.frk_rescore.argtypes = [vp, vp, u32, vp, u32, vp, vp, vp]
    L.frk_rescore_rows.restype = i32
    L.frk_rescore_rows.argtypes = [vp, u32, u64, vp, u32, vp, u32, vp, vp, vp]
    L.frk_table_load.restype = i32
    L.frk_table_load.argtypes = [i32, cp, u32, ctypes.POINTER(vp)]
    L.frk_table_free.restype = None
    L.frk_table_free.argtypes = [vp]
    L.frk_table_info.restype = i32
    L.frk_table_info.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(u32), ctypes.POINTER(u32)]
    L.frk_builder_create.restype = i32
    L.frk_builder_create.argtypes = [i32, u32, u32, ctypes.POINTER(vp)]
    L.frk_builder_free.restype = None
    L.frk_builder_free.argtypes = [vp]
    L.frk_builder_add_vector_field.restype = i32
    L.frk_builder_add_vector_field.argtypes = [vp, cp, u32, u32]
    L.frk_builder_set_vectors.restype = i32
    L.frk_builder_set_vectors.argtypes = [vp, cp, vp, u64, u64, i32]
    L.frk_builder_add_posting_field.restype = i32
    L.frk_builder_add_posting_field.argtypes = [vp, cp, i32, f32, i32]
    L.frk_builder_add_tag_postings.restype = i32
    L.frk_builder_add_tag_postings.argtypes = [vp, cp, u64, vp, vp, vp]
    L.frk_builder_set_gdocids.restype = i32
    L.frk_builder_set_gdocids.argtypes = [vp, vp, u64]
    L.frk_builder_finish.restype = i32
    L.frk_builder_finish.argtypes = [vp, ctypes.POINTER(vp)]
    L.frk_write_index.restype = i32
    L.frk_write_index.argtypes = [cp, u32, u32, u32, u32, u32, ctypes.POINTER(IndexVectorField), u32,
                                  ctypes.POINTER(IndexPostingField), vp]
    L.frk_f32_to_f16.restype = ctypes.c_uint16
    L.frk_f32_to_f16.argtypes = [f32]
    L.frk_hash64.restype = u64
    L.frk_hash64.argtypes = [cp, u32]
    L.frk_serialize_posting.restype = ctypes.c_int64
    L.frk_serialize_posting.argtypes = [vp, u32, u32, i32, f32, i32, vp, u64]
    L.frk_parse_filter.restype = i32
    L.frk_parse_filter.argtypes = [cp, vp, ctypes.c_int32]
    L.frk_compile_filter.restype = i32
    L.frk_compile_filter.argtypes = [vp, cp, cp, vp, ctypes.c_int32, ctypes.POINTER(u32), ctypes.POINTER(u32),
                                     ctypes.POINTER(ctypes.c_int32)]
    L.frk_fill_u32.restype = i32
    L.frk_fill_u32.argtypes = [vp, u64, u32, vp]
    L.frk_floor_features.restype = i32
    L.frk_floor_features.argtypes = [vp, cp, u32, vp, vp, vp, u32, vp]
    L.frk_filter_prepare.restype = i32
    L.frk_filter_prepare.argtypes = [cp, cp, ctypes.POINTER(vp)]
    L.frk_filter_free.restype = None
    L.frk_filter_free.argtypes = [vp]
    L.frk_search_prepared.restype = i32
    L.frk_search_prepared.argtypes = [vp, cp, u32, vp, u32, vp, vp, ctypes.POINTER(SearchOptions), u32, vp, vp, vp,
                                      vp, ctypes.POINTER(SearchStats)]
    L.frk_search.restype = i32
    L.frk_search.argtypes = [vp, cp, cp, u32, vp, u32, vp, vp, ctypes.POINTER(SearchOptions), u32, vp, vp, vp, vp,
                             ctypes.POINTER(SearchStats)]
