// capi_engine.cpp -- the handle layer of the C ABI (include/frk_capi.h, layer 2): DataTable loading
// and building, and batch search with filter strings, for callers that cannot link C++.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/frk_capi.h"
#include "../frk_common.h"
#include "data_table.h"
#include "full_recall_searcher.h"
#include "index_writer.h"
#include "search_server.h"
#include "log.h"
#include "query_node.h"

using namespace NpuRetrieval;

struct frk_table {
    std::shared_ptr<DataTable> table;
};

struct frk_builder {
    std::unique_ptr<TableBuilder> b;
};

struct frk_server {
    std::unique_ptr<SearchServer> s;
};

struct frk_filter {
    std::unique_ptr<QueryNode> tree;  // built once; adjusted on first use, then shared read-only
};

namespace {
int Fail(int code, const char* what) {
    frk::SetError("%s", what);
    return code;
}
int Fail(int code, const std::string& what) { return Fail(code, what.c_str()); }

void FilterToPairs(const FilterAst* n, std::vector<int64_t>& out) {
    switch (n->op) {
        case FilterAst::Term:
            out.push_back(0);
            out.push_back(static_cast<int64_t>(n->tag));
            return;
        case FilterAst::And:
        case FilterAst::Or:
            out.push_back(n->op == FilterAst::And ? 1 : 2);
            out.push_back(static_cast<int64_t>(n->children.size()));
            break;
        case FilterAst::Not:
            out.push_back(3);
            out.push_back(1);
            break;
    }
    for (const auto& c : n->children) FilterToPairs(c.get(), out);
}
}  // namespace

extern "C" {

int frk_table_load(int32_t device, const char* dir, uint32_t dtype_mask, frk_table** out) {
    if (!dir || !out) return Fail(FRK_ERR_INVALID, "frk_table_load: null argument");
    auto t = std::make_shared<DataTable>();
    if (!t->LoadData(device, dir, dtype_mask)) return FRK_ERR_IO;
    *out = new frk_table{t};
    return FRK_OK;
}

void frk_table_free(frk_table* t) { delete t; }

int frk_table_info(const frk_table* t, uint64_t* n_docs, uint32_t* n_segments, uint32_t* docs_per_segment) {
    if (!t) return Fail(FRK_ERR_INVALID, "frk_table_info: null table");
    if (n_docs) *n_docs = t->table->GetDocNum();
    if (n_segments) *n_segments = t->table->GetSegmentNum();
    if (docs_per_segment) *docs_per_segment = t->table->GetDocNumPerSegment();
    return FRK_OK;
}

int frk_builder_create(int32_t device, uint32_t n_docs, uint32_t docs_per_segment, frk_builder** out) {
    if (!out || n_docs == 0) return Fail(FRK_ERR_INVALID, "frk_builder_create: bad arguments");
    auto* b = new frk_builder{std::make_unique<TableBuilder>(device, n_docs, docs_per_segment)};
    if (!b->b->Error().empty()) {
        std::string e = b->b->Error();
        delete b;
        return Fail(FRK_ERR_INVALID, e.c_str());
    }
    *out = b;
    return FRK_OK;
}

void frk_builder_free(frk_builder* b) { delete b; }

int frk_builder_add_vector_field(frk_builder* b, const char* name, uint32_t dim, uint32_t dtype_mask) {
    if (!b || !name || dim == 0) return Fail(FRK_ERR_INVALID, "frk_builder_add_vector_field: bad arguments");
    return b->b->AddVectorField(name, dim, dtype_mask) ? FRK_OK : Fail(FRK_ERR_NOMEM, b->b->Error().c_str());
}

int frk_builder_set_vectors(frk_builder* b, const char* name, const float* rows, uint64_t first, uint64_t n,
                            int rows_on_device) {
    if (!b || !name || !rows) return Fail(FRK_ERR_INVALID, "frk_builder_set_vectors: bad arguments");
    return b->b->SetVectors(name, rows, first, n, rows_on_device != 0) ? FRK_OK
                                                                       : Fail(FRK_ERR_INVALID, b->b->Error().c_str());
}

int frk_builder_add_posting_field(frk_builder* b, const char* name, int is_matrix, float density_threshold,
                                  int packed) {
    if (!b || !name) return Fail(FRK_ERR_INVALID, "frk_builder_add_posting_field: bad arguments");
    return b->b->AddPostingField(name, is_matrix != 0, density_threshold, packed != 0) ? FRK_OK : FRK_ERR_INVALID;
}

int frk_builder_add_tag_postings(frk_builder* b, const char* field, uint64_t n_tags, const uint64_t* tags,
                                 const uint64_t* offsets, const uint32_t* doc_ids) {
    if (!b || !field || !tags || !offsets || (!doc_ids && offsets[n_tags] != 0))
        return Fail(FRK_ERR_INVALID, "frk_builder_add_tag_postings: bad arguments");
    for (uint64_t t = 0; t < n_tags; t++) {
        const uint64_t lo = offsets[t], hi = offsets[t + 1];
        if (hi < lo) return Fail(FRK_ERR_INVALID, "frk_builder_add_tag_postings: offsets must not decrease");
        for (uint64_t i = lo + 1; i < hi; i++)
            if (doc_ids[i] <= doc_ids[i - 1]) return Fail(FRK_ERR_INVALID, "posting doc ids must be strictly ascending");
        if (hi == lo) continue;
        if (!b->b->AddTokenPostings(field, TagToken(tags[t]), doc_ids + lo, hi - lo))
            return Fail(FRK_ERR_INVALID, b->b->Error().c_str());
    }
    return FRK_OK;
}

int frk_builder_set_gdocids(frk_builder* b, const uint64_t* ids, uint64_t n) {
    if (!b || !ids) return Fail(FRK_ERR_INVALID, "frk_builder_set_gdocids: bad arguments");
    b->b->SetGDocIds(ids, n);
    return FRK_OK;
}

int frk_builder_finish(frk_builder* b, frk_table** out) {
    if (!b || !out) return Fail(FRK_ERR_INVALID, "frk_builder_finish: null argument");
    auto t = b->b->Finish();
    if (!t) return Fail(FRK_ERR_HIP, b->b->Error().c_str());
    *out = new frk_table{t};
    return FRK_OK;
}

int frk_write_index(const char* dir, uint32_t version, uint32_t n_docs, uint32_t docs_per_segment,
                    uint32_t split_doc_num_zn, uint32_t n_vector_fields, const frk_index_vector_field* vector_fields,
                    uint32_t n_posting_fields, const frk_index_posting_field* posting_fields, const uint64_t* gdocids) {
    if (!dir || (n_vector_fields && !vector_fields) || (n_posting_fields && !posting_fields))
        return Fail(FRK_ERR_INVALID, "frk_write_index: null argument");
    IndexWriter w(n_docs, docs_per_segment, version, split_doc_num_zn);
    for (uint32_t i = 0; i < n_vector_fields; i++) {
        const frk_index_vector_field& f = vector_fields[i];
        if (!f.name || f.file_type > FRK_VECTOR_FILE_FP8_ROWS)
            return Fail(FRK_ERR_INVALID, "frk_write_index: bad vector field");
        w.AddVectorField(f.name, f.dim, f.rows, static_cast<VectorFileType>(f.file_type));
    }
    for (uint32_t i = 0; i < n_posting_fields; i++) {
        const frk_index_posting_field& f = posting_fields[i];
        if (!f.name) return Fail(FRK_ERR_INVALID, "frk_write_index: bad posting field");
        w.AddPostingField(f.name, f.is_matrix != 0, f.density_threshold, f.packed != 0, f.n_tokens, f.tags, f.terms,
                          f.offsets, f.doc_ids);
    }
    if (gdocids) w.SetGDocIds(gdocids, n_docs);
    if (!w.Write(dir)) return Fail(FRK_ERR_INVALID, "frk_write_index: " + w.Error());
    return FRK_OK;
}

uint16_t frk_f32_to_f16(float f) { return Float32ToFloat16(f); }

uint64_t frk_hash64(const char* data, uint32_t len) { return Hash64(data, len); }

int64_t frk_serialize_posting(const uint32_t* ids, uint32_t n_ids, uint32_t doc_num, int is_matrix, float threshold,
                              int packed, uint8_t* out, uint64_t cap) {
    std::vector<uint32_t> v(ids, ids + n_ids);
    std::string buf;
    if (!SerializePosting(v, doc_num, is_matrix != 0, threshold, packed != 0, buf))
        return Fail(FRK_ERR_INVALID, "posting does not serialize");
    if (buf.size() > cap) return Fail(FRK_ERR_CAPACITY, "frk_serialize_posting: output buffer too small");
    std::memcpy(out, buf.data(), buf.size());
    return static_cast<int64_t>(buf.size());
}

int frk_parse_filter(const char* expr, int64_t* pairs, int32_t cap_pairs) {
    if (!expr) return Fail(FRK_ERR_INVALID, "frk_parse_filter: null expression");
    bool ok = true;
    auto ast = ParseFilter(expr, ok);
    if (!ok) return Fail(FRK_ERR_INVALID, "filter does not parse");
    if (!ast) return 0;
    std::vector<int64_t> p;
    FilterToPairs(ast.get(), p);
    const int32_t n = static_cast<int32_t>(p.size() / 2);
    if (n > cap_pairs) return Fail(FRK_ERR_CAPACITY, "frk_parse_filter: pair buffer too small");
    std::memcpy(pairs, p.data(), p.size() * sizeof(int64_t));
    return n;
}

int frk_compile_filter(const frk_table* t, const char* field, const char* expr, uint32_t* out_expr, int32_t cap,
                       uint32_t* n_postings, uint32_t* op_num, int32_t* match_all) {
    if (!t || !field || !expr || !out_expr || !n_postings || !op_num || !match_all)
        return Fail(FRK_ERR_INVALID, "frk_compile_filter: null argument");
    bool ok = true;
    auto ast = ParseFilter(expr, ok);
    if (!ok) return Fail(FRK_ERR_INVALID, "filter does not parse");
    auto tree = BuildQueryNode(ast.get(), field);
    FilterExpr fe;
    if (!PrepareFilterExpr(*t->table, tree.get(), fe)) return FRK_ERR_INVALID;
    if (static_cast<int32_t>(fe.postExpr.size()) > cap) return Fail(FRK_ERR_CAPACITY, "program buffer too small");
    std::memcpy(out_expr, fe.postExpr.data(), fe.postExpr.size() * 4);
    *n_postings = static_cast<uint32_t>(fe.postingDeviceAddrs.size());
    *op_num = fe.opNum;
    *match_all = fe.emptyTree ? 1 : 0;
    return static_cast<int>(fe.postExpr.size());
}

int frk_floor_features(const frk_table* t, const char* posting_field, uint32_t n, const char* const* filters,
                       const uint32_t* ks, const float* queries, uint32_t dim, float* out) {
    if (!t || !posting_field || !ks || !out || (n && !queries))
        return Fail(FRK_ERR_INVALID, "frk_floor_features: null argument");
    for (uint32_t i = 0; i < n; i++) {
        bool ok = true;
        std::unique_ptr<FilterAst> ast;
        if (filters && filters[i]) {
            ast = ParseFilter(filters[i], ok);
            if (!ok) {
                frk::SetError("filter %u does not parse: %s", i, filters[i]);
                return FRK_ERR_INVALID;
            }
        }
        auto tree = BuildQueryNode(ast.get(), posting_field);
        FilterExpr fe;  // adjusts the tree exactly as a search would
        if (!PrepareFilterExpr(*t->table, tree.get(), fe)) return FRK_ERR_INVALID;
        FloorFeatures(*t->table, fe.emptyTree ? nullptr : tree.get(), ks[i], queries + static_cast<size_t>(i) * dim,
                      dim, out + static_cast<size_t>(i) * kFloorFeatures);
    }
    return FRK_OK;
}

namespace {
// AddQuery x n + BatchSearch over prebuilt trees; the results are FullRecallSearcher's own
// (BatchSearchExtract: scores from the keys, global ids through the id mapping).
int SearchTrees(const frk_table* t, const char* vector_field, uint32_t n_queries, const float* queries, uint32_t dim,
                QueryNode* const* trees, const uint32_t* ks, const frk_search_options* opts, uint32_t k_stride,
                int64_t* out_ids, float* out_scores, uint32_t* out_counts, uint32_t* out_passing,
                frk_search_stats* stats, std::chrono::steady_clock::time_point tCall) {
    FullRecallSearcher searcher(t->table->GetDeviceId(), t->table);
    for (uint32_t q = 0; q < n_queries; q++) {
        if (ks[q] > k_stride) return Fail(FRK_ERR_CAPACITY, "frk_search: k larger than k_stride");
        searcher.AddQuery(trees[q], queries + static_cast<size_t>(q) * dim, dim, ks[q]);
    }
    if (opts) {
        SearchOptions o;
        o.dtype = opts->dtype;
        o.rescoreFactor = opts->rescore_factor;
        o.probe = opts->probe != 0;
        o.probeStep = opts->probe_step;
        o.floors = opts->floors;
        o.candCap = opts->cand_cap;
        o.floorModel = opts->floor_model;
        o.graph = opts->graph != 0;
        searcher.SetOptions(o);
    }
    // BatchSearchDevice + BatchSearchExtract, the two halves of BatchSearch, so the passing counts the
    // device half reports are at hand for the caller as well
    std::vector<AggrDeviceResult> dev;
    if (!searcher.BatchSearchDevice(vector_field, false, dev)) return FRK_ERR_INVALID;
    std::vector<std::shared_ptr<FullRecallResult>> results;
    if (!searcher.BatchSearchExtract(dev, results, false) || results.size() != n_queries)
        return Fail(FRK_ERR_INVALID, "frk_search: result extraction failed");
    for (uint32_t q = 0; q < n_queries; q++) {
        const FullRecallResult& r = *results[q];
        const size_t n = r.docIds.size();
        std::memcpy(out_ids + static_cast<size_t>(q) * k_stride, r.docIds.data(), n * sizeof(int64_t));
        if (out_scores) std::memcpy(out_scores + static_cast<size_t>(q) * k_stride, r.scores.data(), n * sizeof(float));
        out_counts[q] = static_cast<uint32_t>(n);
        if (out_passing) out_passing[q] = dev[q].passing;
    }
    if (stats) {
        const BatchStats& s = searcher.LastStats();
        stats->ms_filter = s.ms_filter;
        stats->ms_probe = s.ms_probe;
        stats->ms_score = s.ms_score;
        stats->ms_select = s.ms_select;
        stats->ms_rescore = s.ms_rescore;
        stats->ms_total = s.ms_total;
        stats->candidates = s.candidates;
        stats->reruns = s.reruns;
        stats->overflowed = s.overflowed;
        stats->probe_step = s.probeStep;
        stats->graph_replayed = s.graphReplayed ? 1u : 0u;
        stats->ms_host_prepare = s.ms_host_prepare;
        stats->ms_host_finish = s.ms_host_finish;
        stats->cand_cap = s.cap;
        stats->ms_call =
            std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - tCall).count();
    }
    return FRK_OK;
}
}  // namespace

int frk_search(const frk_table* t, const char* vector_field, const char* posting_field, uint32_t n_queries,
               const float* queries, uint32_t dim, const char* const* filters, const uint32_t* ks,
               const frk_search_options* opts, uint32_t k_stride, int64_t* out_ids, float* out_scores,
               uint32_t* out_counts, uint32_t* out_passing, frk_search_stats* stats) {
    const auto tCall = std::chrono::steady_clock::now();
    if (!t || !vector_field || !posting_field || (!queries && n_queries) || !ks || !out_ids || !out_counts)
        return Fail(FRK_ERR_INVALID, "frk_search: null argument");
    std::vector<std::unique_ptr<QueryNode>> trees(n_queries);
    std::vector<QueryNode*> raw(n_queries);
    for (uint32_t q = 0; q < n_queries; q++) {
        bool ok = true;
        std::unique_ptr<FilterAst> ast;
        if (filters && filters[q]) {
            ast = ParseFilter(filters[q], ok);
            if (!ok) {
                frk::SetError("filter %u does not parse: %s", q, filters[q]);
                return FRK_ERR_INVALID;
            }
        }
        trees[q] = BuildQueryNode(ast.get(), posting_field);
        raw[q] = trees[q].get();
    }
    return SearchTrees(t, vector_field, n_queries, queries, dim, raw.data(), ks, opts, k_stride, out_ids, out_scores,
                       out_counts, out_passing, stats, tCall);
}

int frk_filter_prepare(const char* posting_field, const char* expr, frk_filter** out) {
    if (!posting_field || !out) return Fail(FRK_ERR_INVALID, "frk_filter_prepare: null argument");
    *out = nullptr;
    bool ok = true;
    std::unique_ptr<FilterAst> ast;
    if (expr) {
        ast = ParseFilter(expr, ok);
        if (!ok) {
            frk::SetError("filter does not parse: %s", expr);
            return FRK_ERR_INVALID;
        }
    }
    auto f = std::make_unique<frk_filter>();
    f->tree = BuildQueryNode(ast.get(), posting_field);
    *out = f.release();
    return FRK_OK;
}

void frk_filter_free(frk_filter* f) { delete f; }

int frk_search_prepared(const frk_table* t, const char* vector_field, uint32_t n_queries, const float* queries,
                        uint32_t dim, const frk_filter* const* filters, const uint32_t* ks,
                        const frk_search_options* opts, uint32_t k_stride, int64_t* out_ids, float* out_scores,
                        uint32_t* out_counts, uint32_t* out_passing, frk_search_stats* stats) {
    const auto tCall = std::chrono::steady_clock::now();
    if (!t || !vector_field || (!queries && n_queries) || !filters || !ks || !out_ids || !out_counts)
        return Fail(FRK_ERR_INVALID, "frk_search_prepared: null argument");
    std::vector<QueryNode*> raw(n_queries);
    for (uint32_t q = 0; q < n_queries; q++) {
        if (!filters[q]) return Fail(FRK_ERR_INVALID, "frk_search_prepared: null filter handle");
        raw[q] = filters[q]->tree.get();
    }
    return SearchTrees(t, vector_field, n_queries, queries, dim, raw.data(), ks, opts, k_stride, out_ids, out_scores,
                       out_counts, out_passing, stats, tCall);
}

int frk_server_create(const frk_table* t, const char* vector_field, const char* posting_field,
                      const frk_search_options* opts, const frk_server_config* cfg, frk_server** out) {
    if (!t || !vector_field || !posting_field || !out) return Fail(FRK_ERR_INVALID, "frk_server_create: null argument");
    ServerConfig c;
    c.vectorField = vector_field;
    c.postingField = posting_field;
    if (opts) {
        c.options.dtype = opts->dtype;
        c.options.rescoreFactor = opts->rescore_factor;
        c.options.probe = opts->probe != 0;
        c.options.probeStep = opts->probe_step;
        c.options.floors = nullptr;  // per-request floors have no meaning across batches
        c.options.candCap = opts->cand_cap;
        c.options.floorModel = opts->floor_model;
        c.options.graph = opts->graph != 0;
    }
    if (cfg) {
        if (cfg->max_batch_queries) c.maxBatchQueries = cfg->max_batch_queries;
        if (cfg->workers) c.workers = cfg->workers;
        c.hbmBudgetBytes = cfg->hbm_budget_bytes;
        if (cfg->pad_queries) c.padQueries = cfg->pad_queries;
        c.splitRequests = cfg->split_requests != 0;
    }
    VectorFieldData* vf = nullptr;
    if (!t->table->GetVectorFieldData(c.vectorField, vf) || !vf)
        return Fail(FRK_ERR_INVALID, "frk_server_create: no such vector field");
    auto* srv = new frk_server;
    srv->s = std::make_unique<SearchServer>(t->table, t->table->GetDeviceId(), c);
    *out = srv;
    return FRK_OK;
}

void frk_server_free(frk_server* s) { delete s; }

int frk_server_submit(frk_server* s, uint32_t n_queries, const float* queries, uint32_t dim,
                      const char* const* filters, const uint32_t* ks, uint64_t* ticket) {
    if (!s || !queries || !ks || !ticket) return Fail(FRK_ERR_INVALID, "frk_server_submit: null argument");
    std::vector<std::string> f;
    if (filters)
        for (uint32_t i = 0; i < n_queries; i++) f.emplace_back(filters[i] ? filters[i] : "");
    std::vector<uint32_t> k(ks, ks + n_queries);
    *ticket = s->s->Submit(n_queries, queries, dim, f, k);
    return *ticket ? FRK_OK : Fail(FRK_ERR_INVALID, "frk_server_submit: request rejected");
}

int frk_server_wait(frk_server* s, uint64_t ticket, uint32_t k_stride, int64_t* out_ids, float* out_scores,
                    uint32_t* out_counts, uint32_t* out_passing, double* times, uint32_t* batch_queries) {
    if (!s || !out_ids || !out_counts) return Fail(FRK_ERR_INVALID, "frk_server_wait: null argument");
    RequestResult r;
    const bool ok = s->s->Wait(ticket, r);
    if (times) {
        times[0] = r.submitted;
        times[1] = r.started;
        times[2] = r.finished;
    }
    if (batch_queries) *batch_queries = r.batchQueries;
    if (!ok) return Fail(FRK_ERR_INVALID, "frk_server_wait: request failed");
    for (size_t q = 0; q < r.counts.size(); q++) {
        const uint32_t n = std::min<uint32_t>(r.counts[q], k_stride);
        for (uint32_t i = 0; i < k_stride; i++) {
            out_ids[q * k_stride + i] = i < n ? r.ids[q][i] : -1;
            if (out_scores) out_scores[q * k_stride + i] = i < n ? r.scores[q][i] : 0.0f;
        }
        out_counts[q] = n;
        if (out_passing) out_passing[q] = r.passing[q];
    }
    return FRK_OK;
}

int frk_server_stats_get(const frk_server* s, frk_server_stats* out) {
    if (!s || !out) return Fail(FRK_ERR_INVALID, "frk_server_stats_get: null argument");
    const ServerStats st = s->s->Stats();
    out->requests = st.requests;
    out->queries = st.queries;
    out->batches = st.batches;
    out->graph_replays = st.graphReplays;
    out->reruns = st.reruns;
    out->overflows = st.overflows;
    out->device_ms = st.deviceMs;
    out->doc_bytes_streamed = st.docBytesStreamed;
    return FRK_OK;
}

double frk_server_now(const frk_server* s) { return s ? s->s->Now() : 0.0; }

uint64_t frk_server_batch_bytes(const frk_server* s, uint32_t n_queries) { return s ? s->s->BatchBytes(n_queries) : 0; }

int frk_workspace_pool_stats(frk_pool_stats* out) {
    if (!out) return Fail(FRK_ERR_INVALID, "frk_workspace_pool_stats: null argument");
    const WorkspacePoolStats st = GetWorkspacePoolStats();
    out->live_bytes = st.liveBytes;
    out->peak_bytes = st.peakBytes;
    out->mallocs = st.mallocs;
    out->frees = st.frees;
    out->retain_limit = st.retainLimit;
    out->workspaces = st.workspaces;
    out->pooled = st.pooled;
    return FRK_OK;
}

}  // extern "C"
