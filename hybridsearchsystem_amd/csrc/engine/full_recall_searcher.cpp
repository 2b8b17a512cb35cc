// full_recall_searcher.cpp -- the batch pipeline (see full_recall_searcher.h).
#include "full_recall_searcher.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "filter_stats.h"
#include "log.h"
#include "../frk_common.h"

namespace NpuRetrieval {

// ------------------------------------------------------------------------------------------ workspace
// Device and pinned buffers reused across batches: grown to the high-water mark and never freed while
// the process runs (the reference's GmMemoryManager pools, searcher/runtime/gm_memory_manager.cpp:9-75,
// without per-call malloc/free). One workspace per concurrent searcher, each with its own stream.
// Process-wide accounting of the workspace pool's device memory (the [POOLSTAT] report, and the
// retention limit a server sets from its HBM budget).
namespace {
struct PoolCounters {
    std::atomic<uint64_t> liveBytes{0}, peakBytes{0}, mallocs{0}, frees{0};
    std::atomic<uint32_t> workspaces{0}, pooled{0};
    std::atomic<uint64_t> retainLimit{0};  // 0 = keep every released workspace
};
PoolCounters& Counters() {
    static PoolCounters c;
    return c;
}
void NoteDeviceAlloc(uint64_t bytes) {
    PoolCounters& c = Counters();
    const uint64_t live = c.liveBytes.fetch_add(bytes) + bytes;
    uint64_t peak = c.peakBytes.load();
    while (live > peak && !c.peakBytes.compare_exchange_weak(peak, live)) {
    }
    c.mallocs++;
}
void NoteDeviceFree(uint64_t bytes) {
    Counters().liveBytes.fetch_sub(bytes);
    Counters().frees++;
}
}  // namespace

WorkspacePoolStats GetWorkspacePoolStats() {
    PoolCounters& c = Counters();
    WorkspacePoolStats s;
    s.liveBytes = c.liveBytes.load();
    s.peakBytes = c.peakBytes.load();
    s.mallocs = c.mallocs.load();
    s.frees = c.frees.load();
    s.workspaces = c.workspaces.load();
    s.pooled = c.pooled.load();
    s.retainLimit = c.retainLimit.load();
    return s;
}

void ReportWorkspacePool(const char* why) {
    const WorkspacePoolStats s = GetWorkspacePoolStats();
    std::printf("[POOLSTAT] pool=search_workspace %s liveMB=%.1f peakMB=%.1f workspaces=%u pooled=%u mallocs=%llu "
                "frees=%llu retainLimitMB=%.1f\n",
                why, s.liveBytes / 1048576.0, s.peakBytes / 1048576.0, s.workspaces, s.pooled,
                static_cast<unsigned long long>(s.mallocs), static_cast<unsigned long long>(s.frees),
                s.retainLimit / 1048576.0);
    std::fflush(stdout);
}

class SearchWorkspace {
   public:
    struct Buf {
        void* p = nullptr;
        size_t n = 0;
        bool pinned = false;
        void Free() {
            if (!p) return;
            if (pinned) {
                (void)hipHostFree(p);
            } else {
                (void)hipFree(p);
                NoteDeviceFree(n);
            }
            p = nullptr;
            n = 0;
        }
        bool Ensure(size_t bytes) {
            if (bytes <= n) return true;
            size_t want = std::max(bytes, n + n / 4);
            want = (want + 4095) / 4096 * 4096;
            Free();
            hipError_t e = pinned ? hipHostMalloc(&p, want, hipHostMallocDefault) : hipMalloc(&p, want);
            if (e != hipSuccess) {
                p = nullptr;
                FRK_LOG_ERROR("workspace allocation of " << want << " B failed");
                return false;
            }
            n = want;
            if (!pinned) NoteDeviceAlloc(want);
            return true;
        }
        template <typename T>
        T* As() const {
            return static_cast<T*>(p);
        }
    };

    explicit SearchWorkspace(int32_t device) : m_device(device) {
        Counters().workspaces++;
        (void)hipSetDevice(device);
        (void)hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
        for (auto& e : ev) (void)hipEventCreate(&e);
        hbatch.pinned = hout.pinned = true;
    }
    ~SearchWorkspace() {
        (void)hipSetDevice(m_device);
        for (auto& g : m_graphs) (void)hipGraphExecDestroy(g.exec);
        for (Buf* b : {&batch, &bitmaps, &bmasks, &counters, &cand, &probe, &lp, &rescored, &out, &mlp, &hbatch, &hout})
            b->Free();
        for (auto& e : ev) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(stream);
        Counters().workspaces--;
    }

    // Instantiated pipeline graphs of this workspace, keyed by every scalar and address the captured
    // sequence bakes in; a key is captured on its second batch (the first runs eagerly).
    hipGraphExec_t FindGraph(const std::vector<uint64_t>& key) const {
        for (const auto& g : m_graphs)
            if (g.key == key) return g.exec;
        return nullptr;
    }
    bool SeenBefore(const std::vector<uint64_t>& key) const {
        return std::find(m_seen.begin(), m_seen.end(), key) != m_seen.end();
    }
    void MarkSeen(const std::vector<uint64_t>& key) {
        if (m_seen.size() >= 128) m_seen.erase(m_seen.begin());
        m_seen.push_back(key);
    }
    void AddGraph(const std::vector<uint64_t>& key, hipGraphExec_t exec) {
        if (m_graphs.size() >= 64) {
            (void)hipGraphExecDestroy(m_graphs.front().exec);
            m_graphs.erase(m_graphs.begin());
        }
        m_graphs.push_back({key, exec});
    }

    static std::unique_ptr<SearchWorkspace> Acquire(int32_t device) {
        std::lock_guard<std::mutex> g(PoolMutex());
        auto& pool = Pool();
        for (auto it = pool.begin(); it != pool.end(); ++it) {
            if ((*it)->m_device == device) {
                auto w = std::move(*it);
                pool.erase(it);
                Counters().pooled--;
                return w;
            }
        }
        return std::make_unique<SearchWorkspace>(device);
    }
    // Back to the pool, unless the workspaces then hold more device memory than the retention limit:
    // such a workspace is freed (its stream synchronised first), so idle workspaces never keep the
    // process above the limit.
    // Frees idle workspaces while the pool holds more than the retention limit.
    static void TrimPool() {
        std::vector<std::unique_ptr<SearchWorkspace>> drop;
        {
            std::lock_guard<std::mutex> g(PoolMutex());
            const uint64_t limit = Counters().retainLimit.load();
            auto& pool = Pool();
            uint64_t live = Counters().liveBytes.load();
            while (limit && live > limit && !pool.empty()) {
                live -= std::min<uint64_t>(live, pool.back()->DeviceBytes());
                drop.push_back(std::move(pool.back()));
                pool.pop_back();
                Counters().pooled--;
            }
        }
        for (auto& w : drop) (void)hipStreamSynchronize(w->stream);
    }
    uint64_t DeviceBytes() const {
        uint64_t n = 0;
        for (const Buf* b : {&batch, &bitmaps, &bmasks, &counters, &cand, &probe, &lp, &rescored, &out, &mlp})
            n += b->n;
        return n;
    }
    static void Release(std::unique_ptr<SearchWorkspace> w) {
        const uint64_t limit = Counters().retainLimit.load();
        if (limit && Counters().liveBytes.load() > limit) {
            (void)hipStreamSynchronize(w->stream);
            w.reset();
            return;
        }
        std::lock_guard<std::mutex> g(PoolMutex());
        Pool().push_back(std::move(w));
        Counters().pooled++;
    }

    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    Buf batch, bitmaps, bmasks, counters, cand, probe, lp, rescored, out, mlp, hbatch, hout;

   private:
    static std::mutex& PoolMutex() {
        static std::mutex m;
        return m;
    }
    static std::vector<std::unique_ptr<SearchWorkspace>>& Pool() {
        static auto* p = new std::vector<std::unique_ptr<SearchWorkspace>>();  // lives until exit
        return *p;
    }
    struct GraphEntry {
        std::vector<uint64_t> key;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> m_graphs;
    std::vector<std::vector<uint64_t>> m_seen;
    int32_t m_device;
};

void SetWorkspaceRetainLimit(uint64_t bytes) {
    Counters().retainLimit.store(bytes);
    SearchWorkspace::TrimPool();
}

// ------------------------------------------------------------------------------------------ searcher
FullRecallSearcher::FullRecallSearcher(int32_t deviceId, std::shared_ptr<DataTable> dataTable)
    : m_deviceId(deviceId), m_table(std::move(dataTable)) {}

FullRecallSearcher::~FullRecallSearcher() = default;

void FullRecallSearcher::Clear() {
    m_trees.clear();
    m_queries.clear();
    m_topKs.clear();
    m_dim = 0;
}

void FullRecallSearcher::AddQuery(QueryNode* queryTree, const std::vector<float>& query, uint32_t topK) {
    AddQuery(queryTree, query.data(), static_cast<uint32_t>(query.size()), topK);
}

void FullRecallSearcher::AddQuery(QueryNode* queryTree, const float* query, uint32_t dim, uint32_t topK) {
    if (m_dim == 0) m_dim = dim;
    m_trees.push_back(queryTree);
    const size_t at = m_queries.size();
    m_queries.resize(at + m_dim, 0.0f);
    std::memcpy(m_queries.data() + at, query, sizeof(float) * std::min(dim, m_dim));
    m_topKs.push_back(topK);
}

bool FullRecallSearcher::BatchSearch(std::vector<std::shared_ptr<FullRecallResult>>& results,
                                     const std::string& vectorFieldName, bool isMultiShard) {
    std::vector<AggrDeviceResult> dev;
    if (!BatchSearchDevice(vectorFieldName, isMultiShard, dev)) return false;
    return BatchSearchExtract(dev, results, isMultiShard);
}

bool FullRecallSearcher::BatchSearchDevice(const std::string& vectorFieldName, bool,
                                           std::vector<AggrDeviceResult>& deviceResults) {
    if (!m_table) {
        FRK_LOG_ERROR("searcher has no data table");
        return false;
    }
    VectorFieldData* vf = nullptr;
    if (!m_table->GetVectorFieldData(vectorFieldName, vf)) {
        FRK_LOG_ERROR("vector field not in the index: " << vectorFieldName);
        return false;
    }
    if (!m_topKs.empty() && m_dim != vf->Dimension()) {
        FRK_LOG_ERROR("query dimension " << m_dim << " != field dimension " << vf->Dimension());
        return false;
    }
    if (hipSetDevice(m_deviceId) != hipSuccess) return false;
    return RunDevice(*vf, deviceResults);
}

bool FullRecallSearcher::FillResult(const std::vector<uint64_t>& keys, bool isMultiShard, FullRecallResult& r) const {
    const DocIdMapping* map = m_table->GetDocIdMapping();
    r.isMultiShard = isMultiShard;
    for (uint64_t key : keys) {
        const uint32_t ldoc = frk::KeyDoc(key);
        const float score = frk::KeyScore(key);
        uint64_t gdoc = ldoc;
        if (map && !map->GetGDocId(ldoc, gdoc)) {
            FRK_LOG_ERROR("no global id for local doc " << ldoc);
            return false;
        }
        if (isMultiShard) {
            r.scoreWithIndexes.emplace_back(score, ldoc);
            r.scoreWithIndexes.back().id = gdoc;
        } else {
            r.docIds.push_back(static_cast<int64_t>(gdoc));
            r.scores.push_back(score);
        }
    }
    return true;
}

bool FullRecallSearcher::BatchSearchExtract(std::vector<AggrDeviceResult>& dev,
                                            std::vector<std::shared_ptr<FullRecallResult>>& results,
                                            bool isMultiShard) {
    results.clear();
    results.reserve(dev.size());
    for (const AggrDeviceResult& d : dev) {
        auto r = std::make_shared<FullRecallResult>();
        if (!FillResult(d.keys, isMultiShard, *r)) return false;
        results.push_back(std::move(r));
    }
    return true;
}

// ------------------------------------------------------------------------------------- unfused path
namespace {
void FreeBlock(GmBlock& b) {
    if (b.data) (void)hipFree(b.data);
    b.data = nullptr;
    b.size = 0;
}
}  // namespace

bool FullRecallSearcher::OneFilter(GmBlock& filterResultChunk) {
    if (!m_table || m_trees.size() != 1) {
        FRK_LOG_ERROR("OneFilter needs a data table and exactly one query (has " << m_trees.size() << ")");
        return false;
    }
    if (hipSetDevice(m_deviceId) != hipSuccess) return false;
    const uint32_t dnps = m_table->GetDocNumPerSegment(), segs = m_table->GetSegmentNum();
    if (dnps % 64 != 0) {
        FRK_LOG_ERROR("docNumPerSegment " << dnps << " is not a multiple of 64 (required by the device filter)");
        return false;
    }
    FilterExpr fe;
    if (!PrepareFilterExpr(*m_table, m_trees[0], fe)) return false;
    frk_filter_query d{};
    d.expr_len = fe.emptyTree ? 0 : static_cast<uint32_t>(fe.postExpr.size());
    d.n_postings = fe.emptyTree ? 0 : static_cast<uint32_t>(fe.postingDeviceAddrs.size());
    d.op_num = fe.opNum;
    d.flags = fe.emptyTree ? FRK_FILTER_MATCH_ALL : 0;
    std::vector<uint64_t> table;
    for (uint32_t sg = 0; sg < segs && !fe.emptyTree; sg++)
        for (const auto* addrs : fe.postingDeviceAddrs) table.push_back((*addrs)[sg]);
    // one staging buffer: [descriptor | counters | program | operand table]
    const size_t oExpr = 64, oTab = oExpr + (fe.postExpr.size() * 4 + 15) / 16 * 16;
    std::vector<uint8_t> host(oTab + table.size() * 8 + 8, 0);
    std::memcpy(host.data(), &d, sizeof(d));
    if (!fe.postExpr.empty()) std::memcpy(host.data() + oExpr, fe.postExpr.data(), fe.postExpr.size() * 4);
    if (!table.empty()) std::memcpy(host.data() + oTab, table.data(), table.size() * 8);
    const size_t bytes = static_cast<size_t>(segs) * dnps / 8;
    uint8_t* stage = nullptr;
    char* out = nullptr;
    if (hipMalloc(&stage, host.size()) != hipSuccess || hipMalloc(&out, bytes) != hipSuccess) {
        (void)hipFree(stage);
        FRK_LOG_ERROR("OneFilter: device allocation failed");
        return false;
    }
    uint32_t cnt[2] = {0, 0};
    bool ok = hipMemcpy(stage, host.data(), host.size(), hipMemcpyHostToDevice) == hipSuccess &&
              frk_filter_eval(reinterpret_cast<const frk_filter_query*>(stage), 1,
                              reinterpret_cast<const uint32_t*>(stage + oExpr),
                              reinterpret_cast<const uint64_t*>(stage + oTab), segs, dnps, m_table->GetDocNum(),
                              reinterpret_cast<uint64_t*>(out), bytes / 8, reinterpret_cast<uint32_t*>(stage + 32),
                              reinterpret_cast<uint32_t*>(stage + 36), d.n_postings, nullptr) == FRK_OK &&
              hipMemcpy(cnt, stage + 32, sizeof(cnt), hipMemcpyDeviceToHost) == hipSuccess;
    (void)hipFree(stage);
    if (!ok || cnt[1] != 0) {
        FRK_LOG_ERROR("OneFilter: filter failed (" << frk_last_error() << ", status " << cnt[1] << ")");
        (void)hipFree(out);
        return false;
    }
    m_passing.assign(1, cnt[0]);
    filterResultChunk.poolName = GmPoolName::TEXT_FILTER_RESULT_POOL;
    filterResultChunk.data = out;
    filterResultChunk.size = static_cast<uint32_t>(bytes);
    return true;
}

bool FullRecallSearcher::BatchScore(const std::string& vectorFieldName, GmBlock& resultChunk) {
    VectorFieldData* vf = nullptr;
    if (!m_table || !m_table->GetVectorFieldData(vectorFieldName, vf)) {
        FRK_LOG_ERROR("vector field not in the index: " << vectorFieldName);
        return false;
    }
    const uint32_t Q = static_cast<uint32_t>(m_topKs.size());
    if (Q == 0 || m_dim != vf->Dimension()) {
        FRK_LOG_ERROR("BatchScore needs queries of the field's dimension " << vf->Dimension());
        return false;
    }
    if (hipSetDevice(m_deviceId) != hipSuccess) return false;
    const frk_doc_store* store = vf->Store(m_options.dtype);
    if (!store) {
        FRK_LOG_ERROR("vector store for dtype " << m_options.dtype << " was not built");
        return false;
    }
    const uint32_t dp = store->dim;
    const uint64_t stride = m_table->GetScoreExtendDocNum();
    std::vector<float> q(static_cast<size_t>(Q) * dp, 0.0f);
    for (uint32_t i = 0; i < Q; i++) std::memcpy(&q[static_cast<size_t>(i) * dp], &m_queries[static_cast<size_t>(i) * m_dim], 4 * m_dim);
    float* dq = nullptr;
    char* out = nullptr;
    const size_t bytes = sizeof(float) * Q * stride;
    if (bytes > 0xFFFFFFFFull) {
        FRK_LOG_ERROR("BatchScore: a " << bytes << "-byte score matrix exceeds the block size field");
        return false;
    }
    if (hipMalloc(&dq, q.size() * 4) != hipSuccess || hipMalloc(&out, bytes) != hipSuccess) {
        (void)hipFree(dq);
        FRK_LOG_ERROR("BatchScore: device allocation failed");
        return false;
    }
    const bool ok = hipMemcpy(dq, q.data(), q.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
                    frk_score_matrix(store, dq, Q, reinterpret_cast<float*>(out), stride, nullptr) == FRK_OK &&
                    hipDeviceSynchronize() == hipSuccess;
    (void)hipFree(dq);
    if (!ok) {
        FRK_LOG_ERROR("BatchScore: scoring failed: " << frk_last_error());
        (void)hipFree(out);
        return false;
    }
    resultChunk.poolName = GmPoolName::VECTOR_SCORE_RESULT_POOL;
    resultChunk.data = out;
    resultChunk.size = static_cast<uint32_t>(bytes);
    resultChunk.readCount = std::make_shared<std::atomic<uint32_t>>(Q);
    return true;
}

bool FullRecallSearcher::OneAggregator(GmBlock& resultChunk, GmBlock& filterResultChunk, uint32_t resultInDeviceOffset,
                                       bool isMultiShard, std::vector<std::shared_ptr<FullRecallResult>>& results) {
    if (!m_table || m_topKs.empty() || !resultChunk.data || !filterResultChunk.data) {
        FRK_LOG_ERROR("OneAggregator needs a query, a score matrix and a filter block");
        return false;
    }
    if (hipSetDevice(m_deviceId) != hipSuccess) return false;
    const uint32_t k = m_topKs[0];
    const uint32_t nDocs = m_table->GetDocNum();
    const uint64_t stride = m_table->GetScoreExtendDocNum();
    if (k > FRK_SELECT_MAX_K || (static_cast<uint64_t>(resultInDeviceOffset) + 1) * stride * 4 > resultChunk.size) {
        FRK_LOG_ERROR("OneAggregator: k " << k << " or row " << resultInDeviceOffset << " out of range");
        return false;
    }
    const float* row = reinterpret_cast<const float*>(resultChunk.data) + static_cast<uint64_t>(resultInDeviceOffset) * stride;
    const uint32_t cap = std::max<uint32_t>(nDocs, 1);
    uint64_t* cand = nullptr;
    uint8_t* small = nullptr;  // [counter | out count | pad][k keys]
    if (hipMalloc(&cand, sizeof(uint64_t) * cap) != hipSuccess ||
        hipMalloc(&small, 256 + sizeof(uint64_t) * std::max<uint32_t>(k, 1)) != hipSuccess) {
        (void)hipFree(cand);
        FRK_LOG_ERROR("OneAggregator: device allocation failed");
        return false;
    }
    uint32_t* cnt = reinterpret_cast<uint32_t*>(small);
    uint64_t* outKeys = reinterpret_cast<uint64_t*>(small + 256);
    std::vector<uint64_t> keys(k);
    uint32_t counts[2] = {0, 0};
    bool ok = hipMemset(small, 0, 256) == hipSuccess &&
              frk_aggregate(row, reinterpret_cast<const uint64_t*>(filterResultChunk.data), nDocs, cand, cap, cnt,
                            nullptr) == FRK_OK &&
              frk_select_topk(cand, cap, cnt, 1, 1, nullptr, k, outKeys, std::max<uint32_t>(k, 1), cnt + 1, nullptr,
                              nullptr) == FRK_OK &&
              hipMemcpy(counts, cnt, sizeof(counts), hipMemcpyDeviceToHost) == hipSuccess &&
              hipMemcpy(keys.data(), outKeys, sizeof(uint64_t) * k, hipMemcpyDeviceToHost) == hipSuccess;
    (void)hipFree(cand);
    (void)hipFree(small);
    if (ok) {
        keys.resize(std::min<uint32_t>(counts[1], k));
        auto r = std::make_shared<FullRecallResult>();
        ok = FillResult(keys, isMultiShard, *r);
        if (ok) results.push_back(std::move(r));
        m_passing.assign(1, counts[0]);
    } else {
        FRK_LOG_ERROR("OneAggregator: aggregation failed: " << frk_last_error());
    }
    // the score matrix is freed by its last reader, the filter block by its one reader
    if (resultChunk.readCount && resultChunk.readCount->fetch_sub(1) == 1) FreeBlock(resultChunk);
    FreeBlock(filterResultChunk);
    return ok;
}

namespace {
template <typename T>
size_t Pad(size_t n) {
    return (n * sizeof(T) + 255) / 256 * 256;
}
}  // namespace

bool FullRecallSearcher::RunDevice(const VectorFieldData& vf, std::vector<AggrDeviceResult>& out) {
    using Clock = std::chrono::steady_clock;
    const auto tStart = Clock::now();
    auto msSince = [](Clock::time_point a) {
        return std::chrono::duration<float, std::milli>(Clock::now() - a).count();
    };
    const uint32_t Q = static_cast<uint32_t>(m_topKs.size());
    out.assign(Q, AggrDeviceResult());
    m_stats = BatchStats();
    m_passing.assign(Q, 0);
    if (Q == 0) return true;
    const uint32_t dtype = m_options.dtype;
    const frk_doc_store* store = vf.Store(dtype);
    const frk_doc_store* f32 = vf.Store(FRK_DTYPE_F32);
    const float* rowsF32 = vf.RowsF32();
    if (!store || !f32) {
        FRK_LOG_ERROR("vector store for dtype " << dtype << " was not built for field " << vf.Name());
        return false;
    }
    const uint32_t nDocs = m_table->GetDocNum();
    const uint32_t dnps = m_table->GetDocNumPerSegment();
    const uint32_t segs = m_table->GetSegmentNum();
    if (dnps % 64 != 0) {
        FRK_LOG_ERROR("docNumPerSegment " << dnps << " is not a multiple of 64 (required by the device filter)");
        return false;
    }
    const uint64_t maskStride = static_cast<uint64_t>(segs) * (dnps / 64);
    // BF16 and FP8 stores select k * m candidates by their own scores and rescore them in FP32; the FP32
    // store and its F16X2 split score to FP32 accuracy and select the k directly
    const bool lowp = dtype == FRK_DTYPE_BF16 || dtype == FRK_DTYPE_FP8E4M3;
    const uint32_t m = lowp ? std::max(1u, m_options.rescoreFactor) : 1;
    uint32_t kmax = 1;
    for (uint32_t k : m_topKs) kmax = std::max(kmax, k);
    if (kmax > FRK_SELECT_MAX_K) {
        FRK_LOG_ERROR("topK " << kmax << " exceeds " << FRK_SELECT_MAX_K);
        return false;
    }
    // keys a query selects on the main pass: k * m, at most what one select takes (never below k)
    auto kPrime = [&](uint32_t k) { return std::min<uint32_t>(k * m, std::max<uint32_t>(k, FRK_SELECT_MAX_K)); };
    const uint32_t kSel = kPrime(kmax);

    // ---- compile the filters
    std::vector<frk_filter_query> descs(Q);
    std::vector<uint32_t> expr;
    std::vector<uint64_t> table;
    uint32_t maxPostings = 0;
    const bool filterStats = FilterStatsEnabled() || EmptyStatsEnabled();
    std::vector<FilterExpr> kept;  // the batch's compiled filters, kept only for the statistics
    for (uint32_t q = 0; q < Q; q++) {
        FilterExpr fe;
        if (!PrepareFilterExpr(*m_table, m_trees[q], fe)) return false;
        if (filterStats) kept.push_back(fe);
        frk_filter_query& d = descs[q];
        d.expr_off = static_cast<uint32_t>(expr.size());
        d.table_off = static_cast<uint32_t>(table.size());
        d.flags = fe.emptyTree ? FRK_FILTER_MATCH_ALL : 0;
        d.expr_len = fe.emptyTree ? 0 : static_cast<uint32_t>(fe.postExpr.size());
        d.n_postings = fe.emptyTree ? 0 : static_cast<uint32_t>(fe.postingDeviceAddrs.size());
        d.op_num = fe.opNum;
        maxPostings = std::max(maxPostings, d.n_postings);
        if (fe.emptyTree) continue;
        expr.insert(expr.end(), fe.postExpr.begin(), fe.postExpr.end());
        for (uint32_t s = 0; s < segs; s++)
            for (const auto* addrs : fe.postingDeviceAddrs) table.push_back((*addrs)[s]);
    }
    if (table.size() > 0xFFFFFFFFull || expr.size() > 0xFFFFFFFFull) {
        FRK_LOG_ERROR("filter tables too large for one batch");
        return false;
    }
    if (filterStats) {
        std::vector<const FilterExpr*> ptrs;
        for (const FilterExpr& fe : kept) ptrs.push_back(&fe);
        NoteFilterBatch(*m_table, ptrs, segs, sizeof(uint64_t) * Q * maskStride);
    }

    // ---- pack the batch. Fixed-size regions first (descriptors | queries per store dim | k | k*m |
    // floors | model inputs | floor offsets: their offsets depend only on the batch shape), then the
    // filter programs and operand table in power-of-two capacities, so a captured pipeline graph sees
    // the same device addresses for every batch of its shape.
    const uint32_t dpMain = store->dim, dpF32 = f32->dim;
    const frk_floor_model* model = m_options.floorModel;
    const bool learned = model != nullptr && m_options.floors == nullptr;
    const uint32_t inDim = learned ? static_cast<uint32_t>(model->mlp.in_dim) : 0;
    if (learned && (inDim != m_dim + kFloorFeatures || model->mlp.hidden <= 0 || !model->feat_mean ||
                    !model->feat_std)) {
        FRK_LOG_ERROR("floor model expects " << model->mlp.in_dim << " inputs, queries have " << m_dim << " + "
                                             << kFloorFeatures << " features");
        return false;
    }
    auto pow2 = [](size_t n) {
        size_t c = 256;
        while (c < n) c <<= 1;
        return c;
    };
    const size_t exprCap = pow2(expr.size() + 1), tableCap = pow2(table.size() + 1);
    const size_t oDesc = 0;
    const size_t oQMain = oDesc + Pad<frk_filter_query>(Q);
    const size_t oQF32 = oQMain + Pad<float>(static_cast<size_t>(Q) * dpMain);
    const size_t oK = oQF32 + (lowp ? Pad<float>(static_cast<size_t>(Q) * dpF32) : 0);
    const size_t oKm = oK + Pad<uint32_t>(Q);
    const size_t oFloor = oKm + Pad<uint32_t>(Q);
    const size_t oX = oFloor + Pad<float>(Q);
    const size_t oOff = oX + (learned ? Pad<float>(static_cast<size_t>(Q) * inDim) : 0);
    const size_t oExpr = oOff + (learned ? Pad<float>(Q) : 0);
    const size_t oTab = oExpr + Pad<uint32_t>(exprCap);
    const size_t batchBytes = oTab + Pad<uint64_t>(tableCap);
    std::unique_ptr<SearchWorkspace> ws = SearchWorkspace::Acquire(m_deviceId);
    struct Releaser {
        std::unique_ptr<SearchWorkspace>& w;
        ~Releaser() {
            if (w) SearchWorkspace::Release(std::move(w));
        }
    } releaser{ws};
    if (!ws->hbatch.Ensure(batchBytes) || !ws->batch.Ensure(batchBytes)) return false;
    uint8_t* hb = ws->hbatch.As<uint8_t>();
    std::memcpy(hb + oDesc, descs.data(), sizeof(frk_filter_query) * Q);
    std::memcpy(hb + oExpr, expr.data(), sizeof(uint32_t) * expr.size());
    std::memcpy(hb + oTab, table.data(), sizeof(uint64_t) * table.size());
    auto putQueries = [&](size_t off, uint32_t dp) {
        float* dst = reinterpret_cast<float*>(hb + off);
        std::memset(dst, 0, sizeof(float) * Q * dp);
        for (uint32_t q = 0; q < Q; q++) std::memcpy(dst + static_cast<size_t>(q) * dp, &m_queries[q * m_dim], 4 * m_dim);
    };
    putQueries(oQMain, dpMain);
    if (lowp) putQueries(oQF32, dpF32);
    uint32_t* hk = reinterpret_cast<uint32_t*>(hb + oK);
    uint32_t* hkm = reinterpret_cast<uint32_t*>(hb + oKm);
    for (uint32_t q = 0; q < Q; q++) {
        hk[q] = m_topKs[q];
        hkm[q] = kPrime(m_topKs[q]);
    }
    const bool callerFloor = m_options.floors != nullptr;
    if (callerFloor) std::memcpy(hb + oFloor, m_options.floors, sizeof(float) * Q);
    if (learned) {
        float* hx = reinterpret_cast<float*>(hb + oX);
        float* hoff = reinterpret_cast<float*>(hb + oOff);
        float f[kFloorFeatures];
        for (uint32_t q = 0; q < Q; q++) {
            const float* qv = &m_queries[static_cast<size_t>(q) * m_dim];
            QueryNode* tree = (descs[q].flags & FRK_FILTER_MATCH_ALL) ? nullptr : m_trees[q];
            FloorFeatures(*m_table, tree, kPrime(m_topKs[q]), qv, m_dim, f);
            float* row = hx + static_cast<size_t>(q) * inDim;
            std::memcpy(row, qv, sizeof(float) * m_dim);
            for (uint32_t j = 0; j < kFloorFeatures; j++) row[m_dim + j] = (f[j] - model->feat_mean[j]) / model->feat_std[j];
            // too few docs expected to pass for a floor to pay off (f[2] = log10 expected passing)
            const bool allow = std::pow(10.0, static_cast<double>(f[2])) >=
                               static_cast<double>(model->min_passing_ratio) * kPrime(m_topKs[q]);
            // margin of this k's decade (f[1] = log10 k) when the model has per-decade margins
            float margin = model->margin;
            if (model->margin_by_k) {
                const int b = std::min(3, std::max(0, static_cast<int>(std::floor(f[1] + 1e-4f))));
                margin = model->margin_by_k[b];
            }
            hoff[q] = allow ? -margin : -std::numeric_limits<float>::infinity();
        }
    }
    const bool floorChecked = callerFloor || learned;  // a floor that may cut below k: verify, rerun

    // ---- device buffers
    const uint64_t nBlocks = (static_cast<uint64_t>(nDocs) + 31) / 32;
    // Probe sample: every probeStep-th doc block, reduced to one max per (query, block). The floor is the
    // k'-th largest sampled block maximum, about the (k' * step)-th best passing score, so the main pass
    // keeps ~k' * step survivors per query while the probe reads 1 / step of the store. Automatic step:
    // ~8,192 survivors for the batch's largest k' (a survivor costs the main pass far less than a
    // sampled block costs the probe), at least 4 sampled blocks per key, every 2nd to every 64th block.
    // Batches below 32 queries (the reruns of overshooting floors among them) aim at proportionally
    // more survivors per query: their main pass reads the whole store for a few queries anyway, so
    // the probe's share is what a larger step saves (one query at k' = 4,000 over 20M docs: the probe
    // reads 1/39 of the store instead of half of it).
    uint32_t probeStep = m_options.probeStep;
    const bool autoStep = probeStep == 0;
    if (autoStep) {
        const uint64_t target = 8192ull * std::max<uint64_t>(1, 32 / std::max<uint32_t>(Q, 1));
        probeStep = static_cast<uint32_t>(std::min<uint64_t>(
            64, std::max<uint64_t>(2, std::min<uint64_t>(target / kSel, nBlocks / (4ull * kSel)))));
    }
    // a query whose k*m exceeds the sampled blocks gets a -inf floor (no pruning), which stays exact
    uint32_t kSelMin = kSel;
    for (uint32_t k : m_topKs) kSelMin = std::min(kSelMin, kPrime(k));
    const bool useProbe = m_options.probe && !floorChecked && probeStep > 1 && nBlocks / probeStep >= kSelMin;
    uint32_t cap = m_options.candCap;
    if (cap == 0) {
        cap = useProbe ? static_cast<uint32_t>(std::max<uint64_t>(65536, 16ull * kSel * probeStep))
                       : static_cast<uint32_t>(std::min<uint64_t>(nDocs, std::max<uint64_t>(65536, 64ull * kSel)));
        // a caller or learned floor sits a calibrated margin below the k-th score, so the survivors
        // above it can number many times k (where scores crowd near the k-th best); room for 256 k
        // keys (at least 256 Ki) keeps overflow reruns, each a whole extra pass, rare
        if (floorChecked) cap = static_cast<uint32_t>(std::max<uint64_t>(cap, std::max<uint64_t>(262144, 256ull * kSel)));
        cap = std::min<uint32_t>(cap, std::max<uint32_t>(nDocs, 1));
    }
    const uint64_t probeBlocks = useProbe ? (nBlocks + probeStep - 1) / probeStep : 0;
    m_stats.probeStep = useProbe ? probeStep : 0;

    // counters: [filter counts Q][status 64][out counts Q][lp counts Q][cand counts Q x kCandStride]; the
    // scorer's per-query candidate counters sit one 128-byte line apart, so their atomics spread over
    // the L2 channels
    constexpr uint32_t kCandStride = 32;
    const size_t cFilter = 0, cStatus = Q, cOut = Q + 64, cLp = 2 * Q + 64, cCand = 3 * Q + 64;
    const size_t nCounters = 3 * Q + 64 + static_cast<size_t>(kCandStride) * Q;
    float* dKth = nullptr;
    const uint32_t qStride = (Q + 31) / 32 * 32;  // per-doc filter words: qStride u32 per doc block
    if (!ws->bitmaps.Ensure(sizeof(uint64_t) * Q * maskStride) || !ws->bmasks.Ensure(sizeof(uint32_t) * qStride * nBlocks) || !ws->counters.Ensure(sizeof(uint32_t) * nCounters + 4 * Q) ||
        !ws->cand.Ensure(sizeof(uint64_t) * Q * cap) || !ws->out.Ensure(sizeof(uint64_t) * Q * kmax) ||
        (useProbe && !ws->probe.Ensure(sizeof(float) * Q * probeBlocks)) ||
        (lowp && (!ws->lp.Ensure(sizeof(uint64_t) * Q * kSel) || !ws->rescored.Ensure(sizeof(uint64_t) * Q * kSel))) ||
        (learned && !ws->mlp.Ensure(sizeof(float) * (4ull * Q * static_cast<uint32_t>(model->mlp.hidden) + 2ull * Q))) ||
        !ws->hout.Ensure(sizeof(uint64_t) * Q * kmax + sizeof(uint32_t) * nCounters))
        return false;
    uint32_t* dCnt = ws->counters.As<uint32_t>();
    dKth = reinterpret_cast<float*>(dCnt + nCounters);
    uint8_t* db = ws->batch.As<uint8_t>();
    hipStream_t s = ws->stream;
    auto fail = [&](const char* what) {
        FRK_LOG_ERROR(what << ": " << frk_last_error());
        (void)hipStreamSynchronize(s);
        return false;
    };

    // ---- the pipeline: one stream-ordered sequence, launched eagerly or replayed from a graph
    const float* dQMain = reinterpret_cast<const float*>(db + oQMain);
    const float* dQF32 = lowp ? reinterpret_cast<const float*>(db + oQF32) : dQMain;
    const uint32_t* dK = reinterpret_cast<const uint32_t*>(db + oK);
    const uint32_t* dKm = reinterpret_cast<const uint32_t*>(db + oKm);
    uint64_t* dMask = ws->bitmaps.As<uint64_t>();
    uint32_t* dBMask = ws->bmasks.As<uint32_t>();
    uint8_t* ho = ws->hout.As<uint8_t>();
    // the filter launch only distinguishes up to 4, up to 8 and more operands (its LDS slot sizing)
    const uint32_t mpLaunch = maxPostings <= 4 ? 4 : maxPostings <= 8 ? 8 : 16;
    auto enqueue = [&](bool events) -> const char* {
        auto rec = [&](int i) {
            if (events) (void)hipEventRecord(ws->ev[i], s);
        };
        rec(0);
        if (hipMemcpyAsync(db, hb, batchBytes, hipMemcpyHostToDevice, s) != hipSuccess) return "batch H2D";
        if (frk_fill_u32(dCnt, nCounters, 0, s) != 0) return "counter reset";
        if (frk_filter_eval(reinterpret_cast<const frk_filter_query*>(db + oDesc), Q,
                            reinterpret_cast<const uint32_t*>(db + oExpr), reinterpret_cast<const uint64_t*>(db + oTab),
                            segs, dnps, nDocs, dMask, maskStride, dCnt + cFilter, dCnt + cStatus, mpLaunch, s) != 0)
            return "filter";
        if (frk_masks_to_blocks(dMask, maskStride, Q, nDocs, dBMask, qStride, s) != 0) return "mask transpose";
        rec(1);
        const float* dFloor = callerFloor ? reinterpret_cast<const float*>(db + oFloor) : nullptr;
        if (useProbe) {
            // the k-th (k*m on low precision) largest block maximum of the sample is the floor
            if (frk_score_blockmax(store, dQMain, Q, dBMask, qStride, 0, probeStep, probeBlocks, ws->probe.As<float>(),
                                   static_cast<uint32_t>(probeBlocks), s) != 0)
                return "probe score";
            if (frk_kth_largest(ws->probe.As<float>(), static_cast<uint32_t>(probeBlocks),
                                static_cast<uint32_t>(probeBlocks), Q, dKm, 0, dKth, s) != 0)
                return "probe select";
            dFloor = dKth;
        }
        if (learned) {
            // the floor model on this stream: prediction - margin, or -inf where no floor applies
            float* work = ws->mlp.As<float>();
            float* y = work + 4ull * Q * static_cast<uint32_t>(model->mlp.hidden);
            float* fl = y + Q;
            if (frk_mlp_forward(&model->mlp, static_cast<int32_t>(Q), reinterpret_cast<const float*>(db + oX),
                                static_cast<int32_t>(inDim), work, y, s) != 0 ||
                frk_mlp_floors(static_cast<int32_t>(Q), y, reinterpret_cast<const float*>(db + oOff), fl, s) != 0)
                return "floor model";
            dFloor = fl;
        }
        rec(2);
        if (frk_score_select(store, dQMain, Q, dBMask, qStride, dFloor, 0, 1, nBlocks, ws->cand.As<uint64_t>(), cap,
                             dCnt + cCand, kCandStride, s) != 0)
            return "score";
        rec(3);
        if (!lowp) {
            if (frk_select_topk(ws->cand.As<uint64_t>(), cap, dCnt + cCand, kCandStride, Q, dK, 0, ws->out.As<uint64_t>(), kmax,
                                dCnt + cOut, nullptr, s) != 0)
                return "select";
            rec(4);
            rec(5);
        } else {
            // the k*m best by the low-precision score, unsorted: they are rescored and sorted after
            if (frk_select_top_set(ws->cand.As<uint64_t>(), cap, dCnt + cCand, kCandStride, Q, dKm, 0, ws->lp.As<uint64_t>(), kSel,
                                dCnt + cLp, nullptr, s) != 0)
                return "select k*m";
            rec(4);
            // from the row-major FP32 copy when the field keeps one (one contiguous read per candidate)
            const int rc = rowsF32 ? frk_rescore_rows(rowsF32, dpF32, nDocs, dQF32, Q, ws->lp.As<uint64_t>(), kSel,
                                                      dCnt + cLp, ws->rescored.As<uint64_t>(), s)
                                   : frk_rescore(f32, dQF32, Q, ws->lp.As<uint64_t>(), kSel, dCnt + cLp,
                                                 ws->rescored.As<uint64_t>(), s);
            if (rc != 0) return "rescore";
            if (frk_select_topk(ws->rescored.As<uint64_t>(), kSel, dCnt + cLp, 1, Q, dK, 0, ws->out.As<uint64_t>(), kmax,
                                dCnt + cOut, nullptr, s) != 0)
                return "select after rescore";
            rec(5);
        }
        if (hipMemcpyAsync(ho, ws->out.p, sizeof(uint64_t) * Q * kmax, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(ho + sizeof(uint64_t) * Q * kmax, dCnt, sizeof(uint32_t) * nCounters, hipMemcpyDeviceToHost,
                           s) != hipSuccess)
            return "result D2H";
        rec(6);
        return nullptr;
    };
    bool replayed = false;
    if (!m_options.graph) {
        if (const char* what = enqueue(true)) return fail(what);
    } else {
        // every scalar and address the sequence bakes into its launches
        const uint64_t key[] = {Q, kmax, kSel, m, dtype, reinterpret_cast<uint64_t>(store->tiles),
                                reinterpret_cast<uint64_t>(f32->tiles), reinterpret_cast<uint64_t>(rowsF32), nDocs,
                                segs, dnps, useProbe, probeStep,
                                probeBlocks, learned, cap, mpLaunch, batchBytes, oExpr, oTab, callerFloor, qStride,
                                reinterpret_cast<uint64_t>(ws->batch.p), reinterpret_cast<uint64_t>(ws->hbatch.p),
                                reinterpret_cast<uint64_t>(ws->bitmaps.p), reinterpret_cast<uint64_t>(ws->bmasks.p),
                                reinterpret_cast<uint64_t>(ws->counters.p), reinterpret_cast<uint64_t>(ws->cand.p),
                                reinterpret_cast<uint64_t>(ws->probe.p), reinterpret_cast<uint64_t>(ws->lp.p),
                                reinterpret_cast<uint64_t>(ws->rescored.p), reinterpret_cast<uint64_t>(ws->out.p),
                                reinterpret_cast<uint64_t>(ws->mlp.p), reinterpret_cast<uint64_t>(ws->hout.p),
                                learned ? reinterpret_cast<uint64_t>(model->mlp.w1) : 0,
                                learned ? reinterpret_cast<uint64_t>(model->mlp.w2) : 0,
                                learned ? reinterpret_cast<uint64_t>(model->mlp.w3) : 0,
                                learned ? reinterpret_cast<uint64_t>(model->mlp.b1) : 0,
                                learned ? reinterpret_cast<uint64_t>(model->mlp.b2) : 0,
                                learned ? reinterpret_cast<uint64_t>(model->mlp.b3) : 0,
                                learned ? static_cast<uint64_t>(model->mlp.hidden) : 0};
        const std::vector<uint64_t> k(std::begin(key), std::end(key));
        hipGraphExec_t exec = ws->FindGraph(k);
        if (!exec && ws->SeenBefore(k)) {
            // second batch of this shape: capture the sequence (the first ran eagerly, so one-time
            // kernel setup is done) and keep the instantiated graph
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) return fail("graph capture");
            const char* what = enqueue(false);
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (what || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(what ? what : "graph capture");
            }
            const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) return fail("graph instantiate");
            ws->AddGraph(k, exec);
        }
        if (exec) {
            (void)hipEventRecord(ws->ev[0], s);
            if (hipGraphLaunch(exec, s) != hipSuccess) return fail("graph launch");
            (void)hipEventRecord(ws->ev[6], s);
            replayed = true;
        } else {
            ws->MarkSeen(k);
            if (const char* what = enqueue(true)) return fail(what);
        }
    }
    m_stats.ms_host_prepare = msSince(tStart);
    if (hipStreamSynchronize(s) != hipSuccess) return fail("batch synchronize");
    const auto tSynced = Clock::now();

    const uint64_t* hKeys = reinterpret_cast<const uint64_t*>(ho);
    const uint32_t* hCnt = reinterpret_cast<const uint32_t*>(ho + sizeof(uint64_t) * Q * kmax);
    if (hCnt[cStatus] != 0) {
        FRK_LOG_ERROR("filter kernel rejected a program (status " << hCnt[cStatus] << ")");
        return false;
    }
    auto ms = [&](int a, int b) {
        float t = 0;
        (void)hipEventElapsedTime(&t, ws->ev[a], ws->ev[b]);
        return t;
    };
    if (!replayed) {
        m_stats.ms_filter = ms(0, 1);
        m_stats.ms_probe = ms(1, 2);
        m_stats.ms_score = ms(2, 3);
        m_stats.ms_select = ms(3, 4);
        m_stats.ms_rescore = ms(4, 5);
    }
    m_stats.ms_total = ms(0, 6);
    m_stats.graphReplayed = replayed;
    m_stats.cap = cap;

    // ---- exactness checks: overflowed candidate lists, and a caller or learned floor that cut below
    // the k-th
    std::vector<uint32_t> redo;
    uint32_t needCap = cap;
    for (uint32_t q = 0; q < Q; q++) {
        const uint32_t nCand = hCnt[cCand + static_cast<size_t>(kCandStride) * q];
        m_passing[q] = hCnt[cFilter + q];
        m_stats.candidates += nCand;
        const uint32_t want = std::min(kPrime(m_topKs[q]), m_passing[q]);
        const bool overflow = nCand > cap;
        const bool shortFloor = floorChecked && std::min(nCand, cap) < want;
        if (overflow) m_stats.overflowed++;
        if (overflow || shortFloor) {
            redo.push_back(q);
            needCap = std::max(needCap, std::min(nCand, m_passing[q]));
        }
    }
    for (uint32_t q = 0; q < Q; q++) {
        const uint32_t n = hCnt[cOut + q];
        out[q].keys.assign(hKeys + static_cast<size_t>(q) * kmax, hKeys + static_cast<size_t>(q) * kmax + n);
        out[q].passing = m_passing[q];
    }
    m_stats.ms_host_finish = msSince(tSynced);
    if (redo.empty()) return true;

    // ---- rerun only the affected queries: no caller or learned floor (the exact probe floor instead)
    // and room for every survivor; their results replace the first pass's
    m_stats.reruns = static_cast<uint32_t>(redo.size());
    const BatchStats first = m_stats;
    const std::vector<uint32_t> passing = m_passing;
    std::vector<QueryNode*> trees;
    std::vector<float> queries;
    std::vector<uint32_t> ks;
    for (uint32_t q : redo) {
        trees.push_back(m_trees[q]);
        queries.insert(queries.end(), m_queries.begin() + static_cast<size_t>(q) * m_dim,
                       m_queries.begin() + static_cast<size_t>(q + 1) * m_dim);
        ks.push_back(m_topKs[q]);
    }
    std::swap(trees, m_trees);
    std::swap(queries, m_queries);
    std::swap(ks, m_topKs);
    const SearchOptions saved = m_options;
    m_options.floors = nullptr;
    m_options.floorModel = nullptr;
    m_options.candCap = std::max<uint32_t>(needCap, cap);
    if (floorChecked) m_options.probe = true;
    // the batch's workspace goes back to the pool first, so the rerun reuses it
    SearchWorkspace::Release(std::move(ws));
    std::vector<AggrDeviceResult> sub;
    const bool ok = RunDevice(vf, sub);
    m_options = saved;
    std::swap(trees, m_trees);
    std::swap(queries, m_queries);
    std::swap(ks, m_topKs);
    if (!ok) return false;
    for (size_t i = 0; i < redo.size(); i++) out[redo[i]] = std::move(sub[i]);
    BatchStats merged = first;
    merged.ms_total += m_stats.ms_total;
    merged.candidates += m_stats.candidates;
    merged.ms_host_prepare += m_stats.ms_host_prepare;
    merged.ms_host_finish += m_stats.ms_host_finish;
    m_stats = merged;
    m_passing = passing;
    return true;
}

}  // namespace NpuRetrieval
