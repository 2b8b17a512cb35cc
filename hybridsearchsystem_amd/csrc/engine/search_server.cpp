// search_server.cpp -- continuous-batching request server (see search_server.h).
#include "search_server.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>

#include "log.h"

namespace NpuRetrieval {

SearchServer::SearchServer(std::shared_ptr<DataTable> table, int32_t deviceId, const ServerConfig& cfg)
    : m_table(std::move(table)), m_deviceId(deviceId), m_cfg(cfg), m_t0(std::chrono::steady_clock::now()) {
    VectorFieldData* vf = nullptr;
    if (!m_table || !m_table->GetVectorFieldData(m_cfg.vectorField, vf) || !vf) {
        FRK_LOG_ERROR("server: no vector field " << m_cfg.vectorField);
    }
    m_vf = vf;
    // an inert token: absent from the posting field, so a term on it matches no doc
    PostingFieldData* pf = nullptr;
    if (m_table && m_table->GetPostingFieldData(m_cfg.postingField, pf) && pf) {
        for (uint64_t tok = 0xFFFFFFFFFFFFFFFFull;; tok--) {
            bool exist = true;
            (void)pf->GetPostingDeviceAddrs(tok, exist);
            if (!exist) {
                m_inertToken = tok;
                break;
            }
        }
    } else {
        m_cfg.padQueries = 1;
    }
    m_cfg.padQueries = std::max<uint32_t>(1, m_cfg.padQueries);
    m_inertTrees.resize(m_cfg.padQueries);
    for (auto& t : m_inertTrees) t = std::make_unique<TermNode>(m_cfg.postingField, m_inertToken);
    m_cfg.workers = std::max<uint32_t>(1, m_cfg.workers);
    m_cfg.maxBatchQueries = std::max<uint32_t>(1, m_cfg.maxBatchQueries);
    // workspaces are grow-only and pooled: without a retention limit, idle ones would keep their
    // high-water memory on top of what the budget admits for the batches in flight
    if (m_cfg.hbmBudgetBytes) SetWorkspaceRetainLimit(m_cfg.hbmBudgetBytes);
    m_slot.resize(m_cfg.workers);
    m_workerCv = std::vector<std::condition_variable>(m_cfg.workers);
    m_scheduler = std::thread([this] { SchedulerLoop(); });
    for (uint32_t w = 0; w < m_cfg.workers; w++) m_workers.emplace_back([this, w] { WorkerLoop(w); });
}

SearchServer::~SearchServer() {
    Drain();
    {
        std::lock_guard<std::mutex> g(m_mu);
        m_stop = true;
    }
    m_cv.notify_all();
    for (auto& c : m_workerCv) c.notify_all();
    if (m_scheduler.joinable()) m_scheduler.join();
    for (auto& t : m_workers)
        if (t.joinable()) t.join();
    if (m_cfg.hbmBudgetBytes) SetWorkspaceRetainLimit(0);  // the limit was this server's
}

double SearchServer::Now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - m_t0).count();
}

uint64_t SearchServer::BatchBytes(uint32_t queries) const {
    // the searcher's per-batch buffers: query-major filter words, block-major masks, candidate lists
    // (up to 256 k keys per query with a learned floor), the k*m / rescored / output keys (up to
    // 4,096 per query)
    const uint64_t n = m_table ? m_table->GetDocNum() : 0;
    const uint64_t q = queries, q32 = (q + 31) / 32 * 32;
    const uint64_t segWords = m_table ? static_cast<uint64_t>(m_table->GetSegmentNum()) *
                                            (m_table->GetDocNumPerSegment() / 64)
                                      : 0;
    const uint64_t cand = std::min<uint64_t>(std::max<uint64_t>(n, 1), 256ull * 4096);
    return q * segWords * 8 + q32 * ((n + 31) / 32) * 4 + q * cand * 8 + 3 * q * 4096 * 8;
}

uint64_t SearchServer::Submit(uint32_t n, const float* queries, uint32_t dim, const std::vector<std::string>& filters,
                              const std::vector<uint32_t>& ks) {
    if (n == 0 || !queries || ks.size() != n || (!filters.empty() && filters.size() != n)) {
        FRK_LOG_ERROR("server: malformed request (" << n << " queries, " << filters.size() << " filters, " << ks.size()
                                                    << " k)");
        return 0;
    }
    auto r = std::make_shared<Request>();
    r->n = n;
    r->dim = dim;
    r->queries.assign(queries, queries + static_cast<size_t>(n) * dim);
    r->ks = ks;
    r->trees.resize(n);
    for (uint32_t i = 0; i < n; i++) {
        std::unique_ptr<FilterAst> ast;
        if (!filters.empty() && !filters[i].empty()) {
            bool ok = true;
            ast = ParseFilter(filters[i], ok);
            if (!ok) {
                FRK_LOG_ERROR("server: filter " << i << " does not parse: " << filters[i]);
                return 0;
            }
        }
        r->trees[i] = BuildQueryNode(ast.get(), m_cfg.postingField);
    }
    RequestResult& o = r->result;
    o.ok = true;  // until a batch holding some of its queries fails
    o.counts.assign(n, 0);
    o.ids.assign(n, {});
    o.scores.assign(n, {});
    o.passing.assign(n, 0);
    o.started = -1.0;
    std::lock_guard<std::mutex> g(m_mu);
    r->ticket = m_nextTicket++;
    o.submitted = Now();
    m_tickets[r->ticket] = r;
    m_pending.push_back(r);
    m_cv.notify_all();
    return r->ticket;
}

bool SearchServer::Wait(uint64_t ticket, RequestResult& out) {
    std::unique_lock<std::mutex> lk(m_mu);
    auto it = m_tickets.find(ticket);
    if (it == m_tickets.end()) {
        FRK_LOG_ERROR("server: unknown ticket " << ticket);
        return false;
    }
    std::shared_ptr<Request> r = it->second;
    m_doneCv.wait(lk, [&] { return r->done; });
    out = std::move(r->result);
    m_tickets.erase(ticket);
    return out.ok;
}

void SearchServer::Drain() {
    std::unique_lock<std::mutex> lk(m_mu);
    m_doneCv.wait(lk, [&] { return m_pending.empty() && m_running == 0; });
}

ServerStats SearchServer::Stats() const {
    std::lock_guard<std::mutex> g(m_mu);
    return m_stats;
}

void SearchServer::SchedulerLoop() {
    std::unique_lock<std::mutex> lk(m_mu);
    for (;;) {
        int free = -1;
        m_cv.wait(lk, [&] {
            free = -1;
            for (uint32_t w = 0; w < m_slot.size(); w++)
                if (!m_slot[w]) {
                    free = static_cast<int>(w);
                    break;
                }
            return (m_stop && m_pending.empty()) || (!m_pending.empty() && free >= 0);
        });
        if (m_stop && m_pending.empty()) return;
        // the oldest pending queries that fit one batch's query slots: whole requests (one larger
        // than a batch runs alone), or with splitRequests the remainder of the oldest request first
        // and the last request cut at the batch's end ...
        std::vector<Piece> pieces;
        uint32_t q = 0;
        for (const auto& r : m_pending) {
            const uint32_t left = r->n - r->scheduled;
            uint32_t take = left;
            if (q + left > m_cfg.maxBatchQueries) {
                if (!m_cfg.splitRequests && !pieces.empty()) break;
                if (m_cfg.splitRequests) take = m_cfg.maxBatchQueries - q;
            }
            pieces.push_back(Piece{r, r->scheduled, take});
            q += take;
            if (q >= m_cfg.maxBatchQueries) break;
        }
        // ... and the HBM left by the batches in flight (one batch always runs on an idle device)
        uint64_t bytes = BatchBytes(q);
        while (m_cfg.hbmBudgetBytes && pieces.size() > 1 && m_bytesInFlight + bytes > m_cfg.hbmBudgetBytes) {
            q -= pieces.back().count;
            pieces.pop_back();
            bytes = BatchBytes(q);
        }
        if (m_cfg.hbmBudgetBytes && m_bytesInFlight + bytes > m_cfg.hbmBudgetBytes && m_running > 0) {
            m_cv.wait(lk, [&] { return m_stop || m_running == 0 || m_bytesInFlight + bytes <= m_cfg.hbmBudgetBytes; });
            continue;
        }
        auto b = std::make_unique<Batch>();
        for (const Piece& p : pieces) {
            p.req->scheduled += p.count;
            p.req->inFlight++;
            // pieces come from the front in order; a request leaves the queue once all of it is taken
            if (p.req->scheduled == p.req->n) m_pending.pop_front();
        }
        b->pieces = std::move(pieces);
        b->queries = q;
        b->bytes = bytes;
        m_bytesInFlight += bytes;
        m_running++;
        m_slot[free] = std::move(b);
        m_workerCv[free].notify_one();
    }
}

void SearchServer::WorkerLoop(uint32_t w) {
    (void)hipSetDevice(m_deviceId);
    std::unique_lock<std::mutex> lk(m_mu);
    for (;;) {
        m_workerCv[w].wait(lk, [&] { return m_stop || m_slot[w] != nullptr; });
        if (!m_slot[w]) return;  // stopping with nothing to run
        Batch* b = m_slot[w].get();
        lk.unlock();
        RunBatch(*b);
        lk.lock();
        m_bytesInFlight -= b->bytes;
        m_running--;
        for (const Piece& p : b->pieces) {
            Request& r = *p.req;
            if (--r.inFlight == 0 && r.scheduled == r.n) {
                r.done = true;
                m_stats.requests++;
            }
        }
        m_slot[w].reset();
        m_doneCv.notify_all();
        m_cv.notify_all();
    }
}

void SearchServer::RunBatch(Batch& b) {
    const double started = Now();
    FullRecallSearcher s(m_deviceId, m_table);
    s.SetOptions(m_cfg.options);
    uint32_t dim = 0;
    for (const Piece& p : b.pieces) {
        const Request& r = *p.req;
        dim = r.dim;
        for (uint32_t i = p.first; i < p.first + p.count; i++)
            s.AddQuery(r.trees[i].get(), &r.queries[static_cast<size_t>(i) * r.dim], r.dim, r.ks[i]);
    }
    // padding: zero vectors with k = 1 over an inert term (no doc passes, no survivors, no floor)
    const uint32_t pad = (m_cfg.padQueries - b.queries % m_cfg.padQueries) % m_cfg.padQueries;
    const std::vector<float> zeros(dim, 0.0f);
    for (uint32_t i = 0; i < pad; i++) s.AddQuery(m_inertTrees[i].get(), zeros.data(), dim, 1);
    std::vector<std::shared_ptr<FullRecallResult>> res;
    const bool ok =
        m_vf != nullptr && s.BatchSearch(res, m_cfg.vectorField, false) && res.size() == b.queries + pad;
    const double finished = Now();
    const BatchStats st = s.LastStats();
    const std::vector<uint32_t>& passing = s.LastPassing();
    // this batch's queries of each request (entries no other batch writes; the vectors were sized at
    // submission)
    size_t at = 0;
    for (const Piece& p : b.pieces) {
        RequestResult& o = p.req->result;
        for (uint32_t i = p.first; i < p.first + p.count && ok; i++, at++) {
            o.ids[i] = res[at]->docIds;
            o.scores[i] = res[at]->scores;
            o.counts[i] = static_cast<uint32_t>(res[at]->docIds.size());
            o.passing[i] = at < passing.size() ? passing[at] : 0;
        }
    }
    uint64_t docBytes = 0;
    if (const frk_doc_store* store = m_vf ? m_vf->Store(m_cfg.options.dtype) : nullptr) {
        const uint64_t elem = store->dtype == FRK_DTYPE_BF16 ? 2 : store->dtype == FRK_DTYPE_FP8E4M3 ? 1 : 4;
        docBytes = (store->n_docs + 31) / 32 * 32 * store->dim * elem * (st.reruns ? 2 : 1);
    }
    std::lock_guard<std::mutex> g(m_mu);
    for (const Piece& p : b.pieces) {
        RequestResult& o = p.req->result;
        o.ok = o.ok && ok;
        o.started = o.started < 0 ? started : std::min(o.started, started);
        if (finished >= o.finished) {
            o.finished = finished;
            o.batchQueries = b.queries;
        }
    }
    m_stats.queries += b.queries;
    m_stats.batches++;
    m_stats.graphReplays += st.graphReplayed ? 1 : 0;
    m_stats.reruns += st.reruns;
    m_stats.overflows += st.overflowed;
    m_stats.deviceMs += st.ms_total;
    m_stats.docBytesStreamed += docBytes;
}

}  // namespace NpuRetrieval
