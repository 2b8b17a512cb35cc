// search_server.h -- continuous-batching request server over FullRecallSearcher.
//
// Extends the search driver's serial mode (harness/run/schedule_serial.cpp:23-51 with --target_qps
// pacing, run_context.cpp:66-69) into a request loop: callers submit requests of any number of
// queries (each with its own filter and k) at any time; a scheduler thread packs pending requests,
// oldest first, into batches of at most maxBatchQueries queries and hands each to the first free
// worker as soon as one is free and the batch's device working set fits the HBM budget left by the
// batches in flight (optionally splitting a request over consecutive batches to fill each one).
// Each worker owns a searcher (its own workspace and stream), so several batches run on the device
// at once; the index stays resident and is shared read-only. Results come back per request with
// submit, start and finish times.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "full_recall_searcher.h"
#include "query_node.h"

namespace NpuRetrieval {

struct ServerConfig {
    std::string vectorField = "vec";
    std::string postingField = "tag";
    SearchOptions options;            // store, rescoring, floors, graph replay
    uint32_t maxBatchQueries = 256;   // queries per device batch
    uint32_t workers = 2;             // batches in flight on the device
    uint64_t hbmBudgetBytes = 0;      // device working memory for all batches in flight; 0 = unlimited
    uint32_t padQueries = 64;         // batches are padded with inert queries to a multiple of this,
                                      // so replayable graph shapes repeat; 1 = no padding
    bool splitRequests = false;       // a request's queries may span consecutive batches, so that under
                                      // load every batch is filled to maxBatchQueries
};

struct RequestResult {
    bool ok = false;
    std::vector<uint32_t> counts;                // per query: results returned
    std::vector<std::vector<int64_t>> ids;       // per query: global doc ids, best first
    std::vector<std::vector<float>> scores;
    std::vector<uint32_t> passing;               // per query: docs that passed the filter
    double submitted = 0, started = 0, finished = 0;  // seconds on the server's steady clock
    uint32_t batchQueries = 0;                   // queries in the device batch that served it (the one
                                                 // that finished last, for a request split over several)
};

struct ServerStats {
    uint64_t requests = 0, queries = 0, batches = 0, graphReplays = 0, reruns = 0, overflows = 0;
    double deviceMs = 0;       // sum of per-batch device time (events around each batch)
    uint64_t docBytesStreamed = 0;  // doc-store bytes the main score passes read
};

class SearchServer {
   public:
    SearchServer(std::shared_ptr<DataTable> table, int32_t deviceId, const ServerConfig& cfg);
    ~SearchServer();  // finishes every submitted request, then stops the threads
    SearchServer(const SearchServer&) = delete;
    SearchServer& operator=(const SearchServer&) = delete;

    // Thread-safe. Returns a ticket (> 0) or 0 if the request is malformed (error logged).
    uint64_t Submit(uint32_t n, const float* queries, uint32_t dim, const std::vector<std::string>& filters,
                    const std::vector<uint32_t>& ks);
    // Blocks until the request is done, moves its result out and forgets the ticket.
    bool Wait(uint64_t ticket, RequestResult& out);
    void Drain();  // blocks until nothing is pending or running
    ServerStats Stats() const;
    double Now() const;  // seconds on the server's clock

    // Device working set of one batch of `queries` queries (the searcher's workspace high-water mark).
    uint64_t BatchBytes(uint32_t queries) const;

   private:
    struct Request {
        uint64_t ticket = 0;
        uint32_t n = 0, dim = 0;
        std::vector<float> queries;
        std::vector<std::unique_ptr<QueryNode>> trees;
        std::vector<uint32_t> ks;
        RequestResult result;    // per-query vectors sized at submission; batches fill their queries
        uint32_t scheduled = 0;  // queries handed to batches so far
        uint32_t inFlight = 0;   // batches holding some of its queries that have not finished
        bool done = false;
    };
    // queries [first, first + count) of one request
    struct Piece {
        std::shared_ptr<Request> req;
        uint32_t first = 0, count = 0;
    };
    struct Batch {
        std::vector<Piece> pieces;
        uint32_t queries = 0;
        uint64_t bytes = 0;
    };
    void SchedulerLoop();
    void WorkerLoop(uint32_t w);
    void RunBatch(Batch& b);

    std::shared_ptr<DataTable> m_table;
    int32_t m_deviceId;
    uint64_t m_inertToken = 0;                 // a token with no postings: padding queries match nothing
    std::vector<std::unique_ptr<QueryNode>> m_inertTrees;
    ServerConfig m_cfg;
    const VectorFieldData* m_vf = nullptr;
    std::chrono::steady_clock::time_point m_t0;

    mutable std::mutex m_mu;
    std::condition_variable m_cv;       // scheduler: new request or a worker / memory freed
    std::condition_variable m_doneCv;   // waiters: a request finished
    std::deque<std::shared_ptr<Request>> m_pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> m_tickets;
    std::vector<std::unique_ptr<Batch>> m_slot;  // per worker: the batch it runs, or null
    std::vector<std::condition_variable> m_workerCv;
    uint64_t m_nextTicket = 1;
    uint64_t m_bytesInFlight = 0;
    uint32_t m_running = 0;
    bool m_stop = false;
    ServerStats m_stats;
    std::thread m_scheduler;
    std::vector<std::thread> m_workers;
};

}  // namespace NpuRetrieval
