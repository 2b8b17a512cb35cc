// full_recall_searcher.h -- the per-batch filtered top-K search with the reference's interface
// (retrieval/searcher/full_recall_searcher.h:12-70, core/full_recall_result.h:38-51).
//
// One batch is one stream-ordered device pipeline with a single synchronisation at the end:
//   H2D of the packed batch (programs, operand tables, queries, k) -> filter (all queries) ->
//   [probe: score a strided sample of doc blocks, k-th best per query = the score floor, or the
//   learned floor model] -> fused score + mask + floor + compaction over all docs -> exact top-k
//   select -> [low precision: the top k*m as a set, FP32 rescore, sorted select] -> D2H of the
//   top-k keys.
// The reference instead syncs after the scorer, per query after the filter and per query after
// each aggregation step (full_recall_searcher.cpp:107-231, result_aggregator.cpp:363-777).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "data_table.h"
#include "full_recall_result.h"
#include "query_node.h"

namespace NpuRetrieval {

// What BatchSearchDevice hands to BatchSearchExtract for one query: the sorted top-k keys.
struct AggrDeviceResult {
    std::vector<uint64_t> keys;
    uint32_t passing = 0;  // docs that passed the filter
};

// Scoring precision and pruning for a batch.
struct SearchOptions {
    uint32_t dtype = FRK_DTYPE_F32;  // store scored on the main pass
    uint32_t rescoreFactor = 4;      // low precision: FP32 rescoring of the top min(k * m, 4096) (m = this)
    bool probe = true;               // sampled k-th best score as the pruning floor (exact)
    uint32_t probeStep = 0;          // sample every probeStep-th doc block; 0 = chosen from k and N
    const float* floors = nullptr;   // optional caller floor per query (e.g. a learned bound); the
                                     // batch reruns unpruned for any query left with < k survivors
    uint32_t candCap = 0;            // candidate capacity per query; 0 = automatic
    const frk_floor_model* floorModel = nullptr;  // learned floors, predicted on the device per batch
                                                  // (same rerun rule as caller floors)
    bool graph = false;  // replay the batch's device sequence from a captured hipGraph per batch shape
};

struct BatchStats {
    float ms_filter = 0, ms_probe = 0, ms_score = 0, ms_select = 0, ms_rescore = 0, ms_total = 0;
    uint64_t candidates = 0;     // survivors written on the main pass
    uint32_t reruns = 0;         // queries rerun (overflow or an unsafe caller floor)
    uint32_t overflowed = 0;
    uint32_t probeStep = 0;
    uint32_t cap = 0;
    bool graphReplayed = false;  // the device sequence ran as a graph replay (stage times then 0)
    float ms_host_prepare = 0;   // host: from the batch's start to its device sequence enqueued
    float ms_host_finish = 0;    // host: from the device sync to the per-query results
};

// The device memory the searchers' workspace pool holds (grow-only per-batch buffers, one stream per
// workspace; the reference's named GM pools, searcher/runtime/gm_memory_manager.cpp:9-75).
struct WorkspacePoolStats {
    uint64_t liveBytes = 0, peakBytes = 0, mallocs = 0, frees = 0, retainLimit = 0;
    uint32_t workspaces = 0;  // alive (in use or pooled)
    uint32_t pooled = 0;      // idle in the pool
};
WorkspacePoolStats GetWorkspacePoolStats();
// A workspace released while the pool holds more than `bytes` of device memory is freed instead of
// pooled (0 = no limit).
void SetWorkspaceRetainLimit(uint64_t bytes);
// One [POOLSTAT] line on stdout (gm_memory_pool.h:172-180 reports its pools the same way).
void ReportWorkspacePool(const char* why);

class SearchWorkspace;

class FullRecallSearcher {
   public:
    FullRecallSearcher(int32_t deviceId, std::shared_ptr<DataTable> dataTable);
    ~FullRecallSearcher();

    void AddQuery(QueryNode* queryTree, const std::vector<float>& query, uint32_t topK);
    void AddQuery(QueryNode* queryTree, const float* query, uint32_t dim, uint32_t topK);
    void SetOptions(const SearchOptions& o) { m_options = o; }
    const BatchStats& LastStats() const { return m_stats; }
    void Clear();

    bool BatchSearch(std::vector<std::shared_ptr<FullRecallResult>>& results, const std::string& vectorFieldName,
                     bool isMultiShard);
    bool BatchSearchDevice(const std::string& vectorFieldName, bool isMultiShard,
                           std::vector<AggrDeviceResult>& deviceResults);
    bool BatchSearchExtract(std::vector<AggrDeviceResult>& deviceResults,
                            std::vector<std::shared_ptr<FullRecallResult>>& results, bool isMultiShard);

    // The unfused path of the reference (full_recall_searcher.h:43-47): OneFilter evaluates the one
    // query this searcher holds into a device bitmap block (segments x docNumPerSegment / 8 bytes);
    // BatchScore writes the [queries][GetScoreExtendDocNum()] FP32 score matrix of every query added;
    // OneAggregator compacts row resultInDeviceOffset of a score matrix by a filter block and selects
    // the top-k of this searcher's first query, then frees the filter block and, after the last
    // reader, the score matrix. The fused BatchSearch never takes this path.
    bool OneFilter(GmBlock& filterResultChunk);
    bool BatchScore(const std::string& vectorFieldName, GmBlock& resultChunk);
    bool OneAggregator(GmBlock& resultChunk, GmBlock& filterResultChunk, uint32_t resultInDeviceOffset,
                       bool isMultiShard, std::vector<std::shared_ptr<FullRecallResult>>& fullRecallResults);

    // Per-query filter pass counts of the last batch (for diagnostics and the learned bound).
    const std::vector<uint32_t>& LastPassing() const { return m_passing; }

   private:
    bool RunDevice(const VectorFieldData& vf, std::vector<AggrDeviceResult>& out);
    // One query's keys (best first) as a result in either mode (result_aggregator.cpp:846-1032).
    bool FillResult(const std::vector<uint64_t>& keys, bool isMultiShard, FullRecallResult& r) const;

    int32_t m_deviceId;
    std::shared_ptr<DataTable> m_table;
    std::vector<QueryNode*> m_trees;
    std::vector<float> m_queries;  // [n][dim]
    std::vector<uint32_t> m_topKs;
    uint32_t m_dim = 0;
    SearchOptions m_options;
    BatchStats m_stats;
    std::vector<uint32_t> m_passing;
};

}  // namespace NpuRetrieval
