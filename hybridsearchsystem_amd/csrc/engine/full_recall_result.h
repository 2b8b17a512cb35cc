// full_recall_result.h -- result and device-block types of the searcher's interface, with the
// reference's names and fields (core/full_recall_result.h:9-51), so callers written against it compile
// unchanged.
#pragma once

#include <atomic>
#include <cstdint>
#include <memory>
#include <vector>

namespace NpuRetrieval {

// The device memory a GmBlock came from. The reference keeps one pool per name
// (searcher/runtime/gm_memory_manager.cpp:9-46); here the name records what the block holds, and the
// blocks the unfused path hands out (OneFilter, BatchScore) are plain device allocations freed by
// OneAggregator.
enum class GmPoolName {
    VECTOR_SCORE_RESULT_POOL = 0,
    TEXT_FILTER_RESULT_POOL = 1,
    TEXT_FILTER_STACK_POOL = 2,
    TEXT_FILTER_BITLIST2SET_POOL = 3,
    AGGREGATOR_POOL = 4,
    AGGREGATOR_SCRATCH_POOL = 5,
    TEXT_FILTER_POSTINGS_POOL = 6,
    SCORER_QUERY_POOL = 7,
    TEXT_FILTER_EXPR_POOL = 8,
    VALID_POOL_NUM,
};

struct GmBlock {
    GmPoolName poolName = GmPoolName::VECTOR_SCORE_RESULT_POOL;
    char* data = nullptr;  // device memory
    uint32_t size = 0;     // bytes
    // readers left before the block may be freed (a score matrix is read by one OneAggregator per query)
    std::shared_ptr<std::atomic<uint32_t>> readCount = nullptr;
};

struct ScoreWithIndex {
    float score;
    uint32_t index;  // local doc id
    uint64_t id{0};  // global doc id
    ScoreWithIndex(float s, uint32_t i) : score(s), index(i) {}
};

// Single shard: docIds (gdocid) and scores, best first, min(k, passing docs) long. Multi-shard:
// scoreWithIndexes {score, ldocid, gdocid} in the same order (result_aggregator.cpp:846-1032).
struct FullRecallResult {
    bool isMultiShard = false;
    std::vector<ScoreWithIndex> scoreWithIndexes;
    std::vector<float> scores;
    std::vector<int64_t> docIds;
    uint32_t scoreResultIndex = 0;                 // kept for interface compatibility; unused here
    std::shared_ptr<GmBlock> resultChunk = nullptr;  // kept for interface compatibility; unused here
};

}  // namespace NpuRetrieval
