// filter_stats.h -- the filter stage's optional instrumentation, switched on by environment variables
// and off the query path otherwise (the reference's TextFilter statistics, engine/retrieval/searcher/
// stage/filter_stats.h:1-56). Reported per batch of the fused pipeline:
//   FRK_FILTER_STATS=1  [BITSTAT]   operand (posting, segment) pairs of the batch's programs by on-disk
//                                   layout, payload bytes the filter kernel reads and bitmap bytes it
//                                   writes; every FRK_FILTER_STATS_EVERY batches (default 1000).
//   FRK_EMPTY_STATS=1   [EMPTYSTAT] pairs whose posting is absent from its segment (the all-zero
//                                   operand), overall and among OR operands; every FRK_EMPTY_STATS_EVERY.
//   FRK_EXPR_STATS=1    [EXPRSTAT]  shape of the postfix programs: per op type nodes, posting operands
//                                   and stack operands, the widest AND/OR, the largest CONJ, stack depth;
//                                   every FRK_EXPR_STATS_EVERY programs.
// The first report prints at once; counters are process-wide and thread-safe.
#pragma once

#include <cstdint>
#include <vector>

#include "query_node.h"

namespace NpuRetrieval {

class DataTable;

bool FilterStatsEnabled();
bool EmptyStatsEnabled();
bool ExprStatsEnabled();

// One batch's compiled filters (match-all queries have no program).
void NoteFilterBatch(const DataTable& table, const std::vector<const FilterExpr*>& exprs, uint32_t segments,
                     uint64_t bitmapBytes);
void NoteExprShape(const std::vector<uint32_t>& postExpr);

}  // namespace NpuRetrieval
