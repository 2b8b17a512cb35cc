// ref_shim.cpp -- C entry points over the reference's own host sources, compiled from where they lie
// in the reference checkout into oracle/_ref/ (see oracle/Makefile). TEST INFRASTRUCTURE ONLY: the
// parity tests use it to pin the CPU oracle (frk_oracle.c) against the reference's code, and the
// bench's CPU-baseline leg times the reference's own CpuTopKBatch through it.
//
// The reference sources linked in:
//   engine/format/posting_serializer.cpp, engine/core/hash/hash.cpp, platform/murmurhash/MurmurHash3.cpp,
//   engine/core/number/trans_number.cpp, 10.13/load/src/harness/recall/recall_ref.cpp
// and the header-only harness/query/filter_expr.h and 10.13/load/src/harness/query/filter_flat.h.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include "src/full_recall/core/hash/hash.h"
#include "src/full_recall/core/number/trans_number.h"
#include "src/full_recall/format/posting_serializer.h"
#include "10.13/load/src/harness/query/filter_expr.h"
#include "10.13/load/src/harness/query/filter_flat.h"
#include "10.13/load/src/harness/recall/recall_ref.h"

namespace {
using npur_harness::FilterNode;
using npur_harness::FilterOp;

// Our flat pre-order (type, arg) pairs: 0 TERM(tag), 1 AND(n), 2 OR(n), 3 NOT(1).
std::unique_ptr<FilterNode> FromPairs(const int64_t* p, int n, int& at) {
    auto node = std::make_unique<FilterNode>();
    int64_t type = p[2 * at], arg = p[2 * at + 1];
    at++;
    if (type == 0) {
        node->op = FilterOp::Term;
        node->tag = static_cast<uint64_t>(arg);
        return node;
    }
    node->op = type == 1 ? FilterOp::And : (type == 2 ? FilterOp::Or : FilterOp::Not);
    int64_t kids = type == 3 ? 1 : arg;
    for (int64_t c = 0; c < kids && at < n; c++) node->children.push_back(FromPairs(p, n, at));
    return node;
}

std::unique_ptr<FilterNode> BuildTree(const int64_t* p, int n) {
    if (n == 0) return nullptr;
    int at = 0;
    return FromPairs(p, n, at);
}

void ToPairs(const FilterNode* node, std::vector<int64_t>& out) {
    switch (node->op) {
        case FilterOp::Term:
            out.push_back(0);
            out.push_back(static_cast<int64_t>(node->tag));
            return;
        case FilterOp::And:
        case FilterOp::Or:
            out.push_back(node->op == FilterOp::And ? 1 : 2);
            out.push_back(static_cast<int64_t>(node->children.size()));
            break;
        case FilterOp::Not:
            out.push_back(3);
            out.push_back(1);
            break;
    }
    for (const auto& c : node->children) ToPairs(c.get(), out);
}
}  // namespace

extern "C" {

// NPUR_SPARSE_PACKED is read once per process by the serializer (posting_serializer.cpp:58-61), so
// the packed and unpacked builds are two copies of this library, each initialised before use.
void ref_init(int packed) {
    setenv("NPUR_SPARSE_PACKED", packed ? "1" : "0", 1);
    setenv("NPUR_LOG_LEVEL", "ERROR", 0);
    // one sparse encode latches the switch in this copy while the environment says what it should
    std::unordered_set<uint32_t> one{0};
    std::string buf;
    NpuRetrieval::PostingSerializer().Serialize(one, 1000, false, 0.05f, buf);
}

uint64_t ref_hash64(const char* data, uint32_t len) { return NpuRetrieval::Hash64(data, len); }

uint16_t ref_f32_to_f16(float f) { return NpuRetrieval::Float32ToFloat16(f); }

float ref_f16_to_f32(uint16_t h) { return NpuRetrieval::Float16ToFloat32(h); }

int64_t ref_posting_serialize(const uint32_t* ids, uint32_t n_ids, uint32_t doc_num, int is_matrix, float threshold,
                              uint8_t* out, uint64_t cap) {
    std::unordered_set<uint32_t> set(ids, ids + n_ids);
    std::string buf;
    NpuRetrieval::PostingSerializer ser;
    if (!ser.Serialize(set, doc_num, is_matrix != 0, threshold, buf)) return -1;
    if (buf.size() > cap) return -2;
    std::memcpy(out, buf.data(), buf.size());
    return static_cast<int64_t>(buf.size());
}

// ParseFilter (filter_expr.h:169-183); returns the pair count, 0 for match-all, -1 on a parse error.
int ref_parse_filter(const char* s, int64_t* out, int cap_pairs) {
    bool ok = true;
    auto tree = npur_harness::ParseFilter(s, ok);
    if (!ok) return -1;
    if (!tree) return 0;
    std::vector<int64_t> pairs;
    ToPairs(tree.get(), pairs);
    int n = static_cast<int>(pairs.size() / 2);
    if (n > cap_pairs) return -2;
    std::memcpy(out, pairs.data(), pairs.size() * sizeof(int64_t));
    return n;
}

// EvalFilterBitmap (filter_expr.h:212-232) over n_docs doc tag-bitmaps of `stride` words.
void ref_filter_bitmap(const int64_t* ast, int n_pairs, const uint64_t* doc_bitmaps, uint64_t n_docs, uint32_t stride,
                       uint64_t* out_words) {
    auto tree = BuildTree(ast, n_pairs);
    std::memset(out_words, 0, ((n_docs + 63) / 64) * 8);
    for (uint64_t d = 0; d < n_docs; d++)
        if (npur_harness::EvalFilterBitmap(tree.get(), doc_bitmaps + d * stride, stride))
            out_words[d / 64] |= 1ULL << (d % 64);
}

// FlattenFilter + EvalFlatFilter (10.13 filter_flat.h:36-140), the evaluator CpuTopKBatch uses.
void ref_flat_filter_bitmap(const int64_t* ast, int n_pairs, const uint64_t* doc_bitmaps, uint64_t n_docs,
                            uint32_t stride, uint64_t* out_words) {
    auto tree = BuildTree(ast, n_pairs);
    npur_harness::FlatFilter flat = npur_harness::FlattenFilter(tree.get(), stride);
    std::memset(out_words, 0, ((n_docs + 63) / 64) * 8);
    for (uint64_t d = 0; d < n_docs; d++)
        if (npur_harness::EvalFlatFilter(flat, doc_bitmaps + d * stride)) out_words[d / 64] |= 1ULL << (d % 64);
}

// CpuTopKBatch (10.13 recall_ref.cpp:90-140) over docs [doc_offset, doc_offset + n_scan). The
// per-query filters are given as concatenated pair arrays (filter_off[q], filter_len[q]).
// out_ids: nq x k_cap, unordered id sets padded with -1; out_counts: set sizes.
int ref_cpu_topk_batch(const float* docs, const uint64_t* doc_bitmaps, uint64_t n_docs, uint32_t dim, uint32_t stride,
                       const float* queries, uint32_t nq, const int64_t* filters, const int32_t* filter_off,
                       const int32_t* filter_len, const int32_t* ks, int fp32, uint64_t doc_offset, uint64_t n_scan,
                       int32_t k_cap, int64_t* out_ids, int32_t* out_counts) {
    npur_port::HwDataset hw;
    hw.vectors = docs;
    hw.bitmaps = doc_bitmaps;
    hw.dim = dim;
    hw.stride = stride;
    hw.doc_num = n_docs;
    std::vector<std::vector<float>> qs(nq);
    std::vector<std::unique_ptr<FilterNode>> trees(nq);
    std::vector<npur_harness::RecallRefJob> jobs(nq);
    for (uint32_t q = 0; q < nq; q++) {
        qs[q].assign(queries + static_cast<size_t>(q) * dim, queries + static_cast<size_t>(q + 1) * dim);
        trees[q] = BuildTree(filters + 2 * static_cast<size_t>(filter_off[q]), filter_len[q]);
        jobs[q] = {&qs[q], trees[q].get(), ks[q], q};
    }
    const auto quantized = npur_harness::QuantizeCorpus(hw, n_scan, doc_offset, fp32 != 0, 1);
    std::vector<std::unordered_set<int64_t>> out(nq);
    npur_harness::CpuTopKBatch(jobs, hw, quantized, n_scan, doc_offset, fp32 != 0, out);
    hw.vectors = nullptr;  // not ours: keep the destructor away from the caller's buffers
    hw.bitmaps = nullptr;
    for (uint32_t q = 0; q < nq; q++) {
        int32_t i = 0;
        for (int64_t id : out[q]) {
            if (i < k_cap) out_ids[static_cast<size_t>(q) * k_cap + i] = id;
            i++;
        }
        for (int32_t j = i; j < k_cap; j++) out_ids[static_cast<size_t>(q) * k_cap + j] = -1;
        out_counts[q] = i;
    }
    return 0;
}

}  // extern "C"
