// filter_stats.cpp -- see filter_stats.h.
#include "filter_stats.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <sstream>

#include "data_table.h"

namespace NpuRetrieval {
namespace {

bool EnvOn(const char* name) {
    const char* v = std::getenv(name);
    return v && *v && std::strtol(v, nullptr, 10) > 0;
}

uint64_t EnvEvery(const char* name, uint64_t fallback) {
    const char* v = std::getenv(name);
    const long long n = v ? std::strtoll(v, nullptr, 10) : 0;
    return n > 0 ? static_cast<uint64_t>(n) : fallback;
}

bool Due(uint64_t seen, uint64_t every) { return seen == 1 || seen % every == 0; }

// layouts 0..4 of the posting format, then the all-zero operand
constexpr int kSlots = 6;
const char* const kLayoutNames[kSlots] = {"denseBitmap", "sparseBitmap", "denseBytemap", "sparseIdList",
                                          "sparsePacked", "absent"};

struct BatchTotals {
    std::mutex mu;
    uint64_t batches = 0, queries = 0, matchAll = 0, pairs = 0, payloadBytes = 0, bitmapBytes = 0;
    uint64_t byLayout[kSlots] = {};
    uint64_t orPairs = 0, orAbsent = 0;
} g_batch;

struct ExprTotals {
    std::mutex mu;
    uint64_t exprs = 0;
    uint64_t nodes[4] = {}, postings[4] = {}, stacks[4] = {};  // and, or, not, conj
    uint64_t andOrHist[8] = {};                                // operands 1, 2, 3, 4, 5-8, 9-16, 17-32, >32
    uint64_t maxOperands = 0, maxConjParts = 0, maxStack = 0;
} g_expr;

int HistBucket(uint64_t n) {
    if (n <= 4) return n == 0 ? 0 : static_cast<int>(n) - 1;
    if (n <= 8) return 4;
    if (n <= 16) return 5;
    if (n <= 32) return 6;
    return 7;
}

// Which posting operands of a program are read by an OR (the only ops the reference's sparse-direct
// path serves, and where an absent posting costs a pass over nothing).
std::vector<bool> OrOperands(const std::vector<uint32_t>& e, size_t nPostings) {
    std::vector<bool> isOr(nPostings, false);
    size_t pc = 0;
    for (size_t i = 0; i < e.size();) {
        const uint32_t op = e[i];
        if (op == CONJUNCTION) {
            const uint32_t parts = e[i + 1];
            for (uint32_t x = 0; x < parts; x++) pc += e[i + 2 + x];
            i += 2 + parts;
        } else {
            const uint32_t pn = e[i + 1];
            for (uint32_t x = 0; x < pn && pc + x < nPostings; x++) isOr[pc + x] = op == OR;
            pc += pn;
            i += 3;
        }
    }
    return isOr;
}

}  // namespace

bool FilterStatsEnabled() {
    static const bool on = EnvOn("FRK_FILTER_STATS");
    return on;
}
bool EmptyStatsEnabled() {
    static const bool on = EnvOn("FRK_EMPTY_STATS");
    return on;
}
bool ExprStatsEnabled() {
    static const bool on = EnvOn("FRK_EXPR_STATS");
    return on;
}

void NoteFilterBatch(const DataTable& table, const std::vector<const FilterExpr*>& exprs, uint32_t segments,
                     uint64_t bitmapBytes) {
    uint64_t byLayout[kSlots] = {}, pairs = 0, payload = 0, matchAll = 0, orPairs = 0, orAbsent = 0;
    for (const FilterExpr* fe : exprs) {
        if (fe->emptyTree) {
            matchAll++;
            continue;
        }
        const std::vector<bool> isOr = OrOperands(fe->postExpr, fe->postingDeviceAddrs.size());
        for (size_t p = 0; p < fe->postingDeviceAddrs.size(); p++) {
            const std::vector<uint64_t>& addrs = *fe->postingDeviceAddrs[p];
            for (uint32_t s = 0; s < segments && s < addrs.size(); s++) {
                pairs++;
                uint8_t layout = 0;
                uint32_t bytes = 0;
                const bool present = addrs[s] != 0 && table.DescribePosting(addrs[s], layout, bytes) && layout < 5;
                byLayout[present ? layout : 5]++;
                payload += present ? 8ull + bytes : 0ull;
                if (isOr[p]) {
                    orPairs++;
                    orAbsent += present ? 0 : 1;
                }
            }
        }
    }
    std::lock_guard<std::mutex> lock(g_batch.mu);
    BatchTotals& t = g_batch;
    t.batches++;
    t.queries += exprs.size();
    t.matchAll += matchAll;
    t.pairs += pairs;
    t.payloadBytes += payload;
    t.bitmapBytes += bitmapBytes;
    t.orPairs += orPairs;
    t.orAbsent += orAbsent;
    for (int i = 0; i < kSlots; i++) t.byLayout[i] += byLayout[i];
    const double b = static_cast<double>(t.batches);
    if (FilterStatsEnabled() && Due(t.batches, EnvEvery("FRK_FILTER_STATS_EVERY", 1000))) {
        std::ostringstream o;
        o << "[BITSTAT] batches=" << t.batches << " queries=" << t.queries << " matchAll=" << t.matchAll
          << " segments=" << segments << " operandPairsPerBatch=" << t.pairs / b;
        for (int i = 0; i < kSlots; i++) o << " " << kLayoutNames[i] << "=" << t.byLayout[i];
        o << " postingMBPerBatch=" << t.payloadBytes / b / 1048576.0 << " bitmapMBPerBatch=" << t.bitmapBytes / b / 1048576.0;
        std::printf("%s\n", o.str().c_str());
        std::fflush(stdout);
    }
    if (EmptyStatsEnabled() && Due(t.batches, EnvEvery("FRK_EMPTY_STATS_EVERY", 1000))) {
        std::ostringstream o;
        o << "[EMPTYSTAT] batches=" << t.batches << " pairs=" << t.pairs << " absent=" << t.byLayout[5]
          << " orPairs=" << t.orPairs << " orAbsent=" << t.orAbsent;
        if (t.pairs) o << " absentOfAll=" << 100.0 * t.byLayout[5] / t.pairs << "%";
        if (t.orPairs) o << " absentOfOr=" << 100.0 * t.orAbsent / t.orPairs << "%";
        std::printf("%s\n", o.str().c_str());
        std::fflush(stdout);
    }
}

void NoteExprShape(const std::vector<uint32_t>& e) {
    std::lock_guard<std::mutex> lock(g_expr.mu);
    ExprTotals& t = g_expr;
    const uint64_t depth = PostExprMaxStack(e);
    t.maxStack = depth > t.maxStack ? depth : t.maxStack;
    for (size_t i = 0; i < e.size();) {
        const uint32_t op = e[i];
        if (op == CONJUNCTION) {
            const uint32_t parts = e[i + 1];
            uint64_t sum = 0;
            for (uint32_t x = 0; x < parts; x++) sum += e[i + 2 + x];
            t.nodes[3]++;
            t.postings[3] += sum;
            t.maxConjParts = parts > t.maxConjParts ? parts : t.maxConjParts;
            t.maxOperands = sum > t.maxOperands ? sum : t.maxOperands;
            i += 2 + parts;
            continue;
        }
        const int slot = op == AND ? 0 : op == OR ? 1 : 2;
        const uint64_t pn = e[i + 1], sn = e[i + 2];
        t.nodes[slot]++;
        t.postings[slot] += pn;
        t.stacks[slot] += sn;
        if (op != NOT) t.andOrHist[HistBucket(pn + sn)]++;
        t.maxOperands = pn + sn > t.maxOperands ? pn + sn : t.maxOperands;
        i += 3;
    }
    t.exprs++;
    if (!Due(t.exprs, EnvEvery("FRK_EXPR_STATS_EVERY", 5000))) return;
    static const char* const kNames[4] = {"and", "or", "not", "conj"};
    std::ostringstream o;
    o << "[EXPRSTAT] exprs=" << t.exprs;
    for (int i = 0; i < 4; i++)
        o << " " << kNames[i] << "={nodes=" << t.nodes[i] << ",postings=" << t.postings[i] << ",stacks=" << t.stacks[i]
          << "}";
    o << " maxOperands=" << t.maxOperands << " maxConjParts=" << t.maxConjParts << " maxStack=" << t.maxStack
      << " andOrOperands(1,2,3,4,5-8,9-16,17-32,>32)=[";
    for (int b = 0; b < 8; b++) o << (b ? "," : "") << t.andOrHist[b];
    o << "]";
    std::printf("%s\n", o.str().c_str());
    std::fflush(stdout);
}

}  // namespace NpuRetrieval
