// fr_search.cpp -- the search driver of this build: loads one index (or several disjoint shards of one
// corpus) onto one GPU, answers a query file batch by batch through FullRecallSearcher, and reports
// latency percentiles, QPS and recall against a CPU brute force over the source dataset.
//
// It is the counterpart of the reference's fr_search (harness/main.cpp:24-119, run/run_context.cpp,
// run/schedule_serial.cpp:23-51, run/report.cpp:61-206), written against the reference's include paths
// (the forwarding headers under include/ref_paths), so a caller of the reference API compiles
// unchanged. What it keeps from the reference: the input files (.fvecs queries, one filter expression
// per line assigned one-to-one or by seeded draw, an optional top-K file, the HYDSET2 dataset for the
// recall check), the serial schedule with open-loop --target_qps pacing (latency counted from each
// query's arrival), the shard merge by score, and the report lines ([Result] latency / CPU-recall,
// [RecallRef], [POOLSTAT]). What it does not build: multi-card schedules (one GPU here), NPU load
// disturbance, CPU affinity. Shards given with --shard_index_dirs all live on --device_id and are
// searched concurrently (one thread, workspace and stream each).
//
// Extra flags of this build: --dtype f32|f16x2|bf16|fp8 (+ --rescore_factor m), --probe, --graph,
// --split_extract (BatchSearchDevice + BatchSearchExtract instead of BatchSearch), --unfused (the
// reference's unfused order: BatchScore writes the score matrix, then OneFilter + OneAggregator per
// query), --result_file (per-query ids and scores), --pool_stats.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_set>
#include <vector>

#include <hip/hip_runtime.h>

#include "src/full_recall/core/full_recall_result.h"
#include "src/full_recall/core/hash/hash.h"
#include "src/full_recall/index/data_table.h"
#include "src/full_recall/retrieval/query/query_node_imp.h"
#include "src/full_recall/retrieval/searcher/full_recall_searcher.h"

using Clock = std::chrono::steady_clock;
using NpuRetrieval::FilterAst;

namespace {

// ------------------------------------------------------------------------------------------ flags
struct Flags {
    std::map<std::string, std::string> kv;
    std::string Str(const std::string& k, const std::string& d = "") const {
        auto it = kv.find(k);
        return it == kv.end() ? d : it->second;
    }
    long long Int(const std::string& k, long long d) const {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::strtoll(it->second.c_str(), nullptr, 10);
    }
    double Num(const std::string& k, double d) const {
        auto it = kv.find(k);
        return it == kv.end() ? d : std::strtod(it->second.c_str(), nullptr);
    }
    bool Bool(const std::string& k, bool d) const {
        auto it = kv.find(k);
        if (it == kv.end()) return d;
        return it->second != "false" && it->second != "0";
    }
};

// --name=value, --name value, or a bare --name (true)
bool ParseFlags(int argc, char** argv, Flags& f) {
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        if (a.rfind("--", 0) != 0) {
            std::fprintf(stderr, "unexpected argument: %s\n", a.c_str());
            return false;
        }
        a = a.substr(2);
        const size_t eq = a.find('=');
        if (eq != std::string::npos) {
            f.kv[a.substr(0, eq)] = a.substr(eq + 1);
        } else if (i + 1 < argc && std::strncmp(argv[i + 1], "--", 2) != 0) {
            f.kv[a] = argv[++i];
        } else {
            f.kv[a] = "true";
        }
    }
    return true;
}

std::vector<std::string> SplitCsv(const std::string& s) {
    std::vector<std::string> out;
    std::string cur;
    std::istringstream in(s);
    while (std::getline(in, cur, ','))
        if (!cur.empty()) out.push_back(cur);
    return out;
}

// ------------------------------------------------------------------------------------------ inputs
// Query vectors: a headered file ([i32 rows][i32 dim][rows x dim f32]) when the size matches, else
// per-vector fvecs ([i32 dim][dim f32] repeated).
bool LoadQueries(const std::string& path, std::vector<std::vector<float>>& out) {
    std::ifstream in(path, std::ios::binary | std::ios::ate);
    if (!in) return false;
    const long long size = in.tellg();
    in.seekg(0);
    int32_t head[2] = {0, 0};
    if (size < 8 || !in.read(reinterpret_cast<char*>(head), 8)) return false;
    if (head[0] > 0 && head[1] > 0 && head[1] <= 4096 && size == 8 + 4ll * head[0] * head[1]) {
        out.assign(head[0], std::vector<float>(head[1]));
        for (auto& v : out) in.read(reinterpret_cast<char*>(v.data()), 4ll * head[1]);
        return static_cast<bool>(in);
    }
    in.seekg(0);
    int32_t dim = 0;
    while (in.read(reinterpret_cast<char*>(&dim), 4) && dim > 0 && dim <= 4096) {
        std::vector<float> v(dim);
        if (!in.read(reinterpret_cast<char*>(v.data()), 4ll * dim)) break;
        out.push_back(std::move(v));
    }
    return !out.empty();
}

struct Filters {
    std::vector<std::string> text;
    std::vector<std::unique_ptr<FilterAst>> ast;  // null = match all
    std::vector<uint32_t> assign;                 // query -> expression
    const FilterAst* For(size_t q) const { return assign.empty() ? nullptr : ast[assign[q]].get(); }
};

// One expression per non-empty line; as many lines as queries are used one-to-one, otherwise each
// query draws a line with a seeded mt19937_64 (harness/query/filter_setup.cpp:13-52).
bool LoadFilters(const std::string& path, size_t nq, uint64_t seed, Filters& f) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    while (std::getline(in, line)) {
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
        if (!line.empty()) f.text.push_back(line);
    }
    if (f.text.empty()) return false;
    for (size_t i = 0; i < f.text.size(); i++) {
        bool ok = true;
        f.ast.push_back(NpuRetrieval::ParseFilter(f.text[i], ok));
        if (!ok) {
            std::fprintf(stderr, "filter line %zu does not parse: %s\n", i + 1, f.text[i].c_str());
            return false;
        }
    }
    f.assign.resize(nq);
    if (f.text.size() == nq) {
        for (size_t i = 0; i < nq; i++) f.assign[i] = static_cast<uint32_t>(i);
    } else {
        std::mt19937_64 rng(seed);
        std::uniform_int_distribution<size_t> pick(0, f.text.size() - 1);
        for (size_t i = 0; i < nq; i++) f.assign[i] = static_cast<uint32_t>(pick(rng));
    }
    return true;
}

bool LoadTopKs(const std::string& path, std::vector<uint32_t>& ks) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    while (std::getline(in, line)) {
        const char* p = line.c_str();
        while (*p == ' ' || *p == '\t') p++;
        if (!*p || *p == '\r') continue;
        char* end = nullptr;
        const unsigned long v = std::strtoul(p, &end, 10);
        if (end == p || v == 0 || v > 0xFFFFFFFFul) return false;
        ks.push_back(static_cast<uint32_t>(v));
    }
    return !ks.empty();
}

// ------------------------------------------------------------------------------------------ query trees
// The driver's tree builder (harness/query/filter_setup.cpp:93-136): a term is a TermNode on the posting
// field with token Hash64(decimal tag); AND / OR / NOT nodes own their children; no filter is an empty
// AND node (match all).
std::unique_ptr<NpuRetrieval::QueryNode> TreeOf(const FilterAst* n, const std::string& field) {
    using namespace NpuRetrieval;
    if (!n) return std::make_unique<AndNode>();
    switch (n->op) {
        case FilterAst::Term:
            return std::make_unique<TermNode>(field, Hash64(std::to_string(n->tag)));
        case FilterAst::Not: {
            auto node = std::make_unique<NotNode>();
            std::unique_ptr<QueryNode> c = TreeOf(n->children[0].get(), field);
            node->AddChild(c);
            return node;
        }
        case FilterAst::And:
        case FilterAst::Or: {
            std::unique_ptr<QueryNode> node;
            if (n->op == FilterAst::And)
                node = std::make_unique<AndNode>();
            else
                node = std::make_unique<OrNode>();
            for (const auto& ch : n->children) {
                std::unique_ptr<QueryNode> c = TreeOf(ch.get(), field);
                node->AddChild(c);
            }
            return node;
        }
    }
    return std::make_unique<AndNode>();
}

// ------------------------------------------------------------------------------------------ HYDSET2
// [40-byte header: "HYDSET2\0", u32 version, u32 pad, u64 docs, u32 dim, u32 tags, u32, u32]
// [docs x dim f32][docs x ceil(tags / 64) u64 tag bitmaps] (io/dataset_hw.h).
struct Dataset {
    void* map = nullptr;
    size_t len = 0;
    uint64_t docs = 0;
    uint32_t dim = 0, tags = 0, stride = 0;
    const float* vectors = nullptr;
    const uint64_t* bits = nullptr;
    ~Dataset() {
        if (map) munmap(map, len);
    }
    bool Open(const std::string& path) {
        const int fd = ::open(path.c_str(), O_RDONLY);
        if (fd < 0) return false;
        struct stat st;
        if (fstat(fd, &st) != 0 || st.st_size < 40) {
            ::close(fd);
            return false;
        }
        len = static_cast<size_t>(st.st_size);
        map = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
        ::close(fd);
        if (map == MAP_FAILED) {
            map = nullptr;
            return false;
        }
        const uint8_t* b = static_cast<const uint8_t*>(map);
        if (std::memcmp(b, "HYDSET2", 7) != 0) return false;
        std::memcpy(&docs, b + 16, 8);
        std::memcpy(&dim, b + 24, 4);
        std::memcpy(&tags, b + 28, 4);
        stride = (tags + 63) / 64;
        if (40 + docs * dim * 4 + docs * stride * 8 > len) return false;
        vectors = reinterpret_cast<const float*>(b + 40);
        bits = reinterpret_cast<const uint64_t*>(b + 40 + docs * dim * 4);
        return true;
    }
};

bool EvalAst(const FilterAst* n, const uint64_t* words, uint32_t tags) {
    if (!n) return true;
    switch (n->op) {
        case FilterAst::Term:
            return n->tag < tags && ((words[n->tag / 64] >> (n->tag % 64)) & 1ull);
        case FilterAst::Not:
            return !EvalAst(n->children[0].get(), words, tags);
        case FilterAst::And:
            for (const auto& c : n->children)
                if (!EvalAst(c.get(), words, tags)) return false;
            return true;
        case FilterAst::Or:
            for (const auto& c : n->children)
                if (EvalAst(c.get(), words, tags)) return true;
            return false;
    }
    return false;
}

// The CPU truth of one query over docs [offset, offset + n) of the dataset: filter per doc, FP32 dot
// product (or over FP16-rounded vectors, as the reference's FP16 engine stores them), the k best by
// score descending then doc id ascending (10.13 harness/recall/recall_ref.cpp:19-25); absolute ids.
std::unordered_set<int64_t> CpuTruth(const Dataset& ds, const std::vector<float>& q, const FilterAst* f, uint64_t offset,
                                     uint64_t n, uint32_t k, bool fp32) {
    struct Hit {
        float s;
        int64_t id;
    };
    auto worse = [](const Hit& a, const Hit& b) { return a.s > b.s || (a.s == b.s && a.id < b.id); };
    std::vector<Hit> heap;  // a min-heap under "worse": its front is the weakest kept hit
    const uint32_t dim = std::min<uint32_t>(ds.dim, static_cast<uint32_t>(q.size()));
    for (uint64_t d = offset; d < offset + n; d++) {
        if (!EvalAst(f, ds.bits + d * ds.stride, ds.tags)) continue;
        const float* v = ds.vectors + d * ds.dim;
        float s = 0.0f;
        if (fp32) {
            for (uint32_t j = 0; j < dim; j++) s += q[j] * v[j];
        } else {
            for (uint32_t j = 0; j < dim; j++)
                s += NpuRetrieval::Float16ToFloat32(NpuRetrieval::Float32ToFloat16(q[j])) *
                     NpuRetrieval::Float16ToFloat32(NpuRetrieval::Float32ToFloat16(v[j]));
        }
        const Hit h{s, static_cast<int64_t>(d)};
        if (heap.size() < k) {
            heap.push_back(h);
            std::push_heap(heap.begin(), heap.end(), worse);
        } else if (k > 0 && worse(h, heap.front())) {
            std::pop_heap(heap.begin(), heap.end(), worse);
            heap.back() = h;
            std::push_heap(heap.begin(), heap.end(), worse);
        }
    }
    std::unordered_set<int64_t> out;
    for (const Hit& h : heap) out.insert(h.id);
    return out;
}

// ------------------------------------------------------------------------------------------ run
struct Shard {
    std::string dir;
    std::shared_ptr<NpuRetrieval::DataTable> table;
};

struct Run {
    Flags flags;
    int32_t device = 0;
    std::vector<Shard> shards;
    std::vector<std::vector<float>> queries;
    Filters filters;
    bool haveFilters = false;
    std::vector<uint32_t> topks;
    size_t nq = 0, bs = 1;
    NpuRetrieval::SearchOptions opts;
    bool splitExtract = false;
    bool unfused = false;
    std::string vecField, postField;
    std::vector<std::vector<int64_t>> ids;
    std::vector<std::vector<float>> scores;
    std::vector<double> latMs;  // per query, end to end
    double wallMs = 0;
    bool paced = false;
    double rate = 0;
    Clock::time_point t0;

    // One searcher per batch: AddQuery per query, then BatchSearch (or its two halves).
    bool SearchOne(const Shard& sh, size_t start, size_t end, bool multiShard,
                   std::vector<std::shared_ptr<NpuRetrieval::FullRecallResult>>& res) {
        NpuRetrieval::FullRecallSearcher searcher(device, sh.table);
        searcher.SetOptions(opts);
        std::vector<std::unique_ptr<NpuRetrieval::QueryNode>> owned;
        for (size_t q = start; q < end; q++) {
            owned.push_back(TreeOf(haveFilters ? filters.For(q) : nullptr, postField));
            searcher.AddQuery(owned.back().get(), queries[q % queries.size()], topks[q]);
        }
        if (unfused) {
            // score every doc of every query first, then filter and aggregate query by query
            // (FullRecallSearcher::BatchScore / OneFilter / OneAggregator)
            NpuRetrieval::GmBlock matrix;
            if (!searcher.BatchScore(vecField, matrix)) return false;
            res.clear();
            bool ok = true;
            for (size_t q = start; q < end; q++) {
                NpuRetrieval::FullRecallSearcher one(device, sh.table);
                one.SetOptions(opts);
                one.AddQuery(owned[q - start].get(), queries[q % queries.size()], topks[q]);
                NpuRetrieval::GmBlock bits;
                if (ok && one.OneFilter(bits)) {
                    ok = one.OneAggregator(matrix, bits, static_cast<uint32_t>(q - start), multiShard, res);
                } else {
                    ok = false;
                    // still release this query's hold on the matrix, so the last reader frees it
                    if (matrix.readCount && matrix.readCount->fetch_sub(1) == 1 && matrix.data) (void)hipFree(matrix.data);
                }
            }
            return ok;
        }
        if (!splitExtract) return searcher.BatchSearch(res, vecField, multiShard);
        std::vector<NpuRetrieval::AggrDeviceResult> dev;
        return searcher.BatchSearchDevice(vecField, multiShard, dev) && searcher.BatchSearchExtract(dev, res, multiShard);
    }

    bool RunBatch(size_t start, size_t end, bool record) {
        const size_t bn = end - start;
        if (shards.size() == 1) {
            std::vector<std::shared_ptr<NpuRetrieval::FullRecallResult>> res;
            if (!SearchOne(shards[0], start, end, false, res) || res.size() != bn) return false;
            if (record)
                for (size_t j = 0; j < bn; j++) {
                    ids[start + j] = res[j]->docIds;
                    scores[start + j] = res[j]->scores;
                }
            return true;
        }
        // every shard concurrently, each with its own searcher (workspace and stream) on the device
        std::vector<std::vector<std::shared_ptr<NpuRetrieval::FullRecallResult>>> per(shards.size());
        std::vector<char> ok(shards.size(), 0);
        std::vector<std::thread> th;
        for (size_t s = 1; s < shards.size(); s++)
            th.emplace_back([&, s] { ok[s] = SearchOne(shards[s], start, end, true, per[s]) && per[s].size() == bn; });
        ok[0] = SearchOne(shards[0], start, end, true, per[0]) && per[0].size() == bn;
        for (auto& t : th) t.join();
        for (char o : ok)
            if (!o) return false;
        if (!record) return true;
        // k-way merge of the shards' sorted top-k lists (run/run_context.cpp:236-272): by score
        // descending, equal scores by global doc id ascending -- the tie rule of the CPU truth, so the
        // merged list equals what one table over the whole corpus returns.
        std::vector<size_t> pos(shards.size());
        for (size_t j = 0; j < bn; j++) {
            const size_t K = topks[start + j];
            std::fill(pos.begin(), pos.end(), 0);
            std::vector<int64_t>& outIds = ids[start + j];
            std::vector<float>& outSc = scores[start + j];
            outIds.clear();
            outSc.clear();
            while (outIds.size() < K) {
                int best = -1;
                for (size_t s = 0; s < shards.size(); s++) {
                    const auto& sw = per[s][j]->scoreWithIndexes;
                    if (pos[s] >= sw.size()) continue;
                    if (best < 0) {
                        best = static_cast<int>(s);
                        continue;
                    }
                    const auto& a = sw[pos[s]];
                    const auto& b = per[best][j]->scoreWithIndexes[pos[best]];
                    if (a.score > b.score || (a.score == b.score && a.id < b.id)) best = static_cast<int>(s);
                }
                if (best < 0) break;
                const auto& e = per[best][j]->scoreWithIndexes[pos[best]++];
                outIds.push_back(static_cast<int64_t>(e.id));
                outSc.push_back(e.score);
            }
        }
        return true;
    }

    // The serial schedule (run/schedule_serial.cpp:23-51): warmup batches, then the queries in batches;
    // with pacing, query i arrives at t0 + i / rate and a batch starts once its first query arrived (or
    // at once when behind); each query's latency runs from its arrival (or the batch start) to the end
    // of its batch.
    bool RunSerial() {
        const long long warm = std::max<long long>(0, flags.Int("warmup", 5));
        for (long long w = 0; w < warm && nq > 0; w++)
            if (!RunBatch(0, std::min(nq, bs), false)) return false;
        t0 = Clock::now();
        for (size_t start = 0; start < nq; start += bs) {
            const size_t end = std::min(nq, start + bs);
            Clock::time_point arrival = Clock::now();
            if (paced) {
                arrival = t0 + std::chrono::duration_cast<Clock::duration>(
                                   std::chrono::duration<double>(static_cast<double>(start) / rate));
                std::this_thread::sleep_until(arrival);
            }
            const Clock::time_point begin = Clock::now();
            if (!RunBatch(start, end, true)) {
                std::fprintf(stderr, "search failed at batch [%zu, %zu)\n", start, end);
                return false;
            }
            const Clock::time_point done = Clock::now();
            for (size_t q = start; q < end; q++) {
                Clock::time_point from = begin;
                if (paced)
                    from = t0 + std::chrono::duration_cast<Clock::duration>(
                                    std::chrono::duration<double>(static_cast<double>(q) / rate));
                latMs[q] = std::chrono::duration<double, std::milli>(done - std::min(from, done)).count();
            }
        }
        wallMs = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        return true;
    }
};

// [Result] latency line (run/report.cpp:61-87): percentile p is the sorted sample at min(n - 1, p n).
void PrintLatency(std::vector<double> lat, double wallMs, size_t bs, double offered) {
    if (lat.empty()) return;
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) { return lat[std::min(lat.size() - 1, static_cast<size_t>(p * lat.size()))]; };
    double sum = 0;
    for (double v : lat) sum += v;
    const double qps = wallMs > 0 ? 1000.0 * lat.size() / wallMs : 0.0;
    if (offered > 0)
        std::printf("[Result] latency ms: avg=%.4f p50=%.4f p95=%.4f p99=%.4f qps=%.1f (per-query e2e incl. queueing, "
                    "offered=%.1f qps, n=%zu, batch_size=%zu)\n",
                    sum / lat.size(), pct(0.5), pct(0.95), pct(0.99), qps, offered, lat.size(), bs);
    else
        std::printf("[Result] latency ms: avg=%.4f p50=%.4f p95=%.4f p99=%.4f qps=%.1f (per-query e2e, n=%zu, "
                    "batch_size=%zu)\n",
                    sum / lat.size(), pct(0.5), pct(0.95), pct(0.99), qps, lat.size(), bs);
}

// [Result] CPU-recall (run/report.cpp:89-206): recall@k of the engine's ids against the CPU truth over
// the docs the index holds ([cpu_doc_offset, + docs)), averaged over queries with a non-empty truth.
void CheckRecall(const Run& r, uint64_t docNum) {
    const std::string path = r.flags.Str("dataset_hw");
    const long long want = r.flags.Int("recall_queries", 0);
    if (path.empty() || want <= 0) return;
    Dataset ds;
    const uint64_t offset = static_cast<uint64_t>(r.flags.Int("cpu_doc_offset", 0));
    if (!ds.Open(path) || ds.docs < offset + docNum) {
        std::fprintf(stderr, "cannot use dataset %s for %llu docs from %llu\n", path.c_str(),
                     (unsigned long long)docNum, (unsigned long long)offset);
        return;
    }
    const bool fp32 = r.flags.Bool("recall_fp32", true);
    const size_t rq = std::min<size_t>(static_cast<size_t>(want), r.nq);
    long long threads = r.flags.Int("recall_ref_threads", 0);
    if (threads <= 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::printf("[RecallRef] computing %zu brute-force refs (%s ground truth, %lld threads)...\n", rq,
                fp32 ? "fp32" : "fp16", threads);
    std::fflush(stdout);
    const Clock::time_point t0 = Clock::now();
    std::vector<std::unordered_set<int64_t>> refs(rq);
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (long long t = 0; t < threads; t++)
        pool.emplace_back([&] {
            for (size_t q = next++; q < rq; q = next++)
                refs[q] = CpuTruth(ds, r.queries[q % r.queries.size()], r.haveFilters ? r.filters.For(q) : nullptr,
                                   offset, docNum, r.topks[q], fp32);
        });
    for (auto& t : pool) t.join();
    std::printf("[RecallRef] built in %lld ms\n",
                static_cast<long long>(std::chrono::duration_cast<std::chrono::milliseconds>(Clock::now() - t0).count()));
    double sum = 0, minRecall = 1.0;
    size_t counted = 0, below = 0;
    for (size_t q = 0; q < rq; q++) {
        if (refs[q].empty()) continue;
        size_t hit = 0;
        for (int64_t id : r.ids[q]) hit += refs[q].count(id);
        const double rec = static_cast<double>(hit) / refs[q].size();
        sum += rec;
        counted++;
        minRecall = std::min(minRecall, rec);
        if (rec < 1.0) below++;
    }
    if (below) std::printf("[Result] %zu quer%s below recall 1.0 (min %.4f)\n", below, below == 1 ? "y" : "ies", minRecall);
    if (counted)
        std::printf("[Result] CPU-recall over %zu queries: avg=%.2f%%\n", counted, 100.0 * sum / counted);
    else
        std::printf("[Result] recall: no queries had non-empty reference results\n");
    std::fflush(stdout);
}

// --effective_filter_file (harness/recall/filter_report.h:1-34): for the first few queries, the filter as
// written, its terms with their token's doc count in the index (0 and "absent" for a token the index
// lacks: it evaluates to no doc), and the postfix program the device runs.
std::string AstText(const FilterAst* n) {
    if (!n) return "MATCH_ALL";
    switch (n->op) {
        case FilterAst::Term:
            return "TAG(" + std::to_string(n->tag) + ")";
        case FilterAst::Not:
            return "NOT(" + AstText(n->children[0].get()) + ")";
        default: {
            std::string s = "(";
            for (size_t i = 0; i < n->children.size(); i++)
                s += (i ? (n->op == FilterAst::And ? " AND " : " OR ") : "") + AstText(n->children[i].get());
            return s + ")";
        }
    }
}

void Terms(const FilterAst* n, std::vector<uint64_t>& out) {
    if (!n) return;
    if (n->op == FilterAst::Term) out.push_back(n->tag);
    for (const auto& c : n->children) Terms(c.get(), out);
}

void EffectiveFilterReport(const Run& r) {
    const std::string path = r.flags.Str("effective_filter_file");
    if (path.empty()) return;
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        return;
    }
    const auto& table = r.shards[0].table;
    NpuRetrieval::PostingFieldData* pf = nullptr;
    (void)table->GetPostingFieldData(r.postField, pf);
    for (size_t q = 0; q < std::min<size_t>(5, r.nq); q++) {
        const FilterAst* ast = r.haveFilters ? r.filters.For(q) : nullptr;
        std::fprintf(f, "query %zu topk=%u\n  written: %s\n  terms:", q, r.topks[q], AstText(ast).c_str());
        std::vector<uint64_t> tags;
        Terms(ast, tags);
        for (uint64_t t : tags) {
            const uint64_t tok = NpuRetrieval::Hash64(std::to_string(t));
            bool exist = false;
            if (pf) (void)pf->GetPostingDeviceAddrs(tok, exist);
            std::fprintf(f, " %llu:%s%llu", (unsigned long long)t, exist ? "docs=" : "absent,docs=",
                         (unsigned long long)(pf && exist ? pf->TokenHits(tok) : 0));
        }
        auto tree = TreeOf(ast, r.postField);
        NpuRetrieval::FilterExpr fe;
        if (!NpuRetrieval::PrepareFilterExpr(*table, tree.get(), fe)) {
            std::fprintf(f, "\n  program: (does not compile)\n");
            continue;
        }
        std::fprintf(f, "\n  program:");
        if (fe.emptyTree) std::fprintf(f, " MATCH_ALL");
        for (size_t i = 0; i < fe.postExpr.size();) {
            const uint32_t op = fe.postExpr[i];
            if (op == NpuRetrieval::CONJUNCTION) {
                const uint32_t parts = fe.postExpr[i + 1];
                std::fprintf(f, " CONJ(parts=%u)", parts);
                i += 2 + parts;
            } else {
                static const char* const kOp[4] = {"CONJ", "AND", "OR", "NOT"};
                std::fprintf(f, " %s(postings=%u,stack=%u)", kOp[op & 3], fe.postExpr[i + 1], fe.postExpr[i + 2]);
                i += 3;
            }
        }
        std::fprintf(f, "\n  operands=%zu\n", fe.postingDeviceAddrs.size());
    }
    std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    Run r;
    if (!ParseFlags(argc, argv, r.flags)) return 1;
    const Flags& F = r.flags;
    const std::vector<std::string> shardDirs = SplitCsv(F.Str("shard_index_dirs"));
    if ((F.Str("index_dir").empty() && shardDirs.empty()) || F.Str("query_file").empty()) {
        std::fprintf(stderr,
                     "usage: fr_search --index_dir DIR --query_file q.fvecs [--filter_file f.txt] [--topk K] "
                     "[--batch_size B] [--target_qps R] [--dataset_hw data.bin --recall_queries N] [...]\n"
                     "   or: fr_search --shard_index_dirs d0,d1,... --query_file q.fvecs [...]\n");
        return 1;
    }
    r.device = static_cast<int32_t>(F.Int("device_id", 0));
    r.vecField = F.Str("vec_field", "content");
    r.postField = F.Str("posting_field", "content");
    const std::string dt = F.Str("dtype", "f32");
    r.opts.dtype = dt == "bf16"    ? FRK_DTYPE_BF16
                   : dt == "fp8"   ? FRK_DTYPE_FP8E4M3
                   : dt == "f16x2" ? FRK_DTYPE_F16X2
                                   : FRK_DTYPE_F32;
    r.opts.rescoreFactor = static_cast<uint32_t>(F.Int("rescore_factor", 4));
    r.opts.probe = F.Bool("probe", true);
    r.opts.graph = F.Bool("graph", false);
    r.splitExtract = F.Bool("split_extract", false);
    r.unfused = F.Bool("unfused", false);
    const uint32_t mask = r.opts.dtype == FRK_DTYPE_F32 ? 0u : (1u << r.opts.dtype);

    for (const std::string& d : shardDirs.empty() ? std::vector<std::string>{F.Str("index_dir")} : shardDirs) {
        const Clock::time_point t0 = Clock::now();
        Shard s{d, std::make_shared<NpuRetrieval::DataTable>()};
        if (!s.table->LoadData(r.device, d, mask)) {
            std::fprintf(stderr, "cannot load index %s\n", d.c_str());
            return 2;
        }
        std::printf("[Load] %s: docs=%u segments=%u in %.1f ms\n", d.c_str(), s.table->GetDocNum(),
                    s.table->GetSegmentNum(), std::chrono::duration<double, std::milli>(Clock::now() - t0).count());
        r.shards.push_back(std::move(s));
    }
    if (!LoadQueries(F.Str("query_file"), r.queries)) {
        std::fprintf(stderr, "cannot read queries from %s\n", F.Str("query_file").c_str());
        return 2;
    }
    const long long numQ = F.Int("num_queries", 0);
    r.nq = numQ > 0 ? static_cast<size_t>(numQ) : r.queries.size();
    r.bs = static_cast<size_t>(std::max<long long>(1, F.Int("batch_size", 1)));
    if (!F.Str("filter_file").empty()) {
        if (!LoadFilters(F.Str("filter_file"), r.nq, static_cast<uint64_t>(F.Int("filter_seed", 42)), r.filters)) {
            std::fprintf(stderr, "cannot read filters from %s\n", F.Str("filter_file").c_str());
            return 2;
        }
        r.haveFilters = true;
    }
    if (!F.Str("topk_file").empty()) {
        std::vector<uint32_t> ks;
        if (!LoadTopKs(F.Str("topk_file"), ks) || ks.size() != r.nq) {
            std::fprintf(stderr, "topk file %s must hold one positive k per query (%zu)\n", F.Str("topk_file").c_str(), r.nq);
            return 2;
        }
        r.topks = ks;
    } else {
        r.topks.assign(r.nq, static_cast<uint32_t>(std::max<long long>(1, F.Int("topk", 100))));
    }
    r.ids.resize(r.nq);
    r.scores.resize(r.nq);
    r.latMs.assign(r.nq, 0.0);
    r.rate = F.Num("target_qps", 0.0);
    r.paced = r.rate > 0.0;
    EffectiveFilterReport(r);
    if (!r.RunSerial()) return 3;
    PrintLatency(r.latMs, r.wallMs, r.bs, r.rate);
    if (!F.Str("latency_dump").empty()) {
        if (std::FILE* lf = std::fopen(F.Str("latency_dump").c_str(), "w")) {
            std::fprintf(lf, "idx\tms\ttopk\n");
            for (size_t q = 0; q < r.nq; q++) std::fprintf(lf, "%zu\t%.4f\t%u\n", q, r.latMs[q], r.topks[q]);
            std::fclose(lf);
        }
    }
    if (!F.Str("result_file").empty()) {
        if (std::FILE* rf = std::fopen(F.Str("result_file").c_str(), "w")) {
            for (size_t q = 0; q < r.nq; q++) {
                std::fprintf(rf, "%zu\t%zu\t", q, r.ids[q].size());
                for (size_t i = 0; i < r.ids[q].size(); i++) std::fprintf(rf, "%s%lld", i ? "," : "", (long long)r.ids[q][i]);
                std::fprintf(rf, "\t");
                for (size_t i = 0; i < r.scores[q].size(); i++) std::fprintf(rf, "%s%.9g", i ? "," : "", r.scores[q][i]);
                std::fprintf(rf, "\n");
            }
            std::fclose(rf);
        }
    }
    uint64_t docNum = 0;
    for (const Shard& s : r.shards) docNum += s.table->GetDocNum();
    CheckRecall(r, docNum);
    if (F.Bool("pool_stats", false) || std::getenv("FRK_POOL_STATS")) NpuRetrieval::ReportWorkspacePool("end");
    return 0;
}
