// query_node.cpp -- query tree, postfix compilation and the filter grammar (see query_node.h).
#include "query_node.h"

#include "../frk_filter_program.h"

#include <algorithm>
#include <cmath>
#include <mutex>

#include "../../../include/frk_capi.h"
#include "filter_stats.h"
#include "log.h"

namespace NpuRetrieval {

void QueryNode::AdjustNode(const std::unordered_set<std::string>& postingFields) {
    for (auto it = m_children.begin(); it != m_children.end();) {
        (*it)->AdjustNode(postingFields);
        if ((*it)->IsNodeValid(postingFields))
            ++it;
        else
            it = m_children.erase(it);
    }
}

bool QueryNode::IsNodeValid(const std::unordered_set<std::string>&) const { return !m_children.empty(); }

void QueryNode::AddChild(std::unique_ptr<QueryNode>& child) {
    if (child) m_children.emplace_back(std::move(child));
}

namespace {
// AND/OR: the non-term children first (each leaves one stack entry), then the direct term children as
// posting operands, then [op, postingNum, stackNum] (query_node_imp.cpp:8-47).
bool EmitAndOr(const std::vector<std::unique_ptr<QueryNode>>& children, FilterOpType op, const DataTable& table,
               std::vector<uint32_t>& postExpr, PostingOperands& addrs, uint32_t& opNum) {
    uint32_t postingNum = 0, stackNum = 0;
    for (const auto& c : children)
        if (c && c->GetNodeType() != QueryNodeType::TermNode && c->GetPostOrderExpression(table, postExpr, addrs, opNum))
            ++stackNum;
    for (const auto& c : children)
        if (c && c->GetNodeType() == QueryNodeType::TermNode && c->GetPostOrderExpression(table, postExpr, addrs, opNum))
            ++postingNum;
    postExpr.push_back(op);
    postExpr.push_back(postingNum);
    postExpr.push_back(stackNum);
    opNum += 1;
    return true;
}

bool FieldOf(const DataTable& table, const std::string& field, PostingFieldData*& fd) {
    if (!table.GetPostingFieldData(field, fd) || fd == nullptr) {
        FRK_LOG_ERROR("posting field not in the index: " << field);
        return false;
    }
    return true;
}
}  // namespace

bool AndNode::GetPostOrderExpression(const DataTable& t, std::vector<uint32_t>& e, PostingOperands& a,
                                     uint32_t& n) const {
    return EmitAndOr(m_children, AND, t, e, a, n);
}

bool OrNode::GetPostOrderExpression(const DataTable& t, std::vector<uint32_t>& e, PostingOperands& a,
                                    uint32_t& n) const {
    return EmitAndOr(m_children, OR, t, e, a, n);
}

// NOT: the child, then [3, 1, 0] for a term child or [3, 0, 1] otherwise (query_node_imp.cpp:65-96).
bool NotNode::GetPostOrderExpression(const DataTable& t, std::vector<uint32_t>& e, PostingOperands& a,
                                     uint32_t& n) const {
    if (m_children.size() != 1 || !m_children[0]) {
        FRK_LOG_ERROR("not node needs exactly one child, has " << m_children.size());
        return false;
    }
    const bool term = m_children[0]->GetNodeType() == QueryNodeType::TermNode;
    if (!m_children[0]->GetPostOrderExpression(t, e, a, n)) return false;
    e.push_back(NOT);
    e.push_back(term ? 1 : 0);
    e.push_back(term ? 0 : 1);
    n += 1;
    return true;
}

bool TermNode::IsNodeValid(const std::unordered_set<std::string>& postingFields) const {
    if (!postingFields.count(m_fieldName)) {
        FRK_LOG_WARN("filter term dropped: posting field '" << m_fieldName << "' is not in this index");
        return false;
    }
    return true;
}

bool TermNode::GetPostOrderExpression(const DataTable& t, std::vector<uint32_t>&, PostingOperands& a,
                                      uint32_t&) const {
    return AppendPostings(t, a);
}

// One operand, the all-zero one if the token is absent (query_node_imp.cpp:124-160).
bool TermNode::AppendPostings(const DataTable& t, PostingOperands& a) const {
    PostingFieldData* fd = nullptr;
    if (!FieldOf(t, m_fieldName, fd)) return false;
    bool exist = true;
    a.push_back(fd->GetPostingDeviceAddrs(m_token, exist));
    return true;
}

bool TermsNode::IsNodeValid(const std::unordered_set<std::string>& postingFields) const {
    if (!postingFields.count(m_fieldName)) {
        FRK_LOG_WARN("filter term dropped: posting field '" << m_fieldName << "' is not in this index");
        return false;
    }
    return true;
}

// OR over the tokens; after one posting is in, further absent tokens add nothing
// (query_node_imp.cpp:180-243).
bool TermsNode::GetPostOrderExpression(const DataTable& t, std::vector<uint32_t>& e, PostingOperands& a,
                                       uint32_t& n) const {
    const size_t base = a.size();
    if (!AppendPostings(t, a)) return false;
    const uint32_t postingNum = static_cast<uint32_t>(a.size() - base);
    if (postingNum == 0) {
        FRK_LOG_ERROR("terms node over posting field '" << m_fieldName << "' resolved no token");
        return false;
    }
    e.push_back(OR);
    e.push_back(postingNum);
    e.push_back(0);
    n += 1;
    return true;
}

bool TermsNode::AppendPostings(const DataTable& t, PostingOperands& a) const {
    PostingFieldData* fd = nullptr;
    if (!FieldOf(t, m_fieldName, fd) || !m_tokens) return false;
    bool added = false;
    for (uint64_t tok : *m_tokens) {
        bool exist = true;
        const std::vector<uint64_t>* addrs = fd->GetPostingDeviceAddrs(tok, exist);
        if (!exist && added) continue;
        a.push_back(addrs);
        added = true;
    }
    return true;
}

bool OrTermsNode::GetPostOrderExpression(const DataTable& t, std::vector<uint32_t>& e, PostingOperands& a,
                                         uint32_t& n) const {
    const size_t base = a.size();
    if (!AppendPostings(t, a)) return false;
    const uint32_t postingNum = static_cast<uint32_t>(a.size() - base);
    if (postingNum == 0) {
        FRK_LOG_ERROR("or-terms node resolved no token");
        return false;
    }
    e.push_back(OR);
    e.push_back(postingNum);
    e.push_back(0);
    n += 1;
    return true;
}

// Term/Terms children only (query_node_imp.cpp:245-293).
bool OrTermsNode::AppendPostings(const DataTable& t, PostingOperands& a) const {
    for (const auto& c : m_children) {
        auto* leaf = dynamic_cast<QueryLeafNode*>(c.get());
        if (!leaf) continue;
        const QueryNodeType ty = leaf->GetNodeType();
        if (ty != QueryNodeType::TermNode && ty != QueryNodeType::TermsNode) continue;
        leaf->AppendPostings(t, a);
    }
    return true;
}

// [0, parts, n_1 .. n_parts]: an AND of one OR per leaf child (query_node_imp.cpp:295-342).
bool ConjunctionNode::GetPostOrderExpression(const DataTable& t, std::vector<uint32_t>& e, PostingOperands& a,
                                             uint32_t& n) const {
    std::vector<uint32_t> groups;
    for (const auto& c : m_children) {
        auto* leaf = dynamic_cast<QueryLeafNode*>(c.get());
        if (!leaf) continue;
        const QueryNodeType ty = leaf->GetNodeType();
        if (ty != QueryNodeType::TermNode && ty != QueryNodeType::TermsNode && ty != QueryNodeType::OrTermsNode) continue;
        const size_t base = a.size();
        if (!leaf->AppendPostings(t, a)) continue;
        const uint32_t num = static_cast<uint32_t>(a.size() - base);
        if (num > 0) groups.push_back(num);
    }
    if (groups.empty()) {
        FRK_LOG_ERROR("conjunction node resolved no token");
        return false;
    }
    e.push_back(CONJUNCTION);
    e.push_back(static_cast<uint32_t>(groups.size()));
    for (uint32_t g : groups) e.push_back(g);
    n += 1;
    return true;
}

// ------------------------------------------------------------------------------------------ prepare
// The structural rule of the reference's check (text_filter.cpp:1220-1352), in the wrap-free form the
// filter kernel applies too (frk_filter_program.h); any stack depth is allowed here, the kernel's
// limit is checked separately with its own message.
bool CheckPostExpr(const std::vector<uint32_t>& e, uint32_t postingNums) {
    if (e.size() > 0xFFFFFFFFull) return false;
    return frk::ValidateProgram(e.data(), static_cast<uint32_t>(e.size()), postingNums, nullptr, 1 << 30);
}

uint32_t PostExprMaxStack(const std::vector<uint32_t>& e) {
    int sp = 0, mx = 0;
    for (size_t i = 0; i < e.size(); i++) {
        if (e[i] == CONJUNCTION) {
            i += e[i + 1] + 1;
            sp += 1;
        } else {
            sp = sp - static_cast<int>(e[i + 2]) + 1;
            i += 2;
        }
        mx = sp > mx ? sp : mx;
    }
    return static_cast<uint32_t>(mx);
}

// ------------------------------------------------------------------------------ floor features
namespace {
double TermDensity(const DataTable& t, const std::string& field, uint64_t token) {
    PostingFieldData* fd = nullptr;
    if (!t.GetPostingFieldData(field, fd) || fd == nullptr || t.GetDocNum() == 0) return 0.0;
    return static_cast<double>(fd->TokenHits(token)) / static_cast<double>(t.GetDocNum());
}
}  // namespace

void QueryNode::Estimate(const DataTable& t, FilterStats& s) const {
    const QueryNodeType ty = GetNodeType();
    double prod = 1.0;  // AND: product of selectivities; OR: product of complements
    uint32_t depth = 0;
    bool first = true;
    double firstSel = 0.0;
    for (const auto& c : m_children) {
        FilterStats cs;
        c->Estimate(t, cs);
        if (first) firstSel = cs.selectivity;
        first = false;
        prod *= (ty == QueryNodeType::OrNode || ty == QueryNodeType::OrTermsNode) ? 1.0 - cs.selectivity
                                                                                   : cs.selectivity;
        depth = std::max(depth, cs.depth);
        s.nAnd += cs.nAnd;
        s.nOr += cs.nOr;
        s.nNot += cs.nNot;
        s.nTerms += cs.nTerms;
        s.termDensities.insert(s.termDensities.end(), cs.termDensities.begin(), cs.termDensities.end());
    }
    s.depth = depth + 1;
    if (ty == QueryNodeType::NotNode) {
        s.selectivity = 1.0 - firstSel;
        s.nNot++;
    } else if (ty == QueryNodeType::OrNode || ty == QueryNodeType::OrTermsNode) {
        s.selectivity = 1.0 - prod;
        s.nOr++;
    } else {  // And, Conjunction (an AND of OR groups)
        s.selectivity = prod;
        s.nAnd++;
    }
}

void TermNode::Estimate(const DataTable& t, FilterStats& s) const {
    const double d = TermDensity(t, m_fieldName, m_token);
    s.selectivity = d;
    s.depth = 1;
    s.nTerms = 1;
    s.termDensities.assign(1, d);
}

void TermsNode::Estimate(const DataTable& t, FilterStats& s) const {
    double comp = 1.0;
    s.termDensities.clear();
    if (m_tokens)
        for (uint64_t tok : *m_tokens) {
            const double d = TermDensity(t, m_fieldName, tok);
            comp *= 1.0 - d;
            s.termDensities.push_back(d);
        }
    s.selectivity = 1.0 - comp;
    s.depth = 2;
    s.nOr = 1;
    s.nTerms = static_cast<uint32_t>(s.termDensities.size());
}

void FloorFeatures(const DataTable& table, QueryNode* tree, uint32_t k, const float* query, uint32_t dim,
                   float* out) {
    std::fill(out, out + kFloorFeatures, 0.0f);
    FilterStats s;
    const bool matchAll = tree == nullptr || (!tree->IsLeaf() && tree->GetChildrenNum() == 0);
    if (matchAll) {
        out[14] = 1.0f;
    } else {
        std::shared_lock<std::shared_mutex> lock(tree->m_adjustNodeMutex);
        tree->Estimate(table, s);
    }
    const double sel = std::min(std::max(s.selectivity, 1e-9), 1.0);
    const double passing = std::max(sel * static_cast<double>(table.GetDocNum()), 1.0);
    out[0] = static_cast<float>(std::log10(sel));
    out[1] = static_cast<float>(std::log10(static_cast<double>(std::max<uint32_t>(k, 1))));
    out[2] = static_cast<float>(std::log10(passing));
    out[3] = std::min(std::max(out[1] - out[2], -9.0f), 3.0f);
    out[4] = static_cast<float>(s.depth);
    out[5] = static_cast<float>(s.nAnd);
    out[6] = static_cast<float>(s.nOr);
    out[7] = static_cast<float>(s.nNot);
    out[8] = static_cast<float>(s.nTerms);
    std::vector<double> rare;
    for (double d : s.termDensities) rare.push_back(std::log10(std::max(d, 1e-9)));
    std::sort(rare.begin(), rare.end());
    for (size_t i = 0; i < rare.size() && i < 4; i++) out[9 + i] = static_cast<float>(rare[i]);
    double norm = 0.0;
    for (uint32_t d = 0; d < dim && query; d++) norm += static_cast<double>(query[d]) * query[d];
    out[13] = static_cast<float>(std::sqrt(norm));
}

bool PrepareFilterExpr(const DataTable& table, QueryNode* tree, FilterExpr& out) {
    out = FilterExpr();
    if (tree == nullptr) {
        out.emptyTree = out.valid = true;
        return true;
    }
    {
        std::unique_lock<std::shared_mutex> lock(tree->m_adjustNodeMutex);
        if (!tree->m_isAdjustNodeDone) {
            tree->AdjustNode(table.GetPostingFields());
            tree->m_isAdjustNodeDone = true;
        }
    }
    std::shared_lock<std::shared_mutex> lock(tree->m_adjustNodeMutex);
    if (tree->IsLeaf()) {
        if (!tree->IsNodeValid(table.GetPostingFields())) {
            out.emptyTree = out.valid = true;  // the only term is on a field the index lacks
            return true;
        }
        if (tree->GetNodeType() == QueryNodeType::TermNode) {
            if (!static_cast<QueryLeafNode*>(tree)->AppendPostings(table, out.postingDeviceAddrs)) return false;
            out.postExpr = {OR, 1, 0};
            out.opNum = 1;
        } else if (!tree->GetPostOrderExpression(table, out.postExpr, out.postingDeviceAddrs, out.opNum)) {
            return false;
        }
    } else if (tree->GetChildrenNum() == 0) {
        out.emptyTree = out.valid = true;  // text_filter.cpp:443-448
        return true;
    } else if (!tree->GetPostOrderExpression(table, out.postExpr, out.postingDeviceAddrs, out.opNum)) {
        FRK_LOG_ERROR("postfix emission of the filter tree failed");
        return false;
    }
    if (!CheckPostExpr(out.postExpr, static_cast<uint32_t>(out.postingDeviceAddrs.size()))) {
        FRK_LOG_ERROR("compiled filter program failed the structural check");
        return false;
    }
    if (ExprStatsEnabled()) NoteExprShape(out.postExpr);
    if (PostExprMaxStack(out.postExpr) > FRK_FILTER_MAX_STACK ||
        out.postingDeviceAddrs.size() > FRK_FILTER_MAX_POSTINGS) {
        FRK_LOG_ERROR("filter needs a stack of " << PostExprMaxStack(out.postExpr) << " (max " << FRK_FILTER_MAX_STACK
                                                 << ") and " << out.postingDeviceAddrs.size() << " operands (max "
                                                 << FRK_FILTER_MAX_POSTINGS << ")");
        return false;
    }
    out.valid = true;
    return true;
}

// ------------------------------------------------------------------------------------------ grammar
namespace {
struct Tok {
    const std::string& s;
    size_t i = 0;
    enum Kind { End, LP, RP, Num, And, Or, Not, Other };
    explicit Tok(const std::string& x) : s(x) {}
    void Ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r' || s[i] == '\n')) i++;
    }
    Kind Peek(uint64_t& num, size_t& len) {
        Ws();
        len = 0;
        if (i >= s.size()) return End;
        const char c = s[i];
        if (c == '(' || c == ')') {
            len = 1;
            return c == '(' ? LP : RP;
        }
        if (c >= '0' && c <= '9') {
            size_t j = i;
            uint64_t v = 0;
            while (j < s.size() && s[j] >= '0' && s[j] <= '9') v = v * 10 + static_cast<uint64_t>(s[j++] - '0');
            num = v;
            len = j - i;
            return Num;
        }
        size_t j = i;
        while (j < s.size() && ((s[j] >= 'A' && s[j] <= 'Z') || (s[j] >= 'a' && s[j] <= 'z'))) j++;
        len = j - i;
        const std::string w = s.substr(i, len);
        return w == "AND" ? And : w == "OR" ? Or : w == "NOT" ? Not : Other;
    }
};

std::unique_ptr<FilterAst> POr(Tok& t, bool& ok);

std::unique_ptr<FilterAst> PAtom(Tok& t, bool& ok) {
    uint64_t num = 0;
    size_t len = 0;
    const Tok::Kind k = t.Peek(num, len);
    if (k == Tok::LP) {
        t.i += len;
        auto n = POr(t, ok);
        if (!ok || t.Peek(num, len) != Tok::RP) {
            ok = false;
            return nullptr;
        }
        t.i += len;
        return n;
    }
    if (k == Tok::Num) {
        t.i += len;
        auto n = std::make_unique<FilterAst>();
        n->op = FilterAst::Term;
        n->tag = num;
        return n;
    }
    ok = false;
    return nullptr;
}

std::unique_ptr<FilterAst> PNot(Tok& t, bool& ok) {
    uint64_t num = 0;
    size_t len = 0;
    if (t.Peek(num, len) == Tok::Not) {
        t.i += len;
        auto c = PNot(t, ok);
        if (!ok) return nullptr;
        auto n = std::make_unique<FilterAst>();
        n->op = FilterAst::Not;
        n->children.push_back(std::move(c));
        return n;
    }
    return PAtom(t, ok);
}

std::unique_ptr<FilterAst> PChain(Tok& t, bool& ok, Tok::Kind sep, FilterAst::Op op,
                                  std::unique_ptr<FilterAst> (*sub)(Tok&, bool&)) {
    std::vector<std::unique_ptr<FilterAst>> parts;
    parts.push_back(sub(t, ok));
    if (!ok) return nullptr;
    uint64_t num = 0;
    size_t len = 0;
    while (t.Peek(num, len) == sep) {
        t.i += len;
        parts.push_back(sub(t, ok));
        if (!ok) return nullptr;
    }
    if (parts.size() == 1) return std::move(parts[0]);
    auto n = std::make_unique<FilterAst>();
    n->op = op;
    n->children = std::move(parts);
    return n;
}

std::unique_ptr<FilterAst> PAnd(Tok& t, bool& ok) { return PChain(t, ok, Tok::And, FilterAst::And, PNot); }
std::unique_ptr<FilterAst> POr(Tok& t, bool& ok) { return PChain(t, ok, Tok::Or, FilterAst::Or, PAnd); }
}  // namespace

std::unique_ptr<FilterAst> ParseFilter(const std::string& expr, bool& ok) {
    ok = true;
    Tok t(expr);
    uint64_t num = 0;
    size_t len = 0;
    if (t.Peek(num, len) == Tok::End) return nullptr;  // empty: match all
    auto n = POr(t, ok);
    if (!ok) return nullptr;
    if (t.Peek(num, len) != Tok::End) {
        ok = false;
        return nullptr;
    }
    return n;
}

std::unique_ptr<QueryNode> BuildQueryNode(const FilterAst* n, const std::string& field) {
    if (n == nullptr) return std::make_unique<AndNode>();  // no filter: an empty AND, i.e. match all
    switch (n->op) {
        case FilterAst::Term:
            return std::make_unique<TermNode>(field, TagToken(n->tag));
        case FilterAst::Not: {
            auto node = std::make_unique<NotNode>();
            auto c = BuildQueryNode(n->children[0].get(), field);
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
            for (const auto& c : n->children) {
                auto child = BuildQueryNode(c.get(), field);
                node->AddChild(child);
            }
            return node;
        }
    }
    return std::make_unique<AndNode>();
}

}  // namespace NpuRetrieval
