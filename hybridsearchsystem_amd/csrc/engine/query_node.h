// query_node.h -- the filter query tree and its postfix compilation, with the reference's interface
// (retrieval/query/query_node.h:24-64, query_node_imp.h): And/Or/Not/Term/Terms/OrTerms/Conjunction
// nodes, AdjustNode pruning of terms on fields the index does not have, and GetPostOrderExpression
// emitting the postfix program the filter kernel runs plus the posting operands it consumes.
//
// Operands are the per-segment device addresses of a token's postings (PostingFieldData); the
// kernel reads each posting's head for its layout, so no separate type/weight arrays are passed.
#pragma once

#include <cstdint>
#include <memory>
#include <shared_mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "data_table.h"

namespace NpuRetrieval {

enum class QueryNodeType { AndNode = 1, OrNode, NotNode, TermNode, TermsNode, OrTermsNode, ConjunctionNode };

// Postfix op codes (retrieval/device/device_common.h:8-13).
enum FilterOpType : uint32_t { CONJUNCTION = 0, AND = 1, OR = 2, NOT = 3 };

using PostingOperands = std::vector<const std::vector<uint64_t>*>;

// What the learned score floor knows of a filter: its selectivity estimated under tag independence
// from the table's per-token doc counts, and its shape.
struct FilterStats {
    double selectivity = 1.0;
    uint32_t depth = 0, nAnd = 0, nOr = 0, nNot = 0, nTerms = 0;
    std::vector<double> termDensities;
};

class QueryNode {
   public:
    QueryNode() = default;
    virtual ~QueryNode() = default;
    // Drops children that end up invalid, depth first (query_node.cpp:7-17).
    virtual void AdjustNode(const std::unordered_set<std::string>& postingFields);
    virtual bool IsNodeValid(const std::unordered_set<std::string>& postingFields) const;
    virtual bool GetPostOrderExpression(const DataTable& dataTable, std::vector<uint32_t>& postExpr,
                                        PostingOperands& postingDeviceAddrs, uint32_t& opNum) const = 0;
    void AddChild(std::unique_ptr<QueryNode>& child);
    virtual QueryNodeType GetNodeType() const = 0;
    virtual std::string GetName() const = 0;
    // Filter statistics of this subtree (AND: product, OR: 1 - product of complements, NOT:
    // complement; a term's density is its doc count over the table's docs).
    virtual void Estimate(const DataTable& dataTable, FilterStats& s) const;
    uint32_t GetChildrenNum() const { return static_cast<uint32_t>(m_children.size()); }
    bool IsLeaf() const {
        QueryNodeType t = GetNodeType();
        return t == QueryNodeType::TermNode || t == QueryNodeType::TermsNode || t == QueryNodeType::OrTermsNode;
    }

    mutable std::shared_mutex m_adjustNodeMutex;
    volatile bool m_isAdjustNodeDone = false;

   protected:
    std::vector<std::unique_ptr<QueryNode>> m_children;
};

class QueryLeafNode : public QueryNode {
   public:
    virtual bool AppendPostings(const DataTable& dataTable, PostingOperands& postingDeviceAddrs) const = 0;
};

class AndNode : public QueryNode {
   public:
    bool GetPostOrderExpression(const DataTable& dataTable, std::vector<uint32_t>& postExpr, PostingOperands& addrs,
                                uint32_t& opNum) const override;
    QueryNodeType GetNodeType() const override { return QueryNodeType::AndNode; }
    std::string GetName() const override { return "and"; }
};

class OrNode : public QueryNode {
   public:
    bool GetPostOrderExpression(const DataTable& dataTable, std::vector<uint32_t>& postExpr, PostingOperands& addrs,
                                uint32_t& opNum) const override;
    QueryNodeType GetNodeType() const override { return QueryNodeType::OrNode; }
    std::string GetName() const override { return "or"; }
};

class NotNode : public QueryNode {
   public:
    bool GetPostOrderExpression(const DataTable& dataTable, std::vector<uint32_t>& postExpr, PostingOperands& addrs,
                                uint32_t& opNum) const override;
    QueryNodeType GetNodeType() const override { return QueryNodeType::NotNode; }
    std::string GetName() const override { return "not"; }
};

class TermNode : public QueryLeafNode {
   public:
    TermNode(const std::string& fieldName, uint64_t token) : m_fieldName(fieldName), m_token(token) {}
    void AdjustNode(const std::unordered_set<std::string>&) override {}
    bool IsNodeValid(const std::unordered_set<std::string>& postingFields) const override;
    bool GetPostOrderExpression(const DataTable& dataTable, std::vector<uint32_t>& postExpr, PostingOperands& addrs,
                                uint32_t& opNum) const override;
    bool AppendPostings(const DataTable& dataTable, PostingOperands& addrs) const override;
    void Estimate(const DataTable& dataTable, FilterStats& s) const override;
    QueryNodeType GetNodeType() const override { return QueryNodeType::TermNode; }
    std::string GetName() const override { return "term:" + m_fieldName + ":" + std::to_string(m_token); }

   private:
    std::string m_fieldName;
    uint64_t m_token;
};

class TermsNode : public QueryLeafNode {
   public:
    TermsNode(const std::string& fieldName, std::unique_ptr<std::unordered_set<uint64_t>> tokens)
        : m_fieldName(fieldName), m_tokens(std::move(tokens)) {}
    void AdjustNode(const std::unordered_set<std::string>&) override {}
    bool IsNodeValid(const std::unordered_set<std::string>& postingFields) const override;
    bool GetPostOrderExpression(const DataTable& dataTable, std::vector<uint32_t>& postExpr, PostingOperands& addrs,
                                uint32_t& opNum) const override;
    bool AppendPostings(const DataTable& dataTable, PostingOperands& addrs) const override;
    void Estimate(const DataTable& dataTable, FilterStats& s) const override;
    QueryNodeType GetNodeType() const override { return QueryNodeType::TermsNode; }
    std::string GetName() const override { return "terms:" + m_fieldName; }

   private:
    std::string m_fieldName;
    std::unique_ptr<std::unordered_set<uint64_t>> m_tokens;
};

class OrTermsNode : public QueryLeafNode {
   public:
    bool GetPostOrderExpression(const DataTable& dataTable, std::vector<uint32_t>& postExpr, PostingOperands& addrs,
                                uint32_t& opNum) const override;
    bool AppendPostings(const DataTable& dataTable, PostingOperands& addrs) const override;
    QueryNodeType GetNodeType() const override { return QueryNodeType::OrTermsNode; }
    std::string GetName() const override { return "or_terms"; }
};

class ConjunctionNode : public QueryNode {
   public:
    bool GetPostOrderExpression(const DataTable& dataTable, std::vector<uint32_t>& postExpr, PostingOperands& addrs,
                                uint32_t& opNum) const override;
    QueryNodeType GetNodeType() const override { return QueryNodeType::ConjunctionNode; }
    std::string GetName() const override { return "conjunction"; }
};

// The 32 features the floor model reads for one query (k = the keys the main pass selects):
// [0] log10 selectivity, [1] log10 k, [2] log10 expected passing docs, [3] clip([1] - [2], -9, 3),
// [4] depth, [5..8] AND / OR / NOT / term counts, [9..12] log10 densities of the rarest terms
// (ascending, 0 when fewer), [13] query L2 norm, [14] 1 for a match-all filter, the rest 0.
constexpr uint32_t kFloorFeatures = 32;
void FloorFeatures(const DataTable& table, QueryNode* tree, uint32_t k, const float* query, uint32_t dim,
                   float* out);

// The compiled filter of one query: the postfix program, its operands, and whether it matches all.
struct FilterExpr {
    std::vector<uint32_t> postExpr;
    PostingOperands postingDeviceAddrs;
    uint32_t opNum = 0;
    bool emptyTree = false;
    bool valid = false;
};

// AdjustNode (once per tree, under its lock) then GetPostOrderExpression, as TextFilter::PrepareExpr
// (searcher/stage/text_filter.cpp:432-467); then the CheckPostExpr invariants (:1220-1352) and this
// kernel's limits. A tree with no children left matches all docs. A root that is itself a leaf is
// compiled as an OR over its postings, so a bare term filters by that term.
bool PrepareFilterExpr(const DataTable& table, QueryNode* tree, FilterExpr& out);
bool CheckPostExpr(const std::vector<uint32_t>& postExpr, uint32_t postingNums);
uint32_t PostExprMaxStack(const std::vector<uint32_t>& postExpr);

// ---------------------------------------------------------------------------------------------
// Filter grammar of the search driver (harness/query/filter_expr.h:1-183): precedence NOT > AND > OR,
// parentheses, tag ids as numbers, AND/OR chains flattened into n-ary nodes; an empty string is
// match-all. BuildQueryNode turns it into a QueryNode tree over one posting field with tokens
// Hash64(tag) (harness/query/filter_setup.cpp:93-136).
struct FilterAst {
    enum Op { Term, And, Or, Not } op = Term;
    uint64_t tag = 0;
    std::vector<std::unique_ptr<FilterAst>> children;
};

std::unique_ptr<FilterAst> ParseFilter(const std::string& expr, bool& ok);
std::unique_ptr<QueryNode> BuildQueryNode(const FilterAst* ast, const std::string& field);

}  // namespace NpuRetrieval
