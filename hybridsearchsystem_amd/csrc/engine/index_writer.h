// index_writer.h -- writes an index directory in the on-disk layout DataTable::LoadData reads: the
// offline builder's output (indexer/builder/fr_builder.cpp with segment/vector_builder.cpp,
// segment/inverted_builder.cpp and the id mapping and meta writers). Host-only: no device memory.
//
// Directory:
//   poissonengine_<v>.meta                          JSON: config, schema, statistics
//   id/poissonengine_<v>_<s>.idm                    header ext "0", then u64 global doc ids
//   <vec>/poissonengine_<v>_<s>_<vec>.vector        header ext [u16 dim] + FP16 ZN body, or
//                                                   [u16 dim][u16 1] + FP32 rows
//   <tag>/poissonengine_<v>_<s>_<tag>.posting       header ext "0" (matrix) / "1" (bitmap), u32 tokenNum,
//                                                   {u64 Hash64(token), u32 postingOff, u32 tokenOff}
//                                                   per token, postings (each padded to 4 bytes), then
//                                                   the token literals {u32 len, bytes}
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "format.h"

namespace NpuRetrieval {

class IndexWriter {
   public:
    IndexWriter(uint32_t docNum, uint32_t docNumPerSegment, uint32_t version = 0, uint32_t splitDocNumZn = 32);

    // rows: [docNum][dim] float, row-major; kept by pointer until Write returns.
    bool AddVectorField(const std::string& name, uint32_t dim, const float* rows, VectorFileType type);
    // Token postings as CSR over nTokens tokens: token t holds the global doc ids
    // docIds[offsets[t] .. offsets[t+1]) (any order, duplicates allowed). Token t's text is terms[t]
    // when terms is given, else the decimal form of tags[t] (how numeric terms become tokens,
    // inverted_builder.cpp:17-35). Arrays are kept by pointer until Write returns.
    bool AddPostingField(const std::string& name, bool isMatrix, float densityThreshold, bool packed, uint64_t nTokens,
                         const uint64_t* tags, const char* const* terms, const uint64_t* offsets,
                         const uint32_t* docIds);
    void SetGDocIds(const uint64_t* ids, uint64_t n);  // default: 0..docNum-1

    bool Write(const std::string& dir);
    const std::string& Error() const { return m_err; }

   private:
    struct VectorField {
        std::string name;
        uint32_t dim;
        const float* rows;
        VectorFileType type;
    };
    struct PostingField {
        std::string name;
        bool isMatrix;
        float threshold;
        bool packed;
        std::vector<std::string> tokens;
        std::vector<std::vector<uint32_t>> docs;  // per token: sorted, unique global doc ids
    };
    bool WriteVectorSegment(const std::string& dir, const VectorField& f, uint32_t seg);
    bool WritePostingSegment(const std::string& dir, const PostingField& f, uint32_t seg,
                             std::vector<size_t>& cursor);
    bool WriteIdSegment(const std::string& dir, uint32_t seg);
    bool Fail(const std::string& e) {
        m_err = e;
        return false;
    }

    uint32_t m_docNum, m_docNumPerSegment, m_version, m_splitDocNumZn, m_segmentNum;
    std::vector<VectorField> m_vectors;
    std::vector<PostingField> m_postings;
    std::vector<uint64_t> m_gdocids;
    std::string m_err;
};

}  // namespace NpuRetrieval
