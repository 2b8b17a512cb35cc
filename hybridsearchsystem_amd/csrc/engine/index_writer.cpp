// index_writer.cpp -- index directory writer (see index_writer.h).
#include "index_writer.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>

#include "../frk_common.h"

namespace NpuRetrieval {

namespace {
bool WriteAll(const std::string& path, const std::string& head, const void* body, size_t bodyLen) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) return false;
    f.write(head.data(), static_cast<std::streamsize>(head.size()));
    if (bodyLen) f.write(static_cast<const char*>(body), static_cast<std::streamsize>(bodyLen));
    return static_cast<bool>(f);
}

bool MakeDir(const std::string& d) {
    std::error_code ec;
    std::filesystem::create_directories(d, ec);
    return std::filesystem::is_directory(d, ec);
}

template <typename T>
void Put(std::string& s, T v) {
    s.append(reinterpret_cast<const char*>(&v), sizeof(T));
}
}  // namespace

IndexWriter::IndexWriter(uint32_t docNum, uint32_t docNumPerSegment, uint32_t version, uint32_t splitDocNumZn)
    : m_docNum(docNum), m_docNumPerSegment(docNumPerSegment), m_version(version), m_splitDocNumZn(splitDocNumZn) {
    m_segmentNum = docNumPerSegment ? (docNum + docNumPerSegment - 1) / docNumPerSegment : 0;
    if (docNumPerSegment == 0 || docNumPerSegment % 8 != 0)
        m_err = "docNumPerSegment must be a positive multiple of 8";
    else if (splitDocNumZn == 0)
        m_err = "splitDocNumZn must be positive";
    else if (version > 255)
        m_err = "the file header holds an 8-bit version";
}

bool IndexWriter::AddVectorField(const std::string& name, uint32_t dim, const float* rows, VectorFileType type) {
    if (!m_err.empty()) return false;
    if (name.empty() || dim == 0 || dim > 0xFFFF || (!rows && m_docNum))
        return Fail("vector field " + name + ": bad name, dimension or rows");
    m_vectors.push_back({name, dim, rows, type});
    return true;
}

bool IndexWriter::AddPostingField(const std::string& name, bool isMatrix, float densityThreshold, bool packed,
                                  uint64_t nTokens, const uint64_t* tags, const char* const* terms,
                                  const uint64_t* offsets, const uint32_t* docIds) {
    if (!m_err.empty()) return false;
    if (name.empty() || (nTokens && (!offsets || (!tags && !terms))))
        return Fail("posting field " + name + ": bad name or token arrays");
    PostingField f{name, isMatrix, densityThreshold, packed, {}, {}};
    f.tokens.resize(nTokens);
    f.docs.resize(nTokens);
    for (uint64_t t = 0; t < nTokens; t++) {
        if (terms) {
            if (!terms[t] || !terms[t][0]) return Fail("posting field " + name + ": empty token text");
            f.tokens[t] = terms[t];
        } else {
            f.tokens[t] = std::to_string(tags[t]);
        }
        if (offsets[t + 1] < offsets[t]) return Fail("posting field " + name + ": offsets decrease");
        auto& d = f.docs[t];
        d.assign(docIds + offsets[t], docIds + offsets[t + 1]);
        std::sort(d.begin(), d.end());
        d.erase(std::unique(d.begin(), d.end()), d.end());  // a doc holds a token once
        if (!d.empty() && d.back() >= m_docNum) return Fail("posting field " + name + ": doc id beyond the table");
    }
    m_postings.push_back(std::move(f));
    return true;
}

void IndexWriter::SetGDocIds(const uint64_t* ids, uint64_t n) { m_gdocids.assign(ids, ids + n); }

bool IndexWriter::WriteIdSegment(const std::string& dir, uint32_t seg) {
    const uint64_t first = static_cast<uint64_t>(seg) * m_docNumPerSegment;
    const uint64_t n = std::min<uint64_t>(m_docNumPerSegment, m_docNum - first);
    std::vector<uint64_t> ids(n);
    for (uint64_t i = 0; i < n; i++) ids[i] = m_gdocids.empty() ? first + i : m_gdocids[first + i];
    const std::string path = PathJoin(PathJoin(dir, kIdMappingDir), SegmentFileName(m_version, seg, "", kIdMappingSuffix));
    if (!WriteAll(path, EncodeFileHeader(static_cast<uint8_t>(m_version), kGdocidType), ids.data(), n * 8))
        return Fail("cannot write " + path);
    return true;
}

bool IndexWriter::WriteVectorSegment(const std::string& dir, const VectorField& f, uint32_t seg) {
    const uint64_t first = static_cast<uint64_t>(seg) * m_docNumPerSegment;
    const uint64_t n = std::min<uint64_t>(m_docNumPerSegment, m_docNum - first);
    const std::string path = PathJoin(PathJoin(dir, f.name), SegmentFileName(m_version, seg, f.name, kVectorSuffix));
    std::string ext;
    Put<uint16_t>(ext, static_cast<uint16_t>(f.dim));
    const float* rows = f.rows + first * f.dim;
    bool ok;
    if (f.type == VectorFileType::FP32_ROWS) {
        Put<uint16_t>(ext, static_cast<uint16_t>(VectorFileType::FP32_ROWS));
        ok = WriteAll(path, EncodeFileHeader(static_cast<uint8_t>(m_version), ext), rows, n * f.dim * 4);
    } else if (f.type == VectorFileType::BF16_ROWS) {
        Put<uint16_t>(ext, static_cast<uint16_t>(VectorFileType::BF16_ROWS));
        std::vector<uint16_t> body(n * f.dim);
        for (uint64_t i = 0; i < n * f.dim; i++) body[i] = frk::F32ToBf16(rows[i]);
        ok = WriteAll(path, EncodeFileHeader(static_cast<uint8_t>(m_version), ext), body.data(), body.size() * 2);
    } else if (f.type == VectorFileType::FP8_ROWS) {
        // per-vector power-of-two scales first, then the scaled E4M3 rows
        Put<uint16_t>(ext, static_cast<uint16_t>(VectorFileType::FP8_ROWS));
        std::vector<uint8_t> body(n * 4 + n * f.dim);
        for (uint64_t i = 0; i < n; i++) {
            const float* r = rows + i * f.dim;
            float mx = 0.0f;
            for (uint32_t d = 0; d < f.dim; d++) mx = std::max(mx, std::fabs(r[d]));
            const float scale = frk::Fp8Scale(mx);
            std::memcpy(body.data() + 4 * i, &scale, 4);
            uint8_t* dst = body.data() + 4 * n + i * f.dim;
            for (uint32_t d = 0; d < f.dim; d++) dst[d] = frk::F32ToE4M3(r[d] / scale);
        }
        ok = WriteAll(path, EncodeFileHeader(static_cast<uint8_t>(m_version), ext), body.data(), body.size());
    } else {
        // ZN: blocks of Z docs; inside a block, the 16-dim slice j of doc o sits at halves
        // ((b * Z) * J + Z * j + o) * 16, a partial last slice zero-filled (vector_builder.cpp:46-66).
        // Every block is written whole, absent docs as zeros (the builder pads its last segment the
        // same way, vector_builder.cpp:162-170).
        const uint64_t Z = m_splitDocNumZn, J = (f.dim + 15) / 16;
        const uint64_t blocks = (n + Z - 1) / Z;
        std::vector<uint16_t> body(blocks * Z * J * 16, 0);
        for (uint64_t i = 0; i < n; i++) {
            const uint64_t b = i / Z, o = i % Z;
            const float* r = rows + i * f.dim;
            for (uint64_t j = 0; j < J; j++) {
                uint16_t* dst = body.data() + ((b * Z) * J + Z * j + o) * 16;
                for (uint64_t e = 0; e < 16 && j * 16 + e < f.dim; e++) dst[e] = Float32ToFloat16(r[j * 16 + e]);
            }
        }
        ok = WriteAll(path, EncodeFileHeader(static_cast<uint8_t>(m_version), ext), body.data(), body.size() * 2);
    }
    if (!ok) return Fail("cannot write " + path);
    return true;
}

bool IndexWriter::WritePostingSegment(const std::string& dir, const PostingField& f, uint32_t seg,
                                      std::vector<size_t>& cursor) {
    const uint32_t first = seg * m_docNumPerSegment;
    const uint32_t segDocs = std::min<uint32_t>(m_docNumPerSegment, m_docNum - first);
    const uint32_t end = first + segDocs;
    const std::string head = EncodeFileHeader(static_cast<uint8_t>(m_version), f.isMatrix ? kMatrixType : "1");
    // postings first (each padded to 4 bytes), token literals after them; tokens with no doc in
    // this segment are left out (inverted_builder.cpp:176-250)
    std::vector<uint64_t> hashes;
    std::vector<uint32_t> postOff, tokOff;
    std::string postings, literals, buf;
    std::vector<uint32_t> local;
    for (size_t t = 0; t < f.tokens.size(); t++) {
        const auto& d = f.docs[t];
        size_t& c = cursor[t];
        local.clear();
        while (c < d.size() && d[c] < end) local.push_back(d[c++] - first);
        if (local.empty()) continue;
        if (!SerializePosting(local, segDocs, f.isMatrix, f.threshold, f.packed, buf))
            return Fail("posting serialization failed for token " + f.tokens[t]);
        buf.resize((buf.size() + 3) / 4 * 4, '\0');
        hashes.push_back(Hash64(f.tokens[t]));
        postOff.push_back(static_cast<uint32_t>(postings.size()));
        postings += buf;
        tokOff.push_back(static_cast<uint32_t>(literals.size()));
        Put<uint32_t>(literals, static_cast<uint32_t>(f.tokens[t].size()));
        literals += f.tokens[t];
    }
    const uint64_t prefix = head.size() + 4 + 16ull * hashes.size();
    if (prefix + postings.size() + literals.size() > 0xFFFFFFFFull)
        return Fail("posting file of field " + f.name + " exceeds the 32-bit offsets of its token dictionary");
    std::string dict;
    dict.reserve(4 + 16 * hashes.size());
    Put<uint32_t>(dict, static_cast<uint32_t>(hashes.size()));
    for (size_t i = 0; i < hashes.size(); i++) {
        Put<uint64_t>(dict, hashes[i]);
        Put<uint32_t>(dict, static_cast<uint32_t>(prefix + postOff[i]));
        Put<uint32_t>(dict, static_cast<uint32_t>(prefix + postings.size() + tokOff[i]));
    }
    const std::string path = PathJoin(PathJoin(dir, f.name), SegmentFileName(m_version, seg, f.name, kPostingSuffix));
    const std::string body = dict + postings + literals;
    if (!WriteAll(path, head, body.data(), body.size())) return Fail("cannot write " + path);
    return true;
}

bool IndexWriter::Write(const std::string& dir) {
    if (!m_err.empty()) return false;
    if (!m_gdocids.empty() && m_gdocids.size() != m_docNum) return Fail("gdocid count differs from the doc count");
    if (!MakeDir(dir) || !MakeDir(PathJoin(dir, kIdMappingDir))) return Fail("cannot create " + dir);
    IndexMeta meta;
    meta.docNumPerSegment = m_docNumPerSegment;
    meta.splitDocNumZn = m_splitDocNumZn;
    meta.docNum = m_docNum;
    meta.segmentNum = m_segmentNum;
    for (const auto& f : m_vectors) {
        meta.vector.push_back(f.name);
        if (!MakeDir(PathJoin(dir, f.name))) return Fail("cannot create the directory of " + f.name);
    }
    for (const auto& f : m_postings) {
        meta.posting.push_back(f.name);
        if (!MakeDir(PathJoin(dir, f.name))) return Fail("cannot create the directory of " + f.name);
    }
    std::vector<std::vector<size_t>> cursors(m_postings.size());
    for (size_t p = 0; p < m_postings.size(); p++) cursors[p].assign(m_postings[p].tokens.size(), 0);
    for (uint32_t s = 0; s < m_segmentNum; s++) {
        if (!WriteIdSegment(dir, s)) return false;
        for (const auto& f : m_vectors)
            if (!WriteVectorSegment(dir, f, s)) return false;
        for (size_t p = 0; p < m_postings.size(); p++)
            if (!WritePostingSegment(dir, m_postings[p], s, cursors[p])) return false;
    }
    const std::string json = IndexMetaToJson(meta);
    const std::string metaPath = PathJoin(dir, std::string(kPoissonEngine) + "_" + std::to_string(m_version) + kMetaSuffix);
    if (!WriteAll(metaPath, "", json.data(), json.size())) return Fail("cannot write " + metaPath);
    return true;
}

}  // namespace NpuRetrieval
