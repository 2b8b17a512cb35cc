// format.h -- on-disk formats of an index directory (engine/format/* and engine/indexer/* in the
// reference): file header, posting layouts and their serializer, token hash, .meta JSON, file names.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_set>
#include <vector>

namespace NpuRetrieval {

// format/posting_layout.h:8 -- the 4-bit code in every posting head word.
enum class PostingLayout { DENSE_BITMAP = 0, SPARSE_BITMAP, DENSE_BYTEMAP, SPARSE_IDLIST, SPARSE_PACKED, UNKNOWN };

// core/constant_definition.h: file naming and header extension tags.
extern const char* const kPoissonEngine;   // "poissonengine"
extern const char* const kIdMappingDir;    // "id"
extern const char* const kVectorSuffix;    // ".vector"
extern const char* const kIdMappingSuffix; // ".idm"
extern const char* const kPostingSuffix;   // ".posting"
extern const char* const kMetaSuffix;      // ".meta"
extern const char* const kMatrixType;      // posting ext "0": matrix codec, "1": bitmap codec
extern const char* const kGdocidType;      // idm ext "0": u64 gdocids
extern const char* const kSdocidType;      // idm ext "1": string docids

// Token ids are Hash64 of the token text: MurmurHash3_x64_128, seed 0x20160713, out[0] ^ out[1]
// (core/hash/hash.cpp:10-16). Tag tokens are the decimal tag id (harness/query/filter_setup.cpp:100).
uint64_t Hash64(const char* data, uint32_t len);
inline uint64_t Hash64(const std::string& s) { return Hash64(s.data(), static_cast<uint32_t>(s.size())); }
uint64_t TagToken(uint64_t tag);

// File header [u8 version][5 reserved][ext bytes] where the ext carries its own u16 length prefix
// (format/file_header.cpp:21-38, indexer/file/file_writer.cpp:4-26).
bool DecodeFileHeader(const uint8_t* data, size_t size, uint8_t& version, std::string& extended, size_t& header_len);
std::string EncodeFileHeader(uint8_t version, const std::string& extended);

// Posting serializer (format/posting_serializer.cpp): head [layout:4][hitCount:28], u32 payload
// length, payload. docIds are local doc ids of one segment of docNum docs. SPARSE_IDLIST writes
// the ids in the given order.
PostingLayout SelectPostingLayout(uint32_t hits, uint32_t docNum, bool isMatrix, float threshold, bool packed);
bool SerializePosting(const std::vector<uint32_t>& docIds, uint32_t docNum, bool isMatrix, float threshold,
                      bool packed, std::string& out);

// .meta JSON (the reference's IndexMeta proto in JSON form, indexer/proto/index_meta.proto).
struct IndexMeta {
    uint32_t docNumPerSegment = 0;
    uint32_t splitDocNumZn = 0;
    std::vector<std::string> posting;
    std::vector<std::string> vector;
    uint32_t docNum = 0;
    uint32_t segmentNum = 0;
};
bool ParseIndexMeta(const std::string& json, IndexMeta& meta, std::string& err);
std::string IndexMetaToJson(const IndexMeta& meta);

// Vector file header extension: the reference writes [u16 len=2][u16 dim] for its FP16 ZN body
// (indexer/builder/segment/vector_builder.cpp:26-32). This build adds [u16 len=4][u16 dim][u16 dtype]
// with dtype 1 = FP32 row-major body ([docs][dim] float), 2 = BF16 rows ([docs][dim] bf16, nearest
// even), 3 = FP8 E4M3 rows with a per-vector scale ([docs] f32 scales, the smallest power of two s with
// max|v| / s <= 448, then [docs][dim] e4m3 bytes of v / s, nearest even, saturating).
enum class VectorFileType : uint16_t { FP16_ZN = 0, FP32_ROWS = 1, BF16_ROWS = 2, FP8_ROWS = 3 };

// Legacy FP16 codec of the reference's vector files: subnormals flush to zero both ways
// (core/number/trans_number.cpp:84-105).
float Float16ToFloat32(uint16_t h);
// FP32 -> FP16 as the reference builder stores vectors (core/number/trans_number.cpp:38-82): round to
// nearest even, no FP16 subnormals (magnitudes below 2^-14 become signed zero), finite overflow to
// infinity, NaN stays NaN with its top payload bits, and an infinite input becomes the NaN 0x7C01
// (sign kept), as the reference's exponent-255 branch does.
uint16_t Float32ToFloat16(float f);

std::string PathJoin(const std::string& a, const std::string& b);
std::string SegmentFileName(uint32_t version, uint32_t segment, const std::string& field, const char* suffix);

}  // namespace NpuRetrieval
