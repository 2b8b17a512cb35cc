// data_table.h -- the device-resident index: postings, vectors and the doc id mapping of one shard,
// loaded from an index directory (DataTable::LoadData, engine/index/data_table.h:10-92) or built in
// memory from arrays (TableBuilder). Read-only after load; shared by searchers through shared_ptr.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../../include/frk_capi.h"
#include "format.h"

namespace NpuRetrieval {

// One posting field: every (token, segment) posting of every segment in one device arena, placed so
// each payload is 16-byte aligned and zero-padded to a 16-byte multiple (the filter kernel reads dense
// payloads as whole u64 words). Mirrors PostingFieldData (index/posting/posting_field_data.cpp:16-522):
// a token absent from a segment, or absent from the shard, resolves to the all-zero operand, encoded
// here as a 0 address.
class PostingFieldData {
   public:
    PostingFieldData(std::string name, uint32_t segmentNum) : m_name(std::move(name)), m_segmentNum(segmentNum) {}
    ~PostingFieldData();
    PostingFieldData(const PostingFieldData&) = delete;
    PostingFieldData& operator=(const PostingFieldData&) = delete;

    // Host staging: the serialized bytes of one posting (head + payload).
    bool AddPosting(uint64_t token, uint32_t segment, const uint8_t* bytes, size_t size);
    bool FinishAdd(int32_t deviceId);  // one H2D of the arena
    // Docs that carry the token, summed over segments (the hitCount of each posting head).
    uint64_t TokenHits(uint64_t token) const;

    // Per-segment device addresses of a token's postings (0 = all-zero); the shared all-zero vector
    // for an absent token (tokenExist = false), as posting_field_data.cpp:487-495.
    const std::vector<uint64_t>* GetPostingDeviceAddrs(uint64_t token, bool& tokenExist) const;
    const std::vector<uint8_t>* GetPostingTypes(uint64_t token) const;  // layout code per segment
    // Layout and payload bytes of the posting whose head sits at device address addr (false for 0 and
    // for addresses that are not posting heads of this field); for diagnostics.
    bool DescribePosting(uint64_t addr, uint8_t& layout, uint32_t& payloadBytes) const;
    const std::string& Name() const { return m_name; }
    size_t TokenNum() const { return m_addrs.size(); }
    uint64_t ArenaBytes() const { return m_arenaBytes; }
    bool IsMatrixCodec() const { return m_isMatrix; }
    void SetMatrixCodec(bool m) { m_isMatrix = m; }

   private:
    struct Pending {
        uint64_t token;
        uint32_t segment;
        uint64_t offset;
    };
    std::string m_name;
    uint32_t m_segmentNum;
    bool m_isMatrix = false;
    std::vector<uint8_t> m_host;  // staging arena
    std::vector<Pending> m_pending;
    std::unordered_map<uint64_t, std::vector<uint64_t>> m_addrs;
    std::unordered_map<uint64_t, std::vector<uint8_t>> m_types;
    std::unordered_map<uint64_t, uint64_t> m_hits;
    struct Described {
        uint8_t layout;
        uint32_t payloadBytes;
    };
    std::unordered_map<uint64_t, Described> m_describe;  // device head address -> posting shape
    std::vector<uint64_t> m_zeroAddrs;
    std::vector<uint8_t> m_zeroTypes;
    uint8_t* m_device = nullptr;
    uint64_t m_arenaBytes = 0;
};

// One vector field: FP32 tiles (the plain path) plus optional BF16 and FP8+scale tiles, all in the
// MFMA-native doc tiling (frk_capi.h), and with a low-precision store a row-major FP32 copy that the
// FP32 rescoring reads (one contiguous row per candidate). Replaces VectorFieldData's ZN FP16 body
// (index/vector/vector_field_data.cpp:10-75).
class VectorFieldData {
   public:
    VectorFieldData(std::string name, uint32_t dim, uint64_t docNum);
    ~VectorFieldData();
    VectorFieldData(const VectorFieldData&) = delete;
    VectorFieldData& operator=(const VectorFieldData&) = delete;

    // Allocates the stores for the dtypes in `dtype_mask` (bit FRK_DTYPE_*; FP32 is always kept).
    bool Allocate(int32_t deviceId, uint32_t dtype_mask);
    // Writes rows [first, first + n) (first a multiple of 32, except that n may end at docNum);
    // `rows` is host or device memory ([n][dim] FP32).
    bool SetRows(const float* rows, uint64_t first, uint64_t n, bool rows_on_device, void* stream);
    const frk_doc_store* Store(uint32_t dtype) const;
    // Row-major FP32 rows, row stride PaddedDim(FRK_DTYPE_F32) (zero padding), or null.
    const float* RowsF32() const { return m_rows; }
    uint32_t Dimension() const { return m_dim; }
    uint32_t PaddedDim(uint32_t dtype) const;
    uint64_t DocNum() const { return m_docNum; }
    const std::string& Name() const { return m_name; }
    uint64_t DeviceBytes() const;

   private:
    std::string m_name;
    uint32_t m_dim;
    uint64_t m_docNum;
    frk_doc_store m_store[4] = {};  // by FRK_DTYPE_*
    void* m_tiles[4] = {nullptr, nullptr, nullptr, nullptr};
    float* m_scales[4] = {nullptr, nullptr, nullptr, nullptr};  // per-doc scales (FP8, F16X2)
    float* m_rows = nullptr;
    uint64_t m_rowsBytes = 0;
    float* m_staging = nullptr;  // device rows staging for host uploads
    uint64_t m_stagingRows = 0;
    uint64_t m_bytes[4] = {0, 0, 0, 0};
};

// ldocid -> gdocid (index/docid/doc_id_mapping.cpp:37-166).
class DocIdMapping {
   public:
    bool GetGDocId(uint32_t ldocid, uint64_t& gdocid) const {
        if (ldocid >= m_gDocIds.size()) return false;
        gdocid = m_gDocIds[ldocid];
        return true;
    }
    const std::vector<uint64_t>& GDocIds() const { return m_gDocIds; }
    std::vector<uint64_t>& MutableGDocIds() { return m_gDocIds; }

   private:
    std::vector<uint64_t> m_gDocIds;
};

class DataTable {
   public:
    DataTable() = default;
    ~DataTable() = default;
    DataTable(const DataTable&) = delete;
    DataTable& operator=(const DataTable&) = delete;

    // Loads <dataDir>/poissonengine_<v>.meta, id/, and one directory per field (data_table.cpp:35-124).
    // vector_dtype_mask selects the extra low-precision stores to build (bits 1 << FRK_DTYPE_*).
    bool LoadData(int32_t deviceId, const std::string& dataDir, uint32_t vector_dtype_mask = 0);

    bool GetPostingFieldData(const std::string& fieldName, PostingFieldData*& fieldData) const;
    bool GetVectorFieldData(const std::string& fieldName, VectorFieldData*& fieldData) const;
    DocIdMapping* GetDocIdMapping() const { return m_docIdMapping.get(); }
    uint32_t GetSegmentNum() const { return m_segmentNum; }
    uint32_t GetSegmentLength() const { return m_docNumPerSegment / 16; }
    uint32_t GetSegmentByteSize() const { return m_docNumPerSegment / 8; }
    uint32_t GetDocNum() const { return m_docNum; }
    uint32_t GetScoreExtendDocNum() const { return (m_docNum + 31) / 32 * 32; }
    uint32_t GetDocNumPerSegment() const { return m_docNumPerSegment; }
    uint8_t* GetDocLocationInDevice() const { return nullptr; }  // the fused scorer derives doc ids itself
    uint32_t GetSplitDocNumZn() const { return m_splitDocNumZn; }
    const std::unordered_set<std::string>& GetPostingFields() const { return m_postingFields; }
    int32_t GetDeviceId() const { return m_deviceId; }
    // Layout and payload bytes of a posting head address of any posting field (diagnostics).
    bool DescribePosting(uint64_t addr, uint8_t& layout, uint32_t& payloadBytes) const {
        for (const auto& kv : m_postings)
            if (kv.second->DescribePosting(addr, layout, payloadBytes)) return true;
        return false;
    }
    uint32_t GetVersion() const { return m_version; }
    std::string Describe() const;

   private:
    friend class TableBuilder;
    bool LoadVectorField(const std::string& dataDir, const std::string& field, uint32_t dtype_mask);
    bool LoadPostingField(const std::string& dataDir, const std::string& field);
    bool LoadDocIds(const std::string& dataDir);

    std::unordered_map<std::string, std::unique_ptr<PostingFieldData>> m_postings;
    std::unordered_map<std::string, std::unique_ptr<VectorFieldData>> m_vectors;
    std::unique_ptr<DocIdMapping> m_docIdMapping;
    uint32_t m_segmentNum = 0;
    uint32_t m_docNum = 0;
    uint32_t m_docNumPerSegment = 0;
    uint32_t m_splitDocNumZn = 0;
    uint32_t m_version = 0;
    std::unordered_set<std::string> m_postingFields;
    int32_t m_deviceId = -1;
};

// In-memory construction of a DataTable (the offline builder's output without the files): posting
// lists are serialized per segment with the reference's layout rules and uploaded; vectors are
// tiled on the device.
class TableBuilder {
   public:
    TableBuilder(int32_t deviceId, uint32_t docNum, uint32_t docNumPerSegment, uint32_t version = 0);
    bool AddVectorField(const std::string& name, uint32_t dim, uint32_t dtype_mask);
    bool SetVectors(const std::string& name, const float* rows, uint64_t first, uint64_t n, bool rows_on_device);
    bool AddPostingField(const std::string& name, bool isMatrix, float densityThreshold, bool packed);
    // One token's posting list over the whole shard: sorted global local-doc ids.
    bool AddTokenPostings(const std::string& field, uint64_t token, const uint32_t* docIds, uint64_t n);
    void SetGDocIds(const uint64_t* ids, uint64_t n);
    std::shared_ptr<DataTable> Finish();
    const std::string& Error() const { return m_err; }

   private:
    struct PostingCfg {
        bool isMatrix;
        float threshold;
        bool packed;
    };
    std::shared_ptr<DataTable> m_table;
    std::unordered_map<std::string, PostingCfg> m_cfg;
    std::string m_err;
};

}  // namespace NpuRetrieval
