// data_table.cpp -- loading and building the device-resident index (see data_table.h).
#include "data_table.h"
#include "../frk_common.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>

#include "log.h"

namespace NpuRetrieval {

namespace {
bool ReadFile(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    f.seekg(0, std::ios::end);
    const std::streamoff n = f.tellg();
    f.seekg(0, std::ios::beg);
    out.resize(static_cast<size_t>(n));
    f.read(reinterpret_cast<char*>(out.data()), n);
    return static_cast<bool>(f);
}

uint32_t DimPad(uint32_t dim, uint32_t dtype) { return frk_store_dim(dim, dtype); }
}  // namespace

// ------------------------------------------------------------------------------------------ postings
PostingFieldData::~PostingFieldData() {
    if (m_device) (void)hipFree(m_device);
}

bool PostingFieldData::AddPosting(uint64_t token, uint32_t segment, const uint8_t* bytes, size_t size) {
    if (segment >= m_segmentNum || size < 8) {
        FRK_LOG_ERROR("posting for segment " << segment << " of " << m_segmentNum << " size " << size);
        return false;
    }
    uint32_t plen = 0;
    std::memcpy(&plen, bytes + 4, 4);
    if (8ull + plen > size) {
        FRK_LOG_ERROR("posting payload length " << plen << " exceeds its " << size << " bytes");
        return false;
    }
    // head at an offset = 8 (mod 16): the payload then starts 16-byte aligned
    if (m_host.empty()) m_host.resize(8, 0);
    const uint64_t off = m_host.size();
    const uint64_t padded = 8 + (plen + 15ull) / 16 * 16 + 8;
    m_host.resize(off + padded, 0);
    std::memcpy(m_host.data() + off, bytes, 8 + plen);
    m_pending.push_back({token, segment, off});
    auto& types = m_types[token];
    if (types.empty()) types.assign(m_segmentNum, 0xFF);
    types[segment] = static_cast<uint8_t>(bytes[3] >> 4);
    uint32_t head = 0;
    std::memcpy(&head, bytes, 4);
    m_hits[token] += head & 0x0FFFFFFFu;  // [layout:4][hitCount:28]
    return true;
}

uint64_t PostingFieldData::TokenHits(uint64_t token) const {
    auto it = m_hits.find(token);
    return it == m_hits.end() ? 0 : it->second;
}

bool PostingFieldData::FinishAdd(int32_t deviceId) {
    (void)deviceId;
    m_host.resize(m_host.size() + 64, 0);  // slack past the last posting
    m_arenaBytes = m_host.size();
    if (hipMalloc(&m_device, m_arenaBytes) != hipSuccess) {
        FRK_LOG_ERROR("hipMalloc " << m_arenaBytes << " B for posting field " << m_name);
        m_device = nullptr;
        return false;
    }
    if (hipMemcpy(m_device, m_host.data(), m_arenaBytes, hipMemcpyHostToDevice) != hipSuccess) {
        FRK_LOG_ERROR("posting H2D failed for " << m_name);
        return false;
    }
    const uint64_t base = reinterpret_cast<uint64_t>(m_device);
    for (const Pending& p : m_pending) {
        auto& v = m_addrs[p.token];
        if (v.empty()) v.assign(m_segmentNum, 0);
        // a posting with no hits is the all-zero operand (its weight is 0, posting_field_data.cpp:345-360)
        uint32_t head = 0;
        std::memcpy(&head, m_host.data() + p.offset, 4);
        v[p.segment] = (head & 0x0FFFFFFFu) ? base + p.offset : 0;
        if (head & 0x0FFFFFFFu) {
            uint32_t len = 0;
            std::memcpy(&len, m_host.data() + p.offset + 4, 4);
            m_describe[base + p.offset] = {static_cast<uint8_t>(head >> 28), len};
        }
    }
    m_pending.clear();
    m_pending.shrink_to_fit();
    std::vector<uint8_t>().swap(m_host);
    m_zeroAddrs.assign(m_segmentNum, 0);
    m_zeroTypes.assign(m_segmentNum, 0);
    return true;
}

const std::vector<uint64_t>* PostingFieldData::GetPostingDeviceAddrs(uint64_t token, bool& tokenExist) const {
    auto it = m_addrs.find(token);
    if (it == m_addrs.end()) {
        tokenExist = false;
        return &m_zeroAddrs;
    }
    tokenExist = true;
    return &it->second;
}

bool PostingFieldData::DescribePosting(uint64_t addr, uint8_t& layout, uint32_t& payloadBytes) const {
    auto it = m_describe.find(addr);
    if (it == m_describe.end()) return false;
    layout = it->second.layout;
    payloadBytes = it->second.payloadBytes;
    return true;
}

const std::vector<uint8_t>* PostingFieldData::GetPostingTypes(uint64_t token) const {
    auto it = m_types.find(token);
    return it == m_types.end() ? &m_zeroTypes : &it->second;
}

// ------------------------------------------------------------------------------------------ vectors
VectorFieldData::VectorFieldData(std::string name, uint32_t dim, uint64_t docNum)
    : m_name(std::move(name)), m_dim(dim), m_docNum(docNum) {}

VectorFieldData::~VectorFieldData() {
    for (void* p : m_tiles)
        if (p) (void)hipFree(p);
    for (float* p : m_scales)
        if (p) (void)hipFree(p);
    if (m_rows) (void)hipFree(m_rows);
    if (m_staging) (void)hipFree(m_staging);
}

uint32_t VectorFieldData::PaddedDim(uint32_t dtype) const { return DimPad(m_dim, dtype); }

bool VectorFieldData::Allocate(int32_t deviceId, uint32_t dtype_mask) {
    (void)deviceId;
    dtype_mask |= 1u << FRK_DTYPE_F32;
    for (uint32_t dt = 0; dt < 4; dt++) {
        if (!(dtype_mask & (1u << dt))) continue;
        const uint32_t dp = DimPad(m_dim, dt);
        m_bytes[dt] = frk_tiles_bytes(m_docNum, dp, dt);
        if (hipMalloc(&m_tiles[dt], m_bytes[dt]) != hipSuccess) {
            FRK_LOG_ERROR("hipMalloc " << m_bytes[dt] << " B for vector store dtype " << dt);
            m_tiles[dt] = nullptr;
            return false;
        }
        const bool scaled = dt == FRK_DTYPE_FP8E4M3 || dt == FRK_DTYPE_F16X2;
        if (scaled && hipMalloc(reinterpret_cast<void**>(&m_scales[dt]), m_docNum * 4 + 4) != hipSuccess) {
            m_scales[dt] = nullptr;
            return false;
        }
        m_store[dt].tiles = m_tiles[dt];
        m_store[dt].scales = m_scales[dt];
        m_store[dt].dtype = dt;
        m_store[dt].dim = dp;
        m_store[dt].n_docs = m_docNum;
    }
    if (dtype_mask & ((1u << FRK_DTYPE_BF16) | (1u << FRK_DTYPE_FP8E4M3))) {
        m_rowsBytes = m_docNum * m_store[FRK_DTYPE_F32].dim * sizeof(float);
        if (hipMalloc(reinterpret_cast<void**>(&m_rows), m_rowsBytes + 16) != hipSuccess) {
            FRK_LOG_ERROR("hipMalloc " << m_rowsBytes << " B for the row-major FP32 copy of " << m_name);
            m_rows = nullptr;
            return false;
        }
        if (hipMemset(m_rows, 0, m_rowsBytes + 16) != hipSuccess) return false;  // the padding dims stay 0
    }
    return true;
}

bool VectorFieldData::SetRows(const float* rows, uint64_t first, uint64_t n, bool rows_on_device, void* stream) {
    if (first % 32 != 0 || first + n > m_docNum || (n % 32 != 0 && first + n != m_docNum)) {
        FRK_LOG_ERROR("SetRows [" << first << ", +" << n << ") must start and end on 32-doc blocks of " << m_docNum);
        return false;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto tile = [&](const float* d_rows, uint64_t at, uint64_t cnt) {
        for (uint32_t dt = 0; dt < 4; dt++) {
            if (!m_tiles[dt]) continue;
            if (frk_tile_rows_device(d_rows, cnt, m_dim, at, dt, m_store[dt].dim, m_tiles[dt], m_scales[dt], s) != 0) {
                FRK_LOG_ERROR("tiling failed: " << frk_last_error());
                return false;
            }
        }
        if (m_rows) {
            const uint32_t dp = m_store[FRK_DTYPE_F32].dim;
            if (hipMemcpy2DAsync(m_rows + at * dp, dp * sizeof(float), d_rows, m_dim * sizeof(float),
                                 m_dim * sizeof(float), cnt, hipMemcpyDeviceToDevice, s) != hipSuccess) {
                FRK_LOG_ERROR("row copy failed for " << m_name);
                return false;
            }
        }
        return true;
    };
    if (rows_on_device) return tile(rows, first, n);
    const uint64_t chunk = std::max<uint64_t>(32, (64ull << 20) / (4ull * m_dim) / 32 * 32);
    if (!m_staging) {
        m_stagingRows = chunk;
        if (hipMalloc(reinterpret_cast<void**>(&m_staging), chunk * m_dim * 4) != hipSuccess) {
            m_staging = nullptr;
            return false;
        }
    }
    for (uint64_t at = 0; at < n; at += m_stagingRows) {
        const uint64_t cnt = std::min<uint64_t>(m_stagingRows, n - at);
        if (hipMemcpyAsync(m_staging, rows + at * m_dim, cnt * m_dim * 4, hipMemcpyHostToDevice, s) != hipSuccess)
            return false;
        if (!tile(m_staging, first + at, cnt)) return false;
        if (hipStreamSynchronize(s) != hipSuccess) return false;  // the staging buffer is reused
    }
    return true;
}

const frk_doc_store* VectorFieldData::Store(uint32_t dtype) const {
    if (dtype > FRK_DTYPE_F16X2 || !m_tiles[dtype]) return nullptr;
    return &m_store[dtype];
}

uint64_t VectorFieldData::DeviceBytes() const {
    uint64_t b = m_rowsBytes;
    for (uint32_t dt = 0; dt < 4; dt++) b += m_bytes[dt] + (m_scales[dt] ? m_docNum * 4 : 0);
    return b;
}

// ------------------------------------------------------------------------------------------ table
bool DataTable::GetPostingFieldData(const std::string& fieldName, PostingFieldData*& fieldData) const {
    auto it = m_postings.find(fieldName);
    fieldData = it == m_postings.end() ? nullptr : it->second.get();
    return fieldData != nullptr;
}

bool DataTable::GetVectorFieldData(const std::string& fieldName, VectorFieldData*& fieldData) const {
    auto it = m_vectors.find(fieldName);
    fieldData = it == m_vectors.end() ? nullptr : it->second.get();
    return fieldData != nullptr;
}

std::string DataTable::Describe() const {
    std::ostringstream o;
    o << "docs=" << m_docNum << " segments=" << m_segmentNum << " docNumPerSegment=" << m_docNumPerSegment;
    for (const auto& kv : m_vectors) o << " vector:" << kv.first << "(dim " << kv.second->Dimension() << ")";
    for (const auto& kv : m_postings) o << " posting:" << kv.first << "(" << kv.second->TokenNum() << " tokens)";
    return o.str();
}

bool DataTable::LoadData(int32_t deviceId, const std::string& dataDir, uint32_t vector_dtype_mask) {
    namespace fs = std::filesystem;
    std::error_code ec;
    if (!fs::is_directory(dataDir, ec)) {
        FRK_LOG_ERROR("index directory not found: " << dataDir);
        return false;
    }
    if (hipSetDevice(deviceId) != hipSuccess) {
        FRK_LOG_ERROR("hipSetDevice(" << deviceId << ") failed");
        return false;
    }
    m_deviceId = deviceId;
    // the single *.meta file; its name carries the version: poissonengine_<v>.meta (data_table.cpp:147-186)
    std::string metaPath, stem;
    for (const auto& e : fs::directory_iterator(dataDir, ec)) {
        if (e.path().extension().string() == kMetaSuffix) {
            metaPath = e.path().string();
            stem = e.path().stem().string();
            break;
        }
    }
    if (metaPath.empty()) {
        FRK_LOG_ERROR("meta file not found in " << dataDir);
        return false;
    }
    const size_t us = stem.find('_');
    if (us == std::string::npos) {
        FRK_LOG_ERROR("meta file name has no version: " << stem);
        return false;
    }
    try {
        m_version = static_cast<uint32_t>(std::stoul(stem.substr(us + 1)));
    } catch (...) {
        FRK_LOG_ERROR("meta file name has no version: " << stem);
        return false;
    }
    std::vector<uint8_t> raw;
    if (!ReadFile(metaPath, raw)) return false;
    IndexMeta meta;
    std::string err;
    if (!ParseIndexMeta(std::string(raw.begin(), raw.end()), meta, err)) {
        FRK_LOG_ERROR(err);
        return false;
    }
    m_segmentNum = meta.segmentNum;
    m_docNum = meta.docNum;
    m_docNumPerSegment = meta.docNumPerSegment;
    m_splitDocNumZn = meta.splitDocNumZn;
    if (m_docNumPerSegment == 0 || m_docNumPerSegment % 8 != 0) {
        FRK_LOG_ERROR("docNumPerSegment " << m_docNumPerSegment << " must be a positive multiple of 8");
        return false;
    }
    if (static_cast<uint64_t>(m_segmentNum) * m_docNumPerSegment < m_docNum) {
        FRK_LOG_ERROR("segments do not cover the docs");
        return false;
    }
    if (!LoadDocIds(dataDir)) return false;
    for (const auto& f : meta.vector)
        if (!LoadVectorField(dataDir, f, vector_dtype_mask)) return false;
    for (const auto& f : meta.posting) {
        if (!LoadPostingField(dataDir, f)) return false;
        m_postingFields.insert(f);
    }
    FRK_LOG_INFO("[Load] device " << deviceId << ": " << Describe());
    return true;
}

bool DataTable::LoadDocIds(const std::string& dataDir) {
    m_docIdMapping = std::make_unique<DocIdMapping>();
    auto& ids = m_docIdMapping->MutableGDocIds();
    ids.reserve(m_docNum);
    for (uint32_t s = 0; s < m_segmentNum; s++) {
        const std::string path = PathJoin(PathJoin(dataDir, kIdMappingDir), SegmentFileName(m_version, s, "", kIdMappingSuffix));
        std::vector<uint8_t> raw;
        if (!ReadFile(path, raw)) {
            FRK_LOG_ERROR("id mapping file missing: " << path);
            return false;
        }
        uint8_t version = 0;
        std::string ext;
        size_t hl = 0;
        if (!DecodeFileHeader(raw.data(), raw.size(), version, ext, hl) || version != m_version) {
            FRK_LOG_ERROR("bad id mapping header: " << path);
            return false;
        }
        if (ext == kSdocidType) {
            FRK_LOG_ERROR("string doc ids (sdocid) are not supported: " << path);
            return false;
        }
        const size_t n = (raw.size() - hl) / 8;
        const size_t at = ids.size();
        ids.resize(at + n);
        std::memcpy(ids.data() + at, raw.data() + hl, n * 8);
    }
    if (ids.size() != m_docNum) {
        FRK_LOG_ERROR("id mapping holds " << ids.size() << " ids for " << m_docNum << " docs");
        return false;
    }
    return true;
}

bool DataTable::LoadVectorField(const std::string& dataDir, const std::string& field, uint32_t dtype_mask) {
    std::unique_ptr<VectorFieldData> vf;
    std::vector<float> rows;
    for (uint32_t s = 0; s < m_segmentNum; s++) {
        const std::string path = PathJoin(PathJoin(dataDir, field), SegmentFileName(m_version, s, field, kVectorSuffix));
        std::vector<uint8_t> raw;
        if (!ReadFile(path, raw)) {
            FRK_LOG_ERROR("vector file missing: " << path);
            return false;
        }
        uint8_t version = 0;
        std::string ext;
        size_t hl = 0;
        if (!DecodeFileHeader(raw.data(), raw.size(), version, ext, hl) || (ext.size() != 2 && ext.size() != 4)) {
            FRK_LOG_ERROR("bad vector header: " << path);
            return false;
        }
        uint16_t dim = 0, type = 0;
        std::memcpy(&dim, ext.data(), 2);
        if (ext.size() == 4) std::memcpy(&type, ext.data() + 2, 2);
        if (dim == 0) return false;
        if (!vf) {
            vf = std::make_unique<VectorFieldData>(field, dim, m_docNum);
            if (!vf->Allocate(m_deviceId, dtype_mask)) return false;
        } else if (dim != vf->Dimension()) {
            FRK_LOG_ERROR("dimension changes between segments of " << field);
            return false;
        }
        const uint64_t first = static_cast<uint64_t>(s) * m_docNumPerSegment;
        const uint64_t n = std::min<uint64_t>(m_docNumPerSegment, m_docNum - first);
        rows.assign(n * dim, 0.0f);
        const uint8_t* body = raw.data() + hl;
        const size_t bodyLen = raw.size() - hl;
        if (type == static_cast<uint16_t>(VectorFileType::FP32_ROWS)) {
            if (bodyLen < n * dim * 4) {
                FRK_LOG_ERROR("short FP32 vector body: " << path);
                return false;
            }
            std::memcpy(rows.data(), body, n * dim * 4);
        } else if (type == static_cast<uint16_t>(VectorFileType::BF16_ROWS)) {
            if (bodyLen < n * dim * 2) {
                FRK_LOG_ERROR("short BF16 vector body: " << path);
                return false;
            }
            const uint16_t* h = reinterpret_cast<const uint16_t*>(body);
            for (uint64_t i = 0; i < n * dim; i++) rows[i] = frk::Bf16ToF32(h[i]);
        } else if (type == static_cast<uint16_t>(VectorFileType::FP8_ROWS)) {
            // the stored values, scale x e4m3: a low-precision store built from them quantizes each
            // back to the same value (the scale is a power of two and the E4M3 grid scales with it)
            if (bodyLen < n * 4 + n * dim) {
                FRK_LOG_ERROR("short FP8 vector body: " << path);
                return false;
            }
            for (uint64_t i = 0; i < n; i++) {
                float scale = 0.0f;
                std::memcpy(&scale, body + 4 * i, 4);
                const uint8_t* b = body + 4 * n + i * dim;
                for (uint32_t d = 0; d < dim; d++) rows[i * dim + d] = scale * frk::E4M3ToF32(b[d]);
            }
        } else if (type != static_cast<uint16_t>(VectorFileType::FP16_ZN)) {
            FRK_LOG_ERROR("unknown vector body type " << type << ": " << path);
            return false;
        } else {
            // ZN FP16 blocks of Z docs; inside a block, dim block j of doc i at ((j * Z) + i) * 16
            // halves (indexer/builder/segment/vector_builder.cpp:46-66). The tail block may be cut short.
            const uint32_t Z = m_splitDocNumZn;
            const uint32_t J = (dim + 15) / 16;
            if (Z == 0) return false;
            const uint16_t* h = reinterpret_cast<const uint16_t*>(body);
            const size_t nh = bodyLen / 2;
            for (uint64_t i = 0; i < n; i++) {
                const uint64_t b = i / Z, o = i % Z;
                for (uint32_t j = 0; j < J; j++) {
                    const uint64_t base = ((b * Z) * J + static_cast<uint64_t>(Z) * j + o) * 16;
                    for (uint32_t e = 0; e < 16 && j * 16 + e < dim; e++)
                        rows[i * dim + j * 16 + e] = base + e < nh ? Float16ToFloat32(h[base + e]) : 0.0f;
                }
            }
        }
        if (!vf->SetRows(rows.data(), first, n, false, nullptr)) return false;
    }
    if (hipDeviceSynchronize() != hipSuccess) return false;
    m_vectors[field] = std::move(vf);
    return true;
}

bool DataTable::LoadPostingField(const std::string& dataDir, const std::string& field) {
    auto pf = std::make_unique<PostingFieldData>(field, m_segmentNum);
    for (uint32_t s = 0; s < m_segmentNum; s++) {
        const std::string path = PathJoin(PathJoin(dataDir, field), SegmentFileName(m_version, s, field, kPostingSuffix));
        std::vector<uint8_t> raw;
        if (!ReadFile(path, raw)) {
            FRK_LOG_ERROR("posting file missing: " << path);
            return false;
        }
        uint8_t version = 0;
        std::string ext;
        size_t hl = 0;
        if (!DecodeFileHeader(raw.data(), raw.size(), version, ext, hl) || raw.size() < hl + 4) {
            FRK_LOG_ERROR("bad posting header: " << path);
            return false;
        }
        pf->SetMatrixCodec(ext == kMatrixType);
        uint32_t tokenNum = 0;
        std::memcpy(&tokenNum, raw.data() + hl, 4);
        const size_t dict = hl + 4;
        if (raw.size() < dict + 16ull * tokenNum) return false;
        // {u64 token hash, u32 posting offset, u32 token offset}, offsets file-absolute
        // (indexer/builder/segment/inverted_builder.cpp:193-250; index/posting/posting_field_data.cpp:60-118)
        for (uint32_t t = 0; t < tokenNum; t++) {
            uint64_t token;
            uint32_t poff, toff;
            std::memcpy(&token, raw.data() + dict + 16ull * t, 8);
            std::memcpy(&poff, raw.data() + dict + 16ull * t + 8, 4);
            std::memcpy(&toff, raw.data() + dict + 16ull * t + 12, 4);
            if (poff >= toff || poff + 8ull > raw.size()) {
                FRK_LOG_ERROR("bad posting offsets in " << path);
                return false;
            }
            uint32_t plen = 0;
            std::memcpy(&plen, raw.data() + poff + 4, 4);
            if (poff + 8ull + plen > raw.size()) {
                FRK_LOG_ERROR("posting overruns its file: " << path);
                return false;
            }
            if (!pf->AddPosting(token, s, raw.data() + poff, 8 + plen)) return false;
        }
    }
    if (!pf->FinishAdd(m_deviceId)) return false;
    m_postings[field] = std::move(pf);
    return true;
}

// ------------------------------------------------------------------------------------------ builder
TableBuilder::TableBuilder(int32_t deviceId, uint32_t docNum, uint32_t docNumPerSegment, uint32_t version) {
    m_table = std::make_shared<DataTable>();
    m_table->m_deviceId = deviceId;
    m_table->m_docNum = docNum;
    m_table->m_docNumPerSegment = docNumPerSegment;
    m_table->m_segmentNum = docNumPerSegment ? (docNum + docNumPerSegment - 1) / docNumPerSegment : 0;
    m_table->m_splitDocNumZn = 32;
    m_table->m_version = version;
    m_table->m_docIdMapping = std::make_unique<DocIdMapping>();
    if (hipSetDevice(deviceId) != hipSuccess) m_err = "hipSetDevice failed";
    if (docNumPerSegment == 0 || docNumPerSegment % 8 != 0) m_err = "docNumPerSegment must be a positive multiple of 8";
}

bool TableBuilder::AddVectorField(const std::string& name, uint32_t dim, uint32_t dtype_mask) {
    auto vf = std::make_unique<VectorFieldData>(name, dim, m_table->m_docNum);
    if (!vf->Allocate(m_table->m_deviceId, dtype_mask)) {
        m_err = "vector store allocation failed";
        return false;
    }
    m_table->m_vectors[name] = std::move(vf);
    return true;
}

bool TableBuilder::SetVectors(const std::string& name, const float* rows, uint64_t first, uint64_t n,
                              bool rows_on_device) {
    VectorFieldData* vf = nullptr;
    if (!m_table->GetVectorFieldData(name, vf)) {
        m_err = "no vector field " + name;
        return false;
    }
    if (!vf->SetRows(rows, first, n, rows_on_device, nullptr)) {
        m_err = "SetRows failed";
        return false;
    }
    return true;
}

bool TableBuilder::AddPostingField(const std::string& name, bool isMatrix, float densityThreshold, bool packed) {
    m_table->m_postings[name] = std::make_unique<PostingFieldData>(name, m_table->m_segmentNum);
    m_table->m_postings[name]->SetMatrixCodec(isMatrix);
    m_table->m_postingFields.insert(name);
    m_cfg[name] = {isMatrix, densityThreshold, packed};
    return true;
}

bool TableBuilder::AddTokenPostings(const std::string& field, uint64_t token, const uint32_t* docIds, uint64_t n) {
    auto it = m_table->m_postings.find(field);
    if (it == m_table->m_postings.end()) {
        m_err = "no posting field " + field;
        return false;
    }
    const PostingCfg& cfg = m_cfg[field];
    const uint32_t dnps = m_table->m_docNumPerSegment;
    std::vector<uint32_t> local;
    std::string buf;
    uint64_t i = 0;
    while (i < n) {
        const uint32_t seg = docIds[i] / dnps;
        if (seg >= m_table->m_segmentNum) {
            m_err = "doc id beyond the table";
            return false;
        }
        local.clear();
        while (i < n && docIds[i] / dnps == seg) local.push_back(docIds[i++] - seg * dnps);
        const uint32_t segDocs = std::min<uint32_t>(dnps, m_table->m_docNum - seg * dnps);
        // empty postings are never written (inverted_builder.cpp:180-189); absent segments read as zero
        if (!SerializePosting(local, segDocs, cfg.isMatrix, cfg.threshold, cfg.packed, buf)) {
            m_err = "posting serialization failed";
            return false;
        }
        if (!it->second->AddPosting(token, seg, reinterpret_cast<const uint8_t*>(buf.data()), buf.size())) {
            m_err = "AddPosting failed";
            return false;
        }
    }
    return true;
}

void TableBuilder::SetGDocIds(const uint64_t* ids, uint64_t n) {
    m_table->m_docIdMapping->MutableGDocIds().assign(ids, ids + n);
}

std::shared_ptr<DataTable> TableBuilder::Finish() {
    if (!m_err.empty()) return nullptr;
    auto& ids = m_table->m_docIdMapping->MutableGDocIds();
    if (ids.empty()) {
        ids.resize(m_table->m_docNum);
        for (uint32_t i = 0; i < m_table->m_docNum; i++) ids[i] = i;
    }
    for (auto& kv : m_table->m_postings)
        if (!kv.second->FinishAdd(m_table->m_deviceId)) {
            m_err = "posting upload failed";
            return nullptr;
        }
    if (hipDeviceSynchronize() != hipSuccess) {
        m_err = "device synchronize failed";
        return nullptr;
    }
    return m_table;
}

}  // namespace NpuRetrieval
