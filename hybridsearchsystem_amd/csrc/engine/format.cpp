// format.cpp -- index file formats (see format.h for the reference files each part follows).
#include "format.h"

#include <cstring>
#include <map>
#include <memory>
#include <sstream>

namespace NpuRetrieval {

const char* const kPoissonEngine = "poissonengine";
const char* const kIdMappingDir = "id";
const char* const kVectorSuffix = ".vector";
const char* const kIdMappingSuffix = ".idm";
const char* const kPostingSuffix = ".posting";
const char* const kMetaSuffix = ".meta";
const char* const kMatrixType = "0";
const char* const kGdocidType = "0";
const char* const kSdocidType = "1";

namespace {
inline uint64_t Rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

inline uint64_t Fmix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}
}  // namespace

// MurmurHash3_x64_128 (Austin Appleby, public domain), folded to 64 bits as the reference does.
uint64_t Hash64(const char* key, uint32_t len) {
    const uint8_t* data = reinterpret_cast<const uint8_t*>(key);
    const uint64_t c1 = 0x87c37b91114253d5ULL, c2 = 0x4cf5ad432745937fULL;
    uint64_t h1 = 0x20160713u, h2 = 0x20160713u;
    const uint32_t nblocks = len / 16;
    for (uint32_t i = 0; i < nblocks; i++) {
        uint64_t k1, k2;
        std::memcpy(&k1, data + 16 * i, 8);
        std::memcpy(&k2, data + 16 * i + 8, 8);
        k1 *= c1;
        k1 = Rotl(k1, 31);
        k1 *= c2;
        h1 ^= k1;
        h1 = Rotl(h1, 27) + h2;
        h1 = h1 * 5 + 0x52dce729;
        k2 *= c2;
        k2 = Rotl(k2, 33);
        k2 *= c1;
        h2 ^= k2;
        h2 = Rotl(h2, 31) + h1;
        h2 = h2 * 5 + 0x38495ab5;
    }
    const uint8_t* tail = data + 16 * nblocks;
    uint64_t k1 = 0, k2 = 0;
    const uint32_t rem = len & 15;
    for (uint32_t i = rem; i > 8; i--) k2 ^= static_cast<uint64_t>(tail[i - 1]) << (8 * (i - 9));
    if (rem > 8) {
        k2 *= c2;
        k2 = Rotl(k2, 33);
        k2 *= c1;
        h2 ^= k2;
    }
    for (uint32_t i = rem < 8 ? rem : 8; i > 0; i--) k1 ^= static_cast<uint64_t>(tail[i - 1]) << (8 * (i - 1));
    if (rem > 0) {
        k1 *= c1;
        k1 = Rotl(k1, 31);
        k1 *= c2;
        h1 ^= k1;
    }
    h1 ^= len;
    h2 ^= len;
    h1 += h2;
    h2 += h1;
    h1 = Fmix(h1);
    h2 = Fmix(h2);
    h1 += h2;
    h2 += h1;
    return h1 ^ h2;
}

uint64_t TagToken(uint64_t tag) { return Hash64(std::to_string(tag)); }

bool DecodeFileHeader(const uint8_t* data, size_t size, uint8_t& version, std::string& extended, size_t& header_len) {
    if (size < 8) return false;
    version = data[0];
    uint16_t ext_len = 0;
    std::memcpy(&ext_len, data + 6, 2);
    if (size < 8u + ext_len) return false;
    extended.assign(reinterpret_cast<const char*>(data + 8), ext_len);
    header_len = 8u + ext_len;
    return true;
}

std::string EncodeFileHeader(uint8_t version, const std::string& extended) {
    std::string h(6, '\0');
    h[0] = static_cast<char>(version);
    uint16_t n = static_cast<uint16_t>(extended.size());
    h.append(reinterpret_cast<const char*>(&n), 2);
    h += extended;
    return h;
}

PostingLayout SelectPostingLayout(uint32_t hits, uint32_t docNum, bool isMatrix, float threshold, bool packed) {
    if (docNum == 0) return PostingLayout::DENSE_BITMAP;
    const float density = static_cast<float>(hits * 1.0 / docNum);
    if (density < threshold) {
        if (isMatrix) return PostingLayout::SPARSE_IDLIST;
        return packed ? PostingLayout::SPARSE_PACKED : PostingLayout::SPARSE_BITMAP;
    }
    return isMatrix ? PostingLayout::DENSE_BYTEMAP : PostingLayout::DENSE_BITMAP;
}

namespace {
template <typename T>
void Put(std::string& out, T v) {
    out.append(reinterpret_cast<const char*>(&v), sizeof(T));
}
}  // namespace

bool SerializePosting(const std::vector<uint32_t>& docIds, uint32_t docNum, bool isMatrix, float threshold, bool packed,
                      std::string& out) {
    const uint32_t hits = static_cast<uint32_t>(docIds.size());
    if (hits > 0x0FFFFFFFu) return false;
    const PostingLayout layout = SelectPostingLayout(hits, docNum, isMatrix, threshold, packed);
    const uint32_t units = (docNum + 15) / 16;
    const uint32_t head = (static_cast<uint32_t>(layout) << 28) | hits;
    out.clear();
    if (layout == PostingLayout::SPARSE_IDLIST) {
        Put(out, head);
        Put(out, hits * 4u);
        for (uint32_t id : docIds) Put(out, id);
        return true;
    }
    std::vector<uint16_t> masks(units, 0);
    for (uint32_t id : docIds)
        if (id / 16 < units) masks[id / 16] |= static_cast<uint16_t>(1u << (id % 16));
    Put(out, head);
    if (layout == PostingLayout::DENSE_BITMAP) {
        Put(out, units * 2u);
        for (uint16_t m : masks) Put(out, m);
    } else if (layout == PostingLayout::DENSE_BYTEMAP) {
        Put(out, docNum);
        for (uint32_t d = 0; d < docNum; d++) Put(out, static_cast<uint8_t>((masks[d / 16] >> (d % 16)) & 1u));
    } else {
        uint32_t k = 0;
        for (uint16_t m : masks) k += m != 0;
        if (layout == PostingLayout::SPARSE_BITMAP) {
            Put(out, k * 6u);
            for (uint32_t u = 0; u < units; u++)
                if (masks[u]) Put(out, u * 2u);
            for (uint32_t u = 0; u < units; u++)
                if (masks[u]) Put(out, masks[u]);
        } else {
            Put(out, k * 4u);
            for (uint32_t u = 0; u < units; u++) {
                if (!masks[u]) continue;
                if (u > 0xFFFFu) return false;
                Put(out, (u << 16) | masks[u]);
            }
        }
    }
    return true;
}

// ------------------------------------------------------------------------------------------------ JSON
namespace {
struct JVal {
    enum Kind { Null, Num, Str, Arr, Obj, Bool } kind = Null;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
};

struct JParser {
    const std::string& s;
    size_t i = 0;
    std::string err;
    explicit JParser(const std::string& x) : s(x) {}
    void Ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\n' || s[i] == '\r' || s[i] == '\t')) i++;
    }
    bool Str(std::string& out) {
        if (i >= s.size() || s[i] != '"') return false;
        i++;
        out.clear();
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) {
                i++;
                char c = s[i];
                out += c == 'n' ? '\n' : c == 't' ? '\t' : c;
            } else {
                out += s[i];
            }
            i++;
        }
        if (i >= s.size()) return false;
        i++;
        return true;
    }
    bool Val(JVal& v) {
        Ws();
        if (i >= s.size()) return false;
        char c = s[i];
        if (c == '{') {
            v.kind = JVal::Obj;
            i++;
            Ws();
            if (i < s.size() && s[i] == '}') {
                i++;
                return true;
            }
            for (;;) {
                Ws();
                std::string key;
                if (!Str(key)) return false;
                Ws();
                if (i >= s.size() || s[i] != ':') return false;
                i++;
                if (!Val(v.obj[key])) return false;
                Ws();
                if (i < s.size() && s[i] == ',') {
                    i++;
                    continue;
                }
                if (i < s.size() && s[i] == '}') {
                    i++;
                    return true;
                }
                return false;
            }
        }
        if (c == '[') {
            v.kind = JVal::Arr;
            i++;
            Ws();
            if (i < s.size() && s[i] == ']') {
                i++;
                return true;
            }
            for (;;) {
                v.arr.emplace_back();
                if (!Val(v.arr.back())) return false;
                Ws();
                if (i < s.size() && s[i] == ',') {
                    i++;
                    continue;
                }
                if (i < s.size() && s[i] == ']') {
                    i++;
                    return true;
                }
                return false;
            }
        }
        if (c == '"') {
            v.kind = JVal::Str;
            return Str(v.str);
        }
        if (!s.compare(i, 4, "true") || !s.compare(i, 5, "false") || !s.compare(i, 4, "null")) {
            v.kind = s[i] == 'n' ? JVal::Null : JVal::Bool;
            v.num = s[i] == 't';
            i += s[i] == 'f' ? 5 : 4;
            return true;
        }
        size_t j = i;
        while (j < s.size() && (isdigit(static_cast<unsigned char>(s[j])) || s[j] == '-' || s[j] == '+' ||
                                s[j] == '.' || s[j] == 'e' || s[j] == 'E'))
            j++;
        if (j == i) return false;
        v.kind = JVal::Num;
        v.num = std::stod(s.substr(i, j - i));
        i = j;
        return true;
    }
};

const JVal* Field(const JVal& o, const char* camel, const char* snake) {
    if (o.kind != JVal::Obj) return nullptr;
    auto it = o.obj.find(camel);
    if (it == o.obj.end()) it = o.obj.find(snake);
    return it == o.obj.end() ? nullptr : &it->second;
}

uint32_t U32Field(const JVal& o, const char* camel, const char* snake) {
    const JVal* v = Field(o, camel, snake);
    if (!v) return 0;
    if (v->kind == JVal::Num) return static_cast<uint32_t>(v->num);
    if (v->kind == JVal::Str) return static_cast<uint32_t>(std::stoul(v->str));  // proto3 JSON may quote ints
    return 0;
}

void StrList(const JVal& o, const char* name, std::vector<std::string>& out) {
    const JVal* v = Field(o, name, name);
    out.clear();
    if (!v || v->kind != JVal::Arr) return;
    for (const auto& e : v->arr)
        if (e.kind == JVal::Str) out.push_back(e.str);
}
}  // namespace

bool ParseIndexMeta(const std::string& json, IndexMeta& meta, std::string& err) {
    JParser p(json);
    JVal root;
    if (!p.Val(root) || root.kind != JVal::Obj) {
        err = "malformed meta JSON";
        return false;
    }
    static const JVal empty;
    const JVal* config = Field(root, "config", "config");
    const JVal* schema = Field(root, "schema", "schema");
    const JVal* stats = Field(root, "statistics", "statistics");
    meta.docNumPerSegment = U32Field(config ? *config : empty, "docNumPerSegment", "doc_num_per_segment");
    meta.splitDocNumZn = U32Field(config ? *config : empty, "splitDocNumZn", "split_doc_num_zn");
    if (schema) {
        StrList(*schema, "posting", meta.posting);
        StrList(*schema, "vector", meta.vector);
    }
    meta.docNum = U32Field(stats ? *stats : empty, "docNum", "doc_num");
    meta.segmentNum = U32Field(stats ? *stats : empty, "segmentNum", "segment_num");
    return true;
}

std::string IndexMetaToJson(const IndexMeta& m) {
    std::ostringstream o;
    auto list = [&](const std::vector<std::string>& v) {
        o << "[";
        for (size_t i = 0; i < v.size(); i++) o << (i ? ", " : "") << "\"" << v[i] << "\"";
        o << "]";
    };
    o << "{\n  \"config\": {\"docNumPerSegment\": " << m.docNumPerSegment << ", \"splitDocNumZn\": "
      << m.splitDocNumZn << "},\n  \"schema\": {\"posting\": ";
    list(m.posting);
    o << ", \"vector\": ";
    list(m.vector);
    o << "},\n  \"statistics\": {\"docNum\": " << m.docNum << ", \"segmentNum\": " << m.segmentNum << "}\n}\n";
    return o.str();
}

float Float16ToFloat32(uint16_t h) {
    const uint32_t sign = (h >> 15) & 1u, exponent = (h >> 10) & 0x1Fu, mantissa = h & 0x3FFu;
    uint32_t bits;
    if (exponent == 0) {
        bits = sign << 31;
    } else if (exponent == 0x1F) {
        bits = mantissa == 0 ? ((sign << 31) | 0x7F800000u) : ((sign << 31) | 0x7FC00000u | (mantissa << 13));
    } else {
        bits = (sign << 31) | ((exponent - 15 + 127) << 23) | (mantissa << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

std::string PathJoin(const std::string& a, const std::string& b) {
    if (a.empty()) return b;
    return a.back() == '/' ? a + b : a + "/" + b;
}

uint16_t Float32ToFloat16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t sign = (u >> 16) & 0x8000u;
    const uint32_t a = u & 0x7FFFFFFFu;
    if (a >= 0x7F800000u) {  // NaN keeps its top payload bits; infinity (zero payload) becomes NaN 1
        const uint32_t payload = (a >> 13) & 0x3FFu;
        return static_cast<uint16_t>(sign | 0x7C00u | (payload ? payload : 1u));
    }
    if (a < 0x38800000u) return static_cast<uint16_t>(sign);  // below 2^-14 (FP32 subnormals too)
    // rebias the exponent (127 -> 15), then round the 13 dropped bits to nearest even; a carry out of
    // the mantissa moves into the exponent, and exponent 31 is infinity
    uint32_t r = a - (112u << 23);
    r += 0xFFFu + ((r >> 13) & 1u);
    if (r >= (31u << 23)) return static_cast<uint16_t>(sign | 0x7C00u);
    return static_cast<uint16_t>(sign | (r >> 13));
}

std::string SegmentFileName(uint32_t version, uint32_t segment, const std::string& field, const char* suffix) {
    std::string n = std::string(kPoissonEngine) + "_" + std::to_string(version) + "_" + std::to_string(segment);
    if (!field.empty()) n += "_" + field;
    return n + suffix;
}

}  // namespace NpuRetrieval
