// Binary PLY reader with the behaviour of the reference's PlyMesh (Shape/PlyMesh.swift:24-344):
//   header   "ply"; "comment" lines; "format binary_little_endian 1.0" or
//            "format binary_big_endian 1.0" (nothing else); "element vertex N", "element face N";
//            "property float x|y|z|nx|ny|nz|s|t|u|v" (vertex element only);
//            "property list uint|uchar|uint8 <type> <name>" (the count type; indices are int32);
//            "property int face_indices"; "end_header". Any other line is an error. Lines end at
//            the first \n or \r (so a CRLF header fails at its second line, as there).
//   vertices the property sets x y z | x y z u v | x y z nx ny nz | x y z nx ny nz (u|s) (v|t) |
//            x y z u v nx ny nz, float32, byte-swapped for big endian;
//   faces    a uint8 or uint32 count, that many int32 indices appended as they are (faces that are
//            not triangles are appended too: the reference only warns), then the int32 face index
//            when declared.
// Unlike the reference, a truncated body or a header line past the buffer is an error instead of
// a read past the end.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gonzales_amd.h"

namespace gz {
void set_error(const char *fmt, ...);
void clear_error();
}  // namespace gz

namespace {

enum Prop { kX, kY, kZ, kNX, kNY, kNZ, kS, kT, kU, kV };

struct Reader {
        const uint8_t *data;
        int64_t size;
        int64_t at = 0;
        bool big = false;
        bool ok = true;

        bool line(std::string &out) {
                out.clear();
                while (at < size) {
                        const char c = (char)data[at++];
                        if (c == '\n' || c == '\r') return true;
                        out.push_back(c);
                }
                return false;
        }
        template <class T>
        T get() {
                T v{};
                if (at + (int64_t)sizeof(T) > size) {
                        ok = false;
                        return v;
                }
                uint8_t b[sizeof(T)];
                std::memcpy(b, data + at, sizeof(T));
                if (big)
                        for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
                std::memcpy(&v, b, sizeof(T));
                at += (int64_t)sizeof(T);
                return v;
        }
};

std::vector<std::string> split_spaces(const std::string &s) {
        // components(separatedBy: " "): consecutive spaces give empty words
        std::vector<std::string> w;
        size_t start = 0;
        for (;;) {
                const size_t k = s.find(' ', start);
                if (k == std::string::npos) {
                        w.push_back(s.substr(start));
                        return w;
                }
                w.push_back(s.substr(start, k - start));
                start = k + 1;
        }
}

struct Parsed {
        int64_t vertexCount = 0, faceCount = 0;
        std::vector<Prop> props;
        int listBits = 8;
        bool hasFaceIndices = false;
        int layout = 0;  // 1 xyz, 2 xyz uv, 3 xyz n, 4 xyz n uv, 5 xyz uv n
};

int fail(const char *msg, const std::string &detail = std::string()) {
        gz::set_error("gz_ply_parse: %s%s%s", msg, detail.empty() ? "" : " ", detail.c_str());
        return GZ_E_INVALID;
}

int parse_header(Reader &r, Parsed &p) {
        std::string ln;
        if (!r.line(ln) || ln != "ply") return fail("first line must be ply");
        enum { kNone, kVertex, kFace } state = kNone;
        for (;;) {
                if (!r.line(ln)) return fail("header without end_header");
                const std::vector<std::string> w = split_spaces(ln);
                const std::string &w0 = w[0];
                auto word = [&](size_t i) { return i < w.size() ? w[i] : std::string(); };
                if (w0 == "comment") continue;
                if (w0 == "format") {
                        if (word(1) == "binary_little_endian")
                                r.big = false;
                        else if (word(1) == "binary_big_endian")
                                r.big = true;
                        else
                                return fail("unknown format", word(1));
                        if (word(2) != "1.0") return fail("format version must be 1.0");
                } else if (w0 == "element") {
                        char *end = nullptr;
                        const std::string n = word(2);
                        const long long v = std::strtoll(n.c_str(), &end, 10);
                        if (n.empty() || *end != '\0' || v < 0) return fail("bad element count", n);
                        if (word(1) == "vertex") {
                                state = kVertex;
                                p.vertexCount = v;
                        } else if (word(1) == "face") {
                                state = kFace;
                                p.faceCount = v;
                        } else {
                                return fail("unknown element", word(1));
                        }
                } else if (w0 == "property") {
                        const std::string t = word(1), name = word(2);
                        if (t == "float") {
                                if (state != kVertex) return fail("float property outside the vertex element");
                                static const char *names[] = {"x", "y", "z", "nx", "ny", "nz", "s", "t", "u", "v"};
                                int k = -1;
                                for (int i = 0; i < 10; ++i)
                                        if (name == names[i]) k = i;
                                if (k < 0) return fail("unknown float property", name);
                                p.props.push_back((Prop)k);
                        } else if (t == "list") {
                                if (name == "uint")
                                        p.listBits = 32;
                                else if (name == "uchar" || name == "uint8")
                                        p.listBits = 8;
                                else
                                        return fail("unknown list property", name);
                        } else if (t == "int") {
                                if (name != "face_indices") return fail("unknown int property", name);
                                p.hasFaceIndices = true;
                        } else {
                                return fail("unknown property", t);
                        }
                } else if (w0 == "end_header") {
                        break;
                } else {
                        return fail("unknown header word", w0);
                }
        }
        const std::vector<Prop> &q = p.props;
        auto is = [&](std::initializer_list<Prop> want) {
                if (q.size() != want.size()) return false;
                size_t i = 0;
                for (Prop x : want)
                        if (q[i++] != x) return false;
                return true;
        };
        if (is({kX, kY, kZ}))
                p.layout = 1;
        else if (is({kX, kY, kZ, kU, kV}))
                p.layout = 2;
        else if (is({kX, kY, kZ, kNX, kNY, kNZ}))
                p.layout = 3;
        else if (q.size() == 8 && q[0] == kX && q[1] == kY && q[2] == kZ && q[3] == kNX && q[4] == kNY && q[5] == kNZ &&
                 (q[6] == kU || q[6] == kS) && (q[7] == kV || q[7] == kT))
                p.layout = 4;
        else if (is({kX, kY, kZ, kU, kV, kNX, kNY, kNZ}))
                p.layout = 5;
        else if (p.vertexCount > 0)
                return fail("unsupported vertex property set");
        return GZ_OK;
}

}  // namespace

// Parses a binary PLY held in memory. Call once with null arrays to get the counts (info), then
// again with arrays of those sizes: points 3 x vertexCount floats, normals 3 x vertexCount (when
// hasNormals), uvs 2 x vertexCount (when hasUvs), indices indexCount int64, faceIndices
// faceCount int64 (when hasFaceIndices).
extern "C" int gz_ply_parse(const uint8_t *data, int64_t size, GzPlyInfo *info, float *points, float *normals,
                            float *uvs, int64_t *indices, int64_t *faceIndices) {
        gz::clear_error();
        if (!data || size < 0 || !info) return fail("bad arguments");
        Reader r{data, size};
        Parsed p;
        const int rc = parse_header(r, p);
        if (rc != GZ_OK) return rc;
        const bool hasN = p.layout >= 3, hasUv = p.layout == 2 || p.layout >= 4;
        const int64_t vbytes = (int64_t)p.props.size() * 4;
        if (p.vertexCount > 0 && (size - r.at) / (vbytes > 0 ? vbytes : 1) < p.vertexCount)
                return fail("truncated vertex data");
        const int64_t vertexStart = r.at;
        r.at += p.vertexCount * vbytes;
        // faces: count the indices (the counts are interleaved with the data)
        int64_t nIdx = 0;
        const int64_t faceStart = r.at;
        for (int64_t f = 0; f < p.faceCount; ++f) {
                const uint32_t n = p.listBits == 8 ? (uint32_t)r.get<uint8_t>() : r.get<uint32_t>();
                if (!r.ok || (size - r.at) / 4 < (int64_t)n + (p.hasFaceIndices ? 1 : 0)) return fail("truncated face data");
                r.at += 4 * (int64_t)n + (p.hasFaceIndices ? 4 : 0);
                nIdx += n;
        }
        info->vertexCount = p.vertexCount;
        info->faceCount = p.faceCount;
        info->indexCount = nIdx;
        info->hasNormals = hasN;
        info->hasUvs = hasUv;
        info->hasFaceIndices = p.hasFaceIndices;
        info->bigEndian = r.big;
        if (!points) return GZ_OK;
        if ((hasN && !normals) || (hasUv && !uvs) || (nIdx > 0 && !indices) ||
            (p.hasFaceIndices && p.faceCount > 0 && !faceIndices))
                return fail("missing output array");
        r.at = vertexStart;
        for (int64_t v = 0; v < p.vertexCount; ++v) {
                float *P = points + 3 * v;
                auto xyz = [&](float *o) {
                        o[0] = r.get<float>();
                        o[1] = r.get<float>();
                        o[2] = r.get<float>();
                };
                auto uv = [&]() {
                        uvs[2 * v] = r.get<float>();
                        uvs[2 * v + 1] = r.get<float>();
                };
                xyz(P);
                if (p.layout == 2) uv();
                if (p.layout == 3 || p.layout == 4) xyz(normals + 3 * v);
                if (p.layout == 4) uv();
                if (p.layout == 5) {
                        uv();
                        xyz(normals + 3 * v);
                }
        }
        r.at = faceStart;
        int64_t k = 0;
        for (int64_t f = 0; f < p.faceCount; ++f) {
                const uint32_t n = p.listBits == 8 ? (uint32_t)r.get<uint8_t>() : r.get<uint32_t>();
                for (uint32_t j = 0; j < n; ++j) indices[k++] = (int64_t)r.get<int32_t>();
                if (p.hasFaceIndices) faceIndices[f] = (int64_t)r.get<int32_t>();
        }
        return r.ok ? GZ_OK : fail("truncated data");
}
