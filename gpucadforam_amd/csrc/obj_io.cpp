// .obj output.
//
// gpucad_file_write_obj reproduces File_output::file_write_obj
// (src/File_output.cu:5-81) byte for byte: the non-indexed float4 positions
// are copied to the host, each component is truncated to the weld key
// (float)(int(v*1000) * 0.001), vertices are numbered 1.. in first-seen order,
// a face is kept when its three ids are distinct and the ordered triple has not
// been written before, and it is written as " f  a c b".  The reference looks
// keys up in std::map (O(V log V)); this uses hash tables with the same key
// equality (float ==, so -0 and +0 are one key) and the same first-seen order.
//
// gpucad_write_obj_indexed writes an indexed device mesh directly.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "gc_common.hpp"

namespace gc {
namespace {

struct Key3 {
    float x, y, z;
    bool operator==(const Key3 &o) const { return x == o.x && y == o.y && z == o.z; }
};

inline uint32_t canon_bits(float f) {
    if (f == 0.0f) f = 0.0f;  // -0 and +0 compare equal, hash them alike
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

struct Key3Hash {
    size_t operator()(const Key3 &k) const {
        uint64_t h = canon_bits(k.x);
        h = h * 0x9E3779B97F4A7C15ull ^ canon_bits(k.y);
        h = h * 0x9E3779B97F4A7C15ull ^ canon_bits(k.z);
        h ^= h >> 29;
        return (size_t)(h * 0xBF58476D1CE4E5B9ull);
    }
};

struct Tri3 {
    uint32_t a, b, c;
    bool operator==(const Tri3 &o) const { return a == o.a && b == o.b && c == o.c; }
};

struct Tri3Hash {
    size_t operator()(const Tri3 &t) const {
        uint64_t h = ((uint64_t)t.a << 32) ^ ((uint64_t)t.b << 16) ^ t.c;
        h ^= (uint64_t)t.b * 0x9E3779B97F4A7C15ull;
        h ^= h >> 31;
        return (size_t)(h * 0xBF58476D1CE4E5B9ull);
    }
};

// Appends the text std::ostream << float produces (default floatfield,
// precision 6 == "%g" of the value as double).
inline void put_float(std::string &out, float v) {
    char buf[32];
    int n = snprintf(buf, sizeof(buf), "%g", (double)v);
    out.append(buf, (size_t)n);
}

inline float weld_key(float v) {
    // int(v * 1000) * 0.001 with v * 1000 in float and the product in double
    float s = v * 1000.0f;
    int i = (int)s;
    return (float)((double)i * 0.001);
}

}  // namespace
}  // namespace gc

extern "C" {

int gpucad_file_write_obj(const gc_float4 *d_pos, uint32_t total_verts, const char *filename) {
    using namespace gc;
    clear_error();
    GC_ARG(filename, "filename is NULL");
    GC_ARG(d_pos || total_verts == 0, "positions are NULL");
    std::vector<gc_float4> h(total_verts);
    if (total_verts) {
        GC_HIP(hipMemcpy(h.data(), d_pos, (size_t)total_verts * sizeof(gc_float4),
                         hipMemcpyDeviceToHost));
    }
    FILE *f = std::fopen(filename, "wb");
    if (!f) return fail(GPUCAD_E_IO, "cannot open %s", filename);
    std::string out;
    out.reserve((size_t)total_verts * 24 + 64);
    out += "##Sample latttice new Obj \n";
    out += "o Solid \n";
    std::unordered_map<Key3, uint32_t, Key3Hash> ids;
    ids.reserve(total_verts / 2 + 16);
    std::vector<uint32_t> faces(total_verts);
    uint32_t index = 0;
    for (uint32_t i = 0; i < total_verts; ++i) {
        Key3 k{weld_key(h[i].x), weld_key(h[i].y), weld_key(h[i].z)};
        auto it = ids.find(k);
        if (it == ids.end()) {
            ++index;
            ids.emplace(k, index);
            faces[i] = index;
            out += "v ";
            put_float(out, k.x);
            out += ' ';
            put_float(out, k.y);
            out += ' ';
            put_float(out, k.z);
            out += '\n';
        } else {
            faces[i] = it->second;
        }
    }
    out += "\n";
    out += "\n";
    std::unordered_set<Tri3, Tri3Hash> seen;
    seen.reserve(total_verts / 3 + 16);
    char buf[64];
    for (size_t i = 0; i + 2 < faces.size(); i += 3) {
        uint32_t a = faces[i], b = faces[i + 1], c = faces[i + 2];
        if (a != b && a != c && b != c) {
            if (seen.insert(Tri3{a, b, c}).second) {
                int n = snprintf(buf, sizeof(buf), " f  %u %u %u\n", a, c, b);
                out.append(buf, (size_t)n);
            }
        }
    }
    size_t w = std::fwrite(out.data(), 1, out.size(), f);
    int cl = std::fclose(f);
    if (w != out.size() || cl != 0) return fail(GPUCAD_E_IO, "short write to %s", filename);
    return GPUCAD_OK;
}

int gpucad_write_obj_indexed(const float *d_vertices, uint64_t n_vertices,
                             const uint32_t *d_triangles, uint64_t n_triangles,
                             const char *filename, int32_t flip_winding) {
    using namespace gc;
    clear_error();
    GC_ARG(filename, "filename is NULL");
    GC_ARG(d_vertices || n_vertices == 0, "vertices are NULL");
    GC_ARG(d_triangles || n_triangles == 0, "triangles are NULL");
    std::vector<float> v(3 * n_vertices);
    std::vector<uint32_t> t(3 * n_triangles);
    if (n_vertices)
        GC_HIP(hipMemcpy(v.data(), d_vertices, v.size() * sizeof(float), hipMemcpyDeviceToHost));
    if (n_triangles)
        GC_HIP(hipMemcpy(t.data(), d_triangles, t.size() * sizeof(uint32_t),
                         hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < t.size(); ++i)
        GC_ARG(t[i] < n_vertices, "triangle index %u out of range", t[i]);
    FILE *f = std::fopen(filename, "wb");
    if (!f) return fail(GPUCAD_E_IO, "cannot open %s", filename);
    std::string out;
    out.reserve((size_t)(n_vertices * 40 + n_triangles * 30 + 64));
    out += "o Solid\n";
    char buf[128];
    for (uint64_t i = 0; i < n_vertices; ++i) {
        int n = snprintf(buf, sizeof(buf), "v %.9g %.9g %.9g\n", (double)v[3 * i],
                         (double)v[3 * i + 1], (double)v[3 * i + 2]);
        out.append(buf, (size_t)n);
    }
    for (uint64_t i = 0; i < n_triangles; ++i) {
        uint32_t a = t[3 * i] + 1, b = t[3 * i + 1] + 1, c = t[3 * i + 2] + 1;
        int n = flip_winding ? snprintf(buf, sizeof(buf), "f %u %u %u\n", a, c, b)
                             : snprintf(buf, sizeof(buf), "f %u %u %u\n", a, b, c);
        out.append(buf, (size_t)n);
    }
    size_t w = std::fwrite(out.data(), 1, out.size(), f);
    int cl = std::fclose(f);
    if (w != out.size() || cl != 0) return fail(GPUCAD_E_IO, "short write to %s", filename);
    return GPUCAD_OK;
}

}  // extern "C"
