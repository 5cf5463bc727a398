// Types shared by the extraction kernels (extract.hip) and the batched and
// served forms (batch.hip): grid descriptors, per-grid results and the
// arguments of a many-grid extraction.
#pragma once

#include <algorithm>
#include <cstring>

#include "gc_common.hpp"

namespace gc {

constexpr int kBlock = 256;  // words per extraction block (4 waves)
constexpr int kGroups = 8;   // triangle-count groups (32 words each) per block
// per block: first vertex number of each 32-point half word (uint32), and the
// triangle count of each group (uint32)
constexpr uint64_t kBaseBytesPerBlock = (uint64_t)kBlock * 2 * 4;
constexpr uint64_t kTwaveBytesPerBlock = (uint64_t)kGroups * 4;

// n / d for n < 2^32 through a double reciprocal and one correction step
struct FastDiv {
    uint32_t d;
    double inv;
};

inline FastDiv make_div(uint32_t d) { return FastDiv{d, 1.0 / (double)d}; }

struct Grid {
    int nx, ny, nz, W;  // W = sign words per row
    uint64_t words;     // W * ny * nz
    FastDiv divW, divNy;
};

inline Grid make_grid(int nx, int ny, int nz) {
    Grid g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.W = (int)gpucad_sign_words(nx, 1, 1);
    g.words = gpucad_sign_words(nx, ny, nz);
    g.divW = make_div((uint32_t)g.W);
    g.divNy = make_div((uint32_t)ny);
    return g;
}

// One grid of an extraction launch.  Its words occupy the global blocks
// [block_off, block_off + ceil(words / kBlock)); per-word arrays are indexed
// by global block * kBlock + word of the block.
struct Seg {
    Grid g;
    uint64_t field_off;  // first point of the grid in the field buffer (floats)
    uint64_t sign_off;   // first sign word of the grid in the sign buffer
    uint32_t block_off;
    uint32_t pad;
    float iso;
    float ox, oy, oz;    // position of point (0, 0, 0)
    float sx, sy, sz;    // point spacing
};

// Mesh of one grid: vertices [v_off, v_off + nv) and triangles
// [t_off, t_off + nt) of the launch's output buffers; triangle corners are
// vertex numbers relative to v_off.  ovf: vertex numbers do not fit 32 bits
// (the grid's vertices and triangles are then not written).
struct SegResult {
    uint64_t v_off, nv, t_off, nt;
    uint32_t ovf, pad;
};

// Extraction of many grids (one workgroup per global block).  The sign words
// must already hold [f < iso] for every grid (the batched field kernel writes
// them).
struct BatchExtract {
    const float *fields;       // field buffer base
    const uint64_t *signs;     // sign buffer base
    const Seg *segs;           // device, nseg
    const uint32_t *blk2seg;   // device, nb: grid of each global block
    uint32_t nseg, nb;
    uint32_t *base, *twave;    // nb * kBlock * 2, nb * kGroups
    uint64_t *bsum;            // nb * 2
    uint64_t *ctl;             // 8 x uint64: totals, any overflow
    SegResult *res;            // device, nseg
    float *vertices;
    uint64_t vcap;
    uint32_t *triangles;
    uint64_t tcap;
};

enum { EXTRACT_COUNT = 1, EXTRACT_EMIT = 2, EXTRACT_ALL = 3 };

// Enqueues the stages `phases` (count+scan, vert+tri) of a many-grid
// extraction on `stream`.
int enqueue_extract_batch(const BatchExtract &b, int phases, hipStream_t stream);

}  // namespace gc
