// Launch plans of the many-grid forms (batched: batch.hip, served: server.cpp):
// which grids, where their fields and sign words live, the brick list of the
// field launches and the block -> grid table of the extraction launches.
#pragma once

#include <vector>

#include "extract_types.hpp"
#include "lattice_body.hpp"

namespace gc {

constexpr int kVariants = 12;  // field variants: rotation x tpms x thickness modes

inline uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

struct Shape {
    int bx = 64, lz = 16;  // field brick: x points per wave row, z planes
};

struct HostPlan {
    std::vector<FieldConsts> consts;
    std::vector<Seg> segs;
    std::vector<uint32_t> blk2seg;
    std::vector<uint64_t> items;     // seg | brick x << 16 | brick y << 32 | brick z << 48
    uint32_t vbegin[kVariants + 1] = {};
    uint32_t nb = 0;
};

// Device buffers of one plan.
struct DevPlan {
    FieldConsts *consts = nullptr;
    uint64_t *items = nullptr;
    Seg *segs = nullptr;
    uint32_t *blk2seg = nullptr;
    SegResult *res = nullptr;
    uint32_t *base = nullptr, *twave = nullptr;
    uint64_t *bsum = nullptr, *ctl = nullptr;
};

int validate_batch_grid(const gc_lattice_params &p, int i);
uint64_t brick_count(const gc_lattice_params &p, int bx, int lz);
void plan_segments(const gc_lattice_params *p, int n, const uint64_t *field_off,
                   const uint64_t *sign_off, float iso, int unit_domain, HostPlan &h);
void plan_items(const gc_lattice_params *p, int n, const Shape &sh, HostPlan &h);
int enqueue_field_plan(const HostPlan &h, const DevPlan &d, const Shape &sh, float *fields,
                       uint64_t *signs, hipStream_t stream);
BatchExtract make_batch_extract(const HostPlan &h, const DevPlan &d, const float *fields,
                                const uint64_t *signs, float *vertices, uint64_t vcap,
                                uint32_t *triangles, uint64_t tcap);
int device_cus(int *cus);
// Default shape: brick 64 x 4 x 16.
int default_shape(Shape *s);

}  // namespace gc
