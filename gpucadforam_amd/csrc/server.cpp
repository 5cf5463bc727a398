// Served form: a request loop with continuous batching under a device-memory
// budget (gpucad_server_*), and the plain per-request baseline it is compared
// with (gpucad_plain_run_trace).
//
// Memory: one hipMalloc'd arena of budget bytes, carved by a first-fit
// allocator (256-byte granules, neighbours coalesced on free).  A request
// holds its field from admission until release; a launch group holds its
// work region (plan tables, sign words, extraction bookkeeping) until it
// completes, and one mesh region (all its members' meshes) until every member
// is released.  Field and sign offsets are taken relative to the arena base,
// so one launch addresses every member through the arena base pointer.
//
// Loop (gpucad_server_poll, never waits): finished groups (event of the emit
// stage) complete their members; groups whose count stage finished read their
// per-request mesh sizes (copied to pinned host memory in-stream), take a
// mesh region and enqueue the emit stage; free streams take the next group
// from the FIFO (admission: the request's field and its share of the group's
// work must fit in the arena now; a request larger than the whole budget is
// rejected with GPUCAD_E_OOM).  While a group waits for mesh memory no new
// group is admitted; if no group in flight can progress, the youngest waiting
// group is put back in the queue, so the loop cannot stall on memory held by
// groups that wait for each other.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <unordered_map>
#include <vector>

#include "plan.hpp"

namespace gc {
namespace {

constexpr uint64_t kGranule = 256;

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

class Arena {
public:
    int init(uint64_t bytes) {
        size_ = round_up(bytes, kGranule);
        hipError_t e = hipMalloc(&base_, size_);
        if (e != hipSuccess) {
            base_ = nullptr;
            return fail(GPUCAD_E_OOM, "server arena of %llu bytes: %s", (unsigned long long)size_,
                        hipGetErrorString(e));
        }
        free_[0] = size_;
        return GPUCAD_OK;
    }
    ~Arena() {
        if (base_) (void)hipFree(base_);
    }
    char *base() const { return base_; }
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }
    // first fit; returns false when no free range is large enough
    bool alloc(uint64_t bytes, uint64_t *off) {
        bytes = round_up(std::max<uint64_t>(bytes, 1), kGranule);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const uint64_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        bytes = round_up(std::max<uint64_t>(bytes, 1), kGranule);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t largest_free() const {
        uint64_t m = 0;
        for (auto &f : free_) m = std::max(m, f.second);
        return m;
    }

private:
    char *base_ = nullptr;
    uint64_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<uint64_t, uint64_t> free_;
};

// bytes of one request's share of a group's work region (plan tables, sign
// words, extraction bookkeeping), each array rounded up to a granule
uint64_t work_bytes(const gc_lattice_params &p, const Shape &sh) {
    const uint64_t words = gpucad_sign_words(p.nx, p.ny, p.nz);
    const uint64_t nb = (words + kBlock - 1) / kBlock;
    uint64_t b = 0;
    b += round_up(sizeof(FieldConsts), kGranule) + round_up(sizeof(Seg), kGranule) +
         round_up(sizeof(SegResult), kGranule);
    b += round_up(brick_count(p, sh.bx, sh.lz) * 8, kGranule);
    b += round_up(nb * 4, kGranule);                    // blk2seg
    b += round_up(nb * kBaseBytesPerBlock, kGranule);   // base
    b += round_up(nb * kTwaveBytesPerBlock, kGranule);  // twave
    b += round_up(nb * 16, kGranule);                   // bsum
    b += round_up(round_up(words, 32) * 8, kGranule);   // sign words
    return b;
}

uint64_t field_bytes(const gc_lattice_params &p) {
    return round_up((uint64_t)p.nx * p.ny * p.nz * 4, kGranule);
}

enum ReqState { Q_WAIT, Q_RUN, Q_DONE };

struct Request {
    uint64_t ticket = 0;
    gc_lattice_params p;
    double submit_s = 0, done_s = 0;
    ReqState state = Q_WAIT;
    int status = GPUCAD_OK;
    uint64_t field_off = 0;
    bool has_field = false;
    int group = -1;
    uint64_t v_off = 0, nv = 0, t_off = 0, nt = 0;  // within the group's mesh region
    bool reported = false;
};

enum GroupPhase { G_COUNT, G_WAIT_MESH, G_EMIT };

struct Group {
    int id = 0, slot = -1;
    std::vector<uint64_t> tickets;
    HostPlan plan;
    DevPlan d;
    float *fields = nullptr;    // arena base
    uint64_t *signs = nullptr;  // arena base
    uint64_t work_off = 0, work_size = 0;
    bool has_work = false;
    uint64_t mesh_off = 0, mesh_size = 0;
    bool has_mesh = false;
    uint64_t vtot = 0, ttot = 0;
    int live = 0;  // members not yet released
    GroupPhase phase = G_COUNT;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    char *pinned = nullptr;
    size_t pinned_bytes = 0;
    int group = -1;
};

}  // namespace

struct Server {
    gc_server_options opt = {};
    Shape shape;
    int cus = 0;
    Arena arena;
    double t0 = 0;
    std::vector<Slot> slots;
    std::deque<uint64_t> fifo;
    std::unordered_map<uint64_t, Request> reqs;
    std::map<int, Group> groups;
    std::deque<uint64_t> finished;  // not yet reported
    uint64_t next_ticket = 1;
    int next_group = 0;
    int group_cap = 1 << 30;  // lowered after a group that could never get its mesh memory
    uint64_t submitted = 0, completed = 0, rejected = 0, n_groups = 0, grouped = 0;

    ~Server() {
        for (auto &s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.ev) (void)hipEventDestroy(s.ev);
            if (s.stream) (void)hipStreamDestroy(s.stream);
            if (s.pinned) (void)hipHostFree(s.pinned);
        }
    }

    double clock() const { return now_s() - t0; }

    int pinned(Slot &s, size_t bytes) {
        if (s.pinned_bytes >= bytes) return GPUCAD_OK;
        if (s.pinned) {
            GC_HIP(hipStreamSynchronize(s.stream));
            GC_HIP(hipHostFree(s.pinned));
            s.pinned = nullptr;
            s.pinned_bytes = 0;
        }
        const size_t want = std::max<size_t>(bytes + bytes / 2, 1 << 20);
        GC_HIP(hipHostMalloc(&s.pinned, want, hipHostMallocDefault));
        s.pinned_bytes = want;
        return GPUCAD_OK;
    }

    void complete(Request &r, int status) {
        r.state = Q_DONE;
        r.status = status;
        r.done_s = clock();
        finished.push_back(r.ticket);
        if (status == GPUCAD_OK) ++completed;
        else ++rejected;
    }

    // ---- new groups on free slots
    int launch_groups() {
        for (int si = 0; si < (int)slots.size() && !fifo.empty(); ++si) {
            Slot &sl = slots[si];
            if (sl.group >= 0) continue;
            Group g;
            g.id = next_group;
            uint64_t work = 2 * kGranule;  // ctl and one granule of slack
            while (!fifo.empty() && (int)g.tickets.size() < std::min(opt.max_batch, group_cap)) {
                Request &r = reqs[fifo.front()];
                const uint64_t fb = field_bytes(r.p), wb = work_bytes(r.p, shape);
                if (fb + wb + 2 * kGranule > arena.size()) {  // can never fit
                    fifo.pop_front();
                    complete(r, GPUCAD_E_OOM);
                    continue;
                }
                uint64_t off = 0;
                if (!arena.alloc(fb, &off)) break;
                if (arena.largest_free() < work + wb) {
                    arena.release(off, fb);
                    break;
                }
                r.field_off = off;
                r.has_field = true;
                r.group = g.id;
                r.state = Q_RUN;
                work += wb;
                g.tickets.push_back(r.ticket);
                fifo.pop_front();
            }
            if (g.tickets.empty()) break;  // wait for memory
            if (!arena.alloc(work, &g.work_off)) {
                // fragmentation: give the fields back and retry later
                for (auto t : g.tickets) {
                    Request &r = reqs[t];
                    arena.release(r.field_off, field_bytes(r.p));
                    r.has_field = false;
                    r.state = Q_WAIT;
                    r.group = -1;
                }
                fifo.insert(fifo.begin(), g.tickets.begin(), g.tickets.end());
                break;
            }
            g.work_size = work;
            g.has_work = true;
            ++next_group;
            int st = start_group(g, si);
            if (st) return st;
            groups.emplace(g.id, std::move(g));
        }
        return GPUCAD_OK;
    }

    int start_group(Group &g, int si) {
        Slot &sl = slots[si];
        g.slot = si;
        sl.group = g.id;
        ++n_groups;
        grouped += g.tickets.size();
        const int n = (int)g.tickets.size();
        std::vector<gc_lattice_params> ps(n);
        std::vector<uint64_t> foff(n), soff(n);
        // work region layout: per-array ranges inside [work_off, work_off + work_size)
        uint64_t at = g.work_off;
        auto take = [&](uint64_t bytes) {
            const uint64_t o = at;
            at += round_up(std::max<uint64_t>(bytes, 1), kGranule);
            return o;
        };
        for (int i = 0; i < n; ++i) {
            const Request &r = reqs[g.tickets[i]];
            ps[i] = r.p;
            foff[i] = r.field_off / 4;
            soff[i] = take(round_up(gpucad_sign_words(r.p.nx, r.p.ny, r.p.nz), 32) * 8) / 8;
        }
        plan_segments(ps.data(), n, foff.data(), soff.data(), opt.iso, 1, g.plan);
        plan_items(ps.data(), n, shape, g.plan);
        char *A = arena.base();
        const uint64_t nb = g.plan.nb;
        const uint64_t o_consts = take(n * sizeof(FieldConsts));
        const uint64_t o_items = take(g.plan.items.size() * 8);
        const uint64_t o_segs = take(n * sizeof(Seg));
        const uint64_t o_b2s = take(nb * 4);
        const uint64_t o_res = take(n * sizeof(SegResult));
        const uint64_t o_base = take(nb * kBaseBytesPerBlock);
        const uint64_t o_twave = take(nb * kTwaveBytesPerBlock);
        const uint64_t o_bsum = take(nb * 16);
        const uint64_t o_ctl = take(64);
        if (at > g.work_off + g.work_size)
            return fail(GPUCAD_E_ARG, "internal: work region overrun");
        g.d.consts = (FieldConsts *)(A + o_consts);
        g.d.items = (uint64_t *)(A + o_items);
        g.d.segs = (Seg *)(A + o_segs);
        g.d.blk2seg = (uint32_t *)(A + o_b2s);
        g.d.res = (SegResult *)(A + o_res);
        g.d.base = (uint32_t *)(A + o_base);
        g.d.twave = (uint32_t *)(A + o_twave);
        g.d.bsum = (uint64_t *)(A + o_bsum);
        g.d.ctl = (uint64_t *)(A + o_ctl);
        g.fields = (float *)A;
        g.signs = (uint64_t *)A;
        // tables through pinned staging (the staging also receives the results)
        const size_t b_consts = n * sizeof(FieldConsts), b_items = g.plan.items.size() * 8,
                     b_segs = n * sizeof(Seg), b_b2s = nb * 4, b_res = n * sizeof(SegResult);
        int st = pinned(sl, b_consts + b_items + b_segs + b_b2s + b_res + 1024);
        if (st) return st;
        char *P = sl.pinned;
        size_t pat = 0;
        auto stage = [&](const void *src, size_t bytes, uint64_t dst) -> int {
            std::memcpy(P + pat, src, bytes);
            GC_HIP(hipMemcpyAsync(A + dst, P + pat, bytes, hipMemcpyHostToDevice, sl.stream));
            pat += round_up(bytes, 16);
            return GPUCAD_OK;
        };
        if ((st = stage(g.plan.consts.data(), b_consts, o_consts)) ||
            (st = stage(g.plan.items.data(), b_items, o_items)) ||
            (st = stage(g.plan.segs.data(), b_segs, o_segs)) ||
            (st = stage(g.plan.blk2seg.data(), b_b2s, o_b2s)))
            return st;
        if (shape.bx < 64) {
            // narrow bricks leave the tail bytes of a row's last sign word unwritten
            for (int i = 0; i < n; ++i) {
                const gc_lattice_params &p = ps[i];
                GC_HIP(hipMemsetAsync(g.signs + soff[i], 0,
                                      gpucad_sign_words(p.nx, p.ny, p.nz) * 8, sl.stream));
            }
        }
        if ((st = enqueue_field_plan(g.plan, g.d, shape, g.fields, g.signs, sl.stream)))
            return st;
        const BatchExtract be = make_batch_extract(g.plan, g.d, g.fields, g.signs,
                                                   nullptr, 0, nullptr, 0);
        if ((st = enqueue_extract_batch(be, EXTRACT_COUNT, sl.stream))) return st;
        GC_HIP(hipMemcpyAsync(P + pat, g.d.res, b_res, hipMemcpyDeviceToHost, sl.stream));
        GC_HIP(hipEventRecord(sl.ev, sl.stream));
        g.phase = G_COUNT;
        g.live = n;
        return GPUCAD_OK;
    }

    // pinned location of a group's results (after the staged tables)
    const SegResult *staged_results(const Group &g) const {
        const size_t n = g.tickets.size();
        size_t pat = round_up(n * sizeof(FieldConsts), 16) + round_up(g.plan.items.size() * 8, 16) +
                     round_up(n * sizeof(Seg), 16) + round_up((size_t)g.plan.nb * 4, 16);
        return (const SegResult *)(slots[g.slot].pinned + pat);
    }

    int start_emit(Group &g) {
        Slot &sl = slots[g.slot];
        const uint64_t bytes = g.vtot * 12 + g.ttot * 12 + 2 * kGranule;
        if (!arena.alloc(bytes, &g.mesh_off)) {
            g.phase = G_WAIT_MESH;
            return GPUCAD_OK;
        }
        g.mesh_size = bytes;
        g.has_mesh = true;
        float *verts = (float *)(arena.base() + g.mesh_off);
        uint32_t *tris =
            (uint32_t *)(arena.base() + g.mesh_off + round_up(g.vtot * 12, kGranule));
        const BatchExtract be = make_batch_extract(g.plan, g.d, g.fields, g.signs,
                                                   verts, g.vtot, tris, g.ttot);
        int st = enqueue_extract_batch(be, EXTRACT_EMIT, sl.stream);
        if (st) return st;
        GC_HIP(hipEventRecord(sl.ev, sl.stream));
        g.phase = G_EMIT;
        return GPUCAD_OK;
    }

    int advance() {
        for (auto &kv : groups) {
            Group &g = kv.second;
            if (g.slot < 0) continue;
            Slot &sl = slots[g.slot];
            if (g.phase == G_WAIT_MESH) {
                int st = start_emit(g);
                if (st) return st;
                continue;
            }
            const hipError_t q = hipEventQuery(sl.ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess)
                return fail(GPUCAD_E_HIP, "server stream failed: %s", hipGetErrorString(q));
            if (g.phase == G_COUNT) {
                const SegResult *r = staged_results(g);
                for (size_t i = 0; i < g.tickets.size(); ++i) {
                    Request &rq = reqs[g.tickets[i]];
                    rq.v_off = r[i].v_off;
                    rq.nv = r[i].nv;
                    rq.t_off = r[i].t_off;
                    rq.nt = r[i].nt;
                    if (r[i].ovf) return fail(GPUCAD_E_OVERFLOW, "mesh exceeds 32-bit numbers");
                }
                const size_t last = g.tickets.size() - 1;
                g.vtot = r[last].v_off + r[last].nv;
                g.ttot = r[last].t_off + r[last].nt;
                int st = start_emit(g);
                if (st) return st;
            } else if (g.phase == G_EMIT) {
                for (auto t : g.tickets) complete(reqs[t], GPUCAD_OK);
                arena.release(g.work_off, g.work_size);
                g.has_work = false;
                sl.group = -1;
                g.slot = -1;
                group_cap = 1 << 30;
            }
        }
        // Groups waiting for mesh memory are fed by frees only.  When no group
        // in flight can make progress, the youngest waiting group gives its
        // memory back and its requests go back to the front of the queue; a
        // lone group that could not fit even in an otherwise empty arena is
        // split (or, a single request, rejected).
        int waiting = 0, youngest = -1, oldest = -1;
        bool progress = false;
        for (auto &kv : groups) {
            const Group &g = kv.second;
            if (g.slot < 0) continue;
            if (g.phase == G_WAIT_MESH) {
                ++waiting;
                youngest = std::max(youngest, g.id);
                oldest = oldest < 0 ? g.id : std::min(oldest, g.id);
            } else {
                progress = true;
            }
        }
        if (waiting && !progress) {
            if (waiting > 1) {
                requeue(groups.at(youngest));
            } else {
                Group &w = groups.at(oldest);
                uint64_t own = w.work_size;
                for (auto t : w.tickets) own += field_bytes(reqs[t].p);
                if (arena.used() == own) {
                    if (w.tickets.size() > 1) {
                        group_cap = std::max(1, (int)w.tickets.size() / 2);
                        requeue(w);
                    } else {
                        const uint64_t t = w.tickets[0];
                        requeue(w);
                        fifo.pop_front();
                        complete(reqs[t], GPUCAD_E_OOM);
                    }
                }
            }
        }
        if (waiting) return GPUCAD_OK;  // frees go to the waiting groups first
        return launch_groups();
    }

    // Returns a group's work region and its members' fields to the arena and
    // puts the members back at the front of the queue, in order.
    void requeue(Group &g) {
        arena.release(g.work_off, g.work_size);
        for (auto t : g.tickets) {
            Request &r = reqs[t];
            if (r.has_field) arena.release(r.field_off, field_bytes(r.p));
            r.has_field = false;
            r.state = Q_WAIT;
            r.group = -1;
        }
        fifo.insert(fifo.begin(), g.tickets.begin(), g.tickets.end());
        slots[g.slot].group = -1;
        groups.erase(g.id);
    }

    void fill_result(const Request &r, gc_request_result *o) const {
        std::memset(o, 0, sizeof(*o));
        o->ticket = r.ticket;
        o->status = r.status;
        o->submit_s = r.submit_s;
        o->done_s = r.done_s;
        if (r.status != GPUCAD_OK) return;
        const Group &g = groups.at(r.group);
        char *A = arena.base();
        o->field = (const float *)(A + r.field_off);
        o->vertices = (const float *)(A + g.mesh_off) + 3 * r.v_off;
        o->triangles =
            (const uint32_t *)(A + g.mesh_off + round_up(g.vtot * 12, kGranule)) + 3 * r.t_off;
        o->n_vertices = r.nv;
        o->n_triangles = r.nt;
    }

    int release(uint64_t ticket) {
        auto it = reqs.find(ticket);
        GC_ARG(it != reqs.end(), "unknown ticket %llu", (unsigned long long)ticket);
        Request &r = it->second;
        if (r.state != Q_DONE) return fail(GPUCAD_E_BUSY, "request %llu is not finished",
                                           (unsigned long long)ticket);
        if (r.has_field) arena.release(r.field_off, field_bytes(r.p));
        if (r.group >= 0) {
            auto gi = groups.find(r.group);
            if (gi != groups.end() && --gi->second.live == 0) {
                if (gi->second.has_mesh) arena.release(gi->second.mesh_off, gi->second.mesh_size);
                groups.erase(gi);
            }
        }
        reqs.erase(it);
        return GPUCAD_OK;
    }
};

}  // namespace gc

struct gc_server {
    gc::Server *impl;
};

extern "C" {

int gpucad_server_create(const gc_server_options *o, gc_server **out) {
    gc::clear_error();
    GC_ARG(o && out, "NULL argument");
    *out = nullptr;
    GC_ARG(o->budget_bytes >= (1ull << 20), "budget must be at least 1 MiB");
    GC_ARG(o->max_batch >= 0 && o->max_batch <= 65535, "max_batch must be 0..65535");
    GC_ARG(o->slots >= 0 && o->slots <= 8, "slots must be 0..8");
    GC_ARG(o->brick_x == 0 || o->brick_x == 64 || o->brick_x == 32 || o->brick_x == 16 ||
               o->brick_x == 8,
           "brick_x must be 0, 64, 32, 16 or 8");
    GC_ARG(o->brick_lz >= 0 && o->brick_lz <= gc::kMaxLZ, "brick_lz out of range");
    GC_ARG(o->waves_per_cu == 0,
           "waves_per_cu must be 0: the field runs one workgroup per brick");
    const gc::DeviceTables *tabs = nullptr;
    int st = gc::device_tables(&tabs);
    if (st) return st;
    gc::Server *s = new gc::Server();
    s->opt = *o;
    if (!s->opt.max_batch) s->opt.max_batch = 256;
    if (!s->opt.slots) s->opt.slots = 3;
    if ((st = gc::default_shape(&s->shape)) || (st = gc::device_cus(&s->cus)) ||
        (st = s->arena.init(o->budget_bytes))) {
        delete s;
        return st;
    }
    if (o->brick_x) s->shape.bx = o->brick_x;
    if (o->brick_lz) s->shape.lz = o->brick_lz;
    s->slots.resize(s->opt.slots);
    for (auto &sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
            delete s;
            return gc::fail(GPUCAD_E_HIP, "server stream creation failed");
        }
    }
    s->t0 = gc::now_s();
    *out = new gc_server{s};
    return GPUCAD_OK;
}

int gpucad_server_submit(gc_server *h, const gc_lattice_params *p, uint64_t *ticket) {
    gc::clear_error();
    GC_ARG(h && h->impl && p && ticket, "NULL argument");
    int st = gc::validate_batch_grid(*p, 0);
    if (st) return st;
    gc::Server *s = h->impl;
    gc::Request r;
    r.ticket = s->next_ticket++;
    r.p = *p;
    r.submit_s = s->clock();
    s->reqs.emplace(r.ticket, r);
    s->fifo.push_back(r.ticket);
    ++s->submitted;
    *ticket = r.ticket;
    return GPUCAD_OK;
}

int gpucad_server_poll(gc_server *h, gc_request_result *done, int32_t max, int32_t *n_done) {
    gc::clear_error();
    GC_ARG(h && h->impl && n_done && (done || max == 0), "NULL argument");
    gc::Server *s = h->impl;
    *n_done = 0;
    int st = s->advance();
    if (st) return st;
    while (*n_done < max && !s->finished.empty()) {
        const uint64_t t = s->finished.front();
        s->finished.pop_front();
        auto it = s->reqs.find(t);
        if (it == s->reqs.end()) continue;
        s->fill_result(it->second, &done[*n_done]);
        ++*n_done;
    }
    return GPUCAD_OK;
}

int gpucad_server_release(gc_server *h, uint64_t ticket) {
    gc::clear_error();
    GC_ARG(h && h->impl, "server is NULL");
    return h->impl->release(ticket);
}

int gpucad_server_stats(const gc_server *h, gc_server_stats *o) {
    gc::clear_error();
    GC_ARG(h && h->impl && o, "NULL argument");
    const gc::Server *s = h->impl;
    o->submitted = s->submitted;
    o->completed = s->completed;
    o->rejected = s->rejected;
    o->groups = s->n_groups;
    o->bytes_in_use = s->arena.used();
    o->peak_bytes_in_use = s->arena.peak();
    o->budget_bytes = s->arena.size();
    o->mean_group_size = s->n_groups ? (double)s->grouped / (double)s->n_groups : 0.0;
    return GPUCAD_OK;
}

int gpucad_server_destroy(gc_server *h) {
    gc::clear_error();
    if (!h) return GPUCAD_OK;
    delete h->impl;
    delete h;
    return GPUCAD_OK;
}

int gpucad_server_run_trace(gc_server *h, const gc_lattice_params *reqs, const double *arrival_s,
                            int32_t n, double *latency_s, double *wall_s) {
    gc::clear_error();
    GC_ARG(h && h->impl && reqs && arrival_s && latency_s && wall_s && n >= 0, "NULL argument");
    for (int i = 1; i < n; ++i) GC_ARG(arrival_s[i] >= arrival_s[i - 1], "arrivals must be sorted");
    gc::Server *s = h->impl;
    std::unordered_map<uint64_t, int> index;
    const double start = s->clock();
    int next = 0, done = 0;
    std::vector<gc_request_result> out(64);
    while (done < n) {
        const double t = s->clock() - start;
        while (next < n && arrival_s[next] <= t) {
            uint64_t ticket = 0;
            int st = gpucad_server_submit(h, &reqs[next], &ticket);
            if (st) return st;
            s->reqs[ticket].submit_s = start + arrival_s[next];
            index[ticket] = next;
            ++next;
        }
        int32_t got = 0;
        int st = gpucad_server_poll(h, out.data(), (int32_t)out.size(), &got);
        if (st) return st;
        for (int k = 0; k < got; ++k) {
            const int i = index[out[k].ticket];
            latency_s[i] = out[k].status == GPUCAD_OK ? out[k].done_s - out[k].submit_s : -1.0;
            if (out[k].status == GPUCAD_OK) {
                st = gpucad_server_release(h, out[k].ticket);
                if (st) return st;
            } else {
                s->reqs.erase(out[k].ticket);
            }
            ++done;
        }
    }
    *wall_s = s->clock() - start;
    return GPUCAD_OK;
}

int gpucad_plain_run_trace(const gc_lattice_params *reqs, const double *arrival_s, int32_t n,
                           float iso, double *latency_s, double *wall_s) {
    gc::clear_error();
    GC_ARG(reqs && arrival_s && latency_s && wall_s && n >= 0, "NULL argument");
    uint64_t most = 0, most_words = 0, most_ws = 0;
    for (int i = 0; i < n; ++i) {
        int st = gc::validate_batch_grid(reqs[i], i);
        if (st) return st;
        const gc_lattice_params &p = reqs[i];
        most = std::max(most, (uint64_t)p.nx * p.ny * p.nz);
        most_words = std::max(most_words, gpucad_sign_words(p.nx, p.ny, p.nz));
        most_ws = std::max(most_ws, gpucad_extract_workspace_bytes(p.nx, p.ny, p.nz));
    }
    if (!n) {
        *wall_s = 0;
        return GPUCAD_OK;
    }
    uint64_t vcap = std::max<uint64_t>(most / 2, 1 << 16), tcap = 2 * vcap;
    float *field = nullptr, *verts = nullptr;
    uint64_t *signs = nullptr;
    uint32_t *tris = nullptr;
    void *ws = nullptr;
    auto cleanup = [&]() {
        if (field) (void)hipFree(field);
        if (signs) (void)hipFree(signs);
        if (verts) (void)hipFree(verts);
        if (tris) (void)hipFree(tris);
        if (ws) (void)hipFree(ws);
    };
    if (hipMalloc(&field, most * 4) != hipSuccess || hipMalloc(&signs, most_words * 8) != hipSuccess ||
        hipMalloc(&verts, vcap * 12) != hipSuccess || hipMalloc(&tris, tcap * 12) != hipSuccess ||
        hipMalloc(&ws, most_ws) != hipSuccess) {
        cleanup();
        return gc::fail(GPUCAD_E_OOM, "plain buffers");
    }
    hipStream_t stream = nullptr;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
        cleanup();
        return gc::fail(GPUCAD_E_HIP, "stream");
    }
    const double start = gc::now_s();
    int st = GPUCAD_OK;
    for (int i = 0; i < n && !st; ++i) {
        while (gc::now_s() - start < arrival_s[i]) {
        }
        const gc_lattice_params &p = reqs[i];
        st = gpucad_lattice_field(&p, field, signs, iso, stream);
        if (st) break;
        gc_extract_args a;
        std::memset(&a, 0, sizeof(a));
        a.field = field;
        a.signs = signs;
        a.nx = p.nx;
        a.ny = p.ny;
        a.nz = p.nz;
        a.iso = iso;
        a.spacing = gc_float3{(float)(1.0 / (p.nx - 1)), (float)(1.0 / (p.ny - 1)),
                              (float)(1.0 / (p.nz - 1))};
        a.vertices = verts;
        a.vert_capacity = vcap;
        a.triangles = tris;
        a.tri_capacity = tcap;
        a.workspace = ws;
        a.workspace_bytes = most_ws;
        gc_mesh_counts c;
        st = gpucad_extract(&a, &c, stream);
        if (st == GPUCAD_E_OVERFLOW && c.n_vertices < 0xffffffffull) {
            // grow once and redo this request
            (void)hipFree(verts);
            (void)hipFree(tris);
            verts = nullptr;
            tris = nullptr;
            vcap = std::max(vcap, c.n_vertices);
            tcap = std::max(tcap, c.n_triangles);
            if (hipMalloc(&verts, vcap * 12) != hipSuccess ||
                hipMalloc(&tris, tcap * 12) != hipSuccess) {
                st = gc::fail(GPUCAD_E_OOM, "plain mesh buffers");
                break;
            }
            a.vertices = verts;
            a.vert_capacity = vcap;
            a.triangles = tris;
            a.tri_capacity = tcap;
            st = gpucad_extract(&a, &c, stream);
        }
        if (st) break;
        latency_s[i] = (gc::now_s() - start) - arrival_s[i];
    }
    *wall_s = gc::now_s() - start;
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
    cleanup();
    return st;
}

}  // extern "C"
