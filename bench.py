#!/usr/bin/env python3
"""Benchmark of the implicit-geometry hot path on one MI355X.

A step is one pass of the path over one synthetic grid: evaluate the
spatially varying Schwarz-P ∩ sphere lattice at 512^3 (FP32, sign words as a
by-product) and extract its isosurface into indexed vertex/index buffers
(marching cubes with shared-edge dedup).  Inputs are resident in HBM; output
buffers are preallocated from a warm-up run's counts.

Before the warm-up steps, untimed steps run in chunks of 10 until the time
per step is steady (the GPU's clocks ramp up over the first tens of
milliseconds of load); the timed steps then measure the steady state.

Prints one JSON line:
  value       Gvoxels/s = grid points / (field + extraction) time per step
  roofline    the step's longest launch, measured live with HIP events on its
              stream (both long kernels are under "kernels"):
              - the emit launch of the extraction (HBM-bound): the
                extraction's algorithmic bytes (4 B per point for the field
                read once, 12 B per vertex and per triangle written) / its
                time from an event the extraction records just before it to
                the event after the pass, against the HBM peak;
              - the field kernel (VALU-bound): the FP32 flops it executes per
                launch (rocprofv3 VALU instruction counters,
                profiles/valu_counters_per_dispatch.txt) / its launch time,
                against the FP32 vector peak, with the flop/point contract's
                count (hoisted per-axis terms included) as "accounted";
              the whole extraction pass's HBM figure is reported beside them,
              with the rates a plain device fill and copy reach in the same run
  cpu_baseline the oracle (host restatement of the reference arithmetic, OpenMP)
              timed on this host on a bounded z-slab of the same workload
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

# thread placement of the CPU baseline's OpenMP runtime (read when the oracle
# library loads, so set before anything imports it)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import gpucadforam_amd as gc  # noqa: E402
from gpucadforam_amd import configs as CF  # noqa: E402

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# algorithmic flops per point of the varying Schwarz-P workload (charges: FMA 2,
# add/mul/min/max/abs/div/sqrt 1, sin or cos 16, sincos 24): coordinates 3,
# rotation angle 1 + sincos 24, rotate 6, wavenumber 3, arguments 3, 3 cos 48,
# sum 2, thickness 8 x (5 + 1 + 16 + 2) + 2 = 194, sheet 2, sphere 10, max 1
FLOP_PER_POINT = {"varying_schwarz_p": 297, "plain_gyroid": 82, "bent_diamond": 144}


def measured_traffic(kernel, workload):
    """HBM bytes per dispatch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (profiles/hbm_traffic_per_dispatch.txt), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_per_dispatch.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 6 and not line.startswith("#") and f[0] == kernel and f[1] == workload:
            return int(round((float(f[2]) * float(f[4]) + float(f[3]) * float(f[5])) * 1024))
    return None


def executed_flops(kernel, workload):
    """FP32 flops one dispatch of `kernel` on `workload` executes, from the
    committed rocprofv3 VALU counters (profiles/valu_counters_per_dispatch.txt):
    64 lanes x (2 FMA + ADD + MUL + TRANS wave-instructions), or None."""
    path = os.path.join(ROOT, "profiles", "valu_counters_per_dispatch.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 10 and not line.startswith("#") and f[0] == kernel and f[1] == workload:
            fma, add, mul, trans = (int(v) for v in f[3:7])
            return {"flops": 64 * (2 * fma + add + mul + trans), "valu_wave_instructions": int(f[2]),
                    "fma_f32": fma, "add_f32": add, "mul_f32": mul, "trans_f32": trans}
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--settle", type=float, default=0.5,
                    help="seconds of untimed steps before the warm-up (0: none, for profiler runs)")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-slab", type=int, default=64)
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--calib", action="store_true",
                    help="also run a 512 MiB device copy (counter calibration)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched form")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the drop-in meshing of the synthesised lattice")
    ap.add_argument("--no-serve", action="store_true", help="skip the served form")
    ap.add_argument("--no-svl", action="store_true",
                    help="skip the spatially variant lattice synthesis")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 128^3 plain gyroid and 1024^3 bent diamond workloads")
    ap.add_argument("--batch-runs", type=int, default=10)
    ap.add_argument("--serve-requests", type=int, default=1000)
    return ap.parse_args()


class HipEvent:
    """A HIP timing event created with hipEventDisableSystemFence: recording
    it writes a timestamp once the stream's earlier work is done, without the
    system-scope cache writeback and invalidation of an ordinary event, so
    events between the passes of a step time the passes without lengthening
    them.  (The process has one HIP runtime, torch's, which libgpucad.so
    binds too.)"""
    _hip = None
    DISABLE_SYSTEM_FENCE = 0x20000000

    def __init__(self):
        if HipEvent._hip is None:
            h = C.CDLL("libamdhip64.so.7")
            h.hipEventCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
            h.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
            h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
            h.hipEventDestroy.argtypes = [C.c_void_p]
            HipEvent._hip = h
        self.e = C.c_void_p()
        st = self._hip.hipEventCreateWithFlags(C.byref(self.e), self.DISABLE_SYSTEM_FENCE)
        if st != 0:
            raise RuntimeError("hipEventCreateWithFlags failed: %d" % st)

    def record(self, stream_handle):
        st = self._hip.hipEventRecord(self.e, stream_handle)
        if st != 0:
            raise RuntimeError("hipEventRecord failed: %d" % st)

    def ms_to(self, other):
        """Milliseconds from this event to `other` (both complete)."""
        f = C.c_float()
        st = self._hip.hipEventElapsedTime(C.byref(f), self.e, other.e)
        if st != 0:
            raise RuntimeError("hipEventElapsedTime failed: %d" % st)
        return f.value

    def __del__(self):
        if self._hip is not None and self.e:
            self._hip.hipEventDestroy(self.e)


def settle(pipe, stream, chunk=10, min_s=0.5, max_s=3.0, tol=0.01):
    """Runs untimed steps until the GPU's clocks have ramped up: chunks of
    `chunk` steps, each timed, for at least `min_s` seconds and then until the
    last three chunks' times per step agree within `tol` (at most `max_s`
    seconds).  A GPU coming out of idle runs its first tens of milliseconds
    of work at lower clocks (the first 20 steps of the 512^3 step read 0.43 ms
    per step, the steady state 0.36 ms).  Returns the number of chunks and the
    first and last chunks' times per step (ms)."""
    times = []
    t0 = time.perf_counter()
    while True:
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(chunk):
            pipe.field_pass()
            pipe.extract_pass()
        b.record(stream)
        b.synchronize()
        times.append(a.elapsed_time(b) / chunk)
        el = time.perf_counter() - t0
        last = times[-3:]
        if el >= max_s or (el >= min_s and len(last) == 3 and
                           max(last) - min(last) <= tol * min(last)):
            break
    return {"chunks": len(times), "steps_per_chunk": chunk, "first_ms": round(times[0], 4),
            "last_ms": round(times[-1], 4)}


def timed_steps(pipe, stream, steps, rounds=3):
    """(ms per step, field ms, extraction ms, emit ms): first `steps` steps
    bracketed by two events only, then `rounds` times as many again with a
    fenceless event before and after each pass and one that the extraction
    records just before its emit launch (its last launch); each time is the
    median over the rounds of its mean per step (one round can be disturbed
    on its own)."""
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(steps):
        pipe.field_pass()
        pipe.extract_pass()
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / steps
    h = pipe.h
    field, extract, emit = [], [], []
    for _ in range(rounds):
        ev = [[HipEvent() for _ in range(4)] for _ in range(steps)]
        for k in range(steps):
            ev[k][0].record(h)
            pipe.field_pass()
            ev[k][1].record(h)
            pipe.args.emit_event = ev[k][3].e.value
            pipe.extract_pass()
            ev[k][2].record(h)
        pipe.args.emit_event = None
        torch.cuda.synchronize()
        field.append(sum(e[0].ms_to(e[1]) for e in ev) / steps)
        extract.append(sum(e[1].ms_to(e[2]) for e in ev) / steps)
        emit.append(sum(e[3].ms_to(e[2]) for e in ev) / steps)
    return (ms, statistics.median(field), statistics.median(extract),
            statistics.median(emit))


class Pipeline:
    def __init__(self, p, stream):
        self.p = p
        self.stream = stream
        n = p.npoints()
        self.field = torch.empty(n, dtype=torch.float32, device="cuda")
        self.signs = torch.empty(gc.sign_words(p.nx, p.ny, p.nz), dtype=torch.int64,
                                 device="cuda")
        wsb = gc.extract_workspace_bytes(p.nx, p.ny, p.nz)
        self.ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
        self.counts = torch.zeros(2, dtype=torch.int64, device="cuda")
        # size outputs from one counted run
        gc.lattice_field(p, out=self.field, signs=self.signs, stream=stream)
        m = gc.extract(self.field, p.nx, p.ny, p.nz, signs=self.signs, workspace=self.ws,
                       stream=stream)
        self.nv, self.nt = m.n_vertices, m.n_triangles
        self.verts = torch.empty((self.nv + 1024, 3), dtype=torch.float32, device="cuda")
        self.tris = torch.empty((self.nt + 1024, 3), dtype=torch.int32, device="cuda")
        a = gc.ExtractArgs()
        a.field, a.signs = self.field.data_ptr(), self.signs.data_ptr()
        a.nx, a.ny, a.nz = p.nx, p.ny, p.nz
        a.iso = 0.0
        a.origin = gc.Float3(0, 0, 0)
        a.spacing = gc.Float3(1 / (p.nx - 1), 1 / (p.ny - 1), 1 / (p.nz - 1))
        a.vertices, a.vert_capacity = self.verts.data_ptr(), self.verts.shape[0]
        a.triangles, a.tri_capacity = self.tris.data_ptr(), self.tris.shape[0]
        a.normals = None
        a.workspace, a.workspace_bytes = self.ws.data_ptr(), wsb
        a.d_counts = self.counts.data_ptr()
        self.args = a
        self.h = gc.stream_handle(stream)

    def field_pass(self):
        gc.check(gc.lib.gpucad_lattice_field(gc.C.byref(self.p), gc.ptr(self.field),
                                             gc.ptr(self.signs), 0.0, self.h))

    def extract_pass(self):
        gc.check(gc.lib.gpucad_extract(gc.C.byref(self.args), None, self.h))


def bench_pipelined(p, steps, warmup, depth):
    """Whole-job throughput with `depth` grids in flight: grid k runs field +
    extraction on stream k % depth with its own buffers, so one grid's
    VALU-bound field pass overlaps another's latency-bound extraction.  Every
    grid is computed in full; the per-grid device time of the headline is not
    this."""
    streams = [torch.cuda.Stream() for _ in range(depth)]
    pipes = [Pipeline(p, s) for s in streams]
    for _ in range(warmup):
        for q in pipes:
            q.field_pass()
            q.extract_pass()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    cur = torch.cuda.current_stream()
    start.record(cur)
    for q in pipes:  # every stream starts after the start event
        q.stream.wait_event(start)
    for k in range(steps):
        q = pipes[k % depth]
        q.field_pass()
        q.extract_pass()
    for q in pipes:
        cur.wait_stream(q.stream)
    end.record(cur)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / steps
    counts = [q.counts.cpu().tolist() for q in pipes]
    del pipes
    torch.cuda.empty_cache()
    return {"grids_in_flight": depth, "steps": steps, "ms_per_grid": round(ms, 4),
            "gvoxels_per_s": round(p.npoints() / (ms * 1e-3) / 1e9, 3),
            "meshes_identical": all(c == counts[0] for c in counts)}


def bench_config(p, name, steps, warmup):
    """Field + indexed extraction of one more single-grid workload (same
    timing as the headline step)."""
    stream = torch.cuda.Stream()
    pipe = Pipeline(p, stream)
    for _ in range(warmup):
        pipe.field_pass()
        pipe.extract_pass()
    ms, field_ms, extract_ms, emit_ms = timed_steps(pipe, stream, steps)
    counts = pipe.counts.cpu().tolist()
    npts = p.npoints()
    flops = FLOP_PER_POINT[name] * npts
    ext_bytes = 4 * npts + 12 * counts[0] + 12 * counts[1]
    out = {"workload": name, "grid": [p.nx, p.ny, p.nz], "steps": steps,
           "ms_per_step": round(ms, 4),
           "gvoxels_per_s": round(npts / (ms * 1e-3) / 1e9, 3),
           "field_ms": round(field_ms, 4),
           "field_tflops_accounted": round(flops / (field_ms * 1e-3) / 1e12, 2),
           "field_frac_accounted": round(flops / (field_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 4),
           "extract_ms": round(extract_ms, 4),
           "extract_GBs": round(ext_bytes / (extract_ms * 1e-3) / 1e9, 1),
           "extract_frac_hbm": round(ext_bytes / (extract_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
           "mtri_per_s": round(counts[1] / (extract_ms * 1e-3) / 1e6, 1),
           "n_vertices": counts[0], "n_triangles": counts[1]}
    del pipe
    torch.cuda.empty_cache()
    return out


def bench_svl(n=120, period=30.0, lattice_type="b"):
    """Spatially variant lattice synthesis at the reference's largest grid
    (coarse n^3, fine (2n)^3; 61^3 unit cell, 62 of 125 coefficients):
    the reference's per-harmonic loop (finding_phi, CG(500, 0.01), grating,
    svl, normalise_four per harmonic; src/main.cu:3949-4033 without the
    display extraction) through the drop-in entry points, against the
    one-call synthesis (three unit-phase solves + one fused pass) followed
    by one normalise_four."""
    dev = "cuda"
    cell = torch.empty(61 ** 3, dtype=torch.float32, device=dev)
    gc.check(gc.lib.gpucad_create_lattice(gc.ptr(cell), 61, 61, 61, 61 ** 3, 0))
    coeffs = gc.unit_spectrum(cell, 61, 2)
    p = gc.svl_params(n, lattice_type=lattice_type, uniform_type=0, const_period=period,
                      cg_end_res=0.01, cg_max_iter=500)
    fn = p.fx * p.fy * p.fz
    ind = torch.empty(fn, dtype=torch.float32, device=dev)
    nrm = torch.empty(fn, dtype=torch.float32, device=dev)

    def one_call():
        svl, res = gc.svl_field(p, coeffs)
        gc.check(gc.lib.gpucad_normalise_four(gc.ptr(svl), gc.ptr(ind), gc.ptr(nrm), fn, p.fx,
                                              p.fy, p.fz, 0.2, 0.3))
        torch.cuda.synchronize()
        return res

    one_call()
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        res = one_call()
        times.append(time.perf_counter() - t0)
    one_ms = 1e3 * min(times)

    phi = torch.empty(n ** 3, dtype=torch.float32, device=dev)
    grating = torch.empty(2 * fn, dtype=torch.float32, device=dev)
    svl = torch.zeros(fn, dtype=torch.float32, device=dev)
    it, r = C.c_int32(0), C.c_float(0)
    iters = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for h in range(p.harmonics):
        i, j, k = h % 5 - 2, (h // 5) % 5 - 2, h // 25 - 2
        gc.check(gc.lib.gpucad_svl_finding_phi(gc.ptr(phi), None, n, n, n, i, j, k, 1.0, 1.0, 1.0,
                                               lattice_type.encode(), 0, period, 30.0, 30.0, 30.0,
                                               0.16, 0.6, 0))
        gc.check(gc.lib.gpucad_svl_cg(gc.ptr(phi), n, n, n, 500, 0.01, C.byref(it), C.byref(r)))
        iters.append(it.value)
        gc.check(gc.lib.gpucad_svl_grating(gc.ptr(phi), n, n, n, gc.ptr(grating), p.fx, p.fy,
                                           p.fz, 0.5, 0.5, 0.5))
        gc.check(gc.lib.gpucad_svl_accumulate(gc.ptr(svl), gc.ptr(grating), fn, h,
                                              gc.ptr(coeffs)))
        gc.check(gc.lib.gpucad_normalise_four(gc.ptr(svl), gc.ptr(ind), gc.ptr(nrm), fn, p.fx,
                                              p.fy, p.fz, 0.2, 0.3))
    torch.cuda.synchronize()
    loop_ms = 1e3 * (time.perf_counter() - t0)
    return {"workload": "SVL synthesis, %s lattice, period %g, coarse %d^3, fine %d^3, "
                        "61^3 unit cell (type 0), %d harmonics" % (lattice_type, period, n,
                                                                   2 * n, p.harmonics),
            "one_call_ms": round(one_ms, 2),
            "one_call_cg_iters": list(res.cg_iters),
            "loop_ms": round(loop_ms, 1),
            "loop_cg_iters_mean": round(sum(iters) / len(iters), 1),
            "loop_cg_iters_max": max(iters),
            "speedup_vs_loop": round(loop_ms / one_ms, 1),
            "fine_gvoxels_per_s": round(fn / (one_ms * 1e-3) / 1e9, 3)}


def bench_dropin(n=120, period=30.0, runs=30, cpu=True):
    """The reference application's meshing of a synthesised sheet lattice
    (src/main.cu:3949-4010) through the drop-in entry points: the one-call
    synthesis on the fine (2n)^3 grid (not timed), then normalise_four (band
    0.20-0.30) and Isosurface::computeIsosurface_lattice with the caller's
    buffers as the reference allocates them (five uint[numVoxels] arrays,
    float4 pos/norm[maxVerts], maxVerts = 4 N^3), isoValue 0.25 on the
    indicator, vol_two zeros, voxel size 0.5 (src/main.cu:2137).  Both calls
    are synchronous, as the reference makes them; timed on the host clock.
    Algorithmic bytes: the three fields read once and 96 B per triangle
    (float4 pos + norm per corner).  CPU beside it: the oracle's
    restatement of the same two calls on one thread."""
    dev = "cuda"
    cell = torch.empty(61 ** 3, dtype=torch.float32, device=dev)
    gc.check(gc.lib.gpucad_create_lattice(gc.ptr(cell), 61, 61, 61, 61 ** 3, 0))
    coeffs = gc.unit_spectrum(cell, 61, 2)
    p = gc.svl_params(n, lattice_type="b", uniform_type=0, const_period=period,
                      cg_end_res=0.01, cg_max_iter=500)
    svl, _ = gc.svl_field(p, coeffs)
    nx, ny, nz = p.fx, p.fy, p.fz
    npts = nx * ny * nz
    cells = (nx - 1) * (ny - 1) * (nz - 1)
    max_verts = max(4 * npts, 300000)
    ind = torch.empty(npts, dtype=torch.float32, device=dev)
    nrm = torch.empty(npts, dtype=torch.float32, device=dev)
    zeros = torch.zeros(npts, dtype=torch.float32, device=dev)
    u = [torch.empty(cells, dtype=torch.int32, device=dev) for _ in range(5)]
    pos = torch.empty((max_verts, 4), dtype=torch.float32, device=dev)
    norm = torch.empty((max_verts, 4), dtype=torch.float32, device=dev)
    active, total = C.c_uint32(0), C.c_uint32(0)
    gs = gc.Uint3(nx, ny, nz)
    sh = gc.Uint3(1, nx - 1, (nx - 1) * (ny - 1))
    mk = gc.Uint3(nx - 1, ny - 1, nz - 1)
    vs, ce = gc.Float3(0.5, 0.5, 0.5), gc.Float3(0.0, 0.0, 0.0)

    def normalise():
        gc.check(gc.lib.gpucad_normalise_four(gc.ptr(svl), gc.ptr(ind), gc.ptr(nrm), npts, nx, ny,
                                              nz, 0.2, 0.3))

    def isosurface():
        gc.check(gc.lib.gpucad_compute_isosurface_lattice(
            gc.ptr(ind), gc.ptr(pos), gc.ptr(norm), 0.25, cells, gc.ptr(u[0]), gc.ptr(u[1]),
            gc.ptr(u[2]), gc.ptr(u[3]), gs, sh, mk, vs, ce, C.byref(active), C.byref(total),
            gc.ptr(u[4]), max_verts, gc.ptr(nrm), gc.ptr(zeros), 0.2, 0.3, 0.0, 0.0))

    # untimed calls first (about 0.3 s), as the reference application makes
    # them one after another; then `runs` timed pairs, medians reported
    t_end = time.perf_counter() + 0.3
    while time.perf_counter() < t_end:
        normalise()
        isosurface()
    torch.cuda.synchronize()
    t_norm, t_iso = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        normalise()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        isosurface()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_norm.append(t1 - t0)
        t_iso.append(t2 - t1)
    ms_norm = 1e3 * statistics.median(t_norm)
    ms_iso = 1e3 * statistics.median(t_iso)
    tris = total.value // 3
    nbytes = 3 * 4 * npts + 96 * tris
    lat_counts = (active.value, total.value)

    # boolean meshing (Isosurface::computeIsosurface after classify_copy_Voxel,
    # src/main.cu:3304-3465): a sphere primitive's grid_points united with the
    # lattice's band, "dynamic" mode (corner inside: F < 0 or 0.2 < L < 0.3)
    flags = 1 | 16
    idx = torch.arange(npts, device=dev)
    zc, yc, xc = idx // (nx * ny), (idx // nx) % ny, idx % nx
    sph = (torch.sqrt((xc - 0.5 * (nx - 1)) ** 2 + (yc - 0.5 * (ny - 1)) ** 2 +
                      (zc - 0.5 * (nz - 1)) ** 2) - 0.3 * nx).float().contiguous()
    del idx, zc, yc, xc
    g0 = torch.zeros((npts, 4), dtype=torch.int32, device=dev)  # grid_points {val, t_x, t_y, t_z}
    g0[:, 0] = torch.where(sph < 0, -1, 1).int()
    d_g = torch.empty_like(g0)
    nfac = C.c_uint64(0)

    def copy_pass():
        gc.check(gc.lib.gpucad_classify_copy_voxel(gc.ptr(d_g), gc.ptr(sph), gc.ptr(nrm), gs, 0.2,
                                                   0.3, 0.0, flags))

    def boolean():
        gc.check(gc.lib.gpucad_compute_isosurface(
            gc.ptr(pos), gc.ptr(norm), 0.0, cells, gc.ptr(u[0]), gc.ptr(u[1]), gc.ptr(u[2]),
            gc.ptr(u[3]), gs, sh, mk, vs, ce, C.byref(active), C.byref(total), gc.ptr(u[4]),
            max_verts, gc.ptr(d_g), gc.ptr(sph), gc.ptr(nrm), 0.2, 0.3, flags, C.byref(nfac)))

    t_copy, t_bool = [], []
    for r in range(runs + 1):
        d_g.copy_(g0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        copy_pass()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        boolean()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if r:
            t_copy.append(t1 - t0)
            t_bool.append(t2 - t1)
    ms_copy = 1e3 * statistics.median(t_copy)
    ms_bool = 1e3 * statistics.median(t_bool)
    btris = total.value // 3
    bool_counts = (active.value, total.value)
    out_bool = {"workload": "computeIsosurface: sphere primitive (r = 0.3 N, grid_points) united "
                            "with the lattice band 0.20-0.30, dynamic mode, %d^3" % nx,
                "classify_copy_ms": round(ms_copy, 3), "isosurface_ms": round(ms_bool, 3),
                "active_voxels": active.value, "total_verts": total.value,
                "mtri_per_s": round(btris / (ms_bool * 1e-3) / 1e6, 1)}
    out = {"workload": "drop-in computeIsosurface_lattice of the synthesised b lattice "
                       "(period %g, fine %d^3), band 0.20-0.30, isoValue 0.25" % (period, nx),
           "grid": [nx, ny, nz], "runs": runs,
           "normalise_four_ms": round(ms_norm, 3),
           "isosurface_ms": round(ms_iso, 3),
           "ms": round(ms_norm + ms_iso, 3),
           "gvoxels_per_s": round(npts / ((ms_norm + ms_iso) * 1e-3) / 1e9, 3),
           "active_voxels": lat_counts[0], "total_verts": lat_counts[1],
           "mtri_per_s": round(tris / (ms_iso * 1e-3) / 1e6, 1),
           "isosurface_GBs": round(nbytes / (ms_iso * 1e-3) / 1e9, 1),
           "isosurface_frac_hbm": round(nbytes / (ms_iso * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)}
    if cpu:
        import numpy as np
        from oracle import oracle as orc
        v = svl.cpu().numpy()
        t0 = time.perf_counter()
        r_ind, r_nrm = orc.normalise_four(v, nx, ny, nz, 0.2, 0.3)
        t1 = time.perf_counter()
        ref = orc.compute_isosurface(orc.MODE_BAND2, r_ind, nx, ny, nz, 0.25,
                                     voxel_size=(0.5, 0.5, 0.5), vol1=r_nrm,
                                     vol2=np.zeros_like(r_nrm), lo1=0.2, hi1=0.3,
                                     max_verts=max_verts)
        t2 = time.perf_counter()
        out["cpu_oracle"] = {"threads": orc.threads(), "normalise_four_ms": round(1e3 * (t1 - t0), 1),
                             "isosurface_ms": round(1e3 * (t2 - t1), 1),
                             "same_counts": (ref["active"], ref["total"]) == lat_counts}
        out["speedup_vs_cpu"] = round(1e3 * (t2 - t0) / (ms_norm + ms_iso), 1)
        s_np, l_np = sph.cpu().numpy(), nrm.cpu().numpy()
        g_np = np.zeros(npts, orc.GRID_POINT)
        g_np["val"] = np.where(s_np < 0, -1, 1)
        t0 = time.perf_counter()
        g_ref = orc.bool_copy(flags, g_np, s_np, l_np, nx, ny, nz, 0.0, 0.2, 0.3)
        t1 = time.perf_counter()
        b_ref = orc.bool_compute_isosurface(flags, g_ref, s_np, l_np, nx, ny, nz, 0.0, 0.2, 0.3,
                                            voxel_size=(0.5, 0.5, 0.5), max_verts=max_verts)
        t2 = time.perf_counter()
        out_bool["cpu_oracle"] = {"threads": orc.threads(), "classify_copy_ms": round(1e3 * (t1 - t0), 1),
                                  "isosurface_ms": round(1e3 * (t2 - t1), 1),
                                  "same_counts": (b_ref["active"], b_ref["total"]) == bool_counts}
        out_bool["speedup_vs_cpu"] = round(1e3 * (t2 - t0) / (ms_copy + ms_bool), 1)
    out["boolean"] = out_bool
    del pos, norm, u, d_g, g0
    torch.cuda.empty_cache()
    return out


def device_rates(mib=512, runs=20):
    """What plain streaming kernels reach on this device: torch's fill
    (writes only) and copy (reads and writes) of mib MiB, medians of runs
    launches timed with events; bytes moved / time."""
    n = mib << 18
    src = torch.empty(n, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)

    def median_ms(fn):
        for _ in range(3):
            fn()
        ts = []
        for _ in range(runs):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b))
        return statistics.median(ts)

    fill_ms = median_ms(lambda: dst.fill_(2.0))
    copy_ms = median_ms(lambda: dst.copy_(src))
    del src, dst
    torch.cuda.empty_cache()
    return {"mib": mib, "fill_write_GBs": round(4 * n / (fill_ms * 1e-3) / 1e9, 1),
            "copy_GBs": round(8 * n / (copy_ms * 1e-3) / 1e9, 1)}


def cpu_model():
    """The host CPU's model name (what lscpu prints as "Model name")."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(p, slab, threads, budget_s=12.0):
    """Oracle field + indexed extraction on z-slabs of the same workload: one
    warm-up slab, then timed slabs (median) until about budget_s of CPU work."""
    from oracle import oracle as orc
    nthreads = orc.threads(threads or None)
    spacing = (1 / (p.nx - 1), 1 / (p.ny - 1), 1 / (p.nz - 1))
    z1 = min(p.nz, slab + 1)
    f = orc.lattice_field_f32(p, 0, z1, nthreads)
    orc.extract_indexed(f, p.nx, p.ny, z1, spacing=spacing, nthreads=nthreads)
    runs, spent, z0 = [], 0.0, 0
    while spent < budget_s and len(runs) < 15:
        z0 = (z0 + slab) % max(1, p.nz - slab - 1)
        t0 = time.perf_counter()
        f = orc.lattice_field_f32(p, z0, z0 + slab + 1, nthreads)
        t1 = time.perf_counter()
        orc.extract_indexed(f, p.nx, p.ny, slab + 1, spacing=spacing, nthreads=nthreads)
        t2 = time.perf_counter()
        runs.append((t2 - t0, t1 - t0, t2 - t1))
        spent += t2 - t0
    runs.sort()
    tot, tf, te = runs[len(runs) // 2]
    pts = p.nx * p.ny * (slab + 1)
    return {
        "value": round(pts / tot / 1e9, 5),
        "unit": "Gvoxels/s",
        "cores": nthreads,
        "cpu_model": cpu_model(),
        "nproc": os.cpu_count(),
        "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
        "build": "g++ -O3 -march=%s -fopenmp -ffp-contract=off (oracle/Makefile; the ISA "
                 "level of this host's CPU)" % orc.ISA,
        "kind": "port",
        "sample": "median of %d z-slabs of %d planes of the %dx%dx%d workload (%d points each, "
                  "%.1f s total): oracle field %.3f s + indexed extraction %.3f s per slab"
                  % (len(runs), slab + 1, p.nx, p.ny, p.nz, pts, spent, tf, te),
    }


def fp64_errors(p, field_gpu, z0, planes):
    """Max |error| of the GPU's FP32 field and of its interpolated vertices
    against FP64 evaluation (the oracle's double-precision twin of the field,
    and the crossing parameter recomputed in double from it) on the z-slab
    [z0, z0 + planes) of the workload.  Vertex error in units of the point
    spacing; only edges whose sign pattern agrees in FP32 and FP64 compare."""
    import numpy as np
    from oracle import oracle as orc
    nx, ny = p.nx, p.ny
    plane = nx * ny
    f32 = field_gpu[z0 * plane:(z0 + planes) * plane].cpu().numpy()
    f64 = orc.lattice_field_f64(p, z0, z0 + planes)
    field_err = float(np.max(np.abs(f32.astype(np.float64) - f64)))
    verts, _, eids = orc.extract_indexed(f32, nx, ny, planes, spacing=(1.0, 1.0, 1.0),
                                         edge_ids=True)
    eids = eids.astype(np.int64)
    pt, ax = eids // 3, eids % 3
    step = np.array([1, nx, plane], np.int64)[ax]
    a, b = f64[pt], f64[pt + step]
    same = (a < 0) != (b < 0)
    t = (0.0 - a[same]) / (b[same] - a[same])
    x, y, z = pt % nx, (pt // nx) % ny, pt // plane
    ref = np.stack([x, y, z], 1).astype(np.float64)[same]
    ref[np.arange(len(t)), ax[same]] += t
    dv = np.max(np.abs(verts[same].astype(np.float64) - ref), axis=1) if len(t) else np.zeros(1)
    return {"slab": "z %d..%d of %dx%dx%d" % (z0, z0 + planes - 1, nx, ny, p.nz),
            "field_max_abs": field_err, "vertex_max_abs_voxels": float(dv.max()),
            "vertex_p99_abs_voxels": float(np.percentile(dv, 99)),
            "vertices_compared": int(len(t)), "vertices_sign_differs": int((~same).sum())}


def percentile(xs, q):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
    return xs[k]


def bench_batched(runs):
    """256 grids of uneven sizes and types: one graph replay per run (field +
    extraction of every grid) against 256 plain per-grid calls."""
    grids = CF.batched_grids(256)
    pts = sum(g.npoints() for g in grids)
    t0 = time.perf_counter()
    b = gc.Batch(grids, use_graph=True)
    create_s = time.perf_counter() - t0
    stream = torch.cuda.Stream()
    b.run(stream)
    torch.cuda.synchronize()
    res = b.results()
    tris = sum(r["n_triangles"] for r in res)
    verts = sum(r["n_vertices"] for r in res)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(runs):
        b.run(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / runs
    tuning = b.tuning()
    b.close()
    # plain: one gpucad_lattice_field + gpucad_extract (synchronous counts) per grid
    gc.plain_run_trace(grids[:8], [0.0] * 8)  # warm-up
    _, plain_s = gc.plain_run_trace(grids, [0.0] * len(grids))
    plain_ms = plain_s * 1e3
    return {
        "workload": "256 grids, per-axis sizes U{32,40,..,256}, type U{gyroid, Schwarz-P, "
                    "diamond}, period U[8,48] voxels, sheet thickness U[0.1,0.5], seed 0xBA7C",
        "points": pts,
        "n_vertices": verts,
        "n_triangles": tris,
        "ms_per_run": round(ms, 4),
        "gvoxels_per_s": round(pts / (ms * 1e-3) / 1e9, 3),
        "mtri_per_s": round(tris / (ms * 1e-3) / 1e6, 1),
        "plain_ms": round(plain_ms, 3),
        "plain_gvoxels_per_s": round(pts / (plain_ms * 1e-3) / 1e9, 3),
        "speedup_vs_plain": round(plain_ms / ms, 2),
        "graph": True,
        "tuning": {k: tuning[k] for k in ("brick_x", "brick_lz")},
        "create_s": round(create_s, 3),
    }


def bench_served(n_req):
    """Lattice-edit requests (cubic edge lognormal median 96, cap 384; random
    lattice + sphere/box boolean): saturation throughput (all requests queued
    at t=0) and p50/p99 latency under Poisson arrivals at fractions of it,
    for the server (continuous batching, 64 GB budget) and for plain
    per-request calls."""
    reqs = CF.served_requests(n_req)
    pts = sum(r.npoints() for r in reqs)
    budget = 64 << 30
    srv = gc.Server(budget)
    srv.run_trace(reqs[:32], [0.0] * 32)  # warm-up
    lat0, wall0 = srv.run_trace(reqs, [0.0] * n_req)
    sat_rps = n_req / wall0
    out = {
        "workload": "%d requests, cubic edge round8(min(384, lognormal(median 96, sigma 0.5))), "
                    "gyroid/Schwarz-P/diamond sheet + sphere or box boolean, seed 0x5E7E; "
                    "64 GB admission budget" % n_req,
        "points": pts,
        "saturation": {"requests_per_s": round(sat_rps, 1),
                       "gvoxels_per_s": round(pts / wall0 / 1e9, 3)},
        "rates": [],
    }
    gc.plain_run_trace(reqs[:16], [0.0] * 16)  # warm-up
    plat0, pwall0 = gc.plain_run_trace(reqs, [0.0] * n_req)
    plain_rps = n_req / pwall0
    out["plain_saturation"] = {"requests_per_s": round(plain_rps, 1),
                               "gvoxels_per_s": round(pts / pwall0 / 1e9, 3)}
    out["throughput_vs_plain"] = round(sat_rps / plain_rps, 2)
    # the two lowest rates lie below the plain path's saturation, where its
    # latency can be measured beside the server's
    for frac in (0.05, 0.1, 0.25, 0.5, 0.8):
        rate = frac * sat_rps
        arr = CF.poisson_arrivals(n_req, rate)
        lat, wall = srv.run_trace(reqs, arr)
        row = {"rate_rps": round(rate, 1), "fraction_of_saturation": frac,
               "p50_ms": round(1e3 * percentile(lat, 50), 3),
               "p99_ms": round(1e3 * percentile(lat, 99), 3),
               "achieved_rps": round(n_req / wall, 1)}
        if rate < plain_rps:  # the plain path can only be compared where it keeps up
            plat, pwall = gc.plain_run_trace(reqs, arr)
            row["plain_p50_ms"] = round(1e3 * percentile(plat, 50), 3)
            row["plain_p99_ms"] = round(1e3 * percentile(plat, 99), 3)
        out["rates"].append(row)
    st = srv.stats()
    out["server"] = {"groups": st["groups"], "mean_group_size": round(st["mean_group_size"], 2),
                     "peak_bytes_in_use": st["peak_bytes_in_use"], "budget_bytes": budget}
    srv.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    torch.cuda.init()
    gc.init(0)
    p = CF.varying_schwarz_p(args.n)
    stream = torch.cuda.Stream()
    pipe = Pipeline(p, stream)
    settled = settle(pipe, stream, min_s=args.settle) if args.settle > 0 else None
    for _ in range(args.warmup):
        pipe.field_pass()
        pipe.extract_pass()
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    ms, field_ms, extract_ms, emit_ms = timed_steps(pipe, stream, args.steps)
    wall = time.perf_counter() - t_wall
    if args.calib:
        src = torch.empty(128 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    counts = pipe.counts.cpu().tolist()
    rates = device_rates()
    npts = p.npoints()
    flops = FLOP_PER_POINT["varying_schwarz_p"] * npts
    field_tflops = flops / (field_ms * 1e-3) / 1e12
    exe = executed_flops("lattice_field_kernel", "varying_schwarz_p_%d" % p.nx)
    exe_tflops = None if exe is None else exe["flops"] / (field_ms * 1e-3) / 1e12
    ext_bytes = 4 * npts + 12 * counts[0] + 12 * counts[1]
    ext_gbs = ext_bytes / (extract_ms * 1e-3) / 1e9
    # the field kernel is bound by the FP32 vector ALUs (no matrix work on
    # this path): its roofline is the FP32 vector peak, against the flops the
    # kernel executes (counters)
    field_rl = {"bound": "valu",
                "achieved": None if exe is None else round(exe_tflops, 2),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": None if exe is None else round(exe_tflops / PEAK_FP32_TFLOPS, 4),
                "traffic": measured_traffic("lattice_field_kernel", "varying_schwarz_p_%d" % p.nx),
                "kernel": "lattice_field_kernel",
                "flops_per_launch": None if exe is None else exe["flops"],
                "flops_source": "rocprofv3 SQ_INSTS_VALU_{FMA,ADD,MUL,TRANS}_F32 x 64 "
                                "(profiles/valu_counters_per_dispatch.txt)",
                "valu_wave_instructions": None if exe is None else exe["valu_wave_instructions"],
                "launch_ms": round(field_ms, 4),
                # VALU utilisation: the launch's VALU wave-instructions over the
                # issue slots of 1024 SIMDs in the live launch time at one wave64
                # instruction per 2 cycles and the 2.4 GHz peak clock (a lower
                # bound: the clock under load is lower)
                "valu_issue_frac": None if exe is None else round(
                    exe["valu_wave_instructions"] / (1024 * field_ms * 1e-3 * 2.4e9 / 2), 4),
                "accounted": {"flop_per_point": FLOP_PER_POINT["varying_schwarz_p"],
                              "tflops": round(field_tflops, 2),
                              "frac": round(field_tflops / PEAK_FP32_TFLOPS, 4)}}
    # the emit launch (vertex and triangle stages) is bound by HBM: one launch
    # processes the grid's points, vertices and triangles, credited at the
    # extraction's per-unit bytes (the field read once, 4 B per point; 12 B
    # per vertex and per triangle written); its time runs from an event the
    # extraction records before the launch to the event after the pass
    emit_gbs = ext_bytes / (emit_ms * 1e-3) / 1e9
    emit_rl = {"bound": "hbm", "achieved": round(emit_gbs, 1), "peak": PEAK_HBM_GBS,
               "unit": "GB/s", "frac": round(emit_gbs / PEAK_HBM_GBS, 4),
               "traffic": measured_traffic("emit_kernel", "varying_schwarz_p_%d" % p.nx),
               "kernel": "emit_kernel", "algorithmic_bytes": ext_bytes,
               "bytes_per_unit": {"point": 4, "vertex": 12, "triangle": 12},
               "launch_ms": round(emit_ms, 4)}
    rec = {
        "metric": "Gvoxels/s field+marching-cubes at 512^3",
        "value": round(npts / (ms * 1e-3) / 1e9, 3),
        "unit": "Gvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded lattice parameters, seed 0x6770)",
        "config": {"workload": "varying_schwarz_p %dx%dx%d: period 16->64 voxels along x, "
                               "rotation 0->90 deg along z, 8-mode Fourier thickness 0.2-0.6, "
                               "sheet, intersect sphere r=0.45; indexed extraction"
                               % (p.nx, p.ny, p.nz),
                   "grid": [p.nx, p.ny, p.nz]},
        # the step's longest launch (measured live): the roofline of the
        # contract is that kernel's; both long kernels are under "kernels"
        "roofline": field_rl if field_ms > emit_ms else emit_rl,
        "kernels": {"lattice_field_kernel": field_rl, "emit_kernel": emit_rl},
        # beside the HBM peak, what a plain copy (reads and writes) and fill
        # (writes) reach on this device in the same run
        "extraction": {"ms": round(extract_ms, 4), "algorithmic_bytes": ext_bytes,
                       "achieved_GBs": round(ext_gbs, 1),
                       "frac_hbm": round(ext_gbs / PEAK_HBM_GBS, 4),
                       "device_rates": rates,
                       "frac_of_copy": round(ext_gbs / rates["copy_GBs"], 4)},
        "mtri_per_s": round(counts[1] / (extract_ms * 1e-3) / 1e6, 1),
        "n_vertices": counts[0],
        "n_triangles": counts[1],
        "wall_s": round(wall, 3),
        # untimed chunks of 10 steps run before the warm-up steps until the
        # time per step is steady (the GPU's clocks ramp up under load)
        "settle": settled,
    }
    if not args.no_cpu:
        rec["fp64_error"] = fp64_errors(p, pipe.field, p.nz // 2 - 16, 33)
    del pipe
    torch.cuda.empty_cache()
    if not args.no_configs:
        print("bench: pipelined grids", file=sys.stderr, flush=True)
        rec["pipelined"] = [bench_pipelined(p, args.steps, args.warmup, d) for d in (2, 3)]
        print("bench: other single-grid workloads", file=sys.stderr, flush=True)
        rec["other_configs"] = [bench_config(CF.plain_gyroid(128), "plain_gyroid", 20, 3),
                                bench_config(CF.bent_diamond(1024), "bent_diamond", 5, 1)]
    if not args.no_svl:
        print("bench: spatially variant lattice synthesis", file=sys.stderr, flush=True)
        rec["svl"] = bench_svl()
    if not args.no_dropin:
        print("bench: drop-in meshing of the synthesised lattice", file=sys.stderr, flush=True)
        rec["dropin"] = bench_dropin(cpu=not args.no_cpu)
    if not args.no_batch:
        print("bench: batched form", file=sys.stderr, flush=True)
        rec["batched"] = bench_batched(args.batch_runs)
    if not args.no_serve:
        print("bench: served form", file=sys.stderr, flush=True)
        rec["served"] = bench_served(args.serve_requests)
    if not args.no_cpu:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        rec["cpu_baseline"] = cpu_baseline(p, args.cpu_slab, args.cpu_threads, args.cpu_budget)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
