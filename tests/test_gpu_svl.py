"""Spatially variant lattice synthesis on the GPU against the oracle.

Bit-exact where the kernel and the oracle share the operation sequence
(period/angle data, normalisations, finding_phi, the operator, upsampling,
grating, accumulation, FFT fill/scale); conjugate gradients and the one-call
synthesis to a tolerance (parallel dot products round differently), with the
residual criterion checked in FP64.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def to_dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def test_period_angle_data_bit_exact(gc, orc, dev):
    nx, ny, nz = 20, 17, 9
    m = (10.0, 8.5, 4.0)
    for axis in "zyx":
        for angle in (False, True):
            out = torch.empty(nx * ny * nz, dtype=torch.float32, device=dev)
            fn = gc.lib.gpucad_svl_angle_data if angle else gc.lib.gpucad_svl_period_data
            gc.check(fn(gc.ptr(out), nx, ny, nz, 1.0, 0.5, 2.0, *m, axis.encode()))
            ref = orc.svl_period_angle(nx, ny, nz, 1.0, 0.5, 2.0, m, axis, angle)
            assert np.array_equal(out.cpu().numpy(), ref), (axis, angle)


def test_normalise_three_and_buffer_bit_exact(gc, orc, dev):
    rng = np.random.default_rng(11)
    for lo, hi in ((-3.0, 5.0), (1.0, 4.0), (-6.0, -1.0)):
        a = rng.uniform(lo, hi, 10007).astype(np.float32)
        d = to_dev(a, dev)
        o = torch.empty_like(d)
        gc.check(gc.lib.gpucad_normalise_three(gc.ptr(d), gc.ptr(o), a.size, 6.0, 15.0))
        assert np.array_equal(o.cpu().numpy(), orc.svl_rescale(a, 6.0, 15.0))
        gc.check(gc.lib.gpucad_normalise_buffer(gc.ptr(d), gc.ptr(o), a.size))
        assert np.array_equal(o.cpu().numpy(), orc.svl_rescale(a, affine=False))
        # in place, as the reference calls it
        gc.check(gc.lib.gpucad_normalise_three(gc.ptr(d), gc.ptr(d), a.size, 6.0, 15.0))
        assert np.array_equal(d.cpu().numpy(), orc.svl_rescale(a, 6.0, 15.0))


CASES = [
    ("n", 0, False), ("n", 1, False), ("b", 0, False), ("r", 0, False), ("r", 2, False),
    ("s", 0, False), ("s", 1, True), ("b", 2, False),
]


@pytest.mark.parametrize("lt,ut,sine", CASES)
def test_finding_phi_bit_exact(gc, orc, dev, lt, ut, sine):
    nx, ny, nz = 13, 11, 9
    period = None
    dper = None
    if ut == 2:
        period = orc.svl_rescale(orc.svl_period_angle(nx, ny, nz, 1.0, 1.0, 1.0, (0, 0, 0), "z"),
                                 1.0, 3.0)
        dper = to_dev(period, dev)
    for ijk in ((1, 0, 0), (-2, 1, 2), (0, -1, 1)):
        out = torch.empty(nx * ny * nz, dtype=torch.float32, device=dev)
        gc.check(gc.lib.gpucad_svl_finding_phi(gc.ptr(out), gc.ptr(dper), nx, ny, nz, *ijk, 1.0,
                                               1.0, 1.0, lt.encode(), ut, 12.0, 9.0, 10.0, 11.0,
                                               0.16, 0.6, 1 if sine else 0))
        ref = orc.svl_finding_phi(nx, ny, nz, *ijk, lattice_type=lt, uniform_type=ut,
                                  const_period=12.0, periods=(9.0, 10.0, 11.0), period=period,
                                  sinewave_zaxis=sine)
        assert np.array_equal(out.cpu().numpy(), ref), (lt, ut, ijk)


@pytest.mark.parametrize("dims", [(4, 4, 4), (16, 9, 7), (61, 61, 61)])
def test_operator_bit_exact(gc, orc, dev, dims):
    nx, ny, nz = dims
    d = np.random.default_rng(12).standard_normal(nx * ny * nz).astype(np.float32)
    dd = to_dev(d, dev)
    q = torch.empty_like(dd)
    gc.check(gc.lib.gpucad_svl_matvec(gc.ptr(dd), gc.ptr(q), nx, ny, nz))
    assert np.array_equal(q.cpu().numpy(), orc.svl_matvec(d, nx, ny, nz))


def residual(orc, b, x, dims):
    q = orc.svl_matvec(x, *dims).astype(np.float64)
    return np.linalg.norm(b.astype(np.float64) - q)


@pytest.mark.parametrize("lt,n", [("n", 12), ("b", 16), ("r", 24)])
def test_cg_against_oracle(gc, orc, dev, lt, n):
    dims = (n, n, n)
    b = orc.svl_finding_phi(n, n, n, 1, 2, 1, lattice_type=lt, uniform_type=0, const_period=9.0)
    xo, ito, reso = orc.svl_cg(b, *dims, iters=500, end_res=0.01)
    d = to_dev(b, dev)
    it, res = C.c_int32(0), C.c_float(0)
    gc.check(gc.lib.gpucad_svl_cg(gc.ptr(d), *dims, 500, 0.01, C.byref(it), C.byref(res)))
    x = d.cpu().numpy()
    # the reference's criterion, and what it reports
    if it.value < 500:
        assert res.value <= 0.01
    assert abs(residual(orc, b, x, dims) - res.value) < 1e-3 + 0.05 * res.value
    assert abs(it.value - ito) <= max(2, ito // 20)
    scale = np.max(np.abs(xo))
    assert np.max(np.abs(x - xo)) < 2e-2 * scale


def test_cg_iteration_cap_and_immediate_stop(gc, orc, dev):
    n = 8
    b = orc.svl_finding_phi(n, n, n, 1, 0, 0, lattice_type="b", uniform_type=0, const_period=8.0)
    for iters, end_res in ((1, 0.0), (4, 0.0), (500, 1e9)):
        d = to_dev(b, dev)
        it, res = C.c_int32(0), C.c_float(0)
        gc.check(gc.lib.gpucad_svl_cg(gc.ptr(d), n, n, n, iters, end_res, C.byref(it),
                                      C.byref(res)))
        xo, ito, reso = orc.svl_cg(b, n, n, n, iters=iters, end_res=end_res)
        assert it.value == ito
        if ito == 1:
            assert np.all(d.cpu().numpy() == 0)
        assert abs(res.value - reso) <= 1e-4 * max(1.0, reso)


def test_refine_and_grating_bit_exact(gc, orc, dev):
    nx, ny, nz = 9, 7, 6
    c = np.random.default_rng(13).uniform(-20, 20, nx * ny * nz).astype(np.float32)
    dc = to_dev(c, dev)
    for fine, h in (((18, 14, 12), (0.5, 0.5, 0.5)), ((13, 20, 7), (0.7, 0.33, 0.9))):
        n = fine[0] * fine[1] * fine[2]
        out = torch.empty(n, dtype=torch.float32, device=dev)
        gc.check(gc.lib.gpucad_svl_refine(gc.ptr(dc), nx, ny, nz, gc.ptr(out), *fine, *h))
        assert np.array_equal(out.cpu().numpy(), orc.svl_upsample(c, nx, ny, nz, fine, h))
        g = torch.empty(2 * n, dtype=torch.float32, device=dev)
        gc.check(gc.lib.gpucad_svl_grating(gc.ptr(dc), nx, ny, nz, gc.ptr(g), *fine, *h))
        assert np.array_equal(g.cpu().numpy(), orc.svl_upsample(c, nx, ny, nz, fine, h, mode=1))


def test_accumulate_bit_exact(gc, orc, dev):
    rng = np.random.default_rng(14)
    n = 5000
    svl = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(2 * n).astype(np.float32)
    c = rng.standard_normal(2 * 27).astype(np.float32)
    ds, dg, dc = to_dev(svl, dev), to_dev(g, dev), to_dev(c, dev)
    gc.check(gc.lib.gpucad_svl_accumulate(gc.ptr(ds), gc.ptr(dg), n, 13, gc.ptr(dc)))
    assert np.array_equal(ds.cpu().numpy(), orc.svl_accumulate(svl, g, 13, c))


def test_axpby_fillfrac_dot(gc, dev):
    rng = np.random.default_rng(17)
    n = 70001
    v = rng.standard_normal(n).astype(np.float32)
    w = rng.standard_normal(n).astype(np.float32)
    dv, dw = to_dev(v, dev), to_dev(w, dev)
    gc.check(gc.lib.gpucad_svl_axpby(gc.ptr(dv), 1.0, gc.ptr(dw), 0.37, n))
    # fused a2 w + a1 v, computed exactly in double then rounded once
    ref = (np.float64(np.float32(0.37)) * w + v).astype(np.float32)
    assert np.array_equal(dv.cpu().numpy(), ref)
    gc.check(gc.lib.gpucad_svl_fillfrac(gc.ptr(dv), gc.ptr(dw), n, 1, 1))
    assert np.array_equal(dv.cpu().numpy(), (w * ref).astype(np.float32))
    out = torch.zeros(1, dtype=torch.float32, device=dev)
    gc.check(gc.lib.gpucad_svl_dot(gc.ptr(out), gc.ptr(dv), gc.ptr(dw), n))
    exact = np.dot(dv.cpu().numpy().astype(np.float64), w.astype(np.float64))
    assert abs(out.item() - exact) <= 1e-6 * abs(exact) + 1e-6


def test_fft_fill_and_scale_bit_exact(gc, orc, dev):
    """The x = 0 plane is written twice (direct and mirrored, as in the
    reference kernel); a real field's spectrum is Hermitian there, so both
    writes agree.  The input is made exactly Hermitian on that plane."""
    n = 9
    half = np.random.default_rng(15).standard_normal((n, n, n // 2 + 1, 2)).astype(np.float32)
    for z in range(n):
        for y in range(n):
            zm, ym = (n - z) % n, (n - y) % n
            if (zm, ym) < (z, y):
                half[z, y, 0, 0] = half[zm, ym, 0, 0]
                half[z, y, 0, 1] = -half[zm, ym, 0, 1]
    half[0, 0, 0, 1] = 0.0
    half = half.ravel()
    dh = to_dev(half, dev)
    full = torch.zeros(2 * n ** 3, dtype=torch.float32, device=dev)
    gc.check(gc.lib.gpucad_fft_fill(gc.ptr(dh), gc.ptr(full), n, n, n))
    assert np.array_equal(full.cpu().numpy(), orc.fft_fill(half, n, n, n))
    gc.check(gc.lib.gpucad_fft_scale(gc.ptr(dh), 729.0, half.size // 2))
    assert np.array_equal(dh.cpu().numpy(), (half / np.float32(729.0)).astype(np.float32))


@pytest.mark.parametrize("cell_type", [0, 3])
def test_unit_spectrum_against_oracle(gc, orc, dev, cell_type):
    n = 61
    cell = torch.empty(n ** 3, dtype=torch.float32, device=dev)
    gc.check(gc.lib.gpucad_create_lattice(gc.ptr(cell), n, n, n, n ** 3, cell_type))
    coeffs = gc.unit_spectrum(cell, n, 2).cpu().numpy()
    ref = orc.unit_spectrum(cell.cpu().numpy(), n, n, n, 2)
    assert np.max(np.abs(coeffs - ref)) < 1e-6 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("lt,ut,n", [("n", 0, 10), ("b", 0, 14), ("r", 2, 12), ("s", 1, 10)])
def test_svl_field_against_oracle(gc, orc, dev, lt, ut, n):
    p = gc.svl_params(n, lattice_type=lt, uniform_type=ut, const_period=7.0, x_period=6.0,
                      y_period=7.0, z_period=8.0, range_=2, cg_end_res=1e-3, cg_max_iter=3000)
    period = None
    if ut == 2:
        period = orc.svl_rescale(
            orc.svl_period_angle(n, n, n, 1.0, 1.0, 1.0, (n / 2, n / 2, n / 2), "z"),
            n / 10, n / 4)
    c = np.random.default_rng(16).standard_normal((125, 2)).astype(np.float32)
    dcoef = to_dev(c, dev)
    dper = None if period is None else to_dev(period, dev)
    svl, res = gc.svl_field(p, dcoef, dper)
    ref, it, r = orc.svl_field(p, c, period)
    out = svl.cpu().numpy()
    for a in range(3):
        assert res.cg_res[a] <= 1e-3 / 6 and res.cg_iters[a] < 3000
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(out - ref)) < 2e-2 * scale, (lt, ut)
    assert np.mean(np.abs(out - ref)) < 2e-3 * scale


def test_svl_argument_errors(gc, dev):
    p = gc.svl_params(3)
    c = torch.zeros(250, dtype=torch.float32, device=dev)
    out = torch.empty(6 ** 3, dtype=torch.float32, device=dev)
    r = gc.SvlResult()
    assert gc.lib.gpucad_svl_field(C.byref(p), None, gc.ptr(c), gc.ptr(out), C.byref(r),
                                   None) == gc.E_ARG
    p = gc.svl_params(8, lattice_type="q")
    assert gc.lib.gpucad_svl_field(C.byref(p), None, gc.ptr(c), gc.ptr(out), C.byref(r),
                                   None) == gc.E_ARG
    p = gc.svl_params(8, uniform_type=2)
    assert gc.lib.gpucad_svl_field(C.byref(p), None, gc.ptr(c), gc.ptr(out), C.byref(r),
                                   None) == gc.E_ARG
    p = gc.svl_params(8, range_=5)
    assert gc.lib.gpucad_svl_field(C.byref(p), None, gc.ptr(c), gc.ptr(out), C.byref(r),
                                   None) == gc.E_ARG
    assert gc.lib.gpucad_svl_matvec(gc.ptr(c), gc.ptr(out), 3, 4, 4) == gc.E_ARG


def test_compat_spatial_lattice_loop(gc, dev, tmp_path):
    """tests/cpp/svl_harness runs the reference caller's synthesis loop
    (62 x finding_phi, GPUCG_lattice(500, 1, 0.01), texture, grating, svl,
    normalise_four; then computeIsosurface_lattice and file_write_obj)
    through the compat classes.  Each solve meets the reference's stopping
    rule, the mesh is written, and the field agrees with the one-call
    synthesis (both solve to a residual of 0.01 per harmonic, so they differ
    by the solves' error, not by construction)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "cpp", "svl_harness")
    n, period = 16, 6.0
    r = subprocess.run([exe, str(tmp_path), str(n), str(period), "b"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in r.stdout.strip().splitlines()]
    per_h, last = lines[:-1], lines[-1]
    assert len(per_h) == 62
    for h, it, res in per_h:
        assert int(it) == 500 or float(res) <= 0.01, (h, it, res)
    active, total = int(last[0]), int(last[1])
    assert active > 0 and total > 0 and total % 3 == 0
    assert os.path.getsize(os.path.join(str(tmp_path), "spatial_lattice.obj")) > 0
    coeffs = np.fromfile(os.path.join(str(tmp_path), "coeffs.bin"), np.float32)
    loop = np.fromfile(os.path.join(str(tmp_path), "svl.bin"), np.float32)
    p = gc.svl_params(n, lattice_type="b", uniform_type=0, const_period=period,
                      cg_end_res=0.01, cg_max_iter=500)
    one, res = gc.svl_field(p, to_dev(coeffs, dev))
    one = one.cpu().numpy()
    rms = np.sqrt(np.mean(loop.astype(np.float64) ** 2))
    diff = np.sqrt(np.mean((one.astype(np.float64) - loop) ** 2))
    assert diff < 0.05 * rms, (diff, rms)
    assert np.corrcoef(one, loop)[0, 1] > 0.995
