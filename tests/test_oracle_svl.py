"""The oracle's restatement of the spatially variant lattice synthesis
(src/lattice_files/Gratings.cu, src/main.cu:3578-3711, 3904-4037), pinned by
facts independent of it:

  * the phase operator (GPUMatvec_lattice) equals Dt Dt^T per axis, with Dt
    built from finding_phi's own stencil weights, summed over the axes as
    Kronecker products (built here with numpy);
  * finding_phi with a constant period is Dt K for the constant grating
    vector K = (2 pi / P)(i, j, k);
  * that system's solution is the linear phase (2 pi / P)(i x + j y + k z)
    minus its mean (the stencil differentiates linear fields exactly;
    CG from 0 stays orthogonal to the constants);
  * the unit-cell coefficients against numpy.fft.fftn;
  * the 2x upsampling against its closed form (even points copy, odd points
    average, the last sample repeats);
  * period/angle data against FP64 numpy; the normalisations against their
    definitions with the reference's 0-seeded min/max;
  * the one-call synthesis against the reference's per-harmonic loop
    (finding_phi, CG, grating, svl per harmonic) and, for a constant-period
    lattice, against the closed-form sum of plane waves.
A planted mistake (a dropped boundary weight) must fail the operator pin.
"""
import numpy as np
import pytest


def weights(c, n):
    if c == 0:
        return -1.0, -0.5, c, c + 1
    if c == 1:
        return 1.0, -0.5, c - 1, c + 1
    if c == n - 2:
        return 0.5, -1.0, c - 1, c + 1
    if c == n - 1:
        return 0.5, 1.0, c - 1, c
    return 0.5, -0.5, c - 1, c + 1


def stencil(n):
    """Dt of one axis from finding_phi's weights: b[x] = a1 K[x1] + a2 K[x2]."""
    dt = np.zeros((n, n))
    for x in range(n):
        w1, w2, x1, x2 = weights(x, n)
        dt[x, x1] += w1
        dt[x, x2] += w2
    return dt


def operator(nx, ny, nz):
    """sum over axes of Dt Dt^T as a dense matrix on the x-fastest layout."""
    ax = [stencil(n) @ stencil(n).T for n in (nx, ny, nz)]
    ix, iy, iz = np.eye(nx), np.eye(ny), np.eye(nz)
    return (np.kron(iz, np.kron(iy, ax[0])) + np.kron(iz, np.kron(ax[1], ix)) +
            np.kron(ax[2], np.kron(iy, ix)))


@pytest.mark.parametrize("dims", [(4, 4, 4), (5, 7, 6), (9, 4, 8)])
def test_operator_is_dt_dtT_of_the_phi_stencil(orc, dims):
    nx, ny, nz = dims
    rng = np.random.default_rng(1)
    d = rng.standard_normal(nx * ny * nz).astype(np.float32)
    q = orc.svl_matvec(d, nx, ny, nz)
    ref = operator(nx, ny, nz) @ d.astype(np.float64)
    assert np.max(np.abs(q - ref)) < 1e-5 * (1 + np.max(np.abs(ref)))
    # symmetric, constants in the null space (exactly)
    assert np.all(orc.svl_matvec(np.ones(nx * ny * nz, np.float32), nx, ny, nz) == 0.0)
    e = rng.standard_normal(nx * ny * nz).astype(np.float32)
    qe = orc.svl_matvec(e, nx, ny, nz)
    assert abs(np.dot(e, q.astype(np.float64)) - np.dot(d, qe.astype(np.float64))) < 1e-3


def test_planted_mistake_in_operator_fails_the_pin(orc):
    nx = ny = nz = 6
    a = operator(nx, ny, nz)
    planted = a.copy()
    planted[0, 1] = -0.5  # a dropped part of the x == 0 boundary weight
    d = np.random.default_rng(2).standard_normal(nx * ny * nz).astype(np.float32)
    q = orc.svl_matvec(d, nx, ny, nz)
    assert np.max(np.abs(q - planted @ d)) > 1e-3
    assert np.max(np.abs(q - a @ d)) < 1e-4


@pytest.mark.parametrize("ijk", [(1, 0, 0), (0, 1, 0), (0, 0, 1), (2, -1, 1)])
def test_finding_phi_constant_period_is_dt_k(orc, ijk):
    nx, ny, nz, P = 8, 6, 5, 12.0
    b = orc.svl_finding_phi(nx, ny, nz, *ijk, lattice_type="n", uniform_type=0, const_period=P)
    k = 2 * np.pi / P * np.array(ijk, float)
    fx = stencil(nx) @ np.full(nx, k[0])
    fy = stencil(ny) @ np.full(ny, k[1])
    fz = stencil(nz) @ np.full(nz, k[2])
    ref = (fx[None, None, :] + fy[None, :, None] + fz[:, None, None]).ravel()
    assert np.max(np.abs(b - ref)) < 2e-6


def test_finding_phi_is_linear_in_the_harmonic(orc):
    nx = ny = nz = 7
    for lt in ("n", "b", "r", "s"):
        base = [orc.svl_finding_phi(nx, ny, nz, *u, lattice_type=lt, uniform_type=1,
                                    periods=(9.0, 11.0, 13.0), sinewave_zaxis=(lt == "s"))
                for u in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]
        b = orc.svl_finding_phi(nx, ny, nz, 2, -1, 2, lattice_type=lt, uniform_type=1,
                                periods=(9.0, 11.0, 13.0), sinewave_zaxis=(lt == "s"))
        comb = 2 * base[0].astype(np.float64) - base[1] + 2 * base[2]
        assert np.max(np.abs(b - comb)) < 1e-5 * (1 + np.max(np.abs(comb))), lt


def test_finding_phi_bent_against_fp64(orc):
    """'b' and 'r' types: the angles atan2 of the neighbour positions (FP64
    numpy restatement of Gratings.cu:254-305)."""
    nx, ny, nz, P = 9, 8, 5, 10.0
    i, j, k = 1, 2, 1
    for lt in ("b", "r"):
        b = orc.svl_finding_phi(nx, ny, nz, i, j, k, lattice_type=lt, uniform_type=0,
                                const_period=P)
        mx = (nx + 1) / 2.0 if lt == "r" else 0.0
        my = (ny + 1) / 2.0 if lt == "r" else 0.0
        ref = np.empty((nz, ny, nx))
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    a1, a2, x1, x2 = weights(x, nx)
                    b1, b2, y1, y2 = weights(y, ny)
                    c1, c2, _, _ = weights(z, nz)
                    t1 = np.arctan2(y + 1 - my, x1 + 1 - mx)
                    t2 = np.arctan2(y + 1 - my, x2 + 1 - mx)
                    t3 = np.arctan2(y1 + 1 - my, x + 1 - mx)
                    t4 = np.arctan2(y2 + 1 - my, x + 1 - mx)
                    w = 2 * np.pi / P
                    kx1 = w * (i * np.cos(t1) - j * np.sin(t1))
                    kx2 = w * (i * np.cos(t2) - j * np.sin(t2))
                    ky1 = w * (i * np.sin(t3) + j * np.cos(t3))
                    ky2 = w * (i * np.sin(t4) + j * np.cos(t4))
                    kz = w * k
                    ref[z, y, x] = a1 * kx1 + a2 * kx2 + b1 * ky1 + b2 * ky2 + (c1 + c2) * kz
        assert np.max(np.abs(b - ref.ravel())) < 5e-6, lt


def test_cg_solves_constant_period_to_the_linear_phase(orc):
    nx, ny, nz, P = 10, 8, 6, 9.0
    ijk = (1, -1, 2)
    b = orc.svl_finding_phi(nx, ny, nz, *ijk, lattice_type="n", uniform_type=0, const_period=P)
    x, it, res = orc.svl_cg(b, nx, ny, nz, iters=2000, end_res=1e-4)
    assert res <= 1e-4 and 1 < it < 2000
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    lin = 2 * np.pi / P * (ijk[0] * xx + ijk[1] * yy + ijk[2] * zz)
    lin = (lin - lin.mean()).ravel()
    assert np.max(np.abs(x - lin)) < 2e-3
    # the residual it reports is the residual of what it returns
    r = b.astype(np.float64) - operator(nx, ny, nz) @ x.astype(np.float64)
    assert np.linalg.norm(r) < 2e-4


def test_cg_stops_where_the_reference_does(orc):
    nx = ny = nz = 6
    b = orc.svl_finding_phi(nx, ny, nz, 1, 0, 0, lattice_type="b", uniform_type=0,
                            const_period=8.0)
    # iteration cap: iCounter starts at 1, so `iter` = 1 runs no iteration
    x, it, res = orc.svl_cg(b, nx, ny, nz, iters=1, end_res=0.0)
    assert it == 1 and np.all(x == 0) and abs(res - np.linalg.norm(b.astype(np.float64))) < 1e-3
    x, it, res = orc.svl_cg(b, nx, ny, nz, iters=4, end_res=0.0)
    assert it == 4
    # tolerance already met: no iteration
    x, it, res = orc.svl_cg(b, nx, ny, nz, iters=500, end_res=1e9)
    assert it == 1


def test_upsample_closed_form(orc):
    nx, ny, nz = 5, 4, 3
    c = np.random.default_rng(3).standard_normal(nx * ny * nz).astype(np.float32)
    f = orc.svl_upsample(c, nx, ny, nz, (2 * nx, 2 * ny, 2 * nz), (0.5, 0.5, 0.5))
    cc = c.reshape(nz, ny, nx)

    def up(a, axis):
        n = a.shape[axis]
        a0 = np.take(a, np.arange(n), axis=axis)
        a1 = np.take(a, np.minimum(np.arange(n) + 1, n - 1), axis=axis)
        mid = (a0 + a1) * np.float32(0.5)
        out = np.stack([a0, mid], axis=axis + 1)
        shape = list(a.shape)
        shape[axis] = 2 * n
        return out.reshape(shape)

    ref = up(up(up(cc, 2), 1), 0)
    assert np.array_equal(f.reshape(2 * nz, 2 * ny, 2 * nx), ref)
    g = orc.svl_upsample(c, nx, ny, nz, (2 * nx, 2 * ny, 2 * nz), (0.5, 0.5, 0.5), mode=1)
    assert np.max(np.abs(g[0::2] - np.cos(ref.ravel()))) < 3e-7 * 8
    assert np.max(np.abs(g[1::2] - np.sin(ref.ravel()))) < 3e-7 * 8


def test_unit_spectrum_against_numpy_fft(orc):
    n = 13
    cell = np.random.default_rng(4).standard_normal(n ** 3).astype(np.float32)
    c = orc.unit_spectrum(cell, n, n, n, 2)
    F = np.fft.fftn(cell.reshape(n, n, n).astype(np.float64)) / n ** 3  # axes z, y, x
    ref = [F[k % n, j % n, i % n] for k in range(-2, 3) for j in range(-2, 3)
           for i in range(-2, 3)]
    assert np.max(np.abs(c - np.array(ref))) < 1e-12
    assert abs(c[62] - cell.astype(np.float64).mean()) < 1e-12  # the centre is the mean


def test_fft_fill_is_the_hermitian_completion(orc):
    n = 7
    x = np.random.default_rng(5).standard_normal((n, n, n))
    F = np.fft.fftn(x)  # [z, y, x]
    half = F[:, :, : n // 2 + 1].astype(np.complex64)
    h = np.empty(2 * half.size, np.float32)
    h[0::2], h[1::2] = half.real.ravel(), half.imag.ravel()
    full = orc.fft_fill(h, n, n, n)
    full = (full[0::2] + 1j * full[1::2]).reshape(n, n, n)
    assert np.max(np.abs(full - F)) < 1e-4 * np.max(np.abs(F))


def test_period_angle_and_normalise(orc):
    nx, ny, nz = 6, 5, 4
    m = (3.0, 2.5, 2.0)
    per = orc.svl_period_angle(nx, ny, nz, 1.0, 1.0, 1.0, m, "z")
    ang = orc.svl_period_angle(nx, ny, nz, 1.0, 1.0, 1.0, m, "z", angle=True)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x, y = xx - m[0] + 1, yy - m[1] + 1
    assert np.max(np.abs(per - np.hypot(x, y).ravel())) < 1e-6 * np.max(per)
    assert np.max(np.abs(ang - np.arctan2(y, x).ravel())) < 5e-7
    # normalise_three maps [min(0, f), max(0, f)] to [a1, a1 + b1]
    t = orc.svl_rescale(per, 6.0, 15.0)
    ref = 6.0 + 15.0 * (per / per.max())
    assert np.max(np.abs(t - ref)) < 1e-5
    v = np.array([-2.0, 1.0, 3.0], np.float32)
    assert np.allclose(orc.svl_rescale(v, affine=False), (v + 2) / 5)
    w = np.array([1.0, 2.0, 4.0], np.float32)  # min seeded with 0
    assert np.allclose(orc.svl_rescale(w, affine=False), w / 4)


class P:
    """gc_svl_params-like record for the oracle."""

    def __init__(self, n, lattice_type="n", uniform_type=0, const_period=8.0, range_=1,
                 harmonics=None, iters=2000, end_res=1e-3):
        self.nx = self.ny = self.nz = n
        self.dx = self.dy = self.dz = 1.0
        self.lattice_type = lattice_type
        self.uniform_type = uniform_type
        self.const_period = const_period
        self.x_period = self.y_period = self.z_period = const_period
        self.lcon, self.lcon_1, self.sinewave_zaxis = 0.16, 0.6, 0
        self.range = range_
        side = 2 * range_ + 1
        self.harmonics = side ** 3 // 2 if harmonics is None else harmonics
        self.cg_max_iter, self.cg_end_res = iters, end_res
        self.fx = self.fy = self.fz = 2 * n
        self.fine_dx = self.fine_dy = self.fine_dz = 0.5


def coefficients(range_, seed):
    side = 2 * range_ + 1
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((side ** 3, 2)).astype(np.float32)
    return c


def test_one_call_synthesis_matches_the_per_harmonic_loop(orc):
    """The unit-phase decomposition against the reference's loop: for each of
    the first `harmonics` coefficients, finding_phi + CG + grating + svl."""
    p = P(8, lattice_type="b", const_period=7.0, range_=1, end_res=1e-4)
    c = coefficients(1, 6)
    svl, it, res = orc.svl_field(p, c)
    assert np.all(res <= p.cg_end_res / 3 + 1e-9)
    side, n = 3, 8
    acc = np.zeros(p.fx * p.fy * p.fz, np.float32)
    for h in range(p.harmonics):
        i, j, k = h % side - 1, (h // side) % side - 1, h // (side * side) - 1
        b = orc.svl_finding_phi(n, n, n, i, j, k, lattice_type="b", uniform_type=0,
                                const_period=7.0)
        x, _, r = orc.svl_cg(b, n, n, n, iters=2000, end_res=1e-4)
        assert r <= 1e-4
        g = orc.svl_upsample(x, n, n, n, (p.fx, p.fy, p.fz), (0.5, 0.5, 0.5), mode=1)
        acc = orc.svl_accumulate(acc, g, h, c)
    scale = np.max(np.abs(acc))
    assert np.max(np.abs(svl - acc)) < 2e-3 * scale


def test_one_call_synthesis_constant_period_is_a_plane_wave_sum(orc):
    n, P_ = 8, 6.0
    p = P(n, lattice_type="n", const_period=P_, range_=1, end_res=1e-5)
    c = coefficients(1, 7)
    svl, it, res = orc.svl_field(p, c)
    # unit phases: the linear phase minus its mean; upsampled linearly
    f = np.arange(2 * n) * 0.5
    f = np.minimum(f, n - 1)  # the clamp of the last fine point
    w = 2 * np.pi / P_
    mean = (n - 1) / 2.0
    zz, yy, xx = np.meshgrid(f, f, f, indexing="ij")
    ref = np.zeros(zz.shape)
    for h in range(p.harmonics):
        i, j, k = h % 3 - 1, (h // 3) % 3 - 1, h // 9 - 1
        ph = w * (i * (xx - mean) + j * (yy - mean) + k * (zz - mean))
        ref += c[h, 0] * np.cos(ph) - c[h, 1] * np.sin(ph)
    assert np.max(np.abs(svl - ref.ravel())) < 1e-3 * np.max(np.abs(ref))
