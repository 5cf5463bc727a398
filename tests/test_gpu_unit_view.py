"""The reference caller's unit-lattice steps (src/main.cu:3500-3711) and its
approximate view (3720-3895, without the frame drawing) replayed by
tests/cpp/unit_view_harness through the drop-in classes, against the
oracle's chain on the same unit cell: the forward transform, fft_scalar,
fft_fill and the (2r+1)^3 coefficient pick; the inverse transform of the
unscaled spectrum, normalise_buffer and normalise_four; the sum of the
inverse transforms of the first 62 single coefficients (each a plane wave,
pinned in tests/test_oracle_fft.py).

Tolerances: the transforms are FP32 on both sides of FP64 sums, so the
coefficients, the normalised fields and the raw grating sum are compared
within 2e-5 of their largest magnitude; indicator values are equal except at
points whose normalised value lies within 2e-5 of a band level (0.2, 0.3)."""
import os
import subprocess

import numpy as np
import pytest
import torch

from tests.helpers import to_np

pytestmark = pytest.mark.gpu

N = 61
R = 2
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def unshift(i):
    mid = N // 2
    u = mid + i
    return 0 if u == mid else (u - mid if u > mid else u + mid + 1)


def pick_order():
    """(i, j, k) in the caller's loop order and the unshifted index."""
    out = []
    for k in range(-R, R + 1):
        for j in range(-R, R + 1):
            for i in range(-R, R + 1):
                out.append((i, j, k, unshift(i) + unshift(j) * N + unshift(k) * N * N))
    return out


def normalise_four_band(orc, a):
    b = orc.svl_rescale(a.astype(np.float32), affine=False)
    ind, nrm = orc.normalise_four(b, N, N, N, 0.2, 0.3)
    return ind, nrm


def check_band(ind, nrm, r_ind, r_nrm, tol=2e-5):
    assert np.abs(nrm - r_nrm).max() <= tol
    near = (np.abs(r_nrm - 0.2) <= tol) | (np.abs(r_nrm - 0.3) <= tol)
    assert np.array_equal(ind[~near], r_ind[~near])


@pytest.mark.parametrize("ltype", [0, 1, 5])
def test_unit_lattice_and_approximate_view(gc, orc, dev, tmp_path, ltype):
    exe = os.path.join(ROOT, "tests", "cpp", "unit_view_harness")
    print("unit view harness: type %d" % ltype, flush=True)
    r = subprocess.run([exe, str(tmp_path), str(ltype)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["coefficients", "125", "gratings", "62"]
    rd = lambda name, dt: np.fromfile(str(tmp_path / name), dt)

    # the same unit cell the harness transformed (same kernel, same bits)
    cellt = torch.empty(N ** 3, dtype=torch.float32, device=dev)
    gc.check(gc.lib.gpucad_create_lattice(gc.ptr(cellt), N, N, N, N ** 3, ltype))
    cell = to_np(cellt).reshape(N, N, N)
    h = N // 2 + 1
    buf = np.zeros((N, N, 2 * h), np.float32)
    buf[:, :, :N] = cell
    X = orc.fft_r2c_padded(buf.ravel(), N, N, N)

    # coefficients: fft_scalar (float32 division) of the half spectrum, fft_fill, the pick
    half = (X.astype(np.complex64) / np.float32(N ** 3)).astype(np.complex64)
    full = orc.fft_fill(half.view(np.float32).ravel(), N, N, N).view(np.complex64)
    order = pick_order()
    ref_coef = np.array([full[0] if (i, j, k) == (0, 0, 0) else full[idx]
                         for i, j, k, idx in order])
    coef = rd("lattice_data.bin", np.complex64)
    assert np.abs(coef - ref_coef).max() <= 2e-5 * np.abs(ref_coef).max()

    # unit_lattice's field: inverse of the unscaled spectrum, normalised
    back = orc.fft_c2r(X, N, N, N)
    r_ind, r_nrm = normalise_four_band(orc, back.ravel())
    check_band(rd("unit_ind.bin", np.float32), rd("unit_norm.bin", np.float32), r_ind, r_nrm)

    # the approximate view: sum of the first 62 coefficients' plane waves
    j0, j1, j2 = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    acc = np.zeros((N, N, N))
    for i, j, k, idx in order[:62]:
        c = complex(full[idx])
        # phase of e^{+2 pi i (t j0 + s j1 + u j2) / N}, jk products reduced mod N
        ang = 2 * np.pi * (((unshift(k) * j0) % N) / N + ((unshift(j) * j1) % N) / N +
                           ((unshift(i) * j2) % N) / N)
        acc += c.real * np.cos(ang) - c.imag * np.sin(ang)
    s = rd("approx_sum.bin", np.float32).reshape(N, N, N)
    assert np.abs(s - acc).max() <= 2e-5 * max(1.0, np.abs(acc).max())
    a_ind, a_nrm = normalise_four_band(orc, acc.ravel())
    check_band(rd("approx_ind.bin", np.float32), rd("approx_norm.bin", np.float32), a_ind, a_nrm)
