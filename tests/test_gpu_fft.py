"""GPU parity of the unit cell's transforms (Fft_lattice::fft_func /
ifft_func / ifft_func_complex on the shape of gpucad_fft_plan3d,
src/lattice_files/Fft_lattice.cu:107-126) and of add_fft_constants /
indiviual_grating_sum (255-289) against the oracle's restatement of cuFFT's
semantics (pinned by numpy.fft in tests/test_oracle_fft.py).

Tolerance: the device accumulates each output in FP64 and rounds it to FP32
once, so every output is within 1e-6 of the largest output magnitude (plus
the FP32 input rounding the oracle shares); the element kernels are exact."""
import numpy as np
import pytest
import torch

from tests.helpers import to_np

pytestmark = pytest.mark.gpu

SHAPES = [(61, 61, 61), (5, 7, 8), (6, 4, 9), (1, 3, 2), (32, 16, 64)]


def padded(x):
    n0, n1, n2 = x.shape
    buf = np.zeros((n0, n1, 2 * (n2 // 2 + 1)), np.float32)
    buf[:, :, :n2] = x
    return buf.ravel()


def close(got, ref, rel=1e-6):
    return np.abs(got - ref).max() <= rel * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_then_c2r(gc, orc, dev, shape):
    n0, n1, n2 = shape
    h = n2 // 2 + 1
    x = np.random.default_rng(7).standard_normal(shape).astype(np.float32)
    buf = torch.from_numpy(padded(x)).to(dev)
    gc.check(gc.lib.gpucad_fft_plan3d(n0, n1, n2))
    gc.check(gc.lib.gpucad_fft_r2c(gc.ptr(buf)))
    got = to_np(buf).view(np.complex64).reshape(n0, n1, h).astype(np.complex128)
    ref = orc.fft_r2c_padded(padded(x), n0, n1, n2)
    assert close(got, ref)
    # the inverse of the device's own spectrum (in place), as unit_lattice does
    spec = to_np(buf).view(np.complex64).reshape(n0, n1, h)
    gc.check(gc.lib.gpucad_fft_c2r(gc.ptr(buf)))
    back = to_np(buf).reshape(n0, n1, 2 * h)[:, :, :n2].astype(np.float64)
    assert close(back, orc.fft_c2r(spec, n0, n1, n2))
    assert close(back, x.astype(np.float64) * (n0 * n1 * n2), 1e-5)


@pytest.mark.parametrize("shape", SHAPES)
def test_c2r_of_a_non_hermitian_half_spectrum(gc, orc, dev, shape):
    n0, n1, n2 = shape
    h = n2 // 2 + 1
    rng = np.random.default_rng(8)
    X = (rng.standard_normal((n0, n1, h)) + 1j * rng.standard_normal((n0, n1, h))).astype(
        np.complex64)
    buf = torch.from_numpy(X.view(np.float32).ravel().copy()).to(dev)
    gc.check(gc.lib.gpucad_fft_plan3d(n0, n1, n2))
    gc.check(gc.lib.gpucad_fft_c2r(gc.ptr(buf)))
    got = to_np(buf).reshape(n0, n1, 2 * h)[:, :, :n2].astype(np.float64)
    assert close(got, orc.fft_c2r(X, n0, n1, n2))


@pytest.mark.parametrize("shape", SHAPES)
def test_c2c_inverse(gc, orc, dev, shape):
    rng = np.random.default_rng(9)
    X = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    buf = torch.from_numpy(X.view(np.float32).ravel().copy()).to(dev)
    gc.check(gc.lib.gpucad_fft_plan3d(*shape))
    gc.check(gc.lib.gpucad_fft_c2c_inverse(gc.ptr(buf)))
    got = to_np(buf).view(np.complex64).reshape(shape).astype(np.complex128)
    assert close(got, orc.fft_c2c_inverse(X, *shape))


def test_single_coefficient_grating(gc, orc, dev):
    """The unit-lattice view's per-coefficient step (src/main.cu:3774-3845):
    zero the grating, set one coefficient, inverse transform, add the real
    part into the running sum."""
    n = 61
    g = torch.zeros(2 * n ** 3, dtype=torch.float32, device=dev)
    acc = torch.full((n ** 3,), 0.25, dtype=torch.float32, device=dev)
    gc.check(gc.lib.gpucad_fft_plan3d(n, n, n))
    v = gc.Float2(0.125, -0.5)
    idx = 2 + 60 * n + 1 * n * n  # (u, s, t) = (2, 60, 1): i = 2, j = -1, k = 1
    gc.check(gc.lib.gpucad_add_fft_constant(gc.ptr(g), v, idx))
    host = to_np(g).view(np.complex64)
    assert host[idx] == np.complex64(0.125 - 0.5j) and np.count_nonzero(host) == 1
    gc.check(gc.lib.gpucad_fft_c2c_inverse(gc.ptr(g)))
    gc.check(gc.lib.gpucad_grating_sum(gc.ptr(g), gc.ptr(acc), n, n, n))
    wave = to_np(g).view(np.complex64).reshape(n, n, n)
    ref = orc.fft_c2c_inverse(host.reshape(n, n, n), n, n, n)
    assert close(wave.astype(np.complex128), ref)
    assert np.array_equal(to_np(acc), (np.float32(0.25) + wave.real.astype(np.float32)).ravel())


def test_transform_without_a_shape_is_refused(gc, dev):
    gc.check(gc.lib.gpucad_fft_destroy())
    buf = torch.zeros(16, device=dev)
    assert gc.lib.gpucad_fft_r2c(gc.ptr(buf)) == gc.E_ARG
    assert gc.lib.gpucad_fft_plan3d(0, 4, 4) == gc.E_ARG
    assert gc.lib.gpucad_fft_plan3d(4, 4, 4096) == gc.E_ARG
