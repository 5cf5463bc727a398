"""Pins of the oracle's restatement of the unit-cell transforms
(Fft_lattice::fft_func / ifft_func / ifft_func_complex on cuFFT plans,
src/lattice_files/Fft_lattice.cu:107-126, src/main.cu:3522-3526) by an
independent implementation (numpy.fft, pocketfft) and by transforms fixed by
the mathematics: a delta, a constant, a single plane wave, Parseval; planted
mistakes (conjugate sign, swapped axes, unpadded rows) fail them.  cuFFT
itself is absent, so bit parity with it is unpinned; these fix the
definition (sign, normalisation, axis order, in-place layout)."""
import numpy as np
import pytest

from oracle import oracle as orc

SHAPES = [(61, 61, 61), (5, 7, 8), (6, 4, 9), (1, 3, 2)]


def padded(x):
    n0, n1, n2 = x.shape
    h = n2 // 2 + 1
    buf = np.zeros((n0, n1, 2 * h), np.float32)
    buf[:, :, :n2] = x
    return buf.ravel()


@pytest.mark.parametrize("shape", SHAPES)
def test_r2c_matches_numpy(shape):
    x = np.random.default_rng(1).standard_normal(shape).astype(np.float32)
    X = orc.fft_r2c_padded(padded(x), *shape)
    ref = np.fft.rfftn(x.astype(np.float64))
    assert np.abs(X - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("shape", SHAPES)
def test_c2r_matches_numpy_for_any_half_spectrum(shape):
    n0, n1, n2 = shape
    rng = np.random.default_rng(2)
    X = rng.standard_normal((n0, n1, n2 // 2 + 1)) + 1j * rng.standard_normal((n0, n1, n2 // 2 + 1))
    x = orc.fft_c2r(X, *shape)
    ref = np.fft.irfftn(X, s=shape, axes=(0, 1, 2)) * (n0 * n1 * n2)
    assert np.abs(x - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("shape", SHAPES)
def test_c2c_inverse_matches_numpy(shape):
    rng = np.random.default_rng(3)
    X = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    Y = orc.fft_c2c_inverse(X, *shape)
    ref = np.fft.ifftn(X) * np.prod(shape)
    assert np.abs(Y - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())


def test_known_transforms():
    n0, n1, n2 = 5, 7, 8
    N = n0 * n1 * n2
    delta = np.zeros((n0, n1, n2), np.float32)
    delta[0, 0, 0] = 1
    assert np.allclose(orc.fft_r2c_padded(padded(delta), n0, n1, n2), 1.0, atol=1e-12)
    const = np.ones((n0, n1, n2), np.float32)
    X = orc.fft_r2c_padded(padded(const), n0, n1, n2)
    assert abs(X[0, 0, 0] - N) < 1e-9 and np.abs(X.ravel()[1:]).max() < 1e-9
    # a plane wave with frequency (1, 2, 3) forward: cos -> N/2 at +-k
    j0, j1, j2 = np.meshgrid(np.arange(n0), np.arange(n1), np.arange(n2), indexing="ij")
    wave = np.cos(2 * np.pi * (j0 * 1 / n0 + j1 * 2 / n1 + j2 * 3 / n2))
    X = orc.fft_r2c_padded(padded(wave.astype(np.float32)), n0, n1, n2)
    assert abs(X[1, 2, 3] - N / 2) < 1e-4
    mask = np.ones(X.shape, bool)
    mask[1, 2, 3] = False
    assert np.abs(X[mask]).max() < 1e-4  # the -k partner lies in the dropped half
    # inverse of a single coefficient: e^{+2 pi i k.x/n} (the unit-lattice view's gratings)
    C = np.zeros((n0, n1, n2), complex)
    C[1, 2, 3] = 2 - 1j
    Y = orc.fft_c2c_inverse(C, n0, n1, n2)
    expect = (2 - 1j) * np.exp(2j * np.pi * (j0 / n0 + 2 * j1 / n1 + 3 * j2 / n2))
    assert np.abs(Y - expect).max() < 1e-12
    # Parseval: sum |x|^2 = sum |X|^2 / N over the full spectrum
    x = np.random.default_rng(4).standard_normal((n0, n1, n2))
    F = np.fft.fftn(x)
    Xh = orc.fft_r2c_padded(padded(x.astype(np.float32)), n0, n1, n2)
    assert np.allclose(Xh, F[:, :, :n2 // 2 + 1], atol=1e-5)
    assert abs((x.astype(np.float32).astype(np.float64) ** 2).sum() -
               (np.abs(np.fft.fftn(x.astype(np.float32).astype(np.float64))) ** 2).sum() / N) < 1e-8
    # round trip: c2r(r2c(x)) = N x
    back = orc.fft_c2r(Xh, n0, n1, n2)
    assert np.abs(back - N * x.astype(np.float32).astype(np.float64)).max() < 1e-9 * N


@pytest.mark.parametrize("mistake", ["sign", "axes", "unpadded"])
def test_planted_mistakes_fail(mistake):
    shape = (5, 7, 8)
    n0, n1, n2 = shape
    x = np.random.default_rng(5).standard_normal(shape).astype(np.float32)
    ref = np.fft.rfftn(x.astype(np.float64))
    if mistake == "sign":
        h = n2 // 2 + 1
        X = orc._along(x.astype(np.float64), 2, orc._dft_matrix(n2, 1)[:h])
        X = orc._along(orc._along(X, 1, orc._dft_matrix(n1, 1)), 0, orc._dft_matrix(n0, 1))
    elif mistake == "axes":  # n0 taken as the fastest axis
        X = np.fft.rfftn(x.astype(np.float64).transpose(2, 1, 0)).transpose(2, 1, 0)
        X = X[: ref.shape[0], : ref.shape[1], : ref.shape[2]] if X.shape == ref.shape else X
    else:  # rows read without the 2 (n2/2 + 1) padding
        buf = np.zeros(n0 * n1 * 2 * (n2 // 2 + 1), np.float32)
        buf[: x.size] = x.ravel()
        X = orc.fft_r2c_padded(buf, n0, n1, n2)
    assert X.shape != ref.shape or np.abs(X - ref).max() > 1e-3
