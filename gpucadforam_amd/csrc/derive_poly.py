#!/usr/bin/env python3
"""Derive the minimax polynomials used by the field kernels (and restated by the oracle).

Method: the Remez exchange algorithm (E. Ya. Remez, 1934; see e.g. Muller,
"Elementary Functions: Algorithms and Implementation", ch. 3), run in
multiprecision with mpmath on the reduced intervals below, followed by a
rounding of each coefficient to binary32.

  sin_rev:  sin(2*pi*d) ~= d * P(d^2),  d in [-1/4, 1/4],  P of degree 4
            (odd polynomial of degree 9 in d, absolute error)
  atan:     atan(a)     ~= a * Q(a^2),  a in [0, 1],     Q of degree 10
            (odd polynomial of degree 21 in a, absolute error)

With a = r - rint(r) (exact), cos(2*pi*r) = sin(2*pi*(1/4 - |a|)) and
sin(2*pi*r) = sin(2*pi*d), d = a for |a| <= 1/4 else sign(a)/2 - a, so one
polynomial serves both; arguments are in revolutions, and every reduction step
is exact (Sterbenz) for |r| < 2^22.

Running this script prints the binary32 coefficients as C hex-float literals
together with the measured maximum error of the binary32 evaluation (Horner,
fused multiply-add) on a dense grid.  The printed constants are the ones
written into gc_math.h and oracle/oracle.cpp.
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 50


def remez_odd(f, a, b, nterms, iters=40):
    """Minimax fit f(x) ~ sum_k c_k x^(2k+1), k < nterms, for x in (0, b].

    The absolute error e(x) = f(x) - x P(x^2) of an odd fit to an odd f is odd
    and vanishes at x = 0, so the n+1 alternation points are sought in (0, b].
    """
    n = nterms
    b = mp.mpf(b)

    def err(c, x):
        u = x * x
        p = mp.mpf(0)
        for k in reversed(range(n)):
            p = p * u + c[k]
        return f(x) - x * p

    # initial reference: Chebyshev extrema mapped to (0, b]
    pts = [b * (1 + mp.cos(mp.pi * (n - i) / (n + mp.mpf(0.5)))) / 2 for i in range(n + 1)]
    pts = sorted(set(pts))
    grid = [b * mp.mpf(j) / 6000 for j in range(1, 6001)]
    c = None
    for _ in range(iters):
        A = mp.matrix(n + 1, n + 1)
        rhs = mp.matrix(n + 1, 1)
        for i, x in enumerate(pts):
            for k in range(n):
                A[i, k] = x ** (2 * k + 1)
            A[i, n] = (-1) ** i
            rhs[i] = f(x)
        sol = mp.lu_solve(A, rhs)
        c = [sol[k] for k in range(n)]
        vals = [err(c, x) for x in grid]
        ext = []
        j = 0
        while j < len(grid):
            s = mp.sign(vals[j])
            k = j
            best = j
            while k < len(grid) and mp.sign(vals[k]) == s:
                if abs(vals[k]) > abs(vals[best]):
                    best = k
                k += 1
            ext.append(grid[best])
            j = k
        if len(ext) < n + 1:
            break
        best_set, best_min = None, -1
        for s0 in range(len(ext) - n):
            cand = ext[s0:s0 + n + 1]
            m = min(abs(err(c, x)) for x in cand)
            if m > best_min:
                best_min, best_set = m, cand
        if best_set == pts:
            break
        pts = best_set
    maxerr = max(abs(err(c, b * mp.mpf(j) / 20000)) for j in range(20001))
    return c, maxerr


def f32(x):
    return np.float32(float(x))


def hexf(x):
    return float(np.float32(x)).hex()


def main():
    two_pi = 2 * mp.pi
    c_sin, e_sin = remez_odd(lambda d: mp.sin(two_pi * d), 0, mp.mpf(1) / 4, 5)
    c_atan, e_atan = remez_odd(lambda a: mp.atan(a), 0, 1, 11)
    cs = [f32(c) for c in c_sin]
    ca = [f32(c) for c in c_atan]
    print("sin_rev minimax (multiprecision) max abs error: %.3e" % float(e_sin))
    for i, c in enumerate(cs):
        print("  GC_SINREV_C%d = %sf  /* %.9e */" % (2 * i + 1, hexf(c), float(c)))
    print("atan minimax (multiprecision) max abs error: %.3e" % float(e_atan))
    for i, c in enumerate(ca):
        print("  GC_ATAN_C%d = %sf  /* %.9e */" % (2 * i + 1, hexf(c), float(c)))

    # binary32 evaluation error (Horner with fused multiply-add emulated in float64,
    # which is exact for the product and rounds once to binary32)
    def fma32(a, b, c):
        return np.float32(np.float64(a) * np.float64(b) + np.float64(c))

    d = np.linspace(-0.25, 0.25, 2_000_001, dtype=np.float32)
    d2 = (d * d).astype(np.float32)
    p = np.full_like(d, cs[-1])
    for c in reversed(cs[:-1]):
        p = fma32(p, d2, np.float32(c))
    s = (d * p).astype(np.float32)
    ref = np.sin(2 * np.pi * d.astype(np.float64))
    print("binary32 sin_rev evaluation max abs error on [-1/4,1/4]: %.3e" % np.max(np.abs(s - ref)))

    a = np.linspace(0, 1, 2_000_001, dtype=np.float32)
    a2 = (a * a).astype(np.float32)
    q = np.full_like(a, ca[-1])
    for c in reversed(ca[:-1]):
        q = fma32(q, a2, np.float32(c))
    t = (a * q).astype(np.float32)
    print("binary32 atan evaluation max abs error on [0,1]: %.3e" % np.max(np.abs(t - np.arctan(a.astype(np.float64)))))


if __name__ == "__main__":
    main()
