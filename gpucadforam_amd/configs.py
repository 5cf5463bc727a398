"""The workloads of the hot path, as gc_lattice_params, with seeded generation.

All randomness comes from splitmix64 so the same seed gives the same grids in
every language and on every machine.

  plain_gyroid      gyroid sheet |f| - 0.3, period 32 voxels, 128^3 (one call)
  varying_schwarz_p Schwarz-P sheet at 512^3: period 16 -> 64 voxels along x,
                    rotation about z 0 -> 90 deg along z, thickness 0.4 +- 0.2
                    from 8 seeded Fourier modes, intersected with a sphere
                    (r = 0.45)
  bent_diamond      diamond at 1024^3, period 24 voxels, bent along x with arc
                    radius 2.0, unioned with a box (half extents .45,.45,.1)
  batched_grids     256 grids, per-axis sizes U{32,40,..,256}, type
                    U{gyroid, P, diamond}, period U[8,48], thickness U[.1,.5]
  served_requests   edit requests: cubic edge round8(min(384, lognormal(median
                    96, sigma 0.5))), random lattice + sphere/box boolean
"""
import math

from . import (LatticeParams, OP_INTERSECT, OP_SUBTRACT, OP_UNION, PRIM_BOX, PRIM_NONE,
               PRIM_SPHERE, TPMS_DIAMOND, TPMS_GYROID, TPMS_SCHWARZ_P)

_M64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & _M64

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next() >> 11) * (1.0 / (1 << 53)))

    def randint(self, n):
        return self.next() % n

    def normal(self):
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def make(nx, ny, nz, tpms, period0, period1=None, rot0=0.0, rot1=0.0, bend_radius=0.0,
         sheet=True, t0=0.0, modes=(), prim=PRIM_NONE, op=0, center=(0.5, 0.5, 0.5), radius=0.0,
         half=(0.0, 0.0, 0.0)):
    p = LatticeParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.tpms = tpms
    p.period0 = period0
    p.period1 = period0 if period1 is None else period1
    p.rot0, p.rot1 = rot0, rot1
    p.bend_radius = bend_radius
    p.sheet = 1 if sheet else 0
    p.t0 = t0
    p.n_modes = len(modes)
    for m, (k, a, phi) in enumerate(modes):
        for ax in range(3):
            p.mode_k[m][ax] = k[ax]
        p.mode_a[m] = a
        p.mode_phi[m] = phi
    p.prim = prim
    p.op = op
    for ax in range(3):
        p.prim_center[ax] = center[ax]
        p.prim_half[ax] = half[ax]
    p.prim_radius = radius
    return p


def fourier_modes(seed, n=8, amplitude=0.2):
    """n modes (k in {0..3}^3 minus 0, a ~ U(-1,1), phase ~ U(0,1) revolutions),
    amplitudes scaled so that sum_m |a_m| = amplitude."""
    rng = SplitMix64(seed)
    raw = []
    while len(raw) < n:
        k = (rng.randint(4), rng.randint(4), rng.randint(4))
        if k == (0, 0, 0):
            continue
        a = rng.uniform(-1.0, 1.0)
        phi = rng.uniform(0.0, 1.0)
        raw.append((k, a, phi))
    s = sum(abs(a) for _, a, _ in raw)
    return [(k, amplitude * a / s, phi) for k, a, phi in raw]


def plain_gyroid(n=128):
    return make(n, n, n, TPMS_GYROID, 32.0, sheet=True, t0=0.3)


def varying_schwarz_p(n=512, seed=0x6770):
    return make(n, n, n, TPMS_SCHWARZ_P, 16.0, 64.0, rot0=0.0, rot1=math.pi / 2, sheet=True,
                t0=0.4, modes=fourier_modes(seed), prim=PRIM_SPHERE, op=OP_INTERSECT,
                center=(0.5, 0.5, 0.5), radius=0.45)


def bent_diamond(n=1024):
    return make(n, n, n, TPMS_DIAMOND, 24.0, bend_radius=2.0, sheet=False, prim=PRIM_BOX,
                op=OP_UNION, center=(0.5, 0.5, 0.5), half=(0.45, 0.45, 0.1))


def batched_grids(count=256, seed=0xBA7C):
    rng = SplitMix64(seed)
    sizes = list(range(32, 257, 8))
    out = []
    for _ in range(count):
        nx, ny, nz = (sizes[rng.randint(len(sizes))] for _ in range(3))
        tpms = (TPMS_GYROID, TPMS_SCHWARZ_P, TPMS_DIAMOND)[rng.randint(3)]
        period = rng.uniform(8.0, 48.0)
        t = rng.uniform(0.1, 0.5)
        out.append(make(nx, ny, nz, tpms, period, sheet=True, t0=t))
    return out


def served_requests(count, seed=0x5E7E):
    """Request parameter sets of the served workload (arrival times are drawn
    separately by the caller at a chosen rate)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        edge = math.exp(math.log(96.0) + 0.5 * rng.normal())
        edge = int(round(min(384.0, edge) / 8.0)) * 8
        edge = max(edge, 16)
        tpms = (TPMS_GYROID, TPMS_SCHWARZ_P, TPMS_DIAMOND)[rng.randint(3)]
        period = rng.uniform(8.0, 48.0)
        t = rng.uniform(0.1, 0.5)
        op = (OP_INTERSECT, OP_UNION, OP_SUBTRACT)[rng.randint(3)]
        if rng.randint(2) == 0:
            prim = dict(prim=PRIM_SPHERE, radius=rng.uniform(0.3, 0.5))
        else:
            h = rng.uniform(0.25, 0.45)
            prim = dict(prim=PRIM_BOX, half=(h, h, rng.uniform(0.1, 0.45)))
        c = tuple(rng.uniform(0.45, 0.55) for _ in range(3))
        out.append(make(edge, edge, edge, tpms, period, sheet=True, t0=t, op=op, center=c, **prim))
    return out


def poisson_arrivals(count, rate_hz, seed=0x5E7E + 1):
    rng = SplitMix64(seed)
    t, out = 0.0, []
    for _ in range(count):
        t += -math.log(max(rng.uniform(), 1e-300)) / rate_hz
        out.append(t)
    return out
