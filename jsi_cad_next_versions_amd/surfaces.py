"""Surface inputs in the JSI_CAD_NEXT NurbsFace format, and the seeded
synthetic workloads the benchmarks and tests run on.

A NurbsFace (JSI_CAD_NEXT_VERSION5/include/nurbs_data.hpp:1-10) is
``ivec[7] = {flags, degU, nCPu, nKu, degV, nCPv, nKv}`` plus a float array
``U[nKu] | V[nKv] | CP xyz (u fastest) | w (if flags & 2)``.

Random numbers come from SplitMix64 so that every machine builds bit-identical
inputs from the same seed.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field

import numpy as np

FLAG_FLAT = 1
FLAG_RATIONAL = 2

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive SplitMix64 outputs for the given seed (state += golden before each draw)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + idx * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Rng:
    """Stream of uniforms/normals built on splitmix64."""

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.pos = 0

    def _raw(self, n):
        with np.errstate(over="ignore"):
            out = splitmix64((self.seed + self.pos * int(_GOLDEN)) & 0xFFFFFFFFFFFFFFFF, n)
        self.pos += n
        return out

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self._raw(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n, sigma=1.0):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])[:n]
        return sigma * z

    def integers(self, n, lo, hi):
        """uniform integers in [lo, hi]"""
        return lo + (self._raw(n) % np.uint64(hi - lo + 1)).astype(np.int64)


@dataclass
class Surface:
    ivec: np.ndarray  # int32[7]
    data: np.ndarray  # float32[ndata]
    name: str = ""
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self.ivec = np.ascontiguousarray(self.ivec, dtype=np.int32)
        self.data = np.ascontiguousarray(self.data, dtype=np.float32)
        check_surface(self.ivec, self.data)

    @property
    def ndata(self):
        return int(self.data.size)

    # NurbsFace::get_ubegin/get_uend/get_vbegin/get_vend (nurbs_data.hpp:6-9)
    def ubegin(self):
        return float(self.data[0])

    def uend(self):
        return float(self.data[self.ivec[3] - 1])

    def vbegin(self):
        return float(self.data[self.ivec[3]])

    def vend(self):
        return float(self.data[self.ivec[3] + self.ivec[6] - 1])

    @property
    def rational(self):
        return bool(self.ivec[0] & FLAG_RATIONAL)

    def control_points(self):
        nu, nv = int(self.ivec[2]), int(self.ivec[5])
        o = int(self.ivec[3] + self.ivec[6])
        return self.data[o:o + 3 * nu * nv].reshape(nv, nu, 3)

    def weights(self):
        nu, nv = int(self.ivec[2]), int(self.ivec[5])
        if not self.rational:
            return np.ones((nv, nu), np.float32)
        o = int(self.ivec[3] + self.ivec[6] + 3 * nu * nv)
        return self.data[o:o + nu * nv].reshape(nv, nu)

    def scaled(self, factor: float) -> "Surface":
        """Control-point xyz times factor, as the 4x hard-wave case does
        (distance_analysis_example.cu:51-64)."""
        d = self.data.copy()
        nu, nv = int(self.ivec[2]), int(self.ivec[5])
        o = int(self.ivec[3] + self.ivec[6])
        d[o:o + 3 * nu * nv] *= np.float32(factor)
        return Surface(self.ivec.copy(), d, self.name + f"x{factor:g}", dict(self.meta))


def check_surface(ivec, data):
    if ivec.shape != (7,):
        raise ValueError("ivec must have 7 entries")
    flags, du, nu, nku, dv, nv, nkv = (int(x) for x in ivec)
    if du < 1 or dv < 1 or nu <= du or nv <= dv:
        raise ValueError(f"bad degrees/net size {ivec.tolist()}")
    if nku != nu + du + 1 or nkv != nv + dv + 1:
        raise ValueError(f"knot counts do not match {ivec.tolist()}")
    need = nku + nkv + 3 * nu * nv + (nu * nv if flags & FLAG_RATIONAL else 0)
    if data.size < need:
        raise ValueError(f"data has {data.size} floats, surface needs {need}")


def clamped_uniform_knots(ncp: int, deg: int, lo=0.0, hi=1.0) -> np.ndarray:
    """[lo]*(deg+1), interior lo + (hi-lo)*i/(ncp-deg), [hi]*(deg+1)."""
    nint = ncp - deg - 1
    inner = [lo + (hi - lo) * i / (nint + 1) for i in range(1, nint + 1)]
    return np.array([lo] * (deg + 1) + inner + [hi] * (deg + 1), dtype=np.float64)


def greville(knots: np.ndarray, deg: int, ncp: int) -> np.ndarray:
    return np.array([knots[i + 1:i + deg + 1].mean() for i in range(ncp)])


def make_surface(degu, degv, knots_u, knots_v, xyz, weights=None, name="", meta=None) -> Surface:
    """xyz: [nv, nu, 3] (u fastest in memory), weights: [nv, nu] or None."""
    nv, nu, _ = xyz.shape
    flags = FLAG_RATIONAL if weights is not None else 0
    ivec = np.array([flags, degu, nu, len(knots_u), degv, nv, len(knots_v)], np.int32)
    parts = [np.asarray(knots_u, np.float32), np.asarray(knots_v, np.float32),
             np.asarray(xyz, np.float32).reshape(-1)]
    if weights is not None:
        parts.append(np.asarray(weights, np.float32).reshape(-1))
    return Surface(ivec, np.concatenate(parts), name, meta or {})


# ---------------------------------------------------------------------------
# Seeded workloads (one per benchmark configuration)
# ---------------------------------------------------------------------------

def wave_pair_small(seed_a=1, seed_b=2, n=16, deg=3, noise=0.01, dz=0.3):
    """Two non-rational bicubic 16x16 surfaces over [0,1]^2:
    z = 0.1 sin(2 pi x) sin(2 pi y) + N(0, noise); B re-noised and shifted +dz."""
    ku = clamped_uniform_knots(n, deg)
    g = np.arange(n) / (n - 1)
    X, Y = np.meshgrid(g, g)  # [v, u]
    base = 0.1 * np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y)
    out = []
    for seed, off in ((seed_a, 0.0), (seed_b, dz)):
        z = base + Rng(seed).normal(n * n, noise).reshape(n, n) + off
        xyz = np.stack([X, Y, z], -1)
        out.append(make_surface(deg, deg, ku, ku, xyz, None, f"wave16_s{seed}"))
    return tuple(out)


# z offset of the hard-wave pair's surface B, chosen so that the surfaces'
# true minimum gap is 1.000e-3 (to 1e-7): FP64 grids of 1024 x 1024 points per
# surface, the closest point pairs of the two clouds (k-d tree), then the FP64
# closest-point polish inside small regions around the 12 best distinct
# candidates.  The minimum lies on the u = 1 edge (x = 1, y ~ 0.47), inside
# the last 1/nuv strip of parameter space that the reference's sampling never
# reaches, so the screening rounds at nuv 512 settle on an interior local
# minimum (~1.43e-3) instead.
HARD_WAVE_LIFT = 0.0216352


def hard_wave_pair(n=64, deg=3, seed_a=3, seed_b=4, lift=HARD_WAVE_LIFT, scale=1.0):
    """Rational bicubic 64x64 'hard wave' surfaces over [0,1]^2 with weights U[0.8,1.2].
    A: z = 0.05 sin(16 pi x) sin(16 pi y); B: the same wave with a pi/7 phase
    shift in x, lifted by `lift` (the default makes the surfaces' minimum gap
    1e-3).  scale multiplies every control-point coordinate (the 4x-scaled
    variant has a 4e-3 gap)."""
    ku = clamped_uniform_knots(n, deg)
    g = np.arange(n) / (n - 1)
    X, Y = np.meshgrid(g, g)
    za = 0.05 * np.sin(16 * np.pi * X) * np.sin(16 * np.pi * Y)
    zb = 0.05 * np.sin(16 * np.pi * X + np.pi / 7) * np.sin(16 * np.pi * Y) + lift
    wa = Rng(seed_a).uniform(n * n, 0.8, 1.2).reshape(n, n)
    wb = Rng(seed_b).uniform(n * n, 0.8, 1.2).reshape(n, n)
    a = make_surface(deg, deg, ku, ku, np.stack([X, Y, za], -1) * scale, wa, "hardwave_A", {"lift": lift})
    b = make_surface(deg, deg, ku, ku, np.stack([X, Y, zb], -1) * scale, wb, "hardwave_B", {"lift": lift})
    return a, b


def saddle_wave_pair(n=32, deg=3, seed_b=5):
    """A: rational (w = 1) bicubic saddle z = x^2 - y^2 on [-1,1]^2, exact via
    blossoms (x = 2u-1 at the Greville abscissae, z control values from the
    polar form of (2u-1)^2).  B: z = 0.2 x + 0.03 sin(10 pi y), weights U[0.9,1.1]."""
    k = clamped_uniform_knots(n, deg)
    gr = greville(k, deg, n)
    # polar form of f(u) = (2u-1)^2 = 4u^2 - 4u + 1 at (t_{i+1},...,t_{i+deg})
    b2 = []
    for i in range(n):
        t = k[i + 1:i + deg + 1]
        s2 = sum(t[a] * t[b] for a in range(deg) for b in range(a + 1, deg)) * 2.0 / (deg * (deg - 1))
        b2.append(4 * s2 - 4 * t.mean() + 1)
    b2 = np.array(b2)
    xs = 2 * gr - 1
    X, Y = np.meshgrid(xs, xs)
    ZA = b2[None, :] - b2[:, None]
    a = make_surface(deg, deg, k, k, np.stack([X, Y, ZA], -1), np.ones((n, n)), "saddle")
    ZB = 0.2 * X + 0.03 * np.sin(10 * np.pi * Y)
    wb = Rng(seed_b).uniform(n * n, 0.9, 1.1).reshape(n, n)
    b = make_surface(deg, deg, k, k, np.stack([X, Y, ZB], -1), wb, "waveplane")
    return a, b


def random_pair(seed: int, kind: str = "dist"):
    """One pair of the uneven batch workload: degrees U{2..5}, nets log-uniform
    8..96 per direction, weights U[0.7,1.3], sinusoidal height fields with 1-16
    periods plus N(0, 0.005).  'dist' pairs are separated in z, 'isect' pairs cross."""
    r = Rng(1000003 * seed + 17)
    surfs = []
    for side in range(2):
        du, dv = (int(x) for x in r.integers(2, 2, 5))
        nu = int(round(math.exp(r.uniform(1, math.log(8), math.log(96))[0])))
        nv = int(round(math.exp(r.uniform(1, math.log(8), math.log(96))[0])))
        nu, nv = max(nu, du + 1), max(nv, dv + 1)
        per = r.integers(2, 1, 16)
        ku, kv = clamped_uniform_knots(nu, du), clamped_uniform_knots(nv, dv)
        X, Y = np.meshgrid(np.arange(nu) / (nu - 1), np.arange(nv) / (nv - 1))
        amp = 0.05
        z = amp * np.sin(np.pi * per[0] * X + r.uniform(1, 0, np.pi)[0]) * np.sin(np.pi * per[1] * Y)
        z = z + r.normal(nu * nv, 0.005).reshape(nv, nu)
        if side == 1:
            if kind == "dist":
                z = z + 0.15 + r.uniform(1, 0.0, 0.2)[0]
            else:
                z = z + 0.3 * (X - 0.5)
        w = r.uniform(nu * nv, 0.7, 1.3).reshape(nv, nu)
        surfs.append(make_surface(du, dv, ku, kv, np.stack([X, Y, z], -1), w, f"rand{seed}_{side}"))
    nuv = int([64, 128, 256, 512][int(r.integers(1, 0, 3)[0])])
    return surfs[0], surfs[1], nuv


def uneven_workload(n: int, seed0: int = 0):
    """n pairs of the uneven batch workload, as (kind, A, B, nuv): random_pair's
    distribution, 70% distance ('dist') and 30% intersection ('isect') pairs."""
    out = []
    for i in range(n):
        kind = "isect" if (i * 7 + seed0) % 10 < 3 else "dist"
        a, b, nuv = random_pair(seed0 + i, kind)
        out.append((kind, a, b, nuv))
    return out


def full_region(s: Surface, nuv: int, side: int):
    """The task setup of the drivers: the whole knot domain
    (distance_analysis_example.cu:105-110, intersection_example.cu:227-232)."""
    from ._abi import Region
    return Region(s.ubegin(), s.vbegin(), s.uend(), s.vend(), int(nuv), int(side))


# ---------------------------------------------------------------------------
# Ingest of the reference's C-array surface files (examples/data/*.cu)
# ---------------------------------------------------------------------------
_NUM = r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?f?"


def parse_nurbs_c_array(text: str, name: str = "") -> Surface:
    """Parses ``int ivec[7]={...}; int ndata = N; float data[N] = {...};`` as
    written by the reference's surface files."""
    m = re.search(r"ivec\s*\[\s*7\s*\]\s*=\s*\{([^}]*)\}", text)
    if not m:
        raise ValueError("no ivec[7] initialiser")
    ivec = np.array([int(x) for x in m.group(1).replace(" ", "").split(",") if x], np.int32)
    m = re.search(r"float\s+data\s*\[\s*(\d+)\s*\]\s*=\s*\{", text)
    if not m:
        raise ValueError("no float data[] initialiser")
    n = int(m.group(1))
    end = text.index("}", m.end())
    vals = np.array([float(t.rstrip("f")) for t in re.findall(_NUM, text[m.end():end])], np.float32)
    if vals.size != n:
        raise ValueError(f"data has {vals.size} values, declared {n}")
    return Surface(ivec, vals, name)
