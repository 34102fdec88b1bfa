"""Exactly representable test surfaces for closed-form checks.

Biquadratic Bezier patches over [-1,1]^2 (u, v in [0,1], x = 2u-1,
y = 2v-1) whose control values are dyadic, so the float32 input format holds
them exactly and the surfaces are the quadrics and planes they stand for;
rational sphere patches (exact up to the float32 rounding of their weights
and coordinates).
"""
import math

import numpy as np

from jsi_cad_next_versions_amd import surfaces as S

KNOTS = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0])
XS = np.array([-1.0, 0.0, 1.0])   # Greville abscissae 0, 1/2, 1 mapped to x
BERN = np.array([1.0, -1.0, 1.0])  # Bezier ordinates of (2u-1)^2


def quadric(sx, sy, c=0.0, rational=True):
    """z = sx x^2 + sy y^2 + c."""
    X, Y = np.meshgrid(XS, XS)  # [v, u]
    Z = sx * BERN[None, :] + sy * BERN[:, None] + c
    w = np.ones((3, 3)) if rational else None
    return S.make_surface(2, 2, KNOTS, KNOTS, np.stack([X, Y, Z], -1), w, f"quadric({sx},{sy},{c})")


def plane(a, b, c):
    """z = a x + b y + c."""
    X, Y = np.meshgrid(XS, XS)
    return S.make_surface(2, 2, KNOTS, KNOTS, np.stack([X, Y, a * X + b * Y + c], -1), None, f"plane({a},{b},{c})")


def sphere_patch(c, r, lon0):
    """Sphere of centre c and radius r over longitude [lon0, lon0 + 90 deg]
    (lon0 a multiple of 90 degrees) and latitude [-30, 30] deg: the tensor
    product of two rational quadratic circular arcs, a parallel (weights 1,
    sqrt(1/2), 1) and a meridian (weights 1, cos 30, 1), which represents the
    surface of revolution exactly.  The irrational weights and coordinates are
    stored as float32, so points lie on the sphere to ~1e-7."""
    ca, sa = math.cos(math.radians(lon0)), math.sin(math.radians(lon0))
    par = [(ca, sa, 1.0), (ca - sa, sa + ca, math.sqrt(0.5)), (-sa, ca, 1.0)]
    c30, s30 = math.sqrt(3.0) / 2.0, 0.5
    mer = [(c30, -s30, 1.0), (1.0 / c30, 0.0, c30), (c30, s30, 1.0)]
    xyz = np.array([[[c[0] + r * rho * px, c[1] + r * rho * py, c[2] + r * z] for (px, py, _) in par]
                    for (rho, z, _) in mer])
    w = np.array([[wm * wp for (_, _, wp) in par] for (_, _, wm) in mer])
    return S.make_surface(2, 2, KNOTS, KNOTS, xyz, w, f"sphere({c},{r},{lon0})")
