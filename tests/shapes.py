"""Exactly representable test surfaces for closed-form checks.

Biquadratic Bezier patches over [-1,1]^2 (u, v in [0,1], x = 2u-1,
y = 2v-1) whose control values are dyadic, so the float32 input format holds
them exactly and the surfaces are the quadrics and planes they stand for.
"""
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
