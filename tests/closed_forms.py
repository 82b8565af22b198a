"""Closed-form nonlinear terms of the Q-tensor model for band-limited fields
(test infrastructure; derived symbolically with sympy, independently of the
oracle's pseudo-spectral evaluation).

The fields are trigonometric polynomials whose every product stays below the
mesh's Nyquist mode, so the pseudo-spectral right-hand side equals the closed
form to rounding:

* Q in polar form, r1 = eps A cos(theta), r2 = eps A sin(theta), with
  A = 1 + 3/10 cos(x) + 1/5 sin(y) > 0 and theta = x + 2 y: the scalar order
  parameter S = 2 sqrt(r1^2 + r2^2) = 2 eps A is NOT constant (every
  reference KAT, UnitTest/Nonlinear/Qsolver.cu:46-77 and nlterms1-4, has
  S = 2).
* omega: a few Fourier modes plus a mean (the (0,0) mode, which carries no
  velocity, QFldfuncs.cuh:204-224); psi = omega/k^2 mode by mode,
  u = d_y psi, v = -d_x psi.
* alpha(x, y): a uniform activity, or a spatial field 1 + 3/10 cos(x - y) +
  1/10 sin(2 x) (the config-4 control, friction/Qsolver.cuh:75-83).

Equations (friction/Qsolver.cuh:63-300 with the Q-equation coefficients D_Q,
a_Q, c_Q; D_Q = a_Q = c_Q = 1 is the reference):
  Pi11 = -lam D_Q S lap r1 + lam S (c_Q S^2 - a_Q) r1 - alpha r1
  Pi12 = -lam D_Q S lap r2 + lam S (c_Q S^2 - a_Q) r2 + 2 D_Q (r2 lap r1 - r1 lap r2) - alpha r2
  Pi21 = -lam D_Q S lap r2 + lam S (c_Q S^2 - a_Q) r2 - 2 D_Q (r2 lap r1 - r1 lap r2) - alpha r2
  N_w  = -(d_x(u w) + d_y(v w)) + (d_xx Pi12 - 2 d_xy Pi11 - d_yy Pi21) / (Er Re)
  N_r1 = -c_Q S^2 r1 - w r2 - (d_x(u r1) + d_y(v r1)) + lam S d_x u
  N_r2 = -c_Q S^2 r2 + w r1 - (d_x(u r2) + d_y(v r2)) + lam/2 S (d_y u + d_x v)
(Pi: :124-151 with single :96-105 and cross :107-122; PiDeriv :154-164;
N_w: wnonl_func :166-201; N_r1, N_r2: r1nonl/r2nonl_func :217-300.)

`mutation` evaluates a deliberately wrong variant (a dropped D_Q in the cross
term, or the activity taken at one point instead of the field), so tests can
show how far such a mistake moves the result.
"""
import functools

import numpy as np
import sympy as sp

X, Y = sp.symbols("x y", real=True)
R = sp.Rational

# omega = mean + modes (coefficient, kx, ky, cos|sin)
OMEGA_MODES = [(R(1, 2), 1, 1, "sin"), (R(1, 5), 2, 0, "cos"), (R(-3, 20), 1, -2, "cos"), (R(1, 10), 0, 3, "sin")]
OMEGA_MEAN = R(1, 20)


def _mode(kx, ky, kind):
    arg = kx * X + ky * Y
    return sp.cos(arg) if kind == "cos" else sp.sin(arg)


@functools.lru_cache(maxsize=None)
def _terms(eps, lam, Re, Er, DQ, aQ, cQ, alpha_field, mutation):
    lap = lambda f: sp.diff(f, X, 2) + sp.diff(f, Y, 2)
    A = 1 + R(3, 10) * sp.cos(X) + R(1, 5) * sp.sin(Y)
    th = X + 2 * Y
    r1 = eps * A * sp.cos(th)
    r2 = eps * A * sp.sin(th)
    S = 2 * eps * A                       # = 2 sqrt(r1^2 + r2^2), A > 0
    w = OMEGA_MEAN + sum(c * _mode(kx, ky, k) for c, kx, ky, k in OMEGA_MODES)
    psi = sum(c / (kx * kx + ky * ky) * _mode(kx, ky, k) for c, kx, ky, k in OMEGA_MODES)
    u, v = sp.diff(psi, Y), -sp.diff(psi, X)
    if alpha_field:
        al = 1 + R(3, 10) * sp.cos(X - Y) + R(1, 10) * sp.sin(2 * X)
        if mutation == "alpha_at_origin":
            al = al.subs({X: 0, Y: 0})
    else:
        al = sp.Integer(1)
    L1, L2 = lap(r1), lap(r2)
    g = lam * S * (cQ * S**2 - aQ)
    crossc = 1 if mutation == "cross_without_DQ" else DQ
    cr = 2 * crossc * (L1 * r2 - r1 * L2)
    p11 = -lam * DQ * S * L1 + g * r1 - al * r1
    p12 = -lam * DQ * S * L2 + g * r2 + cr - al * r2
    p21 = -lam * DQ * S * L2 + g * r2 - cr - al * r2
    d = sp.diff
    Nw = -(d(u * w, X) + d(v * w, Y)) + (d(p12, X, 2) - 2 * d(p11, X, Y) - d(p21, Y, 2)) / (Er * Re)
    Nr1 = -cQ * S**2 * r1 - w * r2 - (d(u * r1, X) + d(v * r1, Y)) + lam * S * d(u, X)
    Nr2 = -cQ * S**2 * r2 + w * r1 - (d(u * r2, X) + d(v * r2, Y)) + lam / 2 * S * (d(u, Y) + d(v, X))
    exprs = [w, r1, r2, al, Nw, Nr1, Nr2]
    return [sp.lambdify((X, Y), e, "numpy") for e in exprs]


def case(Xg, Yg, eps=0.4, lam=1.0, Re=1.0, Er=1.0, DQ=0.05, aQ=1.0, cQ=1.0, alpha_field=True, mutation=None):
    """Physical fields on the grid: (w, r1, r2, alpha, N_w, N_r1, N_r2),
    evaluated in extended precision (x87 long double) and rounded to FP64.

    Parameters enter sympy as exact rationals where they are simple decimals,
    so the closed form carries no extra rounding."""
    q = lambda a: sp.nsimplify(a, rational=True)
    fs = _terms(q(eps), q(lam), q(Re), q(Er), q(DQ), q(aQ), q(cQ), bool(alpha_field), mutation)
    out = []
    for f in fs:
        val = f(np.asarray(Xg, dtype=np.longdouble), np.asarray(Yg, dtype=np.longdouble))
        out.append(np.broadcast_to(np.asarray(val, dtype=np.float64), Xg.shape).copy())
    return out


# Parameter sets of the pins: the benchmark's (D_Q = 0.05), an unusual one
# exercising every coefficient, and the reference's own (D_Q = a_Q = c_Q = 1).
PARAM_SETS = {
    "benchmark": dict(lam=1.0, Re=1.0, Er=1.0, DQ=0.05, aQ=1.0, cQ=1.0),
    "coefficients": dict(lam=0.7, Re=0.5, Er=2.0, DQ=0.3, aQ=1.5, cQ=0.8),
    "reference": dict(lam=0.1, Re=0.1, Er=0.1, DQ=1.0, aQ=1.0, cQ=1.0),
}
