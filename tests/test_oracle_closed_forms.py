"""Pins of the oracle's right-hand side where the reference's own KATs leave it
open (UnitTest/Nonlinear/*: S = 2 everywhere, unit Q coefficients, uniform
activity): a non-constant S (polar-form Q), D_Q != 1 inside single/cross
(friction/Qsolver.cuh:96-122, generalised), a_Q, c_Q != 1, and a spatially
varying activity Ra(x) (:75-83), each against the sympy closed form of
tests/closed_forms.py.  Each parameter set also shows that the two plausible
mistakes -- dropping D_Q from the cross term, reading the activity at one
point (Ra[0]) instead of the field -- move the result by many orders of
magnitude more than the tolerance.

Tolerances (relative L2 of the spectra): 1e-13 for N_r1, N_r2; 1e-11 for N_w,
whose k^4-weighted stress term (factor 1/(Er Re) up to 100) carries the
rounding of the pseudo-spectral derivatives.
"""
import numpy as np
import pytest

import oracle_py as O
import closed_forms as CF

N = 32
TOL = (1e-11, 1e-13, 1e-13)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_rhs(om, p, fields, alpha_field):
    w, r1, r2, al = fields[:4]
    s = O.Solver(om, O.params_array(lam=p["lam"], Re=p["Re"], Er=p["Er"], Rf=0.1, alpha=1.0, DQ=p["DQ"],
                                    aQ=p["aQ"], cQ=p["cQ"], dt=1e-3))
    if alpha_field:
        s.set_activity(al)
    return s.rhs(om.fwd(w), om.fwd(r1), om.fwd(r2))


@pytest.mark.parametrize("alpha_field", [False, True])
@pytest.mark.parametrize("pset", list(CF.PARAM_SETS))
def test_rhs_matches_closed_form(pset, alpha_field):
    p = CF.PARAM_SETS[pset]
    om = O.Mesh(N, N, 2 * np.pi)
    Xg, Yg = om.grid()
    f = CF.case(Xg, Yg, alpha_field=alpha_field, **p)
    S = 2 * np.sqrt(f[1] ** 2 + f[2] ** 2)
    assert S.max() - S.min() > 0.5          # S genuinely varies (0.4 .. 1.2)
    got = oracle_rhs(om, p, f, alpha_field)
    want = [om.fwd(e) for e in f[4:]]
    for g, e, tol in zip(got, want, TOL):
        assert rel(g, e) < tol
    # the plausible mistakes are far outside the tolerance
    if p["DQ"] != 1.0:
        bad = CF.case(Xg, Yg, alpha_field=alpha_field, mutation="cross_without_DQ", **p)
        assert rel(om.fwd(bad[4]), want[0]) > 1e6 * TOL[0]
    if alpha_field:
        bad = CF.case(Xg, Yg, alpha_field=True, mutation="alpha_at_origin", **p)
        assert rel(om.fwd(bad[4]), want[0]) > 1e6 * TOL[0]


def test_closed_form_reduces_to_reference_kat_structure():
    """Sanity of the closed form itself: with eps -> 0 the Q terms vanish and
    N_w reduces to the advection of omega alone, -(u w_x + v w_y) (u, v
    divergence-free), checked on the grid."""
    om = O.Mesh(N, N, 2 * np.pi)
    Xg, Yg = om.grid()
    w, r1, r2, al, Nw, Nr1, Nr2 = CF.case(Xg, Yg, eps=0.0, alpha_field=False, **CF.PARAM_SETS["benchmark"])
    assert np.abs(r1).max() == 0 and np.abs(Nr1).max() == 0 and np.abs(Nr2).max() == 0
    W = om.fwd(w)
    u, v = (om.bwd(a) for a in om.vel(W))
    kx, ky = 1j * om.kx, 1j * om.ky
    wx, wy = om.bwd(kx * W), om.bwd(ky * W)
    assert np.abs(Nw + u * wx + v * wy).max() < 1e-13 * np.abs(Nw).max()
