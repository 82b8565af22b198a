"""The GPU right-hand sides against the sympy closed forms of
tests/closed_forms.py: polar-form Q with non-constant S, D_Q, a_Q, c_Q != 1
and a spatially varying activity (friction/Qsolver.cuh:63-322 generalised).

* Fused path: the row pass's forward set (qs_solver_debug_phase 0 and 1,
  through the C-ABI) combined by the column stage's y formulas
  (tests/fused_ref.py), for every parameter set.
* Per-op path (the reference's calls, unit Q coefficients): curr_func,
  wnonl_func, r1nonl_func, r2nonl_func with the activity field.

Tolerances as for the oracle pins: 1e-11 (N_w), 1e-13 (N_r1, N_r2) relative L2.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import oracle_py as O  # noqa: E402
import qsolver as Q  # noqa: E402
import closed_forms as CF  # noqa: E402
import fused_ref as FR  # noqa: E402

N = 32
TOL = (1e-11, 1e-13, 1e-13)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("alpha_field", [False, True])
@pytest.mark.parametrize("pset", list(CF.PARAM_SETS))
def test_fused_rhs_vs_closed_form(pset, alpha_field):
    print(f"[gpu-test] fused rhs closed form {pset} alpha_field={alpha_field}", flush=True)
    p = CF.PARAM_SETS[pset]
    om = O.Mesh(N, N, 2 * np.pi)
    Xg, Yg = om.grid()
    w, r1, r2, al, *exact = CF.case(Xg, Yg, alpha_field=alpha_field, **p)
    want = [om.fwd(e) for e in exact]
    m = Q.Mesh(N, N, 2 * np.pi)
    s = Q.Solver(m, Q.params(lam=p["lam"], Re=p["Re"], Er=p["Er"], Rf=0.1, alpha=1.0, D_Q=p["DQ"], a_Q=p["aQ"],
                             c_Q=p["cQ"], dt=1e-3))
    if alpha_field:
        s.set_activity(dev(al))
    s.set_state(dev(om.fwd(w)), dev(om.fwd(r1)), dev(om.fwd(r2)))
    s.debug_phase(0)
    s.debug_phase(1)
    F = [s.debug_array(k).cpu().numpy() for k in range(7)]
    kx, ky = FR.tables(N, N)
    got = FR.nonlinear_from_forward_set(F, kx, ky, N, N)
    for g, e, tol in zip(got, want, TOL):
        assert rel(g, e) < tol
    s.close()


def test_per_op_rhs_vs_closed_form():
    """The reference's own call sequence with the reference physics, a
    non-constant S and a spatial activity field."""
    print("[gpu-test] per-op rhs closed form", flush=True)
    p = CF.PARAM_SETS["reference"]
    om = O.Mesh(N, N, 2 * np.pi)
    Xg, Yg = om.grid()
    w, r1, r2, al, *exact = CF.case(Xg, Yg, alpha_field=True, **p)
    m = Q.Mesh(N, N, 2 * np.pi)
    names = "w u v r1 r2 S p11 p12 p21 Ra aux aux1 wn r1n r2n".split()
    F = {k: Q.Field(m) for k in names}
    h = {k: v.handle for k, v in F.items()}
    for nm, val in (("w", w), ("r1", r1), ("r2", r2)):
        F[nm].spec.copy_(dev(om.fwd(val)))
    F["Ra"].phys.copy_(dev(al))
    lam, Re, Er = p["lam"], p["Re"], p["Er"]
    Q.call("qs_curr_func", h["r1"], h["r2"], h["w"], h["u"], h["v"], h["S"])
    Q.call("qs_wnonl_func", h["wn"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["Ra"], h["S"],
           h["p11"], h["p12"], h["p21"], Re, Er, lam, h["aux"], h["aux1"])
    Q.call("qs_r1nonl_func", h["r1n"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["S"], lam, h["aux"], h["aux1"])
    Q.call("qs_r2nonl_func", h["r2n"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["S"], lam, h["aux"], h["aux1"])
    torch.cuda.synchronize()
    S = F["S"].phys.cpu().numpy()
    assert np.abs(S - 2 * np.sqrt(r1**2 + r2**2)).max() < 1e-14
    for k, e, tol in zip(("wn", "r1n", "r2n"), exact, TOL):
        assert rel(F[k].spec.cpu().numpy(), om.fwd(e)) < tol
