"""Per-op C-ABI entry points not exercised elsewhere, the non-finite state
handling, and object lifetimes, all through the C-ABI on the GPU.

* Elementwise ops of FldOp/QFldfuncs.cuh:85-159 (FldAddScalar, FldCopy,
  SpecSetConst, SpecMul; FldSet, SpecSet): each is one reference line, checked
  against that line evaluated by numpy (1e-15 relative; these ops are exact or
  one rounding).
* The per-op integrator stages (TimeIntegration/IFE.cuh:15-25,
  IFRK2.cuh:16-51, RK4.cuh:12-74) against the same lines, and an IF-Euler step
  composed from the per-op calls (curr_func, wnonl/r1nonl/r2nonl_func,
  ife_func1D, dealias) against the oracle's IF-Euler step.
* A non-finite state: the fused step raises the flag and the next call
  returns QS_ERR_NONFINITE (friction.cu:280 checks the snapshot's S for NaN).
* A Mesh destroyed before its solver (with adjoint and per-op state held).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import oracle_py as O  # noqa: E402
import qsolver as Q  # noqa: E402
from qsolver import synth  # noqa: E402
from kat_cases import KAT_PARAMS  # noqa: E402


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def close(a, b, tol=1e-15):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)


def test_remaining_elementwise_ops():
    banner("FldAddScalar, FldCopy, FldSet, SpecSetConst, SpecSet, SpecMul")
    N = 48
    m = Q.Mesh(N, 2 * N, 2 * np.pi)
    rng = np.random.default_rng(12)
    a = rng.standard_normal(m.phys_shape)
    A, C = dev(a), m.phys()
    m.op("FldAddScalar", C, -1.5, A, 0.25)          # QFldfuncs.cuh:85-93: pc = a*pa + b
    assert close(host(C), -1.5 * a + 0.25)
    m.op("FldCopy", C, A)                            # :107-115
    assert np.array_equal(host(C), a)
    m.op("FldSet", C, 3.25)                          # :96-104
    assert np.all(host(C) == 3.25)
    sa = rng.standard_normal(m.spec_shape) + 1j * rng.standard_normal(m.spec_shape)
    SA, SC = dev(sa), m.spec()
    m.op("SpecMul", SC, SA, -0.7)                    # :151-159: spc = C*spa
    assert close(host(SC), -0.7 * sa)
    m.op("SpecMul", SA, SA, 2.0)                     # in place, as the reference aliases
    assert np.array_equal(host(SA), 2.0 * sa)
    m.op("SpecSetConst", SC, Q.Qcomp(0.5, -2.0))     # :129-137
    assert np.all(host(SC) == 0.5 - 2.0j)
    m.op("SpecSet", SC, SA)                          # :118-126
    assert np.array_equal(host(SC), 2.0 * sa)


def test_per_op_integrator_stages():
    """Every per-op stage kernel against its reference line."""
    banner("per-op IFE / IFRK2 / RK4 stage kernels")
    N = 32
    m = Q.Mesh(N, N, 2 * np.pi)
    rng = np.random.default_rng(5)
    sh = m.spec_shape
    cplx = lambda: rng.standard_normal(sh) + 1j * rng.standard_normal(sh)
    old, curr0, new0, nonl = cplx(), cplx(), cplx(), cplx()
    IF, IFh = rng.uniform(0.5, 1.0, sh), rng.uniform(0.5, 1.0, sh)
    dt = 0.01
    an = nonl * dt
    cases = {   # name: (expected new, expected curr or None = untouched)
        "ife_func1D": ((old + an) * IF, None),
        "ifrk2_func0D": (old * IF, old),
        "ifrk2_func1D": (new0 + 0.5 * IF * an, (old + an) * IFh),
        "ifrk2_func2D": (new0 + 0.5 * an, None),
        "rk4_func0D": (old * IF, old),
        "rk4_func1D": (new0 + 1.0 / 6.0 * IF * an, (old + an / 2.0) * IFh),
        "rk4_func2D": (new0 + 1.0 / 3.0 * IFh * an, old * IFh + an / 2.0),
        "rk4_func3D": (new0 + 1.0 / 3.0 * IFh * an, old * IF + an * IFh),
        "rk4_func4D": (new0 + 1.0 / 6.0 * an, None),
    }
    for name, (want_new, want_curr) in cases.items():
        O_, C_, N_, NL = dev(old), dev(curr0), dev(new0), dev(nonl)
        T, Th = dev(IF), dev(IFh)
        if name == "ife_func1D":
            Q.call("qs_" + name, Q.ptr(O_), Q.ptr(N_), Q.ptr(NL), Q.ptr(T), Q.ptr(Th), dt, m.handle)
        elif name.endswith("0D"):
            Q.call("qs_" + name, Q.ptr(O_), Q.ptr(C_), Q.ptr(N_), Q.ptr(T), Q.ptr(Th), dt, m.handle)
        else:
            Q.call("qs_" + name, Q.ptr(O_), Q.ptr(C_), Q.ptr(N_), Q.ptr(NL), Q.ptr(T), Q.ptr(Th), dt, m.handle)
        assert close(host(N_), want_new, 1e-15), name
        assert np.array_equal(host(O_), old), name
        if want_curr is None:
            assert np.array_equal(host(C_), curr0), name
        else:
            assert close(host(C_), want_curr, 1e-15), name


@pytest.mark.parametrize("nsteps", [1, 3])
def test_per_op_ife_step_vs_oracle(nsteps):
    """An IF-Euler step of the reference's call granularity (curr_func, the
    three nonlinear functions, IFE::integrate_func1D with the reference's IF
    tables, dealias) against the oracle's IF-Euler step (reference
    parameters: D_Q = a_Q = c_Q = 1)."""
    banner(f"per-op IF-Euler x{nsteps} vs oracle")
    N, dt = 64, 1e-3
    p = dict(KAT_PARAMS)
    om = O.Mesh(N, N, 2 * np.pi)
    q11, q12 = synth.bandlimited_q(N, N, seed=8)
    rng = np.random.default_rng(8)
    w0 = om.fwd(0.2 * np.cos(om.grid()[0] + 2 * om.grid()[1]) + 0.01 * rng.standard_normal((N, N)))
    st = [w0, om.fwd(q11), om.fwd(q12)]
    s = O.Solver(om, O.params_array(dt=dt, **p))
    s.set_state(*st)
    s.step(nsteps, O.IFE)
    ref = s.get_state()

    m = Q.Mesh(N, N, 2 * np.pi)
    names = "w r1 r2 u v S p11 p12 p21 Ra aux aux1 wn r1n r2n".split()
    F = {k: Q.Field(m) for k in names}
    h = {k: v.handle for k, v in F.items()}
    F["Ra"].phys.fill_(p["alpha"])
    old = [dev(a) for a in st]
    new = [m.spec() for _ in range(3)]
    IFw, IFwh, IFq, IFqh = (torch.zeros(m.spec_shape, dtype=torch.float64, device="cuda") for _ in range(4))
    Q.call("qs_wlin_func", Q.ptr(IFw), Q.ptr(IFwh), p["Re"], p["Rf"], dt, m.handle)     # Qsolver.cuh:20-30
    Q.call("qs_r1lin_func", Q.ptr(IFq), Q.ptr(IFqh), dt, m.handle)                      # :32-42
    lam, Re, Er = p["lam"], p["Re"], p["Er"]
    for _ in range(nsteps):
        for k, nm in enumerate(("w", "r1", "r2")):
            m.op("SpecSet", F[nm].spec, old[k])
        Q.call("qs_curr_func", h["r1"], h["r2"], h["w"], h["u"], h["v"], h["S"])
        Q.call("qs_wnonl_func", h["wn"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["Ra"], h["S"],
               h["p11"], h["p12"], h["p21"], Re, Er, lam, h["aux"], h["aux1"])
        Q.call("qs_r1nonl_func", h["r1n"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["S"], lam,
               h["aux"], h["aux1"])
        Q.call("qs_r2nonl_func", h["r2n"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["S"], lam,
               h["aux"], h["aux1"])
        for k, (nl, T, Th) in enumerate((("wn", IFw, IFwh), ("r1n", IFq, IFqh), ("r2n", IFq, IFqh))):
            Q.call("qs_ife_func1D", Q.ptr(old[k]), Q.ptr(new[k]), Q.ptr(F[nl].spec), Q.ptr(T), Q.ptr(Th), dt,
                   m.handle)
            m.op("dealias", new[k])
        for k in range(3):
            old[k].copy_(new[k])
    for g, r in zip((host(t) for t in old), ref):
        assert rel_l2(g, r) < 1e-12


@pytest.mark.parametrize("method", ["step", "step_rk4", "step_ife"])
def test_nonfinite_state_is_reported(method):
    """A NaN injected into the state: the fused step's final column stage
    raises the flag, the next step call returns QS_ERR_NONFINITE without
    running, qs_solver_nonfinite reports and clears it, and a finite state
    then steps normally."""
    banner(f"non-finite state, {method}")
    N = 64
    om = O.Mesh(N, N, 2 * np.pi)
    q11, q12 = synth.uniform_q(N, N, seed=4)
    st = [np.zeros(om.spec_shape, complex), om.fwd(q11), om.fwd(q12)]
    bad = [a.copy() for a in st]
    bad[1][3, 5] = np.nan
    m = Q.Mesh(N, N, 2 * np.pi)
    s = Q.Solver(m, Q.params(D_Q=0.05, dt=1e-3))
    step = getattr(s, method)
    s.set_state(*(dev(a) for a in bad))
    step(1)                        # runs; the flag copy lands asynchronously
    torch.cuda.synchronize()
    with pytest.raises(Q.QsolverError) as ei:
        step(1)
    assert ei.value.status == Q.QS_ERR_NONFINITE
    assert s.nonfinite()           # reported ...
    assert not s.nonfinite()       # ... and cleared
    s.set_state(*(dev(a) for a in st))
    step(2)
    torch.cuda.synchronize()
    step(1)                        # no refusal after a finite run
    assert not s.nonfinite()
    got = [host(t) for t in s.get_state()]
    assert all(np.all(np.isfinite(a)) for a in got)
    s.close()


def test_mesh_destroyed_before_its_solver():
    """Teardown order through the C-ABI: the solver holds adjoint state,
    per-op fields and events when its Mesh is destroyed first; destroying the
    solver afterwards must not touch the freed Mesh."""
    banner("mesh destroyed before its solver")
    L = Q.load()
    N = 32
    mp = ctypes.POINTER(Q.MeshStruct)()
    Q.call("qs_mesh_create", N, N, 2 * np.pi, 2 * np.pi, 16, ctypes.byref(mp))
    Q.call("qs_mesh_set_stream", mp, Q.current_stream_handle())
    p = Q.params(D_Q=0.05, dt=1e-3)
    sp = ctypes.c_void_p()
    Q.call("qs_solver_create", mp, ctypes.byref(p), ctypes.byref(sp))
    om = O.Mesh(N, N, 2 * np.pi)
    q11, q12 = synth.uniform_q(N, N, seed=9)
    st = [dev(np.zeros(om.spec_shape, complex)), dev(om.fwd(q11)), dev(om.fwd(q12))]
    Q.call("qs_solver_set_state", sp, *(Q.ptr(t) for t in st))
    Q.call("qs_solver_profile", sp, 1)
    Q.call("qs_step_ifrk2", sp, 2)
    Q.call("qs_step_ifrk2_unfused", sp, 1)
    Q.call("qs_adj_forward", sp, 4, 2, None)
    g = [torch.zeros_like(t) for t in st]
    g[0].fill_(1e-3)
    ab = torch.empty((N, N), dtype=torch.float64, device="cuda")
    Q.call("qs_adj_backward", sp, *(Q.ptr(t) for t in g), Q.ptr(ab), None, None, None, None)
    torch.cuda.synchronize()
    assert L.qs_mesh_destroy(mp) == Q.QS_OK
    assert L.qs_solver_destroy(sp) == Q.QS_OK
    torch.cuda.synchronize()
    assert bool(torch.isfinite(ab).all())
