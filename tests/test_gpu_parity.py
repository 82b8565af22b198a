"""GPU parity of the HIP path (through the C-ABI) against the FP64 oracle.

Tolerances: transforms and elementwise ops 1e-13 relative; nonlinear terms and
the IFRK2 step 1e-12 relative L2 per field (the bar of the step's spec).
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import oracle_py as O  # noqa: E402
import qsolver as Q  # noqa: E402
from qsolver import synth  # noqa: E402
from kat_cases import KAT_PARAMS, KAT_TOL, N_MIN, nonlinear_case  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIG = dict(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, DQ=0.05, aQ=1.0, cQ=1.0)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def banner(msg):
    print(f"[gpu-test] {msg}", flush=True)


def qparams(d, dt, advect_w=1):
    return Q.params(lam=d["lam"], Re=d["Re"], Er=d["Er"], Rf=d["Rf"], alpha=d["alpha"],
                    D_Q=d.get("DQ", 1.0), a_Q=d.get("aQ", 1.0), c_Q=d.get("cQ", 1.0), dt=dt,
                    advect_w=advect_w)


def oparams(d, dt, advect_w=1):
    return O.params_array(lam=d["lam"], Re=d["Re"], Er=d["Er"], Rf=d["Rf"], alpha=d["alpha"],
                          DQ=d.get("DQ", 1.0), aQ=d.get("aQ", 1.0), cQ=d.get("cQ", 1.0), dt=dt,
                          advect_w=advect_w)


@pytest.mark.parametrize("N", [16, 64, 256, 1024])
def test_transforms_vs_oracle(N):
    banner(f"transforms N={N}")
    rng = np.random.default_rng(N)
    m = Q.Mesh(N, N, 2 * np.pi)
    om = O.Mesh(N, N, 2 * np.pi)
    f = rng.standard_normal((N, N))
    ft = m.spec()
    fd = dev(f)
    m.FwdTrans(ft, fd)
    ref = om.fwd(f)
    assert rel_l2(host(ft), ref) < 1e-14
    assert np.array_equal(host(fd), f)
    F = rng.standard_normal((N, N // 2 + 1)) + 1j * rng.standard_normal((N, N // 2 + 1))
    Fd = dev(F)
    g = m.phys()
    m.BwdTrans(g, Fd)
    F_or = F.copy()
    gref = om.bwd(F_or)
    assert rel_l2(host(g), gref) < 1e-14
    # side effect: the caller's spectrum is symmetrised in column 0
    assert np.abs(host(Fd) - F_or).max() < 1e-15 * np.abs(F).max() * 4


def test_tables_and_elementwise():
    banner("tables and elementwise")
    N = 64
    m = Q.Mesh(N, N, 2 * np.pi)
    om = O.Mesh(N, N, 2 * np.pi)
    kx, ky, k2, cut = (host(t) for t in m.tables())
    assert np.array_equal(kx, om.kx) and np.array_equal(ky, om.ky)
    assert np.array_equal(k2, om.k_squared) and np.array_equal(cut, om.cutoff)
    rng = np.random.default_rng(1)
    a, b = rng.standard_normal((N, N)), rng.standard_normal((N, N))
    A, B, C = dev(a), dev(b), m.phys()
    m.op("FldMul", C, A, B, -0.3)
    assert np.allclose(host(C), -0.3 * a * b, rtol=1e-15, atol=0)
    m.op("FldAdd", C, 0.5, A, 2.0, B)
    assert np.allclose(host(C), 0.5 * a + 2.0 * b, rtol=1e-15, atol=1e-16)
    sa = rng.standard_normal(m.spec_shape) + 1j * rng.standard_normal(m.spec_shape)
    SA, SC = dev(sa), m.spec()
    m.op("xDeriv", SC, SA)
    assert np.allclose(host(SC), 1j * om.kx * sa, rtol=1e-15, atol=1e-15)
    m.op("yDeriv", SC, SA)
    assert np.allclose(host(SC), 1j * om.ky * sa, rtol=1e-15, atol=1e-15)
    m.op("Laplacian", SC, SA)
    assert np.allclose(host(SC), -om.k_squared * sa, rtol=1e-15, atol=1e-15)
    U, V = m.spec(), m.spec()
    m.op("vel_func", SA, U, V)
    k2s = np.where(om.k_squared == 0, 1.0, om.k_squared)
    u_ref = np.where(om.k_squared == 0, 0, 1j * om.ky * sa / k2s)
    v_ref = np.where(om.k_squared == 0, 0, -1j * om.kx * sa / k2s)
    assert np.allclose(host(U), u_ref, rtol=1e-14, atol=1e-15)
    assert np.allclose(host(V), v_ref, rtol=1e-14, atol=1e-15)
    m.op("dealias", SA)
    assert np.array_equal(host(SA), sa * om.cutoff)


NL_NAMES = "w u v r1 r2 S p11 p12 p21 Ra aux aux1 wn r1n r2n".split()


@pytest.mark.parametrize("c", [0, 1, 2, 3, 4])
def test_per_op_nonlinear_kats(c):
    """Reference KATs (UnitTest/Nonlinear) through the per-op C-ABI composites."""
    banner(f"nonlinear KAT case {c}")
    N = N_MIN[c]
    m = Q.Mesh(N, N, 2 * np.pi)
    om = O.Mesh(N, N, 2 * np.pi)
    X, Y = om.grid()
    (w, r1, r2), exact = nonlinear_case(c, X, Y)
    F = {k: Q.Field(m) for k in NL_NAMES}
    for nm, val in (("w", w), ("r1", r1), ("r2", r2)):
        F[nm].spec.copy_(dev(om.fwd(val)))
    F["Ra"].phys.fill_(KAT_PARAMS["alpha"])
    h = {k: v.handle for k, v in F.items()}
    lam, Re, Er = KAT_PARAMS["lam"], KAT_PARAMS["Re"], KAT_PARAMS["Er"]
    Q.call("qs_curr_func", h["r1"], h["r2"], h["w"], h["u"], h["v"], h["S"])
    Q.call("qs_wnonl_func", h["wn"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["Ra"], h["S"],
           h["p11"], h["p12"], h["p21"], Re, Er, lam, h["aux"], h["aux1"])
    Q.call("qs_r1nonl_func", h["r1n"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["S"], lam,
           h["aux"], h["aux1"])
    Q.call("qs_r2nonl_func", h["r2n"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["S"], lam,
           h["aux"], h["aux1"])
    got = [om.bwd(host(F[k].spec).copy()) for k in ("wn", "r1n", "r2n")]
    for g, e, tol in zip(got, exact, KAT_TOL[c]):
        assert np.abs(g - e).max() <= tol * np.abs(e).max()


def oracle_run(N, d, dt, state, nsteps, Ra=None, advect_w=1):
    om = O.Mesh(N, N, 2 * np.pi)
    s = O.Solver(om, oparams(d, dt, advect_w))
    if Ra is not None:
        s.set_activity(Ra)
    s.set_state(*state)
    s.step(nsteps, O.IFRK2)
    return s.get_state()


def gpu_run(N, d, dt, state, nsteps, Ra=None, advect_w=1, fused=True):
    m = Q.Mesh(N, N, 2 * np.pi)
    s = Q.Solver(m, qparams(d, dt, advect_w))
    if Ra is not None:
        s.set_activity(dev(Ra))
    s.set_state(*(dev(a) for a in state))
    if fused:
        s.step(nsteps)
    else:
        s.step_unfused(nsteps)
    out = [host(t) for t in s.get_state()]
    assert not s.nonfinite()
    return out


def init_state(N, kind, seed):
    om = O.Mesh(N, N, 2 * np.pi)
    q11, q12 = synth.uniform_q(N, N, seed) if kind == "uniform" else synth.bandlimited_q(N, N, seed)
    return [np.zeros(om.spec_shape, complex), om.fwd(q11), om.fwd(q12)]


@pytest.mark.parametrize("N,nsteps,kind", [(16, 1, "bandlimited"), (64, 1, "bandlimited"),
                                           (64, 10, "bandlimited"), (128, 100, "bandlimited"),
                                           (64, 100, "uniform")])
def test_fused_step_vs_oracle(N, nsteps, kind):
    banner(f"fused IFRK2 N={N} steps={nsteps} init={kind}")
    st = init_state(N, kind, seed=3)
    ref = oracle_run(N, CONFIG, 1e-3, st, nsteps)
    got = gpu_run(N, CONFIG, 1e-3, st, nsteps)
    for g, r in zip(got, ref):
        assert rel_l2(g, r) < 1e-12


def test_fused_step_kat_params_and_omega():
    """Reference parameters (Re = Er = 0.1, lambda = 0.1, D_Q = 1) with a
    non-zero vorticity and a mean vorticity component."""
    banner("fused IFRK2 with reference parameters")
    N = 64
    om = O.Mesh(N, N, 2 * np.pi)
    X, Y = om.grid()
    w = -2 * np.sin(X + Y) + 0.3 * np.cos(3 * X - 2 * Y) + 0.05
    r1 = 0.2 * np.cos(2 * X + Y) + 0.1 * np.sin(X)
    r2 = 0.2 * np.sin(2 * X + Y) - 0.1 * np.cos(3 * Y)
    st = [om.fwd(w), om.fwd(r1), om.fwd(r2)]
    d = dict(KAT_PARAMS)
    ref = oracle_run(N, d, 1e-4, st, 20)
    got = gpu_run(N, d, 1e-4, st, 20)
    for g, r in zip(got, ref):
        assert rel_l2(g, r) < 1e-12


def test_fused_activity_field_and_no_advection():
    banner("fused IFRK2 with activity field / advect_w=0")
    N = 64
    st = init_state(N, "bandlimited", seed=5)
    Ra = synth.activity_field(N, N, seed=7)
    ref = oracle_run(N, CONFIG, 1e-3, st, 10, Ra=Ra)
    got = gpu_run(N, CONFIG, 1e-3, st, 10, Ra=Ra)
    for g, r in zip(got, ref):
        assert rel_l2(g, r) < 1e-12
    ref = oracle_run(N, CONFIG, 1e-3, st, 10, advect_w=0)
    got = gpu_run(N, CONFIG, 1e-3, st, 10, advect_w=0)
    for g, r in zip(got, ref):
        assert rel_l2(g, r) < 1e-12


def test_unfused_step_vs_oracle():
    banner("unfused (per-op) IFRK2 N=64")
    N = 64
    st = init_state(N, "bandlimited", seed=9)
    ref = oracle_run(N, CONFIG, 1e-3, st, 10)
    got = gpu_run(N, CONFIG, 1e-3, st, 10, fused=False)
    for g, r in zip(got, ref):
        assert rel_l2(g, r) < 1e-12


def test_config1_golden():
    """64^2 CPU-path workload: 2000 IFRK2 steps against the oracle's fixture."""
    banner("config-1 golden 2000 steps")
    g = np.load(os.path.join(HERE, "golden", "config1_64.npz"))
    st = [g["w0"], g["r10"], g["r20"]]
    got10 = gpu_run(64, CONFIG, 1e-3, st, 10)
    for a, k in zip(got10, ("w_10", "r1_10", "r2_10")):
        assert rel_l2(a, g[k]) < 1e-12
    got = gpu_run(64, CONFIG, 1e-3, st, 2000)
    for a, k in zip(got, ("w_2000", "r1_2000", "r2_2000")):
        assert rel_l2(a, g[k]) < 1e-11


def test_1024_fused_vs_unfused_and_determinism():
    """Full-size 1024^2: fused == unfused (reference call structure) to 1e-12,
    bitwise-deterministic, finite, and the 2/3-rule mask holds."""
    banner("1024^2 fused vs unfused, determinism")
    N = 1024
    q11, q12 = synth.bandlimited_q(N, N, seed=2)
    m = Q.Mesh(N, N, 2 * np.pi)
    st = [m.spec(), m.spec(), m.spec()]
    m.FwdTrans(st[1], dev(q11))
    m.FwdTrans(st[2], dev(q12))
    states = []
    for fused in (True, True, False):
        s = Q.Solver(m, qparams(CONFIG, 5e-4))
        s.set_state(*st)
        if fused:
            s.step(5)
        else:
            s.step_unfused(5)
        states.append([host(t) for t in s.get_state()])
        assert not s.nonfinite()
        s.close()
    for a, b in zip(states[0], states[1]):
        assert np.array_equal(a, b)
    for a, b in zip(states[0], states[2]):
        assert rel_l2(a, b) < 1e-12
    om_cut = host(m.tables()[3])
    for a in states[0]:
        assert np.all(a[om_cut == 0] == 0)


@pytest.mark.parametrize("method,N,nsteps,kind", [("rk4", 64, 20, "bandlimited"), ("rk4", 128, 50, "uniform"),
                                                  ("ife", 64, 30, "bandlimited"), ("ife", 32, 100, "uniform")])
def test_fused_rk4_and_ife_vs_oracle(method, N, nsteps, kind):
    """Fused IF-RK4 (the reference driver's integrator, RK4.cuh with IF, IFh)
    and IF-Euler steps against the oracle's stage kernels."""
    banner(f"fused {method} N={N} steps={nsteps} init={kind}")
    st = init_state(N, kind, seed=5)
    om = O.Mesh(N, N, 2 * np.pi)
    s = O.Solver(om, oparams(CONFIG, 1e-3, 1))
    s.set_state(*st)
    s.step(nsteps, O.RK4 if method == "rk4" else O.IFE)
    ref = s.get_state()
    m = Q.Mesh(N, N, 2 * np.pi)
    g = Q.Solver(m, qparams(CONFIG, 1e-3, 1))
    g.set_state(*(dev(a) for a in st))
    (g.step_rk4 if method == "rk4" else g.step_ife)(nsteps)
    got = [host(t) for t in g.get_state()]
    assert not g.nonfinite()
    for a, b in zip(got, ref):
        assert rel_l2(a, b) < 1e-12


@pytest.mark.parametrize("method,Nx,Ny,nsteps", [("ifrk2", 64, 32, 20), ("ifrk2", 32, 128, 20),
                                                 ("ifrk2", 256, 64, 10), ("ifrk2", 16, 64, 20),
                                                 ("ifrk2", 1024, 16, 3), ("rk4", 128, 32, 10),
                                                 ("ife", 32, 64, 20)])
def test_fused_steps_non_square_vs_oracle(method, Nx, Ny, nsteps):
    """Fused steps on non-square power-of-two meshes (the row pass sized by
    Nx, the column stages by Ny) against the oracle."""
    banner(f"fused {method} {Nx}x{Ny} steps={nsteps}")
    rng = np.random.default_rng(Nx + 3 * Ny)
    om = O.Mesh(Nx, Ny, 2 * np.pi)
    st = [om.fwd(0.1 * rng.standard_normal((Ny, Nx))), om.fwd(rng.uniform(-0.01, 0.01, (Ny, Nx))),
          om.fwd(rng.uniform(-0.01, 0.01, (Ny, Nx)))]
    s = O.Solver(om, oparams(CONFIG, 1e-3))
    s.set_state(*st)
    s.step(nsteps, {"ifrk2": O.IFRK2, "rk4": O.RK4, "ife": O.IFE}[method])
    ref = s.get_state()
    m = Q.Mesh(Nx, Ny, 2 * np.pi)
    g = Q.Solver(m, qparams(CONFIG, 1e-3))
    g.set_state(*(dev(a) for a in st))
    {"ifrk2": g.step, "rk4": g.step_rk4, "ife": g.step_ife}[method](nsteps)
    got = [host(t) for t in g.get_state()]
    assert not g.nonfinite()
    for a, b in zip(got, ref):
        assert rel_l2(a, b) < 1e-12
    g.close()


@pytest.mark.parametrize("method", ["ifrk2", "rk4", "ife"])
@pytest.mark.parametrize("N", [64, 96, 256])
def test_restart_from_state_continues_bit_for_bit(method, N):
    """A run restarted from its state (set_state: the column prologue rebuilds
    the inverse set) continues exactly as the uninterrupted run, whose inverse
    set was chained by the previous step's final column stage: the prologue is
    that same stage kernel with its update skipped."""
    banner(f"restart {method} N={N}")
    st = [dev(a) for a in init_state(N, "bandlimited", seed=7)]
    m = Q.Mesh(N, N, 2 * np.pi)
    a = Q.Solver(m, qparams(CONFIG, 1e-3))
    b = Q.Solver(m, qparams(CONFIG, 1e-3))
    step = {"ifrk2": "step", "rk4": "step_rk4", "ife": "step_ife"}[method]
    a.set_state(*st)
    getattr(a, step)(3)
    b.set_state(*a.get_state())
    getattr(a, step)(2)
    getattr(b, step)(2)
    for x, y in zip(a.get_state(), b.get_state()):
        assert torch.equal(x, y)
    a.close()
    b.close()
