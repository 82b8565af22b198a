"""Meshes with sides 3 * 2^a (the reference driver's production grid is
768 = 512 * 3/2, friction.cu:109): per-op transforms with a radix-3 Stockham
pass, the wave tables against the reference's own 768^2 fixture, and the
unfused and fused IFRK2 / RK4 / IFE steps against the oracle (mixed-radix
FFT)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import oracle_py as O  # noqa: E402
import qsolver as Q  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
CONFIG = dict(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, DQ=0.05, aQ=1.0, cQ=1.0)
# friction.cu:100-105, 109-112: the driver's own parameters and box
DRIVER = dict(lam=0.1, Re=0.1, Er=0.1, Rf=0.00075, alpha=0.2)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_l2(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def qparams(d, dt):
    return Q.params(lam=d["lam"], Re=d["Re"], Er=d["Er"], Rf=d["Rf"], alpha=d["alpha"], D_Q=d.get("DQ", 1.0),
                    a_Q=d.get("aQ", 1.0), c_Q=d.get("cQ", 1.0), dt=dt)


def oparams(d, dt):
    return O.params_array(lam=d["lam"], Re=d["Re"], Er=d["Er"], Rf=d["Rf"], alpha=d["alpha"],
                          DQ=d.get("DQ", 1.0), aQ=d.get("aQ", 1.0), cQ=d.get("cQ", 1.0), dt=dt)


@pytest.mark.parametrize("Nx,Ny", [(48, 48), (96, 96), (768, 768), (96, 48), (64, 192), (1536, 48)])
def test_transforms_3x2a_vs_oracle_and_numpy(Nx, Ny):
    print(f"[gpu-test] radix-3 transforms {Nx}x{Ny}", flush=True)
    rng = np.random.default_rng(Nx + Ny)
    m = Q.Mesh(Nx, Ny, 2 * np.pi)
    om = O.Mesh(Nx, Ny, 2 * np.pi)
    f = rng.standard_normal((Ny, Nx))
    ft = m.spec()
    m.FwdTrans(ft, dev(f))
    got = host(ft)
    assert rel_l2(got, om.fwd(f)) < 1e-14
    assert rel_l2(got, np.fft.rfft2(f)) < 1e-14
    F = rng.standard_normal((Ny, Nx // 2 + 1)) + 1j * rng.standard_normal((Ny, Nx // 2 + 1))
    g = m.phys()
    Fd = dev(F)
    m.BwdTrans(g, Fd)
    F_or = F.copy()
    assert rel_l2(host(g), om.bwd(F_or)) < 1e-14
    assert np.abs(host(Fd) - F_or).max() < 4e-15 * np.abs(F).max()


def test_tables_768_against_the_reference_fixture():
    """post_process/kx.txt, ky.txt (768^2, L = 66 pi, 6 digits), as reduced
    into tests/golden/wave_768_L66pi.json, now on the GPU mesh itself."""
    print("[gpu-test] 768 tables", flush=True)
    with open(os.path.join(HERE, "golden", "wave_768_L66pi.json")) as fh:
        fx = json.load(fh)
    m = Q.Mesh(768, 768, 33 * 2 * np.pi)
    kx, ky, k2, cut = (host(t) for t in m.tables())
    for a, name in ((kx, "kx"), (ky, "ky")):
        got = a[np.ix_(fx["rows"], fx["cols"])]
        assert np.allclose(got, np.array(fx[name]), rtol=1e-5, atol=1e-6), name
    assert ky[384, 0] > 0 and ky[385, 0] < 0
    om = O.Mesh(768, 768, 33 * 2 * np.pi)
    assert np.array_equal(kx, om.kx) and np.array_equal(ky, om.ky)
    assert np.array_equal(k2, om.k_squared) and np.array_equal(cut, om.cutoff)


def init_q(N, Nx=None, seed=1):
    Nx = N if Nx is None else Nx
    u = Q.synth_uniform(seed, -0.01, 0.01, 2 * N * Nx)
    return u[: N * Nx].reshape(N, Nx), u[N * Nx:].reshape(N, Nx)


@pytest.mark.parametrize("method,Nx,Ny,nsteps", [("ifrk2", 96, 96, 10), ("ifrk2", 96, 48, 10),
                                                 ("rk4", 96, 96, 6), ("rk4", 48, 192, 6)])
def test_unfused_steps_3x2a_vs_oracle(method, Nx, Ny, nsteps):
    print(f"[gpu-test] unfused {method} {Nx}x{Ny}", flush=True)
    om = O.Mesh(Nx, Ny, 2 * np.pi)
    q11, q12 = init_q(Ny, Nx, seed=3)
    rng = np.random.default_rng(4)
    st = [om.fwd(0.1 * rng.standard_normal((Ny, Nx))), om.fwd(q11), om.fwd(q12)]
    s = O.Solver(om, oparams(CONFIG, 1e-3))
    s.set_state(*st)
    s.step(nsteps, O.RK4 if method == "rk4" else O.IFRK2)
    ref = s.get_state()
    m = Q.Mesh(Nx, Ny, 2 * np.pi)
    g = Q.Solver(m, qparams(CONFIG, 1e-3))
    g.set_state(*(dev(a) for a in st))
    (g.step_rk4_unfused if method == "rk4" else g.step_unfused)(nsteps)
    for a, b in zip((host(t) for t in g.get_state()), ref):
        assert rel_l2(a, b) < 1e-12
    if Nx != Ny:   # fused kernels of 3 * 2^a sides are square only: an error, no launch
        with pytest.raises(Q.QsolverError) as e:
            g.step(1)
        assert e.value.status == Q.QS_ERR_UNSUPPORTED_SIZE
    g.close()


@pytest.mark.parametrize("method,N,nsteps", [("ifrk2", 48, 10), ("ifrk2", 96, 20), ("ifrk2", 192, 5),
                                             ("ifrk2", 384, 3), ("ifrk2", 768, 2), ("rk4", 96, 6),
                                             ("rk4", 768, 1), ("ife", 96, 10)])
def test_fused_steps_3x2a_vs_oracle(method, N, nsteps):
    """The fused kernels on square 3 * 2^a meshes (radix-12 first pass, then
    radix-4/2 passes per wavefront) against the oracle."""
    print(f"[gpu-test] fused {method} {N}^2", flush=True)
    om = O.Mesh(N, N, 2 * np.pi)
    q11, q12 = init_q(N, seed=5)
    rng = np.random.default_rng(6)
    st = [om.fwd(0.1 * rng.standard_normal((N, N))), om.fwd(q11), om.fwd(q12)]
    s = O.Solver(om, oparams(CONFIG, 1e-3))
    s.set_state(*st)
    s.step(nsteps, {"ifrk2": O.IFRK2, "rk4": O.RK4, "ife": O.IFE}[method])
    ref = s.get_state()
    m = Q.Mesh(N, N, 2 * np.pi)
    g = Q.Solver(m, qparams(CONFIG, 1e-3))
    g.set_state(*(dev(a) for a in st))
    {"ifrk2": g.step, "rk4": g.step_rk4, "ife": g.step_ife}[method](nsteps)
    assert not g.nonfinite()
    for a, b in zip((host(t) for t in g.get_state()), ref):
        assert rel_l2(a, b) < 1e-12
    g.close()


def test_reference_driver_grid_rk4_768():
    """The reference driver's configuration: 768^2, L = 66 pi, its parameters
    and dt = 0.002, RK4 through the per-op calls, 2 steps against the oracle."""
    print("[gpu-test] 768^2 driver RK4", flush=True)
    N, L, dt = 768, 33 * 2 * np.pi, 0.002
    om = O.Mesh(N, N, L)
    q11, q12 = init_q(N, seed=1)
    st = [np.zeros(om.spec_shape, complex), om.fwd(q11), om.fwd(q12)]
    s = O.Solver(om, oparams(DRIVER, dt))
    s.set_state(*st)
    s.step(2, O.RK4)
    ref = s.get_state()
    m = Q.Mesh(N, N, L)
    g = Q.Solver(m, qparams(DRIVER, dt))
    g.set_state(*(dev(a) for a in st))
    g.step_rk4_unfused(2)
    got = [host(t) for t in g.get_state()]
    # the same two steps fused
    f = Q.Solver(m, qparams(DRIVER, dt))
    f.set_state(*(dev(a) for a in st))
    f.step_rk4(2)
    fused = [host(t) for t in f.get_state()]
    for x in (got, fused):
        for a, b in zip(x[1:], ref[1:]):
            assert rel_l2(a, b) < 1e-12
        # omega starts at 0 and is driven by the Q stress: compare on the Q scale
        assert np.linalg.norm(x[0] - ref[0]) <= 1e-12 * np.linalg.norm(ref[1])
    g.close()
    f.close()
