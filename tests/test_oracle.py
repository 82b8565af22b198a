"""Pins the FP64 CPU oracle before it is trusted as the checker.

Each test names what fixes the answer independently of the oracle's code:
brute-force DFT sums, numpy's pocketfft, the reference's own known-answer
tests and data files, and closed-form solutions.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle_py as O
from kat_cases import KAT_PARAMS, KAT_TOL, N_MIN, fld_case, nonlinear_case

HERE = os.path.dirname(os.path.abspath(__file__))


def brute_dft(x, sign):
    n = len(x)
    k = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(k, k) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 12, 16, 27, 64])
def test_fft1d_bruteforce(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        ref = brute_dft(x, sign)
        got = O.fft1d(x, sign)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("n", [128, 768, 1024, 4096])
def test_fft1d_numpy(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.abs(O.fft1d(x, -1) - np.fft.fft(x)).max() <= 1e-15 * np.abs(x).sum()
    assert np.abs(O.fft1d(x, +1) - n * np.fft.ifft(x)).max() <= 1e-15 * np.abs(x).sum()


def brute_r2c2d(f):
    Ny, Nx = f.shape
    Nxh = Nx // 2 + 1
    y = np.arange(Ny)[:, None]
    x = np.arange(Nx)[None, :]
    F = np.empty((Ny, Nxh), complex)
    for j in range(Ny):
        for i in range(Nxh):
            F[j, i] = (f * np.exp(-2j * np.pi * (i * x / Nx + j * y / Ny))).sum()
    return F


def brute_c2r2d(F, Nx):
    """The fixed C2R semantics by definition: inverse DFT along y of each
    column, then the x-direction sum that keeps only the real part of the
    DC and Nyquist bins; divided by Nx*Ny (FldOp/QFldfuncs.cuh:8-16)."""
    Ny, Nxh = F.shape
    ys = np.arange(Ny)
    G = np.array([[(F[:, i] * np.exp(2j * np.pi * np.arange(Ny) * yy / Ny)).sum()
                   for i in range(Nxh)] for yy in ys])
    f = np.zeros((Ny, Nx))
    xs = np.arange(Nx)
    for yy in ys:
        acc = G[yy, 0].real * np.ones(Nx)
        for i in range(1, Nxh):
            if 2 * i == Nx:
                acc = acc + G[yy, i].real * np.cos(np.pi * xs)
            else:
                acc = acc + 2 * (G[yy, i] * np.exp(2j * np.pi * i * xs / Nx)).real
        f[yy] = acc
    return f / (Nx * Ny)


@pytest.mark.parametrize("Nx,Ny", [(8, 8), (12, 12), (8, 6), (16, 8)])
def test_fwd_bruteforce(Nx, Ny):
    rng = np.random.default_rng(Nx * 100 + Ny)
    m = O.Mesh(Nx, Ny)
    f = rng.standard_normal((Ny, Nx))
    f0 = f.copy()
    ft = m.fwd(f)
    assert np.array_equal(f, f0), "FwdTrans must preserve its input"
    ref = brute_r2c2d(f)
    assert np.abs(ft - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("Nx,Ny", [(8, 8), (12, 12), (8, 6)])
def test_bwd_nonhermitian_semantics(Nx, Ny):
    """Non-Hermitian column 0 / Nyquist column (Basic/Mesh.cuh:36-42 makes the
    dealiased state non-Hermitian in column 0)."""
    rng = np.random.default_rng(7 + Nx + Ny)
    m = O.Mesh(Nx, Ny)
    F = rng.standard_normal((Ny, Nx // 2 + 1)) + 1j * rng.standard_normal((Ny, Nx // 2 + 1))
    ref = brute_c2r2d(F.copy(), Nx)
    got = m.bwd(F.copy())
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
    # numpy/pocketfft irfft2 applies the same semantics (independent library)
    assert np.abs(got - np.fft.irfft2(F, s=(Ny, Nx))).max() <= 1e-14 * np.abs(ref).max()


def test_bwd_side_effect_symmetrises_column0_only():
    """FldOp/QFldfuncs.cuh:18-44,50-55: the caller's spectrum is symmetrised in
    column 0 for 0<j<Ny/2 pairs; j=0, j=Ny/2 and other columns untouched."""
    Nx = Ny = 16
    rng = np.random.default_rng(3)
    m = O.Mesh(Nx, Ny)
    F = rng.standard_normal((Ny, 9)) + 1j * rng.standard_normal((Ny, 9))
    F0 = F.copy()
    m.bwd(F)
    assert np.array_equal(F[:, 1:], F0[:, 1:])
    assert F[0, 0] == F0[0, 0] and F[Ny // 2, 0] == F0[Ny // 2, 0]
    for j in range(1, Ny // 2):
        mean = 0.5 * (F0[j, 0] + np.conj(F0[Ny - j, 0]))
        assert abs(F[j, 0] - mean) < 1e-15 and abs(F[Ny - j, 0] - np.conj(mean)) < 1e-15


def test_roundtrip_and_parseval():
    N = 64
    rng = np.random.default_rng(11)
    m = O.Mesh(N, N)
    f = rng.standard_normal((N, N))
    ft = m.fwd(f)
    assert np.abs(m.bwd(ft.copy()) - f).max() < 1e-14
    # Parseval with half-spectrum weights (columns 1..N/2-1 count twice)
    wgt = np.full(N // 2 + 1, 2.0)
    wgt[0] = wgt[-1] = 1.0
    assert abs((np.abs(ft) ** 2 * wgt).sum() / N**2 - (f**2).sum()) < 1e-10 * (f**2).sum()


def test_wave_tables_reference_fixture():
    """Basic/Mesh.cuh:6-21 against the reference's dump kx.txt / ky.txt
    (768^2, L = 66 pi; Nyquist row j=384 carries +ky)."""
    with open(os.path.join(HERE, "golden", "wave_768_L66pi.json")) as fh:
        fx = json.load(fh)
    m = O.Mesh(768, 768, 33 * 2 * np.pi)
    for a, name in ((m.kx, "kx"), (m.ky, "ky")):
        want = np.array(fx[name])
        got = a[np.ix_(fx["rows"], fx["cols"])]
        assert np.allclose(got, want, rtol=1e-5, atol=1e-6), name
    assert m.ky[384, 0] > 0 and m.ky[385, 0] < 0


@pytest.mark.parametrize("N", [12, 64, 1024])
def test_cutoff_closed_form(N):
    """Basic/Mesh.cuh:36-42: keeps i <= (Nxh-1)*2/3 and j <= Ny/3 or j > Ny-Ny/3."""
    m = O.Mesh(N, N)
    Nxh = N // 2 + 1
    i = np.arange(Nxh)[None, :]
    j = np.arange(N)[:, None]
    want = ((i < (Nxh - 1) * 2 // 3 + 1) & ((j < N // 3 + 1) | ((N - j) < N // 3))).astype(float)
    assert np.array_equal(m.cutoff, want)
    if N == 1024:
        assert m.cutoff[341, 341] == 1 and m.cutoff[342, 0] == 0 and m.cutoff[683, 0] == 0
        assert m.cutoff[684, 0] == 1 and m.cutoff[0, 342] == 0


def test_basictest_12():
    """UnitTest/Fld/BasicTest.cu:34-143: omega = -2 cos x cos y on 12x12."""
    m = O.Mesh(12, 12)
    X, Y = m.grid()
    ft = m.fwd(-2 * np.cos(X) * np.cos(Y))
    assert abs(ft[1, 1] + 72) < 1e-12 and abs(ft[11, 1] + 72) < 1e-12
    others = np.abs(ft).copy()
    others[1, 1] = others[11, 1] = 0
    assert others.max() < 1e-12


def _fields(m, names):
    return {n: O.Field(m) for n in names}


def test_fld_kat():
    """UnitTest/Fld/Fld.cu:65-160 (L = 66 pi, phi = sin(x+y))."""
    N = 128
    m = O.Mesh(N, N, 33 * 2 * np.pi)
    X, Y = m.grid()
    kat = fld_case(X, Y)
    phi = m.fwd(kat["phi"])
    assert np.abs(m.bwd(phi.copy()) - kat["phi"]).max() < 1e-13
    k2, kx, ky = m.k_squared, m.kx, m.ky
    lap = -1.0 * k2 * phi
    assert np.abs(m.bwd(lap.copy()) - kat["lap"]).max() < 1e-12
    assert np.abs(m.bwd(1j * kx * phi) - kat["dx"]).max() < 1e-12
    assert np.abs(m.bwd(-(1j * ky * phi)) - kat["mdy"]).max() < 1e-12
    # vel_func through the oracle's curr_func path: u = i ky w/k^2, v = -i kx w/k^2
    f = _fields(m, ["r1", "r2", "w", "u", "v", "S"])
    f["w"].spec[:] = lap
    O.lib().or_curr_func(f["r1"].h, f["r2"].h, f["w"].h, f["u"].h, f["v"].h, f["S"].h)
    assert np.abs(f["u"].phys - kat["velu"]).max() < 1e-12
    assert np.abs(f["v"].phys - kat["velv"]).max() < 1e-12
    # divergence-free and curl(vel(w)) = w
    div = m.bwd(1j * kx * f["u"].spec + 1j * ky * f["v"].spec)
    assert np.abs(div).max() < 1e-12
    curl = m.bwd(1j * kx * f["v"].spec - 1j * ky * f["u"].spec)
    assert np.abs(curl - kat["lap"]).max() < 1e-12


NL_NAMES = "w u v r1 r2 S p11 p12 p21 Ra aux aux1 wn r1n r2n".split()


def oracle_nonlinear_kat(c, N):
    m = O.Mesh(N, N)
    X, Y = m.grid()
    (w, r1, r2), exact = nonlinear_case(c, X, Y)
    f = _fields(m, NL_NAMES)
    for nm, val in (("w", w), ("r1", r1), ("r2", r2)):
        f[nm].spec[:] = m.fwd(val)
        f[nm].phys[:] = m.bwd(f[nm].spec)
    f["Ra"].phys[:] = KAT_PARAMS["alpha"]
    L = O.lib()
    lam, Re, Er = KAT_PARAMS["lam"], KAT_PARAMS["Re"], KAT_PARAMS["Er"]
    h = {k: v.h for k, v in f.items()}
    L.or_curr_func(h["r1"], h["r2"], h["w"], h["u"], h["v"], h["S"])
    L.or_wnonl_func(h["wn"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["Ra"], h["S"], h["p11"],
                    h["p12"], h["p21"], Re, Er, lam, h["aux"], h["aux1"])
    L.or_r1nonl_func(h["r1n"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["S"], lam, h["aux"], h["aux1"])
    L.or_r2nonl_func(h["r2n"], h["w"], h["u"], h["v"], h["r1"], h["r2"], h["S"], lam, h["aux"], h["aux1"])
    got = [m.bwd(f[k].spec.copy()) for k in ("wn", "r1n", "r2n")]
    return got, exact


@pytest.mark.parametrize("c", [0, 1, 2, 3, 4])
def test_nonlinear_kats(c):
    got, exact = oracle_nonlinear_kat(c, N_MIN[c])
    for g, e, tol in zip(got, exact, KAT_TOL[c]):
        assert np.abs(g - e).max() <= tol * np.abs(e).max()


def test_solver_rhs_matches_composites():
    """or_solver_rhs with coefficients 1 is the same composition as the KAT."""
    c = 2
    N = 32
    got, exact = oracle_nonlinear_kat(c, N)
    m = O.Mesh(N, N)
    X, Y = m.grid()
    (w, r1, r2), _ = nonlinear_case(c, X, Y)
    s = O.Solver(m, O.params_array(**{k: v for k, v in KAT_PARAMS.items()}))
    outs = s.rhs(m.fwd(w), m.fwd(r1), m.fwd(r2))
    for o, g in zip(outs, got):
        assert np.abs(m.bwd(o.copy()) - g).max() <= 1e-12 * np.abs(g).max()


@pytest.mark.parametrize("method", [O.IFE, O.IFRK2, O.RK4])
def test_taylor_green(method):
    """UnitTest/TimeIntegration/Taylor/Taylor.cu: omega = -2 cos x cos y, Q = 0,
    Re = 40, Rf = 0, dt = 0.02; exact omega_n = omega_0 exp(-n dt 2/Re)."""
    N = 32
    m = O.Mesh(N, N)
    X, Y = m.grid()
    w0 = -2 * np.cos(X) * np.cos(Y)
    p = O.params_array(lam=0.1, Re=40.0, Er=40.0, Rf=0.0, alpha=0.2, dt=0.02)
    s = O.Solver(m, p)
    z = np.zeros(m.spec_shape, complex)
    s.set_state(m.fwd(w0), z, z)
    n = 50
    s.step(n, method)
    w, r1, r2 = s.get_state()
    wn = m.bwd(w.copy())
    exact = w0 * math.exp(-n * 0.02 * 2 / 40.0)
    assert np.abs(wn - exact).max() < 1e-12
    assert np.abs(r1).max() == 0 and np.abs(r2).max() == 0


def helical_exact(eps0, a, b, t):
    e = math.exp(2 * a * t)
    return math.sqrt(a * eps0**2 * e / (a + b * eps0**2 * (e - 1)))


def helical_run(method, dt, T, eps0=0.3, DQ=1.0, aQ=2.0, cQ=1.0, N=16):
    m = O.Mesh(N, N)
    X, Y = m.grid()
    p = O.params_array(lam=0.0, Re=1.0, Er=1.0, Rf=0.1, alpha=0.0, DQ=DQ, aQ=aQ, cQ=cQ, dt=dt)
    s = O.Solver(m, p)
    z = np.zeros(m.spec_shape, complex)
    s.set_state(z, m.fwd(eps0 * np.cos(X)), m.fwd(eps0 * np.sin(X)))
    s.step(int(round(T / dt)), method)
    w, r1, r2 = s.get_state()
    r1p = m.bwd(r1.copy())
    r2p = m.bwd(r2.copy())
    amp = np.sqrt(r1p**2 + r2p**2)
    assert np.abs(m.bwd(w.copy())).max() < 1e-14
    return amp.mean(), amp.max() - amp.min()


@pytest.mark.parametrize("method,order", [(O.IFE, 1), (O.IFRK2, 2), (O.RK4, 4)])
def test_helical_mode_order(method, order):
    """Single helical Q mode, lambda = alpha = 0: omega stays 0 and the
    amplitude obeys eps' = (aQ - DQ) eps - 4 cQ eps^3 (closed form)."""
    T = 1.0
    a, b = 2.0 - 1.0, 4.0
    exact = helical_exact(0.3, a, b, T)
    errs = []
    for dt in (0.05, 0.025):
        amp, spread = helical_run(method, dt, T)
        assert spread < 1e-13
        errs.append(abs(amp - exact))
    ratio = errs[0] / errs[1]
    assert 2 ** order * 0.8 < ratio < 2 ** order * 1.25, (errs, ratio)


def test_linear_mode_exact():
    """Tiny single Q mode with lambda = alpha = 0: r(t) = r0 exp((aQ - DQ k^2) t)."""
    N = 32
    m = O.Mesh(N, N)
    X, Y = m.grid()
    eps = 1e-9
    DQ, aQ, dt, n = 0.05, 1.0, 1e-3, 200
    p = O.params_array(lam=0.0, Re=1.0, Er=1.0, Rf=0.1, alpha=0.0, DQ=DQ, aQ=aQ, dt=dt)
    s = O.Solver(m, p)
    z = np.zeros(m.spec_shape, complex)
    r0 = eps * np.cos(3 * X + 2 * Y)
    s.set_state(z, m.fwd(r0), z)
    s.step(n, O.IFRK2)
    _, r1, _ = s.get_state()
    got = m.bwd(r1.copy())
    want = r0 * math.exp((aQ - DQ * 13) * n * dt)
    assert np.abs(got - want).max() < 1e-12 * eps * 1e3
    IFw, IFwh, IFr1, IFr1h, IFr2, IFr2h = s.lin_tables()
    k2 = m.k_squared
    assert np.allclose(IFr1, np.exp((aQ - DQ * k2) * dt), rtol=1e-15, atol=0)
    assert np.allclose(IFwh, np.exp((-k2 - 0.1) * dt / 2), rtol=1e-15, atol=0)
