"""Unit tests of the fused kernels' building blocks on the GPU.

Each fused kernel is checked on its own against a numpy restatement of what
it computes (tests/fused_ref.py); the per-wavefront FFT against numpy's FFT.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import qsolver as Q  # noqa: E402
import fused_ref as FR  # noqa: E402

P = dict(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, D_Q=0.05, a_Q=1.0, c_Q=1.0)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("N", [16, 32, 64, 128, 256, 512, 1024, 48, 96, 192, 384, 768])
def test_wave_fft(N):
    """Per-wavefront FFT: 1024/N vectors (768/N for N = 3 * 2^a), transformed
    one wave's worth (WaveLay<N>::NV vectors) at a time, both directions, all
    twiddle quadrants."""
    print(f"[gpu-test] wave fft N={N}", flush=True)
    m = Q.Mesh(64, 64, 2 * np.pi)
    rng = np.random.default_rng(N)
    nv = 768 // N if N % 3 == 0 else 1024 // N
    x = rng.standard_normal((nv, N)) + 1j * rng.standard_normal((nv, N))
    for sign in (-1, 1):
        d = torch.from_numpy(x.copy()).cuda()
        o = torch.zeros_like(d)
        Q.call("qs_debug_wave_fft", m.handle, N, sign, Q.ptr(d), Q.ptr(o))
        ref = np.fft.fft(x, axis=1) if sign < 0 else np.fft.ifft(x, axis=1) * N
        assert np.abs(o.cpu().numpy() - ref).max() < 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("N", [32, 128, 256, 1024, 96, 768])
def test_fused_phases(N):
    """Column prologue -> inverse set; row pass -> forward set; column stage 1
    -> IFRK2 partial sums and the stage's inverse set; row pass again; column
    stage 2 -> the dealiased next state and its inverse set, against the numpy
    model of each kernel."""
    print(f"[gpu-test] fused phases N={N}", flush=True)
    rng = np.random.default_rng(N)
    kx, ky = FR.tables(N, N)
    st = [np.fft.rfft2(0.05 * rng.standard_normal((N, N))) for _ in range(3)]
    dt = 1e-3
    m = Q.Mesh(N, N, 2 * np.pi)
    g = Q.Solver(m, Q.params(dt=dt, **P))
    g.set_state(*(torch.from_numpy(a).cuda() for a in st))
    g.debug_phase(0)
    M = [g.debug_array(k).cpu().numpy() for k in range(7)]
    for a, b in zip(M, FR.inverse_set(st[0], st[1], st[2], kx, ky)):
        assert rel(a, b) < 1e-14
    g.debug_phase(1)
    F = [g.debug_array(k).cpu().numpy() for k in range(7)]
    # white-noise inputs: the k^2-weighted products amplify roundoff
    for a, b in zip(F, FR.row_pass(M, st[0][0, 0], P, N, N, kx)):
        assert rel(a, b) < 1e-12
    g.debug_phase(2)
    Nw, Nr1, Nr2 = FR.nonlinear_from_forward_set(F, kx, ky, N, N)
    k2 = kx**2 + ky**2
    IFw = np.exp((-k2 - P["Rf"]) / P["Re"] * dt)
    IFq = np.exp((-P["D_Q"] * k2 + P["a_Q"]) * dt)
    for k, (old, n, IF) in enumerate(zip(st, (Nw, Nr1, Nr2), (IFw, IFq, IFq))):
        want = old * IF + 0.5 * IF * dt * n
        assert rel(g.debug_array(10 + k).cpu().numpy(), want) < 1e-12
    Mc = [g.debug_array(k).cpu().numpy() for k in range(7)]
    curr = [(old + dt * n) * IF for old, n, IF in zip(st, (Nw, Nr1, Nr2), (IFw, IFq, IFq))]
    for a, b in zip(Mc, FR.inverse_set(curr[0], curr[1], curr[2], kx, ky)):
        assert rel(a, b) < 1e-12
    # second right-hand side: row pass on the stage-1 inverse set, then column
    # stage 2 (IFRK2.cuh:41-51, dealias, the next state over old, its inverse set)
    new1 = [g.debug_array(10 + k).cpu().numpy() for k in range(3)]
    g.debug_phase(1)
    F2 = [g.debug_array(k).cpu().numpy() for k in range(7)]
    for a, b in zip(F2, FR.row_pass(Mc, curr[0][0, 0], P, N, N, kx)):
        assert rel(a, b) < 1e-12
    g.debug_phase(3)
    N2 = FR.nonlinear_from_forward_set(F2, kx, ky, N, N)
    i = np.arange(N // 2 + 1)[None, :]
    j = np.arange(N)[:, None]
    mask = (i < (N // 2) * 2 // 3 + 1) & ((j < N // 3 + 1) | (N - j < N // 3))
    nxt = [np.where(mask, a + 0.5 * dt * n, 0) for a, n in zip(new1, N2)]
    for k in range(3):
        assert rel(g.debug_array(7 + k).cpu().numpy(), nxt[k]) < 1e-12
    M2 = [g.debug_array(k).cpu().numpy() for k in range(7)]
    for a, b in zip(M2, FR.inverse_set(nxt[0], nxt[1], nxt[2], kx, ky)):
        assert rel(a, b) < 1e-12
