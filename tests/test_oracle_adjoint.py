"""The CPU discrete adjoint of the IFRK2 step (oracle), pinned without any
adjoint code of the reference (it has none):

* exact dot-product identities <A x, y> = <x, A^T y> for the linear transform
  pieces (C2R with the A10 semantics, R2C), to roundoff;
* central finite differences of the forward oracle for the right-hand side's
  VJP, one step's VJP and the whole-horizon gradient with respect to the
  activity field and to a per-step activity amplitude.
"""
import numpy as np
import pytest

import oracle_py as O

P = dict(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, DQ=0.05, aQ=1.0, cQ=1.0)


def rdot(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.sum(a.real * b.real + a.imag * b.imag))


def bandlimited(m, rng, kmax, rms):
    """Real field with modes |k| <= kmax, random phases, given rms."""
    Ny, Nx = m.phys_shape
    G = np.zeros(m.spec_shape, complex)
    for j in range(Ny):
        ky = j if j <= Ny // 2 else j - Ny
        for i in range(Nx // 2 + 1):
            if 0 < ky * ky + i * i <= kmax * kmax:
                G[j, i] = np.exp(2j * np.pi * rng.random())
    f = np.fft.irfft2(G, s=(Ny, Nx))
    return f * rms / np.sqrt(np.mean(f * f))


def setup(N=32, dt=1e-3, seed=3, activity=None, advect_w=1, Ny=None):
    m = O.Mesh(N, N if Ny is None else Ny, 2 * np.pi)
    s = O.Solver(m, O.params_array(lam=P["lam"], Re=P["Re"], Er=P["Er"], Rf=P["Rf"], alpha=P["alpha"],
                                   DQ=P["DQ"], aQ=P["aQ"], cQ=P["cQ"], dt=dt, advect_w=advect_w))
    rng = np.random.default_rng(seed)
    q11, q12 = bandlimited(m, rng, 6, 0.05), bandlimited(m, rng, 6, 0.05)
    w = bandlimited(m, rng, 5, 0.5)
    X = [m.fwd(w), m.fwd(q11), m.fwd(q12)]
    if activity is not None:
        s.set_activity(activity)
    return m, s, X, rng


def rand_spec(m, rng):
    return rng.standard_normal(m.spec_shape) + 1j * rng.standard_normal(m.spec_shape)


@pytest.mark.parametrize("N,Ny", [(8, 8), (12, 12), (32, 32), (16, 8), (8, 32)])
def test_transform_transposes(N, Ny):
    m, s, _, rng = setup(N, Ny=Ny)
    A = O.Adjoint(s)
    for _ in range(3):
        G = rand_spec(m, rng)            # arbitrary, non-Hermitian edge columns included
        f = rng.standard_normal(m.phys_shape)
        lhs, rhs = rdot(A.B(G), f), rdot(G, A.BT(f))
        assert abs(lhs - rhs) <= 1e-13 * (abs(lhs) + np.linalg.norm(G) * np.linalg.norm(f) / N)
        lhs, rhs = rdot(A.F(f), G), rdot(f, A.FT(G))
        assert abs(lhs - rhs) <= 1e-13 * (abs(lhs) + np.linalg.norm(G) * np.linalg.norm(f) * N)


def fd_dir(fun, eps):
    return (fun(eps) - fun(-eps)) / (2 * eps)


@pytest.mark.parametrize("advect_w", [1, 0])
def test_rhs_vjp_finite_difference(advect_w):
    m, s, X, rng = setup(32, advect_w=advect_w)
    Ra = 1.0 + 0.1 * bandlimited(m, rng, 4, 1.0)
    s.set_activity(Ra)
    A = O.Adjoint(s)
    Nb = [rand_spec(m, rng) * 1e-3 for _ in range(3)]
    Xb, ab = A.rhs_vjp(X, Nb)
    for trial in range(3):
        d = [rand_spec(m, rng) * np.linalg.norm(x) / np.sqrt(x.size) * 0.1 for x in X]

        def g(e):
            Ns = s.rhs(*(x + e * dx for x, dx in zip(X, d)))
            return sum(rdot(n, b) for n, b in zip(Ns, Nb))
        fd = fd_dir(g, 1e-4)
        an = sum(rdot(xb, dx) for xb, dx in zip(Xb, d))
        assert abs(fd - an) <= 2e-7 * abs(an), (trial, fd, an)
    # activity direction
    da = bandlimited(m, rng, 5, 1.0)

    def h(e):
        s.set_activity(Ra + e * da)
        Ns = s.rhs(*X)
        return sum(rdot(n, b) for n, b in zip(Ns, Nb))
    fd = fd_dir(h, 1e-4)
    an = rdot(ab, da)
    assert abs(fd - an) <= 1e-8 * abs(an), (fd, an)


@pytest.mark.parametrize("N,Ny", [(32, 32), (32, 16), (16, 64)])
def test_step_vjp_finite_difference(N, Ny):
    m, s, X, rng = setup(N, dt=2e-3, Ny=Ny)
    A = O.Adjoint(s)
    gb = [rand_spec(m, rng) for _ in range(3)]
    ub, ab = A.step_vjp(X, gb)

    def step(Xs):
        s.set_state(*Xs)
        s.step(1, O.IFRK2)
        return s.get_state()
    for trial in range(2):
        d = [rand_spec(m, rng) * np.linalg.norm(x) / np.sqrt(x.size) * 0.1 for x in X]
        fd = fd_dir(lambda e: sum(rdot(a, b) for a, b in zip(step([x + e * dx for x, dx in zip(X, d)]), gb)),
                    1e-4)
        an = sum(rdot(a, b) for a, b in zip(ub, d))
        assert abs(fd - an) <= 2e-7 * abs(an), (trial, fd, an)


@pytest.mark.parametrize("kind", [O.Adjoint.ENSTROPHY, O.Adjoint.KINETIC])
def test_gradient_wrt_activity_field(kind):
    """dL/dalpha(x) after 10 steps against central differences.  Over this
    short horizon the loss moves by ~1e-4 of itself per unit activity change,
    so the step is 1e-4 (at 1e-6 the differences of two rounded losses limit
    the check to ~1e-6)."""
    m, s, X, rng = setup(32, dt=1e-3)
    Ra = 1.0 + 0.1 * bandlimited(m, rng, 6, 1.0)
    A = O.Adjoint(s)

    def loss(Rv):
        s.set_activity(Rv)
        s.set_state(*X)
        s.step(10, O.IFRK2)
        W = s.get_state()[0]
        return A.loss_grad(W, kind)[0]
    s.set_activity(Ra)
    s.set_state(*X)
    L, ab, _ = A.gradient(10, kind)
    assert L == pytest.approx(loss(Ra), rel=1e-13)
    for trial in range(3):
        d = rng.standard_normal(m.phys_shape)
        fd = fd_dir(lambda e: loss(Ra + e * d), 1e-4)
        an = rdot(ab, d)
        assert abs(fd - an) <= 1e-6 * abs(an), (trial, fd, an)


def test_gradient_wrt_activity_schedule():
    """Per-step uniform activity a(t_n): dL/da_n against central differences
    (step 1e-3: one step's amplitude moves the loss by ~1e-5 of itself)."""
    m, s, X, rng = setup(32, dt=1e-3)
    n = 8
    a_t = rng.uniform(0.8, 1.2, n)
    A = O.Adjoint(s)
    s.set_state(*X)
    L, _, gt = A.gradient(n, O.Adjoint.KINETIC, alpha_t=a_t)

    def loss(at):
        s.set_state(*X)
        return A.gradient(n, O.Adjoint.KINETIC, alpha_t=at)[0]
    for k in (0, n // 2, n - 1):
        e = np.zeros(n)
        e[k] = 1.0
        fd = fd_dir(lambda h: loss(a_t + h * e), 1e-3)
        assert abs(fd - gt[k]) <= 1e-6 * abs(gt[k]), (k, fd, gt[k])
