"""GPU discrete adjoint (qs_adj_*) against the CPU FP64 adjoint of the oracle
and against central finite differences of the GPU forward itself."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import qsolver as Q  # noqa: E402
import oracle_py as O  # noqa: E402

P = dict(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, D_Q=0.05, a_Q=1.0, c_Q=1.0)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def bandlimited(N, rng, kmax, rms):
    G = np.zeros((N, N // 2 + 1), complex)
    for j in range(N):
        ky = j if j <= N // 2 else j - N
        for i in range(N // 2 + 1):
            if 0 < ky * ky + i * i <= kmax * kmax:
                G[j, i] = np.exp(2j * np.pi * rng.random())
    f = np.fft.irfft2(G, s=(N, N))
    return f * rms / np.sqrt(np.mean(f * f))


def case(N, dt, seed=5):
    rng = np.random.default_rng(seed)
    m = O.Mesh(N, N, 2 * np.pi)
    q11, q12 = bandlimited(N, rng, 6, 0.05), bandlimited(N, rng, 6, 0.05)
    w = bandlimited(N, rng, 5, 0.5)
    X = [m.fwd(w), m.fwd(q11), m.fwd(q12)]
    Ra = 1.0 + 0.1 * bandlimited(N, rng, 6, 1.0)
    return m, X, Ra, rng


def oracle(m, dt):
    return O.Solver(m, O.params_array(lam=P["lam"], Re=P["Re"], Er=P["Er"], Rf=P["Rf"], alpha=P["alpha"],
                                      DQ=P["D_Q"], aQ=P["a_Q"], cQ=P["c_Q"], dt=dt))


def gpu_solver(N, dt):
    mesh = Q.Mesh(N, N, 2 * np.pi)
    return mesh, Q.Solver(mesh, Q.params(dt=dt, **P))


def cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gpu_gradient(N, dt, X, Ra, nsteps, k, kind, alpha_t=None, want_state=False):
    mesh, s = gpu_solver(N, dt)
    s.set_state(*(cu(x) for x in X))
    if Ra is not None:
        s.set_activity(cu(Ra))
    s.adj_forward(nsteps, k, None if alpha_t is None else cu(alpha_t))
    W = s.get_state()[0]
    L, Wb = Q.adj_loss(mesh, W, kind)
    z = torch.zeros_like(Wb)
    ab, at, u = s.adj_backward(Wb, z, z.clone(), want_field=True, want_sched=alpha_t is not None,
                               want_state=want_state, nsteps=nsteps)
    torch.cuda.synchronize()
    out = (L, ab.cpu().numpy(), None if at is None else at.cpu().numpy(),
           None if u is None else [t.cpu().numpy() for t in u])
    s.close()
    return out


@pytest.mark.parametrize("kind", [Q.ENSTROPHY, Q.KINETIC])
def test_loss_and_gradient_vs_oracle(kind):
    print(f"[gpu-test] adj loss kind={kind}", flush=True)
    m, X, _, _ = case(64, 1e-3)
    A = O.Adjoint(oracle(m, 1e-3))
    mesh = Q.Mesh(64, 64, 2 * np.pi)
    L, Wb = Q.adj_loss(mesh, cu(X[0]), kind)
    Lc, Wbc = A.loss_grad(X[0], kind)
    assert L == pytest.approx(Lc, rel=1e-13)
    assert rel(Wb.cpu().numpy(), Wbc) < 1e-13


@pytest.mark.parametrize("kind", [Q.ENSTROPHY, Q.KINETIC])
def test_activity_field_gradient_vs_cpu_adjoint(kind):
    """10 steps at 64^2, checkpoints every 4 steps (segments 4, 4, 2)."""
    print(f"[gpu-test] adj field gradient kind={kind}", flush=True)
    N, dt, n = 64, 1e-3, 10
    m, X, Ra, _ = case(N, dt)
    s = oracle(m, dt)
    s.set_activity(Ra)
    s.set_state(*X)
    Lc, abc, _ = O.Adjoint(s).gradient(n, kind)
    L, ab, _, _ = gpu_gradient(N, dt, X, Ra, n, 4, kind)
    assert L == pytest.approx(Lc, rel=1e-11)
    assert rel(ab, abc) < 1e-9


def test_initial_state_gradient_vs_cpu_adjoint():
    print("[gpu-test] adj initial-state gradient", flush=True)
    N, dt, n = 32, 1e-3, 6
    m, X, Ra, _ = case(N, dt, seed=8)
    s = oracle(m, dt)
    s.set_activity(Ra)
    A = O.Adjoint(s)
    states = [X]
    s.set_state(*X)
    for _ in range(n):
        s.step(1, O.IFRK2)
        states.append(s.get_state())
    _, g0 = A.loss_grad(states[-1][0], O.Adjoint.ENSTROPHY)
    g = [g0, np.zeros_like(g0), np.zeros_like(g0)]
    for k in range(n - 1, -1, -1):
        g, _ = A.step_vjp(states[k], g)
    _, _, _, u = gpu_gradient(N, dt, X, Ra, n, 0, Q.ENSTROPHY, want_state=True)
    for a, b in zip(u, g):
        assert rel(a, b) < 1e-9


def test_activity_schedule_gradient_vs_cpu_adjoint():
    print("[gpu-test] adj schedule gradient", flush=True)
    N, dt, n = 32, 1e-3, 8
    m, X, _, rng = case(N, dt, seed=9)
    a_t = rng.uniform(0.8, 1.2, n)
    s = oracle(m, dt)
    s.set_state(*X)
    Lc, _, gtc = O.Adjoint(s).gradient(n, O.Adjoint.KINETIC, alpha_t=a_t)
    L, _, gt, _ = gpu_gradient(N, dt, X, None, n, 3, Q.KINETIC, alpha_t=a_t)
    assert L == pytest.approx(Lc, rel=1e-11)
    assert rel(gt, gtc) < 1e-9


def test_checkpoint_interval_does_not_change_the_gradient():
    """Recomputed segments reproduce the forward: the same loss and gradient
    for checkpoint intervals 1, 2, 4 and none.  A segment restarts from its
    checkpoint through the column prologue, a different kernel instantiation
    from the stage that chains the inverse set in the forward, so the two agree
    to rounding (1e-12 relative), not necessarily bit for bit."""
    print("[gpu-test] adj checkpoint invariance", flush=True)
    N, dt, n = 64, 1e-3, 9
    _, X, Ra, _ = case(N, dt, seed=11)
    ref = gpu_gradient(N, dt, X, Ra, n, 0, Q.ENSTROPHY)
    for k in (1, 2, 4):
        got = gpu_gradient(N, dt, X, Ra, n, k, Q.ENSTROPHY)
        assert got[0] == pytest.approx(ref[0], rel=1e-13)
        assert rel(got[1], ref[1]) < 1e-12, k


@pytest.mark.parametrize("k", [1, 3, 4])
def test_overlapped_recompute_and_repeated_backward(monkeypatch, k):
    """Segments recomputed on a second stream beside the reverse sweep
    (double-buffered stage slots, QS_ADJ_OVERLAP=1) give the gradient of the
    same segments recomputed in turn with the sweep on one stream
    (QS_ADJ_OVERLAP=0) bit for bit: the same kernels on the same data.  A
    second reverse sweep after the same forward (after the first overwrote the
    last segment's recorded stages, when there are more segments than
    buffers) repeats it bit for bit too."""
    print(f"[gpu-test] adj overlap k={k}", flush=True)
    N, dt, n = 64, 1e-3, 11
    _, X, Ra, _ = case(N, dt, seed=13)
    monkeypatch.delenv("QS_ADJ_OVERLAP", raising=False)
    one = gpu_gradient(N, dt, X, Ra, n, 0, Q.ENSTROPHY)
    out = {}
    for overlap in ("0", "1"):
        monkeypatch.setenv("QS_ADJ_OVERLAP", overlap)
        mesh, s = gpu_solver(N, dt)
        s.set_state(*(cu(x) for x in X))
        s.set_activity(cu(Ra))
        s.adj_forward(n, k)
        L, Wb = Q.adj_loss(mesh, s.get_state()[0], Q.ENSTROPHY)
        z = torch.zeros_like(Wb)
        reps = []
        for _ in range(2):
            ab, _, _ = s.adj_backward(Wb, z, z.clone(), want_field=True)
            reps.append(ab.cpu().numpy())
        s.close()
        assert np.array_equal(reps[0], reps[1]), overlap
        assert L == pytest.approx(one[0], rel=1e-13)
        assert rel(reps[0], one[1]) < 1e-12
        out[overlap] = reps[0]
    assert np.array_equal(out["0"], out["1"])


def test_gradient_vs_finite_differences_on_gpu():
    """Central differences of the GPU forward along random activity fields."""
    print("[gpu-test] adj finite differences", flush=True)
    N, dt, n = 64, 1e-3, 40
    _, X, Ra, rng = case(N, dt, seed=12)
    L, ab, _, _ = gpu_gradient(N, dt, X, Ra, n, 10, Q.ENSTROPHY)
    mesh, s = gpu_solver(N, dt)

    def loss(Rv):
        s.set_state(*(cu(x) for x in X))
        s.set_activity(cu(Rv))
        s.step(n)
        return Q.adj_loss(mesh, s.get_state()[0], Q.ENSTROPHY)[0]
    for _ in range(4):
        d = rng.standard_normal((N, N))
        eps = 1e-4
        fd = (loss(Ra + eps * d) - loss(Ra - eps * d)) / (2 * eps)
        an = float(np.sum(ab * d))
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    s.close()


@pytest.mark.parametrize("Nx,Ny", [(64, 32), (32, 128)])
def test_non_square_gradient_vs_cpu_adjoint_and_finite_differences(Nx, Ny):
    """The adjoint on non-square meshes (row kernels sized by Nx, column
    kernels by Ny): the activity-field gradient against the CPU adjoint and
    against central differences of the GPU forward."""
    print(f"[gpu-test] adj non-square {Nx}x{Ny}", flush=True)
    dt, n, k = 1e-3, 8, 3
    rng = np.random.default_rng(Nx + Ny)
    om = O.Mesh(Nx, Ny, 2 * np.pi)
    X = [om.fwd(0.3 * rng.standard_normal((Ny, Nx))), om.fwd(rng.uniform(-0.05, 0.05, (Ny, Nx))),
         om.fwd(rng.uniform(-0.05, 0.05, (Ny, Nx)))]
    Ra = 1.0 + 0.1 * rng.uniform(-1.0, 1.0, (Ny, Nx))
    s = oracle(om, dt)
    s.set_activity(Ra)
    s.set_state(*X)
    Lc, abc, _ = O.Adjoint(s).gradient(n, Q.ENSTROPHY)
    mesh = Q.Mesh(Nx, Ny, 2 * np.pi)
    g = Q.Solver(mesh, Q.params(dt=dt, **P))
    g.set_state(*(cu(x) for x in X))
    g.set_activity(cu(Ra))
    g.adj_forward(n, k)
    L, Wb = Q.adj_loss(mesh, g.get_state()[0], Q.ENSTROPHY)
    z = torch.zeros_like(Wb)
    ab, _, _ = g.adj_backward(Wb, z, z.clone(), want_field=True)
    ab = ab.cpu().numpy()
    assert L == pytest.approx(Lc, rel=1e-11)
    assert rel(ab, abc) < 1e-9

    def loss(Rv):
        g.set_state(*(cu(x) for x in X))
        g.set_activity(cu(Rv))
        g.step(n)
        return Q.adj_loss(mesh, g.get_state()[0], Q.ENSTROPHY)[0]
    for _ in range(2):
        d = rng.standard_normal((Ny, Nx))
        eps = 1e-4
        fd = (loss(Ra + eps * d) - loss(Ra - eps * d)) / (2 * eps)
        an = float(np.sum(ab * d))
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    g.close()
