"""torch.autograd.Function wrappers (qsolver.torch_ops) over the C-ABI adjoint,
including the 256^2 / 200-step activity-field gradient checked by central
finite differences along 16 random directions."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import qsolver as Q  # noqa: E402
from qsolver import synth, torch_ops as T  # noqa: E402

P = dict(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, D_Q=0.05, a_Q=1.0, c_Q=1.0)


def initial_state(mesh, N, seed=2):
    q11, q12 = synth.bandlimited_q(N, N, seed=seed)
    st = [mesh.spec(), mesh.spec(), mesh.spec()]
    mesh.FwdTrans(st[1], torch.from_numpy(q11).cuda())
    mesh.FwdTrans(st[2], torch.from_numpy(q12).cuda())
    return st


def test_autograd_matches_direct_c_abi():
    print("[gpu-test] torch adjoint vs C-ABI", flush=True)
    N, n, dt = 64, 12, 1e-3
    mesh = Q.Mesh(N, N, 2 * np.pi)
    s = Q.Solver(mesh, Q.params(dt=dt, **P))
    W0, R10, R20 = initial_state(mesh, N)
    alpha = torch.from_numpy(synth.activity_field(N, N, seed=7)).cuda().requires_grad_(True)
    W, R1, R2 = T.rollout(s, alpha, W0, R10, R20, n, 5)
    L = T.spectral_loss(mesh, W, Q.ENSTROPHY)
    L.backward()
    # the same through the C-ABI directly
    s2 = Q.Solver(mesh, Q.params(dt=dt, **P))
    s2.set_state(W0, R10, R20)
    s2.set_activity(alpha.detach())
    s2.adj_forward(n, 5)
    L2, Wb = Q.adj_loss(mesh, s2.get_state()[0], Q.ENSTROPHY)
    z = torch.zeros_like(Wb)
    ab, _, _ = s2.adj_backward(Wb, z, z.clone())
    assert float(L) == L2
    assert torch.equal(alpha.grad, ab)


def test_schedule_control_and_state_gradients():
    """[nsteps] activity schedule, kinetic-energy loss; gradients flow to the
    schedule and to the initial spectra (checked by finite differences)."""
    print("[gpu-test] torch adjoint schedule", flush=True)
    N, n, dt = 32, 10, 1e-3
    mesh = Q.Mesh(N, N, 2 * np.pi)
    s = Q.Solver(mesh, Q.params(dt=dt, **P))
    W0, R10, R20 = initial_state(mesh, N, seed=4)
    rng = np.random.default_rng(3)
    a_t = torch.from_numpy(rng.uniform(0.8, 1.2, n)).cuda().requires_grad_(True)
    R10 = R10.clone().requires_grad_(True)

    def loss(a, r1):
        W, _, _ = T.rollout(s, a, W0, r1, R20, n, 4)
        return T.spectral_loss(mesh, W, Q.KINETIC)
    L = loss(a_t, R10)
    L.backward()
    ga, gr = a_t.grad.clone(), R10.grad.clone()
    with torch.no_grad():
        for k in (0, n - 1):
            e = torch.zeros_like(a_t)
            e[k] = 1e-3
            fd = (loss(a_t + e, R10) - loss(a_t - e, R10)) / 2e-3
            assert abs(float(fd) - float(ga[k])) <= 1e-6 * abs(float(ga[k]))
        d = torch.from_numpy(rng.standard_normal(R10.shape) + 1j * rng.standard_normal(R10.shape)).cuda()
        d *= float(R10.abs().max()) * 1e-6   # small: the loss is far from linear in every mode of R1
        fd = (loss(a_t, R10 + d) - loss(a_t, R10 - d)) / 2
        an = float((gr.real * d.real + gr.imag * d.imag).sum())
        assert abs(float(fd) - an) <= 1e-6 * abs(an)


def test_activity_field_gradient_256_200_steps_finite_differences():
    """256^2, dt = 1e-3, 200 steps, alpha(x) = 1 + 0.1 eta (|k| <= 6, seed 7),
    L = 1/2 mean(omega^2) at the final step, checkpoints every 20 steps;
    central differences with eps = 1e-6 along 16 i.i.d. normal directions."""
    print("[gpu-test] torch adjoint 256^2 x 200 FD", flush=True)
    N, n, dt = 256, 200, 1e-3
    mesh = Q.Mesh(N, N, 2 * np.pi)
    s = Q.Solver(mesh, Q.params(dt=dt, **P))
    W0, R10, R20 = initial_state(mesh, N, seed=2)
    alpha = torch.from_numpy(synth.activity_field(N, N, seed=7)).cuda().requires_grad_(True)
    W, _, _ = T.rollout(s, alpha, W0, R10, R20, n, 20)
    L = T.spectral_loss(mesh, W, Q.ENSTROPHY)
    L.backward()
    g = alpha.grad.clone()
    rng = np.random.default_rng(16)
    errs = []
    with torch.no_grad():
        for _ in range(16):
            d = torch.from_numpy(rng.standard_normal((N, N))).cuda()
            eps = 1e-6
            Lp = T.spectral_loss(mesh, T.rollout(s, alpha + eps * d, W0, R10, R20, n, 0)[0], Q.ENSTROPHY)
            Lm = T.spectral_loss(mesh, T.rollout(s, alpha - eps * d, W0, R10, R20, n, 0)[0], Q.ENSTROPHY)
            fd = float(Lp - Lm) / (2 * eps)
            an = float((g * d).sum())
            errs.append(abs(fd - an) / abs(an))
    print("[gpu-test] FD relative errors: max %.2e median %.2e" % (max(errs), float(np.median(errs))), flush=True)
    assert max(errs) <= 1e-6
