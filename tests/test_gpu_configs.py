"""The ensemble and batched-adjoint workloads at their own sizes, against the
oracle (not only against the GPU's single-run path).

* Uneven ensemble: 64 runs at 128^2, 128 at 256^2 and 16 at 512^2, FP64
  IFRK2, dt = 1e-3, per-run activity ~ U[0.5, 2] and friction ~ U[0.01, 0.5]
  (numpy seed 3), run r: uniform Q in [-0.01, 0.01] (splitmix64 seed 1000 + r),
  omega = 0.  Every run equals its own single run bit for bit; a seeded sample
  that includes every size class matches the oracle (friction.cu:212-260 with
  IFRK2.cuh:16-51) to 1e-12.
* Batched adjoint at 512^2: envs with a per-step activity schedule
  a_e(t_n) ~ U[0.8, 1.2], L = sum_e 1/2 mean(u^2 + v^2) of the final state,
  checkpoints: each env's gradient dL/da_e against the oracle's CPU adjoint
  (1e-9 relative) and central finite differences of the GPU forward (1e-6).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import oracle_py as O  # noqa: E402
import qsolver as Q  # noqa: E402
from qsolver import synth  # noqa: E402

BASE = dict(lam=1.0, Re=1.0, Er=1.0, D_Q=0.05, a_Q=1.0, c_Q=1.0)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def oparams(p):
    return O.params_array(lam=p.lam, Re=p.Re, Er=p.Er, Rf=p.Rf, alpha=p.alpha, DQ=p.D_Q, aQ=p.a_Q, cQ=p.c_Q, dt=p.dt)


def uniform_state(om, N, seed):
    q11, q12 = synth.uniform_q(N, N, seed=seed)
    return [np.zeros(om.spec_shape, complex), om.fwd(q11), om.fwd(q12)]


def test_uneven_ensemble_full_workload():
    sizes = [128] * 64 + [256] * 128 + [512] * 16
    nsteps = 4
    rng = np.random.default_rng(3)
    ps = [Q.params(alpha=float(rng.uniform(0.5, 2.0)), Rf=float(rng.uniform(0.01, 0.5)), dt=1e-3, **BASE)
          for _ in sizes]
    oms = {N: O.Mesh(N, N, 2 * np.pi) for N in (128, 256, 512)}
    inits = [uniform_state(oms[N], N, 1000 + r) for r, N in enumerate(sizes)]
    print("[gpu-test] config-3 ensemble, 208 runs", flush=True)
    e = Q.Ensemble(sizes, ps)
    for r in range(len(sizes)):
        e.set_state(r, *(dev(a) for a in inits[r]))
    e.step(nsteps)
    got = [[t.cpu().numpy() for t in e.get_state(r)] for r in range(len(sizes))]
    e.close()
    print("[gpu-test] config-3 single runs", flush=True)
    meshes = {N: Q.Mesh(N, N, 2 * np.pi) for N in (128, 256, 512)}
    for r, N in enumerate(sizes):
        s = Q.Solver(meshes[N], ps[r])
        s.set_state(*(dev(a) for a in inits[r]))
        s.step(nsteps)
        single = [t.cpu().numpy() for t in s.get_state()]
        assert not s.nonfinite()
        s.close()
        for a, b in zip(got[r], single):
            assert np.array_equal(a, b), r
    print("[gpu-test] config-3 oracle sample", flush=True)
    for r in (0, 41, 64, 150, 191, 192, 207):     # 128^2 x2, 256^2 x3, 512^2 x2
        N = sizes[r]
        s = O.Solver(oms[N], oparams(ps[r]))
        s.set_state(*inits[r])
        s.step(nsteps, O.IFRK2)
        for a, b in zip(got[r], s.get_state()):
            assert rel(a, b) < 1e-12, r


def test_batched_adjoint_512_vs_cpu_adjoint_and_finite_differences():
    N, nenv, n, k, dt = 512, 2, 6, 2, 5e-4
    rng = np.random.default_rng(5)
    ps = [Q.params(alpha=1.0, Rf=0.1, dt=dt, **BASE) for _ in range(nenv)]
    om = O.Mesh(N, N, 2 * np.pi)
    inits = [uniform_state(om, N, 100 + r) for r in range(nenv)]
    a_t = rng.uniform(0.8, 1.2, (nenv, n))
    print("[gpu-test] config-5 batched adjoint 2 x 512^2", flush=True)
    e = Q.Ensemble([N] * nenv, ps)

    def forward(at):
        for r in range(nenv):
            e.set_state(r, *(dev(a) for a in inits[r]))
        e.adj_forward(n, k, dev(at))
        return e.adj_loss(Q.KINETIC)
    losses = forward(a_t)
    gt, _ = e.adj_backward(n)
    g = gt.cpu().numpy()
    assert np.all(np.isfinite(g)) and np.all(losses > 0)
    print("[gpu-test] config-5 CPU adjoint", flush=True)
    for r in range(nenv):
        s = O.Solver(om, oparams(ps[r]))
        s.set_state(*inits[r])
        L, _, abt = O.Adjoint(s).gradient(n, O.Adjoint.KINETIC, alpha_t=a_t[r])
        assert abs(L - losses[r]) <= 1e-12 * L
        assert rel(g[r], abt) < 1e-9, r
    print("[gpu-test] config-5 finite differences", flush=True)
    eps = 1e-6
    drng = np.random.default_rng(17)
    for _ in range(2):
        d = drng.standard_normal((nenv, n))
        fd = (forward(a_t + eps * d).sum() - forward(a_t - eps * d).sum()) / (2 * eps)
        an = float(np.sum(g * d))
        assert abs(fd - an) <= 1e-6 * abs(an)
    e.close()
