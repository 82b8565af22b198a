"""Ensembles of uneven runs (qs_ensemble_*): every run equals its own single
run bit for bit, under every launch configuration the autotuner can pick."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import qsolver as Q  # noqa: E402
from qsolver import synth  # noqa: E402

BASE = dict(lam=1.0, Re=1.0, Er=1.0, D_Q=0.05, a_Q=1.0, c_Q=1.0, dt=1e-3)
SIZES = [32, 64, 32, 16, 64, 128, 256, 512]


def run_params(n, seed=3):
    rng = np.random.default_rng(seed)
    return [Q.params(alpha=float(rng.uniform(0.5, 2.0)), Rf=float(rng.uniform(0.01, 0.5)), **BASE)
            for _ in range(n)]


def init_state(N, seed):
    q11, q12 = synth.uniform_q(N, N, seed=seed)
    mesh = Q.Mesh(N, N, 2 * np.pi)
    st = [mesh.spec(), mesh.spec(), mesh.spec()]
    mesh.FwdTrans(st[1], torch.from_numpy(q11).cuda())
    mesh.FwdTrans(st[2], torch.from_numpy(q12).cuda())
    return st


def single_runs(sizes, ps, inits, nsteps, activity=None):
    out = []
    for r, (N, p) in enumerate(zip(sizes, ps)):
        mesh = Q.Mesh(N, N, 2 * np.pi)
        s = Q.Solver(mesh, p)
        s.set_state(*inits[r])
        if activity is not None and r in activity:
            s.set_activity(activity[r])
        s.step(nsteps)
        out.append([t.clone() for t in s.get_state()])
        s.close()
    return out


def ensemble_runs(sizes, ps, inits, nsteps, cfg=None, activity=None, tune=False, return_config=False):
    e = Q.Ensemble(sizes, ps)
    for r in range(len(sizes)):
        e.set_state(r, *inits[r])
    if activity is not None:
        for r, Ra in activity.items():
            e.set_activity(r, Ra)
    if cfg is not None:
        c = e.config()
        for k, v in cfg.items():
            if isinstance(v, list):
                getattr(c, k)[:] = v
            else:
                setattr(c, k, v)
        e.set_config(c)
    if tune:
        best, ms = e.tune(2)
        assert ms > 0
    e.step(nsteps)
    out = [e.get_state(r) for r in range(len(sizes))]
    used = e.config().as_dict()
    e.close()
    return (out, used) if return_config else out


@pytest.fixture(scope="module")
def reference():
    ps = run_params(len(SIZES))
    inits = [init_state(N, 1000 + r) for r, N in enumerate(SIZES)]
    return ps, inits, single_runs(SIZES, ps, inits, 13)


CONFIGS = {
    "default": None,
    "direct-serial": {"graph_steps": 0, "concurrent": 0},
    "graph1-jobmajor": {"graph_steps": 1, "order": 1},
    "few-blocks": {"graph_steps": 4, "row_blocks": [1, 3, 2, 5, 0, 0, 0, 0], "col_blocks": [2, 1, 3, 7, 0, 0, 0, 0]},
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_ensemble_equals_single_runs(reference, name):
    print(f"[gpu-test] ensemble config {name}", flush=True)
    ps, inits, want = reference
    got = ensemble_runs(SIZES, ps, inits, 13, cfg=CONFIGS[name])
    for r, (g, w) in enumerate(zip(got, want)):
        for a, b in zip(g, w):
            assert torch.equal(a, b), (name, r)


def test_autotuner_restores_states(reference):
    """After tuning (which also captures graphs of several depths and may pick
    other FFT lane layouts, or the merged kernels) the runs start from their
    saved states: bit for bit the single runs for the classes that kept their
    default layout in per-class kernels, to rounding (1e-12) the others."""
    print("[gpu-test] ensemble autotune", flush=True)
    ps, inits, want = reference
    got, cfg = ensemble_runs(SIZES, ps, inits, 13, tune=True, return_config=True)
    lanes = dict(zip(sorted(set(SIZES)), cfg["lanes"]))

    def same_kernels(N):   # the single run's kernels (the class's default layout, not merged)
        return not cfg["merged"] and (lanes[N] or default_lanes(N)) == default_lanes(N)
    for N, g, w in zip(SIZES, got, want):
        for a, b in zip(g, w):
            if same_kernels(N):
                assert torch.equal(a, b), (N, cfg)
            else:
                assert rel(a, b) < 1e-12, (N, cfg)


def default_lanes(N):
    return 8 if N == 256 else 16


def rel(a, b):
    return float(torch.linalg.norm(a - b) / torch.linalg.norm(b))


LANES = {   # elements per lane per class (sizes 16, 32, 64, 128, 256, 512)
    "lanes-small": [0, 0, 4, 4, 4, 8, 0, 0],
    "lanes-mid": [0, 0, 8, 8, 16, 0, 0, 0],
}


@pytest.mark.parametrize("name", list(LANES))
def test_ensemble_other_lane_layouts(reference, name):
    """Other per-wave FFT layouts (first-pass radix, rows per row job, column
    tile width) give the single runs' results to rounding."""
    print(f"[gpu-test] ensemble {name}", flush=True)
    ps, inits, want = reference
    got = ensemble_runs(SIZES, ps, inits, 13, cfg={"lanes": LANES[name], "graph_steps": 4})
    for r, (g, w) in enumerate(zip(got, want)):
        for a, b in zip(g, w):
            assert rel(a, b) < 1e-12, (name, r)


def test_merged_kernels(reference):
    """One persistent kernel per phase over every class (16 elements per lane
    at every side) gives the single runs' and the per-class kernels' results to
    rounding (1e-12), with direct launches and through captured graphs, in
    either work-list order.  Not bit for bit: the compiler contracts multiply-
    adds into FMAs differently when one kernel holds the job code of every
    side (with -ffp-contract=on the results are bitwise equal, at 3.6 % of the
    1024^2 step's speed)."""
    print("[gpu-test] ensemble merged kernels", flush=True)
    ps, inits, want = reference
    classes = sorted(set(SIZES))
    per_class = ensemble_runs(SIZES, ps, inits, 13, cfg={"lanes": [16] * len(classes) + [0] * (8 - len(classes)),
                                                         "graph_steps": 4})
    for cfg in ({"merged": 1, "graph_steps": 4},
                {"merged": 1, "graph_steps": 0, "order": 1, "row_blocks": [37] + [0] * 7,
                 "col_blocks": [53] + [0] * 7}):
        got = ensemble_runs(SIZES, ps, inits, 13, cfg=cfg)
        for r, (g, w, pc) in enumerate(zip(got, want, per_class)):
            for a, b, c in zip(g, w, pc):
                assert rel(a, c) < 1e-12, (cfg, r)
                assert rel(a, b) < 1e-12, (cfg, r)


def test_per_run_activity_field(reference):
    print("[gpu-test] ensemble activity field", flush=True)
    ps, inits, _ = reference
    act = {1: torch.from_numpy(synth.activity_field(64, 64, seed=5)).cuda(),
           3: torch.from_numpy(synth.activity_field(16, 16, seed=6)).cuda()}
    want = single_runs(SIZES, ps, inits, 7, activity=act)
    got = ensemble_runs(SIZES, ps, inits, 7, activity=act)
    for g, w in zip(got, want):
        for a, b in zip(g, w):
            assert torch.equal(a, b)


def test_batched_adjoint_equals_single_run_adjoints():
    """Ensemble adjoint (4 runs at 64^2, per-run activity schedules, kinetic
    loss, checkpoints every 5 of 12 steps) equals each run's own qs_adj_*
    gradient bit for bit."""
    print("[gpu-test] ensemble batched adjoint", flush=True)
    N, n, R = 64, 12, 4
    ps = run_params(R, seed=8)
    inits = [init_state(N, 2000 + r) for r in range(R)]
    rng = np.random.default_rng(9)
    a_t = torch.from_numpy(rng.uniform(0.8, 1.2, (R, n))).cuda()
    e = Q.Ensemble([N] * R, ps)
    for r in range(R):
        e.set_state(r, *inits[r])
    e.adj_forward(n, 5, a_t)
    losses = e.adj_loss(Q.KINETIC)
    gt, _ = e.adj_backward(n)
    assert e.adj_memory() > 0
    gt = gt.clone()
    e.close()
    mesh = Q.Mesh(N, N, 2 * np.pi)
    for r in range(R):
        s = Q.Solver(mesh, ps[r])
        s.set_state(*inits[r])
        s.adj_forward(n, 5, a_t[r].contiguous())
        L, Wb = Q.adj_loss(mesh, s.get_state()[0], Q.KINETIC)
        z = torch.zeros_like(Wb)
        _, at, _ = s.adj_backward(Wb, z, z.clone(), want_field=False, want_sched=True, nsteps=n)
        assert L == losses[r]
        assert torch.equal(at, gt[r]), r
        s.close()
