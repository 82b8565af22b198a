"""The committed golden fixtures still match the oracle as built now (its
compiler flags or code could drift from the build that wrote them):
config1_64.npz (64^2, 10 and 2000 IFRK2 steps, tests/golden/make_golden.py)
and the 10-step part of ifrk2_1024.npz (tests/golden/make_golden_1024.py)."""
import os

import numpy as np

import oracle_py as O
from qsolver import synth

HERE = os.path.dirname(os.path.abspath(__file__))
KB = 24


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_config1_fixture():
    g = np.load(os.path.join(HERE, "golden", "config1_64.npz"))
    p = g["params"]
    m = O.Mesh(64, 64, 2 * np.pi)
    q11, q12 = synth.uniform_q(64, 64, seed=1)
    assert np.array_equal(m.fwd(q11), g["r10"]) and np.array_equal(m.fwd(q12), g["r20"])
    s = O.Solver(m, O.params_array(lam=p[0], Re=p[1], Er=p[2], Rf=p[3], alpha=p[4], DQ=p[5], aQ=p[6], cQ=p[7],
                                   dt=p[8]))
    s.set_state(g["w0"], g["r10"], g["r20"])
    s.step(10, O.IFRK2)
    for a, k in zip(s.get_state(), ("w_10", "r1_10", "r2_10")):
        assert rel(a, g[k]) < 1e-13
    s.step(1990, O.IFRK2)
    for a, k in zip(s.get_state(), ("w_2000", "r1_2000", "r2_2000")):
        assert rel(a, g[k]) < 1e-11


def test_1024_fixture_10_steps():
    g = np.load(os.path.join(HERE, "golden", "ifrk2_1024.npz"))
    N = 1024
    m = O.Mesh(N, N, 2 * np.pi)
    q11, q12 = synth.bandlimited_q(N, N, seed=2)
    s = O.Solver(m, O.params_array(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, DQ=0.05, aQ=1.0, cQ=1.0, dt=5e-4))
    s.set_state(np.zeros(m.spec_shape, complex), m.fwd(q11), m.fwd(q12))
    s.step(10, O.IFRK2)
    rows = np.r_[0:KB + 1, N - KB:N]
    rng = np.random.default_rng(11)
    jj, ii = rng.integers(0, N, 2048), rng.integers(0, N // 2 + 1, 2048)
    for name, a in zip(("w", "r1", "r2"), s.get_state()):
        assert abs(np.linalg.norm(a) - g[f"dt5e-4_10_{name}_norm"]) <= 1e-12 * g[f"dt5e-4_10_{name}_norm"]
        assert rel(a[np.ix_(rows, np.arange(KB + 1))], g[f"dt5e-4_10_{name}_block"]) < 1e-12
        assert rel(a[jj, ii], g[f"dt5e-4_10_{name}_sample"]) < 1e-12
