"""Writes checksum fixtures of the oracle's 1024^2 IFRK2 state.

    python tests/golden/make_golden_1024.py

The full 1024^2 state (3 x 8.4 MB) is too large to commit, so the fixture
keeps, per field: the L2 norm, the block of modes with |kx|, |ky| <= 24, and
2048 spectral coefficients at seeded random positions across the whole half
spectrum.  Workload: the 1024^2 benchmark (L = 2 pi, lambda = 1, Re = Er = 1,
Rf = 0.1, alpha = 1, D_Q = 0.05, a_Q = c_Q = 1; band-limited Q from seed 2,
rms 0.05; omega = 0):
  * 10 IFRK2 steps at dt = 5e-4 (the scheme is unstable at this dt on this
    mesh: roundoff differences grow ~50x per 10 steps and the state turns
    non-finite near step 75, so only a short horizon is comparable at 1e-12);
  * 100 IFRK2 steps at dt = 1e-4 (stable).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "nature-control_amd")]

import oracle_py as O  # noqa: E402
from qsolver import synth  # noqa: E402

N = 1024
CFG = dict(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, DQ=0.05, aQ=1.0, cQ=1.0)
KB = 24


def sample_index(Ny, Nxh, n=2048, seed=11):
    rng = np.random.default_rng(seed)
    return rng.integers(0, Ny, n), rng.integers(0, Nxh, n)


def checksum(a):
    Ny, Nxh = a.shape
    rows = np.r_[0:KB + 1, Ny - KB:Ny]
    jj, ii = sample_index(Ny, Nxh)
    return dict(norm=np.linalg.norm(a), block=a[np.ix_(rows, np.arange(KB + 1))], sample=a[jj, ii])


def run(dt, nsteps):
    O.set_threads(os.cpu_count() or 1)
    m = O.Mesh(N, N, 2 * np.pi)
    q11, q12 = synth.bandlimited_q(N, N, seed=2)
    s = O.Solver(m, O.params_array(dt=dt, **CFG))
    s.set_state(np.zeros(m.spec_shape, complex), m.fwd(q11), m.fwd(q12))
    s.step(nsteps, O.IFRK2)
    return s.get_state()


if __name__ == "__main__":
    out = {}
    for tag, dt, n in (("dt5e-4_10", 5e-4, 10), ("dt1e-4_100", 1e-4, 100)):
        for name, a in zip(("w", "r1", "r2"), run(dt, n)):
            for k, v in checksum(a).items():
                out[f"{tag}_{name}_{k}"] = v
        print("done", tag, flush=True)
    np.savez_compressed(os.path.join(HERE, "ifrk2_1024.npz"), **out)
    print("wrote ifrk2_1024.npz")
