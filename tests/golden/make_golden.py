"""Writes the oracle's golden fixtures for the GPU parity tests.

    python tests/golden/make_golden.py

config1_64.npz: the 64^2 CPU-path workload (L = 2 pi, dt = 1e-3,
lambda = 1, Re = 1, Er = 1, Rf = 0.1, alpha = 1, D_Q = 0.05, a_Q = c_Q = 1;
q11, q12 uniform[-0.01, 0.01] from splitmix64 seed 1, omega = 0), IFRK2:
the initial spectra and the state after 10 and 2000 steps.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "nature-control_amd")]

import oracle_py as O  # noqa: E402
from qsolver import synth  # noqa: E402

CONFIG1 = dict(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, DQ=0.05, aQ=1.0, cQ=1.0, dt=1e-3)


def config1(N=64, steps=(10, 2000)):
    m = O.Mesh(N, N, 2 * np.pi)
    q11, q12 = synth.uniform_q(N, N, seed=1)
    w0 = np.zeros(m.spec_shape, complex)
    r10, r20 = m.fwd(q11), m.fwd(q12)
    s = O.Solver(m, O.params_array(**CONFIG1))
    s.set_state(w0, r10, r20)
    out = dict(w0=w0, r10=r10, r20=r20, params=np.array(list(CONFIG1.values())))
    done = 0
    for n in steps:
        s.step(n - done, O.IFRK2)
        done = n
        w, r1, r2 = s.get_state()
        out[f"w_{n}"], out[f"r1_{n}"], out[f"r2_{n}"] = w, r1, r2
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "config1_64.npz"), **config1())
    print("wrote config1_64.npz")
