"""Known-answer cases taken from the reference's own unit tests.

Each nonlinear case gives the initial physical fields (omega, r1, r2) and the
closed-form nonlinear terms (nl0 = N_omega, nl1 = N_r1, nl2 = N_r2) written in
the reference drivers.  Parameters of every case: lambda=0.1, Rf=7.5e-5,
alpha=0.2 (uniform Ra), Re=Er=0.1, L=2*pi (UnitTest/Nonlinear/Qsolver.cu:79-94).

  case 0: UnitTest/Nonlinear/Qsolver.cu:2-77   (== nlterms0.cu)
  case 1: UnitTest/Nonlinear/nlterms1.cu:2-77
  case 2: UnitTest/Nonlinear/nlterms2.cu:5-84
  case 3: UnitTest/Nonlinear/nlterms3.cu:6-86  (nl1 uses the 6-digit constant
          0.0342282 for 10.2/298, so it is exact only to ~1.2e-8 relative)
  case 4: UnitTest/Nonlinear/nlterms4.cu:7-92

The reference evaluates them at N=512; every field here is a trigonometric
polynomial of low degree, so any N whose Nyquist exceeds the highest product
wavenumber gives the same values to roundoff (N_MIN below).
"""
import numpy as np

KAT_PARAMS = dict(lam=0.1, Rf=0.000075, alpha=0.2, Re=0.1, Er=0.1)

# Smallest power-of-two mesh that resolves all products of each case.
N_MIN = {0: 32, 1: 32, 2: 32, 3: 128, 4: 32}
# Relative max-norm tolerance for (nl0, nl1, nl2).  nl0 carries the spectral
# roundoff floor amplified by k^4/(Er*Re) (PiDeriv after Laplacians); case 3's
# nl1 is limited by the reference's rounded constant.
KAT_TOL = {0: (1e-9, 1e-12, 1e-12), 1: (1e-9, 1e-12, 1e-12), 2: (1e-9, 1e-12, 1e-12),
           3: (1e-8, 2e-8, 1e-12), 4: (1e-9, 1e-12, 1e-12)}


def nonlinear_case(c, X, Y):
    s, co = np.sin, np.cos
    x, y = X, Y
    if c == 0:
        w, r1, r2 = -2 * s(x + y), s(x), co(x)
        nl0 = -60 * co(x)
        nl1 = -4 * s(x) + 0.2 * s(x + y) + co(x) * (co(x + y) + 2 * s(x + y))
        nl2 = -4 * co(x) - s(x) * (co(x + y) + 2 * s(x + y))
    elif c == 1:
        w, r1, r2 = -2 * s(x + y), co(x + y), s(x + y)
        nl0 = 160 * co(x + y)
        nl1 = 1 - 4 * co(x + y) - co(2 * (x + y)) + 0.2 * s(x + y)
        nl2 = s(x + y) * (-4 - 2 * co(x + y))
    elif c == 2:
        w, r1, r2 = -5 * s(2 * x + y), co(x + y), s(x + y)
        nl0 = 160 * co(x + y)
        nl1 = (-4 * co(x) * co(y) + 4 * s(x) * s(y) + co(2 * x + y) * s(x + y)
               + 0.4 * s(2 * x + y) + 5 * s(x + y) * s(2 * x + y))
        nl2 = (-4 * co(y) * s(x) - 4 * co(x) * s(y) - 0.3 * s(2 * x + y)
               - co(x + y) * (co(2 * x + y) + 5 * s(2 * x + y)))
    elif c == 3:
        w, r1, r2 = -298 * s(17 * x + 3 * y), co(5 * x + y), s(5 * x + y)
        nl0 = 5600 * co(5 * x + y) - 13440 * s(5 * x + y)
        nl1 = (-4 * co(5 * x + y) + 2 * co(17 * x + 3 * y) * s(5 * x + y)
               + 298 * (0.0342282 + s(5 * x + y)) * s(17 * x + 3 * y))
        nl2 = (-4 * s(5 * x + y) - 28 * s(17 * x + 3 * y)
               - 2 * co(5 * x + y) * (co(17 * x + 3 * y) + 149 * s(17 * x + 3 * y)))
    elif c == 4:
        w, r1, r2 = -13 * s(2 * x + 3 * y), co(3 * x + 2 * y), s(3 * x + 2 * y)
        nl0 = 3600 * co(3 * x + 2 * y) - 1500 * s(3 * x + 2 * y)
        nl1 = (6.5 * co(x - y) - 6.5 * co(5 * (x + y)) - 4.0 * co(3 * x + 2 * y)
               - 2.5 * s(x - y) - 2.5 * s(5 * (x + y)) + 1.2 * s(2 * x + 3 * y))
        nl2 = (2.5 * co(x - y) + 2.5 * co(5 * (x + y)) + 6.5 * s(x - y)
               - 6.5 * s(5 * (x + y)) - 4.0 * s(3 * x + 2 * y) + 0.5 * s(2 * x + 3 * y))
    else:
        raise ValueError(c)
    return (w, r1, r2), (nl0, nl1, nl2)


def fld_case(X, Y):
    """UnitTest/Fld/Fld.cu:21-63: phi = sin(x+y) and its exact derivatives.

    Returns phi, laplacian (w = -2 sin), d/dx (v = cos), -d/dy (u = -cos),
    and the velocity of w by vel_func (u = -cos(x+y), v = cos(x+y))."""
    phi = np.sin(X + Y)
    return dict(phi=phi, lap=-2.0 * np.sin(X + Y), dx=np.cos(X + Y), mdy=-np.cos(X + Y),
                velu=-np.cos(X + Y), velv=np.cos(X + Y))


def splitmix64_uniform(seed, n):
    """Uniform [0,1) doubles from splitmix64(seed): u = (x >> 11) * 2^-53.

    The same generator is implemented in the library (C++) so that the
    synthetic inputs of the GPU path and of the oracle are identical."""
    out = np.empty(n)
    mask = (1 << 64) - 1
    state = seed & mask
    for k in range(n):
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        z = z ^ (z >> 31)
        out[k] = (z >> 11) * (1.0 / 9007199254740992.0)
    return out
