"""numpy model of the fused kernels' intermediate arrays (test infrastructure).

Restates, with numpy FFTs, what each fused kernel of the library computes, so
the GPU kernels can be checked one at a time (the oracle checks the whole
step).  Arrays use the reference spectral layout [Ny, Nx//2+1].

  prologue / column stages -> inverse set M0..M6 (y-inverse FFT, unnormalised)
  row pass                 -> forward set F0..F6 (x R2C of the products)
"""
import numpy as np


def tables(Nx, Ny, Lx=2 * np.pi, Ly=None):
    Ly = Lx if Ly is None else Ly
    i = np.arange(Nx // 2 + 1)
    j = np.arange(Ny)
    kx = 2 * np.pi / Lx * i
    ky = np.where(j < Ny // 2 + 1, 2 * np.pi / Ly * j, 2 * np.pi / Ly * (j - Ny))
    return kx[None, :], ky[:, None]


def inverse_set(W, R1, R2, kx, ky):
    k2 = kx**2 + ky**2
    k2s = np.where(k2 == 0, 1.0, k2)
    psi = np.where(k2 == 0, 0.0, W / k2s)
    spec = [W, psi, 1j * ky * psi, R1, -ky**2 * R1, R2, -ky**2 * R2]
    Ny = W.shape[0]
    return [Ny * np.fft.ifft(a, axis=0) for a in spec]


def c2r_rows(X, Nx, Ny):
    """x-direction C2R of mixed rows (imag parts of DC/Nyquist discarded) with
    the 1/(Nx Ny) normalisation of BwdTrans."""
    return np.fft.irfft(X, n=Nx, axis=1) / Ny


def row_pass(M, W00, p, Nx, Ny, kx, Ra=None):
    D, aQ, cQ, lam = p["D_Q"], p["a_Q"], p["c_Q"], p["lam"]
    sE = 1.0 / (p["Er"] * p["Re"])
    al = p["alpha"] if Ra is None else Ra
    kxr = kx[0]
    w = c2r_rows(M[0], Nx, Ny)
    r1 = c2r_rows(M[3], Nx, Ny)
    r2 = c2r_rows(M[5], Nx, Ny)
    u = c2r_rows(M[2], Nx, Ny)
    v = c2r_rows(-1j * kxr * M[1], Nx, Ny)
    uy = -M[0] + 2 * kxr**2 * M[1]
    uy[:, 0] += W00.real
    uyvx = c2r_rows(uy, Nx, Ny)
    L1 = c2r_rows(-kxr**2 * M[3] + M[4], Nx, Ny)
    L2 = c2r_rows(-kxr**2 * M[5] + M[6], Nx, Ny)
    ux = c2r_rows(1j * kxr * M[2], Nx, Ny)
    S = 2 * np.sqrt(r1**2 + r2**2)
    cS2 = cQ * S * S
    g = lam * S * (cS2 - aQ)
    s1 = -lam * D * L1 * S + g * r1
    s2 = -lam * D * L2 * S + g * r2
    cr = 2 * D * (L1 * r2 - r1 * L2)
    p11, p12, p21 = s1 - al * r1, s2 + cr - al * r2, s2 - cr - al * r2
    advect = p.get("advect_w", 1)
    A = -u * w if advect else 0 * w
    B = -v * w if advect else 0 * w
    loc1 = -w * r2 - cS2 * r1 + lam * S * ux
    loc2 = w * r1 - cS2 * r2 + 0.5 * lam * S * uyvx
    R = lambda f: np.fft.rfft(f, axis=1)
    kxa = kxr.copy()
    kxa[-1] = 0.0
    F0 = 1j * kxa * R(A) - kxr**2 * sE * R(p12)
    F1 = R(B) - 2j * kxr * sE * R(p11)
    F2 = sE * R(p21)
    F3 = R(loc1) + 1j * kxr * R(-u * r1)
    F4 = R(-v * r1)
    F5 = R(loc2) + 1j * kxr * R(-u * r2)
    F6 = R(-v * r2)
    return [F0, F1, F2, F3, F4, F5, F6]


def nonlinear_from_forward_set(F, kx, ky, Nx, Ny):
    Fy = [np.fft.fft(a, axis=0) for a in F]
    Nw = Fy[0] + 1j * ky * Fy[1] + ky**2 * Fy[2]
    for j in (0, Ny // 2):
        for i in (0, Nx // 2):
            Nw[j, i] = Nw[j, i].real
    return Nw, Fy[3] + 1j * ky * Fy[4], Fy[5] + 1j * ky * Fy[6]
