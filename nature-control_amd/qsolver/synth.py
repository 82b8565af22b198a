"""Synthetic initial conditions of the benchmark workloads (host side).

Random numbers come from the library's splitmix64 stream (qs_synth_uniform),
consumed field by field in row-major (j, i) order, so the GPU path and the
CPU oracle start from identical inputs.
"""
import numpy as np

from . import synth_uniform


def uniform_q(Nx, Ny, seed, lo=-0.01, hi=0.01):
    """q11, q12 i.i.d. uniform[lo, hi] (physical, [Ny, Nx] each)."""
    u = synth_uniform(seed, lo, hi, 2 * Nx * Ny)
    return u[: Nx * Ny].reshape(Ny, Nx).copy(), u[Nx * Ny:].reshape(Ny, Nx).copy()


def _int_wavenumbers(Nx, Ny):
    kx = np.arange(Nx // 2 + 1, dtype=np.float64)[None, :]
    j = np.arange(Ny)
    ky = np.where(j < Ny // 2 + 1, j, j - Ny).astype(np.float64)[:, None]
    return kx, ky


def bandlimited_field(Nx, Ny, u, rms, kscale=32.0):
    """Half-spectrum coefficients exp(-|k|^2/kscale) e^{2 pi i u} with integer
    wavenumbers, (0,0) = 0, to physical space by the C2R that discards the
    imaginary parts of the x DC/Nyquist bins (numpy's irfft2 does), scaled to
    the requested real-space rms."""
    kx, ky = _int_wavenumbers(Nx, Ny)
    F = np.exp(-(kx**2 + ky**2) / kscale) * np.exp(2j * np.pi * u.reshape(Ny, Nx // 2 + 1))
    F[0, 0] = 0.0
    f = np.fft.irfft2(F, s=(Ny, Nx))
    return f * (rms / np.sqrt(np.mean(f**2)))


def bandlimited_q(Nx, Ny, seed, rms=0.05, kscale=32.0):
    n = Ny * (Nx // 2 + 1)
    u = synth_uniform(seed, 0.0, 1.0, 2 * n)
    return (bandlimited_field(Nx, Ny, u[:n], rms, kscale),
            bandlimited_field(Nx, Ny, u[n:], rms, kscale))


def activity_field(Nx, Ny, seed, base=1.0, amp=0.1, kmax=6.0):
    """alpha(x) = base + amp * eta, eta a unit-rms sum of unit-amplitude
    random-phase modes with 0 < |k| <= kmax (integer wavenumbers)."""
    kx, ky = _int_wavenumbers(Nx, Ny)
    n = Ny * (Nx // 2 + 1)
    u = synth_uniform(seed, 0.0, 1.0, n).reshape(Ny, Nx // 2 + 1)
    kk = kx**2 + ky**2
    F = np.where((kk > 0) & (kk <= kmax * kmax), np.exp(2j * np.pi * u), 0.0)
    eta = np.fft.irfft2(F, s=(Ny, Nx))
    eta /= np.sqrt(np.mean(eta**2))
    return base + amp * eta
