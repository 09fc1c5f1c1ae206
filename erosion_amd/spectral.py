"""Spectral routines on periodic grids (host side).

Restates ErosionPackage/src/SpectralMethods.jl (kvec :13-22, specdiff :35-41,
specint :45-54, expsmooth :57-64, gaussfilter :67) with numpy's FFT standing in
for FFTW.jl.  These run on the host inside the time stepper, between solver
calls; they are O(n log n) per body and never on the Stokes-solve hot path.
"""
from __future__ import annotations

import numpy as np


def kvec(nn: int, hmode: int = 0) -> np.ndarray:
    """Wavenumbers in fftshift order; for even ``nn`` the Nyquist slot gets
    ``hmode * nn/2`` (SpectralMethods.jl:13-22)."""
    if nn % 2 == 0:
        nsm = nn // 2 - 1
        return np.concatenate([[hmode * (nsm + 1)], np.arange(-nsm, nsm + 1)]).astype(np.float64)
    nsm = (nn - 1) // 2
    return np.arange(-nsm, nsm + 1, dtype=np.float64)


def _fftnice(fx: np.ndarray) -> np.ndarray:
    return np.fft.fftshift(np.fft.fft(fx))


def _ifftnice(fh: np.ndarray) -> np.ndarray:
    return np.fft.ifft(np.fft.ifftshift(fh))


def specdiff(fx: np.ndarray, intvlen: float = 2 * np.pi) -> np.ndarray:
    """Spectral derivative, Nyquist mode dropped (SpectralMethods.jl:35-41),
    along the last axis."""
    fx = np.asarray(fx, dtype=np.float64)
    fh = np.fft.fftshift(np.fft.fft(fx, axis=-1), axes=-1)
    kv = kvec(fx.shape[-1], 0)
    df = np.fft.ifft(np.fft.ifftshift(2j * np.pi * kv * fh, axes=-1), axis=-1)
    return np.real(df) / intvlen


def specint(fx: np.ndarray, intvlen: float = 2 * np.pi) -> np.ndarray:
    """Mean-free spectral antiderivative (SpectralMethods.jl:45-54), along the
    last axis (a 2-D array is a batch of periodic functions)."""
    fx = np.asarray(fx, dtype=np.float64)
    fh = np.fft.fftshift(np.fft.fft(fx, axis=-1), axes=-1)
    kv = kvec(fx.shape[-1], 0)
    Fh = np.zeros_like(fh)
    nz = kv != 0
    Fh[..., nz] = fh[..., nz] / (2j * np.pi * kv[nz])
    Fx = np.fft.ifft(np.fft.ifftshift(Fh, axes=-1), axis=-1)
    return np.real(Fx) * intvlen


def expsmooth(fx: np.ndarray, factor) -> np.ndarray:
    """Multiply Fourier mode k by exp(-factor k^2) (SpectralMethods.jl:57-64),
    along the last axis; ``factor`` may hold one value per row."""
    fx = np.asarray(fx, dtype=np.float64)
    fh = np.fft.fftshift(np.fft.fft(fx, axis=-1), axes=-1)
    kv = kvec(fx.shape[-1], 1)
    fac = np.asarray(factor, dtype=np.float64)
    if fac.ndim:
        fac = fac[..., None]
    fh = fh * np.exp(-fac * np.abs(kv) ** 2)
    return np.real(np.fft.ifft(np.fft.ifftshift(fh, axes=-1), axis=-1))


def gaussfilter(fx: np.ndarray, sigma: float) -> np.ndarray:
    """Gaussian filter of width sigma (SpectralMethods.jl:67)."""
    return expsmooth(fx, 0.5 * sigma ** 2)
