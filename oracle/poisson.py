"""Oracle: the electrostatic field of a density map (test infrastructure only).

Two independent forms of the same solve:
  * `field_mirrored_fft` restates src/core/electronic_density_layer.py:62-77,133: the spectral multiplier
    -i k / |k|^2 / bin_x on the rfft2 of the map mirrored in both axes (zero frequency and the Nyquist derivative
    zeroed), back with irfft2; only the [:nbx, :nby] quadrant is used downstream (gp_cuda_kernel.cu:116-117).
  * `field_dct` is the cosine / sine-series form on the unmirrored map:
        C  = Cx rho Cy^T                      Cx[k, n] = cos(pi k (2n + 1) / 2N)
        Ax = 4/(NM) w_k inv (l == 0 ? 1/2 : 1) C,  Ay = 4/(NM) v_l inv (k == 0 ? 1/2 : 1) C
        field_x = Sx^T Ax Cy,   field_y = Cx^T Ay Sy,   inv = 1 / ((w_k^2 + v_l^2) bin_x), inv[0, 0] = 0
    with w_k = pi k / (N bin_x), v_l = pi l / (M bin_y). The tests check the two against each other.
"""
import math

import numpy as np
import scipy.fft as sfft


def to_field(nbx, nby, bx, by, dtype=np.float64):
    """[2, 2 nbx, nby + 1] complex multiplier (electronic_density_layer.py:70-77). dtype float32 follows the
    reference's float32 frequency grids."""
    dt = np.dtype(dtype).type
    kx = (dt(2 * math.pi) * np.fft.fftfreq(2 * nbx, d=bx).astype(dtype))[:, None]
    ky = (dt(2 * math.pi) * np.fft.rfftfreq(2 * nby, d=by).astype(dtype))[None, :]
    inv = dt(1) / np.maximum(kx ** 2 + ky ** 2, dt(1e-30))
    inv[0, 0] = 0
    inv = inv / dt(bx)
    kx = kx.copy()
    ky = ky.copy()
    kx[nbx, 0] = 0
    ky[0, nby] = 0
    ctype = np.complex64 if dtype == np.float32 else np.complex128
    return np.stack([-1j * kx * inv, -1j * ky * inv]).astype(ctype)


def field_mirrored_fft(rho_mirrored, tf):
    """irfft2(rfft2(mirrored) * to_field) -> [2, 2 nbx, 2 nby] (x and y planes)."""
    s = rho_mirrored.shape
    spec = sfft.rfft2(rho_mirrored)
    return sfft.irfft2(spec[None] * tf, s=s, axes=(-2, -1))


def dct_mats(n, dtype=np.float64):
    k = np.arange(n)[:, None]
    j = np.arange(n)[None, :]
    ang = np.pi * k * (2 * j + 1) / (2 * n)
    return np.cos(ang).astype(dtype), np.sin(ang).astype(dtype)


def spectral_weights(nbx, nby, bx, by):
    """-> (wx [N, M], wy [N, M]) with Ax = wx * C and Ay = wy * C, float64."""
    k = np.arange(nbx)[:, None]
    l = np.arange(nby)[None, :]
    w = np.pi * k / (nbx * bx)
    v = np.pi * l / (nby * by)
    den = (w ** 2 + v ** 2) * bx
    with np.errstate(divide="ignore"):
        inv = np.where(den > 0, 1.0 / np.where(den > 0, den, 1.0), 0.0)
    s = 4.0 / (nbx * nby)
    wx = s * w * inv * np.where(l == 0, 0.5, 1.0)
    wy = s * v * inv * np.where(k == 0, 0.5, 1.0)
    return wx, wy


def field_dct(rho, bx, by, dtype=np.float64):
    """Cosine / sine-series field of the unmirrored map rho [N, M] -> (field_x, field_y) [N, M]."""
    nbx, nby = rho.shape
    Cx, Sx = dct_mats(nbx, dtype)
    Cy, Sy = dct_mats(nby, dtype)
    C = Cx @ rho.astype(dtype) @ Cy.T
    wx, wy = spectral_weights(nbx, nby, bx, by)
    Ax = (wx * C).astype(dtype)
    Ay = (wy * C).astype(dtype)
    return Sx.T @ Ax @ Cy, Cx.T @ Ay @ Sy


def energy_weights(nbx, nby, bx, by):
    """w_phi[k, l] = 4/(NM) h_k h_l / ((w_k^2 + v_l^2) bx), h_0 = 1/2, h_k = 1, w_phi[0, 0] = 0: the potential of the
    map is phi = Cx^T (w_phi o C) Cy (in the units of the field)."""
    k = np.arange(nbx)[:, None]
    l = np.arange(nby)[None, :]
    w = np.pi * k / (nbx * bx)
    v = np.pi * l / (nby * by)
    den = (w ** 2 + v ** 2) * bx
    with np.errstate(divide="ignore"):
        inv = np.where(den > 0, 1.0 / np.where(den > 0, den, 1.0), 0.0)
    return 4.0 / (nbx * nby) * np.where(k == 0, 0.5, 1.0) * np.where(l == 0, 0.5, 1.0) * inv


def density_energy(rho, bx, by):
    """E = 1/2 sum_b rho_b phi_b = 1/2 sum_kl w_phi C_kl^2 (C = Cx rho Cy^T), float64. Its gradient with respect to a
    node's position is the field gather of the density kernels up to the bin discretisation."""
    nbx, nby = rho.shape
    Cx, _ = dct_mats(nbx)
    Cy, _ = dct_mats(nby)
    C = Cx @ np.asarray(rho, np.float64) @ Cy.T
    return 0.5 * float((energy_weights(nbx, nby, bx, by) * C * C).sum())
