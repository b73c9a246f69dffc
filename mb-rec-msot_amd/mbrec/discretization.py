"""Pixel grid, pixel model and prefilter (reconstruction_model_based/models/Discretization.m).

Arrays follow MATLAB memory order with reversed axes: an image stack that
MATLAB holds as [Nz x Nx x W] (column-major, z fastest) is a C-order array of
shape (W, Nx, Nz) here, so the flat pixel index is x = iz + Nz*ix as in the
reference (Discretization.m:34-35).
"""
from __future__ import annotations

import numpy as np


def matlab_round(x):
    """MATLAB round: half away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def gaussian_pixel_model(r, standard_deviation):
    """Discretization.m:112-118."""
    return np.exp(-np.asarray(r) ** 2 / (2.0 * standard_deviation ** 2))


class Discretization:
    def __init__(self, field_of_view, size_of_pixel_grid, custom_pixel_model=None,
                 prefilter_tol: float = 1e-10):
        n = np.atleast_1d(size_of_pixel_grid).astype(int)
        self.sizeOfPixelGrid = (int(n[0]), int(n[0])) if n.size == 1 else (int(n[0]), int(n[1]))
        Nx, Nz = self.sizeOfPixelGrid                     # sizeOfPixelGrid(1) = N_x, (2) = N_z
        fov = list(field_of_view)
        x_vec = np.linspace(fov[0], fov[1], Nx)
        z_vec = np.linspace(fov[0], fov[1], Nz) if len(fov) <= 2 else np.linspace(fov[2], fov[3], Nz)
        X, Z = np.meshgrid(x_vec, z_vec)                  # (Nz, Nx) like MATLAB meshgrid
        self.region = dict(xPositionsFov=X, zPositionsFov=Z)
        self.numberOfPixels = Nz * Nx
        self.Nx, self.Nz = Nx, Nz
        self.dx = X[0, 1] - X[0, 0] if Nx > 1 else 0.0
        self.dz = Z[1, 0] - Z[0, 0] if Nz > 1 else 0.0
        if custom_pixel_model is not None:
            if not callable(custom_pixel_model):
                raise ValueError("Custom pixel model must be function handle.")
            self.pixelModel = custom_pixel_model
            self.pixelSigma = None
        else:
            # Discretization.m:43.  At h = 125 um, h/2 = 62.5 um is a rounding tie that the
            # grid spacing's last bit decides (min(dx, dz) may be 124.99.. um -> 62 um).
            res = min(self.dx, self.dz)
            sigma = matlab_round(res / 2 * 1e6) / 1e6
            self.pixelSigma = sigma
            self.pixelModel = lambda r, s=sigma: gaussian_pixel_model(r, s)
        self.prefilter = self._prefilter()
        self.prefilterConj = np.conj(self.prefilter)
        self.kernel = np.real(np.fft.ifft2(self.prefilter))   # circular kernel K[iz, ix]
        self.kz, self.kx, self.Kz, self.Kx, self.separable_residual = self._separable_taps(prefilter_tol)

    # Discretization.m:74-86 with refinement 1
    def evaluate_pixel_model_on_grid(self, refinement_factor: int = 1):
        Nx, Nz = self.sizeOfPixelGrid
        gN = refinement_factor * (np.array([Nx, Nz]) - 1) + 1
        sp = np.abs([self.dx, self.dz]) / refinement_factor
        lx = sp[0] * np.arange(np.floor(-gN[0] / 2), np.floor(gN[0] / 2 - 1) + 1)
        lz = sp[1] * np.arange(np.floor(-gN[1] / 2), np.floor(gN[1] / 2 - 1) + 1)
        xg, zg = np.meshgrid(lx, lz)
        return self.pixelModel(np.sqrt(xg ** 2 + zg ** 2))

    # Discretization.m:47-51 -> basic_tools/shift_inv_proj.m:66-71 with lattice [1 1]
    def _prefilter(self):
        phi = np.fft.fftshift(self.evaluate_pixel_model_on_grid(1))
        phi_hat = np.fft.fft2(phi)
        autocorr = np.real(np.fft.ifft2(np.abs(phi_hat) ** 2))
        return np.conj(phi_hat / np.fft.fft2(autocorr))

    def _separable_taps(self, tol):
        K = self.kernel
        k00 = K[0, 0]
        if k00 <= 0:
            raise NotImplementedError("prefilter kernel is not separable (non-positive centre tap)")
        kz = K[:, 0] / np.sqrt(k00)
        kx = K[0, :] / np.sqrt(k00)
        resid = np.abs(K - np.outer(kz, kx)).max() / np.abs(K).max()

        def taps(k):
            # offsets -Kh..Kh of the circular kernel, Kh the largest offset with a tap above
            # tol; for an even N the N/2 tap is one sample reached from both sides, so it
            # is kept once (at +N/2, with 0 at -N/2)
            N = k.shape[0]
            kmax = np.abs(k).max()
            Kh = 0
            for a in range(N // 2, -1, -1):
                if max(abs(k[a % N]), abs(k[-a % N])) > tol * kmax:
                    Kh = a
                    break
            t = np.array([k[a % N] for a in range(-Kh, Kh + 1)])
            if Kh > 0 and 2 * Kh == N:
                t[0] = 0.0
            return t, Kh

        tz, Kz = taps(kz)
        tx, Kx = taps(kx)
        return tz, tx, Kz, Kx, resid

    def project_into_shift_invariant_space(self, initial_pressure):
        """Discretization.m:54-62 (reference FFT form, for host-side use)."""
        c = np.asarray(initial_pressure, dtype=np.float64).reshape(-1, self.Nx, self.Nz)
        f = np.fft.fft2(np.swapaxes(c, 1, 2), axes=(1, 2)) * self.prefilter[None]
        return np.swapaxes(np.real(np.fft.ifft2(f, axes=(1, 2))), 1, 2)

    def transpose_project_into_shift_invariant_space(self, x):
        """Discretization.m:64-72."""
        c = np.asarray(x, dtype=np.float64).reshape(-1, self.Nx, self.Nz)
        f = np.fft.fft2(np.swapaxes(c, 1, 2), axes=(1, 2)) * self.prefilterConj[None]
        return np.swapaxes(np.real(np.fft.ifft2(f, axes=(1, 2))), 1, 2)
