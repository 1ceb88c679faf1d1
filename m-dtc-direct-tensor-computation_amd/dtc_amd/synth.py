"""Seeded synthetic control volumes (there is no network for real uCT data).

  solid_cube(n)            fully solid block (analytic isotropic C)
  trabecular(n, ...)       white noise -> periodic Gaussian smoothing -> threshold
                           at the requested BV/TV -> largest 26-connected component
  greyscale(n, ...)        the smoothed field mapped to 0..255 with the void set to 0,
                           and the per-voxel modulus E0 * (g/255)^2
  uneven_batch(count, ...) domains with edge ~ U{32..96}, BV/TV ~ U[0.08, 0.35],
                           anisotropic smoothing sigma (s, s, 2s) in a random frame
"""
from __future__ import annotations

import numpy as np


def _smooth(n, rng, sigma, frame=None):
    f = rng.standard_normal((n, n, n))
    k = 2.0 * np.pi * np.fft.fftfreq(n)
    kz, ky, kx = np.meshgrid(k, k, k, indexing="ij")
    s = np.broadcast_to(np.asarray(sigma, np.float64), (3,))
    if frame is None:
        q = (s[0] * kx) ** 2 + (s[1] * ky) ** 2 + (s[2] * kz) ** 2
    else:
        # covariance Sigma = R diag(s^2) R^T, H(k) = exp(-k^T Sigma k / 2)
        Sig = frame @ np.diag(s ** 2) @ frame.T
        K = np.stack([kx, ky, kz], axis=-1)
        q = np.einsum("...i,ij,...j->...", K, Sig, K)
    return np.real(np.fft.ifftn(np.fft.fftn(f) * np.exp(-0.5 * q)))


def _largest_component(mask):
    from scipy import ndimage
    lab, nlab = ndimage.label(mask, structure=np.ones((3, 3, 3), bool))
    if nlab == 0:
        return mask
    sizes = np.bincount(lab.ravel())
    sizes[0] = 0
    return lab == int(np.argmax(sizes))


def solid_cube(n=24):
    return np.ones((n, n, n), np.int8)


def trabecular(n=64, seed=2, sigma=2.0, bvtv=0.20, frame=None):
    rng = np.random.default_rng(seed)
    g = _smooth(n, rng, sigma, frame)
    thr = np.quantile(g, 1.0 - bvtv)
    return _largest_component(g >= thr).astype(np.int8)


def greyscale(n=96, seed=4, sigma=2.0, bvtv=0.20, E0=10000.0):
    """Returns (phi int16 0..255, modulus float64): ik2 keeps 128..255 positive."""
    rng = np.random.default_rng(seed)
    g = _smooth(n, rng, sigma)
    thr = np.quantile(g, 1.0 - bvtv)
    keep = _largest_component(g >= thr)
    lo, hi = thr, g.max()
    grey = np.where(keep, np.clip(np.rint(1.0 + 254.0 * (g - lo) / max(hi - lo, 1e-30)), 1, 255), 0).astype(np.int16)
    modulus = E0 * (grey.astype(np.float64) / 255.0) ** 2
    return grey, modulus


def random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def uneven_batch(count=512, seed=3, nmin=32, nmax=96, s=2.0):
    master = np.random.default_rng(seed)
    out = []
    for i in range(count):
        rng = np.random.default_rng([seed, i])
        n = int(rng.integers(nmin, nmax + 1))
        bv = float(rng.uniform(0.08, 0.35))
        frame = random_rotation(rng)
        out.append(trabecular(n, seed=int(master.integers(1 << 31)), sigma=(s, s, 2 * s), bvtv=bv, frame=frame))
    return out


def orthotropic_C(rng):
    """Seeded orthotropic SPD 6x6 Voigt stiffness (engineering shear)."""
    while True:
        E = rng.uniform(500, 5000, 3)
        nu = rng.uniform(0.1, 0.35, 3)    # nu12, nu13, nu23
        G = rng.uniform(200, 2000, 3)     # G12, G13, G23
        S = np.zeros((6, 6))
        S[0, 0], S[1, 1], S[2, 2] = 1 / E
        S[0, 1] = S[1, 0] = -nu[0] / E[0]
        S[0, 2] = S[2, 0] = -nu[1] / E[0]
        S[1, 2] = S[2, 1] = -nu[2] / E[1]
        S[3, 3], S[4, 4], S[5, 5] = 1 / G
        if np.all(np.linalg.eigvalsh(S) > 0):
            return np.linalg.inv(S)
