"""Seeded synthetic FOV stacks for tests and the benchmark.

Puncta are 3D Gaussians (sigmaZ, sigmaXY) of amplitude U(amin, amax) at uniform
continuous positions, convolved analytically with a Gaussian PSF (variances
add, the peak shrinks by sigma_punctum/sigma_total per axis), on a Poisson
background.  An optional smooth illumination ``g(x,y) = 1 - 0.3 r^2`` (r = 1 at
the corner) multiplies background + signal before the Poisson draw and is
returned as the float32 flat field.
"""
import math

import numpy as np


def gaussian_psf(shape, sigma_z, sigma_xy, dtype=np.float32):
    """Centred anisotropic Gaussian PSF of ``shape`` (sum 1)."""
    z = np.arange(shape[0]) - (shape[0] - 1) / 2 if shape[0] % 2 else np.arange(shape[0]) - shape[0] // 2
    x = np.arange(shape[1]) - (shape[1] - 1) / 2 if shape[1] % 2 else np.arange(shape[1]) - shape[1] // 2
    y = np.arange(shape[2]) - (shape[2] - 1) / 2 if shape[2] % 2 else np.arange(shape[2]) - shape[2] // 2
    g = (np.exp(-z[:, None, None] ** 2 / (2 * sigma_z ** 2)) * np.exp(-x[None, :, None] ** 2 / (2 * sigma_xy ** 2))
         * np.exp(-y[None, None, :] ** 2 / (2 * sigma_xy ** 2)))
    return (g / g.sum()).astype(dtype)


def illumination(shape_xy):
    sx, sy = shape_xy
    cx, cy = (sx - 1) / 2.0, (sy - 1) / 2.0
    x = (np.arange(sx) - cx)[:, None]
    y = (np.arange(sy) - cy)[None, :]
    r2 = (x * x + y * y) / (cx * cx + cy * cy)
    return (1.0 - 0.3 * r2).astype(np.float32)


def puncta(rng, shape, n, amp=(2000.0, 10000.0)):
    sz, sx, sy = shape
    pos = rng.uniform([0, 0, 0], [sz, sx, sy], size=(n, 3))
    a = rng.uniform(amp[0], amp[1], size=n)
    return pos, a


def render_np(shape, pos, amp, sz_tot, sxy_tot, peak_scale, radius=(6, 7)):
    """Sum of Gaussians into a float32 volume (NumPy, for small stacks)."""
    vol = np.zeros(shape, dtype=np.float32)
    rz, rxy = radius
    dz = np.arange(-rz, rz + 1)
    dxy = np.arange(-rxy, rxy + 1)
    for (pz, px, py), a in zip(pos, amp):
        cz, cx, cy = int(round(pz)), int(round(px)), int(round(py))
        z = cz + dz
        x = cx + dxy
        y = cy + dxy
        mz = (z >= 0) & (z < shape[0])
        mx = (x >= 0) & (x < shape[1])
        my = (y >= 0) & (y < shape[2])
        if not (mz.any() and mx.any() and my.any()):
            continue
        gz = np.exp(-(z[mz] - pz) ** 2 / (2 * sz_tot ** 2))
        gx = np.exp(-(x[mx] - px) ** 2 / (2 * sxy_tot ** 2))
        gy = np.exp(-(y[my] - py) ** 2 / (2 * sxy_tot ** 2))
        vol[np.ix_(z[mz], x[mx], y[my])] += (a * peak_scale * gz[:, None, None] * gx[None, :, None]
                                             * gy[None, None, :]).astype(np.float32)
    return vol


def stack_np(seed, shape, n_puncta, psf_sigma=(1.5, 1.2), spot_sigma=(1.0, 1.5), background=500.0,
             amp=(2000.0, 10000.0), flat=False):
    """uint16 stack (and float32 flat or None), NumPy generator (small shapes)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sz_tot = math.sqrt(spot_sigma[0] ** 2 + psf_sigma[0] ** 2)
    sxy_tot = math.sqrt(spot_sigma[1] ** 2 + psf_sigma[1] ** 2)
    peak = (spot_sigma[0] / sz_tot) * (spot_sigma[1] / sxy_tot) ** 2
    pos, a = puncta(rng, shape, n_puncta, amp)
    sig = render_np(shape, pos, a, sz_tot, sxy_tot, peak)
    g = illumination(shape[1:]) if flat else None
    rate = (background + sig) * (g[None] if flat else 1.0)
    raw = np.clip(rng.poisson(rate), 0, 65535).astype(np.uint16)
    return raw, g


def stack_torch(seed, shape, density, device, psf_sigma=(1.5, 1.2), spot_sigma=(1.0, 1.5), background=500.0,
                amp=(2000.0, 10000.0), flat=False, chunk=2048):
    """uint16 device stack (and float32 device flat or None) for large FOVs.

    ``density`` is puncta per voxel (1/20**3 for the full hyb FOV)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    sz, sx, sy = shape
    n = int(round(sz * sx * sy * density))
    sz_tot = math.sqrt(spot_sigma[0] ** 2 + psf_sigma[0] ** 2)
    sxy_tot = math.sqrt(spot_sigma[1] ** 2 + psf_sigma[1] ** 2)
    peak = (spot_sigma[0] / sz_tot) * (spot_sigma[1] / sxy_tot) ** 2
    pos = torch.rand((n, 3), generator=g, device=device, dtype=torch.float64) * torch.tensor(
        [sz, sx, sy], dtype=torch.float64, device=device)
    a = amp[0] + (amp[1] - amp[0]) * torch.rand(n, generator=g, device=device, dtype=torch.float64)
    vol = torch.full((sz * sx * sy,), float(background), dtype=torch.float32, device=device)
    rz, rxy = 6, 7
    dz = torch.arange(-rz, rz + 1, device=device)
    dxy = torch.arange(-rxy, rxy + 1, device=device)
    for c0 in range(0, n, chunk):
        p = pos[c0:c0 + chunk]
        aa = a[c0:c0 + chunk]
        c = torch.round(p).long()
        z = c[:, 0:1] + dz[None]
        x = c[:, 1:2] + dxy[None]
        y = c[:, 2:3] + dxy[None]
        gz = torch.exp(-(z - p[:, 0:1]) ** 2 / (2 * sz_tot ** 2)) * ((z >= 0) & (z < sz))
        gx = torch.exp(-(x - p[:, 1:2]) ** 2 / (2 * sxy_tot ** 2)) * ((x >= 0) & (x < sx))
        gy = torch.exp(-(y - p[:, 2:3]) ** 2 / (2 * sxy_tot ** 2)) * ((y >= 0) & (y < sy))
        w = (aa * peak)[:, None, None, None] * gz[:, :, None, None] * gx[:, None, :, None] * gy[:, None, None, :]
        lin = (z.clamp(0, sz - 1)[:, :, None, None] * sx + x.clamp(0, sx - 1)[:, None, :, None]) * sy \
            + y.clamp(0, sy - 1)[:, None, None, :]
        vol.index_add_(0, lin.reshape(-1), w.reshape(-1).to(torch.float32))
    vol = vol.view(sz, sx, sy)
    gflat = None
    if flat:
        gflat = torch.from_numpy(illumination((sx, sy))).to(device)
        vol.mul_(gflat[None])
    raw = torch.poisson(vol, generator=g).clamp_(0, 65535).to(torch.int32).to(torch.uint16)
    return raw, gflat
