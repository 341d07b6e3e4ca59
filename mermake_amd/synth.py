"""Seeded synthetic FOV stacks for tests and the benchmark.

Puncta are 3D Gaussians (sigmaZ, sigmaXY) of amplitude U(amin, amax) at uniform
continuous positions, convolved analytically with a Gaussian PSF (variances
add, the peak shrinks by sigma_punctum/sigma_total per axis), on a Poisson
background.  An optional smooth illumination ``g(x,y) = 1 - 0.3 r^2`` (r = 1 at
the corner) multiplies background + signal before the Poisson draw and is
returned as the float32 flat field.  DAPI-style stacks (``nuclei_torch``) hold ellipsoidal nuclei
textured with chromatin spots, optionally shifted by an integer offset for drift tests.
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


def _add_gaussians(vol, shape, pos, a, sz_tot, sxy_tot, radius=(6, 7), chunk=2048):
    """vol (flat float32 device tensor of ``shape``) += a * exp(-d^2 / 2 sigma^2) around each position."""
    import torch
    sz, sx, sy = shape
    device = vol.device
    rz, rxy = radius
    dz = torch.arange(-rz, rz + 1, device=device)
    dxy = torch.arange(-rxy, rxy + 1, device=device)
    for c0 in range(0, pos.shape[0], chunk):
        p = pos[c0:c0 + chunk]
        aa = a[c0:c0 + chunk]
        c = torch.round(p).long()
        z = c[:, 0:1] + dz[None]
        x = c[:, 1:2] + dxy[None]
        y = c[:, 2:3] + dxy[None]
        gz = torch.exp(-(z - p[:, 0:1]) ** 2 / (2 * sz_tot ** 2)) * ((z >= 0) & (z < sz))
        gx = torch.exp(-(x - p[:, 1:2]) ** 2 / (2 * sxy_tot ** 2)) * ((x >= 0) & (x < sx))
        gy = torch.exp(-(y - p[:, 2:3]) ** 2 / (2 * sxy_tot ** 2)) * ((y >= 0) & (y < sy))
        w = aa[:, None, None, None] * gz[:, :, None, None] * gx[:, None, :, None] * gy[:, None, None, :]
        lin = (z.clamp(0, sz - 1)[:, :, None, None] * sx + x.clamp(0, sx - 1)[:, None, :, None]) * sy \
            + y.clamp(0, sy - 1)[:, None, None, :]
        vol.index_add_(0, lin.reshape(-1), w.reshape(-1).to(torch.float32))


def nuclei_torch(seed, shape, n_blobs, device, shift=(0, 0, 0), noise_seed=None, radius=(20.0, 40.0),
                 zradius=(6.0, 12.0), amp=(1500.0, 3000.0), spots=12, spot_amp=(3000.0, 8000.0),
                 spot_sigma=(1.5, 2.0), background=500.0, chunk=8):
    """uint16 DAPI-style device stack: ``n_blobs`` ellipsoidal nuclei (semi-axes U(radius) in x and y,
    U(zradius) in z, smooth edges, amplitude U(amp)), each textured with ``spots`` Gaussian chromatin spots
    (sigma ``spot_sigma`` in z and xy, amplitude U(spot_amp)) at uniform points inside it, on a Poisson
    background.  The scene (blob and spot positions) is fixed by ``seed``; ``shift`` (dz, dx, dy) moves all of
    it, and ``noise_seed`` (default ``seed``) draws the Poisson noise, so a shifted pair of stacks is
    ``nuclei_torch(s, ..., noise_seed=a)`` and ``nuclei_torch(s, ..., shift=d, noise_seed=b)``."""
    import torch
    sz, sx, sy = shape
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    f64 = dict(generator=g, device=device, dtype=torch.float64)
    lo = torch.tensor([0.0, 0.0, 0.0], dtype=torch.float64, device=device)
    ext = torch.tensor([sz, sx, sy], dtype=torch.float64, device=device)
    cen = lo + torch.rand((n_blobs, 3), **f64) * ext
    rxy = radius[0] + (radius[1] - radius[0]) * torch.rand((n_blobs, 2), **f64)
    rz = zradius[0] + (zradius[1] - zradius[0]) * torch.rand((n_blobs, 1), **f64)
    rad = torch.cat([rz, rxy], 1)
    a = amp[0] + (amp[1] - amp[0]) * torch.rand(n_blobs, **f64)
    # chromatin spots at uniform points of each ellipsoid (direction, radius u^(1/3) * 0.85)
    d = torch.randn((n_blobs, spots, 3), **f64)
    d = d / d.norm(dim=2, keepdim=True).clamp_min(1e-9)
    u = torch.rand((n_blobs, spots, 1), **f64) ** (1.0 / 3.0) * 0.85
    spos = (cen[:, None, :] + d * u * rad[:, None, :]).reshape(-1, 3)
    sa = spot_amp[0] + (spot_amp[1] - spot_amp[0]) * torch.rand(n_blobs * spots, **f64)
    sh = torch.tensor([float(v) for v in shift], dtype=torch.float64, device=device)
    cen = cen + sh
    spos = spos + sh
    vol = torch.full((sz * sx * sy,), float(background), dtype=torch.float32, device=device)
    # blobs: a * sigmoid(6 (1 - q^(1/2))), q the ellipsoid's normalised radius squared, over a box of 1.3 r
    bz, bxy = int(math.ceil(zradius[1] * 1.3)), int(math.ceil(radius[1] * 1.3))
    oz = torch.arange(-bz, bz + 1, device=device)
    oxy = torch.arange(-bxy, bxy + 1, device=device)
    for c0 in range(0, n_blobs, chunk):
        p, r, aa = cen[c0:c0 + chunk], rad[c0:c0 + chunk], a[c0:c0 + chunk]
        c = torch.round(p).long()
        z = c[:, 0:1] + oz[None]
        x = c[:, 1:2] + oxy[None]
        y = c[:, 2:3] + oxy[None]
        qz = ((z - p[:, 0:1]) / r[:, 0:1]) ** 2
        qx = ((x - p[:, 1:2]) / r[:, 1:2]) ** 2
        qy = ((y - p[:, 2:3]) / r[:, 2:3]) ** 2
        q = qz[:, :, None, None] + qx[:, None, :, None] + qy[:, None, None, :]
        ok = (((z >= 0) & (z < sz))[:, :, None, None] & ((x >= 0) & (x < sx))[:, None, :, None]
              & ((y >= 0) & (y < sy))[:, None, None, :])
        w = aa[:, None, None, None] * torch.sigmoid(6.0 * (1.0 - torch.sqrt(q))) * ok
        lin = (z.clamp(0, sz - 1)[:, :, None, None] * sx + x.clamp(0, sx - 1)[:, None, :, None]) * sy \
            + y.clamp(0, sy - 1)[:, None, None, :]
        vol.index_add_(0, lin.reshape(-1), w.reshape(-1).to(torch.float32))
    _add_gaussians(vol, shape, spos, sa, spot_sigma[0], spot_sigma[1], radius=(6, 8))
    gn = torch.Generator(device=device)
    gn.manual_seed(int(seed if noise_seed is None else noise_seed))
    return torch.poisson(vol.view(sz, sx, sy), generator=gn).clamp_(0, 65535).to(torch.int32).to(torch.uint16)


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
    _add_gaussians(vol, shape, pos, a * peak, sz_tot, sxy_tot, chunk=chunk)
    vol = vol.view(sz, sx, sy)
    gflat = None
    if flat:
        gflat = torch.from_numpy(illumination((sx, sy))).to(device)
        vol.mul_(gflat[None])
    raw = torch.poisson(vol, generator=g).clamp_(0, 65535).to(torch.int32).to(torch.uint16)
    return raw, gflat
