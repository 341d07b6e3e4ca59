"""Oracle: 3D local maxima + delta_fit refinement (TEST INFRASTRUCTURE ONLY).

Follows ``mermake/maxima.py:17-98`` and the CUDA kernels
``mermake/maxima.cu:1-145`` (``local_maxima_count`` / ``local_maxima``) and
``mermake/maxima.cu:148-341`` (``delta_fit_cross_corr``).

Maxima (``maxima.cu:19-62``): a voxel is rejected when ``v <= threshold`` (so a
NaN threshold rejects nothing) and is a candidate when ``v >= img[n]`` for every neighbour ``n`` of the sphere
``dz^2+dx^2+dy^2 <= delta^2`` (centre excluded), so ties are maxima.  Neighbour
indices are reflected once: ``-n`` below zero, ``2N-n-2`` above the end.  For
axes shorter than ``delta+1`` that single reflection can still leave the axis;
the reference then reads out of bounds (undefined).  Here such neighbours are
skipped; the HIP kernels do the same.

The reference writes candidates at ``atomicAdd`` positions (order not
deterministic); the oracle returns them in raster order, which is the order the
HIP path produces, and comparisons are made as sets.

Fit (``maxima.cu:189-340``): samples the ``delta_fit`` sphere in dz->dx->dy loop
order with reflected reads (skipping samples still out of range after the single
reflection, at most 600 samples); Gaussian weights
``exp(-(dz^2/2sz^2 + dx^2/2sxy^2 + dy^2/2sxy^2))`` normalised to zero mean and
unit population std (no guard: std 0 gives NaN as in the reference);
``background = min(img samples)``; ``hn = mean(norm(img-bg) * G)``,
``a = mean(norm(raw) * G)`` with std 0 replaced by 1 for those two; centroid
``sum((c0+d)*s)/sum(s)`` over the UNREFLECTED coordinates, ``c0`` when
``sum(s) <= 0``; ``h = img[c0]``, ``habs = raw[c0]``.  Output columns
``[zc, xc, yc, background, a, habs, hn, h]``.

Pins: the reference's own tests ``tests/test_maxima.py:26-166`` (counts, shapes,
threshold/delta behaviour) replayed in ``tests/test_oracle_cpu.py``; a brute-force
triple loop over every voxel and every sphere neighbour on small volumes; for
the fit, closed forms (symmetric peak -> integer centroid, linear ramp -> centroid
shifted by mean(offset^2)/delta_fit, constant image -> ``hn = a = 0``, background =
min), the correlation columns as Pearson correlations (an image or raw equal to
c*G + d on the sphere gives exactly +-1 with population statistics; arbitrary
volumes give ``numpy.corrcoef`` of the sphere samples), and an independent
per-candidate scalar loop that follows the CUDA kernel line by line.
"""
import numpy as np

MAX_KERNEL_POINTS = 600


def _reflect1(n, N):
    """Single reflection of ``maxima.cu:44-52``."""
    n = np.where(n < 0, -n, n)
    return np.where(n >= N, 2 * N - n - 2, n)


def sphere_offsets(delta, include_centre=False):
    """Sphere offsets in the reference loop order dz -> dx -> dy (``maxima.cu:25-36``)."""
    offs = []
    for dz in range(-delta, delta + 1):
        for dx in range(-delta, delta + 1):
            for dy in range(-delta, delta + 1):
                if dz == 0 and dx == 0 and dy == 0 and not include_centre:
                    continue
                if dz * dz + dx * dx + dy * dy > delta * delta:
                    continue
                offs.append((dz, dx, dy))
    return np.array(offs, dtype=np.int64).reshape(-1, 3)


def local_maxima(image, threshold, delta):
    """Raster-ordered (z, x, y) int64 arrays of the maxima of ``maxima.cu:1-145``."""
    img = np.asarray(image)
    D, H, W = img.shape
    flat = img.reshape(-1)
    with np.errstate(invalid="ignore"):   # maxima.cu:19 rejects v <= threshold (NaN rejects nothing)
        cand = np.flatnonzero(~(flat <= np.asarray(threshold, dtype=img.dtype)))
    z, rem = np.divmod(cand, H * W)
    x, y = np.divmod(rem, W)
    v = flat[cand]
    keep = np.ones(cand.size, dtype=bool)
    for dz, dx, dy in sphere_offsets(delta):
        nz = _reflect1(z + dz, D)
        nx = _reflect1(x + dx, H)
        ny = _reflect1(y + dy, W)
        ok = (nz >= 0) & (nz < D) & (nx >= 0) & (nx < H) & (ny >= 0) & (ny < W)
        nb = np.where(ok, flat[np.clip(nz, 0, D - 1) * H * W + np.clip(nx, 0, H - 1) * W + np.clip(ny, 0, W - 1)], v)
        keep &= ~(v < nb)
    return z[keep], x[keep], y[keep]


def local_maxima_bruteforce(image, threshold, delta):
    """Every voxel, every neighbour, scalar loop: pins ``local_maxima``."""
    img = np.asarray(image)
    D, H, W = img.shape
    out = []
    offs = sphere_offsets(delta)
    for z in range(D):
        for x in range(H):
            for y in range(W):
                v = img[z, x, y]
                if v <= threshold:
                    continue
                ismax = True
                for dz, dx, dy in offs:
                    nz, nx, ny = z + dz, x + dx, y + dy
                    nz = -nz if nz < 0 else (2 * D - nz - 2 if nz >= D else nz)
                    nx = -nx if nx < 0 else (2 * H - nx - 2 if nx >= H else nx)
                    ny = -ny if ny < 0 else (2 * W - ny - 2 if ny >= W else ny)
                    if not (0 <= nz < D and 0 <= nx < H and 0 <= ny < W):
                        continue
                    if v < img[nz, nx, ny]:
                        ismax = False
                        break
                if ismax:
                    out.append((z, x, y))
    return out


def delta_fit(image, raw, z0, x0, y0, delta_fit, sigmaZ, sigmaXY, dtype=np.float64):
    """Vectorised restatement of ``delta_fit_cross_corr`` (``maxima.cu:148-341``).

    ``dtype`` selects the arithmetic (float64 for the accuracy oracle).
    Returns an ``(N, 8)`` array in ``dtype``.
    """
    img = np.asarray(image)
    raw = np.asarray(raw)
    D, H, W = img.shape
    n = len(z0)
    out = np.zeros((n, 8), dtype=dtype)
    if n == 0:
        return out
    offs = sphere_offsets(delta_fit, include_centre=True)
    z0 = np.asarray(z0, dtype=np.int64)[:, None]
    x0 = np.asarray(x0, dtype=np.int64)[:, None]
    y0 = np.asarray(y0, dtype=np.int64)[:, None]
    dz, dx, dy = offs[:, 0][None], offs[:, 1][None], offs[:, 2][None]
    zz = _reflect1(z0 + dz, D)
    xx = _reflect1(x0 + dx, H)
    yy = _reflect1(y0 + dy, W)
    valid = (zz >= 0) & (zz < D) & (xx >= 0) & (xx < H) & (yy >= 0) & (yy < W)
    valid &= np.cumsum(valid, axis=1) <= MAX_KERNEL_POINTS
    lin = np.clip(zz, 0, D - 1) * H * W + np.clip(xx, 0, H - 1) * W + np.clip(yy, 0, W - 1)
    iv = img.reshape(-1)[lin].astype(dtype)
    rv = raw.reshape(-1)[lin].astype(dtype)
    cnt = valid.sum(axis=1).astype(dtype)
    sz2 = dtype(2) * dtype(sigmaZ) * dtype(sigmaZ)
    sxy2 = dtype(2) * dtype(sigmaXY) * dtype(sigmaXY)
    g = np.exp(-((dz * dz).astype(dtype) / sz2 + (dx * dx).astype(dtype) / sxy2 + (dy * dy).astype(dtype) / sxy2))
    g = np.broadcast_to(g, valid.shape)
    with np.errstate(invalid="ignore", divide="ignore"):
        mean_g = np.where(valid, g, 0).sum(1) / cnt
        var_g = np.where(valid, (g - mean_g[:, None]) ** 2, 0).sum(1) / cnt
        std_g = np.sqrt(var_g)
        ng = (g - mean_g[:, None]) / std_g[:, None]
        bg = np.where(valid, iv, np.inf).min(1)
        s = iv - bg[:, None]
        mean_s = np.where(valid, s, 0).sum(1) / cnt
        mean_r = np.where(valid, rv, 0).sum(1) / cnt
        std_s = np.sqrt(np.where(valid, (s - mean_s[:, None]) ** 2, 0).sum(1) / cnt)
        std_r = np.sqrt(np.where(valid, (rv - mean_r[:, None]) ** 2, 0).sum(1) / cnt)
        std_s1 = np.where(std_s > 0, std_s, 1)
        std_r1 = np.where(std_r > 0, std_r, 1)
        hn = np.where(valid, (s - mean_s[:, None]) / std_s1[:, None] * ng, 0).sum(1) / cnt
        a = np.where(valid, (rv - mean_r[:, None]) / std_r1[:, None] * ng, 0).sum(1) / cnt
        sw = np.where(valid, s, 0).sum(1)
        # centroid as sum(d*s)/sum(s) + c0 (algebraically the reference's sum((c0+d)*s)/sum(s))
        cz = np.where(valid, dz * s, 0).sum(1)
        cx = np.where(valid, dx * s, 0).sum(1)
        cy = np.where(valid, dy * s, 0).sum(1)
        pos = sw > 0
        zc = np.where(pos, cz / np.where(pos, sw, 1) + z0[:, 0], z0[:, 0])
        xc = np.where(pos, cx / np.where(pos, sw, 1) + x0[:, 0], x0[:, 0])
        yc = np.where(pos, cy / np.where(pos, sw, 1) + y0[:, 0], y0[:, 0])
    c = (z0[:, 0] * H * W + x0[:, 0] * W + y0[:, 0])
    out[:, 0] = zc
    out[:, 1] = xc
    out[:, 2] = yc
    out[:, 3] = bg
    out[:, 4] = a
    out[:, 5] = raw.reshape(-1)[c]
    out[:, 6] = hn
    out[:, 7] = img.reshape(-1)[c]
    out[cnt == 0] = 0
    return out


def delta_fit_scalar(image, raw, z0, x0, y0, delta_fit, sigmaZ, sigmaXY):
    """One candidate, scalar float64 loop in the order of ``maxima.cu:189-340``: pins ``delta_fit``."""
    img = np.asarray(image, dtype=np.float64)
    raw = np.asarray(raw, dtype=np.float64)
    Z, X, Y = img.shape
    vals, raws, ds, gs = [], [], [], []
    for dz in range(-delta_fit, delta_fit + 1):
        for dx in range(-delta_fit, delta_fit + 1):
            for dy in range(-delta_fit, delta_fit + 1):
                if dz * dz + dx * dx + dy * dy > delta_fit * delta_fit:
                    continue
                if len(vals) >= MAX_KERNEL_POINTS:
                    continue
                zz, xx, yy = z0 + dz, x0 + dx, y0 + dy
                zz = -zz if zz < 0 else (2 * Z - zz - 2 if zz >= Z else zz)
                xx = -xx if xx < 0 else (2 * X - xx - 2 if xx >= X else xx)
                yy = -yy if yy < 0 else (2 * Y - yy - 2 if yy >= Y else yy)
                if not (0 <= zz < Z and 0 <= xx < X and 0 <= yy < Y):
                    continue
                vals.append(img[zz, xx, yy])
                raws.append(raw[zz, xx, yy])
                ds.append((dz, dx, dy))
                gs.append(np.exp(-(dz * dz / (2 * sigmaZ * sigmaZ) + dx * dx / (2 * sigmaXY * sigmaXY)
                                   + dy * dy / (2 * sigmaXY * sigmaXY))))
    n = len(vals)
    if n == 0:
        return np.zeros(8)
    g = np.array(gs)
    with np.errstate(invalid="ignore", divide="ignore"):
        g = (g - g.mean()) / np.sqrt(((g - g.mean()) ** 2).mean())
    v = np.array(vals)
    r = np.array(raws)
    bg = v.min()
    s = v - bg
    ss = s.std() if s.std() > 0 else 1.0
    rs = r.std() if r.std() > 0 else 1.0
    hn = np.mean((s - s.mean()) / ss * g)
    a = np.mean((r - r.mean()) / rs * g)
    sw = s.sum()
    d = np.array(ds, dtype=np.float64)
    if sw > 0:
        c = ((np.array([z0, x0, y0], dtype=np.float64)[None] + d) * s[:, None]).sum(0) / sw
    else:
        c = np.array([z0, x0, y0], dtype=np.float64)
    return np.array([c[0], c[1], c[2], bg, a, raw[z0, x0, y0], hn, img[z0, x0, y0]])


_delta_fit = delta_fit


def find_local_maxima(image, threshold=None, delta=1, delta_fit=0, raw=None, sigmaZ=1, sigmaXY=1.5,
                      dtype=np.float64, **kwargs):
    """Restatement of ``find_local_maxima`` (``maxima.py:17-98``); rows in raster order."""
    if delta > 5:
        raise TypeError("Delta must be an less than or equal to 5 due to MAX_KERNEL_POINTS")
    img = np.asarray(image)
    thr = np.float32(threshold)
    z, x, y = local_maxima(img, thr, delta)
    if len(z) == 0:
        return np.zeros((0, 8), dtype=np.float32 if dtype == np.float32 else dtype)
    if raw is None:
        raw = img
    raw = np.asarray(raw)
    if raw.dtype != np.float32:
        raw = raw.astype(np.float32)
    return _delta_fit(img, raw, z, x, y, delta_fit, np.float32(sigmaZ), np.float32(sigmaXY), dtype=dtype)
