"""Kept-row comparison of the hyb caller against the oracle (test helpers).

Given the volumes a HIP pipeline's maxima and fit read (``HybPipeline.tile_volumes``), the oracle's
maxima, fit and keep filter (``maxima.py:17-98``, ``__main__.py:254-265``) must give exactly the
pipeline's kept rows: the same rows in the same order, fitted coordinates within 1e-3 px, background,
habs and h equal, correlations within 1e-4.  A row whose float64 fitted x or y lies within 1e-3 px of a
keep boundary may fall on either side.
"""
import numpy as np

from oracle import maxima as omax

BOUNDARY_TOL = 1e-3


def _close(g, w, tol=1e-3):
    return (np.abs(g[:3].astype(np.float64) - w[:3]).max() <= tol and g[3] == w[3] and g[5] == w[5]
            and g[7] == w[7] and np.abs(g[[4, 6]].astype(np.float64) - w[[4, 6]]).max() <= 1e-4)


def oracle_tile_rows(tiles, overlap, hyb):
    """The hyb loop body of ``__main__.py:254-265`` in the oracle, over tiles (x, y, tile, raw) given as
    host arrays: every fitted candidate as (row offset to stack coordinates, kept, near a keep boundary)."""
    out = []
    for x, y, tile, raw in tiles:
        Xh = omax.find_local_maxima(tile, raw=raw, dtype=np.float64, **hyb)
        hx, hy = tile.shape[1] - overlap, tile.shape[2] - overlap
        for r in Xh:
            keep = overlap <= r[1] < hx and overlap <= r[2] < hy
            amb = min(abs(r[1] - overlap), abs(r[1] - hx), abs(r[2] - overlap), abs(r[2] - hy)) <= BOUNDARY_TOL
            row = r.copy()
            row[1] += x - overlap
            row[2] += y - overlap
            out.append((row, keep, amb))
    return out


def assert_kept_rows(got, cand, what):
    """``got`` (the HIP kept rows, in order) against the oracle's candidates: every kept row not near a
    boundary matches the next HIP row, a row near a boundary may or may not, the others are absent."""
    i = 0
    for row, keep, amb in cand:
        if i < len(got) and (keep or amb) and _close(got[i], row):
            i += 1
        elif keep and not amb:
            raise AssertionError(f"{what}: oracle row {row} missing at HIP row {i} "
                                 f"({got[i] if i < len(got) else None})")
    assert i == len(got), f"{what}: {len(got) - i} HIP rows beyond the oracle's kept rows"
    assert sum(k for _, k, _ in cand) > 0, what
