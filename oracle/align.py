"""Oracle: drift between acquisitions from DAPI maxima (TEST INFRASTRUCTURE ONLY).

Follows ``mermake/align.py:12-177`` and ``mermake/utils.py:10-36`` (``fftconvolve``).

* ``threshold`` keeps rows with ``X[:, -1] > th`` (``align.py:121-129``).  Because
  ``DualAligner`` passes ``Xh[:, :3]`` for the reference set (``align.py:134-135``)
  the reference set is thresholded on column 2 (y), the moving set on column 7 (h).
* ``get_im_from_Xh`` (``align.py:37-59``): ``round(X/5)`` half-to-even as int32, trim
  ``(0, 10, 10)`` against the min/max, binary image of size ``max+1`` with origin 0.
* Correlation ``fftconvolve(im, im_ref[::-1,::-1,::-1])`` in 'full' mode; the values of
  binary images are integers, so the oracle rounds them and takes the first argmax.
* ``get_Xtzxy`` (``align.py:61-77``): 5 rounds at ``dist_th = linspace(5, 3, 5)`` of
  nearest-neighbour matching, mean of the differences of pairs closer than ``dist_th``
  (rounds with no pair are skipped), ``Npts`` from the last round.
* ``DualAligner`` combination (``align.py:138-156``).

Pins: ``fftconvolve`` in all three modes against ``scipy.signal.fftconvolve`` (ranks 1
to 3); the drift known answer of the reference fixture
``tests/examples/single/out/drift_Conv_zscan__020--_set1.pkl`` (N+ = 621, N- = 261,
zero shift; committed as ``tests/golden/drift_known_answer.json``); the branch logic of
``tests/test_align.py:181-233``; synthetic integer shifts that must be recovered exactly.
Nearest-neighbour ties: the reference's cupyx KDTree tie order is parity unpinned;
the oracle (scipy cKDTree) and the HIP path (lowest index among equal distances)
differ only on exact ties.
"""
import numpy as np
from scipy.spatial import cKDTree


def fftconvolve(a, b, mode="full", shape=None):
    """``utils.py:10-36`` in float64; ``shape`` may enlarge the FFT grid (linear result unchanged)."""
    full = [sa + sb - 1 for sa, sb in zip(a.shape, b.shape)]
    s = full if shape is None else shape
    fa = np.fft.fftn(a.astype(np.float64), s)
    fb = np.fft.fftn(b.astype(np.float64), s)
    out = np.fft.ifftn(fa * fb).real[tuple(slice(0, f) for f in full)]
    if mode == "full":
        return out
    if mode == "same":
        st = [(o - sa) // 2 for o, sa in zip(out.shape, a.shape)]
        return out[tuple(slice(st[i], st[i] + a.shape[i]) for i in range(len(st)))]
    if mode == "valid":
        vs = [sa - sb + 1 for sa, sb in zip(a.shape, b.shape)]
        st = [sb - 1 for sb in b.shape]
        return out[tuple(slice(st[i], st[i] + vs[i]) for i in range(len(st)))]
    raise ValueError("mode must be 'full', 'same', or 'valid'")


class Aligner:
    """``align.py:12-129``."""

    def __init__(self, X_ref, resc=5, trim=10, th=None):
        self.resc, self.trim, self.th = resc, trim, th
        X_ref = np.asarray(X_ref)
        if th:
            X_ref = self.threshold(X_ref)
        if X_ref.size == 0:
            raise ValueError("No points left after thresholding; cannot build KDTree. "
                             "Check your data and/or threshold value.")
        self.X_ref = X_ref[:, :3]
        self.tree = cKDTree(self.X_ref)
        self.im_ref, self.Xm_ref = self.get_im_from_Xh(self.X_ref)

    def get_im_from_Xh(self, Xh):
        X = np.round(Xh / self.resc).astype(np.int32)
        Xm = X.min(0)
        XM = X.max(0)
        n = np.array([0, self.trim, self.trim], dtype=np.int32)
        keep = np.all((X <= XM - n) & (X >= Xm + n), axis=-1)
        X = X[keep]
        if X.shape[0] == 0:
            return np.zeros((1, 1, 1), dtype=np.float32), np.zeros(3, dtype=np.int32)
        sz = X.max(0) + 1
        im = np.zeros(sz.tolist(), dtype=np.float32)
        im[tuple(X.T)] = 1
        return im, np.zeros(3, dtype=np.int32)

    def get_Xtzxy(self, X, tzxy0, target=3):
        tzxy = tzxy0
        keep = np.zeros(len(X), dtype=bool)
        for dist_th in np.linspace(self.resc, target, 5):
            XT = X - tzxy
            ds, inds = self.tree.query(XT)
            keep = ds < dist_th
            if not keep.any():
                continue
            tzxy = np.mean(X[keep] - self.X_ref[inds[keep]], axis=0)
        return tzxy, int(np.sum(keep))

    def correlation_peak(self, X):
        """Integer lag of the correlation maximum (first index among equal rounded values)."""
        im, Xm = self.get_im_from_Xh(X)
        cor = fftconvolve(im, self.im_ref[::-1, ::-1, ::-1])
        cor = np.rint(cor)
        idx = np.unravel_index(np.argmax(cor), cor.shape)
        return np.array(idx) - np.array(self.im_ref.shape) + 1 + Xm - self.Xm_ref

    def get_best_translation_points(self, X, target=3, return_counts=False):
        X = np.asarray(X)
        if self.th:
            X = self.threshold(X)
        X = X[:, :3]
        tzxy = self.correlation_peak(X) * self.resc
        tzxy, n = self.get_Xtzxy(X, tzxy, target=target)
        return (tzxy, n) if return_counts else tzxy

    def threshold(self, X):
        dim = X.shape[1]
        if X.size:
            X = X[X[:, -1] > self.th]
            if X.size:
                return X
        return np.zeros([0, dim])


class DualAligner:
    """``align.py:131-156``."""

    def __init__(self, ref, th=None):
        self.plus = Aligner(np.asarray(ref.Xh_plus)[:, :3], th=th)
        self.minus = Aligner(np.asarray(ref.Xh_minus)[:, :3], th=th)
        self.th = th

    def get_best_translation_pointsV2(self, obj):
        tp, n_p = self.plus.get_best_translation_points(obj.Xh_plus, return_counts=True)
        tm, n_m = self.minus.get_best_translation_points(obj.Xh_minus, return_counts=True)
        return combine(tp, n_p, tm, n_m) + (tp, tm, n_p, n_m)


def combine(tp, n_p, tm, n_m):
    """Combination rule of ``align.py:147-154``; returns a 1-tuple ``(tzxyf,)``."""
    tp = np.asarray(tp, dtype=np.float64)
    tm = np.asarray(tm, dtype=np.float64)
    if n_p + n_m == 0:
        f = np.zeros(3)
    elif np.max(np.abs(tm - tp)) <= 2:
        f = -(tp * n_p + tm * n_m) / (n_p + n_m)
    else:
        f = -tp if n_p >= n_m else -tm
    return (f,)
