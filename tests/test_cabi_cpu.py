"""CPU tests of the C ABI: the library loads and exports every declared symbol,
and the host-side logic (tile geometry) matches the oracle.  No GPU calls."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib_path():
    from mermake_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        from mermake_amd.csrc.build import build
        build()
    return _lib.LIBPATH


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mermake_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(mm_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib_path())
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    from mermake_amd import _lib
    assert set(_lib.exported_symbols()) == set(names)


def _declarations():
    """{name: (return type, [parameter types])} of the header's function declarations."""
    src = open(os.path.join(ROOT, "include", "mermake_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    out = {}
    pat = r"^[ \t]*((?:const[ \t]+)?[A-Za-z_]\w*(?:[ \t]*\*)?)[ \t]*\b(mm_[a-z0-9_]+)\s*\(([^)]*)\)\s*;"
    for ret, name, params in re.findall(pat, src, flags=re.M):
        ps = [p.strip() for p in params.split(",") if p.strip() and p.strip() != "void"]
        # drop the parameter name: the type is everything before the last identifier
        types = [re.sub(r"\s*\b\w+$", "", p).replace("const ", "").strip() for p in ps]
        out[name] = (ret.replace("const ", "").strip(), types)
    return out


def _ctype_ok(c_type, ct):
    scalar = {"int": ctypes.c_int, "long long": ctypes.c_longlong, "float": ctypes.c_float,
              "double": ctypes.c_double, "size_t": ctypes.c_size_t}
    c_type = " ".join(c_type.split())
    if c_type in scalar:
        return ct is scalar[c_type]
    if c_type == "char*":
        return ct is ctypes.c_char_p
    # pointers (hipStream_t is one): an address or a typed pointer
    assert c_type.endswith("*") or c_type == "hipStream_t", c_type
    return ct is ctypes.c_void_p or (isinstance(ct, type) and issubclass(ct, ctypes._Pointer))


def test_bindings_match_the_header_declarations():
    """Every ctypes signature of mermake_amd._lib has the header's parameter count and scalar types (a binding
    that drifts from the header passes wrong sizes or strides without any error)."""
    from mermake_amd import _lib
    decl = _declarations()
    assert set(decl) == set(_lib._SIGS)
    for name, (ret, params) in decl.items():
        res, args = _lib._SIGS[name]
        assert len(args) == len(params), (name, params)
        assert _ctype_ok(ret, res), (name, ret)
        for i, (c_type, ct) in enumerate(zip(params, args)):
            assert _ctype_ok(c_type, ct), (name, i, c_type, ct)


def test_version_and_error_string():
    L = ctypes.CDLL(_lib_path())
    L.mm_version.restype = ctypes.c_int
    L.mm_last_error.restype = ctypes.c_char_p
    assert L.mm_version() >= 1
    assert isinstance(L.mm_last_error(), bytes)


def test_blur_maxima_entry_rejects_bad_windows_without_a_gpu():
    """Argument checks of ``mm_tiles_blur_maxima_window`` run before any device call: a missing
    plan is MM_EINVAL with a message."""
    L = ctypes.CDLL(_lib_path())
    L.mm_last_error.restype = ctypes.c_char_p
    f = L.mm_tiles_blur_maxima_window
    f.restype = ctypes.c_int
    vp, i, fl = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    f.argtypes = [vp, vp, i, vp, vp, i, i, i, fl, i, i, i, vp, i, vp, vp, ctypes.c_size_t, vp]
    one = ctypes.c_void_p(1)
    assert f(None, one, 1, one, one, 15, 10, 50, 100.0, 1, 12, 40, one, 8, one, one, 1 << 20, None) == 1000
    assert b"bad argument" in L.mm_last_error()


def test_tile_descriptor_layout():
    from mermake_amd import _lib
    assert _lib.TILE_DTYPE.itemsize == 104
    hdr = open(os.path.join(ROOT, "include", "mermake_hip.h")).read()
    body = hdr[hdr.index("typedef struct mm_tile_desc"):hdr.index("} mm_tile_desc;")]
    body = re.sub(r"/\*.*?\*/", "", body.split("{", 1)[1], flags=re.S)
    fields = re.findall(r"\b([A-Za-z_][A-Za-z_0-9]*)\s*[;,]", body)
    assert fields == list(_lib.TILE_DTYPE.names)


def test_tile_geometry_matches_reference_loop():
    """Host tile table vs the reference tile loop (deconvolver.py:116-144) replayed by the oracle."""
    from mermake_amd.deconvolver import tile_geometry
    from oracle import deconv as odec
    for shape, ts, ov, npsf in [((20, 512, 512), 500, 89, 1), ((40, 2800, 2800), 500, 89, 1),
                                ((40, 2048, 2048), 500, 89, 1), ((40, 3000, 3000), 300, 89, 100),
                                ((40, 2304, 2304), 300, 89, 100), ((10, 50, 50), 30, 5, 1)]:
        g = tile_geometry(shape, ts, ov, npsf)
        T = ts + 2 * ov
        want = list(odec.tiled(shape, ts, ov))
        if npsf > 1:
            want = want[:npsf]
        assert [(t["x"], t["y"]) for t in g] == [(w[0], w[1]) for w in want]
        for t, (x, y, x0, x1, y0, y1) in zip(g, want):
            xend = (x1 - x0) + (ov if (x == 0 or x1 - x0 < T) else 0)
            assert (t["X0"], t["xd"], t["xe"]) == (x0, x1 - x0, min(xend, T))
            assert t["xs"] == (ov if x == 0 else 0)


def test_keyed_psf_assignment_on_the_psf_grid():
    """A 10 x 10 grid of PSFs keyed at the centres of 300-px cells of a 3000^2 FOV, used on a
    2304^2 FOV (8 x 8 tiles of 300, the last 204 wide): every tile takes the key nearest its
    centre, by brute force; the reference's zip pairing keeps only the first 64 PSFs in x-major
    order, so tile (1, 0) gets the PSF of cell (0, 8)."""
    import pytest
    from mermake_amd.deconvolver import tile_geometry
    keys = [(0, 150 + 300 * i, 150 + 300 * j) for i in range(10) for j in range(10)]
    shape, ts, ov = (40, 2304, 2304), 300, 89
    g = tile_geometry(shape, ts, ov, len(keys), keys, "keyed")
    assert len(g) == 64
    for t in g:
        cx = t["x"] + (min(t["x"] + ts, shape[1]) - t["x"]) / 2
        cy = t["y"] + (min(t["y"] + ts, shape[2]) - t["y"]) / 2
        d2 = [(k[1] - cx) ** 2 + (k[2] - cy) ** 2 for k in keys]
        assert t["psf"] == int(np.argmin(d2))
        assert t["psf"] == 10 * (t["x"] // 300) + t["y"] // 300
    z = tile_geometry(shape, ts, ov, len(keys), keys, "zip")
    assert [t["psf"] for t in z] == list(range(64))
    assert (z[8]["x"], z[8]["y"]) == (300, 0) and keys[z[8]["psf"]] == (0, 150, 2550)
    with pytest.raises(ValueError):
        tile_geometry(shape, ts, ov, len(keys), None, "keyed")
    with pytest.raises(ValueError):
        tile_geometry(shape, ts, ov, len(keys), keys, "nearest")


def test_smooth_length():
    from mermake_amd.utils import smooth_length
    for n in [1, 2, 11, 13, 97, 1101, 1119]:
        m = smooth_length(n)
        k = m
        for p in (2, 3, 5, 7):
            while k % p == 0:
                k //= p
        assert k == 1 and m >= n and all(smooth_length(j) != j for j in range(n, m)) or m == n


def test_synthetic_generator_seeded():
    from mermake_amd.synth import gaussian_psf, stack_np
    a, f = stack_np(5, (4, 40, 40), 20, flat=True)
    b, _ = stack_np(5, (4, 40, 40), 20, flat=True)
    assert a.dtype == np.uint16 and np.array_equal(a, b) and f.shape == (40, 40)
    p = gaussian_psf((21, 41, 41), 1.5, 1.2)
    assert abs(p.sum() - 1) < 1e-5 and np.unravel_index(p.argmax(), p.shape) == (10, 20, 20)


def test_squared_u8_stack_type():
    """``SquaredU8`` keeps its type through views and copies, maps to pixel type 2 and holds
    uint8 only."""
    import pytest
    import torch
    from mermake_amd.deconvolver import SquaredU8, pixel_type
    a = SquaredU8.of(np.arange(60, dtype=np.uint8).reshape(3, 4, 5))
    for t in (a, a[:, 1:], a.contiguous(), a.clone()):
        assert isinstance(t, SquaredU8) and pixel_type(t) == 2
    assert pixel_type(torch.zeros(2, dtype=torch.uint16)) == 0 and pixel_type(torch.zeros(2)) == 1
    with pytest.raises(TypeError):
        SquaredU8.of(np.zeros((2, 2, 2), np.uint16))


def test_noalloc_maxima_rejects_oversized_volume_without_a_gpu():
    """``mm_local_maxima_noalloc`` checks the workspace a volume needs against its scratch array before any
    device call: too large is MM_ECAP (1001) with a message."""
    L = ctypes.CDLL(_lib_path())
    L.mm_last_error.restype = ctypes.c_char_p
    f = L.mm_local_maxima_noalloc
    f.restype = ctypes.c_int
    vp, i, ll, fl = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float
    f.argtypes = [vp, i, i, i, ll, ll, fl, fl, i, vp, i, vp, vp]
    one = ctypes.c_void_p(1)
    assert f(one, 48, 3000, 3000, 9000000, 3000, 1.0, 0.5, 1, one, 8, one, None) == 1001
    assert b"scratch" in L.mm_last_error()
    assert f(None, 4, 4, 4, 16, 4, 1.0, 0.5, 1, one, 8, one, None) == 1000
