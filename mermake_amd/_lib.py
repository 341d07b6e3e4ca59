"""ctypes binding of libmermake_hip.so (the C ABI declared in include/mermake_hip.h).

The library is built in-tree by ``mermake_amd.csrc.build``.  There is no
fallback: every operator of this package runs through the HIP library, and a
missing library raises ``MermakeLibraryError`` on first use.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libmermake_hip.so")


class MermakeLibraryError(RuntimeError):
    """The HIP library is missing or a call into it failed."""


_vp = ctypes.c_void_p
_i = ctypes.c_int
_ll = ctypes.c_longlong
_f = ctypes.c_float
_d = ctypes.c_double
_sz = ctypes.c_size_t

# name -> (restype, argtypes)
_SIGS = {
    "mm_version": (_i, []),
    "mm_last_error": (ctypes.c_char_p, []),
    "mm_box_1d": (_i, [_vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "mm_box_1d_sub": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "mm_blur_sub_xy": (_i, [_vp, _vp, _vp, _i, _i, _i, _vp]),
    "mm_blur_sub_xy_window": (_i, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _vp]),
    "mm_reflect": (_i, [_vp, _i, ctypes.POINTER(_ll), ctypes.POINTER(_ll), _i, _i, _i, _vp]),
    "mm_deconv_create": (_i, [_i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "mm_deconv_destroy": (_i, [_vp]),
    "mm_deconv_dims": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "mm_deconv_spec_layout": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "mm_deconv_schedule": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                                ctypes.POINTER(_i)]),
    "mm_deconv_set_schedule": (_i, [_vp, _i, _i]),
    "mm_deconv_wiener_ws_bytes": (_sz, [_vp]),
    "mm_deconv_build_wiener": (_i, [_vp, _vp, _d, _vp, _vp, _sz, _vp]),
    "mm_deconv_build_wiener_box": (_i, [_vp, _vp, _d, _i, _vp, _vp, _sz, _vp]),
    "mm_deconv_spec_bytes": (_sz, [_vp]),
    "mm_deconv_tiles": (_i, [_vp, _vp, _i, _vp, _i, _i, _vp, _vp, _vp]),
    "mm_deconv_tiles_window": (_i, [_vp, _vp, _i, _vp, _i, _i, _vp, _vp, _i, _i, _vp]),
    "mm_deconv_tile_raw": (_i, [_vp, _vp, _i, _vp, _vp]),
    "mm_deconv_set_profiling": (_i, [_vp, _i]),
    "mm_deconv_pass_times": (_i, [_vp, ctypes.POINTER(_f)]),
    "mm_deconv_tiles_stitch": (_i, [_vp, _vp, _i, _vp, _i, _i, _vp, _vp, _ll, _ll, _i, _i, _vp]),
    "mm_deconv_stitch": (_i, [_vp, _vp, _i, _vp, _vp, _ll, _ll, _i, _i, _vp]),
    "mm_fft3d_c2c": (_i, [_vp, _i, _i, _i, _i, _i, _vp]),
    "mm_fft3d_c2c_box": (_i, [_vp, _i, _i, _i, _i, _i, _vp]),
    "mm_median_ws_bytes": (_sz, []),
    "mm_median": (_i, [_vp, _ll, _vp, _vp, _sz, _vp]),
    "mm_rms_ws_bytes": (_sz, []),
    "mm_rms_normalize": (_i, [_vp, _ll, _vp, _vp, _vp]),
    "mm_maxima_ws_bytes": (_sz, [_i, _i, _i, _i]),
    "mm_local_maxima": (_i, [_vp, _i, _i, _i, _ll, _ll, _f, _f, _i, _vp, _i, _vp, _vp, _sz, _vp]),
    "mm_local_maxima_pm": (_i, [_vp, _i, _i, _i, _ll, _ll, _f, _i, _vp, _i, _vp, _vp, _sz, _vp]),
    "mm_local_maxima_noalloc": (_i, [_vp, _i, _i, _i, _ll, _ll, _f, _f, _i, _vp, _i, _vp, _vp]),
    "mm_delta_fit": (_i, [_vp, _i, _i, _i, _ll, _ll, _f, _vp, _i, _ll, _ll, _vp, _vp, _i, _i, _f, _f, _vp, _vp]),
    "mm_tiles_maxima": (_i, [_vp, _vp, _i, _vp, _f, _i, _vp, _i, _vp, _vp, _sz, _vp]),
    "mm_tiles_maxima_window": (_i, [_vp, _vp, _i, _vp, _f, _i, _i, _i, _vp, _i, _vp, _vp, _sz, _vp]),
    "mm_tiles_deconv_maxima_window": (_i, [_vp, _vp, _i, _vp, _i, _i, _vp, _vp, _i, _i, _f, _i, _i, _i, _vp, _i,
                                           _vp, _vp, _sz, _vp]),
    "mm_tiles_blur_maxima_window": (_i, [_vp, _vp, _i, _vp, _vp, _i, _i, _i, _f, _i, _i, _i, _vp, _i, _vp, _vp, _sz,
                                         _vp]),
    "mm_tiles_fit": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _i, _i, _f, _f, _vp, _vp]),
    "mm_tiles_keep": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp, _vp, _i, _vp, _vp, _vp]),
    "mm_points_round_bounds": (_i, [_vp, _i, _i, _f, _vp, _vp, _vp]),
    "mm_points_kept_max": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), _vp, _vp]),
    "mm_points_raster": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), _vp, _i, _i, _i, _i, _i, _i, _i, _vp]),
    "mm_cmul": (_i, [_vp, _vp, _ll, _vp]),
    "mm_argmax_ws_bytes": (_sz, []),
    "mm_argmax_rounded": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    "mm_nn_round": (_i, [_vp, _i, _i, _vp, _i, _i, _vp, _d, _vp, _vp, _vp, _vp]),
}

_lib = None


def lib():
    """The loaded library (loaded and declared on first use)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise MermakeLibraryError(
                f"{LIBPATH} is not built; build it with `python -m mermake_amd.csrc.build` "
                "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    """Call ``name`` and raise ``MermakeLibraryError`` on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().mm_last_error()
        raise MermakeLibraryError(f"{name} failed ({rc}): {msg.decode() if msg else ''}")
    return rc


def exported_symbols():
    return sorted(_SIGS)


def stream_ptr(device=None):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    """Device address of a tensor (None -> NULL)."""
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


# mm_tile_desc (include/mermake_hip.h), 104 bytes
TILE_DTYPE = np.dtype([
    ("img", "<u8"), ("flat", "<u8"), ("img_sz", "<i8"), ("img_sx", "<i8"), ("flat_sx", "<i8"),
    ("X0", "<i4"), ("Y0", "<i4"), ("xd", "<i4"), ("yd", "<i4"), ("xs", "<i4"), ("ys", "<i4"),
    ("xfirst", "<i4"), ("yfirst", "<i4"), ("xe", "<i4"), ("ye", "<i4"), ("psf", "<i4"),
    ("ox", "<i4"), ("oy", "<i4"), ("img_type", "<i4"), ("pad0", "<i4"), ("pad1", "<i4"),
])
assert TILE_DTYPE.itemsize == 104
