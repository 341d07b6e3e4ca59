"""MI355X (gfx950) implementation of MERMAKE's image hot path.

Operators keep the reference's Python signatures (``mermake.blur``,
``mermake.fill``, ``mermake.maxima``, ``mermake.deconvolver``,
``mermake.align``, ``mermake.utils``) and run through ``libmermake_hip.so``
(C ABI in ``include/mermake_hip.h``).  ``mermake_amd.pipeline`` holds the
batched hyb / DAPI callers of ``mermake/__main__.py``.
"""
__version__ = "0.1.0"

from . import _lib  # noqa: F401
