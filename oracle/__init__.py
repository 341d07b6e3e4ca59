"""CPU oracle for the MERMAKE image hot path: TEST INFRASTRUCTURE ONLY.

This package is a from-scratch NumPy/SciPy restatement of the reference algorithm.
It serves two purposes:

* the checker that the HIP path is compared against (tests/, the smoke check in
  ``__graft_entry__.smoke``);
* the CPU baseline leg timed by ``bench.py`` (``cpu_baseline``).

Nothing in the product package (``mermake_amd``) imports it, and the product path
never falls back to it.

Two arithmetic modes are provided where it matters:

* ``"fp64"``: every step is carried out in float64 / complex128. This is the
  accuracy reference (relative L2 targets are stated against it).
* ``"fp32"``: mirrors the reference dtypes (float32 tiles, complex64 spectra,
  float16 Laplacian promoted to complex64). It is used for the CPU throughput
  baseline.

Pins (what each function is checked against) are listed in every module's
header; a function that nothing pins says "parity unpinned" there.
"""

from . import fill, blur, maxima, deconv, align, io  # noqa: F401
