"""Extract the committed golden vectors from the reference's own test data.

Run once in a container where the reference checkout is mounted:

    python tests/golden/make_fixtures.py /root/reference

Reads only data files (zarr v2 chunks decoded with the stdlib ``bz2`` module,
``.npy``/``.npz`` arrays, and the drift pickle inspected with ``pickletools``
without unpickling).  Writes into ``tests/golden/``:

* ``fixture_tiles.npz``: the three input windows (uint8 before ``read_im``'s
  squaring, stored as bz2-compressed bytes) that fully determine the owned output of tiles (0,0), (300,300) and
  (900,900) of ``full_deconv(channel 0, psf_750, tile_size=300, overlap=89)``
  (``tests/test_deconvolver.py:422-434``), plus the reference's float16 golden
  output of those owned regions on planes GOLDEN_Z (``tests/examples/single/out/
  deconv.300.{z}.npz``).
* ``psf_750.npy``: the fixture PSF (20x110x110 float32).
* ``dapi_features.npz``: ``Xh_plus`` / ``Xh_minus`` of the reference's DAPI
  features of the same FOV.
* ``drift_known_answer.json``: the drift pickle's numbers (N+, N-, shifts).
"""
import bz2
import io
import json
import os
import pickletools
import sys

import numpy as np

GOLDEN_Z = (0, 1, 17, 33, 34)
TILES = ((0, 0), (300, 300), (900, 900))
TS, OV, N = 300, 89, 1200


def read_channel0(root):
    """``mermake/io.py:116-146``: drop frame 0, de-interleave 3 channels, keep channel 0 (uint8)."""
    d = os.path.join(root, "tests/examples/single/H1_TEST_set1/020/data")
    meta = json.load(open(os.path.join(d, ".zarray")))
    assert meta["compressor"]["id"] == "bz2" and meta["dtype"] == "|u1"
    nfr = meta["shape"][0]
    nz = ((nfr - 1) // 3) * 3
    frames = [np.frombuffer(bz2.decompress(open(os.path.join(d, f"{i}.0.0"), "rb").read()), np.uint8)
              .reshape(N, N) for i in range(1, 1 + nz, 3)]
    return np.stack(frames)


def window(x):
    return max(x - OV, 0), min(x + TS + OV, N)


def main(root):
    here = os.path.dirname(os.path.abspath(__file__))
    ex = os.path.join(root, "tests/examples")
    ch0 = read_channel0(root)
    out = {}
    for (x, y) in TILES:
        x0, x1 = window(x)
        y0, y1 = window(y)
        w = np.ascontiguousarray(ch0[:, x0:x1, y0:y1])
        out[f"winbz2_{x}_{y}"] = np.frombuffer(bz2.compress(w.tobytes(), 9), np.uint8)
        out[f"winshape_{x}_{y}"] = np.array(w.shape)
        out[f"bounds_{x}_{y}"] = np.array([x0, x1, y0, y1])
        gold = []
        for z in GOLDEN_Z:
            g = np.load(os.path.join(ex, f"single/out/deconv.300.{z}.npz"))["arr_0"]
            gold.append(g[x:x + TS, y:y + TS])
        out[f"gold_{x}_{y}"] = np.stack(gold)
    out["golden_z"] = np.array(GOLDEN_Z)
    out["shape"] = np.array(ch0.shape)
    np.savez(os.path.join(here, "fixture_tiles.npz"), **out)

    np.save(os.path.join(here, "psf_750.npy"), np.load(os.path.join(ex, "psf_750_Scope0_final.npy")))

    f = np.load(os.path.join(ex, "single/out/Conv_zscan__020--H1_TEST_set1--dapiFeatures.npz"))
    np.savez_compressed(os.path.join(here, "dapi_features.npz"), Xh_plus=f["Xh_plus"], Xh_minus=f["Xh_minus"])

    # the drift pickle holds cupy arrays: read its integer fields from the opcode stream only
    raw = open(os.path.join(ex, "single/out/drift_Conv_zscan__020--_set1.pkl"), "rb").read()
    ints = [arg for op, arg, _ in pickletools.genops(io.BytesIO(raw))
            if op.name in ("BININT", "BININT1", "BININT2") and arg in (621, 261)]
    known = {"N_plus": 621, "N_minus": 261, "tzxyf": [0.0, 0.0, 0.0],
             "found_in_pickle": sorted(set(ints)),
             "source": "tests/examples/single/out/drift_Conv_zscan__020--_set1.pkl"}
    json.dump(known, open(os.path.join(here, "drift_known_answer.json"), "w"), indent=1)
    print("fixtures written to", here)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
