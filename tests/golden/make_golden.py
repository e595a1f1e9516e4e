"""Writes tests/golden/ref_records.npz from the reference's own CPU path.

Run in the container that holds the reference checkout, after
`make -C oracle ref` (which compiles the reference's src/cpu/cpu_backend.cpp
and headers into oracle/_ref/). For every configuration of tests/configs.py
PARITY it stores the reference's records (fate, n_boundary, n_scatter, t, x,
y, z, path) and Philox draw counts for photon ids 0..255 and for 256 ids
starting at 2^32 + 5 (so the high word of the photon id is exercised), in both
math modes. The file is data: inputs are the named configs and ids, outputs
the reference's values.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import Reference  # noqa: E402
from tests import configs  # noqa: E402

N = 256
FIRST_IDS = (0, (1 << 32) + 5)


def main():
    out = {}
    for math in ("portable", "libm"):
        ref = Reference(math)
        for name, p in sorted(configs.PARITY.items()):
            for first in FIRST_IDS:
                rows = [ref.one(p, first + i) for i in range(N)]
                key = f"{math}/{name}/{first}"
                out[key + "/fate"] = np.array([r["fate"] for r in rows], np.uint8)
                out[key + "/n_boundary"] = np.array([r["n_boundary"] for r in rows], np.uint32)
                out[key + "/n_scatter"] = np.array([r["n_scatter"] for r in rows], np.uint32)
                out[key + "/draws"] = np.array([r["draws"] for r in rows], np.uint32)
                for f in ("t", "x", "y", "z", "path"):
                    out[key + "/" + f] = np.array([r[f] for r in rows], np.float32)
    np.savez_compressed(os.path.join(HERE, "ref_records.npz"), **out)
    print(f"wrote {len(out)} arrays")


if __name__ == "__main__":
    main()
