"""Write tests/golden/interpolator_vectors.npz: the output of the reference's
gmg.create_interpolator (gmg.py:34-146) on small grids, for the oracle's
prolongator (oracle/mg.py) to be compared with.

The function is never called by the reference itself and none of the
reference's tests or fixtures holds its output, so its vectors are made here
by importing it from the reference checkout (read-only; the reference is not
needed afterwards).  For each grid the fixture keeps either the matrix as
sorted COO triplets (duplicates summed, explicit zeros dropped) or, when the
call raises, the exception's type and message.
Usage: python tests/golden/make_interpolator_fixture.py [reference root]
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "interpolator_vectors.npz"
# (nelx, nely, nelz or None, ndof)
CASES = [
    (2, 2, None, 1), (4, 4, None, 1), (4, 4, None, 2), (8, 8, None, 1),
    (2, 2, 2, 1), (4, 4, 4, 1), (4, 4, 4, 3),
    (4, 2, None, 1), (6, 4, None, 2), (4, 2, 2, 3), (2, 4, 6, 1),
]


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    sys.path.insert(0, str(ref))
    from topoptlab.gmg import create_interpolator
    arrays, meta = {}, []
    for c, (nelx, nely, nelz, ndof) in enumerate(CASES):
        entry = {"nelx": nelx, "nely": nely, "nelz": nelz, "ndof": ndof}
        try:
            P = create_interpolator(nelx, nely, nelz, ndof=ndof, stride=2).tocoo()
            P.sum_duplicates()
            keep = P.data != 0
            r, cc, v = P.row[keep], P.col[keep], P.data[keep]
            o = np.lexsort((cc, r))
            arrays[f"c{c}_row"], arrays[f"c{c}_col"], arrays[f"c{c}_val"] = r[o], cc[o], v[o]
            entry["shape"] = list(P.shape)
        except Exception as e:  # the fixture records what the call does
            entry["error"] = f"{type(e).__name__}: {e}"
        meta.append(entry)
        print(c, entry)
    np.savez(OUT, meta=json.dumps(meta), **arrays)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
