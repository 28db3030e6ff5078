"""Writes the golden fixtures of tests/test_oracle.py::test_golden_direct_solves.

For each case: the case dictionary text (the input) and the FP64 solution of
the assembled coupled system by a sparse direct solve (scipy.sparse.linalg.
spsolve) in this container.  The assembly itself is pinned by the closed-form
tests; these fixtures pin the converged fields end to end.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from multiregionfoam_amd import cases  # noqa: E402
import oracle_bind as ob  # noqa: E402

FIXTURES = [
    dict(name="plate2d_40x10x6", builder="plate2d", kwargs=dict(nx=40, nyf=10, nys=6)),
    dict(name="duct_ggi_12x8x6", builder="duct3d", kwargs=dict(nf=(12, 8, 6), ns=(15, 10, 3))),
    dict(name="exchanger3_8x5x4", builder="exchanger3", kwargs=dict(n=(8, 5, 4))),
]


def main():
    meta = []
    for fx in FIXTURES:
        c = getattr(cases, fx["builder"])(**fx["kwargs"])
        case_file = fx["name"] + ".case"
        field_file = fx["name"] + ".npz"
        with open(os.path.join(HERE, case_file), "w") as f:
            f.write(cases.to_dict_text(c))
        O = ob.OracleSystem(c)
        psi = spla.spsolve(O.to_scipy().tocsc(), O.source_vector())
        np.savez_compressed(os.path.join(HERE, field_file), psi=psi)
        meta.append(dict(fx, case_file=case_file, field_file=field_file, n_cells=int(psi.size),
                         solver="scipy.sparse.linalg.spsolve (SuperLU), FP64"))
        print(fx["name"], psi.size, float(psi.min()), float(psi.max()))
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump({"fixtures": meta}, f, indent=1)


if __name__ == "__main__":
    main()
