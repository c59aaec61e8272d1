"""Writes tests/golden/ref_outputs.json: outputs of the reference's own
Born/sinc_bob.cpp and Born/cgstep.cpp, compiled from its sources into
oracle/_ref by oracle/Makefile (`make -C oracle ref`), on seeded inputs.

The fixture lets the oracle be checked against the reference where the
reference sources are absent (the compiled oracle/_ref is never committed).

    python tests/golden/make_ref_fixtures.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import born_oracle as O  # noqa: E402


def main():
    sinc = O.ref_lib("ref_sinc")
    cg = O.ref_lib("ref_cgstep")
    if sinc is None or cg is None:
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` where the reference is")
    out = {"source": "Born/sinc_bob.cpp and Born/cgstep.cpp built by oracle/Makefile",
           "sinc": {}, "cgstep": []}
    for ntab in (1, 2, 3, 4, 5, 8):
        t = np.zeros((ntab, 8), np.float32)
        sinc.ref_sinc_table(ntab, 8, C.c_void_p(t.ctypes.data))
        out["sinc"][str(ntab)] = t.tolist()
    cg.ref_cgstep_solve.restype = C.c_int
    for seed, (nd, nm, niter) in ((6, (30, 10, 8)), (7, (48, 16, 12)), (8, (20, 20, 6))):
        r = np.random.default_rng(seed)
        A = r.standard_normal((nd, nm)).astype(np.float32)
        d = r.standard_normal(nd).astype(np.float32)
        m = np.zeros(nm, np.float32)
        res = np.zeros(niter)
        n = cg.ref_cgstep_solve(nd, nm, C.c_void_p(A.ctypes.data), C.c_void_p(d.ctypes.data),
                                niter, C.c_void_p(m.ctypes.data), C.c_void_p(res.ctypes.data))
        assert n == niter
        out["cgstep"].append({"A": A.tolist(), "d": d.tolist(), "niter": niter,
                              "m": m.tolist(), "resid": res.tolist()})
    path = os.path.join(HERE, "ref_outputs.json")
    with open(path, "w") as f:
        json.dump(out, f)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
