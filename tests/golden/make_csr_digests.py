"""Writes tests/golden/csr_digests.json: SHA-256 digests of the oracle's CSC
arrays (outer int32, inner int32, values as complex128 bit patterns) for the
sectors the GPU parity tests build at full size.

The oracle (oracle/oracle.c, the CPU restatement of operator.hpp:135-267 and
operator_abel.hpp:25-34) needs about a minute for XXZ N=24, too long to run
inside a GPU test, so the GPU test compares digests of its download with
these.  Run from the repository root:  python tests/golden/make_csr_digests.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from qudrip_amd import operators as Q  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csr_digests.json")


def cases():
    """name -> (sector, operator sum); the same constructions as the GPU test."""
    out = {}
    idx, ham = Q.xxz_chain(24, 1.0, 1.0, True)
    out["xxz24"] = (Q.Constrain(idx, Q.getNset(idx) == 12), ham)
    idx, ham = Q.xxz_chain(22, 1.0, 0.7317, True)
    out["xxz22_jz0.7317"] = (Q.Constrain(idx, Q.getNset(idx) == 11), ham)
    bonds = Q.hubbard_bonds_square(3, 4)
    idx, ham = Q.hubbard(12, bonds, 1.0, 4.0)
    out["hub3x4_66"] = (Q.hubbard_sector(idx, 12, 6, 6), ham)
    idx, ham = Q.hubbard(12, bonds, 1.0, 5.123)
    out["hub3x4_66_u5.123"] = (Q.hubbard_sector(idx, 12, 6, 6), ham)
    return out


def digest(outer, inner, vals):
    h = {}
    for name, a in (("outer", np.ascontiguousarray(outer, dtype="<i4")),
                    ("inner", np.ascontiguousarray(inner, dtype="<i4")),
                    ("vals", np.ascontiguousarray(vals, dtype=np.complex128).view("<u8"))):
        h[name] = hashlib.sha256(a.tobytes()).hexdigest()
    return h


def main():
    res = {}
    for name, (sec, ham) in cases().items():
        t = time.time()
        b = oracle.basis_scan(sec)
        outer, inner, vals = oracle.build_csc(b, ham)
        res[name] = {"n": int(len(b)), "nnz": int(len(inner)), **digest(outer, inner, vals),
                     "basis": hashlib.sha256(np.ascontiguousarray(b, dtype="<u8").tobytes()).hexdigest()}
        print(f"{name}: n={len(b)} nnz={len(inner)} ({time.time() - t:.1f} s)", flush=True)
    with open(OUT, "w") as f:
        json.dump(res, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
