"""Parity at the configured sizes (the largest sectors of the benchmark
configurations), through the C ABI:

* CSR arrays of XXZ N=24 (config 2), XXZ N=22 with a non-dyadic Jz, and the
  3x4 Hubbard 6/6 sector (config 5, U = 4 and a non-dyadic U) bit for bit
  against the oracle, by SHA-256 digests of the arrays the oracle builds
  (tests/golden/make_csr_digests.py writes them: the oracle needs about a
  minute for XXZ N=24);
* the 4x4 Hubbard (5,5) sector (config 3, 19,079,424 states) state by state
  against the construction the mathematics fixes: x = (dn << 16) | up over
  the two ascending tables of 16-bit words with 5 bits set, dn major;
* the delayed-CGS2 Lanczos ground state (Arnoldi::getLanczosSolver, k=150,
  config 2) against the Bethe-ansatz energy.
"""
import hashlib
import itertools
import json
import os

import numpy as np
import pytest

from qudrip_amd import operators as Q
from qudrip_amd.runtime import CSR, Basis, Krylov, LinOp, TermProgram
from tests import physics as P

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "csr_digests.json")
E0_XXX24 = -10.67001451653724  # Bethe ansatz, 12 real roots


def _cases():
    bonds = Q.hubbard_bonds_square(3, 4)
    return {
        "xxz24": lambda: _xxz(24, 1.0),
        "xxz22_jz0.7317": lambda: _xxz(22, 0.7317),
        "hub3x4_66": lambda: _hub(bonds, 4.0),
        "hub3x4_66_u5.123": lambda: _hub(bonds, 5.123),
    }


def _xxz(N, jz):
    idx, ham = Q.xxz_chain(N, 1.0, jz, True)
    return Q.Constrain(idx, Q.getNset(idx) == N // 2), ham, N


def _hub(bonds, U):
    idx, ham = Q.hubbard(12, bonds, 1.0, U)
    return Q.hubbard_sector(idx, 12, 6, 6), ham, 24


def _sha(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(a).astype(dtype, copy=False).tobytes()).hexdigest()


@pytest.mark.parametrize("name", list(_cases()))
def test_csr_full_size_matches_oracle_digest(ctx, name):
    with open(GOLDEN) as f:
        ref = json.load(f)[name]
    sec, ham, nbits = _cases()[name]()
    print(f"[full-size CSR] {name}", flush=True)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, nbits)
    c = CSR(ctx, b, p)
    try:
        assert b.size == ref["n"]
        assert _sha(b.states(), "<u8") == ref["basis"]
        outer, inner, vals = c.download()
        assert len(inner) == ref["nnz"]
        assert _sha(outer, "<i4") == ref["outer"]
        assert _sha(inner, "<i4") == ref["inner"]
        assert hashlib.sha256(np.ascontiguousarray(vals, dtype=np.complex128).view("<u8").tobytes()).hexdigest() \
            == ref["vals"]
    finally:
        c.close(); p.close(); b.close()


def test_hub4x4_55_basis_state_by_state(ctx):
    """Constrain(fermi, nup = 5, ndo = 5) on 32 bits (conservation.hpp:212-231):
    ascending x, i.e. (dn << 16) | up with dn major, both ascending."""
    idx = Q.getQbits(32)
    sec = Q.hubbard_sector(idx, 16, 5, 5)
    words = np.array(sorted(sum(1 << i for i in c) for c in itertools.combinations(range(16), 5)), dtype=np.uint64)
    assert len(words) == 4368
    expect = ((words[:, None] << np.uint64(16)) | words[None, :]).reshape(-1)
    b = Basis(ctx, sec)
    try:
        assert b.size == 19079424
        got = b.states()
        assert got.dtype == np.uint64 and np.array_equal(got, expect)
        # rank of every 997th state, and a miss (6 up bits) ranks to range()
        sample = expect[::997]
        assert np.array_equal(b.rank(sample), np.arange(0, len(expect), 997))
        miss = np.array([(int(words[0]) << 16) | 0b111111], dtype=np.uint64)
        assert b.rank(miss)[0] == len(expect)
    finally:
        b.close()


def test_lanczos_xxz24_k150_delayed_cgs2_bethe(ctx):
    """Arnoldi::getLanczosSolver(chain, 150).getGroundState (sparse_algorithm.hpp:564-585)
    with full reorthogonalisation (delayed CGS2) at n = 2,704,156."""
    import torch
    sec, ham, nbits = _xxz(24, 1.0)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, nbits)
    c = CSR(ctx, b, p)
    op = LinOp.from_csr(ctx, c)
    kr = Krylov(ctx, b.size, 150, herm=True)
    try:
        kr.set_method(1)
        psi = ctx.seeded_state(b.size, 2)
        e0 = kr.ground_state(op, psi)
        info = kr.info()
        assert info["act_dim"] == 150
        assert abs(e0 - E0_XXX24) / abs(E0_XXX24) < 1e-10, (e0, E0_XXX24)
        # the Ritz vector: unit norm, Rayleigh quotient = E0, small residual
        nrm = ctx.norm2(psi)
        assert abs(nrm - 1.0) < 1e-10
        hpsi = op.apply(psi)
        ray = torch.vdot(psi, hpsi).real.item()
        assert abs(ray - E0_XXX24) / abs(E0_XXX24) < 1e-10
        res = torch.linalg.vector_norm(hpsi - e0 * psi).item()
        assert res < 1e-4, res
    finally:
        kr.close(); op.close(); c.close(); p.close(); b.close()
