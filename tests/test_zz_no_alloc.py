"""qd_expect_noalloc (the expectation value that allocates nothing) through
the C ABI against qd_expect (State::eval, state.hpp:33-44) on the same
operator and state: two calls in a row, the second on a smaller operator
(the file-scope buffer keeps the first call's contents), each within 1e-12
of qd_expect (the energies' bar is 1e-10 relative)."""
import ctypes

import numpy as np
import pytest

import oracle
from qudrip_amd import abi
from qudrip_amd import operators as Q
from qudrip_amd.runtime import CSR, Basis, LinOp, TermProgram

pytestmark = pytest.mark.gpu


def _noalloc(ctx, op, psi):
    import torch
    out = torch.full((2,), float("nan"), dtype=torch.float64, device=psi.device)
    ctx.check(ctx.lib.qd_expect_noalloc(ctx.ptr, op.ptr, ctypes.c_void_p(psi.data_ptr()),
                                        ctypes.c_void_p(out.data_ptr())), "qd_expect_noalloc")
    ctx.sync()
    h = out.cpu().numpy()
    return complex(h[0], h[1])


def test_expect_noalloc_matches_expect_on_two_calls(ctx):
    for N, jz, seed in ((20, 0.83, 11), (12, 1.7, 12)):
        print(f"[no-alloc] XXZ N={N} Jz={jz}", flush=True)
        idx, ham = Q.xxz_chain(N, 1.0, jz, pbc=True)
        sec = Q.Constrain(idx, Q.getNset(idx) == N // 2)
        b = Basis(ctx, sec)
        p = TermProgram(ctx, ham, N)
        c = CSR(ctx, b, p)
        op = LinOp.from_csr(ctx, c)
        try:
            psi = ctx.seeded_state(b.size, seed)
            e_ref = op.expect(psi)
            e = _noalloc(ctx, op, psi)
            assert abs(e - e_ref) <= 1e-12 * abs(e_ref), (N, e, e_ref)
            if N == 12:  # small enough for the oracle as a third opinion
                ref = oracle.eval_expect(oracle.build_csc(oracle.basis_scan(sec), ham), oracle.seeded_state(b.size, seed))
                assert abs(e - ref) <= 1e-12 * abs(ref)
        finally:
            for h in (op, c, p, b):
                h.close()


def test_expect_noalloc_rejects_other_operators(ctx):
    """A matrix-free operator is refused with a status (nothing launched)."""
    idx, ham = Q.xxz_chain(10, 1.0, 1.0, pbc=True)
    sec = Q.Constrain(idx, Q.getNset(idx) == 5)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, 10)
    op = LinOp.matfree(ctx, b, p)
    try:
        psi = ctx.seeded_state(b.size, 1)
        with pytest.raises(abi.QDError):
            _noalloc(ctx, op, psi)
    finally:
        for h in (op, p, b):
            h.close()
