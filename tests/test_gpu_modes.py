"""GPU parity for a boson mode next to the qubits (SingleModeIndex, Indices,
modeOp: index.hpp:94-184, operators/mode.hpp) through the C ABI: composite
sector tables and ranks bit-exact, `>>` arrays bit-identical to the oracle,
a ground state against the Jaynes-Cummings closed form."""
import numpy as np
import pytest

import oracle
from qudrip_amd import abi
from qudrip_amd import operators as Q
from qudrip_amd.runtime import CSR, Basis, Krylov, LinOp, TermProgram
from tests.test_oracle import jaynes_cummings

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def holstein(N, nd, w, g):
    return Q.holstein_chain(N, nd, w, g)


def _sector(case):
    if case == "jc9":
        idx, ham = jaynes_cummings(9, 1.0, 0.83, 0.37)
        return Q.Constrain(idx), ham
    if case == "holstein10_nd8":
        q, idx, ham = holstein(10, 8, 0.9, 0.45)
        return Q.Constrain(idx, Q.getNset(q) == 5), ham
    if case == "two_qubits_nonherm":
        q = Q.getQbits(2)
        md = Q.SingleModeIndex(3)
        idx = q * md
        gc, ga = Q.getBoseGate(q), Q.getBoseGate(q)
        gc << [[0, 1], [0, 0]]
        ga << [[0, 0], [1, 0]]
        ad = Q.getModeOp(md)
        ad << Q.LadderMatrix(3, Q.MODE_LOWER)
        ham = 0.7 * (gc(1) * ga(0)) + 0.7 * (gc(0) * ga(1)) + (0.25 - 0.5j) * ad(0)
        return Q.Constrain(idx, Q.getNset(q) == 1), ham
    if case == "hub4_nd1_mode4":  # table-ranked qubit sector (double occupancy fixed) times 4 levels
        return _hub4_mode4()
    # no mode gate at all: the program acts as the identity on the mode
    q = Q.getQbits(6)
    md = Q.SingleModeIndex(5)
    _, ham = Q.xxz_chain(6, 0.8, 1.3, pbc=True)
    return Q.Constrain(q * md, Q.getNset(q) == 3), ham


def _hub4_mode4():
    L = 4
    q = Q.getQbits(2 * L)
    md = Q.SingleModeIndex(4)
    idx = q * md
    _, hub = Q.hubbard(L, Q.hubbard_bonds_chain(L, False), 1.0, 3.0)
    a, ad = Q.getModeOp(md), Q.getModeOp(md)
    a << Q.LadderMatrix(4, Q.MODE_UPPER)
    ad << Q.LadderMatrix(4, Q.MODE_LOWER)
    ham = hub + 0.3 * a(0) + 0.3 * ad(0)
    return Q.Constrain(idx, Q.getNup(q, L) == 2, Q.getNdo(q, L) == 2, Q.getNd(q, L) == 1), ham


CASES = ["jc9", "holstein10_nd8", "two_qubits_nonherm", "hub4_nd1_mode4", "qubits_only_on_mode"]


@pytest.mark.parametrize("case", CASES)
def test_modes_basis_and_csr_bit_exact(ctx, case):
    sec, ham = _sector(case)
    ref_basis = oracle.basis_scan(sec)
    r_outer, r_inner, r_vals = oracle.build_csc(ref_basis, ham, mdim=sec.mdim)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    c = CSR(ctx, b, p)
    try:
        assert b.size == len(ref_basis)
        assert np.array_equal(b.states(), ref_basis)
        assert np.array_equal(b.rank(ref_basis), np.arange(len(ref_basis)))
        outer, inner, vals = c.download()
        assert np.array_equal(outer, r_outer)
        assert np.array_equal(inner, r_inner)
        assert np.array_equal(bits(vals), bits(r_vals))
        assert c.hermitian == (0 if case == "two_qubits_nonherm" else 1)
        # the product y = H x (the non-Hermitian case goes through the adjoint program)
        op = LinOp.from_csr(ctx, c)
        x = oracle.seeded_state(b.size, 11)
        import torch
        y = op.apply(torch.from_numpy(x).to(f"cuda:{ctx.device}")).cpu().numpy()
        yr = oracle.spmv((r_outer, r_inner, r_vals), x)
        assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-14
        op.close()
    finally:
        c.close()
        p.close()
        b.close()


def test_modes_ground_state_jaynes_cummings(ctx):
    nd, w, D, g = 12, 1.0, 0.83, 0.37
    idx, ham = jaynes_cummings(nd, w, D, g)
    sec = Q.Constrain(idx)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    c = CSR(ctx, b, p)
    op = LinOp.from_csr(ctx, c)
    kr = Krylov(ctx, b.size, 2 * nd, herm=True)
    try:
        ref = [0.0, w * (nd - 1) + D]
        for N in range(1, nd):
            ref += list(np.linalg.eigvalsh(np.array([[w * N, g * np.sqrt(N)], [g * np.sqrt(N), w * (N - 1) + D]])))
        e0 = kr.ground_state(op, ctx.seeded_state(b.size, 3))
        assert abs(e0 - min(ref)) < 1e-12
    finally:
        for o in (kr, op, c, p, b):
            o.close()


def test_modes_matrix_free_is_refused(ctx):
    idx, ham = jaynes_cummings(4, 1.0, 0.5, 0.2)
    sec = Q.Constrain(idx)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    try:
        with pytest.raises(abi.QDError) as e:
            LinOp.matfree(ctx, b, p)
        assert abi.STATUS[e.value.code] == "QD_E_UNSUPPORTED"
    finally:
        p.close()
        b.close()


def test_modes_holstein_full_size_matches_kron(ctx):
    """XY chain of 16 qubits (8 up, 12,870 states) times a 16-level mode
    (205,920 composite states): the `>>` matrix built with mode gates equals
    H_q (x) I + I (x) w N + g (sum_i n_i) (x) (a + a+), assembled with
    scipy.sparse from the GPU's own qubit matrix; products agree."""
    import scipy.sparse as sp
    import torch
    N, nd, w, g = 16, 16, 0.9, 0.45
    q, idx, ham = holstein(N, nd, w, g)
    sec = Q.Constrain(idx, Q.getNset(q) == N // 2)
    _, hq = Q.xy_chain(N)
    qsec = Q.Constrain(q, Q.getNset(q) == N // 2)
    bq, pq, cq = Basis(ctx, qsec), TermProgram(ctx, hq, N), None
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, N)
    c = CSR(ctx, b, p)
    try:
        cq = CSR(ctx, bq, pq)
        qo, qi, qv = cq.download()
        nq = bq.size
        Hq = sp.csc_matrix((qv.view(np.complex128), qi, qo), shape=(nq, nq))
        occ = np.array([bin(int(x)).count("1") for x in bq.states()], dtype=float)  # sum_i n_i = 8
        a = sp.diags(np.sqrt(np.arange(1, nd)), 1).astype(np.complex128)
        ref = (sp.kron(Hq, sp.identity(nd)) + sp.kron(sp.identity(nq), w * (a.T @ a))
               + g * sp.kron(sp.diags(occ), a + a.T)).tocsc()
        o, i, v = c.download()
        H = sp.csc_matrix((v.view(np.complex128), i, o), shape=(b.size, b.size))
        assert b.size == nq * nd
        assert abs(H - ref).max() < 1e-14
        op = LinOp.from_csr(ctx, c)
        x = np.random.default_rng(2).standard_normal(b.size) + 0j
        y = op.apply(torch.from_numpy(x).to(f"cuda:{ctx.device}")).cpu().numpy()
        yr = ref @ x
        assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-13
        op.close()
    finally:
        for o_ in (c, p, b, cq, pq, bq):
            if o_ is not None:
                o_.close()
