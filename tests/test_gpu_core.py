"""GPU parity of the core path through the C ABI against the CPU oracle:
basis (bit-exact), CSR build (bit-exact), products, Lanczos, Arnoldi."""
import numpy as np
import pytest

import oracle
from qudrip_amd import abi
from qudrip_amd import operators as Q
from qudrip_amd.runtime import CSR, Basis, Krylov, LinOp, TermProgram
from tests import physics as P

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def xxz(N, Jxy=1.0, Jz=1.0, pbc=True):
    idx, ham = Q.xxz_chain(N, Jxy, Jz, pbc)
    return idx, ham, Q.Constrain(idx, Q.getNset(idx) == N // 2)


def hub(nsite, bonds, nup, ndn, t=1.0, U=4.0):
    idx, ham = Q.hubbard(nsite, bonds, t, U)
    return idx, ham, Q.hubbard_sector(idx, nsite, nup, ndn)


def to_dev(ctx, x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).to(f"cuda:{ctx.device}")


# ------------------------------------------------------------------ basis
@pytest.mark.parametrize("case", ["xy16", "xxz20", "xxz24", "hub3x4_66", "hub8_44", "hub8_31", "free_hi_bits"])
def test_basis_bit_exact(ctx, case):
    if case == "xy16":
        idx = Q.getQbits(16); sec = Q.Constrain(idx, Q.getNset(idx) == 8)
    elif case == "xxz20":
        idx = Q.getQbits(20); sec = Q.Constrain(idx, Q.getNset(idx) == 10)
    elif case == "xxz24":
        idx = Q.getQbits(24); sec = Q.Constrain(idx, Q.getNset(idx) == 12)
    elif case == "hub3x4_66":
        idx = Q.getQbits(24); sec = Q.hubbard_sector(idx, 12, 6, 6)
    elif case == "hub8_44":
        idx = Q.getQbits(16); sec = Q.hubbard_sector(idx, 8, 4, 4)
    elif case == "hub8_31":
        idx = Q.getQbits(16); sec = Q.hubbard_sector(idx, 8, 3, 1)
    else:  # constraint on the low bits only: the high bits are free
        idx = Q.getQbits(14); sec = Q.Constrain(idx, Q.getNup(idx, 6) == 2)
    ref = oracle.basis_scan(sec)
    b = Basis(ctx, sec)
    try:
        assert b.size == len(ref)
        assert np.array_equal(b.states(), ref)
        sample = ref[:: max(1, len(ref) // 5000)]
        assert np.array_equal(b.rank(sample), np.arange(len(ref))[:: max(1, len(ref) // 5000)])
    finally:
        b.close()


def test_basis_rank_miss_and_general_path(ctx):
    idx = Q.getQbits(12)
    # overlapping masks (total count + up count) -> predicate-scan path
    sec = Q.Constrain(idx, Q.getNset(idx) == 6, Q.getNup(idx, 6) == 2)
    ref = oracle.basis_scan(sec)
    b = Basis(ctx, sec)
    try:
        assert np.array_equal(b.states(), ref)
        out = np.array([0b111, int(ref[5]), (1 << 12) | int(ref[0])], dtype=np.uint64)
        r = b.rank(out)
        assert r[0] == len(ref) and r[1] == 5 and r[2] == len(ref)
    finally:
        b.close()


@pytest.mark.parametrize("case", ["hub8_44_nd2", "hub6_33_nh1", "hub6_nd1_scan", "hub7_43_nd1_nh2", "hub6_empty"])
def test_basis_nd_nh_bit_exact(ctx, case):
    """getNd / getNh sectors: the Nset sector enumerated then filtered in
    order (or a scan when there is no Nset part); rank by binary search."""
    L = {"hub8_44_nd2": 8, "hub7_43_nd1_nh2": 7}.get(case, 6)
    idx = Q.getQbits(2 * L)
    nup, ndo, nd, nh = Q.getNup(idx, L), Q.getNdo(idx, L), Q.getNd(idx, L), Q.getNh(idx, L)
    cons = {"hub8_44_nd2": [nup == 4, ndo == 4, nd == 2],
            "hub6_33_nh1": [nup == 3, ndo == 3, nh == 1],
            "hub6_nd1_scan": [nd == 1],
            "hub7_43_nd1_nh2": [nup == 4, ndo == 3, nd == 1, nh == 2],
            "hub6_empty": [nup == 2, ndo == 2, nd == 3]}[case]
    sec = Q.Constrain(idx, *cons)
    ref = oracle.basis_scan(sec)
    b = Basis(ctx, sec)
    try:
        assert b.size == len(ref)
        assert np.array_equal(b.states(), ref)
        if len(ref):
            assert np.array_equal(b.rank(ref), np.arange(len(ref)))
            miss = np.array([int(ref[0]) ^ 1, (1 << (2 * L)) | int(ref[-1])], dtype=np.uint64)
            r = b.rank(miss)
            assert r[1] == len(ref) and (r[0] == len(ref) or ref[r[0]] == miss[0])
    finally:
        b.close()


def test_basis_empty_sector(ctx):
    idx = Q.getQbits(6)
    b = Basis(ctx, Q.Constrain(idx, Q.getNset(idx) == 9))
    assert b.size == 0
    b.close()


# ------------------------------------------------------------------ CSR
def _build(ctx, ham, sec):
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    c = CSR(ctx, b, p)
    return b, p, c


def _close(*hs):
    for h in hs:
        h.close()


def _xy_all_to_all(N, seed):
    """sum_{i<j} J_ij (S+_i S-_j + h.c.) with non-dyadic J_ij: at half filling a
    column has (N/2)^2 off-diagonal targets, above the single-expansion build's
    32 (its count-then-fill fallback runs)."""
    rng = np.random.default_rng(seed)
    idx = Q.getQbits(N)
    gc, ga = Q.getBoseGate(idx), Q.getBoseGate(idx)
    gc << 0.5 * (Q.pauli_x + Q.II * Q.pauli_y)
    ga << 0.5 * (Q.pauli_x - Q.II * Q.pauli_y)
    ham = Q.getEmptyOperator()
    for i in range(N):
        for j in range(i + 1, N):
            ham = ham + float(rng.uniform(0.3, 1.7)) * (gc(j) * ga(i) + gc(i) * ga(j))
    return ham, Q.Constrain(idx, Q.getNset(idx) == N // 2)


CSR_CASES = {
    "xy12_all_to_all": lambda: _xy_all_to_all(12, 3),
    "xy16_obc": lambda: (lambda i, h: (h, Q.Constrain(i, Q.getNset(i) == 8)))(*Q.xy_chain(16, pbc=False)),
    "xxz16_pbc": lambda: xxz(16)[1:],
    "xxz20_pbc": lambda: xxz(20)[1:],
    "xxz18_random_coeffs": lambda: xxz(18, 0.9317, 0.3711)[1:],
    "hub8_obc_44": lambda: hub(8, Q.hubbard_bonds_chain(8, False), 4, 4)[1:],
    "hub8_obc_random_U": lambda: hub(8, Q.hubbard_bonds_chain(8, False), 3, 2, t=0.77, U=5.123)[1:],
    "hub3x4_33": lambda: hub(12, Q.hubbard_bonds_square(3, 4), 3, 3)[1:],
    # doubly occupied sites fixed (getNd): table-ranked sector, hops that
    # change the count leave it and are dropped, as `>>` drops any miss
    "hub8_obc_44_nd2": lambda: (lambda i, h: (h, Q.Constrain(i, Q.getNup(i, 8) == 4, Q.getNdo(i, 8) == 4,
                                                             Q.getNd(i, 8) == 2)))(
        *Q.hubbard(8, Q.hubbard_bonds_chain(8, False), 1.0, 4.0)),
}


@pytest.mark.parametrize("case", list(CSR_CASES))
def test_csr_bit_exact(ctx, case):
    ham, sec = CSR_CASES[case]()
    ref_basis = oracle.basis_scan(sec)
    r_outer, r_inner, r_vals = oracle.build_csc(ref_basis, ham)
    b, p, c = _build(ctx, ham, sec)
    try:
        outer, inner, vals = c.download()
        assert c.nnz == len(r_inner)
        assert np.array_equal(outer, r_outer)
        assert np.array_equal(inner, r_inner)
        assert np.array_equal(bits(vals), bits(r_vals))  # bit-identical, signed zeros included
        assert c.hermitian == 1
    finally:
        _close(c, p, b)


@pytest.mark.parametrize("N,nnz", [(22, 8834696), (24, 36564892)])
def test_csr_xxz_full_size_nnz_and_symmetry(ctx, N, nnz):
    _, ham, sec = xxz(N)
    b, p, c = _build(ctx, ham, sec)
    try:
        assert c.nnz == nnz and c.hermitian == 1
        outer, inner, vals = c.download()
        v = vals.view(np.complex128)
        assert np.all(v.imag == 0)
        # values are Jxy/2 = 0.5 off the diagonal, multiples of 0.25 on it
        assert set(np.unique(np.round(v.real * 4)).astype(int)) <= set(range(-24, 25))
        d = np.diff(inner.astype(np.int64))
        boundary = np.zeros(len(d), dtype=bool)
        st = outer[1:-1].astype(np.int64)
        st = st[(st > 0) & (st < len(inner))]
        boundary[st - 1] = True
        assert np.all((d > 0) | boundary)  # rows ascending inside every column
    finally:
        _close(c, p, b)


def test_csr_hub3x4_66_nnz(ctx):
    ham, sec = hub(12, Q.hubbard_bonds_square(3, 4), 6, 6)[1:]
    b, p, c = _build(ctx, ham, sec)
    try:
        assert b.size == 853776 and c.nnz == 23206260
    finally:
        _close(c, p, b)


def test_csr_non_hermitian_operator(ctx):
    """A lone hop c+_3 c_1 (not Hermitian): exact CSC and the adjoint-built product."""
    idx = Q.getQbits(8)
    g = Q.getHopFermiGate(idx)
    ham = Q.getEmptyOperator() + (0.5 + 0.25j) * g(3, 1) + 0.3 * (g(5, 5) * g(2, 2))
    sec = Q.Constrain(idx, Q.getNset(idx) == 4)
    ref_basis = oracle.basis_scan(sec)
    ref = oracle.build_csc(ref_basis, ham)
    b, p, c = _build(ctx, ham, sec)
    try:
        outer, inner, vals = c.download()
        assert np.array_equal(outer, ref[0]) and np.array_equal(inner, ref[1])
        assert np.array_equal(bits(vals), bits(ref[2]))
        assert c.hermitian == 0
        op = LinOp.from_csr(ctx, c)
        x = oracle.seeded_state(len(ref_basis), 9)
        y = op.apply(to_dev(ctx, x)).cpu().numpy()
        assert np.allclose(y, oracle.spmv(ref, x), atol=1e-14)
        op.close()
    finally:
        _close(c, p, b)


# ------------------------------------------------------------------ products
@pytest.mark.parametrize("case", ["xxz20_pbc", "xxz18_random_coeffs", "hub3x4_33"])
def test_spmv_matches_oracle(ctx, case):
    ham, sec = CSR_CASES[case]()
    ref_basis = oracle.basis_scan(sec)
    ref = oracle.build_csc(ref_basis, ham)
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    mf = LinOp.matfree(ctx, b, p)
    try:
        x = oracle.seeded_state(len(ref_basis), 21)
        yr = oracle.spmv(ref, x)
        xd = to_dev(ctx, x)
        y = op.apply(xd).cpu().numpy()
        ym = mf.apply(xd).cpu().numpy()
        scale = np.linalg.norm(yr)
        assert np.linalg.norm(y - yr) / scale < 1e-14
        assert np.linalg.norm(ym - yr) / scale < 1e-14
        e = op.expect(xd)
        er = oracle.eval_expect(ref, x)
        assert abs(e - er) / abs(er) < 1e-13
        for lanes in (*range(0, -15, -1), 2, 4, 8, 16, 32, 64):
            op.set_lanes(lanes)
            assert np.linalg.norm(op.apply(xd).cpu().numpy() - yr) / scale < 1e-14
    finally:
        _close(mf, op, c, p, b)


def test_matfree_matches_map_acc(ctx):
    """Lazy operator vs the reference's push-form map_acc (operator.hpp:215-232)."""
    ham, sec = hub(8, Q.hubbard_bonds_chain(8, False), 4, 4)[1:]
    ref_basis = oracle.basis_scan(sec)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    mf = LinOp.matfree(ctx, b, p)
    try:
        x = oracle.seeded_state(len(ref_basis), 13)
        to = np.zeros(len(ref_basis), dtype=np.complex128)
        oracle.map_acc(ref_basis, ham, x, to)
        y = mf.apply(to_dev(ctx, x)).cpu().numpy()
        assert np.linalg.norm(y - to) / np.linalg.norm(to) < 1e-14
    finally:
        _close(mf, p, b)


def test_seeded_state_matches_oracle(ctx):
    v = ctx.seeded_state(100003, 2).cpu().numpy()
    r = oracle.seeded_state(100003, 2)
    assert np.max(np.abs(v - r)) < 1e-15
    assert abs(ctx.norm2(to_dev(ctx, r)) - 1.0) < 1e-14


# ------------------------------------------------------------------ solvers
def test_lanczos_xy16_closed_form(ctx):
    """spin.cpp sequence: XY16 OBC, N_up = 8, k = 100."""
    ham, sec = CSR_CASES["xy16_obc"]()
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    kr = Krylov(ctx, b.size, 100, herm=True)
    try:
        psi = to_dev(ctx, oracle.seeded_state(b.size, 1))
        e0 = kr.ground_state(op, psi)
        e_ref = P.xy_free_fermion_e0(16, 8)
        assert abs(e0 - e_ref) / abs(e_ref) < 1e-10
        ev = op.expect(psi)
        nrm = ctx.norm2(psi)
        assert abs(ev.real / nrm - e_ref) / abs(e_ref) < 1e-10
        assert abs(nrm - 1.0) < 1e-10  # full reorthogonalisation keeps V orthonormal
        inf = kr.info()
        assert inf["act_dim"] == 100 and inf["breakdown"] == 0
    finally:
        _close(kr, op, c, p, b)


def test_lanczos_matches_oracle_and_dense(ctx):
    ham, sec = xxz(12, 1.0, 0.6)[1:]
    ref_basis = oracle.basis_scan(sec)
    ref = oracle.build_csc(ref_basis, ham)
    w = np.linalg.eigvalsh(oracle.csc_to_dense(ref))
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    kr = Krylov(ctx, b.size, 120, herm=True)
    try:
        psi0 = oracle.seeded_state(b.size, 2)
        r = oracle.lanczos_gs(ref, 120, psi0)
        psi = to_dev(ctx, psi0)
        e0 = kr.ground_state(op, psi)
        assert abs(e0 - w[0]) / abs(w[0]) < 1e-10
        assert abs(e0 - r["E0"]) / abs(w[0]) < 1e-10
        g = psi.cpu().numpy()
        ov = abs(np.vdot(g, r["psi"])) / (np.linalg.norm(g) * np.linalg.norm(r["psi"]))
        assert ov > 1 - 1e-9
    finally:
        _close(kr, op, c, p, b)


@pytest.mark.parametrize("method", [1, 2])
def test_lanczos_incremental_and_breakdown(ctx, method):
    """getGroundState with add_dim in two calls, then an eigenvector start
    (breakdown at the first step); method 2 carries its orthogonality
    estimate across the calls."""
    ham, sec = xxz(10, 1.0, 1.0)[1:]
    ref_basis = oracle.basis_scan(sec)
    H = oracle.csc_to_dense(oracle.build_csc(ref_basis, ham))
    w, z = np.linalg.eigh(H)
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    kr = Krylov(ctx, b.size, 60, herm=True)
    kr.set_method(method)
    try:
        psi = to_dev(ctx, oracle.seeded_state(b.size, 3))
        e_a = kr.ground_state(op, psi, 20)  # getGroundState(H, psi, 0, 20)
        assert kr.info()["act_dim"] == 20
        e_b = kr.ground_state(op, psi, 40)
        assert kr.info()["act_dim"] == 60
        assert e_b <= e_a + 1e-12 and abs(e_b - w[0]) < 1e-9 * abs(w[0])
        kr.reset()
        psi2 = to_dev(ctx, z[:, 2])  # eigenvector start: breakdown at the first step
        e_c = kr.ground_state(op, psi2)
        inf = kr.info()
        assert inf["breakdown"] == 1 and inf["act_dim"] == 1
        assert abs(e_c - w[2]) < 1e-10
    finally:
        _close(kr, op, c, p, b)


def test_arnoldi_evolve_matches_oracle(ctx):
    ham, sec = hub(8, Q.hubbard_bonds_chain(8, False), 3, 3, U=2.0)[1:]
    ref_basis = oracle.basis_scan(sec)
    ref = oracle.build_csc(ref_basis, ham)
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    mf = LinOp.matfree(ctx, b, p)
    kr = Krylov(ctx, b.size, 30, herm=False)
    try:
        psi0 = oracle.seeded_state(b.size, 4)
        r = oracle.arnoldi_evolve(ref, 30, 0.05, psi0)
        xin = to_dev(ctx, psi0)
        out = ctx.zeros(b.size)
        kr.evolve(op, xin, out, 0.05, 30)
        assert np.max(np.abs(out.cpu().numpy() - r)) < 1e-11
        kr.evolve(mf, xin, out, 0.05, 30)
        assert np.max(np.abs(out.cpu().numpy() - r)) < 1e-11
        # forward then backward returns to the start; the norm is kept
        back = ctx.zeros(b.size)
        kr.evolve(op, out, back, -0.05, 30)
        assert np.max(np.abs(back.cpu().numpy() - psi0)) < 1e-11
    finally:
        _close(kr, mf, op, c, p, b)


def test_evolve_tol_matches_dense_propagator(ctx):
    """Adaptive Krylov dimension: the step agrees with the dense exponential
    to the requested bound, and a loose bound picks a smaller basis."""
    import scipy.linalg
    ham, sec = hub(8, Q.hubbard_bonds_chain(8, False), 3, 3, U=6.0)[1:]
    ref_basis = oracle.basis_scan(sec)
    colptr, rowidx, vals = oracle.build_csc(ref_basis, ham)
    n = len(ref_basis)
    H = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        H[rowidx[colptr[j]:colptr[j + 1]], j] = vals.view(np.complex128)[colptr[j]:colptr[j + 1]]
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    kr = Krylov(ctx, b.size, 40, herm=False)
    try:
        psi0 = oracle.seeded_state(b.size, 9)
        exact = scipy.linalg.expm(-1j * 0.05 * H) @ psi0
        xin = to_dev(ctx, psi0)
        out = ctx.zeros(b.size)
        m_tight, me_tight = kr.evolve_tol(op, xin, out, 0.05, 8, 40, 1e-14)
        assert np.max(np.abs(out.cpu().numpy() - exact)) < 1e-12
        m_loose, me_loose = kr.evolve_tol(op, xin, out, 0.05, 8, 40, 1e-6)
        assert np.max(np.abs(out.cpu().numpy() - exact)) < 1e-5
        assert 8 <= m_loose < m_tight < 39
        assert me_tight <= m_tight and me_loose <= m_loose and me_loose < me_tight
        # starting at the suggested dimension reaches the bound again
        m2, _ = kr.evolve_tol(op, xin, out, 0.05, me_tight, 40, 1e-14)
        assert m2 <= m_tight and np.max(np.abs(out.cpu().numpy() - exact)) < 1e-12
    finally:
        _close(kr, op, c, p, b)


def _structured_case(case):
    if case == "chain8_44":
        idx, ham = Q.hubbard(8, Q.hubbard_bonds_chain(8, False), 1.0, 4.0)
        sec = Q.hubbard_sector(idx, 8, 4, 4)
    elif case == "sq3x4_33":
        idx, ham = Q.hubbard(12, Q.hubbard_bonds_square(3, 4), 1.0, 2.5)
        sec = Q.hubbard_sector(idx, 12, 3, 3)
    elif case == "sq3x4_66":
        idx, ham = Q.hubbard(12, Q.hubbard_bonds_square(3, 4), 1.0, 4.0)
        sec = Q.hubbard_sector(idx, 12, 6, 6)
    elif case == "chain6_nonherm":  # right-moving hops only, complex amplitude: not Hermitian
        idx = Q.getQbits(12)
        g = Q.getHopFermiGate(idx)
        ham = Q.getEmptyOperator()
        for i in range(5):
            for off in (0, 6):
                ham = ham + (0.3 - 0.7j) * g(i + 1 + off, i + off)
        for i in range(6):
            ham = ham + 1.7 * (g(i, i) * g(i + 6, i + 6))
        sec = Q.hubbard_sector(idx, 6, 3, 2)
    else:  # Peierls phases and a chemical potential (bit gates on the down block)
        idx = Q.getQbits(12)
        g = Q.getHopFermiGate(idx)
        nb = Q.getFermiGate(idx)
        nb << np.diag([0.0, 1.0])
        ph = np.exp(0.4j)
        ham = Q.getEmptyOperator()
        for (a, b) in Q.hubbard_bonds_chain(6, True):
            for off in (0, 6):
                ham = ham + (-ph) * g(a + off, b + off) + (-np.conj(ph)) * g(b + off, a + off)
        for i in range(6):
            ham = ham + 3.0 * (g(i, i) * g(i + 6, i + 6)) + (-0.4) * nb(i + 6)
        sec = Q.hubbard_sector(idx, 6, 3, 3)
    return ham, sec


@pytest.mark.parametrize("case", ["chain8_44", "sq3x4_33", "sq3x4_66", "chain6_nonherm", "phase_hops"])
def test_matfree_structured_matches_oracle(ctx, case):
    """Two-block Kronecker form of the matrix-free product (Hubbard layout)."""
    ham, sec = _structured_case(case)
    basis = oracle.basis_scan(sec)
    ref = oracle.build_csc(basis, ham)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    mf = LinOp.matfree(ctx, b, p)
    try:
        assert mf.set_structured(True)
        x = oracle.seeded_state(len(basis), 17)
        yr = oracle.spmv(ref, x)
        xd = to_dev(ctx, x)
        ys = mf.apply(xd).cpu().numpy()
        mf.set_structured(False)
        yg = mf.apply(xd).cpu().numpy()
        scale = np.linalg.norm(yr)
        assert np.linalg.norm(ys - yr) / scale < 1e-14
        assert np.linalg.norm(yg - yr) / scale < 1e-14
        # a diagonal view keeps the two-block form: (H + diag(d)) x, d complex
        rng = np.random.default_rng(5)
        d = rng.standard_normal(len(basis)) + 1j * rng.standard_normal(len(basis))
        yd = yr + d * x
        dd = to_dev(ctx, d)
        for structured in (True, False):
            mf.set_structured(structured)
            view = LinOp.with_diagonal(ctx, mf, dd)
            yv = view.apply(xd).cpu().numpy()
            view.close()
            assert np.linalg.norm(yv - yd) / np.linalg.norm(yd) < 1e-14, structured
    finally:
        _close(mf, p, b)


def _chain_two_amplitudes(nsite, nup, ndn, t1, t2, U):
    """Hubbard ring whose bonds alternate between hopping amplitudes t1 and
    t2: the block operators hold more than one value up to sign."""
    idx = Q.getQbits(2 * nsite)
    gate = Q.getHopFermiGate(idx)
    ham = Q.getEmptyOperator()
    for j, (a, b) in enumerate(Q.hubbard_bonds_chain(nsite, True)):
        t = t1 if j % 2 == 0 else t2
        for off in (0, nsite):
            ham = ham + (-t) * gate(a + off, b + off) + (-t) * gate(b + off, a + off)
    for i in range(nsite):
        ham = ham + U * (gate(i, i) * gate(i + nsite, i + nsite))
    return ham, Q.hubbard_sector(idx, nsite, nup, ndn)


@pytest.mark.parametrize("nup,ndn,amps", [(7, 1, 1), (1, 7, 1), (7, 1, 2)])
def test_matfree_structured_every_variant(ctx, nup, ndn, amps, monkeypatch):
    """Every rows-pass and slab-pass variant of the two-block product, fixed
    through QD_KRON_VARIANTS, against the oracle (rows variants 4 to 8
    split the 3432 columns into four groups of blocks): a 14-site ring with
    one block of 3432 states (more than 128 chunks of 8 rows, so the slab
    variants that size chunks to one XCD's resident blocks differ from the
    128-chunk ones; 54 slabs of 64 columns for slab variants 7 and 8) and
    one of 14 (one partial 64-column slab).  One hopping amplitude
    (amps = 1) puts rows variants 5-8 and slab variants 7-8 in their sign
    form (every block value +c or -c); two alternating amplitudes
    (amps = 2) in their value-table form."""
    if amps == 1:
        ham, sec = hub(14, Q.hubbard_bonds_chain(14, True), nup, ndn, U=3.0)[1:]
    else:
        ham, sec = _chain_two_amplitudes(14, nup, ndn, 1.0, 0.55, 3.0)
    basis = oracle.basis_scan(sec)
    ref = oracle.build_csc(basis, ham)
    x = oracle.seeded_state(len(basis), 23)
    yr = oracle.spmv(ref, x)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    xd = to_dev(ctx, x)
    ran = []
    try:
        for rv in range(9):
            for sv in range(9):
                monkeypatch.setenv("QD_KRON_VARIANTS", f"{rv},{sv}")
                mf = LinOp.matfree(ctx, b, p)
                try:
                    assert mf.set_structured(True)
                    try:
                        y = mf.apply(xd).cpu().numpy()
                    except abi.QDError as e:  # a variant whose LDS does not fit this shape
                        assert abi.STATUS[e.code] == "QD_E_UNSUPPORTED"
                        continue
                finally:
                    mf.close()
                assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-14, (rv, sv)
                ran.append((rv, sv))
    finally:
        _close(p, b)
    assert {s for r, s in ran} == set(range(9)) and (0, 0) in ran
    assert {4, 5, 6, 7, 8} <= {r for r, s in ran}  # the persistent rows kernels (T_A in registers)


@pytest.mark.parametrize("nup,ndn,amps", [(2, 3, 1), (3, 2, 1), (2, 3, 2)])
def test_matfree_fused_matches_oracle(ctx, nup, ndn, amps, monkeypatch):
    """The fused one-pass two-block product (QD_KRON_VARIANTS=15,15: T_A
    from the LDS-staged row, T_B from coalesced loads of the source rows, y
    written once) against the oracle, alone and under a complex diagonal
    view, on a 14-site ring with blocks of 91 and 364 states.  With two
    hopping amplitudes (amps = 2) the blocks are not in sign form and the
    fused product is refused with QD_E_UNSUPPORTED."""
    if amps == 1:
        ham, sec = hub(14, Q.hubbard_bonds_chain(14, True), nup, ndn, U=3.0)[1:]
    else:
        ham, sec = _chain_two_amplitudes(14, nup, ndn, 1.0, 0.55, 3.0)
    basis = oracle.basis_scan(sec)
    ref = oracle.build_csc(basis, ham)
    x = oracle.seeded_state(len(basis), 29)
    yr = oracle.spmv(ref, x)
    rng = np.random.default_rng(7)
    dv = rng.standard_normal(len(basis)) + 1j * rng.standard_normal(len(basis))
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    xd, dd = to_dev(ctx, x), to_dev(ctx, dv)
    monkeypatch.setenv("QD_KRON_VARIANTS", "15,15")
    mf = LinOp.matfree(ctx, b, p)
    try:
        assert mf.set_structured(True)
        if amps == 2:
            with pytest.raises(abi.QDError) as ei:
                mf.apply(xd)
            assert abi.STATUS[ei.value.code] == "QD_E_UNSUPPORTED"
            return
        y = mf.apply(xd).cpu().numpy()
        assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-14
        view = LinOp.with_diagonal(ctx, mf, dd)
        yv = view.apply(xd).cpu().numpy()
        view.close()
        yd = yr + dv * x
        assert np.linalg.norm(yv - yd) / np.linalg.norm(yd) < 1e-14
    finally:
        _close(mf, p, b)


def test_kron_tuned_variants_follow_the_form(ctx):
    """The variant pair chosen for a block shape is cached for the process;
    two forms of one shape (14-site ring, 2 up and 3 down) whose down-spin
    blocks differ -- one hopping amplitude (sign form: the fused product and
    the sign-form kernels can run) and two alternating amplitudes on the down
    spins (they cannot) -- must each get a pair they support: both products
    against the oracle, in either order."""
    idx = Q.getQbits(28)
    gate = Q.getHopFermiGate(idx)
    bonds = Q.hubbard_bonds_chain(14, True)

    def model(t_dn2):
        ham = Q.getEmptyOperator()
        for j, (a, b) in enumerate(bonds):
            for off, t in ((0, 1.0), (14, 1.0 if j % 2 == 0 else t_dn2)):
                ham = ham + (-t) * gate(a + off, b + off) + (-t) * gate(b + off, a + off)
        for i in range(14):
            ham = ham + 3.0 * (gate(i, i) * gate(i + 14, i + 14))
        return ham

    sec = Q.hubbard_sector(idx, 14, 2, 3)
    basis = oracle.basis_scan(sec)
    x = oracle.seeded_state(len(basis), 31)
    b = Basis(ctx, sec)
    xd = to_dev(ctx, x)
    for order in ((1.0, 0.55), (0.4, 1.0)):
        for t_dn2 in order:
            ham = model(t_dn2)
            yr = oracle.spmv(oracle.build_csc(basis, ham), x)
            p = TermProgram(ctx, ham, sec.nbits)
            mf = LinOp.matfree(ctx, b, p)
            try:
                assert mf.set_structured(True)
                y = mf.apply(xd).cpu().numpy()
                assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-14, t_dn2
            finally:
                _close(mf, p)
    _close(b)


@pytest.mark.parametrize("method", [0, 1])
def test_reorth_methods_agree(ctx, method):
    """CGS2 (4 passes) and delayed CGS2 (2 passes): same E0, same evolve, orthonormal basis."""
    ham, sec = hub(8, Q.hubbard_bonds_chain(8, False), 4, 4, U=4.0)[1:]
    ref_basis = oracle.basis_scan(sec)
    ref = oracle.build_csc(ref_basis, ham)
    w = np.linalg.eigvalsh(oracle.csc_to_dense(ref))
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    kr = Krylov(ctx, b.size, 150, herm=True)
    kr.set_method(method)
    try:
        psi = to_dev(ctx, oracle.seeded_state(b.size, 5))
        e0 = kr.ground_state(op, psi)
        assert abs(e0 - w[0]) / abs(w[0]) < 1e-10
        assert abs(ctx.norm2(psi) - 1.0) < 1e-9
        ev = kr.eigenvalues()
        # full reorthogonalisation: no spurious copies of the ground state
        assert np.sum(np.abs(ev - w[0]) < 1e-8) == 1
        kr2 = Krylov(ctx, b.size, 30, herm=False)
        kr2.set_method(method)
        x0 = oracle.seeded_state(b.size, 6)
        out = ctx.zeros(b.size)
        kr2.evolve(op, to_dev(ctx, x0), out, 0.05, 30)
        r = oracle.arnoldi_evolve(ref, 30, 0.05, x0)
        assert np.max(np.abs(out.cpu().numpy() - r)) < 1e-11
        kr2.close()
    finally:
        _close(kr, op, c, p, b)


@pytest.mark.parametrize("case", ["hub8_44", "xxz16"])
def test_partial_reorth_matches_full(ctx, case):
    """Method 2 (local window + partial reorthogonalisation) against method 1
    (full delayed CGS2) over 150 steps: the converged Ritz values and the
    ground state agree, no spurious copy of E0 appears, and only some steps
    pay for a full pass.  Method 3 (window only, no estimate) on the same
    problem does acquire a spurious copy: the check detects lost orthogonality."""
    if case == "hub8_44":
        ham, sec = hub(8, Q.hubbard_bonds_chain(8, False), 4, 4, U=4.0)[1:]
    else:
        ham, sec = xxz(16, 1.0, 0.8)[1:]
    ref_basis = oracle.basis_scan(sec)
    w = np.linalg.eigvalsh(oracle.csc_to_dense(oracle.build_csc(ref_basis, ham)))
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    x0 = oracle.seeded_state(b.size, 7)
    res = {}
    try:
        for method in (1, 2, 3):
            kr = Krylov(ctx, b.size, 150, herm=True)
            kr.set_method(method)
            psi = to_dev(ctx, x0)
            e0 = kr.ground_state(op, psi)
            res[method] = (e0, kr.eigenvalues(), psi.cpu().numpy(), kr.info())
            kr.close()
    finally:
        _close(op, c, p, b)
    e1, ev1, g1, _ = res[1]
    e2, ev2, g2, inf2 = res[2]
    ev3 = res[3][1]
    assert abs(e2 - w[0]) / abs(w[0]) < 1e-10 and abs(e2 - e1) / abs(w[0]) < 1e-10
    # Ritz values that the full scheme converged (to 1e-8) match in the partial one
    for i in range(8):
        if abs(ev1[i] - w[i]) < 1e-8 * abs(w[0]):
            assert abs(ev2[i] - w[i]) < 1e-8 * abs(w[0]), (i, ev2[i], w[i])
    assert np.sum(np.abs(ev2 - w[0]) < 1e-8 * abs(w[0])) == 1
    assert np.sum(np.abs(ev3 - w[0]) < 1e-8 * abs(w[0])) >= 2
    # semi-orthogonal basis: the Ritz vector's norm is 1 to O(sqrt(eps))
    assert abs(np.linalg.norm(g2) - 1.0) < 1e-6
    ov = abs(np.vdot(g1, g2)) / (np.linalg.norm(g1) * np.linalg.norm(g2))
    assert ov > 1 - 1e-9
    assert 0 < inf2["full_steps"] < 75, inf2


def test_stream_bandwidth_probe(ctx):
    """qd_stream_bandwidth: plausible positive rates; bad sizes rejected."""
    import ctypes
    rd, cp = ctypes.c_double(), ctypes.c_double()
    assert ctx.lib.qd_stream_bandwidth(ctx.ptr, 1 << 30, 3, ctypes.byref(rd), ctypes.byref(cp)) == 0
    assert 500.0 < rd.value < 16000.0 and 500.0 < cp.value < 16000.0, (rd.value, cp.value)
    assert ctx.lib.qd_stream_bandwidth(ctx.ptr, 16, 3, ctypes.byref(rd), ctypes.byref(cp)) != 0


@pytest.mark.parametrize("nd,nh", [(2, None), (None, 8), (1, 7)])
def test_basis_nd_nh_full_size_hubbard4x4(ctx, nd, nh):
    """4x4 Hubbard, 5 up / 5 down (19,079,424 states) filtered by the number
    of doubly occupied (d) and empty (h) sites: the size matches the closed
    form (choose the d doublons, then the remaining up and down sites, with
    h = L - N_up - N_dn + d), the table is strictly ascending, sampled states
    carry the right counts, and sampled ranks round-trip."""
    from math import comb
    L, nu, ndn = 16, 5, 5
    idx = Q.getQbits(2 * L)
    cons = [Q.getNup(idx, L) == nu, Q.getNdo(idx, L) == ndn]
    if nd is not None:
        cons.append(Q.getNd(idx, L) == nd)
    if nh is not None:
        cons.append(Q.getNh(idx, L) == nh)
    d = nd if nd is not None else nh - (L - nu - ndn)
    expect = comb(L, d) * comb(L - d, nu - d) * comb(L - nu, ndn - d)
    if nd is not None and nh is not None and nh != L - nu - ndn + nd:
        expect = 0
    b = Basis(ctx, Q.Constrain(idx, *cons))
    try:
        assert b.size == expect
        if expect == 0:
            return
        st = b.states()
        assert np.all(st[1:] > st[:-1])
        rng = np.random.default_rng(3)
        pick = rng.choice(len(st), size=2000, replace=False)
        mask = (1 << L) - 1
        for x in st[pick[:200]]:
            up, dn = int(x) & mask, (int(x) >> L) & mask
            assert bin(up).count("1") == nu and bin(dn).count("1") == ndn
            assert bin(up & dn).count("1") == d
        assert np.array_equal(b.rank(st[pick]), pick)
    finally:
        b.close()
