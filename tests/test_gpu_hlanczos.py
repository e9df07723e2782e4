"""GPU parity of the two-vector Hermitian Lanczos (Hermitian::LanczosSolver,
sparse_algorithm.hpp:236-313) through the C ABI against the CPU oracle, plus
full-size checks (XXZ N=24 ground state, a 3x4 Hubbard step through the
two-block product) and the C++ drop-in class."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from qudrip_amd import operators as Q
from qudrip_amd.runtime import CSR, Basis, HermitianLanczos, Krylov, LinOp, TermProgram
from tests import physics as P

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "qudrip_amd", "programs", "bin")


def to_dev(ctx, x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.complex128)).to(f"cuda:{ctx.device}")


def _build(ctx, ham, sec):
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    c = CSR(ctx, b, p)
    return b, p, c


def _close(*objs):
    for o in objs:
        o.close()


def _phase_aligned(g, r):
    ov = np.vdot(r, g)
    return g * (abs(ov) / ov)


def test_hlanczos_ground_state_matches_oracle_xy16(ctx):
    """XY16 OBC, N_up = 8, kry_dim 100: 101 coefficients; the leading ones
    agree with the oracle to rounding (later ones drift as the recurrence
    loses orthogonality, in both); E0 against the free-fermion value; the
    Ritz vector against the oracle's."""
    idx, ham = Q.xy_chain(16, pbc=False)
    sec = Q.Constrain(idx, Q.getNset(idx) == 8)
    ref = oracle.build_csc(oracle.basis_scan(sec), ham)
    psi0 = oracle.seeded_state(12870, 1)
    r = oracle.hlanczos(ref, 100, psi0)
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    lz = HermitianLanczos(ctx, b.size, 100)
    try:
        psi = to_dev(ctx, psi0)
        e0 = lz.ground_state(op, psi)
        e_ref = P.xy_free_fermion_e0(16, 8)
        assert abs(e0 - e_ref) / abs(e_ref) < 1e-10
        assert abs(e0 - r["evals"][0]) / abs(e_ref) < 1e-12
        inf = lz.info()
        assert inf["dim"] == r["dim"] == 101
        al, be = lz.coefficients()
        assert np.max(np.abs(al[:30] - r["alpha"][:30])) < 1e-11
        assert np.max(np.abs(be[:30] - r["beta"][:30])) < 1e-11
        assert np.max(np.abs(al.imag)) < 1e-13
        # without reorthogonalisation the Ritz vector is only as stable as the
        # recurrence: 1e-15 noise on the start moves its norm by ~3e-8 and its
        # entries by ~2e-9 (oracle against itself)
        g = psi.cpu().numpy()
        assert abs(np.linalg.norm(g) - np.linalg.norm(r["psi"])) < 1e-7
        assert np.max(np.abs(_phase_aligned(g, r["psi"]) - r["psi"])) < 1e-7
        ev = op.expect(psi)
        assert abs(ev.real / ctx.norm2(psi) - e_ref) / abs(e_ref) < 1e-9
    finally:
        _close(lz, op, c, p, b)


def test_hlanczos_evolve_matches_oracle(ctx):
    """Hubbard chain L=8, (3,3), n = 3136, kry_dim 20, h = 0.05; stored and
    matrix-free products; output over the input; a step forward and back."""
    idx, ham = Q.hubbard(8, Q.hubbard_bonds_chain(8, False), 1.0, 2.0)
    sec = Q.hubbard_sector(idx, 8, 3, 3)
    ref = oracle.build_csc(oracle.basis_scan(sec), ham)
    psi0 = 1.7 * oracle.seeded_state(3136, 4)
    r = oracle.hlanczos(ref, 20, psi0, mode=1, h=0.05)
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    mf = LinOp.matfree(ctx, b, p)
    lz = HermitianLanczos(ctx, b.size, 20)
    try:
        xin = to_dev(ctx, psi0)
        out = ctx.zeros(b.size)
        lz.evolve(op, xin, out, 0.05)
        assert np.max(np.abs(out.cpu().numpy() - r["psi"])) < 1e-12
        assert lz.info()["dim"] == 21
        lz.evolve(mf, xin, out, 0.05)
        assert np.max(np.abs(out.cpu().numpy() - r["psi"])) < 1e-12
        y = to_dev(ctx, psi0)
        lz.evolve(op, y, y, 0.05)  # in place
        assert np.max(np.abs(y.cpu().numpy() - r["psi"])) < 1e-12
        back = ctx.zeros(b.size)
        lz.evolve(op, out, back, -0.05)
        assert np.max(np.abs(back.cpu().numpy() - psi0 / np.linalg.norm(psi0))) < 1e-12
    finally:
        _close(lz, mf, op, c, p, b)


def test_hlanczos_closed_forms_and_stopping(ctx):
    """Two-site Hubbard (dim 4 then |r|^2 <= 1e-10, Ritz values = spectrum),
    sigma_x from e_0 (evolve = cos h e_0 - i sin h e_1), a start 1e-6 off an
    eigenvector (one coefficient: the test is on |r|^2), kry_dim 7 (eight)."""
    U = 4.0
    idx, ham = Q.hubbard(2, Q.hubbard_bonds_chain(2, False), 1.0, U)
    sec = Q.hubbard_sector(idx, 2, 1, 1)
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    lz = HermitianLanczos(ctx, 4, 50)
    try:
        psi = to_dev(ctx, oracle.seeded_state(4, 9))
        e0 = lz.ground_state(op, psi)
        spec = sorted([(U - np.sqrt(U * U + 16)) / 2, 0.0, U, (U + np.sqrt(U * U + 16)) / 2])
        assert lz.info()["dim"] == 4 and lz.info()["error"] ** 2 <= 1e-10
        assert np.allclose(lz.eigenvalues(), spec, atol=1e-12) and abs(e0 - spec[0]) < 1e-12
    finally:
        _close(lz, op, c, p, b)

    idx, ham = Q.xy_chain(2, 1.0, pbc=False)
    sec = Q.Constrain(idx, Q.getNset(idx) == 1)
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    lz = HermitianLanczos(ctx, 2, 10)
    try:
        h = 0.3
        out = ctx.zeros(2)
        lz.evolve(op, to_dev(ctx, [2.0, 0.0]), out, h)
        assert lz.info()["dim"] == 2
        assert np.allclose(out.cpu().numpy(), [np.cos(h), -1j * np.sin(h)], atol=1e-14)
    finally:
        _close(lz, op, c, p, b)

    idx, ham = Q.xxz_chain(8, 1.0, 0.7, pbc=True)
    sec = Q.Constrain(idx, Q.getNset(idx) == 4)
    ref = oracle.build_csc(oracle.basis_scan(sec), ham)
    w, z = np.linalg.eigh(oracle.csc_to_dense(ref))
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    lz = HermitianLanczos(ctx, b.size, 30)
    lz7 = HermitianLanczos(ctx, b.size, 7)
    try:
        psi = to_dev(ctx, z[:, 0] + 1e-6 * z[:, -1])
        e0 = lz.ground_state(op, psi)
        assert lz.info()["dim"] == 1 and 1e-7 < lz.info()["error"] < 1e-5 and abs(e0 - w[0]) < 1e-9
        lz7.ground_state(op, to_dev(ctx, oracle.seeded_state(b.size, 3)))
        r7 = oracle.hlanczos(ref, 7, oracle.seeded_state(b.size, 3))
        assert lz7.info()["dim"] == r7["dim"] == 8
        assert np.allclose(lz7.eigenvalues(), r7["evals"], atol=1e-12)
        assert abs(lz7.info()["error"] - r7["err"]) < 1e-12
    finally:
        _close(lz7, lz, op, c, p, b)


def test_hlanczos_xxz24_ground_state_full_size(ctx):
    """XXZ N=24 PBC, Sz = 0 (n = 2,704,156), kry_dim 150: E0 against the
    Bethe-ansatz value; the Ritz vector's Rayleigh quotient."""
    idx, ham = Q.xxz_chain(24, 1.0, 1.0, pbc=True)
    sec = Q.Constrain(idx, Q.getNset(idx) == 12)
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.from_csr(ctx, c)
    lz = HermitianLanczos(ctx, b.size, 150)
    try:
        psi = ctx.seeded_state(b.size, 2)
        e0 = lz.ground_state(op, psi)
        e_bethe = -10.670014516537
        assert abs(e0 - e_bethe) / abs(e_bethe) < 1e-10
        ev = op.expect(psi)
        assert abs(ev.real / ctx.norm2(psi) - e0) / abs(e0) < 1e-9
        assert lz.info()["dim"] == 151
    finally:
        _close(lz, op, c, p, b)


def test_hlanczos_hubbard3x4_step_matches_arnoldi(ctx):
    """3x4 Hubbard PBC, (6,6), n = 853,776, through the two-block product: one
    step h = 0.05 of the two-vector solver (25 coefficients) against the fully
    reorthogonalised Arnoldi step (24 vectors; both converged to rounding at
    this h); norm and energy kept."""
    L = 12
    idx, ham = Q.hubbard(L, Q.hubbard_bonds_square(3, 4), 1.0, 4.0)
    sec = Q.hubbard_sector(idx, L, 6, 6)
    b = Basis(ctx, sec)
    p = TermProgram(ctx, ham, sec.nbits)
    op = LinOp.matfree(ctx, b, p)
    assert op.set_structured(True)
    lz = HermitianLanczos(ctx, b.size, 24)
    kr = Krylov(ctx, b.size, 30, herm=True)
    try:
        psi = ctx.seeded_state(b.size, 5)
        out = ctx.zeros(b.size)
        lz.evolve(op, psi, out, 0.05)
        ref = ctx.zeros(b.size)
        kr.evolve(op, psi, ref, 0.05, 25)
        d = (out - ref).abs().max().item()
        assert d < 1e-11
        assert abs(ctx.norm2(out) - 1.0) < 1e-12
        e_in = op.expect(psi).real
        e_out = op.expect(out).real
        assert abs(e_out - e_in) < 1e-10 * abs(e_in)
    finally:
        _close(kr, lz, op, p, b)


def test_dropin_hermitian_solver(tmp_path):
    """Hermitian::getSolver(sub, k): evolve, getGroundState, storeE and
    error() from a reference-style C++ driver, against the oracle."""
    prefix = str(tmp_path / "herm")
    N, k, h = 12, 40, 0.05
    r = subprocess.run([os.path.join(BIN, "dropin_dump"), "xxz", str(N), prefix, "0.9317", "0.3711", "herm", str(k),
                        repr(h)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    idx, ham = Q.xxz_chain(N, 0.9317, 0.3711, pbc=True)
    ref = oracle.build_csc(oracle.basis_scan(Q.Constrain(idx, Q.getNset(idx) == N // 2)), ham)
    x = np.fromfile(prefix + ".x", dtype=np.complex128)
    rev = oracle.hlanczos(ref, k, x, mode=1, h=h)
    ev = np.fromfile(prefix + ".ev", dtype=np.complex128)
    assert np.max(np.abs(ev - rev["psi"])) < 1e-12
    rgs = oracle.hlanczos(ref, k, x)
    gs = np.fromfile(prefix + ".gs", dtype=np.complex128)
    assert np.max(np.abs(_phase_aligned(gs, rgs["psi"]) - rgs["psi"])) < 1e-8
    E = np.loadtxt(prefix + ".E")
    assert len(E) == rgs["dim"] == k + 1
    assert np.allclose(E, rgs["evals"], rtol=1e-5, atol=1e-5)  # printed with 6 digits, like the reference
    err = float(r.stdout.split("error = ")[1].split()[0])
    assert abs(err - rgs["err"]) < 1e-9


@pytest.mark.parametrize("matfree", [False, True])
def test_hlanczos_non_temporal_form_is_bit_identical(ctx, matfree, monkeypatch):
    """The vector kernels' non-temporal form (chosen above 256 MB vectors)
    forced at n = 924 through QD_HLZ_NT: coefficients, Ritz values and the
    state bit for bit equal to the ordinary form.  Stored matrix: the
    product fused with the dot; matrix-free: product then the dot pass."""
    idx, ham = Q.xxz_chain(12, 1.0, 0.7, pbc=True)
    sec = Q.Constrain(idx, Q.getNset(idx) == 6)
    b, p, c = _build(ctx, ham, sec)
    op = LinOp.matfree(ctx, b, p) if matfree else LinOp.from_csr(ctx, c)
    out = {}
    try:
        for nt in ("0", "1"):
            monkeypatch.setenv("QD_HLZ_NT", nt)
            lz = HermitianLanczos(ctx, b.size, 40)
            psi = ctx.seeded_state(b.size, 3)
            e0 = lz.ground_state(op, psi)
            a, be = lz.coefficients()
            out[nt] = (e0, a, be, psi.cpu().numpy())
            lz.close()
    finally:
        _close(op, c, p, b)
    e0, a0, b0, s0 = out["0"]
    e1, a1, b1, s1 = out["1"]
    assert e0 == e1
    assert np.array_equal(a0.view(np.uint64), a1.view(np.uint64))
    assert np.array_equal(b0.view(np.uint64), b1.view(np.uint64))
    assert np.array_equal(s0.view(np.uint64), s1.view(np.uint64))


def test_hlanczos_fused_product_matches_separate_dot(ctx):
    """The stored-matrix step (product and a = <q|Hq> in one kernel) against
    the matrix-free step (product, then a dot pass) on the same operator:
    the leading coefficients agree to rounding, E0 to 1e-12."""
    idx, ham = Q.xxz_chain(14, 1.0, 1.3, pbc=True)
    sec = Q.Constrain(idx, Q.getNset(idx) == 7)
    b, p, c = _build(ctx, ham, sec)
    ops = [LinOp.from_csr(ctx, c), LinOp.matfree(ctx, b, p)]
    res = []
    try:
        for op in ops:
            lz = HermitianLanczos(ctx, b.size, 60)
            psi = ctx.seeded_state(b.size, 9)
            e0 = lz.ground_state(op, psi)
            a, be = lz.coefficients()
            res.append((e0, a, be))
            lz.close()
    finally:
        _close(*ops, c, p, b)
    (e0, a0, b0), (e1, a1, b1) = res
    assert abs(e0 - e1) <= 1e-12 * abs(e0)
    assert np.max(np.abs(a0[:20] - a1[:20])) < 1e-12 and np.max(np.abs(b0[:20] - b1[:20])) < 1e-12
