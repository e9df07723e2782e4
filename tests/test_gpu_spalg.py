"""SpMatrix algebra on the device (qd_csr_zero / identity / scale / add /
adjoint / multiply) against the oracle's restatement of Eigen's rules
(oracle/spalg.c), bit for bit, and the reference's dissipative driver loop
(programs/hubbard_diss.cpp:106-159,439-470,514-529) through the drop-in C++
headers (qudrip_amd/programs/hubbard_quench.cpp) against the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle
from qudrip_amd import operators as Q
from qudrip_amd.runtime import CSR, Basis, Krylov, LinOp, TermProgram

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "qudrip_amd", "programs", "bin")


def u64(v):
    return np.ascontiguousarray(v).view(np.uint64)


def same(gpu_csr, ref):
    o, i, v = gpu_csr.download()
    ro, ri, rv = oracle._c(ref)
    assert np.array_equal(o, ro)
    assert np.array_equal(i, ri)
    assert np.array_equal(u64(v), u64(rv))


class Chain:
    """Open Hubbard chain parts as hubbard_diss.cpp builds them (:97-159)."""

    def __init__(self, L, nup, ndn):
        self.L = L
        self.idx = Q.getQbits(2 * L)
        self.sec = Q.hubbard_sector(self.idx, L, nup, ndn)
        self.gate = Q.getHopFermiGate(self.idx)
        self.basis = oracle.basis_scan(self.sec)
        self.n = len(self.basis)

    def sums(self, thop=-1.0):
        """(Tp terms, Tm terms, docc terms, h0 terms, lindblad terms) per site, in the driver's order."""
        g, L = self.gate, self.L
        out = []
        for s in range(L):
            s2 = (s + 1) % L
            hops = []
            if s != L - 1:
                hops = [(thop * g(s2, s), thop * g(s, s2)), (thop * g(s2 + L, s + L), thop * g(s + L, s2 + L))]
            nu, nd = g(s, s), g(s + L, s + L)
            # `op >> index` of one product is OperatorSum(op, 1) >> index (operator.hpp:235-267)
            out.append({"hops": hops, "docc": Q.getEmptyOperator() + nu * nd, "h0": nu + nd,
                        "ni": g(s, s) + g(s + L, s + L)})
        return out


def test_algebra_ops_bit_exact(ctx):
    """Each operation on matrices built by `>>` (Hubbard chain L=6, 3/3,
    n = 400) with non-dyadic scalars, device against oracle."""
    ch = Chain(6, 3, 3)
    b = Basis(ctx, ch.sec)
    keep = [b]

    def build(opsum):
        p = TermProgram(ctx, opsum, 2 * ch.L)
        c = CSR(ctx, b, p)
        keep.extend([p, c])
        return c, oracle.build_csc(ch.basis, opsum)

    parts = ch.sums()
    T, T_ref = build(parts[0]["hops"][0][0])
    T2, T2_ref = build(parts[2]["hops"][1][1])
    D, D_ref = build(parts[1]["docc"])
    N, N_ref = build(parts[3]["ni"])
    try:
        n = ch.n
        cases = []
        Z = CSR.zero(ctx, n); keep.append(Z)
        I = CSR.identity(ctx, n); keep.append(I)
        same(Z, oracle.csr_zero(n))
        same(I, oracle.csr_identity(n))
        c1 = complex(math.cos(0.37), math.sin(0.37)) * complex(1.0, 0.15)
        for gpu, ref in (
            (T.scale(c1), oracle.csr_scale(T_ref, c1)),
            (T.scale(-1.3), oracle.csr_scale(T_ref, -1.3)),
            (T.add(T2, 1), oracle.csr_add(T_ref, T2_ref, 1)),
            (T.add(D, -1), oracle.csr_add(T_ref, D_ref, -1)),
            (Z.add(T, 1), oracle.csr_add(oracle.csr_zero(n), T_ref, 1)),
            (Z.add(T, -1), oracle.csr_add(oracle.csr_zero(n), T_ref, -1)),
            (T.adjoint(), oracle.csr_adjoint(T_ref)),
            (T.multiply(T2), oracle.csr_multiply(T_ref, T2_ref)),
            (N.adjoint().multiply(N), oracle.csr_multiply(oracle.csr_adjoint(N_ref), N_ref)),
            (I.add(N, -1).add(D.scale(2.0), 1), oracle.csr_add(oracle.csr_add(oracle.csr_identity(n), N_ref, -1),
                                                               oracle.csr_scale(D_ref, 2.0), 1)),
        ):
            keep.append(gpu)
            same(gpu, ref)
        # T - T: every entry stays, exactly zero (explicit zeros, no pruning)
        d = T.add(T, -1); keep.append(d)
        o, i, v = d.download()
        assert len(i) == T.nnz and np.all(v == 0.0)
    finally:
        for h in reversed(keep):
            h.close()


def test_non_hermitian_combination_evolves_like_the_oracle(ctx):
    """H_eff = Htot - 0.5 i (qj - w Id) (hubbard_diss.cpp:443,514-517) built
    on the device; the Arnoldi step (sparse_algorithm.hpp:548-562) through
    it (its adjoint for the product comes from the device transpose) against
    oracle.arnoldi_evolve on the oracle's H_eff."""
    ch = Chain(4, 2, 2)
    b = Basis(ctx, ch.sec)
    keep = [b]
    parts = ch.sums()
    gpu, ref = {}, {}
    n = ch.n
    gamma, U0, w = 0.1, 4.0, 1.234567
    Tp, Tm, Dc, Qj = CSR.zero(ctx, n), CSR.zero(ctx, n), CSR.zero(ctx, n), CSR.zero(ctx, n)
    keep += [Tp, Tm, Dc, Qj]
    rTp, rTm, rDc, rQj = [oracle.csr_zero(n) for _ in range(4)]

    def mk(opsum):
        p = TermProgram(ctx, opsum, 2 * ch.L)
        c = CSR(ctx, b, p)
        keep.extend([p, c])
        return c, oracle.build_csc(ch.basis, opsum)

    for s in parts:
        for hp, hm in s["hops"]:
            a, ra = mk(hp)
            Tp = Tp.add(a, 1); rTp = oracle.csr_add(rTp, ra, 1); keep.append(Tp)
            a, ra = mk(hm)
            Tm = Tm.add(a, 1); rTm = oracle.csr_add(rTm, ra, 1); keep.append(Tm)
        a, ra = mk(s["docc"])
        Dc = Dc.add(a, 1); rDc = oracle.csr_add(rDc, ra, 1); keep.append(Dc)
        a, ra = mk(s["ni"])
        q = a.adjoint(); keep.append(q)
        q2 = q.scale(gamma); keep.append(q2)
        q3 = q2.multiply(a); keep.append(q3)
        Qj = Qj.add(q3, 1); keep.append(Qj)
        rQj = oracle.csr_add(rQj, oracle.csr_multiply(oracle.csr_scale(oracle.csr_adjoint(ra), gamma), ra), 1)
    try:
        Tdr = complex(math.cos(0.2), math.sin(0.2)) * complex(1.0, 0.15)
        H = Tp.scale(Tdr).add(Tm.scale(Tdr.conjugate()), 1).add(Dc.scale(0.3 + U0), 1)
        rH = oracle.csr_add(oracle.csr_add(oracle.csr_scale(rTp, Tdr), oracle.csr_scale(rTm, Tdr.conjugate()), 1),
                            oracle.csr_scale(rDc, 0.3 + U0), 1)
        keep.append(H)
        same(H, rH)
        I = CSR.identity(ctx, n); keep.append(I)
        X = Qj.add(I.scale(w), -1).scale(0.5j); keep.append(X)
        He = H.add(X, -1); keep.append(He)
        rHe = oracle.csr_add(rH, oracle.csr_scale(oracle.csr_add(rQj, oracle.csr_scale(oracle.csr_identity(n), w),
                                                                 -1), 0.5j), -1)
        same(He, rHe)
        assert He.hermitian == 0 and H.hermitian == 1
        op = LinOp.from_csr(ctx, He); keep.append(op)
        kr = Krylov(ctx, n, 12, herm=False); keep.append(kr)
        psi = ctx.seeded_state(n, 5)
        out = ctx.zeros(n)
        kr.evolve(op, psi, out, 0.05, 12)
        ref = oracle.arnoldi_evolve(oracle._c(rHe), 12, 0.05, oracle.seeded_state(n, 5))
        assert np.max(np.abs(out.cpu().numpy() - ref)) < 1e-11
        # and the product itself: y = H_eff x through the adjoint-based pull
        y = op.apply(psi)
        yr = oracle.spmv(oracle._c(rHe), oracle.seeded_state(n, 5))
        assert np.linalg.norm(y.cpu().numpy() - yr) <= 1e-14 * np.linalg.norm(yr)
    finally:
        for h in reversed(keep):
            h.close()


def _load(prefix, ext, dtype):
    return np.fromfile(prefix + ext, dtype=dtype)


def test_dropin_dissipative_driver_matches_oracle(tmp_path):
    """The driver loop of hubbard_diss.cpp through the drop-in headers:
    SpMatrix arithmetic for Tp, Tm, docc, qj_tot (with SpMatrix(ni.adjoint())
    * ni) and Pc, the two-vector ground state, then H_eff per step and
    l_solver.evolve.  Matrices bit for bit against the oracle; states and
    observables against the oracle's own loop from the same ground state."""
    L, nup, ndn, kd, krt, nt, h, U0, gamma = 4, 2, 2, 30, 12, 20, 0.05, 4.0, 0.1
    prefix = str(tmp_path / "hq")
    r = subprocess.run([os.path.join(BIN, "hubbard_quench"), str(L), str(nup), str(ndn), str(kd), str(krt), str(nt),
                        repr(h), repr(U0), repr(gamma), prefix], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    ch = Chain(L, nup, ndn)
    n = ch.n
    thop = -1.0
    Tp = Tm = Dc = Qj = oracle.csr_zero(n)
    Pc = oracle.csr_identity(n)
    Id = oracle.csr_identity(n)
    for s in ch.sums(thop):
        for hp, hm in s["hops"]:
            Tp = oracle.csr_add(Tp, oracle.build_csc(ch.basis, hp), 1)
            Tm = oracle.csr_add(Tm, oracle.build_csc(ch.basis, hm), 1)
        Dc = oracle.csr_add(Dc, oracle.build_csc(ch.basis, s["docc"]), 1)
        ni = oracle.build_csc(ch.basis, s["ni"])
        Qj = oracle.csr_add(Qj, oracle.csr_multiply(oracle.csr_scale(oracle.csr_adjoint(ni), gamma), ni), 1)
        f = oracle.csr_add(oracle.csr_add(Id, oracle.build_csc(ch.basis, s["h0"]), -1),
                           oracle.csr_scale(oracle.build_csc(ch.basis, s["docc"]), 2.0), 1)
        Pc = oracle.csr_multiply(Pc, f)
    for name, M in (("qj", Qj), ("pc", Pc)):
        ro, ri, rv = oracle._c(M)
        assert np.array_equal(_load(prefix, f".{name}.outer", np.int32), ro)
        assert np.array_equal(_load(prefix, f".{name}.inner", np.int32), ri)
        assert np.array_equal(_load(prefix, f".{name}.vals", np.uint64), u64(rv))
    # Pc projects onto configurations without doubly occupied or empty sites
    # (hubbard_diss.cpp:158): diagonal, entries 0 or 1
    Pd = oracle.csc_to_dense(oracle._c(Pc))
    assert np.allclose(Pd, np.diag(np.diag(Pd))) and set(np.round(np.diag(Pd).real, 12)) <= {0.0, 1.0}
    # ground state of Htot = Tp + Tm + U0 docc by the two-vector solver
    H0 = oracle.csr_add(oracle.csr_add(Tp, Tm, 1), oracle.csr_scale(Dc, U0), 1)
    gs = _load(prefix, ".gs", np.complex128)
    w, z = np.linalg.eigh(oracle.csc_to_dense(oracle._c(H0)))
    assert abs(abs(np.vdot(z[:, 0], gs)) ** 2 / np.vdot(gs, gs).real - 1.0) < 1e-9
    orc_gs = oracle.hlanczos(oracle._c(H0), kd, oracle.seeded_state(n, 1))["psi"]
    assert np.max(np.abs(orc_gs - gs)) < 1e-7
    # the time loop from the driver's own ground state
    obs = np.loadtxt(prefix + ".obs", ndmin=2)
    psi = gs.copy()
    for i in range(1, nt + 1):
        Adr = 0.7 * math.sin(0.9 * i * h)
        Tdr = complex(math.cos(Adr), math.sin(Adr)) * complex(1.0, 0.15)
        ndr = 0.3 * math.cos(2.0 * i * h)
        H = oracle.csr_add(oracle.csr_add(oracle.csr_scale(Tp, Tdr), oracle.csr_scale(Tm, Tdr.conjugate()), 1),
                           oracle.csr_scale(Dc, ndr + U0), 1)
        wtot = 0.0
        for s in ch.sums(thop):
            ni = oracle.build_csc(ch.basis, s["ni"])
            wtot += oracle.eval_expect(oracle.csr_multiply(oracle.csr_adjoint(ni), ni), psi).real
        assert abs(wtot - obs[i - 1, 2]) < 1e-10
        if i == 1:
            wt = float(obs[0, 2])  # the driver's own weight: H_eff of step 1 bit for bit
            He1 = oracle.csr_add(H, oracle.csr_scale(oracle.csr_add(Qj, oracle.csr_scale(Id, gamma * wt), -1),
                                                     0.5j), -1)
            ro, ri, rv = oracle._c(He1)
            assert np.array_equal(_load(prefix, ".H1.outer", np.int32), ro)
            assert np.array_equal(_load(prefix, ".H1.inner", np.int32), ri)
            assert np.array_equal(_load(prefix, ".H1.vals", np.uint64), u64(rv))
        He = oracle.csr_add(H, oracle.csr_scale(oracle.csr_add(Qj, oracle.csr_scale(Id, gamma * wtot), -1), 0.5j),
                            -1)
        psi = oracle.arnoldi_evolve(oracle._c(He), krt, h, psi)
        psi /= math.sqrt(np.vdot(psi, psi).real)
        assert abs(oracle.eval_expect(Dc, psi).real - obs[i - 1, 1]) < 1e-10
    assert np.max(np.abs(_load(prefix, ".psi", np.complex128) - psi)) < 1e-10
