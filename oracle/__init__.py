"""CPU oracle: plain-C restatement of the reference algorithms (oracle.c).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline; the
product (qudrip_amd/, include/) never imports, links or calls it.

Each wrapper names the reference function it restates; oracle.c documents
how each one is pinned independently (tests/test_oracle.py).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_VP = ctypes.c_void_p


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    L.orc_basis_scan.restype = ctypes.c_int64
    L.orc_basis_scan.argtypes = [ctypes.c_int, _VP, ctypes.c_int, _VP, ctypes.c_int64]
    L.orc_basis_scan_q.restype = ctypes.c_int64
    L.orc_basis_scan_q.argtypes = [ctypes.c_int, _VP, ctypes.c_int, _VP, ctypes.c_int64]
    L.orc_build_csc_modes.restype = ctypes.c_int64
    L.orc_build_csc_modes.argtypes = [_VP, ctypes.c_int64, ctypes.c_int, _VP, ctypes.c_int, _VP, _VP, _VP, _VP,
                                      _VP, ctypes.c_int64]
    L.orc_rank.restype = ctypes.c_int64
    L.orc_rank.argtypes = [_VP, ctypes.c_int64, ctypes.c_uint64]
    L.orc_binom.restype = ctypes.c_uint64
    L.orc_binom.argtypes = [ctypes.c_int, ctypes.c_int]
    L.orc_expand_term.restype = ctypes.c_int
    L.orc_expand_term.argtypes = [_VP, ctypes.c_int, ctypes.c_uint64, _VP, _VP, _VP, ctypes.c_int]
    L.orc_build_csc.restype = ctypes.c_int64
    L.orc_build_csc.argtypes = [_VP, ctypes.c_int64, _VP, ctypes.c_int, _VP, _VP, _VP, _VP, ctypes.c_int64]
    L.orc_map_acc.restype = None
    L.orc_map_acc.argtypes = [_VP, ctypes.c_int64, _VP, ctypes.c_int, _VP, _VP, _VP, _VP]
    L.orc_spmv_csc.restype = None
    L.orc_spmv_csc.argtypes = [ctypes.c_int64, _VP, _VP, _VP, _VP, _VP]
    L.orc_norm2.restype = ctypes.c_double
    L.orc_norm2.argtypes = [ctypes.c_int64, _VP]
    L.orc_eval.restype = None
    L.orc_eval.argtypes = [ctypes.c_int64, _VP, _VP, _VP, _VP, _VP]
    L.orc_uniform.restype = ctypes.c_double
    L.orc_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.orc_seeded_state.restype = None
    L.orc_seeded_state.argtypes = [ctypes.c_int64, ctypes.c_uint64, _VP, ctypes.c_int]
    L.orc_tridiag_eigvals.restype = ctypes.c_int
    L.orc_tridiag_eigvals.argtypes = [ctypes.c_int, _VP, _VP, _VP]
    L.orc_tridiag_eigvec.restype = ctypes.c_int
    L.orc_tridiag_eigvec.argtypes = [ctypes.c_int, _VP, _VP, ctypes.c_double, _VP]
    L.orc_hlanczos.restype = ctypes.c_int
    L.orc_hlanczos.argtypes = [ctypes.c_int64, _VP, _VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, _VP,
                               _VP, _VP, _VP, _VP, _VP]
    L.orc_lanczos_gs.restype = ctypes.c_int
    L.orc_lanczos_gs.argtypes = [ctypes.c_int64, _VP, _VP, _VP, ctypes.c_int, _VP, _VP, _VP, _VP, _VP, _VP]
    L.orc_expm.restype = ctypes.c_int
    L.orc_expm.argtypes = [ctypes.c_int, _VP, _VP]
    L.orc_traj_run.restype = ctypes.c_int
    L.orc_traj_run.argtypes = [ctypes.c_int, _VP, ctypes.c_int, _VP, ctypes.c_int, ctypes.c_int, _VP, _VP, _VP, _VP,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                               ctypes.c_uint64, _VP, _VP]
    L.orc_traj_run_rebuild.restype = ctypes.c_int
    L.orc_traj_run_rebuild.argtypes = L.orc_traj_run.argtypes + [_VP, _VP, _VP]
    L.orc_traj_uniform.restype = ctypes.c_double
    L.orc_traj_uniform.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int]
    L.orc_arnoldi_evolve.restype = ctypes.c_int
    L.orc_arnoldi_evolve.argtypes = [ctypes.c_int64, _VP, _VP, _VP, ctypes.c_int, ctypes.c_double, _VP, _VP]
    L.orc_csr_add.restype = ctypes.c_int64
    L.orc_csr_add.argtypes = [ctypes.c_int64, _VP, _VP, _VP, _VP, _VP, _VP, ctypes.c_int, _VP, _VP, _VP]
    L.orc_csr_scale.restype = None
    L.orc_csr_scale.argtypes = [ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int, _VP, _VP]
    L.orc_csr_adjoint.restype = None
    L.orc_csr_adjoint.argtypes = [ctypes.c_int64, _VP, _VP, _VP, _VP, _VP, _VP]
    L.orc_csr_multiply.restype = ctypes.c_int64
    L.orc_csr_multiply.argtypes = [ctypes.c_int64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP]
    _lib = L
    return L


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def basis_scan(sector):
    """Constrain / LooseConstrain (conservation.hpp:212-231): ascending scan."""
    from qudrip_amd.abi import qd_constraint
    if getattr(sector, "mdim", 1) > 1:
        # composite index (Indices, index.hpp:112-184): LooseConstrain walks
        # q * mdim + m in order and tests the quantities on q alone
        import copy
        qsec = copy.copy(sector)
        qsec.mdim = 1
        q = basis_scan(qsec)
        md = sector.mdim
        return (q[:, None] * np.uint64(md) + np.arange(md, dtype=np.uint64)[None, :]).reshape(-1)
    if not all(c.plain for c in sector.constraints):
        arr, nq = sector.ctypes_quantities()
        n = lib().orc_basis_scan_q(sector.nbits, arr, nq, None, 0)
        out = np.empty(max(n, 1), dtype=np.uint64)
        lib().orc_basis_scan_q(sector.nbits, arr, nq, _p(out), n)
        return out[:n]
    cons = (qd_constraint * max(1, len(sector.constraints)))()
    for i, c in enumerate(sector.constraints):
        cons[i].mask = c.mask
        cons[i].count = c.count
    n = lib().orc_basis_scan(sector.nbits, cons, len(sector.constraints), None, 0)
    out = np.empty(max(n, 1), dtype=np.uint64)
    lib().orc_basis_scan(sector.nbits, cons, len(sector.constraints), _p(out), n)
    return out[:n]


def rank(basis, x):
    """SubIndex -> inv_mapping lookup (index.hpp:234-240); miss -> len(basis)."""
    return lib().orc_rank(_p(basis), len(basis), int(x))


def binom(n, k):
    return lib().orc_binom(n, k)


def build_csc(basis, opsum, mdim=1):
    """`ham >> index` (operator_abel.hpp:25-34, operator.hpp:235-267); mdim > 1:
    basis holds composite values q * mdim + m (a program without mode gates
    leaves the mode alone)."""
    if opsum.has_modes() or mdim > 1:
        terms, nt, ops, no, mats, nm, md = opsum.program_modes()
        mdim = md if opsum.has_modes() else mdim
        n = len(basis)
        colptr = np.empty(n + 1, dtype=np.int32)
        nnz = lib().orc_build_csc_modes(_p(basis), n, mdim, terms, nt, ops, _p(mats), _p(colptr), None, None, 0)
        rowidx = np.empty(max(nnz, 1), dtype=np.int32)
        vals = np.empty(2 * max(nnz, 1), dtype=np.float64)
        lib().orc_build_csc_modes(_p(basis), n, mdim, terms, nt, ops, _p(mats), _p(colptr), _p(rowidx), _p(vals), nnz)
        return colptr, rowidx[:nnz], vals[:2 * nnz]
    terms, nt, ops, no = opsum.program()
    n = len(basis)
    colptr = np.empty(n + 1, dtype=np.int32)
    nnz = lib().orc_build_csc(_p(basis), n, terms, nt, ops, _p(colptr), None, None, 0)
    if nnz < 0:
        raise RuntimeError("oracle csc build failed")
    rowidx = np.empty(max(nnz, 1), dtype=np.int32)
    vals = np.empty(2 * max(nnz, 1), dtype=np.float64)
    lib().orc_build_csc(_p(basis), n, terms, nt, ops, _p(colptr), _p(rowidx), _p(vals), nnz)
    return colptr, rowidx[:nnz], vals[:2 * nnz]


def expand_term(opsum, k, state, init=(1.0, 0.0)):
    terms, nt, ops, no = opsum.program()
    t = terms[k]
    cap = 1 << 8
    idx = np.zeros(cap, dtype=np.uint64)
    val = np.zeros(2 * cap, dtype=np.float64)
    ini = np.array(init, dtype=np.float64)
    base = ctypes.addressof(ops) + t.op_begin * ctypes.sizeof(ops[0])
    nl = lib().orc_expand_term(ctypes.c_void_p(base), t.n_ops, int(state), _p(ini), _p(idx), _p(val), cap)
    return idx[:nl], val[:2 * nl].view(np.complex128)


def map_acc(basis, opsum, frm, to, coeff=1.0 + 0j):
    """OperatorSum::map_acc (operator_abel.hpp:36-42; operator.hpp:215-232)."""
    terms, nt, ops, no = opsum.program()
    c = np.array([coeff.real, coeff.imag], dtype=np.float64)
    frm = np.ascontiguousarray(frm, dtype=np.complex128)
    assert to.dtype == np.complex128 and to.flags.c_contiguous
    lib().orc_map_acc(_p(basis), len(basis), terms, nt, ops, _p(frm), _p(to), _p(c))
    return to


def spmv(csc, x):
    """Eigen ColMajor sparse * dense (sparse_algorithm.hpp:378,404)."""
    colptr, rowidx, vals = csc
    n = len(colptr) - 1
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty(n, dtype=np.complex128)
    lib().orc_spmv_csc(n, _p(colptr), _p(rowidx), _p(vals), _p(x), _p(y))
    return y


def eval_expect(csc, psi):
    """State::eval (state.hpp:33-44)."""
    colptr, rowidx, vals = csc
    n = len(colptr) - 1
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.empty(2, dtype=np.float64)
    lib().orc_eval(n, _p(colptr), _p(rowidx), _p(vals), _p(psi), _p(out))
    return complex(out[0], out[1])


def seeded_state(n, seed, normalise=True):
    """Seeded replacement of shuffle_state (utils.hpp:129-140)."""
    out = np.empty(n, dtype=np.complex128)
    lib().orc_seeded_state(n, seed, _p(out), 1 if normalise else 0)
    return out


def tridiag_eigvals(d, e):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    w = np.empty(len(d))
    if lib().orc_tridiag_eigvals(len(d), _p(d), _p(e), _p(w)) != 0:
        raise RuntimeError("tridiag eigvals failed")
    return w


def tridiag_eigvec(d, e, lam):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    z = np.empty(len(d))
    if lib().orc_tridiag_eigvec(len(d), _p(d), _p(e), lam, _p(z)) != 0:
        raise RuntimeError("tridiag eigvec failed")
    return z


def lanczos_gs(csc, k, psi0):
    """getLanczosSolver(idx,k).getGroundState (sparse_algorithm.hpp:395-473,564-585)."""
    colptr, rowidx, vals = csc
    n = len(colptr) - 1
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
    out = np.empty(n, dtype=np.complex128)
    e0 = np.zeros(1)
    ev = np.zeros(k + 1)
    al = np.zeros(k + 1)
    be = np.zeros(k + 1)
    act = lib().orc_lanczos_gs(n, _p(colptr), _p(rowidx), _p(vals), k, _p(psi0), _p(out), _p(e0), _p(ev),
                               _p(al), _p(be))
    if act < 0:
        raise RuntimeError("oracle lanczos failed")
    return {"E0": e0[0], "psi": out, "evals": ev[:act], "alpha": al[:act], "beta": be[:act], "act_dim": act}


def hlanczos(csc, kry_dim, psi0, mode=0, h=0.0):
    """Hermitian::LanczosSolver (sparse_algorithm.hpp:236-313): mode 0 getGroundState, mode 1 evolve by h."""
    colptr, rowidx, vals = csc
    n = len(colptr) - 1
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
    out = np.empty(n, dtype=np.complex128)
    al = np.zeros(kry_dim + 2, dtype=np.complex128)
    be = np.zeros(kry_dim + 2)
    ev = np.zeros(kry_dim + 2)
    err = np.zeros(1)
    dim = lib().orc_hlanczos(n, _p(colptr), _p(rowidx), _p(vals), kry_dim, mode, h, _p(psi0), _p(out), _p(al),
                             _p(be), _p(ev), _p(err))
    if dim < 0:
        raise RuntimeError("oracle Hermitian Lanczos failed")
    return {"psi": out, "alpha": al[:dim], "beta": be[:dim], "evals": ev[:dim], "err": err[0], "dim": dim}


def expm(A):
    A = np.asfortranarray(A, dtype=np.complex128)
    m = A.shape[0]
    E = np.empty((m, m), dtype=np.complex128, order="F")
    if lib().orc_expm(m, _p(A), _p(E)) != 0:
        raise RuntimeError("oracle expm failed")
    return E


def arnoldi_evolve(csc, k_rt, h, psi):
    """getSolver(idx,k).evolve (sparse_algorithm.hpp:372-393,548-562)."""
    colptr, rowidx, vals = csc
    n = len(colptr) - 1
    psi = np.ascontiguousarray(psi, dtype=np.complex128)
    out = np.empty(n, dtype=np.complex128)
    act = lib().orc_arnoldi_evolve(n, _p(colptr), _p(rowidx), _p(vals), k_rt, h, _p(psi), _p(out))
    if act < 0:
        raise RuntimeError("oracle arnoldi failed")
    return out


def csc_to_dense(csc):
    colptr, rowidx, vals = csc
    n = len(colptr) - 1
    v = vals.view(np.complex128)
    M = np.zeros((n, n), dtype=np.complex128)
    for j in range(n):
        for e in range(colptr[j], colptr[j + 1]):
            M[rowidx[e], j] += v[e]
    return M


def traj_run(opsum, nsite, nup0, ndn0, psi0, channels, ntraj, nsteps, dt, k_rt, seed, traj_offset=0):
    """Quantum-jump loop (hubbard_diss.cpp:440-482) for trajectories traj_offset.. ; returns
    (sum of densities [(nsteps+1), 2 nsite], jump log [ntraj, nsteps])."""
    terms, nt, ops, no = opsum.program()
    M = 2 * nsite
    jk = np.array([c[0] for c in channels], dtype=np.int32)
    jm = np.array([c[1] for c in channels], dtype=np.int32)
    jr = np.array([c[2] for c in channels], dtype=np.float64)
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
    dsum = np.zeros((nsteps + 1) * M)
    log = np.zeros(ntraj * max(nsteps, 1), dtype=np.int32)
    rc = lib().orc_traj_run(nsite, terms, nt, ops, nup0, ndn0, _p(psi0), _p(jk), _p(jm), _p(jr), len(channels),
                            ntraj, traj_offset, nsteps, dt, k_rt, seed, _p(dsum), _p(log))
    if rc != 0:
        raise RuntimeError("oracle trajectories failed")
    return dsum.reshape(nsteps + 1, M), log.reshape(ntraj, max(nsteps, 1))


def traj_run_rebuild(opsum, lsums, nsite, nup0, ndn0, psi0, channels, ntraj, nsteps, dt, k_rt, seed,
                     traj_offset=0):
    """traj_run as the reference runs it (hubbard_diss.cpp:443,462-463,514-529):
    every step the jump operators lsums[c] (e.g. sqrt(g) gate(m, m)) are built
    by `>>`, multiplied by their adjoints and evaluated, and H_eff is assembled
    as a stored matrix before the Arnoldi step.  Same outputs as traj_run."""
    import ctypes
    terms, nt, ops, no = opsum.program()
    progs = [ls.program() for ls in lsums]
    assert len(progs) == len(channels)
    tot_t = sum(p[1] for p in progs)
    tot_o = sum(p[3] for p in progs)
    lt = (type(progs[0][0][0]) * max(1, tot_t))() if progs else None
    lo = (type(progs[0][2][0]) * max(1, tot_o))() if progs else None
    off = np.zeros(len(progs) + 1, dtype=np.int32)
    pt = po = 0
    for c, (t, ntc, o, noc) in enumerate(progs):
        for k in range(ntc):
            ctypes.memmove(ctypes.byref(lt, (pt + k) * ctypes.sizeof(lt[0])), ctypes.byref(t[k]),
                           ctypes.sizeof(lt[0]))
            lt[pt + k].op_begin = t[k].op_begin + po
        for k in range(noc):
            ctypes.memmove(ctypes.byref(lo, (po + k) * ctypes.sizeof(lo[0])), ctypes.byref(o[k]),
                           ctypes.sizeof(lo[0]))
        pt += ntc
        po += noc
        off[c + 1] = pt
    M = 2 * nsite
    jk = np.array([c[0] for c in channels], dtype=np.int32)
    jm = np.array([c[1] for c in channels], dtype=np.int32)
    jr = np.array([c[2] for c in channels], dtype=np.float64)
    psi0 = np.ascontiguousarray(psi0, dtype=np.complex128)
    dsum = np.zeros((nsteps + 1) * M)
    log = np.zeros(ntraj * max(nsteps, 1), dtype=np.int32)
    rc = lib().orc_traj_run_rebuild(nsite, terms, nt, ops, nup0, ndn0, _p(psi0), _p(jk), _p(jm), _p(jr),
                                    len(channels), ntraj, traj_offset, nsteps, dt, k_rt, seed, _p(dsum), _p(log),
                                    lt, _p(off), lo)
    if rc != 0:
        raise RuntimeError("oracle trajectories (rebuild) failed")
    return dsum.reshape(nsteps + 1, M), log.reshape(ntraj, max(nsteps, 1))


# ------------------------------------------------------------ SpMatrix algebra
# Matrices are (outer int32[n+1], inner int32[nnz], vals complex128[nnz]),
# the arrays build_csc returns (spalg.c restates Eigen's rules).
def _c(csc):
    """(outer, inner, vals) with vals complex128, or interleaved float64 as
    build_csc returns them."""
    o, i, v = csc
    v = np.ascontiguousarray(v)
    v = v.view(np.complex128) if v.dtype == np.float64 else v.astype(np.complex128, copy=False)
    return (np.ascontiguousarray(o, dtype=np.int32), np.ascontiguousarray(i, dtype=np.int32),
            np.ascontiguousarray(v))


def csr_add(a, b, sign=1):
    """A + B (sign 1) / A - B (sign -1) (hubbard_diss.cpp:135-156,443,516)."""
    ao, ai, av = _c(a)
    bo, bi, bv = _c(b)
    n = len(ao) - 1
    nnz = lib().orc_csr_add(n, _p(ao), _p(ai), _p(av), _p(bo), _p(bi), _p(bv), sign, None, None, None)
    co = np.empty(n + 1, dtype=np.int32)
    ci = np.empty(max(nnz, 1), dtype=np.int32)
    cv = np.empty(max(nnz, 1), dtype=np.complex128)
    lib().orc_csr_add(n, _p(ao), _p(ai), _p(av), _p(bo), _p(bi), _p(bv), sign, _p(co), _p(ci), _p(cv))
    return co, ci[:nnz], cv[:nnz]


def csr_scale(a, c):
    """c * A; a Python float is a real scalar (Eigen's double * complex)."""
    ao, ai, av = _c(a)
    real = not isinstance(c, complex)
    c = complex(c)
    cv = np.empty_like(av)
    lib().orc_csr_scale(len(ai), c.real, c.imag, 1 if real else 0, _p(av), _p(cv))
    return ao.copy(), ai.copy(), cv


def csr_adjoint(a):
    """SpMatrix(A.adjoint()) (hubbard_diss.cpp:156,526)."""
    ao, ai, av = _c(a)
    n = len(ao) - 1
    co = np.empty(n + 1, dtype=np.int32)
    ci = np.empty(max(len(ai), 1), dtype=np.int32)
    cv = np.empty(max(len(ai), 1), dtype=np.complex128)
    lib().orc_csr_adjoint(n, _p(ao), _p(ai), _p(av), _p(co), _p(ci), _p(cv))
    return co, ci[:len(ai)], cv[:len(ai)]


def csr_multiply(a, b):
    """A * B, Eigen's conservative sparse product (hubbard_diss.cpp:156,526)."""
    ao, ai, av = _c(a)
    bo, bi, bv = _c(b)
    n = len(ao) - 1
    nnz = lib().orc_csr_multiply(n, _p(ao), _p(ai), _p(av), _p(bo), _p(bi), _p(bv), None, None, None)
    co = np.empty(n + 1, dtype=np.int32)
    ci = np.empty(max(nnz, 1), dtype=np.int32)
    cv = np.empty(max(nnz, 1), dtype=np.complex128)
    lib().orc_csr_multiply(n, _p(ao), _p(ai), _p(av), _p(bo), _p(bi), _p(bv), _p(co), _p(ci), _p(cv))
    return co, ci[:nnz], cv[:nnz]


def csr_identity(n):
    return (np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.ones(n, dtype=np.complex128))


def csr_zero(n):
    return np.zeros(n + 1, dtype=np.int32), np.zeros(0, dtype=np.int32), np.zeros(0, dtype=np.complex128)
