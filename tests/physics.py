"""Independent dense constructions from the physics definitions (numpy only).

These never touch the operator-tree code: spin models are built from bit
flips of the full 2^N space, fermion models from Jordan-Wigner Kronecker
products c_j = (prod_{k>j} Z_k) a_j.  They pin the oracle's tree expansion
and CSC build on small systems.
"""
import numpy as np


def popcount(x):
    return bin(int(x)).count("1")


def sector_states(nbits, masks_counts):
    xs = np.arange(1 << nbits, dtype=np.uint64)
    ok = np.ones(len(xs), dtype=bool)
    for mask, cnt in masks_counts:
        pc = np.array([popcount(int(x) & mask) for x in xs])
        ok &= pc == cnt
    return xs[ok]


def dense_xxz_full(N, Jxy, Jz, pbc):
    """H = sum_<ij> [Jxy/2 (flip a 1 between i and j) + Jz s_i s_j], s = +1/2 for bit 0."""
    D = 1 << N
    H = np.zeros((D, D), dtype=np.complex128)
    bonds = [(i, (i + 1) % N) for i in range(N if pbc else N - 1)]
    for x in range(D):
        for (i, j) in bonds:
            bi, bj = (x >> i) & 1, (x >> j) & 1
            si, sj = (0.5 if bi == 0 else -0.5), (0.5 if bj == 0 else -0.5)
            H[x, x] += Jz * si * sj
            if bi != bj:
                y = x ^ (1 << i) ^ (1 << j)
                H[y, x] += 0.5 * Jxy
    return H


def dense_xy_full(N, J=1.0, pbc=False):
    return dense_xxz_full(N, 2.0 * J, 0.0, pbc)


def _kron_list(mats):
    out = np.array([[1.0 + 0j]])
    for m in mats:
        out = np.kron(out, m)
    return out


def jw_annihilators(nmodes):
    """c_j with bit j of the state index <-> mode j; string over modes k > j."""
    a = np.array([[0, 1], [0, 0]], dtype=np.complex128)
    Z = np.diag([1.0, -1.0]).astype(np.complex128)
    I2 = np.eye(2, dtype=np.complex128)
    cs = []
    for j in range(nmodes):
        mats = []
        for k in range(nmodes - 1, -1, -1):  # kron order: highest bit first
            mats.append(Z if k > j else (a if k == j else I2))
        cs.append(_kron_list(mats))
    return cs


def dense_hubbard_full(nsite, bonds, t, U):
    cs = jw_annihilators(2 * nsite)
    D = 1 << (2 * nsite)
    H = np.zeros((D, D), dtype=np.complex128)
    for (i, j) in bonds:
        for off in (0, nsite):
            a, b = i + off, j + off
            H += -t * (cs[a].conj().T @ cs[b] + cs[b].conj().T @ cs[a])
    for i in range(nsite):
        nu = cs[i].conj().T @ cs[i]
        nd = cs[i + nsite].conj().T @ cs[i + nsite]
        H += U * (nu @ nd)
    return H


def restrict(H, states):
    idx = np.array(states, dtype=np.int64)
    return H[np.ix_(idx, idx)]


def xy_free_fermion_e0(N, nup):
    eps = sorted(2.0 * np.cos(np.pi * k / (N + 1)) for k in range(1, N + 1))
    return float(sum(eps[:nup]))
