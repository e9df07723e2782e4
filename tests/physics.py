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


# ---------------------------------------------------------------- trajectories
_M64 = (1 << 64) - 1


def splitmix64(z):
    """splitmix64 finaliser (Steele, Lea & Flood, OOPSLA 2014), on Python ints."""
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def traj_uniform(seed, traj, step, k):
    """r_k(step) of trajectory `traj`: key = sm(seed*golden + traj), r = sm(key + 4 step + k) >> 11 * 2^-53."""
    key = splitmix64((seed * 0x9E3779B97F4A7C15 + traj) & _M64)
    return (splitmix64((key + 4 * step + k) & _M64) >> 11) * (1.0 / 9007199254740992.0)


def dense_jump_trajectories(H, nmodes, psi_full, channels, ntraj, nsteps, dt, seed, traj_offset=0):
    """First-order quantum-jump unravelling on the full Fock space with dense
    operators: weights w_c = g_c <n_m>, jump iff r0 < Gamma dt (channel by the
    cumulative weights), else psi <- exp(-i H_eff dt) psi with
    H_eff = H - i/2 (sum_c g_c L_c^+ L_c - Gamma); normalise every step.
    Returns (densities [ntraj, nsteps+1, nmodes], jump log [ntraj, nsteps])."""
    import scipy.linalg
    D = H.shape[0]
    cs = jw_annihilators(nmodes)
    occ = np.array([[(x >> m) & 1 for m in range(nmodes)] for x in range(D)], dtype=np.float64)
    Dsum = np.zeros(D)
    for kind, m, g in channels:
        Dsum += g * occ[:, m]
    # the -Gamma shift only rescales the propagated state, which is renormalised
    prop = scipy.linalg.expm(-1j * H * dt - 0.5 * dt * np.diag(Dsum))
    dens = np.zeros((ntraj, nsteps + 1, nmodes))
    log = np.full((ntraj, max(nsteps, 1)), -1, dtype=np.int64)
    for tr in range(ntraj):
        psi = psi_full / np.linalg.norm(psi_full)
        for step in range(nsteps + 1):
            p = np.abs(psi) ** 2
            d = occ.T @ p / p.sum()
            dens[tr, step] = d
            if step == nsteps:
                break
            w = np.array([g * d[m] for (_, m, g) in channels])
            G = w.sum()
            r0 = traj_uniform(seed, traj_offset + tr, step, 0)
            r1 = traj_uniform(seed, traj_offset + tr, step, 1)
            ch = -1
            if r0 < G * dt:
                cum = np.cumsum(w)
                ok = np.nonzero((r1 * G < cum) & (w > 0))[0]
                ch = int(ok[0]) if len(ok) else int(np.nonzero(w > 0)[0][-1])
            log[tr, step] = ch
            if ch >= 0:
                kind, m, _ = channels[ch]
                psi = occ[:, m] * psi if kind == 0 else cs[m] @ psi
            else:
                psi = prop @ psi
            psi = psi / np.linalg.norm(psi)
    return dens, log


def lindblad_densities(H, nmodes, rho0, channels, times):
    """<n_m>(t) of the Lindblad master equation d rho/dt = -i[H,rho] + sum_c (L rho L^+ - 1/2 {L^+L, rho}),
    L = sqrt(g) n_m or sqrt(g) c_m, by the exponential of the Liouvillian (column-stacked vec)."""
    import scipy.linalg
    D = H.shape[0]
    cs = jw_annihilators(nmodes)
    I = np.eye(D)
    Lv = -1j * (np.kron(I, H) - np.kron(H.T, I))
    for kind, m, g in channels:
        L = np.sqrt(g) * (cs[m].conj().T @ cs[m] if kind == 0 else cs[m])
        LdL = L.conj().T @ L
        Lv += np.kron(L.conj(), L) - 0.5 * (np.kron(I, LdL) + np.kron(LdL.T, I))
    occ = np.array([[(x >> m) & 1 for m in range(nmodes)] for x in range(D)], dtype=np.float64)
    out = []
    v0 = rho0.reshape(-1, order="F")
    for t in times:
        rho = (scipy.linalg.expm(Lv * t) @ v0).reshape(D, D, order="F")
        out.append(occ.T @ np.real(np.diag(rho)))
    return np.array(out)
