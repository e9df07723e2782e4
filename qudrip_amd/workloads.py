"""Benchmark workloads built from the C-ABI handles (used by bench.py and tests).

hubbard_quench: ground state of a square-lattice Hubbard model with Lanczos,
then real-time evolution after a quench of U with the Arnoldi step, recording
the double occupancy <sum_i n_i,up n_i,dn>(t) -- the driver sequence of
programs/hubbard_diss.cpp:377-385,439-470 without dissipation.
"""
import time

from . import operators as Q
from .runtime import CSR, Basis, Krylov, LinOp, TermProgram


def docc_operator(nsite):
    """sum_i n_i,up n_i,dn (number-op products, hubbard_diss.cpp:149-152)."""
    idx = Q.getQbits(2 * nsite)
    g = Q.getHopFermiGate(idx)
    op = Q.getEmptyOperator()
    for i in range(nsite):
        op = op + g(i, i) * g(i + nsite, i + nsite)
    return op


def make_operator(ctx, basis, opsum, nbits, matfree):
    prog = TermProgram(ctx, opsum, nbits)
    if matfree:
        return LinOp.matfree(ctx, basis, prog), [prog]
    csr = CSR(ctx, basis, prog)
    return LinOp.from_csr(ctx, csr), [prog, csr]


def hubbard_quench(ctx, Lx, Ly, nup, ndn, U0, U1, t=1.0, k_gs=120, k_rt=30, dt=0.05, steps=200, seed=3,
                   matfree=True, sync_each_step=False, log=None, two_vector_gs=False, rt_method=1):
    """Ground state at U0, then `steps` real-time steps of H(U1) (Arnoldi
    step, sparse_algorithm.hpp:548-562).  rt_method: reorthogonalisation of
    the real-time basis -- 1 = full (delayed CGS2, the default); 2 = the
    window of the last two vectors with partial reorthogonalisation (Simon's
    estimate; full passes only when it passes the trigger), valid because
    H(U1) is Hermitian."""
    import torch
    nsite = Lx * Ly
    bonds = Q.hubbard_bonds_square(Lx, Ly)
    idx0, h0 = Q.hubbard(nsite, bonds, t, U0)
    sec = Q.hubbard_sector(idx0, nsite, nup, ndn)
    t0 = time.perf_counter()
    basis = Basis(ctx, sec)
    op0, keep0 = make_operator(ctx, basis, h0, 2 * nsite, matfree)
    _, h1 = Q.hubbard(nsite, bonds, t, U1)
    op1, keep1 = make_operator(ctx, basis, h1, 2 * nsite, matfree)
    dop, keepd = make_operator(ctx, basis, docc_operator(nsite), 2 * nsite, True)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    n = basis.size

    psi = ctx.seeded_state(n, seed)
    kr = Krylov(ctx, n, k_gs, herm=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0 = kr.ground_state(op0, psi)
    torch.cuda.synchronize()
    gs_s = time.perf_counter() - t0
    kr.close()
    nrm = ctx.norm2(psi)
    docc = [op_expect(dop, psi) / nrm]
    energy = [op_expect(op1, psi).real / nrm]
    if log:
        log(f"hubbard {Lx}x{Ly} n={n}: E0(U={U0}) = {e0:.12f} in {gs_s:.3f} s; docc(0) = {docc[0]:.10f}")
    two = None
    if two_vector_gs:
        # the same ground state through Hermitian::getSolver (the reference
        # driver's own call, programs/hubbard_diss.cpp:382-384): two-vector
        # Lanczos, no reorthogonalisation, two passes of the recurrence
        from .runtime import HermitianLanczos
        hl = HermitianLanczos(ctx, n, k_gs)
        psi2 = ctx.seeded_state(n, seed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e2 = hl.ground_state(op0, psi2)
        torch.cuda.synchronize()
        two = {"E0": e2, "solve_s": time.perf_counter() - t0, "coefficients": hl.info()["dim"],
               "rayleigh": op_expect(op0, psi2) / ctx.norm2(psi2)}
        hl.close()
        del psi2
        if log:
            log(f"hubbard {Lx}x{Ly}: two-vector Lanczos E0 = {e2:.12f} in {two['solve_s']:.3f} s")
    kr_rt = Krylov(ctx, n, k_rt, herm=rt_method >= 2)
    if rt_method != 1:
        kr_rt.set_method(rt_method)
    full_steps = 0
    nxt = ctx.zeros(n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        kr_rt.evolve(op1, psi, nxt, dt, k_rt)
        if rt_method >= 2:
            full_steps += kr_rt.info()["full_steps"]
        psi, nxt = nxt, psi
        nrm = ctx.norm2(psi)
        docc.append(op_expect(dop, psi) / nrm)
        if s == steps - 1 or sync_each_step:
            energy.append(op_expect(op1, psi).real / nrm)
    torch.cuda.synchronize()
    rt_s = time.perf_counter() - t0
    kr_rt.close()
    for h in [op0, op1, dop] + keep0 + keep1 + keepd + [basis]:
        h.close()
    return {"n": n, "E0": e0, "docc": docc, "energy_start": energy[0], "energy_end": energy[-1],
            "norm_end": nrm, "setup_s": setup_s, "gs_s": gs_s, "rt_s": rt_s, "steps": steps,
            "rt_step_ms": rt_s * 1e3 / max(steps, 1), "psi": psi, "two_vector_gs": two,
            "rt_method": rt_method, "rt_full_steps": full_steps}


def op_expect(op, psi):
    v = op.expect(psi)
    return v.real


def hubbard_trajectories(ctx, L=8, nup=4, ndn=4, U=4.0, t=1.0, dephase=0.1, loss=0.05, ntraj=4096, T=10.0,
                         dt=0.05, k_gs=100, k_rt=13, seed=5, want_log=False, log=None):
    """Dissipative Hubbard chain (open boundaries): ground state of the
    (nup, ndn) sector by Lanczos, then ntraj quantum-jump trajectories with
    dephasing sqrt(dephase) n_m and loss sqrt(loss) c_m on every mode
    (programs/hubbard_diss.cpp:377-385,439-482 with loss channels added)."""
    import torch
    from .runtime import jump_channels, traj_run
    bonds = Q.hubbard_bonds_chain(L, False)
    idx, ham = Q.hubbard(L, bonds, t, U)
    basis = Basis(ctx, Q.hubbard_sector(idx, L, nup, ndn))
    op, keep = make_operator(ctx, basis, ham, 2 * L, False)
    psi = ctx.seeded_state(basis.size, seed)
    kr = Krylov(ctx, basis.size, k_gs, herm=True)
    e0 = kr.ground_state(op, psi)
    kr.close()
    nrm = ctx.norm2(psi) ** 0.5
    psi = psi / nrm
    torch.cuda.synchronize()
    nsteps = int(round(T / dt))
    ch = jump_channels(L, dephase, loss)
    t0 = time.perf_counter()
    res = traj_run(ctx, ham, L, nup, ndn, psi, ch, ntraj, nsteps, dt, k_rt, seed, want_log=want_log)
    torch.cuda.synchronize()
    run_s = time.perf_counter() - t0
    for h in [op] + keep + [basis]:
        h.close()
    res.update({"E0": e0, "psi0": psi, "run_s": run_s, "nsteps": nsteps, "channels": ch, "n": basis.size,
                "ham": ham})
    if log:
        log(f"trajectories L={L} n={basis.size}: {ntraj} x {nsteps} steps in {run_s:.3f} s, "
            f"{res['jumps']} jumps ({res['losses']} losses), {res['sectors']} sectors")
    return res


_M64 = (1 << 64) - 1


def _sm64(z):
    z = (z + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def _traj_uniform(seed, traj, step, k):
    """The trajectory stream of qd_traj_run: key = sm(seed * golden + traj), r = sm(key + 4 step + k) >> 11."""
    key = _sm64((seed * 0x9E3779B97F4A7C15 + traj) & _M64)
    return (_sm64((key + 4 * step + k) & _M64) >> 11) * (1.0 / 9007199254740992.0)


def plain_trajectories(ctx, ham, nsite, nup0, ndn0, psi0, channels, ntraj, nsteps, dt, k_rt, seed, traj_offset=0):
    """One trajectory at a time on the single-vector GPU path, the loop of
    programs/hubbard_diss.cpp:440-482 with the channels of qd_traj_run: per
    step the densities of the state, the jump test on the same random stream,
    then either the jump or evolve(H_eff, psi, dt, k_rt) with
    H_eff = H - i/2 (D - Gamma) as the stored H plus a diagonal, and a
    renormalisation.  Returns (sum of densities [(nsteps+1), 2 nsite],
    jump log [ntraj, nsteps], seconds)."""
    import numpy as np
    import torch
    dev = f"cuda:{ctx.device}"
    M = 2 * nsite
    idx = Q.getQbits(M)
    sectors = {}

    def sector(a, b):
        if (a, b) in sectors:
            return sectors[(a, b)]
        basis = Basis(ctx, Q.hubbard_sector(idx, nsite, a, b))
        prog = TermProgram(ctx, ham, M)
        csr = CSR(ctx, basis, prog)
        op = LinOp.from_csr(ctx, csr)
        st = basis.states()
        occ = ((st[:, None] >> np.arange(M, dtype=np.uint64)[None, :]) & np.uint64(1)).astype(np.float64)
        D = np.zeros(len(st))
        for (kind, m, g) in channels:
            D += g * occ[:, m]
        S = {"basis": basis, "prog": prog, "csr": csr, "op": op, "states": st, "n": len(st),
             "occ": torch.from_numpy(occ).to(dev), "D": torch.from_numpy(D).to(dev),
             "kr": Krylov(ctx, max(1, len(st)), k_rt, herm=False), "loss": {}}
        sectors[(a, b)] = S
        return S

    def loss_map(a, b, m):
        S = sector(a, b)
        if m in S["loss"]:
            return S["loss"][m]
        na, nb = (a - 1, b) if m < nsite else (a, b - 1)
        T = sector(na, nb)
        st = S["states"]
        src = np.nonzero((st >> np.uint64(m)) & np.uint64(1))[0]
        y = st[src] & ~np.uint64(1 << m)
        dst = np.searchsorted(T["states"], y)
        par = np.array([bin(int(x) >> (m + 1)).count("1") & 1 for x in st[src]])
        sign = torch.from_numpy(np.where(par == 1, -1.0, 1.0)).to(dev).to(torch.complex128)
        S["loss"][m] = (na, nb, torch.from_numpy(src).to(dev), torch.from_numpy(dst).to(dev), sign)
        return S["loss"][m]

    rates = np.array([c[2] for c in channels])
    modes = np.array([c[1] for c in channels])
    dsum = np.zeros((nsteps + 1, M))
    log = np.full((ntraj, max(nsteps, 1)), -1, dtype=np.int64)
    psi0 = torch.as_tensor(psi0).to(dev).to(torch.complex128)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for tr in range(ntraj):
        a, b = nup0, ndn0
        S = sector(a, b)
        psi = psi0 / torch.linalg.vector_norm(psi0)
        for step in range(nsteps + 1):
            p = psi.real ** 2 + psi.imag ** 2
            dens = ((p @ S["occ"]) / p.sum()).cpu().numpy()
            dsum[step] += dens
            if step == nsteps:
                break
            w = rates * dens[modes]
            G = float(w.sum())
            r0 = _traj_uniform(seed, traj_offset + tr, step, 0)
            r1 = _traj_uniform(seed, traj_offset + tr, step, 1)
            ch = -1
            if r0 < G * dt:
                cum = 0.0
                for c in range(len(channels)):
                    cum += w[c]
                    if w[c] > 0.0:
                        ch = c
                    if r1 * G < cum and w[c] > 0.0:
                        break
            log[tr, step] = ch
            if ch >= 0:
                kind, m, _ = channels[ch]
                if kind == 0:
                    psi = psi * S["occ"][:, m]
                else:
                    na, nb, src, dst, sign = loss_map(a, b, m)
                    T = sector(na, nb)
                    out = torch.zeros(T["n"], dtype=torch.complex128, device=dev)
                    out[dst] = psi[src] * sign
                    psi, S, a, b = out, T, na, nb
                psi = psi / torch.linalg.vector_norm(psi)
            else:
                diag = (-0.5j) * (S["D"] - G).to(torch.complex128)
                heff = LinOp.with_diagonal(ctx, S["op"], diag)
                nxt = torch.empty_like(psi)
                S["kr"].evolve(heff, psi, nxt, dt, k_rt)
                heff.close()
                psi = nxt / torch.linalg.vector_norm(nxt)
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    for S in sectors.values():
        S["kr"].close()
        for h in (S["op"], S["csr"], S["prog"], S["basis"]):
            h.close()
    return dsum, log, secs
