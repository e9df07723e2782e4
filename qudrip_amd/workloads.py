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
                   matfree=True, sync_each_step=False, log=None):
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
    kr_rt = Krylov(ctx, n, k_rt, herm=False)
    nxt = ctx.zeros(n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        kr_rt.evolve(op1, psi, nxt, dt, k_rt)
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
            "rt_step_ms": rt_s * 1e3 / max(steps, 1), "psi": psi}


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
