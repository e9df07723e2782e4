"""Solve server handle (qd_server_* of include/qudrip_hip.h) and the mixed
request workload built on it.

Workload: Poisson arrivals (rate 200 / s, seeded) of
  - 70 %: ground states of periodic XXZ chains, N uniform in {16, ..., 26},
    Sz = 0, Jxy = 1, Jz ~ U[0, 2]   (programs/spin.cpp's solve);
  - 30 %: real-time solves on the 3 x 4 periodic Hubbard lattice,
    N_up = N_dn = 6 (n = 853,776), t = 1, U ~ U[2, 8], 100 steps of dt = 0.05
    from a seeded start vector (hubbard_diss.cpp:439-470 without dissipation).
"Plain" solves every request alone, one after another, the way a user of the
reference does: `ham >> chain` (basis and stored matrix of the full H), then
getLanczosSolver(chain, 150).getGroundState(psi) or 100 evolve calls with
kry_dim_rt = 30.  "Served" registers one model per lattice and sector, then
streams the same requests into the server.
"""
import ctypes
import time

import numpy as np

from . import abi
from . import operators as Q


class Server:
    def __init__(self, device=0, hbm_budget=200 << 30, workers=6, batch=1, gs_method=2):
        self.lib = abi.load()
        p = ctypes.c_void_p()
        rc = self.lib.qd_server_create(device, hbm_budget, workers, ctypes.byref(p))
        if rc:
            raise abi.QDError(rc, "qd_server_create")
        self.ptr = p
        self._check(self.lib.qd_server_set_batch(self.ptr, batch), "set_batch")
        self._check(self.lib.qd_server_set_gs_method(self.ptr, gs_method), "set_gs_method")

    def _check(self, rc, what):
        if rc:
            msg = self.lib.qd_server_last_error(self.ptr)
            raise abi.QDError(rc, f"{what}: {msg.decode() if msg else ''}")

    def add_model(self, sector, A, D=None, structured=False):
        cons, ncons = sector.ctypes_constraints()
        ta, nta, oa, noa = A.program()
        if D is not None and D.size():
            td, ntd, od, nod = D.program()
        else:
            td, ntd, od, nod = None, 0, None, 0
        mid = ctypes.c_int()
        self._check(self.lib.qd_server_add_model(self.ptr, sector.nbits, cons, ncons, ta, nta, oa, noa, td, ntd, od,
                                                 nod, 1 if structured else 0, ctypes.byref(mid)), "add_model")
        return mid.value

    def submit(self, kind, model, coeff, k, chunk=10, tol=0.0, steps=0, dt=0.0, seed=0, tag=0):
        r = abi.qd_request(kind=kind, model=model, coeff=coeff, k=k, chunk=chunk, tol=tol, steps=steps, dt=dt,
                           seed=seed, tag=tag)
        self._check(self.lib.qd_server_submit(self.ptr, ctypes.byref(r)), "submit")

    def cancel(self, tag):
        """Cancels the queued or running requests with this tag (they finish
        with status QD_E_CANCELLED); False if none is queued or running."""
        rc = self.lib.qd_server_cancel(self.ptr, tag)
        if rc == abi.STATUS_CODE["QD_E_ARG"]:
            return False
        self._check(rc, "cancel")
        return True

    def poll(self, cap=1024):
        buf = (abi.qd_result * cap)()
        n = ctypes.c_int()
        self._check(self.lib.qd_server_poll(self.ptr, buf, cap, ctypes.byref(n)), "poll")
        out = []
        for i in range(n.value):
            r = buf[i]
            out.append({"tag": r.tag, "status": r.status, "iterations": r.iterations, "energy": r.energy,
                        "energy_start": r.energy_start, "observable": r.observable, "t_submit": r.t_submit,
                        "t_start": r.t_start, "t_end": r.t_end, "krylov_mean": r.krylov_mean,
                        "worker": r.worker})
        return out

    def drain(self):
        self._check(self.lib.qd_server_drain(self.ptr), "drain")

    def clock(self):
        return self.lib.qd_server_clock(self.ptr)

    def stats(self):
        st = abi.qd_server_stats()
        self._check(self.lib.qd_server_get_stats(self.ptr, ctypes.byref(st)), "stats")
        return {f: getattr(st, f) for f, _ in abi.qd_server_stats._fields_}

    def close(self):
        if self.ptr:
            self.lib.qd_server_destroy(self.ptr)
            self.ptr = None


# ------------------------------------------------------------------ workload
XXZ_SIZES = (16, 18, 20, 22, 24, 26)


def request_stream(n_req, rate=200.0, seed=4, gs_frac=0.7, sizes=XXZ_SIZES, u_range=(2.0, 8.0),
                   jz_range=(0.0, 2.0)):
    """Seeded Poisson stream: arrival times (exponential gaps) and the mix."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n_req):
        t += rng.exponential(1.0 / rate)
        if rng.random() < gs_frac:
            out.append({"tag": i, "arrival": t, "kind": "gs", "N": int(rng.choice(sizes)),
                        "coeff": float(rng.uniform(*jz_range)), "seed": 1000 + i})
        else:
            out.append({"tag": i, "arrival": t, "kind": "rt", "coeff": float(rng.uniform(*u_range)),
                        "seed": 1000 + i})
    return out


def hubbard_lattice(Lx=3, Ly=4, nup=6, ndn=6):
    nsite = Lx * Ly
    return nsite, Q.hubbard_bonds_square(Lx, Ly), nup, ndn


def run_served(reqs, lattice=None, workers=6, hbm_budget=200 << 30, gs_k=150, gs_chunk=4, gs_tol=1e-12,
               rt_k=30, rt_steps=100, dt=0.05, rt_tol=1e-13, rt_mmin=16, batch=1, gs_method=2, log=None,
               closed_loop=False):
    """Registers the models (timed apart: the structure cache is built once at
    start-up), then submits every request at its arrival time and waits.
    closed_loop: the requests go one at a time instead (each submitted when
    the previous one has finished, arrival times ignored): the server's
    algorithms and cached structure with no overlap between requests."""
    nsite, bonds, nup, ndn = lattice or hubbard_lattice()
    srv = Server(hbm_budget=hbm_budget, workers=workers, batch=batch, gs_method=gs_method)
    try:
        t0 = time.perf_counter()
        models = {}
        for N in sorted({r["N"] for r in reqs if r["kind"] == "gs"}):
            idx, A, D = Q.xxz_parts(N, 1.0, pbc=True)
            models[("gs", N)] = srv.add_model(Q.sz_sector(idx, N, N // 2), A, D)
        if any(r["kind"] == "rt" for r in reqs):
            idx, T, D = Q.hubbard_parts(nsite, bonds, 1.0)
            models["rt"] = srv.add_model(Q.hubbard_sector(idx, nsite, nup, ndn), T, D, structured=True)
        setup_s = time.perf_counter() - t0
        if log:
            log(f"served: {len(models)} models registered in {setup_s:.2f} s")
        c0 = srv.clock()
        for r in sorted(reqs, key=lambda r: r["arrival"]):
            if closed_loop:
                srv.drain()
            wait = c0 + r["arrival"] - srv.clock()
            if wait > 0 and not closed_loop:
                time.sleep(wait)
            if r["kind"] == "gs":
                srv.submit(abi.QD_REQ_GROUND_STATE, models[("gs", r["N"])], r["coeff"], gs_k, chunk=gs_chunk,
                           tol=gs_tol, seed=r["seed"], tag=r["tag"])
            else:
                srv.submit(abi.QD_REQ_EVOLVE, models["rt"], r["coeff"], rt_k, chunk=rt_mmin, tol=rt_tol,
                           steps=rt_steps, dt=dt, seed=r["seed"], tag=r["tag"])
        srv.drain()
        res = {x["tag"]: x for x in srv.poll(cap=max(16, len(reqs)))}
        stats = srv.stats()
    finally:
        srv.close()
    for r in reqs:
        x = res[r["tag"]]
        x["arrival"] = x["t_submit"] if closed_loop else c0 + r["arrival"]
        x["latency"] = x["t_end"] - x["arrival"]
    span = max(x["t_end"] for x in res.values()) - min(x["arrival"] for x in res.values())
    return {"results": res, "stats": stats, "setup_s": setup_s, "span_s": span,
            "solves_per_s": len(reqs) / span}


def run_plain(reqs, ctx, lattice=None, gs_k=150, rt_k=30, rt_steps=100, dt=0.05, log=None):
    """One request at a time through the drop-in path: basis and stored matrix
    of the full H per request, then the fixed-dimension solver calls."""
    import torch

    from .runtime import CSR, Basis, Krylov, LinOp, TermProgram
    nsite, bonds, nup, ndn = lattice or hubbard_lattice()
    out = {}
    for r in sorted(reqs, key=lambda r: r["arrival"]):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if r["kind"] == "gs":
            N = r["N"]
            idx, ham = Q.xxz_chain(N, 1.0, r["coeff"], pbc=True)
            sec = Q.sz_sector(idx, N, N // 2)
            nbits = N
        else:
            idx, ham = Q.hubbard(nsite, bonds, 1.0, r["coeff"])
            sec = Q.hubbard_sector(idx, nsite, nup, ndn)
            nbits = 2 * nsite
        basis = Basis(ctx, sec)
        prog = TermProgram(ctx, ham, nbits)
        csr = CSR(ctx, basis, prog)
        op = LinOp.from_csr(ctx, csr)
        psi = ctx.seeded_state(basis.size, r["seed"])
        if r["kind"] == "gs":
            kr = Krylov(ctx, basis.size, gs_k, herm=True)
            e = kr.ground_state(op, psi)
            kr.close()
            x = {"energy": e, "iterations": gs_k}
        else:
            e0 = op.expect(psi).real / ctx.norm2(psi)
            kr = Krylov(ctx, basis.size, rt_k, herm=False)
            nxt = ctx.zeros(basis.size)
            for _ in range(rt_steps):
                kr.evolve(op, psi, nxt, dt, rt_k)
                psi, nxt = nxt, psi
            kr.close()
            nn = ctx.norm2(psi)
            dprog = TermProgram(ctx, Q.hubbard_parts(nsite, bonds, 1.0)[2], nbits)
            dop = LinOp.matfree(ctx, basis, dprog)
            x = {"energy": op.expect(psi).real / nn, "energy_start": e0, "iterations": rt_steps,
                 "observable": dop.expect(psi).real / nn}
            dop.close()
            dprog.close()
        for h in (op, csr, prog, basis):
            h.close()
        torch.cuda.synchronize()
        x["time_s"] = time.perf_counter() - t0
        out[r["tag"]] = x
        del psi
    # latency if the same stream were queued in front of the plain path (FIFO)
    done = 0.0
    for r in sorted(reqs, key=lambda r: r["arrival"]):
        x = out[r["tag"]]
        done = max(done, r["arrival"]) + x["time_s"]
        x["latency"] = done - r["arrival"]
    total = sum(x["time_s"] for x in out.values())
    return {"results": out, "busy_s": total, "solves_per_s": len(reqs) / total}


def latency_percentiles(results):
    """p50 / p99 of arrival -> result (queueing included)."""
    lat = np.array([x["latency"] for x in results.values()])
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


def service_percentiles(results):
    """p50 / p99 of start -> result on a worker (the solve alone)."""
    t = np.array([x["t_end"] - x["t_start"] if "t_start" in x else x["time_s"] for x in results.values()])
    return float(np.percentile(t, 50)), float(np.percentile(t, 99))
