"""Solve server (qd_server_*) on the GPU: results against the plain one-at-a-
time path and exact diagonalisation, identity alone vs in the mix, and
admission by the HBM budget (deferral and rejection)."""
import numpy as np
import pytest

from qudrip_amd import abi
from qudrip_amd import operators as Q
from qudrip_amd import serving as SV
from tests import physics as P

pytestmark = pytest.mark.gpu

LATTICE = (6, Q.hubbard_bonds_square(2, 3), 3, 3)  # 2 x 3 periodic, N_up = N_dn = 3 (n = 400)


def _reqs():
    out = []
    for i, (N, jz) in enumerate([(10, 0.3), (12, 1.0), (10, 1.7), (12, 0.05), (10, 1.0)]):
        out.append({"tag": i, "arrival": 0.002 * i, "kind": "gs", "N": N, "coeff": jz, "seed": 100 + i})
    for i, u in enumerate([3.0, 6.5, 2.2]):
        out.append({"tag": 10 + i, "arrival": 0.001 + 0.002 * i, "kind": "rt", "coeff": u, "seed": 200 + i})
    return out


KW = dict(gs_k=80, gs_chunk=10, gs_tol=1e-13, rt_k=20, rt_steps=25, dt=0.05)


def _dense_e0(N, jz):
    H = P.dense_xxz_full(N, 1.0, jz, True)
    st = P.sector_states(N, [((1 << N) - 1, N // 2)])
    return np.linalg.eigvalsh(P.restrict(H, st))[0]


@pytest.mark.parametrize("gs_method", [2, 1])
def test_served_matches_plain_and_exact(ctx, gs_method):
    """Ground states with partial (2, default) or full (1) reorthogonalisation."""
    reqs = _reqs()
    print("serving mix of", len(reqs), "gs_method", gs_method, flush=True)
    served = SV.run_served(reqs, lattice=LATTICE, workers=4, hbm_budget=8 << 30, gs_method=gs_method, **KW)
    plain = SV.run_plain(reqs, ctx, lattice=LATTICE, gs_k=KW["gs_k"], rt_k=KW["rt_k"], rt_steps=KW["rt_steps"],
                         dt=KW["dt"])
    for r in reqs:
        s, p = served["results"][r["tag"]], plain["results"][r["tag"]]
        assert s["status"] == 0
        if r["kind"] == "gs":
            e = _dense_e0(r["N"], r["coeff"])
            assert abs(s["energy"] - e) <= 1e-10 * abs(e), (r, s["energy"], e)
            assert abs(p["energy"] - e) <= 1e-10 * abs(e)
            assert s["iterations"] < KW["gs_k"] or s["iterations"] == KW["gs_k"]
        else:
            assert abs(s["energy"] - p["energy"]) <= 1e-10 * abs(p["energy"])
            assert abs(s["observable"] - p["observable"]) <= 1e-10
            assert abs(s["energy"] - s["energy_start"]) <= 1e-9 * abs(s["energy_start"])
    assert served["stats"]["max_running"] >= 2


def test_result_same_alone_and_in_the_mix(ctx):
    reqs = _reqs()
    mix = SV.run_served(reqs, lattice=LATTICE, workers=4, hbm_budget=8 << 30, **KW)["results"]
    for r in reqs:
        alone = SV.run_served([dict(r, arrival=0.0)], lattice=LATTICE, workers=1, hbm_budget=8 << 30,
                              **KW)["results"][r["tag"]]
        for key in ("energy", "energy_start", "observable"):
            assert abs(alone[key] - mix[r["tag"]][key]) <= 1e-12 * max(1.0, abs(alone[key])), (r, key)
        assert alone["iterations"] == mix[r["tag"]]["iterations"]


@pytest.mark.parametrize("batch", [1, 16])
def test_evolve_batching_matches_single(ctx, batch):
    """Lockstep batches of evolve requests (joining at step boundaries) give the
    plain path's results; batch = 1 runs each request alone."""
    reqs = [{"tag": i, "arrival": 0.003 * i, "kind": "rt", "coeff": u, "seed": 300 + i}
            for i, u in enumerate([2.5, 7.0, 4.1, 5.5, 3.3])]
    served = SV.run_served(reqs, lattice=LATTICE, workers=2, hbm_budget=8 << 30, batch=batch, **KW)
    plain = SV.run_plain(reqs, ctx, lattice=LATTICE, gs_k=KW["gs_k"], rt_k=KW["rt_k"], rt_steps=KW["rt_steps"],
                         dt=KW["dt"])
    for r in reqs:
        s, p = served["results"][r["tag"]], plain["results"][r["tag"]]
        assert s["status"] == 0 and s["iterations"] == KW["rt_steps"]
        assert abs(s["energy_start"] - p["energy_start"]) <= 1e-12 * abs(p["energy_start"])
        assert abs(s["energy"] - p["energy"]) <= 1e-10 * abs(p["energy"])
        assert abs(s["observable"] - p["observable"]) <= 1e-10
        assert 2 <= s["krylov_mean"] <= KW["rt_k"] - 1


def test_admission_defers_and_rejects(ctx):
    idx, A, D = Q.xxz_parts(12, 1.0, pbc=True)
    sec = Q.sz_sector(idx, 12, 6)
    probe = SV.Server(hbm_budget=1 << 30, workers=1)
    probe.add_model(sec, A, D)
    model_bytes = probe.stats()["bytes_in_use"]
    probe.close()
    n = 924
    fp = (60 + 6) * n * 16 + (8 << 20)  # footprint of a k = 60 ground-state request
    srv = SV.Server(hbm_budget=model_bytes + fp + fp // 2, workers=4)  # room for one request at a time
    try:
        m = srv.add_model(sec, A, D)
        for i in range(5):
            srv.submit(abi.QD_REQ_GROUND_STATE, m, 0.5 + 0.1 * i, 60, chunk=10, tol=1e-12, seed=7, tag=i)
        srv.submit(abi.QD_REQ_GROUND_STATE, m, 1.0, 200000, seed=7, tag=99)  # can never fit
        srv.drain()
        res = {x["tag"]: x for x in srv.poll()}
        st = srv.stats()
    finally:
        srv.close()
    assert res[99]["status"] == 2  # QD_E_OOM
    assert all(res[i]["status"] == 0 for i in range(5))
    assert st["max_running"] == 1 and st["rejected"] == 1 and st["finished"] == 6
    for i in range(5):
        e = _dense_e0(12, 0.5 + 0.1 * i)
        assert abs(res[i]["energy"] - e) <= 1e-10 * abs(e)


def test_served_config5_sizes_match_plain_and_oracle(ctx):
    """Requests of the serving workload's own sizes: XXZ ground states at N = 16
    and N = 24 (n = 12,870 and 2,704,156; the workload's Lanczos k = 150,
    chunks of 4, Ritz-bound stop 1e-12) and one 3x4 Hubbard real-time solve
    (n = 853,776, 100 steps of dt = 0.05, Krylov dimension by Saad's
    estimate), against the plain one-at-a-time path and, for N = 16, the
    oracle's Lanczos (sparse_algorithm.hpp:395-425,564-585) on the oracle's
    own matrix."""
    import oracle
    reqs = [{"tag": 0, "arrival": 0.0, "kind": "gs", "N": 16, "coeff": 0.8321, "seed": 41},
            {"tag": 1, "arrival": 0.001, "kind": "rt", "coeff": 5.4, "seed": 42},
            {"tag": 2, "arrival": 0.002, "kind": "gs", "N": 24, "coeff": 1.37, "seed": 43}]
    print("serving at the workload's sizes", flush=True)
    served = SV.run_served(reqs, workers=4, hbm_budget=64 << 30)
    plain = SV.run_plain(reqs, ctx)
    for r in reqs:
        s, p = served["results"][r["tag"]], plain["results"][r["tag"]]
        assert s["status"] == 0, r
        assert abs(s["energy"] - p["energy"]) <= 1e-10 * abs(p["energy"]), (r, s["energy"], p["energy"])
        if r["kind"] == "rt":
            assert abs(s["energy_start"] - p["energy_start"]) <= 1e-12 * abs(p["energy_start"])
            assert abs(s["observable"] - p["observable"]) <= 1e-9
            assert abs(s["energy"] - s["energy_start"]) <= 1e-9 * abs(s["energy_start"])
    idx, ham = Q.xxz_chain(16, 1.0, 0.8321, pbc=True)
    basis = oracle.basis_scan(Q.sz_sector(idx, 16, 8))
    csc = oracle.build_csc(basis, ham)
    e_orc = oracle.lanczos_gs(csc, 150, oracle.seeded_state(len(basis), 41))["E0"]
    assert abs(served["results"][0]["energy"] - e_orc) <= 1e-10 * abs(e_orc), (served["results"][0]["energy"], e_orc)


def _wait(pred, limit_s=60.0):
    import time
    t0 = time.perf_counter()
    while not pred():
        assert time.perf_counter() - t0 < limit_s, "server did not reach the expected state"
        time.sleep(0.002)


@pytest.mark.parametrize("batch", [1, 4])
def test_cancel_queued_and_running(ctx, batch):
    """qd_server_cancel: a queued request is answered at once, a running one
    stops at its next time step with QD_E_CANCELLED, an unknown tag is
    refused, and the server keeps serving afterwards."""
    idx, A, D = Q.xxz_parts(12, 1.0, pbc=True)
    sec = Q.sz_sector(idx, 12, 6)
    srv = SV.Server(hbm_budget=4 << 30, workers=1, batch=batch)
    done = {}

    def collect():
        for x in srv.poll():
            done[x["tag"]] = x

    try:
        m = srv.add_model(sec, A, D)
        print("cancel test, batch", batch, flush=True)
        long_steps = 50000
        srv.submit(abi.QD_REQ_EVOLVE, m, 0.7, 16, steps=long_steps, dt=0.02, seed=3, tag=1)
        if batch > 1:  # a short request in the same batch finishes normally
            srv.submit(abi.QD_REQ_EVOLVE, m, 1.3, 16, steps=20, dt=0.02, seed=4, tag=5)
        _wait(lambda: srv.stats()["running"] >= 1)
        srv.submit(abi.QD_REQ_GROUND_STATE, m, 0.5, 60, chunk=10, tol=1e-12, seed=7, tag=2)  # queued behind
        if batch > 1:
            _wait(lambda: (collect(), 5 in done)[1])
            assert done[5]["status"] == 0 and done[5]["iterations"] == 20
        assert srv.cancel(2)
        collect()
        assert done[2]["status"] == abi.STATUS_CODE["QD_E_CANCELLED"] and done[2]["worker"] == -1
        assert srv.cancel(1)
        _wait(lambda: (collect(), 1 in done)[1])
        assert done[1]["status"] == abi.STATUS_CODE["QD_E_CANCELLED"]
        assert 0 < done[1]["iterations"] < long_steps
        assert not srv.cancel(1) and not srv.cancel(12345)
        srv.submit(abi.QD_REQ_GROUND_STATE, m, 0.9, 60, chunk=10, tol=1e-12, seed=7, tag=3)
        srv.drain()
        collect()
        st = srv.stats()
    finally:
        srv.close()
    e = _dense_e0(12, 0.9)
    assert done[3]["status"] == 0 and abs(done[3]["energy"] - e) <= 1e-10 * abs(e)
    assert st["running"] == 0 and st["queued"] == 0
