"""GPU workloads against the oracle / closed forms: Hubbard quench (config 3 shape)."""
import numpy as np
import pytest
import scipy.linalg

import oracle
from qudrip_amd import operators as Q
from qudrip_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("matfree,rt_method", [(True, 1), (False, 1), (True, 2)])
def test_hubbard_quench_matches_dense(ctx, matfree, rt_method):
    """3x2 periodic Hubbard, 2 up / 2 down: GS by Lanczos, quench U 4 -> 2, Arnoldi dt=0.05
    (real-time basis fully reorthogonalised, or by the partial scheme)."""
    Lx, Ly, nup, ndn = 3, 2, 2, 2
    res = WL.hubbard_quench(ctx, Lx, Ly, nup, ndn, U0=4.0, U1=2.0, k_gs=60, k_rt=20, dt=0.05, steps=20,
                            matfree=matfree, two_vector_gs=True, rt_method=rt_method)
    nsite = Lx * Ly
    bonds = Q.hubbard_bonds_square(Lx, Ly)
    idx, h0 = Q.hubbard(nsite, bonds, 1.0, 4.0)
    _, h1 = Q.hubbard(nsite, bonds, 1.0, 2.0)
    sec = Q.hubbard_sector(idx, nsite, nup, ndn)
    basis = oracle.basis_scan(sec)
    H0 = oracle.csc_to_dense(oracle.build_csc(basis, h0))
    H1 = oracle.csc_to_dense(oracle.build_csc(basis, h1))
    D = oracle.csc_to_dense(oracle.build_csc(basis, WL.docc_operator(nsite)))
    w, z = np.linalg.eigh(H0)
    assert abs(res["E0"] - w[0]) / abs(w[0]) < 1e-10
    tv = res["two_vector_gs"]  # Hermitian::getSolver on the same operator
    assert abs(tv["E0"] - w[0]) / abs(w[0]) < 1e-10 and abs(tv["rayleigh"] - w[0]) / abs(w[0]) < 1e-9
    g = z[:, 0]
    U = scipy.linalg.expm(-1j * 0.05 * H1)
    docc = [np.vdot(g, D @ g).real]
    for _ in range(20):
        g = U @ g
        docc.append(np.vdot(g, D @ g).real)
    assert np.max(np.abs(np.array(res["docc"]) - np.array(docc))) < 1e-9
    assert abs(res["energy_end"] - res["energy_start"]) < 1e-9
    assert abs(res["norm_end"] - 1.0) < 1e-9


def test_quench_partial_reorth_matches_full(ctx):
    """3x4 periodic Hubbard (3,3), n = 48,400, k_rt = 30: the real-time basis
    kept semi-orthogonal by partial reorthogonalisation gives the same double
    occupancy at every step and the same final state as the fully
    reorthogonalised basis, to rounding."""
    import torch
    kw = dict(U0=4.0, U1=2.0, k_gs=100, k_rt=30, dt=0.05, steps=10, matfree=True)
    full = WL.hubbard_quench(ctx, 3, 4, 3, 3, **kw)
    part = WL.hubbard_quench(ctx, 3, 4, 3, 3, rt_method=2, **kw)
    assert part["n"] == 48400 and part["rt_method"] == 2
    assert np.max(np.abs(np.array(part["docc"]) - np.array(full["docc"]))) < 1e-12
    d = float(torch.linalg.vector_norm(part["psi"] - full["psi"]) / torch.linalg.vector_norm(full["psi"]))
    assert d < 1e-11
    assert abs(part["energy_end"] - part["energy_start"]) < 1e-10


def test_hubbard_4x4_free_fermions_closed_form(ctx):
    """U = 0, 5 up / 5 down on the 4x4 torus: E0 = -24 (closed shell), n = 19,079,424, matrix-free."""
    res = WL.hubbard_quench(ctx, 4, 4, 5, 5, U0=0.0, U1=0.0, k_gs=120, k_rt=4, dt=0.01, steps=1, matfree=True,
                            two_vector_gs=True)
    assert res["n"] == 19079424
    assert abs(res["E0"] - (-24.0)) / 24.0 < 1e-10
    # the two-vector solver (Hermitian::getSolver) through the two-block product
    assert abs(res["two_vector_gs"]["E0"] - (-24.0)) / 24.0 < 1e-10
    # free fermions: <n_up n_dn> factorises, docc = sum_i <n_i,up><n_i,dn> = 16 * (5/16)^2
    assert abs(res["docc"][0] - 16 * (5 / 16) ** 2) < 1e-8
