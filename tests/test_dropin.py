"""The C++ drop-in API (include/qudrip/): compiles with plain g++, and a
reference-style driver runs on the GPU with the same results as the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from qudrip_amd import abi
from qudrip_amd import operators as Q
from tests import physics as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "qudrip_amd", "programs", "bin")


def test_headers_compile_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "qudrip/qudrip.hpp"\nint main(){ auto q = qudrip::getQbits(4); (void)q; }\n')
    subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"), str(src)],
                   check=True)


def test_headers_compile_with_nd_nh(tmp_path):
    """getNd / getNh (conservation.hpp:248-256) next to getNup / getNdo in Constrain."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "qudrip/qudrip.hpp"\nusing namespace qudrip;\n'
                   'void f(){ auto q = getQbits(8); auto nup = getNup(q, 4); auto ndo = getNdo(q, 4);\n'
                   ' auto nd = getNd(q, 4); auto nh = getNh(q, 4);\n'
                   ' auto s = Constrain(q, nup = 2, ndo = 2, nd = 1, nh = 1); (void)s; }\nint main(){}\n')
    subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"), str(src)],
                   check=True)


def test_headers_compile_solvers_over_a_mode_index(tmp_path):
    """Both solver families, McManager and the lazy product instantiate over a
    qubits-times-mode SubIndex (index.hpp:94-184, operators/mode.hpp)."""
    src = tmp_path / "t.cpp"
    src.write_text('''#include "qudrip/qudrip.hpp"
using namespace qudrip;
int main() {
  auto q = getQbits(4);
  SingleModeIndex<> md(5);
  auto idx = q * md;
  auto nset = getNset(q);
  auto sub = Constrain(idx, nset = 2);
  auto gc = getBoseGate(q), ga = getBoseGate(q);
  gc << 0.5 * (pauli_x + II * pauli_y);
  ga << 0.5 * (pauli_x - II * pauli_y);
  auto a = getModeOp(md), ad = getModeOp(md);
  a << LadderMatrix(5, Mode::Upper);
  ad << LadderMatrix(5, Mode::Lower);
  auto ham = getEmptyOperator();
  for (int i = 0; i < 3; ++i) ham = ham + gc(i + 1) * ga(i) + gc(i) * ga(i + 1);
  ham = ham + 0.3 * (a(0) * ad(0));
  auto H = ham >> sub;
  auto psi = getState(sub, 3);
  shuffle_state(psi(0));
  auto hs = Hermitian::getSolver(sub, 20);
  hs.evolve(H, psi, 0, 0.1);
  hs.getGroundState(H, psi, 0);
  auto ar = Arnoldi::getLanczosSolver(sub, 20);
  ar.getGroundState(H, psi, 1);
  auto mc = getMC(psi);
  mc.addObs("E", 1);
  mc.updateMatrix("E", 0, H);
  mc.incrAllObs(0);
  psi(2) = ham * psi(0);
  return (int)psi.eval(0, H).real();
}
''')
    subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"), str(src)],
                   check=True)


def test_driver_programs_built():
    for p in ("spin", "dropin_dump", "mc_dump", "hubbard_quench", "serve_stress"):
        assert os.access(os.path.join(BIN, p), os.X_OK), p


def _run(args, **kw):
    r = subprocess.run(args, capture_output=True, text=True, timeout=120, **kw)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    return r.stdout


@pytest.mark.gpu
def test_serve_stress_driver_six_workers():
    """The server from C++ alone (programs/serve_stress.cpp): a mixed stream
    submitted at once to six workers, twice; every request succeeds and
    identical requests agree to 1e-12 whatever ran beside them."""
    out = _run([os.path.join(BIN, "serve_stress"), "6", "48", "2", "20"])
    m = re.search(r"serve_stress: (\d+) results, (\d+) distinct requests, (\d+) failures, max running (\d+)", out)
    assert m and m.group(1) == "96" and m.group(3) == "0", out
    assert int(m.group(2)) > 1 and 1 <= int(m.group(4)) <= 6


@pytest.mark.gpu
def test_spin_driver_xy16():
    out = _run([os.path.join(BIN, "spin"), "16", "100"])
    m = re.search(r"E = \(([-0-9.e+]+),([-0-9.e+]+)\)", out)
    nrm = float(re.search(r"norm = ([-0-9.e+]+)", out).group(1))
    e_ref = P.xy_free_fermion_e0(16, 8)
    assert abs(float(m.group(1)) / nrm - e_ref) / abs(e_ref) < 1e-10
    assert abs(float(m.group(2))) < 1e-10
    assert "dimension of Hil space = 12870" in out and "nnz = 102960" in out


REF_SPIN = "/root/reference/programs/spin.cpp"


@pytest.mark.skipif(not os.path.exists(REF_SPIN), reason="the reference tree is not on this machine")
def test_reference_spin_cpp_compiles_unchanged(tmp_path):
    """The reference's own driver (programs/spin.cpp), byte for byte, builds
    against the drop-in headers and links against the HIP library: its
    `#include "qudrip.hpp"` resolves to include/qudrip/qudrip.hpp, and every
    call it makes (getQbits ... get_clock, Arnoldi::getLanczosSolver,
    getGroundState, psi.eval) exists with the reference's signature."""
    exe = tmp_path / "ref_spin"
    r = subprocess.run(["g++", "-std=c++20", "-O1", "-Wall", "-I", os.path.join(ROOT, "include", "qudrip"),
                        "-I", os.path.join(ROOT, "include"), "-o", str(exe), REF_SPIN,
                        "-L", os.path.join(ROOT, "qudrip_amd"), "-lqudrip_hip",
                        "-Wl,-rpath," + os.path.join(ROOT, "qudrip_amd"), "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.access(exe, os.X_OK)


@pytest.mark.gpu
def test_reference_spin_binary_runs_unchanged():
    """bin/ref_spin is the reference's programs/spin.cpp compiled unchanged
    against the drop-in headers (qudrip_amd/programs/Makefile, built where the
    reference tree exists).  It runs the reference's call sequence on the GPU:
    XY chain N=16, Nset=8, Lanczos krylov_dim=100; E printed as the
    reference prints it (std::complex, default precision)."""
    exe = os.path.join(BIN, "ref_spin")
    if not os.access(exe, os.X_OK):
        pytest.skip("bin/ref_spin was not built (no reference tree where the library was built)")
    out = _run([exe, "16"])
    assert "dimension of Hil space = 12870" in out
    assert re.search(r"Time elapsed: [0-9.e+-]+", out)
    m = re.search(r"E = \(([-0-9.e+]+),([-0-9.e+]+)\)", out)
    assert m, out
    # six significant digits (the reference's cout default); the state is
    # the Ritz vector of a normalised start, so <psi|H|psi> = E0 |psi|^2 = E0
    assert abs(float(m.group(1)) - P.xy_free_fermion_e0(16, 8)) < 5e-5
    assert abs(float(m.group(2))) < 1e-6


def _load(prefix, ext, dtype):
    return np.fromfile(prefix + ext, dtype=dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("model,N,a,b", [("xy", 12, 1.0, 1.0), ("xxz", 12, 0.9317, 0.3711), ("hubchain", 6, 0.77, 5.123),
                                         ("hubnd", 8, 0.77, 5.123)])
def test_dropin_matches_python_mirror_and_oracle(tmp_path, model, N, a, b):
    prefix = str(tmp_path / model)
    _run([os.path.join(BIN, "dropin_dump"), model, str(N), prefix, repr(a), repr(b)])
    if model == "xy":
        idx, ham = Q.xy_chain(N)
        sec = Q.Constrain(idx, Q.getNset(idx) == N // 2)
    elif model == "xxz":
        idx, ham = Q.xxz_chain(N, a, b, pbc=True)
        sec = Q.Constrain(idx, Q.getNset(idx) == N // 2)
    else:
        idx, ham = Q.hubbard(N, Q.hubbard_bonds_chain(N, False), t=a, U=b)
        sec = Q.hubbard_sector(idx, N, N // 2, N // 2)
        if model == "hubnd":
            sec = Q.Constrain(idx, *sec.constraints, Q.getNd(idx, N) == N // 4)
    terms, nt, ops, no = ham.program()
    # identical term programs (C++ algebra == Python mirror, term order and coefficients)
    assert _load(prefix, ".terms", np.uint8).tobytes() == bytes(terms)[: nt * 24]
    assert _load(prefix, ".ops", np.uint8).tobytes() == bytes(ops)[: no * 80]
    basis = oracle.basis_scan(sec)
    ref = oracle.build_csc(basis, ham)
    assert np.array_equal(_load(prefix, ".outer", np.int32), ref[0])
    assert np.array_equal(_load(prefix, ".inner", np.int32), ref[1])
    assert np.array_equal(_load(prefix, ".vals", np.uint64), ref[2].view(np.uint64))
    x = _load(prefix, ".x", np.complex128)
    assert np.max(np.abs(x - oracle.seeded_state(len(basis), 7))) < 1e-15
    y = _load(prefix, ".y", np.complex128)
    yr = oracle.spmv(ref, x)
    assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-14


@pytest.mark.gpu
def test_dropin_boson_mode_jaynes_cummings(tmp_path):
    """SingleModeIndex / Indices / getModeOp / LadderMatrix from C++: the
    program (terms, ops, mode matrices), the `>>` arrays and H x against the
    Python mirror and the oracle."""
    prefix = str(tmp_path / "jc")
    nd, g, D = 7, 0.37, 0.83
    _run([os.path.join(BIN, "dropin_dump"), "jc", str(nd), prefix, repr(g), repr(D)])
    q = Q.getQbits(1)
    md = Q.SingleModeIndex(nd)
    idx = q * md
    sp, sm, ne = Q.getBoseGate(q), Q.getBoseGate(q), Q.getBoseGate(q)
    sp << [[0, 1], [0, 0]]
    sm << [[0, 0], [1, 0]]
    ne << [[0, 0], [0, 1]]
    ad, a = Q.getModeOp(md), Q.getModeOp(md)
    ad << Q.LadderMatrix(nd, Q.MODE_LOWER)
    a << Q.LadderMatrix(nd, Q.MODE_UPPER)
    ham = Q.getEmptyOperator() + ad(0) * a(0)
    ham = ham + D * ne(0)
    ham = ham + g * (sp(0) * a(0)) + g * (sm(0) * ad(0))
    terms, nt, ops, no, mats, nm, mdim = ham.program_modes()
    assert _load(prefix, ".terms", np.uint8).tobytes() == bytes(terms)[: nt * 24]
    assert _load(prefix, ".ops", np.uint8).tobytes() == bytes(ops)[: no * 80]
    assert np.array_equal(_load(prefix, ".mats", np.float64), mats)
    basis = oracle.basis_scan(Q.Constrain(idx))
    ref = oracle.build_csc(basis, ham)
    assert np.array_equal(_load(prefix, ".outer", np.int32), ref[0])
    assert np.array_equal(_load(prefix, ".inner", np.int32), ref[1])
    assert np.array_equal(_load(prefix, ".vals", np.uint64), ref[2].view(np.uint64))
    x = _load(prefix, ".x", np.complex128)
    y = _load(prefix, ".y", np.complex128)
    yr = oracle.spmv(ref, x)
    assert np.linalg.norm(y - yr) / np.linalg.norm(yr) < 1e-14
    yl = _load(prefix, ".ylazy", np.complex128)  # psi(1) = ham * psi(0)
    assert np.linalg.norm(yl - yr) / np.linalg.norm(yr) < 1e-14


@pytest.mark.gpu
def test_dropin_lazy_product_matches_map_acc(tmp_path):
    """psi2 = ham * psi1 (the lazy path, operator.hpp:57-66 / operator_abel.hpp:36-49)."""
    prefix = str(tmp_path / "lazy")
    N = 6
    _run([os.path.join(BIN, "dropin_dump"), "lazy", str(N), prefix, "1.0", "4.0"])
    idx, ham = Q.hubbard(N, Q.hubbard_bonds_chain(N, False), t=1.0, U=4.0)
    basis = oracle.basis_scan(Q.hubbard_sector(idx, N, N // 2, N // 2))
    x = _load(prefix, ".x", np.complex128)
    to = np.zeros(len(basis), dtype=np.complex128)
    oracle.map_acc(basis, ham, x, to)
    y = _load(prefix, ".y", np.complex128)
    assert np.linalg.norm(y - to) / np.linalg.norm(to) < 1e-14


@pytest.mark.gpu
def test_dropin_mc_manager_out_file(tmp_path):
    """McManager::outFile (mc_manager.hpp:147-196): '#time' header with
    'name(real)\tname(imag)' for complex and 'name' for real observables,
    '#time\tindex' and a blank line after each time for vector observables,
    std::to_string numbers; values = the oracle's psi.eval along the same
    Arnoldi steps (averaged over two identical samples)."""
    N, nt, h = 8, 5, 0.1
    prefix = str(tmp_path / "mc")
    _run([os.path.join(BIN, "mc_dump"), str(N), str(nt), repr(h), prefix])
    idx, ham = Q.xy_chain(N)
    basis = oracle.basis_scan(Q.Constrain(idx, Q.getNset(idx) == N // 2))
    csc = oracle.build_csc(basis, ham)
    psi = oracle.seeded_state(len(basis), 5)
    states = [psi]
    for _ in range(nt - 1):
        states.append(oracle.arnoldi_evolve(csc, 12, h, states[-1]))
    occ = np.array([[((int(b) >> i) & 1) for i in range(N)] for b in basis], dtype=float)
    E = open(prefix + ".E.txt").read().split("\n")
    assert E[0] == "#time\tE(real)\tE(imag)"
    for t in range(nt):
        f = E[1 + t].split("\t")
        assert f[0] == f"{t * h:f}" and len(f) == 3
        e = oracle.eval_expect(csc, states[t])
        assert abs(float(f[1]) - e.real) < 2e-6 and abs(float(f[2]) - e.imag) < 2e-6
        assert re.fullmatch(r"-?\d+\.\d{6}", f[1])
    lines = open(prefix + ".n.txt").read().split("\n")
    assert lines[0] == "#time\tindex\tn"
    pos = 1
    for t in range(nt):
        dens = (np.abs(states[t]) ** 2) @ occ
        for i in range(N):
            f = lines[pos].split("\t")
            assert f[0] == f"{t * h:f}" and f[1] == str(i) and len(f) == 3
            assert abs(float(f[2]) - dens[i]) < 2e-6
            pos += 1
        assert lines[pos] == ""  # blank line after each time of a vector observable
        pos += 1

