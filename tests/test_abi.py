"""CPU checks of the C-ABI library: it loads, and exports every function that
include/qudrip_hip.h declares (no compute calls -- no GPU here)."""
import ctypes
import os
import re

from qudrip_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**\s*(qd_[a-z0-9_]+)\s*\(",
                             src, flags=re.M):
            names.append(m.group(1))
    return names


def test_header_declares_functions():
    names = declared_functions()
    assert "qd_csr_build" in names and "qd_krylov_ground_state" in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_functions()) <= set(abi.SIGNATURES), set(declared_functions()) - set(abi.SIGNATURES)
    lib = abi.load()
    assert lib.qd_abi_version() >= 1


def test_host_numerics_without_gpu():
    """The FP64 host numerics run on the CPU side of the library."""
    import numpy as np
    import scipy.linalg
    lib = abi.load()
    rng = np.random.default_rng(0)
    m = 40
    d, e = rng.normal(size=m), rng.uniform(0.1, 1.0, size=m - 1)
    w = np.zeros(m)
    z = np.zeros(m * m)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.qd_host_tridiag_eig(m, dp(d), dp(e), dp(w), dp(z)) == 0
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    wr, zr = np.linalg.eigh(T)
    assert np.allclose(w, wr, atol=1e-12)
    Z = z.reshape(m, m).T  # column-major
    assert np.allclose(np.abs(np.sum(Z * zr, axis=0)), 1.0, atol=1e-10)
    A = (rng.normal(size=(29, 29)) + 1j * rng.normal(size=(29, 29))) * 0.6
    Af = np.asfortranarray(A)
    E = np.zeros((29, 29), dtype=np.complex128, order="F")
    assert lib.qd_host_expm(29, dp(Af), dp(E)) == 0
    assert np.allclose(E, scipy.linalg.expm(A), rtol=1e-12, atol=1e-12)
    # NaN input is an error status, not a crash (the reference aborts)
    d[3] = np.nan
    assert lib.qd_host_tridiag_eig(m, dp(d), dp(e), dp(w), dp(z)) == 5  # QD_E_EIGEN
