"""Python handles over the C ABI (include/qudrip_hip.h).

The product is the HIP library; these classes only own its handles and hand
it device pointers.  Device vectors are torch complex128 tensors on the
context's GPU; the context issues its work on torch's current stream, so
the two orders agree without extra synchronisation.
"""
import ctypes

import numpy as np

from . import abi


def _torch():
    import torch
    return torch


class Context:
    def __init__(self, device=0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("no GPU visible: the HIP path has no CPU fallback")
        self.lib = abi.load()
        self.device = device
        torch.cuda.set_device(device)
        self.torch_stream = torch.cuda.current_stream(device)
        p = ctypes.c_void_p()
        abi.check(None, self.lib.qd_ctx_create_on_stream(device, ctypes.c_void_p(self.torch_stream.cuda_stream),
                                                         ctypes.byref(p)), "qd_ctx_create_on_stream")
        self.ptr = p

    def check(self, rc, what=""):
        abi.check(self.ptr, rc, what)

    def sync(self):
        self.check(self.lib.qd_ctx_sync(self.ptr), "sync")

    def close(self):
        if self.ptr:
            self.lib.qd_ctx_destroy(self.ptr)
            self.ptr = None

    def zeros(self, n):
        return _torch().zeros(n, dtype=_torch().complex128, device=f"cuda:{self.device}")

    def seeded_state(self, n, seed):
        v = self.zeros(n)
        self.check(self.lib.qd_seeded_state(self.ptr, n, seed, ctypes.c_void_p(v.data_ptr())), "seeded_state")
        return v

    def norm2(self, v):
        out = ctypes.c_double()
        self.check(self.lib.qd_norm2(self.ptr, v.numel(), ctypes.c_void_p(v.data_ptr()), ctypes.byref(out)),
                   "norm2")
        return out.value


class Basis:
    def __init__(self, ctx, sector):
        self.ctx = ctx
        p = ctypes.c_void_p()
        if sector.mdim > 1:
            arr, n = sector.ctypes_quantities()
            ctx.check(ctx.lib.qd_basis_create_modes(ctx.ptr, sector.nbits, sector.mdim, arr, n, ctypes.byref(p)),
                      "basis_create_modes")
        elif sector.plain:
            arr, n = sector.ctypes_constraints()
            ctx.check(ctx.lib.qd_basis_create(ctx.ptr, sector.nbits, arr, n, ctypes.byref(p)), "basis_create")
        else:
            arr, n = sector.ctypes_quantities()
            ctx.check(ctx.lib.qd_basis_create_q(ctx.ptr, sector.nbits, arr, n, ctypes.byref(p)), "basis_create_q")
        self.ptr = p
        self.nbits = sector.nbits
        self.size = ctx.lib.qd_basis_size(p)

    def states(self):
        out = np.empty(self.size, dtype=np.uint64)
        if self.size:
            self.ctx.check(self.ctx.lib.qd_basis_download(self.ptr, out.ctypes.data), "basis_download")
        return out

    def rank(self, states_np):
        torch = _torch()
        st = torch.from_numpy(np.ascontiguousarray(states_np, dtype=np.uint64).view(np.int64)).to(
            f"cuda:{self.ctx.device}")
        out = torch.empty(st.numel(), dtype=torch.int64, device=st.device)
        self.ctx.check(self.ctx.lib.qd_basis_rank(self.ptr, ctypes.c_void_p(st.data_ptr()), st.numel(),
                                                  ctypes.c_void_p(out.data_ptr())), "basis_rank")
        return out.cpu().numpy()

    def close(self):
        if self.ptr:
            self.ctx.lib.qd_basis_destroy(self.ptr)
            self.ptr = None


class TermProgram:
    def __init__(self, ctx, opsum, nbits):
        self.ctx = ctx
        p = ctypes.c_void_p()
        if opsum.has_modes():
            terms, nt, ops, no, mats, nm, mdim = opsum.program_modes()
            self._keep = (terms, ops, mats)
            ctx.check(ctx.lib.qd_termprog_create_modes(ctx.ptr, nbits, mdim, terms, nt, ops, no,
                                                       mats.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), nm,
                                                       ctypes.byref(p)), "termprog_modes")
        else:
            terms, nt, ops, no = opsum.program()
            self._keep = (terms, ops)
            ctx.check(ctx.lib.qd_termprog_create(ctx.ptr, nbits, terms, nt, ops, no, ctypes.byref(p)), "termprog")
        self.ptr = p

    def close(self):
        if self.ptr:
            self.ctx.lib.qd_termprog_destroy(self.ptr)
            self.ptr = None


class CSR:
    """`ham >> index` on the GPU (reference ColMajor arrays)."""

    def __init__(self, ctx, basis, prog, _ptr=None):
        self.ctx = ctx
        p = ctypes.c_void_p()
        if _ptr is None:
            ctx.check(ctx.lib.qd_csr_build(ctx.ptr, basis.ptr, prog.ptr, ctypes.byref(p)), "csr_build")
        else:
            p = _ptr
        self.ptr = p
        n, nnz, herm = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        ctx.check(ctx.lib.qd_csr_info(p, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(herm)), "csr_info")
        self.n, self.nnz, self.hermitian = n.value, nnz.value, herm.value

    # SpMatrix algebra on the device (qd_csr_zero ... qd_csr_multiply)
    @classmethod
    def _made(cls, ctx, fn, *args):
        p = ctypes.c_void_p()
        ctx.check(fn(ctx.ptr, *args, ctypes.byref(p)), fn.__name__)
        return cls(ctx, None, None, _ptr=p)

    @classmethod
    def zero(cls, ctx, n):
        return cls._made(ctx, ctx.lib.qd_csr_zero, n)

    @classmethod
    def identity(cls, ctx, n):
        return cls._made(ctx, ctx.lib.qd_csr_identity, n)

    def scale(self, c):
        """c * A; a Python float is a real scalar (Eigen's double * complex)."""
        real = not isinstance(c, complex)
        c = complex(c)
        return CSR._made(self.ctx, self.ctx.lib.qd_csr_scale, self.ptr, c.real, c.imag, 1 if real else 0)

    def add(self, other, sign=1):
        return CSR._made(self.ctx, self.ctx.lib.qd_csr_add, self.ptr, other.ptr, sign)

    def adjoint(self):
        return CSR._made(self.ctx, self.ctx.lib.qd_csr_adjoint, self.ptr)

    def multiply(self, other):
        return CSR._made(self.ctx, self.ctx.lib.qd_csr_multiply, self.ptr, other.ptr)

    def download(self):
        outer = np.empty(self.n + 1, dtype=np.int32)
        inner = np.empty(max(self.nnz, 1), dtype=np.int32)
        vals = np.empty(2 * max(self.nnz, 1), dtype=np.float64)
        self.ctx.check(self.ctx.lib.qd_csr_download(self.ptr, outer.ctypes.data, inner.ctypes.data,
                                                    vals.ctypes.data), "csr_download")
        return outer, inner[:self.nnz], vals[:2 * self.nnz]

    def close(self):
        if self.ptr:
            self.ctx.lib.qd_csr_destroy(self.ptr)
            self.ptr = None


class LinOp:
    def __init__(self, ctx, ptr, keep=()):
        self.ctx = ctx
        self.ptr = ptr
        self._keep = keep
        self.n = ctx.lib.qd_linop_dim(ptr)

    @classmethod
    def from_csr(cls, ctx, csr):
        p = ctypes.c_void_p()
        ctx.check(ctx.lib.qd_linop_from_csr(ctx.ptr, csr.ptr, ctypes.byref(p)), "linop_from_csr")
        return cls(ctx, p, (csr,))

    @classmethod
    def matfree(cls, ctx, basis, prog):
        p = ctypes.c_void_p()
        ctx.check(ctx.lib.qd_linop_matfree(ctx.ptr, basis.ptr, prog.ptr, ctypes.byref(p)), "linop_matfree")
        return cls(ctx, p, (basis, prog))

    @classmethod
    def with_diagonal(cls, ctx, base, diag):
        """(base + diag(d)) as a view (qd_linop_with_diagonal); diag is a device complex128 vector."""
        p = ctypes.c_void_p()
        ctx.check(ctx.lib.qd_linop_with_diagonal(ctx.ptr, base.ptr, ctypes.c_void_p(diag.data_ptr()),
                                                 ctypes.byref(p)), "linop_with_diagonal")
        return cls(ctx, p, (base, diag))

    def set_lanes(self, lanes):
        self.ctx.check(self.ctx.lib.qd_linop_set_lanes(self.ptr, lanes), "set_lanes")

    def set_structured(self, on):
        has = ctypes.c_int()
        self.ctx.check(self.ctx.lib.qd_linop_set_structured(self.ptr, 1 if on else 0, ctypes.byref(has)),
                       "set_structured")
        return bool(has.value)

    def lanes(self):
        return self.ctx.lib.qd_linop_get_lanes(self.ptr)

    def autotune(self, x, y, reps=20):
        best = ctypes.c_int()
        us = ctypes.c_float()
        self.ctx.check(self.ctx.lib.qd_linop_autotune(self.ctx.ptr, self.ptr, ctypes.c_void_p(x.data_ptr()),
                                                      ctypes.c_void_p(y.data_ptr()), reps, ctypes.byref(best),
                                                      ctypes.byref(us)), "autotune")
        return best.value, us.value

    def apply(self, x, y=None):
        if y is None:
            y = self.ctx.zeros(self.n)
        self.ctx.check(self.ctx.lib.qd_spmv(self.ctx.ptr, self.ptr, ctypes.c_void_p(x.data_ptr()),
                                            ctypes.c_void_p(y.data_ptr())), "spmv")
        return y

    def expect(self, psi):
        out = (ctypes.c_double * 2)()
        self.ctx.check(self.ctx.lib.qd_expect(self.ctx.ptr, self.ptr, ctypes.c_void_p(psi.data_ptr()), out),
                       "expect")
        return complex(out[0], out[1])

    def close(self):
        if self.ptr:
            self.ctx.lib.qd_linop_destroy(self.ptr)
            self.ptr = None


class Krylov:
    """Arnoldi::getSolver (herm=0) / getLanczosSolver (herm=1)."""

    def __init__(self, ctx, n, kry_dim, herm):
        self.ctx = ctx
        p = ctypes.c_void_p()
        ctx.check(ctx.lib.qd_krylov_create(ctx.ptr, n, kry_dim, 1 if herm else 0, ctypes.byref(p)), "krylov")
        self.ptr = p
        self.n = n

    def ground_state(self, op, psi, add_dim=-1):
        e0 = ctypes.c_double()
        self.ctx.check(self.ctx.lib.qd_krylov_ground_state(self.ptr, op.ptr, ctypes.c_void_p(psi.data_ptr()),
                                                           add_dim, ctypes.byref(e0)), "ground_state")
        return e0.value

    def evolve(self, op, psi_in, psi_out, h, kry_dim_rt=0):
        self.ctx.check(self.ctx.lib.qd_krylov_evolve(self.ptr, op.ptr, ctypes.c_void_p(psi_in.data_ptr()),
                                                     ctypes.c_void_p(psi_out.data_ptr()), float(h), kry_dim_rt),
                       "evolve")

    def evolve_tol(self, op, psi_in, psi_out, h, m_min, kry_max, tol):
        """Returns (dimension used, smallest dimension whose estimate met tol)."""
        m, me = ctypes.c_int(), ctypes.c_int()
        self.ctx.check(self.ctx.lib.qd_krylov_evolve_tol(self.ptr, op.ptr, ctypes.c_void_p(psi_in.data_ptr()),
                                                         ctypes.c_void_p(psi_out.data_ptr()), float(h), m_min,
                                                         kry_max, float(tol), ctypes.byref(m), ctypes.byref(me)),
                       "evolve_tol")
        return m.value, me.value

    def info(self):
        inf = abi.qd_krylov_info()
        self.ctx.check(self.ctx.lib.qd_krylov_get_info(self.ptr, ctypes.byref(inf)), "info")
        return {"act_dim": inf.act_dim, "breakdown": inf.breakdown, "error": inf.error,
                "beta_last": inf.beta_last, "n_spmv": inf.n_spmv, "ritz_resid": inf.ritz_resid,
                "ritz_gap": inf.ritz_gap, "full_steps": inf.full_steps}

    def eigenvalues(self, cap=4096):
        buf = (ctypes.c_double * cap)()
        cnt = ctypes.c_int()
        self.ctx.check(self.ctx.lib.qd_krylov_eigenvalues(self.ptr, buf, cap, ctypes.byref(cnt)), "eigenvalues")
        return np.array(buf[:cnt.value])

    def set_method(self, method):
        self.ctx.check(self.ctx.lib.qd_krylov_set_method(self.ptr, method), "set_method")

    def reset(self):
        self.ctx.check(self.ctx.lib.qd_krylov_reset(self.ptr), "reset")

    def close(self):
        if self.ptr:
            self.ctx.lib.qd_krylov_destroy(self.ptr)
            self.ptr = None


class HermitianLanczos:
    """Hermitian::getSolver / Hermitian::LanczosSolver (sparse_algorithm.hpp:236-313):
    the two-vector Lanczos without reorthogonalisation, two passes per solve."""

    def __init__(self, ctx, n, kry_dim):
        self.ctx = ctx
        p = ctypes.c_void_p()
        ctx.check(ctx.lib.qd_hlanczos_create(ctx.ptr, n, kry_dim, ctypes.byref(p)), "hlanczos")
        self.ptr = p
        self.n = n
        self.kry_dim = kry_dim

    def ground_state(self, op, psi):
        e0 = ctypes.c_double()
        self.ctx.check(self.ctx.lib.qd_hlanczos_ground_state(self.ptr, op.ptr, ctypes.c_void_p(psi.data_ptr()),
                                                             ctypes.byref(e0)), "hlanczos ground_state")
        return e0.value

    def evolve(self, op, psi_in, psi_out, h):
        self.ctx.check(self.ctx.lib.qd_hlanczos_evolve(self.ptr, op.ptr, ctypes.c_void_p(psi_in.data_ptr()),
                                                       ctypes.c_void_p(psi_out.data_ptr()), float(h)),
                       "hlanczos evolve")

    def info(self):
        inf = abi.qd_hlanczos_info()
        self.ctx.check(self.ctx.lib.qd_hlanczos_get_info(self.ptr, ctypes.byref(inf)), "hlanczos info")
        return {"dim": inf.dim, "error": inf.error, "n_spmv": inf.n_spmv}

    def eigenvalues(self):
        cap = self.kry_dim + 2
        buf = (ctypes.c_double * cap)()
        cnt = ctypes.c_int()
        self.ctx.check(self.ctx.lib.qd_hlanczos_eigenvalues(self.ptr, buf, cap, ctypes.byref(cnt)),
                       "hlanczos eigenvalues")
        return np.array(buf[:cnt.value])

    def coefficients(self):
        """(alpha complex[dim], beta[dim]); beta[dim-1] is error()."""
        cap = self.kry_dim + 2
        al = np.zeros(cap, dtype=np.complex128)
        be = np.zeros(cap)
        cnt = ctypes.c_int()
        self.ctx.check(self.ctx.lib.qd_hlanczos_coefficients(
            self.ptr, al.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            be.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cap, ctypes.byref(cnt)), "hlanczos coefficients")
        return al[:cnt.value], be[:cnt.value]

    def close(self):
        if self.ptr:
            self.ctx.lib.qd_hlanczos_destroy(self.ptr)
            self.ptr = None


def jump_channels(nsite, dephase_rate, loss_rate):
    """Dephasing sqrt(g) n_m and loss sqrt(g) c_m on every mode m (0 rates omitted)."""
    ch = []
    for m in range(2 * nsite):
        if dephase_rate > 0:
            ch.append((abi.QD_JUMP_DEPHASE, m, dephase_rate))
        if loss_rate > 0:
            ch.append((abi.QD_JUMP_LOSS, m, loss_rate))
    return ch


def traj_run(ctx, opsum, nsite, nup0, ndn0, psi0, channels, ntraj, nsteps, dt, k_rt, seed, want_log=False):
    """Batched quantum-jump trajectories (qd_traj_run); returns mean/var densities [(nsteps+1), 2 nsite]."""
    terms, nt, ops, no = opsum.program()
    jarr = (abi.qd_jump * max(1, len(channels)))()
    for i, (k, m, g) in enumerate(channels):
        jarr[i].kind, jarr[i].mode, jarr[i].rate = k, m, g
    M = 2 * nsite
    mean = np.zeros((nsteps + 1) * M)
    var = np.zeros((nsteps + 1) * M)
    log = np.zeros(ntraj * max(nsteps, 1), dtype=np.int32) if want_log else None
    st = abi.qd_traj_stats()
    ctx.check(ctx.lib.qd_traj_run(ctx.ptr, nsite, terms, nt, ops, no, nup0, ndn0, ctypes.c_void_p(psi0.data_ptr()),
                                  jarr, len(channels), ntraj, nsteps, float(dt), k_rt, seed, mean.ctypes.data,
                                  var.ctypes.data, log.ctypes.data if log is not None else None,
                                  ctypes.byref(st)), "traj_run")
    out = {"mean": mean.reshape(nsteps + 1, M), "var": var.reshape(nsteps + 1, M),
           "jumps": st.jumps, "losses": st.losses, "sectors": st.sectors, "lanczos": st.lanczos,
           "lanczos_fallback_steps": st.lanczos_fallback_steps}
    if log is not None:
        out["log"] = log.reshape(ntraj, max(nsteps, 1))
    return out
