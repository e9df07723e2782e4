"""Python mirror of the reference's operator interface, producing term programs.

Same names and argument meaning as the reference headers:
  getQbits / getNset / getNup / getNdo / Constrain   (index.hpp, conservation.hpp)
  getBoseGate / getFermiGate / getHopFermiGate / getHopBoseGate, ``gate << U``,
  ``gate(site)``, ``gate(site2, site1)``                (operators/bit.hpp, hopbit.hpp)
  Operator products, OperatorSum algebra, getEmptyOperator (operator.hpp,
  operator_abel.hpp)
The reference's ``cq = value`` assignment becomes ``cq == value``.

``OperatorSum.program()`` flattens the sum into the qd_term / qd_op arrays of
include/qudrip_hip.h, in the reference's term order.  No device work happens
here: this is host logic that both the Python tests/bench and the C++
drop-in headers (include/qudrip/) implement identically.
"""
import ctypes

import numpy as np

from . import abi

II = 1j
pauli_x = np.array([[0, 1], [1, 0]], dtype=np.complex128)       # bit.hpp:42
pauli_y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)    # bit.hpp:43
pauli_z = np.array([[1, 0], [0, -1]], dtype=np.complex128)      # bit.hpp:44
id2 = np.array([[1, 0], [0, 1]], dtype=np.complex128)           # bit.hpp:45


class QbitsIndex:
    """QbitsIndex (index.hpp:80-95): range 1 << N."""

    def __init__(self, nbits):
        self.nbits = int(nbits)

    def range(self):
        return 1 << self.nbits


    def __mul__(self, other):  # operator* of two indices (index.hpp:179-184)
        if isinstance(other, SingleModeIndex):
            return Indices(self, other)
        return NotImplemented


def getQbits(N):
    return QbitsIndex(N)


class SingleModeIndex:
    """SingleModeIndex (index.hpp:94-110): one boson mode, levels 0..n_mode-1."""

    def __init__(self, n_mode):
        self.mdim = int(n_mode)

    def range(self):
        return self.mdim


class Indices:
    """Indices (index.hpp:112-177) of a QbitsIndex and a SingleModeIndex:
    value = q * mdim + m, the mode the least significant digit."""

    def __init__(self, lhs, rhs):
        self.lhs = lhs
        self.rhs = rhs
        self.nbits = lhs.nbits
        self.mdim = rhs.mdim

    def range(self):
        return (1 << self.nbits) * self.mdim


class Constraint:
    """cq = count for popcount(((x >> shift) ^ flip) & (x ^ flip) & mask): a
    plain Nset has shift 0 and flip 0 (the qd_quantity form)."""

    def __init__(self, mask, count, shift=0, flip=0):
        self.mask = int(mask)
        self.count = int(count)
        self.shift = int(shift)
        self.flip = int(flip)

    @property
    def plain(self):
        return self.shift == 0 and self.flip == 0


class Nset:
    """popcount(x & mask) over bits start..end inclusive (conservation.hpp:89-115)."""

    def __init__(self, index, start=-1, end=-1):
        if start == -1:
            start = 0
        if end == -1:
            end = bin(index.range() - 1).count("1")  # N: masks N+1 bits (:112)
        self.index = index
        self.mask = 0
        for i in range(start, end + 1):
            if i < 64:
                self.mask |= 1 << i

    def __eq__(self, value):  # reference: `cq = value`
        return Constraint(self.mask, value)


class Nd:
    """Doubly occupied (hole=False) or empty (hole=True) sites of a spinful
    index, up spins on bits 0..Nsite-1 (conservation.hpp:117-151):
    popcount(((x >> Nsite) ^ m) & (x ^ m)), m = bits 0..Nsite-1 if hole."""

    ALL = (1 << 64) - 1

    def __init__(self, index, Nsite, hole=False):
        self.index = index
        self.Nsite = int(Nsite)
        self.flip = ((1 << self.Nsite) - 1) if hole else 0

    def __eq__(self, value):
        return Constraint(self.ALL, value, shift=self.Nsite, flip=self.flip)


def getNset(index):
    return Nset(index, -1, -1)


def getNd(index, Nsite):
    return Nd(index, Nsite, False)


def getNh(index, Nsite):
    return Nd(index, Nsite, True)


def getNup(index, Nsite):
    return Nset(index, 0, Nsite - 1)


def getNdo(index, Nsite):
    return Nset(index, Nsite, 2 * Nsite - 1)


class Sector:
    """Result of Constrain: the QbitsIndex width plus the constraints."""

    def __init__(self, index, constraints):
        self.index = index
        self.nbits = index.nbits
        self.mdim = getattr(index, "mdim", 1)
        self.constraints = list(constraints)

    @property
    def plain(self):
        return all(c.plain for c in self.constraints)

    def ctypes_constraints(self):
        arr = (abi.qd_constraint * max(1, len(self.constraints)))()
        for i, c in enumerate(self.constraints):
            assert c.plain, "Nd/Nh constraints: use ctypes_quantities"
            arr[i].mask = c.mask
            arr[i].count = c.count
        return arr, len(self.constraints)

    def ctypes_quantities(self):
        arr = (abi.qd_quantity * max(1, len(self.constraints)))()
        for i, c in enumerate(self.constraints):
            arr[i].sel = c.mask
            arr[i].flip = c.flip
            arr[i].shift = c.shift
            arr[i].count = c.count
        return arr, len(self.constraints)


def Constrain(index, *constraints):
    """Constrain(index, cq == v, ...) (conservation.hpp:227-231)."""
    return Sector(index, constraints)


class _OpRecord:
    """One gate instance, i.e. one shared_ptr<elOpBase> (operator.hpp:46)."""

    __slots__ = ("kind", "site_a", "site_b", "string", "U")

    def __init__(self, kind, site_a, site_b, string, U):
        self.kind = kind
        self.site_a = site_a
        self.site_b = site_b
        self.string = string
        self.U = U


class Gate:
    """bitOp / hopBitOp prototype: holds U and the JW string (+1 / -1)."""

    def __init__(self, index, kind, string):
        self.index = index
        self.kind = kind
        self.string = string
        self.U = np.zeros((2, 2), dtype=np.complex128)

    def __lshift__(self, U):  # elOpBase::operator<< (operator.hpp:24-29)
        self.U = np.array(U, dtype=np.complex128).reshape(2, 2).copy()
        return self

    def __call__(self, site, site1=None):
        if self.kind == abi.QD_OP_BIT:  # bit.hpp:17-21
            if site1 is not None:
                raise TypeError("bit gate takes one site")
            return Operator([_OpRecord(self.kind, int(site), int(site), self.string, self.U.copy())])
        if site1 is None:  # hopbit.hpp:17-22: gate(site2, site1) = c+_{site2} c_{site1}
            raise TypeError("hop gate takes (site2, site1)")
        return Operator([_OpRecord(self.kind, int(site), int(site1), self.string, self.U.copy())])


class ModeGate:
    """modeOp prototype (operators/mode.hpp:6-35): `gate << U` stores an
    mdim x mdim matrix, gate(site) makes a fresh instance (the site is not
    used); the instance takes mode m to sum_j U(m, j) |j> (mode.hpp:80-103)."""

    def __init__(self, index):
        self.index = index
        self.mdim = index.mdim
        self.U = np.zeros((self.mdim, self.mdim), dtype=np.complex128)

    def __lshift__(self, U):
        self.U = np.array(U, dtype=np.complex128).reshape(self.mdim, self.mdim).copy()
        return self

    def __call__(self, site=0):
        return Operator([_OpRecord(abi.QD_OP_MODE, 0, 0, 1, self.U.copy())])


def getModeOp(index):
    return ModeGate(index)


MODE_UPPER, MODE_LOWER = 1, -1  # enum class Mode (mode.hpp:36)


def LadderMatrix(ndim, mode=MODE_UPPER):
    """mode.hpp:38-52: Upper puts sqrt(i+1) at (i+1, i), Lower at (i, i+1)."""
    m = np.zeros((ndim, ndim), dtype=np.complex128)
    for i in range(ndim - 1):
        if mode == MODE_UPPER:
            m[i + 1, i] = np.sqrt(i + 1)
        else:
            m[i, i + 1] = np.sqrt(i + 1)
    return m


def CoherentProjector(ndim, amp):
    """mode.hpp:55-67: |a><a| of the truncated coherent state."""
    a = np.zeros(ndim, dtype=np.complex128)
    a[0] = 1.0
    for i in range(1, ndim):
        a[i] = a[i - 1] * amp / np.sqrt(float(i))
    a *= np.exp(-abs(amp) * abs(amp) / 2.0)
    return np.outer(a, a.conj())


def FockProjector(ndim, idx):
    """mode.hpp:70-77: |idx><idx|."""
    m = np.zeros((ndim, ndim), dtype=np.complex128)
    m[idx, idx] = 1.0
    return m


def getBoseGate(index):
    return Gate(index, abi.QD_OP_BIT, 1)


def getFermiGate(index):
    return Gate(index, abi.QD_OP_BIT, -1)


def getHopFermiGate(index):
    return Gate(index, abi.QD_OP_HOP, -1)


def getHopBoseGate(index):
    return Gate(index, abi.QD_OP_HOP, 1)


def _is_scalar(x):
    return isinstance(x, (int, float, complex, np.number))


class Operator:
    """Ordered product of gate instances (operator.hpp:68-132)."""

    def __init__(self, ops):
        self.ops = list(ops)

    def same(self, other):  # operator== (:123-131): shared_ptr identity
        return len(self.ops) == len(other.ops) and all(a is b for a, b in zip(self.ops, other.ops))

    def __mul__(self, other):
        if isinstance(other, Operator):  # :82-91
            return Operator(self.ops + other.ops)
        if isinstance(other, OperatorSum):
            return OperatorSum([(self, 1.0 + 0j)]) * other
        if _is_scalar(other):  # operator_abel.hpp:147-149
            return OperatorSum([(self, complex(other))])
        return NotImplemented

    def __rmul__(self, other):
        if _is_scalar(other):  # operator_abel.hpp:143-145
            return OperatorSum([(self, complex(other))])
        return NotImplemented

    def __add__(self, other):
        if isinstance(other, Operator):  # :159-163
            return OperatorSum([(self, 1.0 + 0j)]).append(other, 1.0 + 0j)
        if isinstance(other, OperatorSum):
            return OperatorSum([(self, 1.0 + 0j)]) + other
        return NotImplemented

    def __sub__(self, other):
        if isinstance(other, Operator):  # :165-169
            return OperatorSum([(self, 1.0 + 0j)]).append(other, -1.0 + 0j)
        if isinstance(other, OperatorSum):
            return OperatorSum([(self, 1.0 + 0j)]) - other
        return NotImplemented

    def __rshift__(self, sector):  # `op >> index`
        return OperatorSum([(self, 1.0 + 0j)]), sector


class OperatorSum:
    """Terms with complex coefficients, in construction order (operator_abel.hpp:9-128)."""

    def __init__(self, terms=None):
        self.terms = [[op, complex(c)] for op, c in (terms or [])]

    def size(self):
        return len(self.terms)

    def copy(self):
        return OperatorSum([(op, c) for op, c in self.terms])

    def append(self, op, coeff):  # :115-127, merges only identical gate instances
        for t in self.terms:
            if t[0].same(op):
                t[1] = t[1] + complex(coeff)
                return self
        self.terms.append([op, complex(coeff)])
        return self

    def __add__(self, other):
        if isinstance(other, Operator):  # :83-87
            return self.copy().append(other, 1.0 + 0j)
        if isinstance(other, OperatorSum):  # :89-96, concatenation
            return OperatorSum([(op, c) for op, c in self.terms] + [(op, c) for op, c in other.terms])
        return NotImplemented

    def __sub__(self, other):
        if isinstance(other, Operator):  # :109-113
            return self.copy().append(other, -1.0 + 0j)
        if isinstance(other, OperatorSum):  # :98-107
            return OperatorSum([(op, c) for op, c in self.terms] + [(op, -c) for op, c in other.terms])
        return NotImplemented

    def __mul__(self, other):
        if _is_scalar(other):  # :65-71
            return OperatorSum([(op, c * complex(other)) for op, c in self.terms])
        if isinstance(other, Operator):  # :51-58
            return OperatorSum([(op * other, c) for op, c in self.terms])
        if isinstance(other, OperatorSum):  # :73-81, i-major
            out = OperatorSum()
            for op_i, c_i in self.terms:
                for op_j, c_j in other.terms:
                    out.append(op_i * op_j, c_i * c_j)
            return out
        return NotImplemented

    def __rmul__(self, other):
        if _is_scalar(other):  # :139-142
            return self * other
        return NotImplemented

    def has_modes(self):
        return any(r.kind == abi.QD_OP_MODE for op, _ in self.terms for r in op.ops)

    def program_modes(self):
        """(terms, nterms, ops, nops, mats, nmats, mdim): as program(), with
        every mode gate's matrix appended to mats (its site_a points there)."""
        recs = [r for op, _ in self.terms for r in op.ops if r.kind == abi.QD_OP_MODE]
        mdim = recs[0].U.shape[0] if recs else 1
        if any(r.U.shape != (mdim, mdim) for r in recs):
            raise ValueError("mode gates of one program must share the mode dimension")
        mats = []
        terms, nt, ops, no = self.program(allow_modes=True, mats=mats)
        flat = np.ascontiguousarray(np.array(mats, dtype=np.complex128).reshape(-1)).view(np.float64) \
            if mats else np.zeros(2)
        return terms, nt, ops, no, flat, len(mats), mdim

    def program(self, allow_modes=False, mats=None):
        """Flatten to (terms, ops) ctypes arrays in the reference's order; a
        mode gate's matrix is appended to mats, its site fields point there."""
        if not allow_modes and self.has_modes():
            raise ValueError("a program with mode gates: use program_modes()")
        nops = sum(len(op.ops) for op, _ in self.terms)
        terms = (abi.qd_term * max(1, len(self.terms)))()
        ops = (abi.qd_op * max(1, nops))()
        pos = 0
        for k, (op, c) in enumerate(self.terms):
            if len(op.ops) > abi.QD_MAX_OPS_PER_TERM:
                raise ValueError(f"term {k} has {len(op.ops)} ops (max {abi.QD_MAX_OPS_PER_TERM})")
            terms[k].coeff[0] = c.real
            terms[k].coeff[1] = c.imag
            terms[k].op_begin = pos
            terms[k].n_ops = len(op.ops)
            for rec in op.ops:
                o = ops[pos]
                o.kind = rec.kind
                o.site_a = rec.site_a
                o.site_b = rec.site_b
                o.string = rec.string
                if rec.kind == abi.QD_OP_MODE:
                    o.site_a = o.site_b = len(mats)
                    mats.append(rec.U)
                else:
                    for q in range(4):
                        u = rec.U[q // 2, q % 2]
                        o.U[2 * q] = u.real
                        o.U[2 * q + 1] = u.imag
                pos += 1
        return terms, len(self.terms), ops, nops

    def nbits_used(self):
        m = 0
        for op, _ in self.terms:
            for r in op.ops:
                if r.kind != abi.QD_OP_MODE:
                    m = max(m, r.site_a + 1, r.site_b + 1)
        return m


def getEmptyOperator():
    return OperatorSum()


# ---------------------------------------------------------------- models
# Hamiltonians of the benchmark configurations, written with the API above in
# the way a reference driver builds them (programs/spin.cpp:23-35,
# programs/hubbard_diss.cpp:125-159).

def xy_chain(N, J=1.0, pbc=False):
    """H = J sum_i (S+_{i+1} S-_i + S+_i S-_{i+1}) on N qubits (spin.cpp:23-35)."""
    idx = getQbits(N)
    gc, ga = getBoseGate(idx), getBoseGate(idx)
    gc << 0.5 * (pauli_x + II * pauli_y)
    ga << 0.5 * (pauli_x - II * pauli_y)
    ham = getEmptyOperator()
    nb = N if pbc else N - 1
    for i in range(nb):
        j = (i + 1) % N
        if J == 1.0:
            ham = ham + gc(j) * ga(i) + gc(i) * ga(j)
        else:
            ham = ham + J * (gc(j) * ga(i) + gc(i) * ga(j))
    return idx, ham


def xxz_chain(N, Jxy=1.0, Jz=1.0, pbc=True):
    """H = sum_i [Jxy/2 (S+_{i+1} S-_i + h.c.) + Jz Sz_{i+1} Sz_i] (S = sigma/2)."""
    idx = getQbits(N)
    gc, ga, gz = getBoseGate(idx), getBoseGate(idx), getBoseGate(idx)
    gc << 0.5 * (pauli_x + II * pauli_y)
    ga << 0.5 * (pauli_x - II * pauli_y)
    gz << 0.5 * pauli_z
    ham = getEmptyOperator()
    nb = N if pbc else N - 1
    for i in range(nb):
        j = (i + 1) % N
        ham = ham + (0.5 * Jxy) * (gc(j) * ga(i) + gc(i) * ga(j)) + Jz * (gz(j) * gz(i))
    return idx, ham


def sz_sector(idx, N, up):
    return Constrain(idx, getNset(idx) == up)


def hubbard_bonds_chain(L, pbc):
    return [(i, (i + 1) % L) for i in range(L if pbc else L - 1)]


def hubbard_bonds_square(Lx, Ly):
    """Nearest-neighbour bonds of an Lx x Ly periodic square lattice, each once."""
    bonds = []
    for y in range(Ly):
        for x in range(Lx):
            s = x + Lx * y
            for dx, dy in ((1, 0), (0, 1)):
                nx, ny = (x + dx) % Lx, (y + dy) % Ly
                t = nx + Lx * ny
                if (Lx == 2 and dx == 1 and x == 1) or (Ly == 2 and dy == 1 and y == 1):
                    continue  # a 2-site ring has one bond
                if s != t:
                    bonds.append((s, t))
    return bonds


def hubbard(nsite, bonds, t=1.0, U=4.0):
    """H = -t sum_<ij>,s (c+_i c_j + c+_j c_i) + U sum_i n_i,up n_i,dn.

    Up spins on bits 0..nsite-1, down spins on nsite..2 nsite-1, fermionic
    hop gates (hubbard_diss.cpp:97-159 layout)."""
    idx = getQbits(2 * nsite)
    gate = getHopFermiGate(idx)
    ham = getEmptyOperator()
    for (a, b) in bonds:
        for off in (0, nsite):
            ham = ham + (-t) * gate(a + off, b + off) + (-t) * gate(b + off, a + off)
    for i in range(nsite):
        ham = ham + U * (gate(i, i) * gate(i + nsite, i + nsite))
    return idx, ham


def holstein_chain(N, nd, w, g):
    """XY chain of N qubits next to one boson mode of nd levels, coupled to
    the total occupation: sum_i (S+_{i+1} S-_i + h.c.) + w a+a
    + g sum_i n_i (a + a+), with mode gates (operators/mode.hpp) in the
    reference's matrix convention.  Returns (qubit index, composite index, H)."""
    q = getQbits(N)
    md = SingleModeIndex(nd)
    idx = q * md
    gc, ga, gn = getBoseGate(q), getBoseGate(q), getBoseGate(q)
    gc << [[0, 1], [0, 0]]
    ga << [[0, 0], [1, 0]]
    gn << [[0, 0], [0, 1]]
    ad, a, num = getModeOp(md), getModeOp(md), getModeOp(md)
    ad << LadderMatrix(nd, MODE_LOWER)
    a << LadderMatrix(nd, MODE_UPPER)
    num << np.diag(np.arange(nd, dtype=float))
    ham = getEmptyOperator()
    for i in range(N - 1):
        ham = ham + gc(i + 1) * ga(i) + gc(i) * ga(i + 1)
    ham = ham + w * num(0)
    for i in range(N):
        ham = ham + g * (gn(i) * a(0)) + g * (gn(i) * ad(0))
    return q, idx, ham


def hubbard_sector(idx, nsite, nup, ndn):
    return Constrain(idx, getNup(idx, nsite) == nup, getNdo(idx, nsite) == ndn)


def xxz_parts(N, Jxy=1.0, pbc=True):
    """XXZ chain split as H(Jz) = A + Jz D: A = sum_i Jxy/2 (S+_{i+1} S-_i + h.c.),
    D = sum_i Sz_{i+1} Sz_i (diagonal), with the terms of xxz_chain."""
    idx = getQbits(N)
    gc, ga, gz = getBoseGate(idx), getBoseGate(idx), getBoseGate(idx)
    gc << 0.5 * (pauli_x + II * pauli_y)
    ga << 0.5 * (pauli_x - II * pauli_y)
    gz << 0.5 * pauli_z
    A = getEmptyOperator()
    D = getEmptyOperator()
    for i in range(N if pbc else N - 1):
        j = (i + 1) % N
        A = A + (0.5 * Jxy) * (gc(j) * ga(i) + gc(i) * ga(j))
        D = D + gz(j) * gz(i)
    return idx, A, D


def hubbard_parts(nsite, bonds, t=1.0):
    """Hubbard model split as H(U) = T + U D: T = -t sum_<ij>,s (c+_is c_js + h.c.),
    D = sum_i n_i,up n_i,dn (diagonal), with the terms of hubbard()."""
    idx = getQbits(2 * nsite)
    g = getHopFermiGate(idx)
    T = getEmptyOperator()
    for (a, b) in bonds:
        for off in (0, nsite):
            T = T + (-t) * g(a + off, b + off) + (-t) * g(b + off, a + off)
    D = getEmptyOperator()
    for i in range(nsite):
        D = D + g(i, i) * g(i + nsite, i + nsite)
    return idx, T, D
