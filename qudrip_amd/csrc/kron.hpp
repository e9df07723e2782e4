// kron.hpp -- two-block (Kronecker) form of a spinful lattice operator:
// with the sector a mixed radix of block A (low bits, nA states) and block B
// (high bits, nB states), H = T_A (x) I + I (x) T_B + D.  The description
// below is what the product kernels read; matfree_kron.hip builds it.
#pragma once

#include "qd_internal.hpp"

struct KronDesc {
  int64_t nA, nB;
  int loB;  // first bit of block B
  // T_A / T_B in pull form: row r = conj of column r of the Hermitian block
  const int32_t* Aouter;
  const int32_t* Ainner;
  const cplx* Avals;
  const int32_t* Bouter;
  const int32_t* Binner;
  const cplx* Bvals;
  const uint64_t* statesA;
  const uint64_t* statesB;
  const DevTerm* dterms;  // diagonal terms other than number products
  const DevOp* dops;
  int ndterms;
  // number-operator products c * prod n_s (hopBitOp(s, s), hopbit.hpp:50-51):
  // value c when every bit of the mask is set
  const uint64_t* nmask;
  const cplx* ncoef;
  int nnum;
  // packed T_A (ELL, entry j of row a at ell[j * nA + a]): column in the high
  // 16 bits, index into vtab in the low 16 bits; vtab[0] = 0 pads short rows.
  const uint32_t* ell;
  int ell_w;
  const cplx* vtab;
  int nvtab;
  // packed T_B, row-major (entry j of row b at ellB[b * ellB_w + j]), same
  // encoding; vtabB[0] = 0 pads short rows
  const uint32_t* ellB;
  int ellB_w;
  const cplx* vtabB;
  int nvtabB;
  // T_A compressed for registers (nA < 8192, at most 8 table values, rows of
  // at most 2 * KRON_WP entries): entry = col << 3 | value id (id 0 pads),
  // two entries per word, word p of row a at ell16[p * nA + a]
  const uint32_t* ell16;
  // T_A for the branch-free rows kernel (nA < 8191, same limits as ell16):
  // entry = (col + 1) << 3 | low, low = the value id, or in the sign form
  // (asign: every table value is +acoef or -acoef) bit 2 set for -acoef;
  // entry 0 pads (slot 0 of the kernel's row buffer holds 0); two entries
  // per word, word p of row a at ell16s[p * nA + a]
  const uint32_t* ell16s;
  int asign;
  cplx acoef;
  // T_B for the wave-per-row slab kernel (nA nB 16 < 2^32): row b at
  // ellB4[b * ellB4_stride]: the entry count, then per entry the byte offset
  // in x of its source row (col nA 16) | a code in the low 4 bits: in the
  // sign form (bsign: every value is +bcoef or -bcoef) 1 for +bcoef, 2 for
  // -bcoef; otherwise an id into vtabB4; code 0 (offset 0) pads the row to
  // ellB4_stride - 1 entries, a multiple of 8
  const uint32_t* ellB4;
  int ellB4_stride;
  int bsign;
  cplx bcoef;
  const cplx* vtabB4;
  // per product: an extra diagonal (n values, state b nA + a), or nullptr
  const cplx* ext_diag;
};
constexpr int KRON_WP = 12;  // words per row of the compressed T_A (24 entries)

// the description of a built Kronecker form (nullptr for nullptr)
const KronDesc* qd_kron_desc(const qd_kron* k);
// choose the launch variants now (x, y: n-vector scratch), so that later
// products from several threads only read them
int qd_kron_prepare(qd_ctx* ctx, qd_kron* k, const cplx* x, cplx* y);
