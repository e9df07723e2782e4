/*
 * qudrip_hip.h -- C ABI of the MI355X (gfx950) sparse-Hamiltonian Krylov path.
 *
 * This is the boundary a QuDrip caller binds to.  The reference library is
 * header-only C++ (no FFI); each entry point below replaces one call of its
 * template API, cited as reference file:line.  The C++ drop-in headers under
 * include/qudrip/ re-create that template API on top of these functions.
 *
 * Conventions
 *  - Every function returns an int status (QD_OK == 0); qd_last_error(ctx)
 *    returns a message for the last failure on that context.  No C++
 *    exception ever crosses this boundary.
 *  - Complex numbers are interleaved (re, im) doubles, the memory layout of
 *    std::complex<double>.
 *  - Arguments named d_* are DEVICE pointers (allocated with qd_malloc, or by
 *    any HIP allocator on the context's device).  Arguments named h_* or
 *    documented as host buffers are caller-owned host memory read or written
 *    only during the call.
 *  - A context owns one HIP stream.  Calls on one context must be serialised
 *    by the caller; distinct contexts are independent.
 *  - Handles (basis, termprog, csr, linop, krylov) own their device memory and
 *    must be destroyed before their context.
 */
#ifndef QUDRIP_HIP_H
#define QUDRIP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
  QD_OK = 0,
  QD_E_ARG = 1,         /* invalid argument                                   */
  QD_E_OOM = 2,         /* device allocation failed                           */
  QD_E_HIP = 3,         /* a HIP runtime call failed                          */
  QD_E_BREAKDOWN = 4,   /* Krylov breakdown (|w| < 1e-10) before any vector   */
  QD_E_EIGEN = 5,       /* small dense eigensolver / expm failed (NaN input)  */
  QD_E_CAPACITY = 6,    /* a row exceeds the kernel's per-row capacity        */
  QD_E_UNSUPPORTED = 7, /* sector or operator shape not supported             */
  QD_E_BUDGET = 8,      /* request does not fit the HBM budget                */
  QD_E_STATE = 9,       /* call not valid in the handle's current state       */
  QD_E_CANCELLED = 10   /* served request stopped by qd_server_cancel         */
};

/* ------------------------------------------------------------------ handles */
typedef struct qd_ctx qd_ctx;
typedef struct qd_basis qd_basis;
typedef struct qd_termprog qd_termprog;
typedef struct qd_csr qd_csr;
typedef struct qd_linop qd_linop;
typedef struct qd_krylov qd_krylov;

/* ------------------------------------------------------------ data formats */

/* One conserved quantity: popcount(x & mask) == count.
 * Replaces Nset (conservation.hpp:89-115) with `cq = value` (:78-81):
 * getNset masks bits 0..N (:233-236), getNup bits 0..N-1 (:238-241),
 * getNdo bits N..2N-1 (:243-246). */
typedef struct {
  uint64_t mask;
  int64_t count;
} qd_constraint;

/* One elementary gate, as stored by the reference after `gate << U` and
 * `gate(site)` / `gate(site2, site1)`.
 *   kind == QD_OP_BIT : bitOp (operators/bit.hpp:6-84).  site_a = site.
 *       U holds U_[i*2+j] = U(i,j) (operator.hpp:24-29) as 4 interleaved
 *       complex numbers; the child with output bit o of input bit b gets
 *       (jw * U(b,o)) * val (bit.hpp:57-63).
 *   kind == QD_OP_HOP : hopBitOp (operators/hopbit.hpp:6-84), gate(site2,
 *       site1) = c+_{site2} c_{site1}: site_a = site2, site_b = site1.
 *       U is unused.
 *   kind == QD_OP_MODE : modeOp (operators/mode.hpp:6-30,80-103) on the
 *       boson mode of a composite index (a QbitsIndex times a
 *       SingleModeIndex of dimension mdim, index.hpp:94-184).  site_a is the
 *       position of its matrix in the table given to
 *       qd_termprog_create_modes; the matrix is stored as elOpBase stores it
 *       (U_[i*mdim+j] = U(i,j)) and takes mode m to sum_j U(m,j) |j>
 *       (mode.hpp:86-90).  string and U are unused.
 *   string = +1 (getBoseGate / getHopBoseGate) or -1 (fermionic, Jordan-
 *       Wigner sign from the bits above the site). */
enum { QD_OP_BIT = 0, QD_OP_HOP = 1, QD_OP_MODE = 2 };
typedef struct {
  int32_t kind;
  int32_t site_a;
  int32_t site_b;
  int32_t string;
  double U[8];
} qd_op;

/* One term of an OperatorSum (operator_abel.hpp:9-128): coefficient and a
 * product of ops ops[op_begin .. op_begin+n_ops-1] in Operator::ops_ order
 * (leftmost factor first; applied right-to-left, operator.hpp:151). */
typedef struct {
  double coeff[2];
  int32_t op_begin;
  int32_t n_ops;
} qd_term;

#define QD_MAX_OPS_PER_TERM 8
#define QD_MAX_BITS 64
#define QD_MAX_MODE_DIM 16

/* ------------------------------------------------------------------ context */
int qd_ctx_create(int device, qd_ctx** out);
/* Same, but the context issues its work on an existing hipStream_t owned by
 * the caller (e.g. a framework's current stream), so the two orders agree. */
int qd_ctx_create_on_stream(int device, void* stream, qd_ctx** out);
int qd_ctx_destroy(qd_ctx* ctx);
const char* qd_last_error(const qd_ctx* ctx);
int qd_ctx_sync(qd_ctx* ctx);
/* The context's hipStream_t, for callers that interleave their own work. */
void* qd_ctx_stream(qd_ctx* ctx);
/* Number of ABI functions; lets a loader check it bound the right library. */
int qd_abi_version(void);

/* ------------------------------------------------------- device memory */
int qd_malloc(qd_ctx* ctx, size_t bytes, void** d_ptr);
int qd_free(qd_ctx* ctx, void* d_ptr);
int qd_memcpy_h2d(qd_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
int qd_memcpy_d2h(qd_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);

/* ------------------------------------------------------------------- basis */
/* Constrain(QbitsIndex(nbits), cq = v, ...) (conservation.hpp:212-231):
 * all x in [0, 2^nbits) meeting every constraint, in ascending x.  The
 * states are produced on the GPU by combinatorial-number-system unranking
 * (disjoint contiguous masks) or by a predicate scan (any other masks). */
int qd_basis_create(qd_ctx* ctx, int nbits, const qd_constraint* cons, int ncons,
                    qd_basis** out);
/* A conserved quantity in the form shared by the reference's Nset and Nd
 * classes: value(x) = popcount(((x >> shift) ^ flip) & (x ^ flip) & sel).
 *   Nset (conservation.hpp:89-115): shift 0, flip 0, sel = its mask;
 *   Nd   (conservation.hpp:117-151; getNd :248-251): shift Nsite, flip 0,
 *        sel all bits: the doubly occupied sites;
 *   Nh   (getNh :253-256): shift Nsite, flip bits 0..Nsite-1, sel all bits:
 *        the empty sites. */
typedef struct {
  uint64_t sel;
  uint64_t flip;
  int32_t shift; /* 0..63 */
  int32_t reserved;
  int64_t count;
} qd_quantity;
/* Constrain(index, cq = v, ...) (conservation.hpp:212-231) over such
 * quantities (at most 16).  Plain Nset lists take the qd_basis_create path;
 * with Nd / Nh the sector of the Nset part is enumerated (or, if that part is
 * not a disjoint set of bit ranges, all 2^nbits states are scanned, nbits <=
 * 40) and filtered in order; such a basis ranks by binary search. */
int qd_basis_create_q(qd_ctx* ctx, int nbits, const qd_quantity* q, int nq, qd_basis** out);
/* Constrain over a composite index QbitsIndex(nbits) * SingleModeIndex(mdim)
 * (Indices, index.hpp:112-184): the quantities act on the qubit part q; the
 * sector is every composite value q * mdim + m with q satisfying them, in
 * ascending order (q-major), for 2 <= mdim <= QD_MAX_MODE_DIM; rank(x) =
 * rank(q) * mdim + m, a miss ranks to the size. */
int qd_basis_create_modes(qd_ctx* ctx, int nbits, int mdim, const qd_quantity* q, int nq, qd_basis** out);
int qd_basis_destroy(qd_basis* basis);
/* SubIndex::range() (index.hpp:261). */
int64_t qd_basis_size(const qd_basis* basis);
/* SubIndex::mapping, i.e. the unrank table (index.hpp:242-248); host buffer
 * of qd_basis_size() entries. */
int qd_basis_download(qd_basis* basis, uint64_t* h_states);
/* Device pointer to the same table (valid while the basis lives). */
const uint64_t* qd_basis_device_states(qd_basis* basis);
/* SubIndex::operator idxv_type() (index.hpp:234-240): rank of each state;
 * a state outside the sector ranks to qd_basis_size().  Device arrays. */
int qd_basis_rank(qd_basis* basis, const uint64_t* d_states, int64_t count,
                  int64_t* d_ranks);

/* -------------------------------------------------------------- term program */
/* An OperatorSum flattened by the drop-in headers.  nbits is the width of
 * the QbitsIndex the gates act on; every site must be < nbits. */
int qd_termprog_create(qd_ctx* ctx, int nbits, const qd_term* terms, int nterms,
                       const qd_op* ops, int nops, qd_termprog** out);
/* The same over a composite index (QbitsIndex * SingleModeIndex(mdim),
 * index.hpp:112-184): states are the composite values q * mdim + m, bit and
 * hop gates act on q, QD_OP_MODE gates on m.  mats: nmats matrices of
 * mdim * mdim interleaved complex numbers, U_[i*mdim+j] = U(i,j) each;
 * 2 <= mdim <= QD_MAX_MODE_DIM.  Such a program builds stored matrices
 * (qd_csr_build) over a basis made by qd_basis_create_modes with the same
 * mdim; it has no matrix-free product.  A program without mode gates
 * (qd_termprog_create) over such a basis acts as the identity on the mode. */
int qd_termprog_create_modes(qd_ctx* ctx, int nbits, int mdim, const qd_term* terms, int nterms,
                             const qd_op* ops, int nops, const double* mats, int nmats, qd_termprog** out);
int qd_termprog_destroy(qd_termprog* prog);

/* --------------------------------------------------------------- CSR build */
/* `ham >> index` (operator_abel.hpp:25-34 over operator.hpp:235-267):
 * two-pass count/fill on the GPU.  The arrays are the reference's Eigen
 * ColMajor storage: outer[n+1] indexed by SOURCE state (column), inner[nnz]
 * target rows ascending, values complex128 -- bit-identical to the
 * reference's summation order (duplicates of one term in leaf order, then
 * acc = (+0,+0); acc += c_k * v_k in term order; explicit zeros kept). */
int qd_csr_build(qd_ctx* ctx, qd_basis* basis, qd_termprog* prog, qd_csr** out);
int qd_csr_destroy(qd_csr* csr);
int qd_csr_info(const qd_csr* csr, int64_t* n, int64_t* nnz, int* hermitian);
/* Host buffers of n+1, nnz and 2*nnz entries. */
int qd_csr_download(qd_csr* csr, int32_t* h_outer, int32_t* h_inner, double* h_vals);
/* Device pointers to the same arrays (valid while the csr lives). */
int qd_csr_device_arrays(qd_csr* csr, const int32_t** d_outer, const int32_t** d_inner,
                         const double** d_vals);

/* ------------------------------------------------------ SpMatrix algebra */
/* Arithmetic between stored n x n matrices, on the device, each returning a
 * new matrix in the same ColMajor storage, with the element rules of the
 * Eigen expressions the reference's callers write on SpMatrix
 * (programs/hubbard_diss.cpp:106-159,443-446,514-517).  Operands must have
 * the same n; they are read only.  A result without a term program gets
 * its adjoint (for the product kernels) from qd_csr_adjoint when needed. */
/* SpMatrix(n, n) / setZero(): no stored entries (hubbard_diss.cpp:106-116). */
int qd_csr_zero(qd_ctx* ctx, int64_t n, qd_csr** out);
/* setIdentity(): one entry (1, 0) per column (hubbard_diss.cpp:117). */
int qd_csr_identity(qd_ctx* ctx, int64_t n, qd_csr** out);
/* c * A: every value times c (the complex product; real_scalar != 0 scales
 * both parts by c_re, Eigen's double * complex), pattern and explicit zeros
 * kept (hubbard_diss.cpp:135,443,516). */
int qd_csr_scale(qd_ctx* ctx, const qd_csr* a, double c_re, double c_im, int real_scalar, qd_csr** out);
/* A + B (sign = 1) or A - B (sign = -1), also A += B and A -= B
 * (hubbard_diss.cpp:135-156,443,516): the union of the patterns; an entry
 * of one side only is a + 0, 0 + b or 0 - b; no pruning. */
int qd_csr_add(qd_ctx* ctx, const qd_csr* a, const qd_csr* b, int sign, qd_csr** out);
/* SpMatrix(A.adjoint()) (hubbard_diss.cpp:156,526): the conjugate transpose,
 * rows ascending in each column. */
int qd_csr_adjoint(qd_ctx* ctx, const qd_csr* a, qd_csr** out);
/* A * B (hubbard_diss.cpp:156,158,526): the conservative sparse product;
 * column j visits B's column j in order and for each (k, y) A's column k in
 * order, the first visit of a row stores x * y and later ones add x * y.
 * QD_E_CAPACITY if a column visits more than 1024 entries of A. */
int qd_csr_multiply(qd_ctx* ctx, const qd_csr* a, const qd_csr* b, qd_csr** out);

/* ------------------------------------------------------- linear operators */
/* y = H x through the stored matrix. */
int qd_linop_from_csr(qd_ctx* ctx, qd_csr* csr, qd_linop** out);
/* y = H x evaluated on the fly from the term program (no stored matrix);
 * the same H as `prog >> basis` (operator.hpp:215-232 is the lazy form). */
int qd_linop_matfree(qd_ctx* ctx, qd_basis* basis, qd_termprog* prog, qd_linop** out);
int qd_linop_destroy(qd_linop* op);
/* A view of `base` whose products add diag(d): (base + diag(d)) x, with
 * d_diag n complex values on the device.  Borrows base's arrays (stored
 * matrix, or matrix-free program and its two-block form) and d_diag (both
 * must outlive the view).  Lets many operators H = A + c D share one A
 * (no reference counterpart; used by the solve server below). */
int qd_linop_with_diagonal(qd_ctx* ctx, qd_linop* base, const double* d_diag, qd_linop** out);
int64_t qd_linop_dim(const qd_linop* op);
/* Matrix-free operators over a two-block sector (up bits low, down bits
 * high) whose terms are hops inside the low block, ops inside the high block
 * or diagonal are applied as H = T_A (x) 1 + 1 (x) T_B + D with small block
 * matrices; on = 0 forces the generic per-row kernel.  *has_structured
 * (may be NULL) reports whether the operator has that form. */
int qd_linop_set_structured(qd_linop* op, int on, int* has_structured);
/* Row mapping of the stored-matrix product: 0, -1, ..., -14 select a
 * CSR-stream variant (a wave streams 64 rows' nonzeros through LDS; the
 * variants differ in LDS chunk, loads in flight, cache hint and software
 * pipelining; all give the same bits), 2..64 the lanes per row of the vector
 * kernel.  Chosen from nnz/n at creation: -13 (pipelined, 512 entries per
 * slice) for rows of up to 48 entries on average. */
int qd_linop_set_lanes(qd_linop* op, int lanes);
int qd_linop_get_lanes(const qd_linop* op);
/* Tile autotuner: times every row mapping on d_x -> d_y (reps launches
 * each, HIP events on the context stream) and keeps the fastest; reports it
 * and its time per launch in microseconds (host pointers, may be NULL). */
int qd_linop_autotune(qd_ctx* ctx, qd_linop* op, const double* d_x, double* d_y, int reps, int* best,
                      float* best_us);

/* ------------------------------------------------------- vector kernels */
/* d_y = H d_x (complex128, n entries). */
int qd_spmv(qd_ctx* ctx, qd_linop* op, const double* d_x, double* d_y);
/* d_y += alpha * H d_x, alpha host[2] (OperatorSum::map_acc,
 * operator_abel.hpp:36-42; with a matrix-free operator this is the lazy
 * apply of operator.hpp:215-232). */
int qd_spmv_acc(qd_ctx* ctx, qd_linop* op, const double* h_alpha, const double* d_x, double* d_y);
/* <psi|H|psi>, unnormalised (State::eval, state.hpp:33-44).  out: host[2]. */
int qd_expect(qd_ctx* ctx, qd_linop* op, const double* d_psi, double* h_out);
/* The same value with no device allocation (no counterpart in the
 * reference): for callers that cannot allocate -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Stored-
 * matrix operators (qd_linop_from_csr) of at most 33,554,432 rows
 * (QD_E_CAPACITY beyond; QD_E_UNSUPPORTED for other operators).  The
 * product is fused with the dot and never stored; block sums use a fixed
 * 4 MiB device buffer per GPU held until the process ends.  d_out: device
 * double[2], written asynchronously on the context's stream (no memset
 * node: the call may be captured in a graph).  Calls that overlap on one GPU
 * share that buffer and race: serialise them. */
int qd_expect_noalloc(qd_ctx* ctx, qd_linop* op, const double* d_psi, double* d_out);
/* Re(psi^H psi) (Hermitian::norm, sparse_algorithm.hpp:7-16). */
int qd_norm2(qd_ctx* ctx, int64_t n, const double* d_psi, double* h_out);
/* Seeded replacement of shuffle_state (utils.hpp:129-140): re, im of entry
 * i from splitmix64((seed << 32) ^ (2i + c)) mapped to U(-1, 1), then
 * normalised. */
int qd_seeded_state(qd_ctx* ctx, int64_t n, uint64_t seed, double* d_psi);
/* Achievable streaming bandwidth (no reference counterpart; the ceiling that
 * HBM-bound kernels are read against): a non-temporal read of `bytes`, and a
 * copy of `bytes` into a second buffer, each averaged over `reps` launches
 * after a warm-up.  Allocates both buffers for the call. */
int qd_stream_bandwidth(qd_ctx* ctx, size_t bytes, int reps, double* read_gbs, double* copy_gbs);

/* ------------------------------------------------------------ Krylov solver */
/* Arnoldi::getSolver (herm = 0) / getLanczosSolver (herm = 1)
 * (sparse_algorithm.hpp:596-604).  Keeps up to kry_dim+1 basis vectors on
 * the device, full (CGS2) reorthogonalisation, FP64 throughout. */
int qd_krylov_create(qd_ctx* ctx, int64_t n, int kry_dim, int herm, qd_krylov** out);
int qd_krylov_destroy(qd_krylov* kr);
/* ArnoldiSolver::getGroundState(H, psi, tstp, add_dim) (:564-585): starts
 * from d_psi if no basis is held yet, extends the basis by add_dim vectors
 * (-1: up to kry_dim), solves the tridiagonal Ritz problem in FP64 and
 * overwrites d_psi with the lowest Ritz vector (not renormalised).
 * E0 (host, may be NULL) receives the lowest Ritz value. */
int qd_krylov_ground_state(qd_krylov* kr, qd_linop* op, double* d_psi, int add_dim,
                           double* h_E0);
/* ArnoldiSolver::evolve(H, psi, tstp, h, kry_dim_rt) (:548-562):
 * d_psi_out = V_m expm(-i h K_m) e_1 with V_m[:,0] = psi_in / |psi_in| and
 * m = kry_dim_rt - 1 (kry_dim_rt = 0: kry_dim); like the reference the
 * input norm is not carried and the result is not renormalised.
 * d_psi_out may equal d_psi_in. */
int qd_krylov_evolve(qd_krylov* kr, qd_linop* op, const double* d_psi_in, double* d_psi_out,
                     double h, int kry_dim_rt);
/* evolve with the Krylov dimension chosen per step (no reference
 * counterpart; the reference fixes kry_dim_rt): the basis grows from m_min by
 * 2 vectors until Saad's a-posteriori estimate of the step error,
 * h_{m+1,m} |[expm(-i h K_m)]_{m,1}|, is <= tol, or it holds kry_max - 1
 * vectors.  m_used (may be NULL) receives the dimension used, m_enough (may
 * be NULL) the smallest leading dimension whose estimate met tol, a start for
 * the next step. */
int qd_krylov_evolve_tol(qd_krylov* kr, qd_linop* op, const double* d_psi_in, double* d_psi_out, double h,
                         int m_min, int kry_max, double tol, int* m_used, int* m_enough);
typedef struct {
  int act_dim;      /* ArnoldiSolver::act_dim() */
  int breakdown;    /* 1 if the last expansion stopped on |w| < 1e-10 */
  double error;     /* ArnoldiSolver::error() */
  double beta_last; /* trailing |w| of the last expansion */
  int64_t n_spmv;   /* operator applications so far */
  double ritz_resid; /* |(H - E0) psi| of the last ground-state Ritz pair,
                        beta_m |z_m| (no reference counterpart) */
  double ritz_gap;   /* second minus lowest Ritz value of that solve */
  int full_steps;    /* method 2: steps that projected against the whole
                        basis since the basis was started */
} qd_krylov_info;
int qd_krylov_get_info(qd_krylov* kr, qd_krylov_info* info);
/* Ritz values of the last ground-state solve, ascending (storeE, :587-593).
 * Writes min(cap, act_dim) values; returns their count through *count. */
int qd_krylov_eigenvalues(qd_krylov* kr, double* h_evals, int cap, int* count);
/* ArnoldiSolver::convergeEV(N, iv) (:494-500). */
int qd_krylov_converge_ev(qd_krylov* kr, int N, double iv_re, double iv_im);
/* Forget the held basis (act_dim = 0). */
int qd_krylov_reset(qd_krylov* kr);
/* Reorthogonalisation scheme: 1 = delayed CGS2 (two passes over the basis
 * per step; default), 0 = CGS2 (four passes).  For Hermitian operators only
 * (no reference counterpart): 2 = delayed CGS2 over the last two basis
 * vectors with partial reorthogonalisation (Simon 1984): a recurrence
 * estimates the loss of orthogonality and, when it passes
 * sqrt(eps / kry_dim) / 4, the next two steps project against the whole
 * basis, so the basis stays semi-orthogonal and the tridiagonal matrix
 * agrees with a fully reorthogonalised one to O(eps |H|); 3 = the local
 * window only (no estimate; Ritz values may acquire spurious copies).
 * Takes effect when a basis is started. */
int qd_krylov_set_method(qd_krylov* kr, int method);

/* ------------------------------------------- Hermitian two-vector Lanczos */
/* Hermitian::LanczosSolver / Hermitian::getSolver(index, kry_dim)
 * (sparse_algorithm.hpp:236-313): the three-term recurrence on two vectors,
 * without reorthogonalisation, in two passes per solve.  Pass 1
 * (getKrylovReprAd, :67-121) builds the tridiagonal: q_0 = psi / |psi|,
 * a_i = q_i^H H q_i (complex), r = H q_i - a_i q_i - b_{i-1} q_{i-1},
 * b_i = |r|; it stops when |r|^2 <= 1e-10 or once kry_dim + 1 coefficients
 * a are held.  The tridiagonal (Re a, b) is solved in FP64 on the host.
 * Pass 2 (getKrylovGS, :166-233) reruns the recurrence with Re a and
 * accumulates sum_i c_i q_i.  Holds three device vectors of n entries. */
typedef struct qd_hlanczos qd_hlanczos;
int qd_hlanczos_create(qd_ctx* ctx, int64_t n, int kry_dim, qd_hlanczos** out);
int qd_hlanczos_destroy(qd_hlanczos* lz);
/* LanczosSolver::getGroundState(H, psi, tstp) (:285-300): d_psi is
 * overwritten with sum_i U[i][0] q_i, U[:,0] the eigenvector of the lowest
 * eigenvalue (not renormalised); E0 (host, may be NULL) receives that
 * eigenvalue. */
int qd_hlanczos_ground_state(qd_hlanczos* lz, qd_linop* op, double* d_psi, double* h_E0);
/* LanczosSolver::evolve(H, psi, tstp, h) (:257-283): d_psi_out =
 * sum_i c_i q_i with c = U diag(exp(-i lam h)) U^T e_0, i.e. the Krylov
 * approximation of exp(-i h H) psi_in / |psi_in|.  d_psi_out may equal
 * d_psi_in. */
int qd_hlanczos_evolve(qd_hlanczos* lz, qd_linop* op, const double* d_psi_in, double* d_psi_out, double h);
typedef struct {
  int dim;         /* coefficients a of the last pass 1 (the Krylov matrix's order) */
  double error;    /* LanczosSolver::error(): |r| after the last step of pass 1 */
  int64_t n_spmv;  /* operator applications of the recurrences so far */
} qd_hlanczos_info;
int qd_hlanczos_get_info(qd_hlanczos* lz, qd_hlanczos_info* info);
/* eigenvalues of the last Krylov matrix, ascending (LanczosSolver::storeE,
 * :302-308); writes min(cap, dim) values */
int qd_hlanczos_eigenvalues(qd_hlanczos* lz, double* h_evals, int cap, int* count);
/* the last pass 1's a_i (interleaved complex, 2 doubles each) and b_i
 * (b_i couples q_i and q_{i+1}; the last one is error()); min(cap, dim) of
 * each */
int qd_hlanczos_coefficients(qd_hlanczos* lz, double* h_alpha, double* h_beta, int cap, int* count);

/* ------------------------------------------------ batched trajectories */
/* Quantum-jump trajectories of a spinful fermion model (up spins on bits
 * 0..nsite-1, down spins on nsite..2nsite-1), the loop of
 * programs/hubbard_diss.cpp:440-482 with mc_manager.hpp:88-128 averaging,
 * generalised to dephasing (L = sqrt(rate) n_mode) and loss (L = sqrt(rate)
 * c_mode, fermionic) channels; a loss moves a trajectory to the sector with
 * one particle less.  All trajectories of one sector advance together
 * (multi-RHS product, per-trajectory Arnoldi with kry_dim_rt, device-side
 * jump sampling).  When every up mode carries the same total jump rate and
 * every down mode likewise, H_eff is H plus a number on each sector and the
 * Arnoldi step runs as the equivalent Lanczos recurrence (same Krylov space
 * and projected exponential; QD_TRAJ_KRYLOV=arnoldi in the environment keeps
 * the Arnoldi form).  Random numbers: key = splitmix64(seed*0x9E3779B97F4A7C15 +
 * traj), r_k(step) = (splitmix64(key + 4 step + k) >> 11) 2^-53. */
enum { QD_JUMP_DEPHASE = 0, QD_JUMP_LOSS = 1 };
typedef struct {
  int32_t kind;
  int32_t mode;
  double rate;
} qd_jump;
typedef struct {
  int64_t jumps;
  int64_t losses;
  int32_t sectors;
  int32_t lanczos; /* 1: the steps ran in the Lanczos form (every sector's
                    * sum_c L_c^+ L_c constant, H_eff = H + c_t there);
                    * 0: the Arnoldi form */
  int32_t lanczos_fallback_steps; /* steps of the Lanczos form recomputed fully
                                   * reorthogonalised because a trajectory's
                                   * estimated loss of orthogonality (Simon's
                                   * recurrence) passed the trigger */
} qd_traj_stats;
/* H: term program over 2*nsite bits conserving each spin's particle number.
 * d_psi0: start state in sector (nup0, ndn0) (device, ascending basis order).
 * h_dens_mean[(nsteps+1) * 2nsite]: <n_mode>(t_i) averaged over trajectories;
 * h_dens_var (optional): their variance over trajectories; h_jump_log
 * (optional, ntraj * nsteps): channel taken per step, -1 for none. */
int qd_traj_run(qd_ctx* ctx, int nsite, const qd_term* terms, int nterms, const qd_op* ops, int nops, int nup0,
                int ndn0, const double* d_psi0, const qd_jump* jumps, int njumps, int ntraj, int nsteps,
                double dt, int kry_dim_rt, uint64_t seed, double* h_dens_mean, double* h_dens_var,
                int* h_jump_log, qd_traj_stats* stats);

/* ------------------------------------------------------------ serving */
/* A solve server: no reference counterpart (the reference runs one solve per
 * process, programs/spin.cpp, programs/hubbard_diss.cpp); every request runs
 * the reference's own solver calls (getGroundState with add_dim and the
 * change of the lowest Ritz value, sparse_algorithm.hpp:564-585; evolve,
 * :548-562) on the device.  A model is a sector and two operators, A and a
 * diagonal D; a request solves H = A + coeff * D.  The server stores the
 * basis, A's matrix and D's diagonal once per model; workers (one thread and
 * stream each) admit queued requests in arrival order whenever a worker is
 * free and the request's device memory fits in hbm_budget (later, smaller
 * requests may pass a waiting one a bounded number of times; a request that
 * can never fit is answered with QD_E_OOM).  A request's Krylov state stays on
 * the device until it finishes; its result does not depend on what else
 * runs.  Request buffers (Krylov bases, vectors) come from a caching pool of
 * the server, capped at hbm_budget and released by qd_server_destroy; the
 * workers' device contexts (one stream each) outlive the server and pass to
 * the next server created in the process. */
typedef struct qd_server qd_server;
enum { QD_REQ_GROUND_STATE = 0, QD_REQ_EVOLVE = 1 };
typedef struct {
  int32_t kind;   /* QD_REQ_GROUND_STATE or QD_REQ_EVOLVE */
  int32_t model;  /* from qd_server_add_model */
  double coeff;   /* H = A + coeff * D */
  int32_t k;      /* ground state: Krylov dimension limit; evolve: kry_dim_rt */
  int32_t chunk;  /* ground state: basis vectors added per getGroundState call */
  double tol;     /* ground state: stop when the Ritz-residual bound resid^2 / gap
                   * or the move of E0 over one chunk is <= tol |E0|
                   * (0: one call with the full dimension, like the plain solver);
                   * evolve: step-error bound of qd_krylov_evolve_tol with m_min =
                   * chunk (0: qd_krylov_evolve with kry_dim_rt = k) */
  int32_t steps;  /* evolve: number of time steps */
  double dt;      /* evolve: time step */
  uint64_t seed;  /* start vector: qd_seeded_state(seed) */
  int64_t tag;    /* caller's id, returned in the result */
} qd_request;
typedef struct {
  int64_t tag;
  int32_t status;      /* QD_OK or an error code (QD_E_OOM: never fits the budget) */
  int32_t iterations;  /* ground state: Krylov dimension used; evolve: steps */
  double energy;       /* ground state: E0; evolve: <H> / <psi|psi> at the end */
  double energy_start; /* evolve: <H> / <psi|psi> of the start vector */
  double observable;   /* <D> / <psi|psi> of the final state */
  double t_submit, t_start, t_end; /* seconds on the server's clock */
  double krylov_mean;  /* Krylov dimension used (mean over steps for evolve) */
  int32_t worker;
} qd_result;
typedef struct {
  int64_t submitted, finished, rejected, queued;
  int32_t running, max_running;
  int64_t bytes_in_use, peak_bytes;
} qd_server_stats;
int qd_server_create(int device, size_t hbm_budget, int workers, qd_server** out);
/* Sector (nbits, cons) with operators A (tA, oA) and diagonal D (tD, oD;
 * ntD = 0: none).  structured != 0: evolve requests use the matrix-free
 * product of the combined program (two-block Kronecker form when present). */
int qd_server_add_model(qd_server* s, int nbits, const qd_constraint* cons, int ncons, const qd_term* tA, int ntA,
                        const qd_op* oA, int noA, const qd_term* tD, int ntD, const qd_op* oD, int noD,
                        int structured, int* model_id);
int qd_server_submit(qd_server* s, const qd_request* req);
/* Evolve requests of one model with equal kry_dim_rt and dt run in lockstep
 * batches of up to `slots` (one multi-RHS Arnoldi step per time step over all
 * of them, on the model's stored matrix; arrivals join at the next step).
 * 1 (default): each request alone on its worker. */
int qd_server_set_batch(qd_server* s, int slots);
/* Orthogonalisation scheme of the Krylov bases of requests submitted from
 * now on, ground states and (unbatched) real-time steps alike
 * (qd_krylov_set_method): 2 (default) = local window with partial
 * reorthogonalisation, 1 = full delayed CGS2 (the scheme of
 * qd_krylov_ground_state and qd_krylov_evolve), 0 = CGS2.  No reference
 * counterpart. */
int qd_server_set_gs_method(qd_server* s, int method);
/* Cancels every request with this tag: a queued one is answered at once
 * (status QD_E_CANCELLED, worker -1); a running one stops at its next
 * cancellation point and is answered with QD_E_CANCELLED (iterations = time
 * steps done).  The cancellation points are the reference's own loop
 * boundaries: between getGroundState extensions of a ground state with a
 * tolerance (sparse_algorithm.hpp:564-585, one per chunk), between evolve
 * calls of a real-time request (:548-562, one per time step), and between
 * the lockstep steps of a batch; a ground state with tol = 0 is one call and
 * is reported cancelled when it returns.  QD_E_ARG if no queued or running
 * request has the tag.  No reference counterpart (the reference runs one
 * solve per process). */
int qd_server_cancel(qd_server* s, int64_t tag);
/* Moves up to cap finished results (in completion order) into out. */
int qd_server_poll(qd_server* s, qd_result* out, int cap, int* count);
/* Blocks until every submitted request has finished. */
int qd_server_drain(qd_server* s);
double qd_server_clock(qd_server* s);
int qd_server_get_stats(qd_server* s, qd_server_stats* st);
const char* qd_server_last_error(qd_server* s);
/* Finishes the queued requests, then stops the workers and frees the models. */
int qd_server_destroy(qd_server* s);

/* ------------------------------------------------- FP64 host numerics */
/* The small dense problems of the solvers, exposed for direct checks.
 * Symmetric tridiagonal eigenpairs (d[m], e[m-1]) -> w[m] ascending and,
 * if z != NULL, column-major eigenvectors z[m*m] (SelfAdjointEigenSolver on
 * the Krylov matrix, sparse_algorithm.hpp:571). */
int qd_host_tridiag_eig(int m, const double* d, const double* e, double* w, double* z);
/* E = expm(A), m x m complex column-major interleaved (MatrixExp,
 * matrix.hpp:20; used at sparse_algorithm.hpp:559). */
int qd_host_expm(int m, const double* A, double* E);

#ifdef __cplusplus
}
#endif

#endif /* QUDRIP_HIP_H */
