// termprog.hip -- flattened OperatorSum on the device, and its adjoint.
//
// A term is c_k * (op_0 op_1 ... op_{m-1}) with the ops applied right-to-left
// (operator.hpp:151).  The adjoint program (used by pull-form products for
// non-Hermitian sums and by the matrix-free kernel) is
//     A^dagger = sum_k conj(c_k) op_{m-1}^dagger ... op_0^dagger,
// where for a bitOp (operators/bit.hpp:48-84) the adjoint gate has
// U'(i,j) = conj(U(j,i)) on the same site (the Jordan-Wigner factor only
// depends on the bits above the site, which the gate leaves unchanged), and
// for hopBitOp (operators/hopbit.hpp:43-84) (c+_a c_b)^dagger = c+_b c_a
// carries the same sign: the two popcounts exchange roles.  A mode gate
// (operators/mode.hpp) on a composite index gets the conjugate transpose of
// its matrix, stored after the program's own matrices.
#include "qd_internal.hpp"

static void finish_terms(std::vector<DevTerm>& terms, const std::vector<DevOp>& ops, int* max_leaves,
                         int mdim = 1) {
  int ml = 1;
  for (auto& t : terms) {
    int nl = 1;
    int dup = 0;
    int nmode = 0;
    for (int i = t.n_ops - 1; i >= 0; --i) {  // application order
      const DevOp& oi = ops[t.op_begin + i];
      if (oi.kind == QD_OP_MODE) {
        nl *= mdim;
        if (++nmode >= 2) dup = 1;  // two mode gates: several paths reach one mode
      } else if (oi.kind == QD_OP_BIT) {
        nl *= 2;
        // a later-applied op touching the same site can merge two leaves
        for (int j = i - 1; j >= 0; --j) {
          const DevOp& oj = ops[t.op_begin + j];
          if (oj.kind == QD_OP_BIT && oj.a == oi.a) dup = 1;
          if (oj.kind == QD_OP_HOP && oj.a != oj.b && (oj.a == oi.a || oj.b == oi.a)) dup = 1;
        }
      }
    }
    t.n_leaves = nl;
    t.may_dup = dup;
    if (nl > ml) ml = nl;
  }
  *max_leaves = ml;
}

static int termprog_create(qd_ctx* ctx, int nbits, int mdim, const qd_term* terms, int nterms, const qd_op* ops,
                           int nops, const double* mats, int nmats, qd_termprog** out) {
  if (!ctx || !out || nterms < 0 || nops < 0 || (nterms && !terms) || (nops && !ops) || nbits < 1 ||
      nbits > QD_MAX_BITS || mdim < 1 || mdim > QD_MAX_MODE_DIM || nmats < 0 || (nmats && !mats) ||
      (mdim == 1 && nmats > 0))
    return qd_set_error(ctx, QD_E_ARG, "qd_termprog_create: bad arguments");
  *out = nullptr;
  qd_termprog* p = new qd_termprog();
  p->ctx = ctx;
  p->nbits = nbits;
  p->mdim = mdim;
  p->nmats = nmats;
  p->h_ops.resize(nops);
  for (int i = 0; i < nops; ++i) {
    const qd_op& o = ops[i];
    if (o.kind == QD_OP_MODE) {
      if (mdim < 2 || o.site_a < 0 || o.site_a >= nmats) {
        delete p;
        return qd_set_error(ctx, QD_E_ARG, "op %d: mode gate without a matrix (or no mode)", i);
      }
      DevOp d{};
      d.kind = QD_OP_MODE;
      d.a = d.b = o.site_a;
      d.string = 1;
      p->h_ops[i] = d;
      continue;
    }
    if ((o.kind != QD_OP_BIT && o.kind != QD_OP_HOP) || o.site_a < 0 || o.site_a >= nbits ||
        (o.kind == QD_OP_HOP && (o.site_b < 0 || o.site_b >= nbits)) || (o.string != 1 && o.string != -1)) {
      delete p;
      return qd_set_error(ctx, QD_E_ARG, "op %d: bad kind/site/string", i);
    }
    DevOp d;
    d.kind = o.kind;
    d.a = o.site_a;
    d.b = (o.kind == QD_OP_HOP) ? o.site_b : o.site_a;
    d.string = o.string;
    for (int q = 0; q < 4; ++q) d.U[q] = cmake(o.U[2 * q], o.U[2 * q + 1]);
    p->h_ops[i] = d;
  }
  p->h_terms.resize(nterms);
  for (int k = 0; k < nterms; ++k) {
    const qd_term& t = terms[k];
    if (t.n_ops < 1 || t.n_ops > QD_MAX_OPS_PER_TERM || t.op_begin < 0 || t.op_begin + t.n_ops > nops) {
      delete p;
      return qd_set_error(ctx, QD_E_ARG, "term %d: bad op range", k);
    }
    DevTerm d{};
    d.coeff = cmake(t.coeff[0], t.coeff[1]);
    d.op_begin = t.op_begin;
    d.n_ops = t.n_ops;
    p->h_terms[k] = d;
  }
  for (int k = 0; k < nterms; ++k) {  // leaves per term: 2 per bit gate, mdim per mode gate
    int64_t nl = 1;
    for (int i = 0; i < p->h_terms[k].n_ops; ++i) {
      const int kind = p->h_ops[p->h_terms[k].op_begin + i].kind;
      nl *= kind == QD_OP_BIT ? 2 : kind == QD_OP_MODE ? mdim : 1;
    }
    if (nl > (1 << 20)) {
      delete p;
      return qd_set_error(ctx, QD_E_UNSUPPORTED, "term %d has %lld leaves (at most 2^20)", k, (long long)nl);
    }
  }
  finish_terms(p->h_terms, p->h_ops, &p->max_leaves, mdim);
  hipSetDevice(ctx->device);
  if (nmats > 0) {
    // the matrices, then their conjugate transposes (for the adjoint program)
    const size_t mm = (size_t)mdim * mdim;
    p->h_mats.resize(2 * mm * nmats);
    for (int k = 0; k < nmats; ++k)
      for (int i = 0; i < mdim; ++i)
        for (int j = 0; j < mdim; ++j) {
          const double* u = mats + 2 * (k * mm + (size_t)i * mdim + j);
          p->h_mats[k * mm + (size_t)i * mdim + j] = cmake(u[0], u[1]);
          p->h_mats[(nmats + k) * mm + (size_t)j * mdim + i] = cmake(u[0], -u[1]);
        }
    if (hipMalloc(&p->d_mats, sizeof(cplx) * p->h_mats.size()) != hipSuccess) {
      delete p;
      return qd_set_error(ctx, QD_E_OOM, "mode matrices");
    }
    hipMemcpyAsync(p->d_mats, p->h_mats.data(), sizeof(cplx) * p->h_mats.size(), hipMemcpyHostToDevice,
                   ctx->stream);
  }
  size_t tb = sizeof(DevTerm) * (nterms ? nterms : 1), ob = sizeof(DevOp) * (nops ? nops : 1);
  if (hipMalloc(&p->d_terms, tb) != hipSuccess || hipMalloc(&p->d_ops, ob) != hipSuccess) {
    hipFree(p->d_terms);
    hipFree(p->d_mats);
    delete p;
    return qd_set_error(ctx, QD_E_OOM, "term program upload");
  }
  if (nterms) hipMemcpyAsync(p->d_terms, p->h_terms.data(), tb, hipMemcpyHostToDevice, ctx->stream);
  if (nops) hipMemcpyAsync(p->d_ops, p->h_ops.data(), ob, hipMemcpyHostToDevice, ctx->stream);
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    hipFree(p->d_terms);
    hipFree(p->d_ops);
    hipFree(p->d_mats);
    delete p;
    return qd_set_error(ctx, QD_E_HIP, "term program upload: %s", hipGetErrorString(e));
  }
  *out = p;
  return QD_OK;
}

extern "C" int qd_termprog_create(qd_ctx* ctx, int nbits, const qd_term* terms, int nterms, const qd_op* ops,
                                  int nops, qd_termprog** out) {
  return termprog_create(ctx, nbits, 1, terms, nterms, ops, nops, nullptr, 0, out);
}

extern "C" int qd_termprog_create_modes(qd_ctx* ctx, int nbits, int mdim, const qd_term* terms, int nterms,
                                        const qd_op* ops, int nops, const double* mats, int nmats,
                                        qd_termprog** out) {
  if (mdim < 2) return qd_set_error(ctx, QD_E_ARG, "qd_termprog_create_modes: mdim %d", mdim);
  return termprog_create(ctx, nbits, mdim, terms, nterms, ops, nops, mats, nmats, out);
}

int qd_termprog_ensure_adjoint(qd_termprog* p) {
  if (p->d_adj_terms) return QD_OK;
  qd_ctx* ctx = p->ctx;
  std::vector<DevTerm> at(p->h_terms.size());
  std::vector<DevOp> ao(p->h_ops.size());
  int pos = 0;
  for (size_t k = 0; k < p->h_terms.size(); ++k) {
    const DevTerm& t = p->h_terms[k];
    DevTerm d = t;
    d.coeff = cconj(t.coeff);
    d.op_begin = pos;
    for (int i = t.n_ops - 1; i >= 0; --i) {
      DevOp o = p->h_ops[t.op_begin + i];
      if (o.kind == QD_OP_MODE) {
        o.a = o.b = p->nmats + o.a;  // its conjugate transpose
      } else if (o.kind == QD_OP_BIT) {
        DevOp a = o;
        for (int r = 0; r < 2; ++r)
          for (int c = 0; c < 2; ++c) a.U[r * 2 + c] = cconj(o.U[c * 2 + r]);
        o = a;
      } else {
        int tmp = o.a;
        o.a = o.b;
        o.b = tmp;
      }
      ao[pos++] = o;
    }
    at[k] = d;
  }
  int ml = 1;
  finish_terms(at, ao, &ml, p->mdim);
  size_t tb = sizeof(DevTerm) * (at.size() ? at.size() : 1), ob = sizeof(DevOp) * (ao.size() ? ao.size() : 1);
  if (hipMalloc(&p->d_adj_terms, tb) != hipSuccess || hipMalloc(&p->d_adj_ops, ob) != hipSuccess) {
    hipFree(p->d_adj_terms);
    p->d_adj_terms = nullptr;
    return qd_set_error(ctx, QD_E_OOM, "adjoint program upload");
  }
  if (!at.empty()) hipMemcpyAsync(p->d_adj_terms, at.data(), tb, hipMemcpyHostToDevice, ctx->stream);
  if (!ao.empty()) hipMemcpyAsync(p->d_adj_ops, ao.data(), ob, hipMemcpyHostToDevice, ctx->stream);
  QD_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return QD_OK;
}

extern "C" int qd_termprog_destroy(qd_termprog* p) {
  if (!p) return QD_OK;
  hipSetDevice(p->ctx->device);
  hipStreamSynchronize(p->ctx->stream);
  hipFree(p->d_terms);
  hipFree(p->d_ops);
  if (p->d_adj_terms) hipFree(p->d_adj_terms);
  if (p->d_adj_ops) hipFree(p->d_adj_ops);
  if (p->d_mats) hipFree(p->d_mats);
  delete p;
  return QD_OK;
}
