// Batched steady states: many problems of one hierarchy shape solved together.
//
// The reference solves one parameter point per process (a shell loop over
// inputs, tools/make_input_didv).  Here problems that share the hierarchy
// (same orbitals, leads, Pade terms, top tier) share every structural device
// table -- tiers, coupling entries, symmetry sectors, gather maps -- and only
// their coefficients are stacked: coefficient records, per-ADO rates gamma_K
// and the commutator values of their own H.  One launch then advances every
// problem: the packed apply and every TFQMR vector kernel carry the problem in
// grid dimension y, each problem keeps its own scalars and reductions, and a
// finished problem is frozen out of every kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <limits>
#include <string>
#include <complex>
#include <cstring>
#include <memory>
#include <vector>

#include "heomq_ctx.hpp"

using namespace heomq;

struct heomq_batch {
  int nprob = 0;
  heomq_model* shape = nullptr;   // problem 0's model (structure and shared host tables)
  heomq_ctx* ctx = nullptr;       // device tables of the shape
  DevBuf rec_all, gam_all, hq_union, hc_all, work, scal, partial, out, mask;
  DevPacked bf;                   // batched forward operator
  std::vector<std::vector<double>> dbsqrt;   // per problem: lscale factors (1 unscaled), for the observables
  int64_t n = 0;                  // reals per problem vector (2 * packed elements)
  hipStream_t stream = nullptr;
  hipGraphExec_t chunk = nullptr;
  void* pinned = nullptr;
  bool precond = false;           // Jacobi scaling of the iteration (every H diagonal)
  ~heomq_batch() {
    if (chunk) (void)hipGraphExecDestroy(chunk);
    if (stream) (void)hipStreamDestroy(stream);
    if (pinned) (void)hipHostFree(pinned);
    if (ctx) heomq_ctx_free(ctx);
    if (shape) heomq_model_free(shape);
  }
};

namespace {

// i(HX - XH) rows of every sector over the union of the problems' H patterns:
// one slot list for all problems (hq), one coefficient row per problem (hc).
int union_commutator(const heomq_ctx* c, const std::vector<std::vector<cplx>>& Hs, std::vector<int16_t>& hq,
                     std::vector<double>& hc, int& width, int& widthp) {
  const int d = c->d, dd = d * d, S = c->nsec, P = (int)Hs.size();
  std::vector<char> pat((size_t)dd, 0);
  for (const auto& H : Hs)
    for (int e = 0; e < dd; ++e) pat[e] |= H[e] != cplx(0, 0);
  std::vector<std::vector<int>> rows((size_t)S * dd);
  width = 1;
  for (int s = 0; s < S; ++s)
    for (int t = 0; t < c->sec_nnz_h[s]; ++t) {
      const int e = c->sec_elems_h[(size_t)s * dd + t];
      const int i = e % d, j = e / d;
      std::vector<int>& r = rows[(size_t)s * dd + t];
      for (int k = 0; k < d; ++k) {
        const int q1 = pat[i + k * d] ? c->sec_pos_h[(size_t)s * dd + k + j * d] : -1;   // H(i,k) X(k,j)
        const int q2 = pat[k + j * d] ? c->sec_pos_h[(size_t)s * dd + i + k * d] : -1;   // X(i,k) H(k,j)
        for (int q : {q1, q2})
          if (q >= 0 && std::find(r.begin(), r.end(), q) == r.end()) r.push_back(q);
      }
      width = std::max<int>(width, (int)r.size());
    }
  if (width > 16) return fail(HEOMQ_UNSUPPORTED, "batch: commutator rows wider than 16 slots");
  widthp = (width + 7) / 8 * 8;
  hq.assign((size_t)S * dd * widthp, 0);
  hc.assign((size_t)P * S * dd * width * 2, 0.0);
  for (int s = 0; s < S; ++s)
    for (int t = 0; t < c->sec_nnz_h[s]; ++t) {
      const std::vector<int>& r = rows[(size_t)s * dd + t];
      const int e = c->sec_elems_h[(size_t)s * dd + t];
      const int i = e % d, j = e / d;
      for (size_t u = 0; u < r.size(); ++u) hq[((size_t)s * dd + t) * widthp + u] = (int16_t)r[u];
      for (int p = 0; p < P; ++p) {
        const auto& H = Hs[p];
        for (size_t u = 0; u < r.size(); ++u) {
          const int q = r[u];
          const int eq = c->sec_elems_h[(size_t)s * dd + q];
          const int qi = eq % d, qj = eq / d;
          cplx v(0, 0);   // coefficient of X(qi, qj) in i(HX - XH)(i, j)
          if (qj == j) v += cplx(0, 1) * H[i + qi * d];
          if (qi == i) v -= cplx(0, 1) * H[qj + j * d];
          const size_t at = (((size_t)p * S + s) * dd + t) * width + u;
          hc[2 * at] = v.real();
          hc[2 * at + 1] = v.imag();
        }
      }
    }
  return HEOMQ_OK;
}

}  // namespace

extern "C" {

int heomq_batch_create(const heomq_params* params, int nprob, int device, heomq_batch** out) {
  if (!params || nprob < 1 || !out) return fail(HEOMQ_BAD_ARG, "null argument or no problems");
  auto B = std::make_unique<heomq_batch>();
  B->nprob = nprob;
  if (int rc = heomq_model_build(&params[0], &B->shape)) return rc;
  if (int rc = heomq_ctx_create(B->shape, device, &B->ctx)) return rc;
  heomq_ctx* c = B->ctx;
  const Model& m0 = B->shape->m;
  Model scratch = m0;
  std::vector<DevRec> rec_all, rec;
  std::vector<double> gam_all, gam;
  std::vector<std::vector<cplx>> Hs;
  for (int p = 0; p < nprob; ++p) {
    if (p > 0)
      if (int rc = set_params(scratch, params[p]))
        return fail(rc, "batch: problem " + std::to_string(p) + " differs in shape: " + last_error());
    build_records(scratch, rec);   // lscale: every problem's upper couplings carry its own |cb|^(1/2)
    build_gamma(scratch, scratch.gdraw, gam, false);
    B->dbsqrt.push_back(scratch.dbsqrt.empty() ? std::vector<double>(scratch.M, 1.0) : scratch.dbsqrt);
    rec_all.insert(rec_all.end(), rec.begin(), rec.end());
    gam_all.insert(gam_all.end(), gam.begin(), gam.end());
    for (int i = 0; i < c->d; ++i)   // the sector layout needs the conserved quantities of problem 0
      for (int k = 0; k < c->d; ++k)
        if (scratch.H[i + (size_t)k * c->d] != cplx(0, 0) &&
            ((c->cons_n && c->fock_n[i] != c->fock_n[k]) || (c->cons_s && c->fock_s[i] != c->fock_s[k])))
          return fail(HEOMQ_UNSUPPORTED, "batch: problem " + std::to_string(p) + " breaks N / S_z conservation");
    Hs.push_back(scratch.H);
  }
  std::vector<int16_t> hq;
  std::vector<double> hc;
  int width = 0, widthp = 0;
  if (int rc = union_commutator(c, Hs, hq, hc, width, widthp)) return rc;
  HIPCHECK(hipSetDevice(device));
  HIPCHECK(upload(B->rec_all, rec_all));
  HIPCHECK(upload(B->gam_all, gam_all));
  HIPCHECK(upload(B->hq_union, hq));
  HIPCHECK(upload(B->hc_all, hc));
  const int64_t nrec = (int64_t)rec.size();
  // the apply form (work granularity) timed on this shape and GPU
  if (int rc = heomq_ctx_autotune(c, 10, nullptr, nullptr)) return rc;
  B->bf = c->pfwd;
  B->bf.rec = B->rec_all.as<DevRec>();
  B->bf.gam = B->gam_all.as<double>();
  B->bf.hell_w = width;
  B->bf.hell_wp = widthp;
  B->bf.hq_ell = B->hq_union.as<int16_t>();
  B->bf.hc_ell = B->hc_all.as<double>();
  B->bf.vstride = c->ptotal;
  B->bf.rec_stride = nrec;
  B->bf.gam_stride = c->N;
  B->bf.hc_stride = (int64_t)c->nsec * c->d * c->d * width;
  B->bf.nprob = nprob;
  B->n = 2 * c->ptotal;
  // workspace: x, b, w, y1, y2, u1, u2, v, d, ax, xbest for every problem
  HIPCHECK(B->work.alloc(sizeof(double) * (size_t)B->n * nprob * 11));
  bool diag = true;
  for (const auto& H : Hs)
    for (int i = 0; i < c->d; ++i)
      for (int k = 0; k < c->d; ++k)
        if (i != k && H[i + (size_t)k * c->d] != cplx(0, 0)) diag = false;
  B->precond = diag;
  HIPCHECK(B->scal.alloc(sizeof(TfqmrScal) * (size_t)nprob));
  HIPCHECK(B->partial.alloc(sizeof(double) * 2 * kMaxRedBlocks * (size_t)nprob));
  HIPCHECK(B->out.alloc(sizeof(double) * 2 * (size_t)nprob));
  HIPCHECK(B->mask.alloc(2 * sizeof(int32_t) * (size_t)nprob));
  HIPCHECK(hipHostMalloc(&B->pinned, (sizeof(TfqmrScal) + 2 * sizeof(double) + 2 * sizeof(int32_t)) * (size_t)nprob + 256,
                         hipHostMallocDefault));
  HIPCHECK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
  B->bf.frozen_flag = reinterpret_cast<const int32_t*>(B->scal.as<char>() + offsetof(TfqmrScal, flag));
  B->bf.frozen_stride = (int)(sizeof(TfqmrScal) / sizeof(int32_t));
  *out = B.release();
  return HEOMQ_OK;
}

void heomq_batch_free(heomq_batch* b) { delete b; }

int heomq_batch_info(const heomq_batch* b, int64_t* info) {
  if (!b || !info) return fail(HEOMQ_BAD_ARG, "null argument");
  info[0] = b->nprob;
  info[1] = b->ctx->N;
  info[2] = b->ctx->ptotal;
  info[3] = b->ctx->d;
  info[4] = b->bf.hell_w;
  return HEOMQ_OK;
}

int heomq_batch_apply_packed(heomq_batch* b, const double* xp, double* yp, int trace, void* stream) {
  if (!b || !xp || !yp) return fail(HEOMQ_BAD_ARG, "null argument");
  if (trace != HEOMQ_TRACE_RDO && trace != HEOMQ_TRACE_NONE) return fail(HEOMQ_BAD_ARG, "packed apply: trace RDO or NONE");
  DevPacked op = b->bf;
  op.frozen_flag = nullptr;   // every problem
  HIPCHECK(launch_apply_packed_f64(op, xp, yp, trace == HEOMQ_TRACE_RDO, (hipStream_t)stream));
  return HEOMQ_OK;
}

// Steady states of every problem of the batch.  rdo: nprob * d * d complex
// (tier-0 block, column-major), current: nprob * nalf * nspin (td/calcj.f90),
// occ: nprob * norbs * nspin, stats: nprob * 4 (TFQMR iterations, relative
// residual, status, restarts).  Returns HEOMQ_OK when every problem converged.
int heomq_batch_solve(heomq_batch* B, double tol, int maxiter, double* rdo, double* current, double* occ,
                      double* stats, void* stream) {
  if (!B) return fail(HEOMQ_BAD_ARG, "null argument");
  heomq_ctx* c = B->ctx;
  const int P = B->nprob, d = c->d;
  const int64_t n = B->n;
  const Batch bt{P, n};
  if (stream) HIPCHECK(hipStreamSynchronize((hipStream_t)stream));
  hipStream_t st = B->stream;
  double* base = B->work.as<double>();
  double *x = base, *b = base + n * P, *w = base + 2 * n * P, *y1 = base + 3 * n * P, *y2 = base + 4 * n * P,
         *u1 = base + 5 * n * P, *u2 = base + 6 * n * P, *v = base + 7 * n * P, *dv = base + 8 * n * P,
         *ax = base + 9 * n * P, *xbest = base + 10 * n * P;
  TfqmrScal* s = B->scal.as<TfqmrScal>();
  double* part = B->partial.as<double>();
  double* outv = B->out.as<double>();
  int32_t* mask = B->mask.as<int32_t>();
  TfqmrScal* hs = static_cast<TfqmrScal*>(B->pinned);
  double* hout = reinterpret_cast<double*>(hs + P);
  int32_t* hmask = reinterpret_cast<int32_t*>(hout + 2 * P);
  int32_t* hmask2 = hmask + P;
  int32_t* mask2 = mask + P;
  DevPacked op_it = B->bf, op0 = B->bf;
  op_it.precond = B->precond ? 1 : 0;   // b = e_trace sits on the unscaled trace row
  op0.precond = 0;
  auto Ait = [&](const double* in, double* o) { return launch_apply_packed_f64(op_it, in, o, true, st); };
  auto A0 = [&](const double* in, double* o) { return launch_apply_packed_f64(op0, in, o, true, st); };

  // x0: trace-one identity / d on tier 0 (a Hermitian physical start); b = e_trace
  HIPCHECK(vec_fill<double>(st, n * P, x, 0.0));
  HIPCHECK(vec_fill<double>(st, n * P, b, 0.0));
  SetPattern px, pb;
  for (int i = 0; i < d; ++i) {
    const int q = c->sec_pos_h[(size_t)c->psec_h[0] * d * d + i + i * d];
    if (q >= 0 && px.m < 64) { px.idx[px.m] = 2 * q; px.val[px.m] = 1.0 / d; px.m++; }
  }
  pb.m = 1;
  pb.idx[0] = 2 * (int64_t)op_it.trace_slot;
  pb.val[0] = 1.0;
  HIPCHECK(vec_set_pattern(st, x, n, P, px));
  HIPCHECK(vec_set_pattern(st, b, n, P, pb));
  HIPCHECK(hipMemcpyAsync(xbest, x, sizeof(double) * n * P, hipMemcpyDeviceToDevice, st));

  auto init = [&](const int32_t* m) -> int {   // (re)start the masked problems from their current x
    HIPCHECK(tfq_reset(st, s, tol, m, P, m == nullptr));   // ||b|| = 1: the stop target is tol
    if (B->precond) {   // true ||r||^2 of every problem, for the scaled-target ratio
      HIPCHECK(A0(x, ax));
      HIPCHECK((vec_resid<double, double>(st, n, b, ax, outv, part, bt)));
    }
    HIPCHECK(Ait(x, ax));
    HIPCHECK(tfq_init_r<double>(st, n, b, ax, w, y1, dv, s, part, bt));
    if (B->precond) HIPCHECK(tfq_scale_targets(st, s, outv, P));
    HIPCHECK(Ait(y1, ax));
    HIPCHECK(tfq_init_v<double>(st, n, ax, v, s, part, bt, u1));
    return HEOMQ_OK;
  };
  HIPCHECK(tfq_set_unscaled(st, s, nullptr, P));   // every problem starts with the scaled iteration
  if (int rc = init(nullptr)) return rc;

  const int check_every = 8, verify_every = 32;
  if (!B->chunk) {
    HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t rec = hipSuccess;
    for (int k = 0; k < check_every && rec == hipSuccess; ++k) {
      if ((rec = tfq_step1<double>(st, n, y1, v, y2, w, u1, dv, s, part, bt))) break;
      if ((rec = Ait(y2, u2))) break;
      if ((rec = tfq_step2<double>(st, n, x, w, u2, dv, y2, s, part, bt))) break;
      if ((rec = tfq_step3<double>(st, n, x, dv, w, y2, y1, s, bt))) break;
      if ((rec = Ait(y1, u1))) break;
      rec = tfq_v<double>(st, n, u1, u2, v, s, part, bt);
    }
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(st, &g);
    if (rec != hipSuccess || end != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(HEOMQ_HIP_ERROR, "batch: TFQMR chunk capture failed");
    }
    const hipError_t inst = hipGraphInstantiate(&B->chunk, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (inst != hipSuccess) return fail(HEOMQ_HIP_ERROR, "batch: TFQMR chunk instantiate failed");
  }

  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int> status(P, 0), restarts(P, 0), total(P, 0), stalls(P, 0), unscaled(P, 0);
  std::vector<double> best(P, std::numeric_limits<double>::infinity()), relres(P, 0.0);
  int active = P;
  for (int replay = 1; active > 0; ++replay) {
    HIPCHECK(hipGraphLaunch(B->chunk, st));
    HIPCHECK(hipMemcpyAsync(hs, s, sizeof(TfqmrScal) * P, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const bool verify = replay % verify_every == 0;
    bool check = verify;
    for (int p = 0; p < P; ++p)
      if (status[p] == 0 && (hs[p].flag == 1 || hs[p].flag == 2 || total[p] + hs[p].iter >= maxiter)) check = true;
    if (!check) continue;
    // true residuals of every unfrozen problem (unscaled operator)
    HIPCHECK(A0(x, ax));
    HIPCHECK((vec_resid<double, double>(st, n, b, ax, outv, part, bt)));
    HIPCHECK(hipMemcpyAsync(hout, outv, sizeof(double) * 2 * P, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    std::vector<int32_t> fr(P, 0), rs(P, 0);
    bool any_freeze = false, any_restart = false;
    for (int p = 0; p < P; ++p) {
      if (status[p] != 0) continue;
      const bool stopped = hs[p].flag == 1 || hs[p].flag == 2;
      const bool out_of_iters = total[p] + hs[p].iter >= maxiter;
      const double rr = hout[2 * p], bb = hout[2 * p + 1];
      const double rel = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
      bool revert = false, restart = false, freeze = false;
      const double prev = best[p];
      if (rr <= tol * tol * bb) {
        status[p] = 1;
        relres[p] = rel;
        freeze = true;
      } else {
        const bool improved = rr < prev;
        if (improved) {
          best[p] = rr;
          relres[p] = rel;
          HIPCHECK(hipMemcpyAsync(xbest + n * p, x + n * p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        }
        revert = rr > 100.0 * prev;   // the recurrences diverged: back to the best iterate
        restart = stopped || revert;
        bool give_up = (restart && !improved && ++stalls[p] > 3) || out_of_iters;
        if (give_up && !out_of_iters && B->precond && !unscaled[p]) {
          unscaled[p] = 1;   // the scaled iteration stalled: continue unscaled from the best iterate
          stalls[p] = 0;
          revert = rr > best[p];
          restart = true;
          give_up = false;
        }
        if (give_up) {
          status[p] = hs[p].flag == 2 ? 2 : 3;
          revert = rr > best[p];
          restart = false;
          freeze = true;
        }
      }
      if (revert) {
        HIPCHECK(hipMemcpyAsync(x + n * p, xbest + n * p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        relres[p] = std::sqrt(best[p] / (bb > 0 ? bb : 1.0));
      }
      if (stopped || restart || freeze) total[p] += hs[p].iter;
      if (freeze) {
        fr[p] = 1;
        any_freeze = true;
        active--;
      } else if (restart) {
        restarts[p]++;
        rs[p] = 1;
        any_restart = true;
      }
    }
    if (any_freeze) {
      std::memcpy(hmask, fr.data(), sizeof(int32_t) * P);
      HIPCHECK(hipMemcpyAsync(mask, hmask, sizeof(int32_t) * P, hipMemcpyHostToDevice, st));
      HIPCHECK(tfq_freeze(st, s, mask, P));
      HIPCHECK(hipStreamSynchronize(st));
    }
    if (any_restart) {
      for (int p = 0; p < P; ++p) {
        if (!rs[p]) continue;
        HIPCHECK(launch_herm_packed<double>(c->pfwd, x + n * p, st));
      }
      for (int p = 0; p < P; ++p) hmask2[p] = unscaled[p];   // the apply reads this word of each problem's scalars
      HIPCHECK(hipMemcpyAsync(mask2, hmask2, sizeof(int32_t) * P, hipMemcpyHostToDevice, st));
      HIPCHECK(tfq_set_unscaled(st, s, mask2, P));
      std::memcpy(hmask, rs.data(), sizeof(int32_t) * P);
      HIPCHECK(hipMemcpyAsync(mask, hmask, sizeof(int32_t) * P, hipMemcpyHostToDevice, st));
      if (int rc = init(mask)) return rc;
      HIPCHECK(hipStreamSynchronize(st));
    }
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  // observables from each problem's tier-0 and tier-1 blocks
  const int64_t n01 = c->L >= 1 ? c->first[2] : 1;
  const int64_t p01 = c->poff_h[n01];
  std::vector<double> hp(2 * (size_t)p01), dense(2 * (size_t)n01 * d * d);
  const int nal = c->nalf * c->nspin, noc = c->norbs * c->nspin;
  for (int p = 0; p < P; ++p) {
    HIPCHECK(hipMemcpy(hp.data(), x + n * p, sizeof(double) * hp.size(), hipMemcpyDeviceToHost));
    std::fill(dense.begin(), dense.end(), 0.0);
    for (int64_t a = 0; a < n01; ++a) {
      const int sct = c->psec_h[a];
      for (int64_t t = 0; t < c->poff_h[a + 1] - c->poff_h[a]; ++t) {
        const int e = c->sec_elems_h[(size_t)sct * d * d + t];
        dense[2 * (a * d * d + e)] = hp[2 * (c->poff_h[a] + t)];
        dense[2 * (a * d * d + e) + 1] = hp[2 * (c->poff_h[a] + t) + 1];
      }
    }
    if (rdo) std::memcpy(rdo + 2 * (size_t)p * d * d, dense.data(), sizeof(double) * 2 * d * d);
    std::vector<double> cur(nal), oc(noc);
    c->dbsqrt = B->dbsqrt[p];   // the currents read tier 1, stored scaled with lscale
    observables_host(c, dense, cur.data(), oc.data());
    if (current) std::memcpy(current + (size_t)p * nal, cur.data(), sizeof(double) * nal);
    if (occ) std::memcpy(occ + (size_t)p * noc, oc.data(), sizeof(double) * noc);
    if (stats) {
      stats[4 * p] = total[p];
      stats[4 * p + 1] = relres[p];
      stats[4 * p + 2] = status[p] == 1 ? HEOMQ_OK : (status[p] == 2 ? HEOMQ_BREAKDOWN : HEOMQ_NOT_CONVERGED);
      stats[4 * p + 3] = restarts[p];
    }
  }
  (void)secs;
  for (int p = 0; p < P; ++p)
    if (status[p] != 1) return fail(HEOMQ_NOT_CONVERGED, "batch: problem " + std::to_string(p) + " did not converge");
  return HEOMQ_OK;
}

}  // extern "C"
