// Batched steady states: many problems solved together on one GPU.
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
//
// A sweep goes one step further: problems of different hierarchy shapes
// (Pade terms, top tier) are grouped into shapes, and every solver step is
// still one launch, over a flat grid that covers every problem of every shape
// (largest shape first).  Both run the same host loop below.
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
  DevBuf scal2, partial2;         // TFQMR's mid-iteration scalar copies and second partial sums
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

constexpr int kTfqVecs = 10;   // x, b and the TFQMR vectors of one solve, plus A x and the best iterate

struct heomq_sweep {
  int nprob = 0, ng = 0;
  std::vector<heomq_batch*> shapes;   // one per shape, largest first (structure and coefficients only)
  std::vector<int> perm;              // global problem q (shape-major) -> caller's problem index
  std::vector<int> shape_of;          // caller's problem -> shape
  std::vector<int64_t> qoff;          // global problem q -> first real of its vectors
  Multi m;                            // vector-kernel layout of every problem
  std::vector<DevPacked> hops_it, hops0;   // per shape: iteration (scaled where the shape is) and plain operators
  std::vector<ApplyMulti> am_it, am0;      // the apply launches (kApplyShapesPerLaunch shapes each)
  int64_t apply_blocks = 0;
  // the launches of a solve cover only its unfinished problems: the layouts
  // below, rebuilt (with the step graph) when a quarter of them has finished
  Multi cur;
  std::vector<ApplyMulti> cur_it, cur0;
  int covered = 0;
  DevBuf pmap;                        // per shape: in-shape indices of the covered problems
  DevBuf work, scal, partial, out, mask;
  DevBuf scal2, partial2;
  int64_t nt = 0;                     // reals of one vector over every problem
  hipStream_t stream = nullptr;
  hipGraphExec_t chunk = nullptr;
  void* pinned = nullptr;
  ~heomq_sweep() {
    if (chunk) (void)hipGraphExecDestroy(chunk);
    if (stream) (void)hipStreamDestroy(stream);
    if (pinned) (void)hipHostFree(pinned);
    for (heomq_batch* b : shapes) delete b;
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

namespace {

// Structure, coefficients and operator of a batch; the solver workspace only
// when `workspace` (a sweep keeps one workspace for all of its shapes).
int batch_build(const heomq_params* params, int nprob, int device, bool workspace, heomq_batch** out) {
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
  bool diag = true;
  for (const auto& H : Hs)
    for (int i = 0; i < c->d; ++i)
      for (int k = 0; k < c->d; ++k)
        if (i != k && H[i + (size_t)k * c->d] != cplx(0, 0)) diag = false;
  B->precond = diag;
  if (workspace) {
    // x, b, w, y2, u2, aw, v, d, ax, xbest for every problem
    HIPCHECK(B->work.alloc(sizeof(double) * (size_t)B->n * nprob * kTfqVecs));
    HIPCHECK(B->scal.alloc(sizeof(TfqmrScal) * (size_t)nprob));
    HIPCHECK(B->partial.alloc(sizeof(double) * 2 * kMaxRedBlocks * (size_t)nprob));
    HIPCHECK(B->scal2.alloc(sizeof(TfqmrScal) * (size_t)nprob));
    HIPCHECK(B->partial2.alloc(sizeof(double) * 2 * kMaxRedBlocks * (size_t)nprob));
    HIPCHECK(B->out.alloc(sizeof(double) * 2 * (size_t)nprob));
    HIPCHECK(B->mask.alloc(2 * sizeof(int32_t) * (size_t)nprob));
    HIPCHECK(hipHostMalloc(&B->pinned,
                           (sizeof(TfqmrScal) + 2 * sizeof(double) + 2 * sizeof(int32_t)) * (size_t)nprob + 256,
                           hipHostMallocDefault));
    HIPCHECK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    B->bf.frozen_flag = reinterpret_cast<const int32_t*>(B->scal.as<char>() + offsetof(TfqmrScal, flag));
    B->bf.frozen_stride = (int)(sizeof(TfqmrScal) / sizeof(int32_t));
  }
  *out = B.release();
  return HEOMQ_OK;
}

// Device vectors, scalars and host mirrors of one solve over P problems.
struct TfqBufs {
  double *x, *b, *w, *y2, *u2, *aw, *v, *dv, *ax, *xbest;
  TfqmrScal *s, *s2;
  double *part, *part2, *outv;
  int32_t *mask, *mask2;
  TfqmrScal* hs;   // pinned host copies
  double* hout;
  int32_t *hmask, *hmask2;
  int64_t nt;      // reals of one vector over every problem
};

TfqBufs carve_bufs(double* base, int64_t nt, TfqmrScal* s, TfqmrScal* s2, double* part, double* part2, double* outv,
                   int32_t* mask, void* pinned, int P) {
  TfqBufs V;
  double** vec[kTfqVecs] = {&V.x, &V.b, &V.w, &V.y2, &V.u2, &V.aw, &V.v, &V.dv, &V.ax, &V.xbest};
  for (int k = 0; k < kTfqVecs; ++k) *vec[k] = base + (int64_t)k * nt;
  V.s = s;
  V.s2 = s2;
  V.part = part;
  V.part2 = part2;
  V.outv = outv;
  V.mask = mask;
  V.mask2 = mask + P;
  V.hs = static_cast<TfqmrScal*>(pinned);
  V.hout = reinterpret_cast<double*>(V.hs + P);
  V.hmask = reinterpret_cast<int32_t*>(V.hout + 2 * P);
  V.hmask2 = V.hmask + P;
  V.nt = nt;
  return V;
}

// Per-problem results of a solve (in the layout's problem order).
struct SolveOut {
  std::vector<int> status, restarts, total;
  std::vector<double> relres;
};

// The TFQMR loop over every problem of a layout.  Lay supplies the problem
// count and vector placement (nprob, off, len), its operators (Ait: the
// iteration's, A0: the plain one, both with the trace row), the fused vector
// steps over all problems, the start vectors and the tier-0
// Hermitianisation of one problem.  Every problem starts from the trace-one
// state; the scaled iteration of a problem that stalls continues unscaled
// from its best iterate; a problem whose true residual meets tol, or that
// runs out of iterations, is frozen out of every kernel.
template <class Lay>
int tfqmr_loop(Lay& L, const TfqBufs& V, hipGraphExec_t& chunk, hipStream_t st, double tol, int maxiter,
               SolveOut& so) {
  const int P = L.nprob();
  HIPCHECK(vec_fill<double>(st, V.nt, V.x, 0.0));
  HIPCHECK(vec_fill<double>(st, V.nt, V.b, 0.0));
  HIPCHECK(L.start(V.x, V.b));
  HIPCHECK(hipMemcpyAsync(V.xbest, V.x, sizeof(double) * V.nt, hipMemcpyDeviceToDevice, st));

  auto init = [&](const int32_t* msk) -> int {   // (re)start the masked problems from their current x
    HIPCHECK(tfq_reset(st, V.s, tol, msk, P, msk == nullptr));   // ||b|| = 1: the stop target is tol
    if (L.any_scaled()) {   // true ||r||^2 of every problem, for the scaled-target ratio
      HIPCHECK(L.A0(V.x, V.ax));
      HIPCHECK(L.resid(V.b, V.ax, V.outv, V.part));
    }
    HIPCHECK(L.Ait(V.x, V.ax));
    HIPCHECK(L.init_r(V.b, V.ax, V.w, V.y2, V.dv, V.s, V.part));
    if (L.any_scaled()) HIPCHECK(tfq_scale_targets(st, V.s, V.outv, P));
    HIPCHECK(L.Ait(V.w, V.ax));   // into scratch: the problems that keep running keep their aw
    HIPCHECK(L.init_v(V.ax, V.v, V.aw, V.u2, V.s, V.part));
    return HEOMQ_OK;
  };
  HIPCHECK(tfq_set_unscaled(st, V.s, nullptr, P));   // every problem starts with the scaled iteration
  if (int rc = init(nullptr)) return rc;

  const int check_every = 8, verify_every = 32;
  auto capture = [&]() -> int {   // the graph of check_every iterations over the layout's problems
    HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t rec = hipSuccess;
    for (int k = 0; k < check_every && rec == hipSuccess; ++k) {
      if ((rec = L.step1(V.w, V.y2, V.v, V.aw, V.u2, V.dv, V.s, V.part))) break;
      if ((rec = L.Ait(V.y2, V.u2))) break;
      if ((rec = L.step2(V.x, V.w, V.u2, V.dv, V.y2, V.s, V.s2, V.part, V.part2))) break;
      if ((rec = L.Ait(V.w, V.aw))) break;
      rec = L.vstep(V.aw, V.u2, V.v, V.x, V.dv, V.s, V.s2, V.part, V.part2);
    }
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(st, &g);
    if (rec != hipSuccess || end != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(HEOMQ_HIP_ERROR, "batch: TFQMR chunk capture failed");
    }
    const hipError_t inst = hipGraphInstantiate(&chunk, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (inst != hipSuccess) {
      chunk = nullptr;
      return fail(HEOMQ_HIP_ERROR, "batch: TFQMR chunk instantiate failed");
    }
    return HEOMQ_OK;
  };

  so.status.assign(P, 0);
  so.restarts.assign(P, 0);
  so.total.assign(P, 0);
  so.relres.assign(P, 0.0);
  std::vector<int> stalls(P, 0), unscaled(P, 0);
  std::vector<double> best(P, std::numeric_limits<double>::infinity());
  TfqmrScal* hs = V.hs;
  double* hout = V.hout;
  int active = P;
  for (int replay = 1; active > 0; ++replay) {
    if (!chunk)
      if (int rc = capture()) return rc;
    HIPCHECK(hipGraphLaunch(chunk, st));
    HIPCHECK(hipMemcpyAsync(hs, V.s, sizeof(TfqmrScal) * P, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const bool verify = replay % verify_every == 0;
    bool check = verify;
    for (int p = 0; p < P; ++p)
      if (so.status[p] == 0 && (hs[p].flag == 1 || hs[p].flag == 2 || so.total[p] + hs[p].iter >= maxiter))
        check = true;
    if (!check) continue;
    // true residuals of every unfrozen problem (unscaled operator)
    HIPCHECK(L.A0(V.x, V.ax));
    HIPCHECK(L.resid(V.b, V.ax, V.outv, V.part));
    HIPCHECK(hipMemcpyAsync(hout, V.outv, sizeof(double) * 2 * P, hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    std::vector<int32_t> fr(P, 0), rs(P, 0);
    bool any_freeze = false, any_restart = false;
    for (int p = 0; p < P; ++p) {
      if (so.status[p] != 0) continue;
      const int64_t o = L.off(p), n = L.len(p);
      const bool stopped = hs[p].flag == 1 || hs[p].flag == 2;
      const bool out_of_iters = so.total[p] + hs[p].iter >= maxiter;
      const double rr = hout[2 * p], bb = hout[2 * p + 1];
      const double rel = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
      bool revert = false, restart = false, freeze = false;
      const double prev = best[p];
      if (rr <= tol * tol * bb) {
        so.status[p] = 1;
        so.relres[p] = rel;
        freeze = true;
      } else {
        const bool improved = rr < prev;
        if (improved) {
          best[p] = rr;
          so.relres[p] = rel;
          HIPCHECK(hipMemcpyAsync(V.xbest + o, V.x + o, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        }
        revert = rr > 100.0 * prev;   // the recurrences diverged: back to the best iterate
        restart = stopped || revert;
        bool give_up = (restart && !improved && ++stalls[p] > 3) || out_of_iters;
        if (give_up && !out_of_iters && L.scaled(p) && !unscaled[p]) {
          unscaled[p] = 1;   // the scaled iteration stalled: continue unscaled from the best iterate
          stalls[p] = 0;
          revert = rr > best[p];
          restart = true;
          give_up = false;
        }
        if (give_up) {
          so.status[p] = hs[p].flag == 2 ? 2 : 3;
          revert = rr > best[p];
          restart = false;
          freeze = true;
        }
      }
      if (revert) {
        HIPCHECK(hipMemcpyAsync(V.x + o, V.xbest + o, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        so.relres[p] = std::sqrt(best[p] / (bb > 0 ? bb : 1.0));
      }
      if (stopped || restart || freeze) so.total[p] += hs[p].iter;
      if (freeze) {
        fr[p] = 1;
        any_freeze = true;
        active--;
      } else if (restart) {
        so.restarts[p]++;
        rs[p] = 1;
        any_restart = true;
      }
    }
    if (any_freeze) {
      std::memcpy(V.hmask, fr.data(), sizeof(int32_t) * P);
      HIPCHECK(hipMemcpyAsync(V.mask, V.hmask, sizeof(int32_t) * P, hipMemcpyHostToDevice, st));
      HIPCHECK(tfq_freeze(st, V.s, V.mask, P));
      HIPCHECK(hipStreamSynchronize(st));
      bool relaid = false;
      HIPCHECK(L.compact(so.status, active, relaid));
      if (relaid) {   // the step graph was captured over the previous layout
        HIPCHECK(hipGraphExecDestroy(chunk));
        chunk = nullptr;
      }
    }
    if (any_restart) {
      for (int p = 0; p < P; ++p)
        if (rs[p]) HIPCHECK(L.herm(p, V.x + L.off(p)));
      for (int p = 0; p < P; ++p) V.hmask2[p] = unscaled[p];   // the apply reads this word of each problem's scalars
      HIPCHECK(hipMemcpyAsync(V.mask2, V.hmask2, sizeof(int32_t) * P, hipMemcpyHostToDevice, st));
      HIPCHECK(tfq_set_unscaled(st, V.s, V.mask2, P));
      std::memcpy(V.hmask, rs.data(), sizeof(int32_t) * P);
      HIPCHECK(hipMemcpyAsync(V.mask, V.hmask, sizeof(int32_t) * P, hipMemcpyHostToDevice, st));
      if (int rc = init(V.mask)) return rc;
      HIPCHECK(hipStreamSynchronize(st));
    }
  }
  return HEOMQ_OK;
}

// Start vectors of one shape's problems: x0 = identity / d on tier 0 (a
// Hermitian trace-one state), b = e_trace.
hipError_t shape_start(const heomq_ctx* c, const DevPacked& op, double* x, double* b, int64_t n, int nprob,
                       hipStream_t st) {
  const int d = c->d;
  SetPattern px, pb;
  for (int i = 0; i < d; ++i) {
    const int q = c->sec_pos_h[(size_t)c->psec_h[0] * d * d + i + i * d];
    if (q >= 0 && px.m < 64) { px.idx[px.m] = 2 * q; px.val[px.m] = 1.0 / d; px.m++; }
  }
  pb.m = 1;
  pb.idx[0] = 2 * (int64_t)op.trace_slot;
  pb.val[0] = 1.0;
  if (hipError_t e = vec_set_pattern(st, x, n, nprob, px)) return e;
  return vec_set_pattern(st, b, n, nprob, pb);
}

// Observables of one problem from its packed state (tier-0 and tier-1 blocks).
int problem_observables(heomq_ctx* c, const std::vector<double>& dbsqrt, const double* xq, double* rdo,
                        double* current, double* occ) {
  const int d = c->d;
  const int64_t n01 = c->L >= 1 ? c->first[2] : 1;
  const int64_t p01 = c->poff_h[n01];
  std::vector<double> hp(2 * (size_t)p01), dense(2 * (size_t)n01 * d * d, 0.0);
  HIPCHECK(hipMemcpy(hp.data(), xq, sizeof(double) * hp.size(), hipMemcpyDeviceToHost));
  for (int64_t a = 0; a < n01; ++a) {
    const int sct = c->psec_h[a];
    for (int64_t t = 0; t < c->poff_h[a + 1] - c->poff_h[a]; ++t) {
      const int e = c->sec_elems_h[(size_t)sct * d * d + t];
      dense[2 * (a * d * d + e)] = hp[2 * (c->poff_h[a] + t)];
      dense[2 * (a * d * d + e) + 1] = hp[2 * (c->poff_h[a] + t) + 1];
    }
  }
  if (rdo) std::memcpy(rdo, dense.data(), sizeof(double) * 2 * d * d);
  std::vector<double> cur(c->nalf * c->nspin), oc(c->norbs * c->nspin);
  c->dbsqrt = dbsqrt;   // the currents read tier 1, stored scaled with lscale
  observables_host(c, dense, cur.data(), oc.data());
  if (current) std::memcpy(current, cur.data(), sizeof(double) * cur.size());
  if (occ) std::memcpy(occ, oc.data(), sizeof(double) * oc.size());
  return HEOMQ_OK;
}

void write_stats(const SolveOut& so, int p, double* st4) {
  st4[0] = so.total[p];
  st4[1] = so.relres[p];
  st4[2] = so.status[p] == 1 ? HEOMQ_OK : (so.status[p] == 2 ? HEOMQ_BREAKDOWN : HEOMQ_NOT_CONVERGED);
  st4[3] = so.restarts[p];
}

// One shape: grid y = problem.
struct BatchLayout {
  heomq_batch* B;
  hipStream_t st;
  DevPacked op_it, op0;
  Batch bt;
  int nprob() const { return B->nprob; }
  int64_t off(int p) const { return B->n * p; }
  int64_t len(int) const { return B->n; }
  bool scaled(int) const { return B->precond; }
  bool any_scaled() const { return B->precond; }
  hipError_t Ait(const double* in, double* o) { return launch_apply_packed_f64(op_it, in, o, true, st); }
  hipError_t A0(const double* in, double* o) { return launch_apply_packed_f64(op0, in, o, true, st); }
  hipError_t resid(const double* b, const double* ax, double* out, double* part) {
    return vec_resid<double, double>(st, B->n, b, ax, out, part, bt);
  }
  hipError_t init_r(const double* b, const double* ax, double* w, double* y2, double* d, TfqmrScal* s, double* part) {
    return tfq_init_r<double>(st, B->n, b, ax, w, y2, d, s, part, bt);
  }
  hipError_t init_v(const double* ax, double* v, double* aw, double* u2, TfqmrScal* s, double* part) {
    return tfq_init_v<double>(st, B->n, ax, v, aw, u2, s, part, bt);
  }
  hipError_t step1(double* w, double* y2, const double* v, const double* aw, const double* u2, double* d,
                   TfqmrScal* s, double* part) {
    return tfq_step1<double>(st, B->n, w, y2, v, aw, u2, d, s, part, bt);
  }
  hipError_t step2(double* x, double* w, const double* u2, double* d, const double* y2, TfqmrScal* s,
                   TfqmrScal* s2, const double* part, double* part2) {
    return tfq_step2<double>(st, B->n, x, w, u2, d, y2, s, s2, part, part2, bt);
  }
  hipError_t vstep(const double* aw, const double* u2, double* v, double* x, const double* d, TfqmrScal* s,
                   const TfqmrScal* s2, double* part, const double* part2) {
    return tfq_v<double>(st, B->n, aw, u2, v, x, d, s, s2, part, part2, bt);
  }
  hipError_t start(double* x, double* b) { return shape_start(B->ctx, op_it, x, b, B->n, B->nprob, st); }
  hipError_t herm(int, double* xq) { return launch_herm_packed<double>(B->ctx->pfwd, xq, st); }
  hipError_t compact(const std::vector<int>&, int, bool& relaid) {
    relaid = false;
    return hipSuccess;
  }
};

// Every shape of a sweep in one flat grid per step.
struct SweepLayout {
  heomq_sweep* W;
  hipStream_t st;
  int nprob() const { return W->nprob; }
  int shape(int q) const {
    int g = 0;
    while (g + 1 < W->ng && W->m.qstart[g + 1] <= q) ++g;
    return g;
  }
  int64_t off(int q) const { return W->qoff[q]; }
  int64_t len(int q) const { return 2 * W->m.n2[shape(q)]; }
  bool scaled(int q) const { return W->shapes[shape(q)]->precond; }
  bool any_scaled() const {
    for (heomq_batch* b : W->shapes)
      if (b->precond) return true;
    return false;
  }
  hipError_t Ait(const double* in, double* o) { return apply(W->cur_it, in, o); }
  hipError_t A0(const double* in, double* o) { return apply(W->cur0, in, o); }
  hipError_t apply(const std::vector<ApplyMulti>& parts, const double* in, double* o) {
    for (const ApplyMulti& am : parts)
      if (am.bstart[am.ng] > 0)   // a launch whose shapes have all finished is left out
        if (hipError_t e = launch_apply_packed_multi_f64(am, in, o, true, st)) return e;
    return hipSuccess;
  }
  hipError_t resid(const double* b, const double* ax, double* out, double* part) {
    return vec_resid(st, W->cur, b, ax, out, part);
  }
  hipError_t init_r(const double* b, const double* ax, double* w, double* y2, double* d, TfqmrScal* s, double* part) {
    return tfq_init_r(st, W->cur, b, ax, w, y2, d, s, part);
  }
  hipError_t init_v(const double* ax, double* v, double* aw, double* u2, TfqmrScal* s, double* part) {
    return tfq_init_v(st, W->cur, ax, v, aw, u2, s, part);
  }
  hipError_t step1(double* w, double* y2, const double* v, const double* aw, const double* u2, double* d,
                   TfqmrScal* s, double* part) {
    return tfq_step1(st, W->cur, w, y2, v, aw, u2, d, s, part);
  }
  hipError_t step2(double* x, double* w, const double* u2, double* d, const double* y2, TfqmrScal* s,
                   TfqmrScal* s2, const double* part, double* part2) {
    return tfq_step2(st, W->cur, x, w, u2, d, y2, s, s2, part, part2);
  }
  hipError_t vstep(const double* aw, const double* u2, double* v, double* x, const double* d, TfqmrScal* s,
                   const TfqmrScal* s2, double* part, const double* part2) {
    return tfq_v(st, W->cur, aw, u2, v, x, d, s, s2, part, part2);
  }
  hipError_t start(double* x, double* b) {
    for (int g = 0; g < W->ng; ++g) {
      heomq_batch* B = W->shapes[g];
      const int64_t o = W->m.voff[g];
      if (hipError_t e = shape_start(B->ctx, W->hops_it[g], x + o, b + o, B->n, B->nprob, st)) return e;
    }
    return hipSuccess;
  }
  hipError_t herm(int q, double* xq) { return launch_herm_packed<double>(W->shapes[shape(q)]->ctx->pfwd, xq, st); }
  // Drop finished problems from every launch once a quarter of the covered
  // ones has finished: per shape, the in-shape indices of the unfinished
  // ones (status 0) become the launches' problem maps.
  hipError_t compact(const std::vector<int>& status, int active, bool& relaid) {
    relaid = false;
    if (active <= 0 || 4 * active > 3 * W->covered) return hipSuccess;
    const Multi& full = W->m;
    std::vector<int32_t> list;
    std::vector<int32_t> start(W->ng + 1, 0), cnt(W->ng, 0);
    for (int g = 0; g < W->ng; ++g) {
      start[g] = (int32_t)list.size();
      for (int p = 0; p < full.qstart[g + 1] - full.qstart[g]; ++p)
        if (status[full.qstart[g] + p] == 0) list.push_back(p);
      cnt[g] = (int32_t)list.size() - start[g];
    }
    const int32_t* dev = W->pmap.as<int32_t>();
    if (hipError_t e = hipMemcpyAsync(W->pmap.as<int32_t>(), list.data(), sizeof(int32_t) * list.size(),
                                      hipMemcpyHostToDevice, st))
      return e;
    if (hipError_t e = hipStreamSynchronize(st)) return e;   // list is pageable host memory
    Multi c = full;
    int32_t vb = 0, fq = 0;
    for (int g = 0; g < W->ng; ++g) {
      c.bstart[g] = vb;
      c.qstart[g] = fq;
      c.q0[g] = full.qstart[g];
      c.pmap[g] = dev + start[g];
      vb += cnt[g] * full.nb[g];
      fq += cnt[g];
    }
    c.bstart[W->ng] = vb;
    c.qstart[W->ng] = fq;
    W->cur = c;
    for (size_t k = 0; k < W->am_it.size(); ++k) {
      ApplyMulti& ai = W->cur_it[k];
      ApplyMulti& a0 = W->cur0[k];
      int32_t ab = 0;
      for (int j = 0; j < ai.ng; ++j) {
        const int g = (int)k * kApplyShapesPerLaunch + j;
        ai.bstart[j] = a0.bstart[j] = ab;
        ai.pmap[j] = a0.pmap[j] = dev + start[g];
        ab += cnt[g] * ai.nb[j];
      }
      ai.bstart[ai.ng] = a0.bstart[a0.ng] = ab;
    }
    W->covered = active;
    relaid = true;
    return hipSuccess;
  }
};

}  // namespace

extern "C" {

int heomq_batch_create(const heomq_params* params, int nprob, int device, heomq_batch** out) {
  return batch_build(params, nprob, device, true, out);
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
  if (!B->stream) return fail(HEOMQ_BAD_ARG, "batch: no solver workspace");
  heomq_ctx* c = B->ctx;
  const int P = B->nprob, d = c->d;
  if (stream) HIPCHECK(hipStreamSynchronize((hipStream_t)stream));
  BatchLayout L{B, B->stream, B->bf, B->bf, Batch{P, B->n}};
  L.op_it.precond = B->precond ? 1 : 0;   // b = e_trace sits on the unscaled trace row
  L.op0.precond = 0;
  const TfqBufs V = carve_bufs(B->work.as<double>(), B->n * P, B->scal.as<TfqmrScal>(), B->scal2.as<TfqmrScal>(),
                               B->partial.as<double>(), B->partial2.as<double>(), B->out.as<double>(),
                               B->mask.as<int32_t>(), B->pinned, P);
  SolveOut so;
  if (int rc = tfqmr_loop(L, V, B->chunk, B->stream, tol, maxiter, so)) return rc;
  const int nal = c->nalf * c->nspin, noc = c->norbs * c->nspin;
  for (int p = 0; p < P; ++p) {
    if (int rc = problem_observables(c, B->dbsqrt[p], V.x + B->n * p, rdo ? rdo + 2 * (size_t)p * d * d : nullptr,
                                     current ? current + (size_t)p * nal : nullptr,
                                     occ ? occ + (size_t)p * noc : nullptr))
      return rc;
    if (stats) write_stats(so, p, stats + 4 * p);
  }
  for (int p = 0; p < P; ++p)
    if (so.status[p] != 1) return fail(HEOMQ_NOT_CONVERGED, "batch: problem " + std::to_string(p) + " did not converge");
  return HEOMQ_OK;
}

// ---------------------------------------------------------------------------
// Sweeps: problems of several hierarchy shapes, one launch per solver step.
// ---------------------------------------------------------------------------
int heomq_sweep_create(const heomq_params* params, int nprob, int device, heomq_sweep** out) {
  if (!params || nprob < 1 || !out) return fail(HEOMQ_BAD_ARG, "null argument or no problems");
  auto W = std::make_unique<heomq_sweep>();
  W->nprob = nprob;
  // group the problems by hierarchy shape (the fields set_params requires to
  // match; batch_build checks every member against its shape again)
  auto key = [](const heomq_params& p) {
    return std::vector<int64_t>{p.norbs, p.nalf, p.nmats, p.ntop, p.lsdraw, p.ham, p.lscale};
  };
  std::vector<std::vector<int64_t>> keys;
  std::vector<std::vector<int>> members;
  for (int p = 0; p < nprob; ++p) {
    const std::vector<int64_t> k = key(params[p]);
    const size_t g = std::find(keys.begin(), keys.end(), k) - keys.begin();
    if (g == keys.size()) {
      if ((int)keys.size() == kMaxShapes)
        return fail(HEOMQ_UNSUPPORTED, "sweep: more than " + std::to_string(kMaxShapes) + " hierarchy shapes");
      keys.push_back(k);
      members.emplace_back();
    }
    members[g].push_back(p);
  }
  const int ng = (int)members.size();
  W->ng = ng;
  std::vector<std::unique_ptr<heomq_batch>> built;
  for (int g = 0; g < ng; ++g) {
    std::vector<heomq_params> ps;
    for (int p : members[g]) ps.push_back(params[p]);
    heomq_batch* b = nullptr;
    if (int rc = batch_build(ps.data(), (int)ps.size(), device, false, &b)) return rc;
    built.emplace_back(b);
  }
  // largest shape first: its long heavy workgroups are dispatched first
  std::vector<int> order(ng);
  for (int g = 0; g < ng; ++g) order[g] = g;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return built[a]->n > built[b]->n; });
  for (int k = 0; k < ng; ++k) W->shapes.push_back(built[order[k]].release());   // owned by W from here
  const heomq_ctx* c0 = W->shapes[0]->ctx;
  W->shape_of.assign(nprob, 0);
  int64_t voff = 0;
  int32_t vb = 0, q = 0;
  W->m.ng = ng;
  for (int k = 0; k < ng; ++k) {
    heomq_batch* B = W->shapes[k];
    const heomq_ctx* c = B->ctx;
    if (c->d != c0->d || c->nalf != c0->nalf || c->nspin != c0->nspin || c->norbs != c0->norbs)
      return fail(HEOMQ_UNSUPPORTED, "sweep: every problem needs the same d, leads, spins and orbitals");
    if (B->bf.n_heavy < 1 || B->bf.trace_slot < 0)
      return fail(HEOMQ_UNSUPPORTED, "sweep: every shape needs a top tier >= 1 and the tier-0 trace slot");
    const int64_t nbv = vec_blocks(c->ptotal), nba = apply_blocks(B->bf);
    if ((int64_t)vb + nbv * B->nprob > INT32_MAX || nba * B->nprob > INT32_MAX / kApplyShapesPerLaunch)
      return fail(HEOMQ_UNSUPPORTED, "sweep: too many blocks for one launch");
    W->m.bstart[k] = vb;
    W->m.nb[k] = (int32_t)nbv;
    W->m.qstart[k] = q;
    W->m.q0[k] = q;
    W->m.n2[k] = c->ptotal;
    W->m.voff[k] = voff;
    for (int j = 0; j < B->nprob; ++j) {
      const int p = members[order[k]][j];
      W->perm.push_back(p);
      W->shape_of[p] = k;
      W->qoff.push_back(voff + (int64_t)j * B->n);
    }
    vb += (int32_t)(nbv * B->nprob);
    q += B->nprob;
    voff += B->n * B->nprob;
  }
  W->m.bstart[ng] = vb;
  W->m.qstart[ng] = q;
  W->nt = voff;
  HIPCHECK(hipSetDevice(device));
  HIPCHECK(W->work.alloc(sizeof(double) * (size_t)W->nt * kTfqVecs));
  HIPCHECK(W->scal.alloc(sizeof(TfqmrScal) * (size_t)nprob));
  HIPCHECK(W->partial.alloc(sizeof(double) * 2 * (size_t)vb));
  HIPCHECK(W->scal2.alloc(sizeof(TfqmrScal) * (size_t)nprob));
  HIPCHECK(W->partial2.alloc(sizeof(double) * 2 * (size_t)vb));
  HIPCHECK(W->out.alloc(sizeof(double) * 2 * (size_t)nprob));
  HIPCHECK(W->mask.alloc(2 * sizeof(int32_t) * (size_t)nprob));
  HIPCHECK(W->pmap.alloc(sizeof(int32_t) * (size_t)nprob));
  HIPCHECK(hipHostMalloc(&W->pinned, (sizeof(TfqmrScal) + 2 * sizeof(double) + 2 * sizeof(int32_t)) * (size_t)nprob + 256,
                         hipHostMallocDefault));
  HIPCHECK(hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking));
  const int32_t* flags = reinterpret_cast<const int32_t*>(W->scal.as<char>() + offsetof(TfqmrScal, flag));
  for (int k = 0; k < ng; ++k) {
    DevPacked op = W->shapes[k]->bf;
    op.xoff = W->m.voff[k];
    op.frozen_flag = flags + (int64_t)W->m.qstart[k] * (int64_t)(sizeof(TfqmrScal) / sizeof(int32_t));
    op.frozen_stride = (int)(sizeof(TfqmrScal) / sizeof(int32_t));
    op.precond = W->shapes[k]->precond ? 1 : 0;   // b = e_trace sits on the unscaled trace row
    W->hops_it.push_back(op);
    op.precond = 0;
    W->hops0.push_back(op);
  }
  // the apply launches: consecutive runs of kApplyShapesPerLaunch shapes
  for (int k0 = 0; k0 < ng; k0 += kApplyShapesPerLaunch) {
    ApplyMulti ai, a0;
    int32_t ab = 0;
    for (int k = k0; k < ng && k < k0 + kApplyShapesPerLaunch; ++k) {
      const int j = k - k0;
      ai.ops[j] = W->hops_it[k];
      a0.ops[j] = W->hops0[k];
      ai.bstart[j] = a0.bstart[j] = ab;
      ai.nb[j] = a0.nb[j] = (int32_t)apply_blocks(W->hops_it[k]);
      ab += ai.nb[j] * W->shapes[k]->nprob;
      ai.ng = a0.ng = j + 1;
    }
    ai.bstart[ai.ng] = a0.bstart[a0.ng] = ab;
    W->apply_blocks += ab;
    W->am_it.push_back(ai);
    W->am0.push_back(a0);
  }
  *out = W.release();
  return HEOMQ_OK;
}

void heomq_sweep_free(heomq_sweep* w) { delete w; }

int heomq_sweep_info(const heomq_sweep* w, int64_t* info) {
  if (!w || !info) return fail(HEOMQ_BAD_ARG, "null argument");
  info[0] = w->nprob;
  info[1] = w->ng;
  info[2] = w->nt / 2;
  info[3] = w->shapes[0]->ctx->d;
  info[4] = w->apply_blocks;
  info[5] = w->m.bstart[w->ng];
  return HEOMQ_OK;
}

int heomq_sweep_problems(const heomq_sweep* w, int32_t* shape, int64_t* off, int64_t* len) {
  if (!w) return fail(HEOMQ_BAD_ARG, "null argument");
  for (int q = 0; q < w->nprob; ++q) {
    const int p = w->perm[q], g = w->shape_of[p];
    if (shape) shape[p] = g;
    if (off) off[p] = w->qoff[q] / 2;
    if (len) len[p] = w->m.n2[g];
  }
  return HEOMQ_OK;
}

int heomq_sweep_apply_packed(heomq_sweep* w, const double* xp, double* yp, int trace, void* stream) {
  if (!w || !xp || !yp) return fail(HEOMQ_BAD_ARG, "null argument");
  if (trace != HEOMQ_TRACE_RDO) return fail(HEOMQ_BAD_ARG, "sweep apply: trace RDO only");
  for (ApplyMulti am : w->am0) {
    for (int g = 0; g < am.ng; ++g) am.ops[g].frozen_flag = nullptr;   // every problem
    HIPCHECK(launch_apply_packed_multi_f64(am, xp, yp, true, (hipStream_t)stream));
  }
  return HEOMQ_OK;
}

int heomq_sweep_solve(heomq_sweep* W, double tol, int maxiter, double* rdo, double* current, double* occ,
                      double* stats, void* stream) {
  if (!W) return fail(HEOMQ_BAD_ARG, "null argument");
  if (stream) HIPCHECK(hipStreamSynchronize((hipStream_t)stream));
  const int P = W->nprob;
  if (W->covered != P) {   // the previous solve ended on a compacted layout
    if (W->chunk) HIPCHECK(hipGraphExecDestroy(W->chunk));
    W->chunk = nullptr;
  }
  W->cur = W->m;
  W->cur_it = W->am_it;
  W->cur0 = W->am0;
  W->covered = P;
  SweepLayout L{W, W->stream};
  const TfqBufs V = carve_bufs(W->work.as<double>(), W->nt, W->scal.as<TfqmrScal>(), W->scal2.as<TfqmrScal>(),
                               W->partial.as<double>(), W->partial2.as<double>(), W->out.as<double>(),
                               W->mask.as<int32_t>(), W->pinned, P);
  SolveOut so;
  if (int rc = tfqmr_loop(L, V, W->chunk, W->stream, tol, maxiter, so)) return rc;
  for (int q = 0; q < P; ++q) {
    const int p = W->perm[q], g = W->shape_of[p];
    heomq_batch* B = W->shapes[g];
    heomq_ctx* c = B->ctx;
    const int d = c->d, nal = c->nalf * c->nspin, noc = c->norbs * c->nspin;
    if (int rc = problem_observables(c, B->dbsqrt[q - W->m.qstart[g]], V.x + W->qoff[q],
                                     rdo ? rdo + 2 * (size_t)p * d * d : nullptr,
                                     current ? current + (size_t)p * nal : nullptr,
                                     occ ? occ + (size_t)p * noc : nullptr))
      return rc;
    if (stats) write_stats(so, q, stats + 4 * p);
  }
  for (int q = 0; q < P; ++q)
    if (so.status[q] != 1)
      return fail(HEOMQ_NOT_CONVERGED, "sweep: problem " + std::to_string(W->perm[q]) + " did not converge");
  return HEOMQ_OK;
}

}  // extern "C"
