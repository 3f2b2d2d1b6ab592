// Solver-level entry points: device TFQMR (the solve_tfqmr.f90 + dutfx.f90
// drop-in), FP32-inside-FP64 iterative refinement, RK4 (td/intstp.f90),
// Chebyshev (td/chebyshev.f90 + prechebyshev.f90), observables
// (td/calcj.f90, main/calcocc.f90) and the host-pointer apply drop-ins.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <type_traits>
#include <limits>
#include <mutex>
#include <vector>

#include "heomq_ctx.hpp"

using namespace heomq;

namespace {

template <typename T>
hipError_t apply_T(heomq_ctx* c, bool trans, const T* x, T* y, int64_t tpos, hipStream_t s, bool proj);
// prec: 0 none; 1 left Jacobi scaling inside the apply; 2 left block-Jacobi
// preconditioning y <- (shift - S_K)^-1 y per ADO after the apply (FP64
// forward operator only; scratch holds the eigenbasis coefficients)
template <typename T>
hipError_t apply_packed_T(heomq_ctx* c, bool trans, const T* x, T* y, bool rdo, hipStream_t s, int prec = 0,
                          T* scratch = nullptr);
// The forward packed operator of the context: the plain apply, or with the
// derivative terminator set the apply of the hierarchy one tier higher with
// that tier eliminated.
hipError_t packed_forward(heomq_ctx* c, const DevPacked& op, const double* x, double* y, bool rdo, hipStream_t s) {
  if (c->trun.built) return trun_apply(c, x, y, rdo, op.precond, op.herm0, s);
  return launch_apply_packed_f64(op, x, y, rdo, s);
}

template <>
hipError_t apply_packed_T<double>(heomq_ctx* c, bool trans, const double* x, double* y, bool rdo, hipStream_t s,
                                  int prec, double* scratch) {
  if (trans && c->trun.built) return hipErrorNotSupported;
  DevPacked op = trans ? c->ptrn : c->pfwd;
  op.precond = prec == 1 ? 1 : 0;
  if (prec != 2) return trans ? launch_apply_packed_f64(op, x, y, rdo, s) : packed_forward(c, op, x, y, rdo, s);
  DevJacobi jc = c->pjac;
  jc.ejac = c->prec_shift;
  if (c->fuse_precond && !c->trun.built && apply_jac_fusable(op, jc))   // one launch, no A x in HBM
    return launch_apply_packed_jac(op, jc, x, y, rdo, s);
  if (jacobi_one_kernel(jc)) {   // A x into the scratch, then y = M^-1 (A x) in one kernel
    if (hipError_t e = packed_forward(c, op, x, scratch, rdo, s)) return e;
    return launch_jacobi_1k(op, jc, scratch, y, false, s);
  }
  if (hipError_t e = packed_forward(c, op, x, y, rdo, s)) return e;
  if (hipError_t e = launch_jacobi_fwd(op, jc, y, scratch, s)) return e;
  return launch_jacobi_back(op, jc, scratch, y, false, s);
}
template <>
hipError_t apply_packed_T<float>(heomq_ctx* c, bool trans, const float* x, float* y, bool rdo, hipStream_t s,
                                 int prec, float* scratch) {
  if (c->trun.built) return hipErrorNotSupported;
  DevPacked op = trans ? c->ptrn : c->pfwd;
  op.precond = prec == 1 ? 1 : 0;
  if (prec != 2) return launch_apply_packed_f32(op, x, y, rdo, s);
  DevJacobi jc = c->pjac;   // FP32 vectors: the ADO-aligned inverse only (checked by the caller)
  jc.ejac = c->prec_shift;
  if (hipError_t e = launch_apply_packed_f32(op, x, scratch, rdo, s)) return e;
  return launch_jacobi_ado_f32(op, jc, scratch, y, s);
}

// y = M^-1 b with the block-Jacobi preconditioner (the right-hand side of a
// left-preconditioned solve); t is scratch for the two-kernel form.
hipError_t precond_rhs(heomq_ctx* c, const double* b, double* y, double* t, hipStream_t st) {
  DevJacobi jc = c->pjac;
  jc.ejac = c->prec_shift;
  if (jacobi_one_kernel(jc)) return launch_jacobi_1k(c->pfwd, jc, b, y, false, st);
  if (hipError_t e = launch_jacobi_fwd(c->pfwd, jc, b, t, st)) return e;
  return launch_jacobi_back(c->pfwd, jc, t, y, false, st);
}
hipError_t precond_rhs(heomq_ctx* c, const float* b, float* y, float*, hipStream_t st) {
  DevJacobi jc = c->pjac;
  jc.ejac = c->prec_shift;
  return launch_jacobi_ado_f32(c->pfwd, jc, b, y, st);
}
// Solvers iterate with the projected operator Pi*A (Pi: parity sectors +
// Hermitian tier 0).  A maps range(Pi) into itself and the physical steady
// state is its unique solution there, so every Krylov vector stays physical
// and round-off in the operator's non-physical null directions is removed at
// every apply.  For the adjoint, (Pi A Pi)^T = Pi A^T Pi.
template <>
hipError_t apply_T<double>(heomq_ctx* c, bool trans, const double* x, double* y, int64_t tpos, hipStream_t s,
                           bool proj) {
  return launch_apply_f64(trans ? c->trn.op : c->fwd.op, x, y, tpos, s, proj);
}
template <>
hipError_t apply_T<float>(heomq_ctx* c, bool trans, const float* x, float* y, int64_t tpos, hipStream_t s,
                          bool proj) {
  return launch_apply_f32(trans ? c->trn.op : c->fwd.op, x, y, tpos, s, proj);
}

int64_t tpos_of(const heomq_ctx* c, int trace) {
  if (trace == HEOMQ_TRACE_LAST) return c->N - 1;
  if (trace == HEOMQ_TRACE_RDO) return 0;
  return -1;
}

template <typename T>
struct Work {
  T *w, *y2, *u2, *aw, *v, *d;        // the Krylov state (A is applied to w: y1, u1 are formed in the kernels)
  T *ax, *xbest;                      // true-residual product, best iterate so far
  T* pt = nullptr;                    // block-Jacobi preconditioning: eigenbasis scratch
  T* bp = nullptr;                    //   and the preconditioned right-hand side
};
constexpr int kWorkVecs = 8;

template <typename T>
int carve(DevBuf& buf, int64_t n, int count, std::vector<T*>& out) {
  HIPCHECK(buf.alloc(sizeof(T) * (size_t)n * count + 256 * count));
  out.clear();
  char* base = static_cast<char*>(buf.p);
  const size_t stride = ((sizeof(T) * (size_t)n + 255) / 256) * 256;
  for (int k = 0; k < count; ++k) out.push_back(reinterpret_cast<T*>(base + k * stride));
  return HEOMQ_OK;
}

struct RunStats {
  int iters = 0, applies = 0;
  double relres = 0;
  bool fused = false;   // some iterations ran with their vector work in the operator launches
};

// Whether the block-Jacobi iteration can run with its vector work carried by
// the fused preconditioned launches (FP64, forward operator, one problem);
// HEOMQ_TFQ_FUSED=0 keeps the five-kernel iteration for comparison.
bool tfq_fusable(heomq_ctx* c, bool trans, bool rdo) {
  static const bool off = [] {
    const char* e = std::getenv("HEOMQ_TFQ_FUSED");
    return e && e[0] == '0';
  }();
  if (off || trans || !c->fuse_precond || c->trun.built) return false;
  DevPacked op = c->pfwd;
  op.precond = 0;
  DevJacobi jc = c->pjac;
  return apply_jac_fusable(op, jc) && (!rdo || op.trace_slot >= 0);
}

// Sums of the fused iteration's operator launches over na workgroups: four
// arrays of na, then three of ceil(na / 16).
size_t tfq_fused_sums(int64_t na) { return 4 * (size_t)na + 3 * (size_t)((na + 15) / 16); }

// The fused iteration, `iters` times: step 1 (with the previous iteration's
// deferred x, d and v work), M^-1 A y2 with step 2, M^-1 A w with the rt.aw
// sums, and the one-block fold of all three reductions.
hipError_t record_fused(heomq_ctx* c, double* x, const Work<double>& W, int64_t n, hipStream_t st, bool rdo,
                        int iters) {
  TfqmrScal* s = c->scal.as<TfqmrScal>();
  double* part = c->partial.as<double>();
  DevPacked op = c->pfwd;
  op.precond = 0;
  DevJacobi jc = c->pjac;
  jc.ejac = c->prec_shift;
  const int na = (int)apply_jac_blocks(op);
  double* pa = c->partial_apply.as<double>();
  if (!pa || c->partial_apply.bytes < sizeof(double) * tfq_fused_sums(na)) return hipErrorInvalidValue;
  for (int k = 0; k < iters; ++k) {
    if (hipError_t e = tfq_step1_fused(st, n, W.w, W.y2, W.v, W.aw, W.u2, W.d, x, s, part)) return e;
    if (hipError_t e = launch_apply_packed_jac_tfq(op, jc, W.y2, W.u2, rdo, 1, TfqEpi{s, W.w, pa}, st)) return e;
    if (hipError_t e = launch_apply_packed_jac_tfq(op, jc, W.w, W.aw, rdo, 2,
                                                   TfqEpi{s, nullptr, pa + 3 * na, pa, pa + 4 * na}, st))
      return e;
    if (hipError_t e = tfq_fold(st, n, part, pa, na, s)) return e;
  }
  return hipSuccess;
}

// The TFQMR iteration, `iters` times: step1, A y2, step2, A w, v (fused >= 0:
// the fused iteration, with the trace row when fused is 1).
template <typename T, typename Apply>
hipError_t record_chunk(heomq_ctx* c, T* x, const Work<T>& W, int64_t n, hipStream_t st, Apply& A, int iters,
                        int fused = -1) {
  if (fused >= 0) {
    if constexpr (std::is_same<T, double>::value) return record_fused(c, x, W, n, st, fused == 1, iters);
    return hipErrorInvalidValue;
  }
  TfqmrScal* s = c->scal.as<TfqmrScal>();
  TfqmrScal* s2 = c->scal2.as<TfqmrScal>();
  double* part = c->partial.as<double>();
  double* part2 = c->partial2.as<double>();
  for (int k = 0; k < iters; ++k) {
    if (hipError_t e = tfq_step1<T>(st, n, W.w, W.y2, W.v, W.aw, W.u2, W.d, s, part)) return e;
    if (hipError_t e = A(W.y2, W.u2)) return e;
    if (hipError_t e = tfq_step2<T>(st, n, x, W.w, W.u2, W.d, W.y2, s, s2, part, part2)) return e;
    if (hipError_t e = A(W.w, W.aw)) return e;
    if (hipError_t e = tfq_v<T>(st, n, W.aw, W.u2, W.v, x, W.d, s, s2, part, part2)) return e;
  }
  return hipSuccess;
}

// The chunk of `iters` iterations as an instantiated graph (captured once per
// set of buffers, then replayed): one launch per chunk instead of ~11 per
// iteration.  The kernels only read the device scalars, so a replay after
// convergence or breakdown does nothing.
template <typename T, typename Apply>
int chunk_graph(heomq_ctx* c, T* x, const T* b, const Work<T>& W, int64_t n, bool trans, int64_t tpos, int mode,
                hipStream_t st, Apply& A, int iters, hipGraphExec_t* out, int fused = -1) {
  for (const auto& g : c->graphs)
    if (g.x == x && g.b == b && g.work == W.w && g.n == n && g.tpos == tpos && g.mode == mode &&
        g.trans == (int)trans && g.bytes == (int)sizeof(T) && g.stream == st) {
      *out = g.exec;
      return HEOMQ_OK;
    }
  if (std::getenv("HEOMQ_SOLVER_LOG"))
    std::fprintf(stderr, "tfqmr: capturing a chunk graph (mode %d, %zu cached)\n", mode, c->graphs.size());
  HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const hipError_t rec = record_chunk<T>(c, x, W, n, st, A, iters, fused);
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(st, &graph);
  if (rec != hipSuccess || end != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return fail(HEOMQ_HIP_ERROR, std::string("TFQMR chunk capture: ") + hipGetErrorString(rec != hipSuccess ? rec : end));
  }
  hipGraphExec_t exec = nullptr;
  const hipError_t inst = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (inst != hipSuccess) return fail(HEOMQ_HIP_ERROR, std::string("TFQMR chunk instantiate: ") + hipGetErrorString(inst));
  if (c->graphs.size() >= 16) c->drop_graphs();
  c->graphs.push_back({x, b, W.w, n, tpos, mode, (int)trans, (int)sizeof(T), 0, st, exec});
  *out = exec;
  return HEOMQ_OK;
}

// TFQMR on A x = b with every vector and scalar on the device.  Restarts from
// the true residual when the quasi-residual bound is met but ||b - Ax|| is not
// below tol*||b|| (at most 3 times; stagnation stops it).
// mode: 0 dense vectors and the plain operator (any trace row); 1 symmetry-
// packed vectors and the projected operator Pi*A (trace row on the tier-0
// block), the iterate's tier-0 block re-Hermitianised at every restart.
template <typename T>
int tfqmr_run(heomq_ctx* c, T* x, const T* b, const Work<T>& W, int64_t n, bool trans, int64_t tpos,
              double tol, int maxiter, hipStream_t st, RunStats& rs, int mode, int prec = 0) {
  // the iteration runs on the left-preconditioned operator M^-1 A (prec 1:
  // Jacobi scaling, 2: block Jacobi; 0: A itself); every true residual
  // ||b - A x|| uses the operator itself.  Jacobi scaling keeps b (the
  // callers enable it for b = e_trace, whose one element sits on the
  // unscaled trace row); block Jacobi iterates with M^-1 b in W.bp.
  auto Ait = [&](const T* in, T* out) -> hipError_t {
    return mode == 1 ? apply_packed_T<T>(c, trans, in, out, tpos >= 0, st, prec, W.pt)
                     : apply_T<T>(c, trans, in, out, tpos, st, false);
  };
  auto A0 = [&](const T* in, T* out) -> hipError_t {
    return mode == 1 ? apply_packed_T<T>(c, trans, in, out, tpos >= 0, st, 0)
                     : apply_T<T>(c, trans, in, out, tpos, st, false);
  };
  if (prec == 2) {
    if (mode != 1 || trans || !W.pt || !W.bp || (sizeof(T) != sizeof(double) && c->pjac.ado_blocks <= 0))
      return fail(HEOMQ_BAD_ARG, "block-Jacobi preconditioning needs the forward packed operator (FP32: the "
                                 "ADO-aligned inverse)");
    HIPCHECK(precond_rhs(c, b, W.bp, W.pt, st));
  }
  bool fuse_ok = false;
  if constexpr (std::is_same<T, double>::value) fuse_ok = mode == 1 && prec == 2 && tfq_fusable(c, trans, tpos >= 0);
  if (fuse_ok) {   // captured chunks hold the buffer's address: a new buffer retires them
    const void* before = c->partial_apply.p;
    HIPCHECK(c->partial_apply.alloc(sizeof(double) * tfq_fused_sums(apply_jac_blocks(c->pfwd))));
    if (c->partial_apply.p != before) c->drop_graphs();
  }
  TfqmrScal* s = c->scal.as<TfqmrScal>();
  double* part = c->partial.as<double>();
  double* red = c->red.as<double>();
  double* hr = static_cast<double*>(c->pinned);
  auto true_resid = [&](double& rr, double& bb) -> int {
    HIPCHECK(A0(x, W.ax));
    rs.applies++;
    HIPCHECK((vec_resid<T, T>(st, n, b, W.ax, red, part)));
    HIPCHECK(hipMemcpyAsync(hr, red, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    rr = hr[0];
    bb = hr[1];
    rs.relres = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    return HEOMQ_OK;
  };
  const int check_every = 8;      // iterations per graph replay / host read of the scalars
  const int verify_every = 32;    // replays between true-residual checks (256 iterations)
  static const bool log = std::getenv("HEOMQ_SOLVER_LOG") != nullptr;
  double best = std::numeric_limits<double>::infinity();
  int base_iters = 0, stalls = 0;
  while (true) {
    if (mode == 1) HIPCHECK(launch_herm_packed<T>(c->pfwd, x, st));
    double rr = 0, bb = 0;
    if (int rc = true_resid(rr, bb)) return rc;
    if (log)
      std::fprintf(stderr, "tfqmr: start iters %d relres %.3e best %.3e prec %d stalls %d\n", base_iters,
                   bb > 0 ? std::sqrt(rr / bb) : 0.0, bb > 0 ? std::sqrt(best / bb) : 0.0, prec, stalls);
    if (rr <= tol * tol * bb || rr == 0.0) return HEOMQ_OK;
    if (base_iters >= maxiter) break;
    const double best_before = best;
    if (rr < best) {
      best = rr;
      HIPCHECK(hipMemcpyAsync(W.xbest, x, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
    } else {
      if (++stalls > 3) {
        if (!prec) break;
        prec = 0;   // the preconditioned iteration stalled: continue unpreconditioned from the best iterate
        stalls = 0;
      }
      if (rr > 100.0 * best) {   // the recurrences diverged: restart from the best iterate
        HIPCHECK(hipMemcpyAsync(x, W.xbest, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
        if (int rc = true_resid(rr, bb)) return rc;
      }
    }
    HIPCHECK(tfq_reset(st, s, tol * std::sqrt(bb), nullptr, 1, base_iters == 0 && std::isinf(best_before),
                       c->shadow_seed));
    if (prec) {
      HIPCHECK(Ait(x, W.ax));
      rs.applies++;
    }
    HIPCHECK(tfq_init_r<T>(st, n, prec == 2 ? W.bp : b, W.ax, W.w, W.y2, W.d, s, part));
    TfqmrScal& hs = *reinterpret_cast<TfqmrScal*>(static_cast<char*>(c->pinned) + 1024);
    if (prec) {   // the quasi-residual bound measures the scaled residual: scale the target with it
      HIPCHECK(hipMemcpyAsync(&hs, s, sizeof(TfqmrScal), hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      if (rr > 0 && hs.rr > 0)
        HIPCHECK(vec_set_elem(st, reinterpret_cast<double*>(s), offsetof(TfqmrScal, target) / sizeof(double),
                              tol * std::sqrt(bb) * std::sqrt(hs.rr / rr)));
    }
    HIPCHECK(Ait(W.w, W.aw));
    rs.applies++;
    HIPCHECK(tfq_init_v<T>(st, n, W.aw, W.v, nullptr, W.u2, s, part));
    hipGraphExec_t chunk = nullptr;
    int no_progress = 0;
    double last_check = rr;
    // the fused iteration leaves x one deferred update behind: flush() before x is read
    const int fused = (prec == 2 && fuse_ok) ? (tpos >= 0 ? 1 : 0) : -1;
    rs.fused = rs.fused || fused >= 0;
    auto flush = [&]() -> hipError_t {
      if constexpr (std::is_same<T, double>::value)
        if (fused >= 0) return tfq_flush(st, n, x, W.d, W.y2, s);
      return hipSuccess;
    };
    if (c->use_graphs)
      if (int rc = chunk_graph<T>(c, x, b, W, n, trans, tpos, mode + 2 * (int)prec + 8 * (fused >= 0), st, Ait,
                                  check_every, &chunk, fused))
        return rc;
    for (int replay = 1;; ++replay) {
      if (chunk)
        HIPCHECK(hipGraphLaunch(chunk, st));
      else
        HIPCHECK((record_chunk<T>(c, x, W, n, st, Ait, check_every, fused)));
      HIPCHECK(hipMemcpyAsync(&hs, s, sizeof(TfqmrScal), hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      rs.iters = base_iters + hs.iter;
      rs.applies += 2 * check_every;
      if (hs.flag != 0 || rs.iters >= maxiter) break;
      if (replay % verify_every == 0) {   // converged already, drifting away or stagnating?
        HIPCHECK(flush());
        double r2 = 0, b2 = 0;
        if (int rc = true_resid(r2, b2)) return rc;
        if (log)
          std::fprintf(stderr, "tfqmr: verify iters %d relres %.3e\n", rs.iters, b2 > 0 ? std::sqrt(r2 / b2) : 0.0);
        if (r2 <= tol * tol * b2 || r2 > 100.0 * best) break;
        // less than a 10 % drop of the residual norm since the last check
        // counts as no progress; two such checks in a row restart the
        // recurrences (fresh shadow vector) from the best iterate
        no_progress = r2 > 0.81 * last_check ? no_progress + 1 : 0;
        last_check = r2;
        if (r2 < best) {
          best = r2;
          HIPCHECK(hipMemcpyAsync(W.xbest, x, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
        }
        if (no_progress >= 2) break;
      }
    }
    HIPCHECK(flush());
    base_iters = rs.iters;
    if (hs.flag == 2 && ++stalls > 3) return fail(HEOMQ_BREAKDOWN, "TFQMR breakdown");
  }
  double rr = 0, bb = 0;
  if (int rc = true_resid(rr, bb)) return rc;
  if (rr > best) {   // hand back the best iterate seen
    HIPCHECK(hipMemcpyAsync(x, W.xbest, sizeof(T) * n, hipMemcpyDeviceToDevice, st));
    if (int rc = true_resid(rr, bb)) return rc;
  }
  if (rr <= tol * tol * bb) return HEOMQ_OK;
  return fail(HEOMQ_NOT_CONVERGED, "TFQMR did not reach the tolerance");
}

int inner_maxit() {
  static const int v = [] {
    const char* e = std::getenv("HEOMQ_INNER_MAXIT");
    const int k = e ? std::atoi(e) : 0;
    return k > 0 ? k : 64;
  }();
  return v;
}

}  // namespace

namespace heomq {
// Device scratch of heomq_apply_packed_precond_noalloc (one per GPU).
__device__ __align__(256) double scratch[2 * kNoAllocElems];

// Address of `scratch` on `device` (made current by the caller), resolved
// with hipGetSymbolAddress once per device and cached; host code passes this
// address, never the array's name, to kernels.
int scratch_address(int device, double** out) {
  static std::mutex mu;
  static double* addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return fail(HEOMQ_BAD_ARG, "device id out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(HEOMQ_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    addr[device] = static_cast<double*>(p);
  }
  *out = addr[device];
  return HEOMQ_OK;
}
}  // namespace heomq

namespace {

int e_trace_into(heomq_ctx* c, double* b, int64_t n, int trace, hipStream_t st) {
  HIPCHECK(vec_fill<double>(st, n, b, 0.0));
  const int64_t pos = tpos_of(c, trace);
  if (pos < 0) return fail(HEOMQ_BAD_ARG, "steady state needs a trace row");
  const int64_t off = 2 * (pos * c->d * c->d + (int64_t)c->d * c->d - 1);
  HIPCHECK(vec_set_elem(st, b, off, 1.0));
  return HEOMQ_OK;
}

}  // namespace

namespace heomq {

// (A + T) x: x goes into the extended iterate with the eliminated tier zeroed;
// the extended apply's top-tier blocks give that tier's right-hand sides B_V
// (lower couplings only, its own elements being zero); the ADO-aligned
// inverse at shift 0 sets x_V = (-S_V)^-1 B_V, the steady solution of V's
// equation without upper couplings; the lower-tier blocks then form rows
// 0..L, the top ones with their upper couplings to x_V.
hipError_t trun_apply(heomq_ctx* c, const double* x, double* y, bool rdo, int precond, int herm0, hipStream_t st) {
  TrunOperator& tr = c->trun;
  double* X = tr.xext.as<double>();
  double* T = tr.text.as<double>();
  const int64_t nv = 2 * tr.vis, nt = 2 * (tr.total - tr.vis);
  if (hipError_t e = hipMemcpyAsync(X, x, sizeof(double) * nv, hipMemcpyDeviceToDevice, st)) return e;
  DevPacked op = tr.op;
  op.rec = c->pfwd.rec;
  if (nt > 0) {
    if (hipError_t e = vec_fill<double>(st, nt, X + nv, 0.0)) return e;
    DevPacked lo = op;   // the eliminated tier's blocks only
    lo.n_heavy = 0;
    if (hipError_t e = launch_apply_packed_f64(lo, X, T, false, st)) return e;
    if (hipError_t e = launch_jacobi_ado(op, tr.jac, T, X, false, st)) return e;
  }
  DevPacked hi = op;     // tiers 0..L only
  hi.total = op.light_begin;
  hi.lt_blk = nullptr;
  hi.n_lt_blocks = 0;
  hi.precond = precond;
  hi.herm0 = herm0;
  return launch_apply_packed_f64(hi, X, y, rdo, st);
}

}  // namespace heomq

extern "C" {

}  // extern "C"

namespace {

// The steady TFQMR solve; x_packed: x already holds the packed iterate (the
// context's packed layout, 2 * ptotal reals), no pack / unpack.
int steady_solve(heomq_ctx* c, double* x, const double* b, double tol, int maxiter, int trace, int transpose,
                 double* stats, void* stream, bool x_packed) {
  if (!c || !x) return fail(HEOMQ_BAD_ARG, "null argument");
  if (trace != HEOMQ_TRACE_LAST && trace != HEOMQ_TRACE_RDO) return fail(HEOMQ_BAD_ARG, "trace row required");
  if (c->trun.built && (trace != HEOMQ_TRACE_RDO || transpose))
    return fail(HEOMQ_UNSUPPORTED, "with the derivative terminator the steady solve is the forward packed one "
                                   "(trace row RDO, no transpose)");
  hipStream_t st = (hipStream_t)stream;
  if (st == nullptr) {   // the iteration is replayed as a graph, which the null stream cannot capture
    HIPCHECK(hipStreamSynchronize(nullptr));
    st = c->solve_stream;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const bool packed = trace == HEOMQ_TRACE_RDO;
  // packed: the solve runs on the symmetry-packed physical sector with Pi*A
  const int64_t n = packed ? 2 * c->ptotal : 2 * c->N * c->d * c->d;
  std::vector<double*> v;
  if (int rc = carve<double>(c->ws64, n, kWorkVecs + (packed ? 4 : 1), v)) return rc;
  Work<double> W{v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]};
  double* bb = v[kWorkVecs];
  double* xs = packed && !x_packed ? v[kWorkVecs + 1] : x;
  if (packed) {
    if (!x_packed) HIPCHECK((launch_pack<double, double>(c->pfwd, x, xs, st)));
    if (b) {
      HIPCHECK((launch_pack<double, double>(c->pfwd, b, bb, st)));
    } else {
      HIPCHECK(vec_fill<double>(st, n, bb, 0.0));
      HIPCHECK(vec_set_elem(st, bb, 2 * (int64_t)c->pfwd.trace_slot, 1.0));
    }
  } else if (b) {
    HIPCHECK(hipMemcpyAsync(bb, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  } else if (int rc = e_trace_into(c, bb, n, trace, st)) {
    return rc;
  }
  RunStats rs;
  // automatic choice: Jacobi scaling when H is diagonal in the occupation
  // basis, block Jacobi otherwise
  int prec = 0;
  if (packed && !transpose) {
    const int want = c->precond >= 0 ? c->precond : (c->h_diagonal ? 1 : 2);
    if (want == 1 && !b) prec = 1;   // the scaling keeps only b = e_trace
    if (want == 2) prec = 2;
  }
  if (prec == 2) {
    W.pt = v[kWorkVecs + 2];
    W.bp = v[kWorkVecs + 3];
  }
  // The steady problem (b = e_trace): a few sweeps of the reference's own
  // convergent block-Jacobi iteration (trace kept) before TFQMR.  They damp
  // the start's tier-1+ residual (3e3 x ||b|| on the two-impurity model), so
  // the recurrences lose less accuracy: 16 shadow seeds took 150 iterations
  // on average instead of 225, the slowest 191 instead of 825.
  int pre = 0;
  if (prec == 2 && !b && !transpose && c->presmooth) {
    if (int rc = jacobi_presmooth(c, xs, kPresmoothSweeps, W.ax, W.pt, st)) return rc;
    pre = kPresmoothSweeps;
  }
  const int rc = tfqmr_run<double>(c, xs, bb, W, n, transpose != 0, tpos_of(c, trace), tol, maxiter, st, rs,
                                   packed ? 1 : 0, prec);
  rs.applies += pre;
  if (packed && !x_packed) HIPCHECK((launch_unpack<double, double>(c->pfwd, xs, x, st)));
  HIPCHECK(hipStreamSynchronize(st));
  if (stats) {
    std::fill(stats, stats + HEOMQ_STATS_LEN, 0.0);
    stats[0] = rs.iters;
    stats[1] = rs.applies;
    stats[2] = rs.relres;
    stats[3] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats[4] = prec;
    stats[5] = (double)n;
    stats[6] = rs.fused ? 1.0 : 0.0;
  }
  return rc;
}

}  // namespace

extern "C" {

int heomq_solve_steady(heomq_ctx* c, double* x, const double* b, double tol, int maxiter, int trace,
                       int transpose, double* stats, void* stream) {
  return steady_solve(c, x, b, tol, maxiter, trace, transpose, stats, stream, false);
}

// The steady state on the sparse pattern (solve_tfqmr.f90 with lsparse = T):
// TFQMR on P A P^T over rho_spa, the trace row on element (d, d) of tier 0
// (which the pattern must hold), tier 0 Hermitianised; preconditioned as the
// packed solve (block Jacobi when the pattern's lists are closed under H's
// groups, Jacobi scaling for diagonal H).  The context's packed operator and
// preconditioner tables are swapped for the pattern's for the duration.
int heomq_spa_solve_steady(heomq_ctx* c, double* x_spa, double tol, int maxiter, double* stats, void* stream) {
  if (!c || !x_spa) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c->spa_lunk < 0 || !c->spa.built) return fail(HEOMQ_BAD_ARG, "no sparse pattern: call heomq_spa_pattern first");
  if (c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_spa_solve_steady: not with the derivative terminator");
  if (c->spa.trace_slot < 0)
    return fail(HEOMQ_UNSUPPORTED, "sparse steady solve: the pattern lacks element (d, d) of tier 0 (the trace row)");
  const int want = c->precond >= 0 ? c->precond : (c->h_diagonal ? 1 : 2);
  if (want == 2 && !c->spa.has_jac)
    return fail(HEOMQ_UNSUPPORTED, "sparse steady solve: block-Jacobi preconditioning needs every ADO's list closed "
                                   "under H's groups (heomq_ctx_set_precond(c, 0) solves unpreconditioned)");
  DevPacked op = spa_operator(c, nullptr);
  op.trace_slot = c->spa.trace_slot;
  op.herm0 = 1;
  const DevPacked keep_op = c->pfwd;
  const DevJacobi keep_jac = c->pjac;
  const int64_t keep_total = c->ptotal;
  c->drop_graphs();   // captured chunks hold the operator by value
  c->pfwd = op;
  if (c->spa.has_jac) {
    DevJacobi jc = c->spa.jac;
    jc.ejac = keep_jac.ejac;
    c->pjac = jc;
  }
  c->ptotal = c->spa_lunk;
  const int rc = steady_solve(c, x_spa, nullptr, tol, maxiter, HEOMQ_TRACE_RDO, 0, stats, stream, true);
  c->drop_graphs();
  c->pfwd = keep_op;
  c->pjac = keep_jac;
  c->ptotal = keep_total;
  return rc;
}

int heomq_solve_steady_mixed(heomq_ctx* c, double* x, const double* b, double tol, double inner_tol,
                             int maxiter, int trace, double* stats, void* stream) {
  if (!c || !x) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_solve_steady_mixed: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (trace != HEOMQ_TRACE_RDO) return fail(HEOMQ_BAD_ARG, "mixed precision solves use the tier-0 trace row");
  hipStream_t st = (hipStream_t)stream;
  if (st == nullptr) {
    HIPCHECK(hipStreamSynchronize(nullptr));
    st = c->solve_stream;
  }
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = 2 * c->ptotal;
  // Preconditioning as the FP64 solver chooses it (automatic: Jacobi scaling
  // for diagonal H, block Jacobi otherwise).  The FP32 inner solves take the
  // block-Jacobi inverse (ADO-aligned form, FP32 vectors, FP64 transform); the
  // Jacobi scaling keeps only b = e_trace, so residual corrections of a
  // diagonal-H model iterate unscaled.
  const int want = c->precond >= 0 ? c->precond : (c->h_diagonal ? 1 : 2);
  const int prec32 = (want == 2 && c->pjac.ado_blocks > 0) ? 2 : 0;
  const int prec64 = want == 2 ? 2 : 0;
  std::vector<double*> v64;
  if (int rc = carve<double>(c->ws64, n, 3, v64)) return rc;
  std::vector<float*> v32;
  if (int rc = carve<float>(c->ws32, n, kWorkVecs + 4, v32)) return rc;
  if (stats) std::fill(stats, stats + HEOMQ_STATS_MIXED_LEN, 0.0);
  Work<float> W{v32[0], v32[1], v32[2], v32[3], v32[4], v32[5], v32[6], v32[7]};
  float* x32 = v32[kWorkVecs];
  float* r32 = v32[kWorkVecs + 1];
  if (prec32 == 2) {
    W.pt = v32[kWorkVecs + 2];
    W.bp = v32[kWorkVecs + 3];
  }
  double* ax = v64[0];
  double* bb = v64[1];
  double* xs = v64[2];
  HIPCHECK((launch_pack<double, double>(c->pfwd, x, xs, st)));
  if (b) {
    HIPCHECK((launch_pack<double, double>(c->pfwd, b, bb, st)));
  } else {
    HIPCHECK(vec_fill<double>(st, n, bb, 0.0));
    HIPCHECK(vec_set_elem(st, bb, 2 * (int64_t)c->pfwd.trace_slot, 1.0));
  }
  double* red = c->red.as<double>();
  double* part = c->partial.as<double>();
  // an inner solve stops at inner_tol or after this many FP32 iterations: the
  // preconditioned recurrences need ~100 per five digits on the two-impurity
  // model, the unpreconditioned ones stall earlier and are cut at 64
  const int inner_cap = std::getenv("HEOMQ_INNER_MAXIT") ? inner_maxit() : (prec32 == 2 ? 400 : 64);
  int outer = 0, inner_total = 0, applies = 0, fp64_iters = 0;
  double relres = 0, prev = 0;
  int rc = HEOMQ_NOT_CONVERGED;
  bool finish64 = false;
  for (outer = 0; outer < 60; ++outer) {
    HIPCHECK(launch_herm_packed<double>(c->pfwd, xs, st));
    HIPCHECK(launch_apply_packed_f64(c->pfwd, xs, ax, true, st));   // FP64 residual
    applies++;
    HIPCHECK((vec_resid<double, double>(st, n, bb, ax, red, part)));
    double* hr = static_cast<double*>(c->pinned);
    HIPCHECK(hipMemcpyAsync(hr, red, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    relres = std::sqrt(hr[0] / hr[1]);
    if (stats && 8 + outer < 48) stats[8 + outer] = relres;   // refinement history, slots 8..47
    if (hr[0] <= tol * tol * hr[1]) { rc = HEOMQ_OK; break; }
    if (inner_total >= maxiter) break;
    // FP32 corrections stop contracting once eps32 * ||A|| ||dx|| reaches the
    // residual; the rest of the way is done by FP64 TFQMR from this iterate.
    if (outer >= 2 && relres > 0.5 * prev) { finish64 = true; break; }
    prev = relres;
    const double rn = std::sqrt(hr[0]);
    HIPCHECK(vec_resid_to_f32(st, n, bb, ax, r32, 1.0 / rn));
    HIPCHECK(vec_fill<float>(st, n, x32, 0.0f));
    RunStats rs;
    const int cap = std::min(inner_cap, maxiter - inner_total);
    const int irc = tfqmr_run<float>(c, x32, r32, W, n, false, 0, inner_tol, cap, st, rs, 1, prec32);
    if (irc != HEOMQ_OK && irc != HEOMQ_NOT_CONVERGED) return irc;
    inner_total += rs.iters;
    applies += rs.applies;
    if (stats && 48 + outer < HEOMQ_STATS_MIXED_LEN) stats[48 + outer] = rs.iters;   // slots 48..63
    HIPCHECK(vec_axpy_f32_to_f64(st, n, xs, x32, rn));   // FP64 update
  }
  if (finish64) {
    std::vector<double*> w64;
    if (int e = carve<double>(c->ws64b, n, kWorkVecs + 2, w64)) return e;
    Work<double> W64{w64[0], w64[1], w64[2], w64[3], w64[4], w64[5], w64[6], w64[7]};
    if (prec64 == 2) {
      W64.pt = w64[kWorkVecs];
      W64.bp = w64[kWorkVecs + 1];
    }
    RunStats rs;
    rc = tfqmr_run<double>(c, xs, bb, W64, n, false, 0, tol, std::max(maxiter - inner_total, 1), st, rs, 1, prec64);
    fp64_iters = rs.iters;
    applies += rs.applies;
    relres = rs.relres;
  }
  HIPCHECK((launch_unpack<double, double>(c->pfwd, xs, x, st)));
  HIPCHECK(hipStreamSynchronize(st));
  if (stats) {
    stats[0] = inner_total;
    stats[1] = applies;
    stats[2] = relres;
    stats[3] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats[4] = outer;
    stats[6] = fp64_iters;
  }
  if (rc != HEOMQ_OK) return fail(HEOMQ_NOT_CONVERGED, "mixed-precision refinement did not converge");
  return rc;
}

// ---- BiCG convention (bicg/) ----
int heomq_apply_bicg(heomq_ctx* c, double* x, double* y, void* stream) {
  if (!c || !x || !y) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_apply_bicg: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  HIPCHECK(launch_apply_bicg_f64(c->fwd.op, x, y, (hipStream_t)stream));
  return HEOMQ_OK;
}

int heomq_apply_transpose_bicg(heomq_ctx* c, double* w, double* z, void* stream) {
  if (!c || !w || !z) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_apply_transpose_bicg: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  HIPCHECK(launch_apply_transpose_bicg_f64(c->trn.op, w, z, (hipStream_t)stream));
  return HEOMQ_OK;
}

// solvebicg.f90 on the device: x (dense, in/out) starts from its tier-0 lower
// triangle (:52-58), b = e_(1,1) (:60-61), stop when ||r|| <= crit (absolute,
// :124-126); x's rho0 is returned Hermitian (:108-113).
}  // extern "C"

namespace {

int dense_in_sector(heomq_ctx* c, const double* x, hipStream_t st, bool& yes);

// The BiCG iteration of bicg/solvebicg.f90 on vectors of n reals: fwd(in, out)
// is A_B (in's tier 0 rebuilt from its lower triangle first), trn(w, z) its
// transpose, e_off the real offset of b = e_(1,1); x holds the initial guess
// with its tier-0 upper triangle zeroed.  The device scalars are read back
// every eight iterations.
template <typename Fwd, typename Trn>
int bicg_core(heomq_ctx* c, double* x, int64_t n, int64_t e_off, const std::vector<double*>& v, double crit,
              int maxiter, hipStream_t st, Fwd&& fwd, Trn&& trn, int& iters, int& applies, int& flag,
              double& rr) {
  double *r = v[0], *sv = v[1], *dv = v[2], *f = v[3], *q = v[4], *qt = v[5];
  BicgScal* sc = reinterpret_cast<BicgScal*>(c->scal.p);
  double* part = c->partial.as<double>();
  if (int rc = fwd(x, q)) return rc;
  HIPCHECK(bicg_init(st, n, q, e_off, r, sv, dv, f, sc, crit, part));
  applies = 1;
  iters = 0;
  BicgScal& hs = *reinterpret_cast<BicgScal*>(static_cast<char*>(c->pinned) + 2048);
  HIPCHECK(hipMemcpyAsync(&hs, sc, sizeof(BicgScal), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  flag = hs.flag;
  const int check_every = 8;
  while (flag == 0 && iters < maxiter) {
    for (int k = 0; k < check_every; ++k) {
      if (int rc = fwd(dv, q)) return rc;
      HIPCHECK(bicg_alpha(st, n, f, q, sc, part));
      if (int rc = trn(f, qt)) return rc;
      HIPCHECK(bicg_update(st, n, x, dv, r, q, sv, qt, sc, part));
      HIPCHECK(bicg_dirs(st, n, r, sv, dv, f, sc));
    }
    HIPCHECK(hipMemcpyAsync(&hs, sc, sizeof(BicgScal), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    iters = hs.iter;
    flag = hs.flag;
    applies = 1 + 2 * iters;
  }
  rr = hs.rr;
  return HEOMQ_OK;
}

}  // namespace

extern "C" {

// A state with nothing outside the symmetry sectors (every physical guess)
// iterates on the packed layout: b = e_(1,1) and A_B map the sectors into
// themselves, so every BiCG vector stays there and the packed iterates are the
// dense ones to rounding.  The packed operators keep the BiCG trace row at
// element (1,1) of tier 0 and skip the tier-0 Hermitianisation.
int heomq_solve_bicg(heomq_ctx* c, double* x, double crit, int maxiter, double* stats, void* stream) {
  if (!c || !x) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_solve_bicg: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  hipStream_t st = (hipStream_t)stream;
  const auto t0 = std::chrono::steady_clock::now();
  bool packed = false;
  if (int rc = dense_in_sector(c, x, st, packed)) return rc;
  std::vector<double*> v;
  int iters = 0, applies = 0, flag = 0;
  double rr = 0.0;
  if (packed) {
    const int64_t n = 2 * c->ptotal;
    if (int rc = carve<double>(c->ws64, n, 7, v)) return rc;
    double* xp = v[6];
    DevPacked fo = c->pfwd, to = c->ptrn;
    fo.precond = to.precond = 0;
    fo.herm0 = to.herm0 = 0;
    const int slot = c->sec_pos_h[(size_t)c->psec_h[0] * c->d * c->d];   // element (1,1) of tier 0
    if (slot < 0) return fail(HEOMQ_UNSUPPORTED, "BiCG: element (1,1) of tier 0 is not in its sector");
    fo.trace_slot = to.trace_slot = slot;
    auto fwd = [&](double* in, double* out) -> int {
      HIPCHECK(launch_bicg_tier0_packed(fo, in, 0, st));
      HIPCHECK(launch_apply_packed_f64(fo, in, out, true, st));   // y(1,1) of tier 0 = tr rho0
      HIPCHECK(launch_bicg_tier0_packed(fo, out, 1, st));
      return HEOMQ_OK;
    };
    auto trn = [&](double* w, double* z) -> int {
      HIPCHECK(launch_bicg_tier0_packed(to, w, 1, st));
      HIPCHECK(launch_apply_packed_f64(to, w, z, true, st));      // w(1,1) moved onto the diagonal of z0
      HIPCHECK(launch_bicg_tier0_packed(to, z, 2, st));
      return HEOMQ_OK;
    };
    HIPCHECK((launch_pack<double, double>(c->pfwd, x, xp, st)));
    HIPCHECK(launch_bicg_tier0_packed(fo, xp, 1, st));
    if (int rc = bicg_core(c, xp, n, 2 * (int64_t)slot, v, crit, maxiter, st, fwd, trn, iters, applies, flag, rr))
      return rc;
    HIPCHECK(launch_bicg_tier0_packed(fo, xp, 0, st));
    HIPCHECK((launch_unpack<double, double>(c->pfwd, xp, x, st)));
  } else {
    const int64_t n = 2 * c->N * c->d * c->d;
    if (int rc = carve<double>(c->ws64, n, 6, v)) return rc;
    auto fwd = [&](double* in, double* out) -> int {
      HIPCHECK(launch_apply_bicg_f64(c->fwd.op, in, out, st));
      return HEOMQ_OK;
    };
    auto trn = [&](double* w, double* z) -> int {
      HIPCHECK(launch_apply_transpose_bicg_f64(c->trn.op, w, z, st));
      return HEOMQ_OK;
    };
    HIPCHECK(launch_lower_tier0_f64(c->d, x, st));
    if (int rc = bicg_core(c, x, n, 0, v, crit, maxiter, st, fwd, trn, iters, applies, flag, rr)) return rc;
    HIPCHECK(launch_herm_from_lower_f64(c->d, x, st));
  }
  HIPCHECK(hipStreamSynchronize(st));
  if (stats) {
    std::fill(stats, stats + HEOMQ_STATS_LEN, 0.0);
    stats[0] = iters;
    stats[1] = applies;
    stats[2] = std::sqrt(rr);
    stats[3] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats[4] = packed ? 1.0 : 0.0;
  }
  if (flag == 2) return fail(HEOMQ_BREAKDOWN, "BiCG breakdown");
  if (flag != 1) return fail(HEOMQ_NOT_CONVERGED, "BiCG did not reach the criterion");
  return HEOMQ_OK;
}

int heomq_solve_bicg_host(heomq_ctx* c, double* rho_host, double crit, int maxiter, double* stats) {
  if (!c || !rho_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_solve_bicg_host: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  const int64_t n = 2 * c->N * c->d * c->d;
  HIPCHECK(c->ws_host.alloc(sizeof(double) * (size_t)n));
  double* x = c->ws_host.as<double>();
  HIPCHECK(hipMemcpy(x, rho_host, sizeof(double) * n, hipMemcpyHostToDevice));
  const int rc = heomq_solve_bicg(c, x, crit, maxiter, stats, nullptr);
  if (rc != HEOMQ_OK && rc != HEOMQ_NOT_CONVERGED) return rc;
  HIPCHECK(hipMemcpy(rho_host, x, sizeof(double) * n, hipMemcpyDeviceToHost));
  return rc;
}

int heomq_ctx_set_precond(heomq_ctx* c, int mode) {
  if (!c || mode < -1 || mode > 2) return fail(HEOMQ_BAD_ARG, "precond mode must be -1, 0, 1 or 2");
  c->precond = mode;
  return HEOMQ_OK;
}

int heomq_ctx_set_precond_shift(heomq_ctx* c, double shift) {
  if (!c) return fail(HEOMQ_BAD_ARG, "null argument");
  if (shift < 0.0) return fail(HEOMQ_BAD_ARG, "the block-Jacobi shift must be positive (0: default)");
  const double v = shift > 0.0 ? shift : kPrecShiftFactor * c->ejac_default;
  for (size_t a = 0; 2 * a + 1 < c->gam_h.size(); ++a)
    if (std::abs(std::complex<double>(v - c->gam_h[2 * a], -c->gam_h[2 * a + 1])) <= 1e-12)
      return fail(HEOMQ_UNSUPPORTED, "block-Jacobi shift equals an ADO's gamma_K");
  c->drop_graphs();   // captured chunks hold the shift as a kernel argument
  c->prec_shift_user = shift;
  c->prec_shift = v;
  return HEOMQ_OK;
}

int heomq_ctx_set_shadow_seed(heomq_ctx* c, uint64_t seed) {
  if (!c) return fail(HEOMQ_BAD_ARG, "null argument");
  c->shadow_seed = seed;
  return HEOMQ_OK;
}

int heomq_ctx_set_presmooth(heomq_ctx* c, int enable) {
  if (!c) return fail(HEOMQ_BAD_ARG, "null argument");
  c->presmooth = enable ? 1 : 0;
  return HEOMQ_OK;
}

int heomq_ctx_set_precond_fusion(heomq_ctx* c, int enable) {
  if (!c) return fail(HEOMQ_BAD_ARG, "null argument");
  c->drop_graphs();   // captured chunks hold the preconditioned apply's kernels
  c->fuse_precond = enable != 0;
  return HEOMQ_OK;
}

int heomq_ctx_set_graphs(heomq_ctx* c, int enable) {
  if (!c) return fail(HEOMQ_BAD_ARG, "null argument");
  c->use_graphs = enable != 0;
  return HEOMQ_OK;
}

int heomq_apply_packed(heomq_ctx* c, const double* x, double* y, int trace, int transpose, void* stream) {
  if (!c || !x || !y) return fail(HEOMQ_BAD_ARG, "null argument");
  if (trace != HEOMQ_TRACE_RDO && trace != HEOMQ_TRACE_NONE) return fail(HEOMQ_BAD_ARG, "packed apply: trace RDO or NONE");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = 2 * c->ptotal;
  std::vector<double*> v;
  if (int rc = carve<double>(c->ws_io, n, 2, v)) return rc;
  const DevPacked& op = transpose ? c->ptrn : c->pfwd;
  if (transpose && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "no transposed apply with the derivative terminator");
  HIPCHECK((launch_pack<double, double>(op, x, v[0], st)));
  if (transpose)
    HIPCHECK(launch_apply_packed_f64(op, v[0], v[1], trace == HEOMQ_TRACE_RDO, st));
  else
    HIPCHECK(packed_forward(c, op, v[0], v[1], trace == HEOMQ_TRACE_RDO, st));
  HIPCHECK((launch_unpack<double, double>(op, v[1], y, st)));
  HIPCHECK(hipStreamSynchronize(st));
  return HEOMQ_OK;
}

int heomq_apply_packed_raw(heomq_ctx* c, const double* xp, double* yp, int trace, int transpose, void* stream) {
  if (!c || !xp || !yp) return fail(HEOMQ_BAD_ARG, "null argument");
  if (trace != HEOMQ_TRACE_RDO && trace != HEOMQ_TRACE_NONE) return fail(HEOMQ_BAD_ARG, "packed apply: trace RDO or NONE");
  if (transpose && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "no transposed apply with the derivative terminator");
  if (transpose)
    HIPCHECK(launch_apply_packed_f64(c->ptrn, xp, yp, trace == HEOMQ_TRACE_RDO, (hipStream_t)stream));
  else
    HIPCHECK(packed_forward(c, c->pfwd, xp, yp, trace == HEOMQ_TRACE_RDO, (hipStream_t)stream));
  return HEOMQ_OK;
}

int heomq_apply_packed_precond(heomq_ctx* c, const double* xp, double* yp, int trace, int form, void* stream) {
  if (!c || !xp || !yp) return fail(HEOMQ_BAD_ARG, "null argument");
  if (trace != HEOMQ_TRACE_RDO && trace != HEOMQ_TRACE_NONE) return fail(HEOMQ_BAD_ARG, "packed apply: trace RDO or NONE");
  if (form < 1 || form > 4) return fail(HEOMQ_BAD_ARG, "form must be 1, 2, 3 or 4");
  if (xp == yp) return fail(HEOMQ_BAD_ARG, "x and y must differ");
  hipStream_t st = (hipStream_t)stream;
  const bool rdo = trace == HEOMQ_TRACE_RDO;
  DevPacked op = c->pfwd;
  op.precond = 0;
  DevJacobi jc = c->pjac;
  jc.ejac = c->prec_shift;
  if (form == 4) {
    if (c->trun.built || !apply_jac_fusable(op, jc))
      return fail(HEOMQ_UNSUPPORTED, "the fused preconditioned apply needs apply form 2 or 6 without the wave form "
                                     "and the ADO-aligned inverse");
    HIPCHECK(launch_apply_packed_jac(op, jc, xp, yp, rdo, st));
    return HEOMQ_OK;
  }
  std::vector<double*> v;
  if (int rc = carve<double>(c->ws64b, 2 * c->ptotal, 1, v)) return rc;
  HIPCHECK(packed_forward(c, op, xp, v[0], rdo, st));
  if (form == 1) {
    if (!jacobi_one_kernel(jc)) return fail(HEOMQ_UNSUPPORTED, "the one-kernel preconditioner does not apply here");
    HIPCHECK(launch_jacobi_1k(op, jc, v[0], yp, false, st));
  } else if (form == 3) {
    if (!jc.windowed) return fail(HEOMQ_UNSUPPORTED, "the LDS-window preconditioner does not apply here");
    HIPCHECK(launch_jacobi_win(op, jc, v[0], yp, false, st));
  } else {
    HIPCHECK(hipMemcpyAsync(yp, v[0], sizeof(double) * 2 * c->ptotal, hipMemcpyDeviceToDevice, st));
    std::vector<double*> w;
    if (int rc = carve<double>(c->ws_io, 2 * c->ptotal, 1, w)) return rc;
    HIPCHECK(launch_jacobi_fwd(op, jc, yp, w[0], st));
    HIPCHECK(launch_jacobi_back(op, jc, w[0], yp, false, st));
  }
  return HEOMQ_OK;
}

// The preconditioned operator without allocating: A x goes to the file-scope
// device array `scratch` (kNoAllocElems complex FP64 elements, 32 MiB, held by
// every GPU the library runs on until the process ends) instead of a
// context-owned workspace.  Its address on each device is resolved once, at
// context creation (scratch_address below), and only that address is passed
// to kernels.  Every call writes the whole of A x before reading it, so
// nothing carries over between calls; calls that overlap on one GPU share the
// array and race.
int heomq_apply_packed_precond_noalloc(heomq_ctx* c, const double* xp, double* yp, int trace, void* stream) {
  if (!c || !xp || !yp) return fail(HEOMQ_BAD_ARG, "null argument");
  if (trace != HEOMQ_TRACE_RDO && trace != HEOMQ_TRACE_NONE) return fail(HEOMQ_BAD_ARG, "packed apply: trace RDO or NONE");
  if (xp == yp) return fail(HEOMQ_BAD_ARG, "x and y must differ");
  if (c->ptotal > kNoAllocElems)
    return fail(HEOMQ_BAD_DIM, "packed vector of " + std::to_string(c->ptotal) +
                                   " complex elements exceeds the static scratch (" + std::to_string(kNoAllocElems) + ")");
  if (!c->scratch_dev) return fail(HEOMQ_HIP_ERROR, "static scratch address not resolved for this device");
  DevPacked op = c->pfwd;
  op.precond = 0;
  DevJacobi jc = c->pjac;
  jc.ejac = c->prec_shift;
  if (!jacobi_one_kernel(jc)) return fail(HEOMQ_UNSUPPORTED, "the one-kernel preconditioner does not apply here");
  hipStream_t st = (hipStream_t)stream;
  HIPCHECK(packed_forward(c, op, xp, c->scratch_dev, trace == HEOMQ_TRACE_RDO, st));
  HIPCHECK(launch_jacobi_1k(op, jc, c->scratch_dev, yp, false, st));
  return HEOMQ_OK;
}

int heomq_apply_packed_raw_f32(heomq_ctx* c, const float* xp, float* yp, int trace, int transpose, void* stream) {
  if (!c || !xp || !yp) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_apply_packed_raw_f32: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (trace != HEOMQ_TRACE_RDO && trace != HEOMQ_TRACE_NONE) return fail(HEOMQ_BAD_ARG, "packed apply: trace RDO or NONE");
  HIPCHECK(launch_apply_packed_f32(transpose ? c->ptrn : c->pfwd, xp, yp, trace == HEOMQ_TRACE_RDO,
                                   (hipStream_t)stream));
  return HEOMQ_OK;
}

int heomq_pack(heomq_ctx* c, const double* dense, double* packed, void* stream) {
  if (!c || !dense || !packed) return fail(HEOMQ_BAD_ARG, "null argument");
  HIPCHECK((launch_pack<double, double>(c->pfwd, dense, packed, (hipStream_t)stream)));
  return HEOMQ_OK;
}

int heomq_unpack(heomq_ctx* c, const double* packed, double* dense, void* stream) {
  if (!c || !dense || !packed) return fail(HEOMQ_BAD_ARG, "null argument");
  HIPCHECK((launch_unpack<double, double>(c->pfwd, packed, dense, (hipStream_t)stream)));
  return HEOMQ_OK;
}

int heomq_packed_info(const heomq_ctx* c, int64_t* info) {
  if (!c || !info) return fail(HEOMQ_BAD_ARG, "null argument");
  info[0] = c->ptotal;
  info[1] = c->nsec;
  info[2] = c->cons_n;
  info[3] = c->cons_s;
  info[4] = c->N * c->d * c->d;
  return HEOMQ_OK;
}

}  // extern "C"

namespace {

// The operator of packed propagation: the forward packed apply with neither
// the trace row, the Jacobi scaling nor the tier-0 Hermitianisation, i.e. A
// restricted to the sector that holds every physical state; the rates are
// gamma_K(t) when a lead field is set.  A time-dependent system Hamiltonian
// is not in the packed tables.
int packed_prop_op(heomq_ctx* c, double t, hipStream_t st, DevPacked& op) {
  if (c->ptotal <= 0) return fail(HEOMQ_UNSUPPORTED, "packed propagation: no packed layout");
  if (c->sys_nterm > 0)
    return fail(HEOMQ_UNSUPPORTED, "packed propagation: a time-dependent system Hamiltonian needs the dense path");
  op = c->pfwd;
  op.precond = 0;
  op.herm0 = 0;
  if (c->field_type >= 0) {
    DevOperator dop;
    if (int rc = operator_at(c, t, st, dop)) return rc;
    op.gam = dop.gam;
  }
  return HEOMQ_OK;
}

// Whether a dense state can be propagated on the packed layout: no
// time-dependent system Hamiltonian, and no nonzero element outside the
// sectors (that part is zero and stays zero: A maps the sector into itself and
// the rest into the rest, so the packed result, unpacked, is the dense one).
int dense_in_sector(heomq_ctx* c, const double* x, hipStream_t st, bool& yes) {
  yes = false;
  if (c->ptotal <= 0 || c->sys_nterm > 0) return HEOMQ_OK;
  HIPCHECK(c->prop_flag.alloc(sizeof(int32_t)));
  HIPCHECK(launch_offsector_any(c->pfwd, x, c->prop_flag.as<int32_t>(), st));
  int32_t* h = reinterpret_cast<int32_t*>(static_cast<char*>(c->pinned) + 3072);
  HIPCHECK(hipMemcpyAsync(h, c->prop_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  yes = *h == 0;
  return HEOMQ_OK;
}

// nstep RK4 steps of td/intstp.f90:20-100 on n reals; deriv(t, in, out) is
// the derivative at stage time t (calcderiv_omp(tt) at tt, tt + dt/2 twice,
// tt + dt).
template <typename Deriv>
int rk4_steps(double* x, int64_t n, double* rho0, double* tmp, double* k, double t0, double dt, int nstep,
              hipStream_t st, Deriv&& deriv) {
  const double dt2 = dt / 2, dt3 = dt / 3, dt6 = dt / 6;
  for (int step = 0; step < nstep; ++step) {
    const double t = t0 + step * dt;
    HIPCHECK(hipMemcpyAsync(rho0, x, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    if (int rc = deriv(t, rho0, k)) return rc;
    HIPCHECK(vec_lincomb3(st, n, tmp, 1.0, rho0, dt2, k, 0.0, nullptr));
    HIPCHECK(vec_lincomb3(st, n, x, 1.0, x, dt6, k, 0.0, nullptr));
    if (int rc = deriv(t + dt2, tmp, k)) return rc;
    HIPCHECK(vec_lincomb3(st, n, tmp, 1.0, rho0, dt2, k, 0.0, nullptr));
    HIPCHECK(vec_lincomb3(st, n, x, 1.0, x, dt3, k, 0.0, nullptr));
    if (int rc = deriv(t + dt2, tmp, k)) return rc;
    HIPCHECK(vec_lincomb3(st, n, tmp, 1.0, rho0, dt, k, 0.0, nullptr));
    HIPCHECK(vec_lincomb3(st, n, x, 1.0, x, dt3, k, 0.0, nullptr));
    if (int rc = deriv(t + dt, tmp, k)) return rc;
    HIPCHECK(vec_lincomb3(st, n, x, 1.0, x, dt6, k, 0.0, nullptr));
  }
  return HEOMQ_OK;
}

// The expansion of td/prechebyshev.f90:71-134: delta is the spread of
// Re(gamma_K) over the hierarchy, the weights are J_n(delta dt), cut where
// |J_n| / sum |J_m| < 1e-21 (the term that meets the threshold is kept).
int cheby_setup(heomq_ctx* c, double dt, std::vector<double>& J, double& delta) {
  std::vector<double> gam(2 * (size_t)c->N);
  HIPCHECK(hipMemcpy(gam.data(), c->fwd.gam.p, sizeof(double) * gam.size(), hipMemcpyDeviceToHost));
  double rmax = 0.0, rmin = 0.0;
  for (int64_t a = 0; a < c->N; ++a) {
    rmax = std::max(rmax, gam[2 * a]);
    rmin = std::min(rmin, gam[2 * a]);
  }
  delta = rmax - rmin;
  if (!(delta > 0)) return fail(HEOMQ_BAD_ARG, "Chebyshev: zero spectral width");
  const double arg = delta * dt;
  J.clear();
  double sum = 0.0;
  for (int k = 0; k < 100000; ++k) {
    const double j = std::cyl_bessel_j((double)k, arg);
    J.push_back(j);
    sum += std::fabs(j);
    if (std::fabs(j) / sum < 1e-21) break;
  }
  return HEOMQ_OK;
}

// nstep Chebyshev steps of td/chebyshev.f90:126-200 on n reals: P_0 = x,
// P_1 = L x / delta, P_k = P_{k-2} + (2 / delta) L P_{k-1}, x = sum_k c_k J_k P_k
// (c_0 = 1, else 2); apply(in, out) is L.  term(p1, p, x, a, b), when given,
// does one k >= 2 term in one go (p += a L p1, then x += b p); without it
// the term is the apply and two vector passes.
template <typename Apply>
int cheby_steps(double* x, int64_t n, double* pm2, double* pm1, double* lp, const std::vector<double>& J,
                double delta, int nstep, hipStream_t st, Apply&& apply,
                const std::function<hipError_t(const double*, double*, double*, double, double)>& term = {}) {
  const int ncheby = (int)J.size();
  for (int step = 0; step < nstep; ++step) {
    HIPCHECK(hipMemcpyAsync(pm2, x, sizeof(double) * n, hipMemcpyDeviceToDevice, st));   // P0
    HIPCHECK(vec_lincomb3(st, n, x, J[0], pm2, 0.0, nullptr, 0.0, nullptr));
    if (int rc = apply(pm2, lp)) return rc;
    HIPCHECK(vec_lincomb3(st, n, pm1, 1.0 / delta, lp, 0.0, nullptr, 0.0, nullptr));        // P1
    if (ncheby > 1) HIPCHECK(vec_lincomb3(st, n, x, 1.0, x, 2.0 * J[1], pm1, 0.0, nullptr));
    for (int k = 2; k < ncheby; ++k) {
      // P_k = P_{k-2} + (2/delta) L P_{k-1}, written over P_{k-2}
      if (term) {
        HIPCHECK(term(pm1, pm2, x, 2.0 / delta, 2.0 * J[k]));
      } else {
        if (int rc = apply(pm1, lp)) return rc;
        HIPCHECK(vec_lincomb3(st, n, pm2, 1.0, pm2, 2.0 / delta, lp, 0.0, nullptr));
        HIPCHECK(vec_lincomb3(st, n, x, 1.0, x, 2.0 * J[k], pm2, 0.0, nullptr));
      }
      std::swap(pm1, pm2);
    }
  }
  return HEOMQ_OK;
}

int rk4_packed_core(heomq_ctx* c, double* xp, double t0, double dt, int nstep, hipStream_t st,
                    const std::vector<double*>& v) {
  auto deriv = [&](double t, const double* in, double* out) -> int {
    DevPacked op;
    if (int rc = packed_prop_op(c, t, st, op)) return rc;
    HIPCHECK(launch_apply_packed_f64(op, in, out, false, st));
    return HEOMQ_OK;
  };
  return rk4_steps(xp, 2 * c->ptotal, v[0], v[1], v[2], t0, dt, nstep, st, deriv);
}

int cheby_packed_core(heomq_ctx* c, double* xp, const std::vector<double>& J, double delta, int nstep,
                      hipStream_t st, const std::vector<double*>& v) {
  DevPacked op;
  if (int rc = packed_prop_op(c, 0.0, st, op)) return rc;
  auto apply = [&](const double* in, double* out) -> int {
    HIPCHECK(launch_apply_packed_f64(op, in, out, false, st));
    return HEOMQ_OK;
  };
  std::function<hipError_t(const double*, double*, double*, double, double)> term;
  if (c->fuse_precond && apply_cheb_fusable(op))   // the term in one launch (the same switch as the solver's fusion)
    term = [&](const double* p1, double* p, double* x, double a, double b) {
      return launch_apply_packed_cheb(op, p1, p, x, a, b, st);
    };
  return cheby_steps(xp, 2 * c->ptotal, v[0], v[1], v[2], J, delta, nstep, st, apply, term);
}

}  // namespace

extern "C" {

// RK4 on the dense layout (the drop-in); a state with nothing outside the
// packed sectors is propagated on the packed layout instead (the same result
// to rounding, d^2 / (packed share) fewer bytes per apply).
int heomq_rk4_at(heomq_ctx* c, double* x, double t0, double dt, int nstep, void* stream) {
  if (!c || !x) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_rk4_at: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (nstep < 0) return fail(HEOMQ_BAD_ARG, "negative step count");
  hipStream_t st = (hipStream_t)stream;
  bool packed = false;
  if (int rc = dense_in_sector(c, x, st, packed)) return rc;
  std::vector<double*> v;
  if (packed) {
    if (int rc = carve<double>(c->ws64, 2 * c->ptotal, 4, v)) return rc;
    HIPCHECK((launch_pack<double, double>(c->pfwd, x, v[3], st)));
    if (int rc = rk4_packed_core(c, v[3], t0, dt, nstep, st, v)) return rc;
    HIPCHECK((launch_unpack<double, double>(c->pfwd, v[3], x, st)));
    return HEOMQ_OK;
  }
  const int64_t n = 2 * c->N * c->d * c->d;
  if (int rc = carve<double>(c->ws64, n, 3, v)) return rc;
  auto deriv = [&](double t, const double* in, double* out) -> int {
    DevOperator op;
    if (int rc = operator_at(c, t, st, op)) return rc;
    HIPCHECK(launch_apply_f64(op, in, out, -1, st));
    return HEOMQ_OK;
  };
  return rk4_steps(x, n, v[0], v[1], v[2], t0, dt, nstep, st, deriv);
}

int heomq_rk4(heomq_ctx* c, double* x, double dt, int nstep, void* stream) {
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_rk4: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  return heomq_rk4_at(c, x, 0.0, dt, nstep, stream);
}

int heomq_rk4_packed(heomq_ctx* c, double* xp, double t0, double dt, int nstep, void* stream) {
  if (!c || !xp) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_rk4_packed: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (nstep < 0) return fail(HEOMQ_BAD_ARG, "negative step count");
  std::vector<double*> v;
  if (int rc = carve<double>(c->ws64, 2 * c->ptotal, 3, v)) return rc;
  return rk4_packed_core(c, xp, t0, dt, nstep, (hipStream_t)stream, v);
}

// Chebyshev propagation with the reference's recurrence and weights (see
// cheby_steps and cheby_setup); like heomq_rk4_at, a state with nothing
// outside the packed sectors runs on the packed layout.
int heomq_chebyshev(heomq_ctx* c, double* x, double dt, int nstep, double* stats, void* stream) {
  if (!c || !x) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_chebyshev: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (nstep < 0) return fail(HEOMQ_BAD_ARG, "negative step count");
  hipStream_t st = (hipStream_t)stream;
  std::vector<double> J;
  double delta = 0.0;
  if (int rc = cheby_setup(c, dt, J, delta)) return rc;
  bool packed = false;
  if (int rc = dense_in_sector(c, x, st, packed)) return rc;
  std::vector<double*> v;
  if (packed) {
    if (int rc = carve<double>(c->ws64, 2 * c->ptotal, 4, v)) return rc;
    HIPCHECK((launch_pack<double, double>(c->pfwd, x, v[3], st)));
    if (int rc = cheby_packed_core(c, v[3], J, delta, nstep, st, v)) return rc;
    HIPCHECK((launch_unpack<double, double>(c->pfwd, v[3], x, st)));
  } else {
    const int64_t n = 2 * c->N * c->d * c->d;
    if (int rc = carve<double>(c->ws64, n, 3, v)) return rc;
    auto apply = [&](const double* in, double* out) -> int {
      HIPCHECK(launch_apply_f64(c->fwd.op, in, out, -1, st));
      return HEOMQ_OK;
    };
    if (int rc = cheby_steps(x, n, v[0], v[1], v[2], J, delta, nstep, st, apply)) return rc;
  }
  if (stats) {
    stats[0] = (double)J.size();
    stats[1] = delta;
    stats[2] = packed ? 1.0 : 0.0;
  }
  return HEOMQ_OK;
}

int heomq_chebyshev_packed(heomq_ctx* c, double* xp, double dt, int nstep, double* stats, void* stream) {
  if (!c || !xp) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_chebyshev_packed: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (nstep < 0) return fail(HEOMQ_BAD_ARG, "negative step count");
  hipStream_t st = (hipStream_t)stream;
  std::vector<double> J;
  double delta = 0.0;
  if (int rc = cheby_setup(c, dt, J, delta)) return rc;
  std::vector<double*> v;
  if (int rc = carve<double>(c->ws64, 2 * c->ptotal, 3, v)) return rc;
  if (int rc = cheby_packed_core(c, xp, J, delta, nstep, st, v)) return rc;
  if (stats) {
    stats[0] = (double)J.size();
    stats[1] = delta;
    stats[2] = 1.0;
  }
  return HEOMQ_OK;
}

int heomq_observables(heomq_ctx* c, const double* x, double* current, double* occ, void* stream) {
  if (!c || !x) return fail(HEOMQ_BAD_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  const int d = c->d;
  const int64_t nado = c->L >= 1 ? c->first[2] : 1;
  std::vector<double> h(2 * (size_t)nado * d * d);
  HIPCHECK(hipStreamSynchronize(st));
  HIPCHECK(hipMemcpy(h.data(), x, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  observables_host(c, h, current, occ);
  return HEOMQ_OK;
}

}  // extern "C"

// Currents and occupations from the dense tier-0 and tier-1 blocks (host copy).
void observables_host(const heomq_ctx* c, const std::vector<double>& h, double* current, double* occ) {
  const int d = c->d;
  using C = std::complex<double>;
  auto X = [&](int64_t a, int i, int j) { return C(h[2 * (a * d * d + i + j * d)], h[2 * (a * d * d + i + j * d) + 1]); };
  auto cv = [&](int orb, int s, int i, int j) { return c->ams[((size_t)(orb * c->nspin + s) * d + j) * d + i]; };
  if (current) {
    for (int t = 0; t < c->nalf * c->nspin; ++t) current[t] = 0.0;
    for (int k = 0; k < c->M; ++k) {
      if (c->sgn[k] != 0 || c->tier1_of[k] < 0) continue;
      const int64_t a = c->tier1_of[k];
      C tr(0, 0);   // tr(c X), X = x_a or fact * x_a^H
      for (int i = 0; i < d; ++i)
        for (int kk = 0; kk < d; ++kk) {
          const double cik = cv(c->orb[k], c->spin[k], i, kk);
          if (cik == 0.0) continue;
          const C xv = c->tier1_dag[k] ? (double)c->tier1_fact[k] * std::conj(X(a, i, kk)) : X(a, kk, i);
          tr += cik * xv;
        }
      // lscale: the stored tier-1 ADO is rho / |cb|^(1/2) of its own drawer
      // (k, or its sign-flipped partner when stored as the Hermitian partner)
      const int ks = c->tier1_dag[k] ? (k < c->M / 2 ? k + c->M / 2 : k - c->M / 2) : k;
      current[c->alf[k] * c->nspin + c->spin[k]] += -2.0 * tr.imag() * kJconst * c->dbsqrt[ks];   // td/calcj.f90:28-37,68
    }
  }
  if (occ)
    for (int o = 0; o < c->norbs; ++o)
      for (int s = 0; s < c->nspin; ++s) {
        double acc = 0.0;   // tr(c^+ c rho0)
        for (int i = 0; i < d; ++i)
          for (int j = 0; j < d; ++j) {
            double nij = 0.0;
            for (int k = 0; k < d; ++k) nij += cv(o, s, k, i) * cv(o, s, k, j);
            if (nij != 0.0) acc += nij * X(0, j, i).real();
          }
        occ[o * c->nspin + s] = acc;
      }
}

extern "C" {

static int host_apply(heomq_ctx* c, const double* xh, double* yh, int trace, bool trans) {
  const int64_t n = 2 * c->N * c->d * c->d;
  std::vector<double*> v;
  if (int rc = carve<double>(c->ws64, n, 2, v)) return rc;
  HIPCHECK(hipMemcpy(v[0], xh, sizeof(double) * n, hipMemcpyHostToDevice));
  HIPCHECK(launch_apply_f64(trans ? c->trn.op : c->fwd.op, v[0], v[1], tpos_of(c, trace), nullptr));
  HIPCHECK(hipMemcpy(yh, v[1], sizeof(double) * n, hipMemcpyDeviceToHost));
  return HEOMQ_OK;
}

int heomq_calc_ax_tfqmr(heomq_ctx* c, int32_t dim0, const double* x_host, double* y_host) {
  if (!c || !x_host || !y_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calc_ax_tfqmr: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (dim0 != c->d) return fail(HEOMQ_BAD_DIM, "error! dim0 != nrho in calc_ax_tfqmr_omp");
  return host_apply(c, x_host, y_host, HEOMQ_TRACE_LAST, false);
}

int heomq_calc_atrx(heomq_ctx* c, int32_t dim0, const double* x_host, double* y_host) {
  if (!c || !x_host || !y_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calc_atrx: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (dim0 != c->d) return fail(HEOMQ_BAD_DIM, "error! dim0 != nrho in calc_atrx_bicg_omp");
  return host_apply(c, x_host, y_host, HEOMQ_TRACE_LAST, true);
}

static int host_bicg(heomq_ctx* c, double* xh, double* yh, bool trans) {
  const int64_t n = 2 * c->N * c->d * c->d;
  std::vector<double*> v;
  if (int rc = carve<double>(c->ws64, n, 2, v)) return rc;
  HIPCHECK(hipMemcpy(v[0], xh, sizeof(double) * n, hipMemcpyHostToDevice));
  if (trans)
    HIPCHECK(launch_apply_transpose_bicg_f64(c->trn.op, v[0], v[1], nullptr));
  else
    HIPCHECK(launch_apply_bicg_f64(c->fwd.op, v[0], v[1], nullptr));
  HIPCHECK(hipMemcpy(xh, v[0], sizeof(double) * n, hipMemcpyDeviceToHost));   // the input is inout
  HIPCHECK(hipMemcpy(yh, v[1], sizeof(double) * n, hipMemcpyDeviceToHost));
  return HEOMQ_OK;
}

int heomq_calc_ax_bicg(heomq_ctx* c, int32_t dim0, double* x_host, double* y_host) {
  if (!c || !x_host || !y_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calc_ax_bicg: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (dim0 != c->d) return fail(HEOMQ_BAD_DIM, "error! dim0 != nrho in calc_ax_bicg_omp");
  return host_bicg(c, x_host, y_host, false);
}

int heomq_calc_atrx_bicg(heomq_ctx* c, int32_t dim0, double* x_host, double* y_host) {
  if (!c || !x_host || !y_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calc_atrx_bicg: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (dim0 != c->d) return fail(HEOMQ_BAD_DIM, "error! dim0 != nrho in calc_atrx_bicg_omp");
  return host_bicg(c, x_host, y_host, true);
}

int heomq_calcderiv(heomq_ctx* c, const double* rhotmp_host, double* rhorhs_host) {
  if (!c || !rhotmp_host || !rhorhs_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calcderiv: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  return host_apply(c, rhotmp_host, rhorhs_host, HEOMQ_TRACE_NONE, false);
}

int heomq_solve_steady_host(heomq_ctx* c, double* rho_host, double tol, int maxiter, double* stats) {
  if (!c || !rho_host) return fail(HEOMQ_BAD_ARG, "null argument");
  const int64_t n = 2 * c->N * c->d * c->d;
  HIPCHECK(c->ws_host.alloc(sizeof(double) * (size_t)n));
  double* x = c->ws_host.as<double>();
  HIPCHECK(hipMemcpy(x, rho_host, sizeof(double) * n, hipMemcpyHostToDevice));
  const int rc = heomq_solve_steady(c, x, nullptr, tol, maxiter, HEOMQ_TRACE_RDO, 0, stats, nullptr);
  if (rc != HEOMQ_OK && rc != HEOMQ_NOT_CONVERGED) return rc;
  HIPCHECK(hipMemcpy(rho_host, x, sizeof(double) * n, hipMemcpyDeviceToHost));
  return rc;
}

}  // extern "C"
