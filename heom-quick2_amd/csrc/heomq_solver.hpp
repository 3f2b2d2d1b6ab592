// Device scalars and launchers of the TFQMR / propagation vector kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "heomq_kernels.hpp"

namespace heomq {

constexpr int kMaxRedBlocks = 1024;

// Device flag values of a TFQMR problem.  0 running, 1 quasi-residual bound
// met, 2 breakdown; a (re)start goes kInitRequested -> kInitialising -> 0;
// kFrozen problems (finished members of a batch) are skipped by every kernel.
constexpr int32_t kInitRequested = 4, kInitialising = 5, kFrozen = 3;

// A batch of problems of one shape: nprob copies of every vector, vs reals apart.
struct Batch {
  int nprob = 1;
  int64_t vs = 0;
};

// Problems of several sizes advanced by one launch of each vector kernel.
// Shape g < ng holds global problems q0[g] + p, each n2[g] complex pairs long
// and stored one after another (problem p at real voff[g] + 2 p n2[g] of every
// vector).  The launch covers qstart[g + 1] - qstart[g] of them: p =
// pmap[g][j] for j = 0, 1, ... (p = j when pmap[g] is null), with blocks
// bstart[g] .. bstart[g + 1] - 1 of the flat grid, nb[g] = vec_blocks(n2[g])
// per problem; finalize block f takes the f-th covered problem.
constexpr int kMaxShapes = 16;
struct Multi {
  int ng = 0;
  int32_t bstart[kMaxShapes + 1] = {};
  int32_t nb[kMaxShapes] = {};
  int32_t qstart[kMaxShapes + 1] = {};
  int32_t q0[kMaxShapes] = {};
  int64_t n2[kMaxShapes] = {};
  int64_t voff[kMaxShapes] = {};
  const int32_t* pmap[kMaxShapes] = {};
};

// Blocks per problem of a vector kernel over n2 complex pairs.  The flat
// multi-size launch gives each problem the same count, so its reductions fold
// the same partial sums in the same order as a launch over its size alone.
int vec_blocks(int64_t n2);

// TFQMR scalars, resident on the device (Freund 1993 notation).
struct TfqmrScal {
  double rho, sigma, alpha, beta;
  double theta, tau, eta, dcoef;
  double rr, target;      // ||r0||^2, stop when tau*sqrt(2m+1) <= target
  int32_t iter, flag;     // flag: 0 running, 1 bound met, 2 breakdown
  int32_t stop;           // bound met this iteration: finish its x update, then flag = 1
  int32_t unscaled;       // batches: 1 = this problem iterates without the Jacobi scaling
  uint64_t seed;          // shadow vector of this (re)start: a new one at every restart
  // Deferred updates of the iteration fused into the operator launches: the
  // previous iteration's x += eta_a d_a + eta_b d_b with d_b = cy y2 + cd d_a
  // runs at the start of the next tfq_step1_fused (or in tfq_flush).
  double eta_a, eta_b, cy, cd;
};

// BiCG scalars (bicg/solvebicg.f90 notation: rho = <s, r>, alpha, beta).
struct BicgScal {
  double rho, alpha, beta, rr, crit;
  int32_t iter, flag;     // flag: 0 running, 1 ||r|| <= crit, 2 breakdown
};

// BiCG vector steps on complex vectors of n reals (each followed by its finalize):
//   bicg_init:  r = e - q (e: 1 at real element e_off), s = d = f = r;  ||r||^2
//   bicg_alpha: <f, q> -> alpha = rho / <f, q>
//   bicg_update: x += alpha d, r -= alpha q, s -= alpha qt;  ||r||^2, <s, r> -> beta
//   bicg_dirs:  d = r + beta d, f = s + beta f
hipError_t bicg_init(hipStream_t, int64_t n, const double* q, int64_t e_off, double* r, double* s, double* d,
                     double* f, BicgScal* sc, double crit, double* partial);
hipError_t bicg_alpha(hipStream_t, int64_t n, const double* f, const double* q, BicgScal* sc, double* partial);
hipError_t bicg_update(hipStream_t, int64_t n, double* x, const double* d, double* r, const double* q, double* s,
                       const double* qt, BicgScal* sc, double* partial);
hipError_t bicg_dirs(hipStream_t, int64_t n, const double* r, const double* s, double* d, double* f,
                     const BicgScal* sc);

template <typename T>
hipError_t vec_project(hipStream_t, int64_t nado, int d, T* x, const int8_t* par, const TierBounds& tb,
                       bool hermitian);

// One TFQMR iteration (Freund 1993) is: tfq_step1, A y2 -> u2, tfq_step2,
// A w -> aw, tfq_v.  The operator is applied to w rather than to y1 = w +
// beta y2 (A y1 = aw + beta u2), so y1 and u1 = A y1 are formed inside
// tfq_step1 and tfq_v and never stored.  A (re)start is tfq_init_r (r = b -
// A x into w; y2 = d = 0), A w -> ax, tfq_init_v (v = ax, aw = ax when aw is
// given; u2 = 0).  The first two reductions of an iteration are folded by the
// kernels that consume them (tfq_step2 folds tfq_step1's partial sums and
// keeps the half step's scalars in a second copy s2; tfq_v folds tfq_step2's
// partial sums in partial2 and writes the scalars back to s); tfq_v's own
// reduction is followed by a one-block finalize.  The shadow vector is
// generated in the kernels.
template <typename T>
hipError_t tfq_init_r(hipStream_t, int64_t n, const T* b, const T* ax, T* w, T* y2, T* d, TfqmrScal* s,
                      double* partial, Batch bt = Batch{});
template <typename T>
hipError_t tfq_init_v(hipStream_t, int64_t n, const T* ax, T* v, T* aw, T* u2, TfqmrScal* s, double* partial,
                      Batch bt = Batch{});
template <typename T>
hipError_t tfq_step1(hipStream_t, int64_t n, T* w, T* y2, const T* v, const T* aw, const T* u2, T* d,
                     TfqmrScal* s, double* partial, Batch bt = Batch{});
template <typename T>
hipError_t tfq_step2(hipStream_t, int64_t n, T* x, T* w, const T* u2, T* d, const T* y2, TfqmrScal* s,
                     TfqmrScal* s2, const double* partial, double* partial2, Batch bt = Batch{});
template <typename T>
hipError_t tfq_v(hipStream_t, int64_t n, const T* aw, const T* u2, T* v, T* x, const T* d, TfqmrScal* s,
                 const TfqmrScal* s2, double* partial, const double* partial2, Batch bt = Batch{});

// The same iteration with the vector work of tfq_step2 and tfq_v moved into
// the two block-Jacobi-preconditioned operator launches and the next
// iteration's first kernel (FP64, one problem):
//   tfq_step1_fused   the previous iteration's deferred x update, d_b and v
//                     (v = aw + beta (2 u2 + beta v)), then step 1; partials
//                     of ||w||^2 and rt.v (vector grid, pv)
//   M^-1 A y2 -> u2   with w -= alpha u2 and partials of ||w||^2, rt.w,
//                     rt.u2 per operator workgroup (pa: three arrays of na)
//   M^-1 A w -> aw    with partials of rt.aw (pa + 3 na, one per block), and
//                     the first launch's sums in runs of 16 (pa + 4 na)
//   tfq_fold          both half steps, the stop test, beta, and
//                     sigma = rt.v_new = rt.aw + beta (2 rt.u2 + beta rt.v)
//                     from the stored v's direct sum, alpha = rho / sigma.
// x lags by the deferred update between iterations: tfq_flush completes it
// (a no-op when nothing is pending) before x is read.
hipError_t tfq_step1_fused(hipStream_t, int64_t n, double* w, double* y2, double* v, const double* aw,
                           const double* u2, double* d, double* x, const TfqmrScal* s, double* pv);
hipError_t tfq_fold(hipStream_t, int64_t n, const double* pv, const double* pa, int na, TfqmrScal* s);
hipError_t tfq_flush(hipStream_t, int64_t n, double* x, const double* d, const double* y2, TfqmrScal* s);

template <typename T>
hipError_t vec_fill(hipStream_t, int64_t n, T* v, T val);
template <typename T>
hipError_t vec_randn(hipStream_t, int64_t n, T* v, uint64_t seed);
template <typename TB, typename TA>
hipError_t vec_resid(hipStream_t, int64_t n, const TB* b, const TA* ax, double* out, double* partial,
                     Batch bt = Batch{});
hipError_t vec_resid_to_f32(hipStream_t, int64_t n, const double* b, const double* ax, float* r32, double scale);
hipError_t vec_axpy_f32_to_f64(hipStream_t, int64_t n, double* x, const float* dx, double alpha);
// fresh: the first start of a solve (shadow seed first_seed); otherwise a restart (next seed)
hipError_t tfq_reset(hipStream_t, TfqmrScal* s, double target, const int32_t* mask = nullptr, int nprob = 1,
                     bool fresh = true, uint64_t first_seed = 0);
hipError_t tfq_freeze(hipStream_t, TfqmrScal* s, const int32_t* mask, int nprob);
// unscaled[p] = mask[p] (0 for all with a null mask)
hipError_t tfq_set_unscaled(hipStream_t, TfqmrScal* s, const int32_t* mask, int nprob);
// target *= sqrt(rr / rr_true[2p]) for problems in kInitialising (scaled iterations)
hipError_t tfq_scale_targets(hipStream_t, TfqmrScal* s, const double* rr_true, int nprob);
// The same steps over every problem of a Multi layout (FP64; partial holds
// two doubles per flat-grid block, out two per problem).
hipError_t tfq_init_r(hipStream_t, const Multi&, const double* b, const double* ax, double* w, double* y2, double* d,
                      TfqmrScal* s, double* partial);
hipError_t tfq_init_v(hipStream_t, const Multi&, const double* ax, double* v, double* aw, double* u2,
                      TfqmrScal* s, double* partial);
hipError_t tfq_step1(hipStream_t, const Multi&, double* w, double* y2, const double* v, const double* aw,
                     const double* u2, double* d, TfqmrScal* s, double* partial);
hipError_t tfq_step2(hipStream_t, const Multi&, double* x, double* w, const double* u2, double* d, const double* y2,
                     TfqmrScal* s, TfqmrScal* s2, const double* partial, double* partial2);
hipError_t tfq_v(hipStream_t, const Multi&, const double* aw, const double* u2, double* v, double* x,
                 const double* d, TfqmrScal* s, const TfqmrScal* s2, double* partial, const double* partial2);
hipError_t vec_resid(hipStream_t, const Multi&, const double* b, const double* ax, double* out, double* partial);
hipError_t vec_set_elem(hipStream_t, double* v, int64_t off, double val);
// A few scattered values written into every problem's copy of a vector.
struct SetPattern {
  int m = 0;
  int64_t idx[64];
  double val[64];
};
hipError_t vec_set_pattern(hipStream_t, double* v, int64_t stride, int nprob, const SetPattern& pat);
hipError_t vec_lincomb3(hipStream_t, int64_t n, double* out, double a, const double* x, double b,
                        const double* y, double c, const double* z);

}  // namespace heomq
