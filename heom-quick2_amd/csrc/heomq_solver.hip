// Vector kernels of the device-resident TFQMR steady-state solver, the RK4 and
// Chebyshev propagators.  The solver keeps every vector and every scalar on
// the GPU: each fused kernel does its vector updates (16-byte complex-pair
// accesses) and writes per-workgroup FP64 partial sums; a one-block finalize
// kernel folds the partials in a fixed order and advances the TFQMR scalars.
// A whole iteration is launch-only (no host round trip) and deterministic.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "heomq_shadow.hpp"
#include "heomq_solver.hpp"

namespace heomq {

constexpr int kRedThreads = 256;

// Block-level sum of (a, b); result valid in thread 0.
__device__ inline void block_sum2(double& a, double& b) {
  __shared__ double sa[kRedThreads / 64], sb[kRedThreads / 64];
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_down(a, off, 64);
    b += __shfl_down(b, off, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) { sa[wid] = a; sb[wid] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = 0; b = 0;
    for (int t = 0; t < kRedThreads / 64; ++t) { a += sa[t]; b += sb[t]; }
  }
}

// Phase 1 of every reduction: each block writes its (a, b) partial sums.
__device__ inline void block_partial(double a, double b, double* partial) {
  block_sum2(a, b);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = b;
  }
}

// Vector kernels walk complex pairs (16-byte accesses).
template <typename T> struct W2;
template <> struct W2<double> { using type = double2; };
template <> struct W2<float> { using type = float2; };

#define GRID_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// Where a block of a TFQMR vector kernel works: its problem q (the index of
// the problem's scalars), the problem's first real in every vector, its
// complex-pair count, the block's rank among the problem's blocks, and the
// slot of the block's partial sums.
struct Span {
  int64_t q, off, n2, blk, nblk, pidx;
};

// Problems of one size: blockIdx.y is the problem, its vectors start vs reals
// after the previous problem's; gridDim.x blocks per problem.
struct UniLoc {
  int64_t n2, vs;
  __device__ Span span() const {
    const int64_t q = blockIdx.y;
    return Span{q, q * vs, n2, (int64_t)blockIdx.x, (int64_t)gridDim.x, q * gridDim.x + blockIdx.x};
  }
  // finalize block q: problem q and its partial sums
  __device__ int64_t fin(int64_t q, int64_t nblk_uni, int64_t& pbase, int64_t& nblk) const {
    pbase = q * nblk_uni;
    nblk = nblk_uni;
    return q;
  }
};

// Problems of several sizes in one flat grid (the shape loop is over at most
// kMaxShapes entries of the kernel argument, block-uniform).
struct MultiLoc {
  Multi m;
  __device__ int shape_of_block(int64_t b) const {
    int g = 0;
    while (g + 1 < m.ng && m.bstart[g + 1] <= b) ++g;
    return g;
  }
  __device__ Span span() const {
    const int64_t b = blockIdx.x;
    const int g = shape_of_block(b);
    const int64_t loc = b - m.bstart[g], j = loc / m.nb[g];
    const int64_t p = m.pmap[g] ? m.pmap[g][j] : j;
    return Span{m.q0[g] + p, m.voff[g] + 2 * p * m.n2[g], m.n2[g], loc - j * m.nb[g], m.nb[g], b};
  }
  // finalize block f: the covered problem's global index and partial sums
  __device__ int64_t fin(int64_t f, int64_t, int64_t& pbase, int64_t& nblk) const {
    int g = 0;
    while (g + 1 < m.ng && m.qstart[g + 1] <= f) ++g;
    const int64_t j = f - m.qstart[g];
    pbase = m.bstart[g] + j * m.nb[g];
    nblk = m.nb[g];
    return m.q0[g] + (m.pmap[g] ? m.pmap[g][j] : j);
  }
};

#define SPAN_LOOP(i, sp) \
  for (int64_t i = (sp).blk * blockDim.x + threadIdx.x; i < (sp).n2; i += (sp).nblk * blockDim.x)

// Phase 1 of a span's reduction: the block's (a, b) partial sums in its slot.
__device__ inline void span_partial(double a, double b, double* partial, const Span& sp) {
  block_sum2(a, b);
  if (threadIdx.x == 0) {
    partial[2 * sp.pidx] = a;
    partial[2 * sp.pidx + 1] = b;
  }
}

// A reduction's phase 2 folded into the kernel that consumes it: every
// workgroup sums its problem's partial sums (nblk pairs) in the order of
// k_finalize's 256-thread fold, so every workgroup (and the separate
// finalize it replaces) gets the same bits.  Every thread receives (a, b).
__device__ inline void fold_partials(const double* __restrict__ partial, int64_t nblk, double& a, double& b) {
  __shared__ double fa, fb;
  double x = 0, y = 0;
  for (int64_t t = threadIdx.x; t < nblk; t += blockDim.x) {
    x += partial[2 * t];
    y += partial[2 * t + 1];
  }
  block_sum2(x, y);
  if (threadIdx.x == 0) {
    fa = x;
    fb = y;
  }
  __syncthreads();
  a = fa;
  b = fb;
}


// One quasi-minimisation half step from ||w||^2 = a.
__device__ inline void half_step(TfqmrScal* s, double a) {
  const double theta = sqrt(a) / s->tau;
  const double c = 1.0 / sqrt(1.0 + theta * theta);
  s->tau = s->tau * theta * c;
  s->theta = theta;
  s->eta = c * c * s->alpha;
  s->dcoef = theta * theta * s->eta / s->alpha;
}

// The second half step of an iteration, then rho_new = rt.w = b: the bound
// met sets stop (the x update of the iteration still runs, k_v; the final
// finalize then stops), rho = 0 is a breakdown.
__device__ inline void half_step_rho(TfqmrScal* s, double a, double b) {
  half_step(s, a);
  s->iter += 1;
  if (s->tau * sqrt(2.0 * s->iter + 1.0) <= s->target) {
    s->stop = 1;
    return;
  }
  if (s->rho == 0.0) {
    s->flag = 2;
    return;
  }
  s->beta = b / s->rho;
  s->rho = b;
}

// The iteration is kept in a form that applies the operator to w instead of
// y1 = w + beta y2: A y1 = A w + beta u2 (u2 = A y2), so y1 and u1 = A y1 are
// formed where they are read and never stored.  One iteration is
//   k_step1, A y2 -> u2, k_step2, A w -> aw, k_v
// (Freund 1993's TFQMR; the same iterates as the y1 form in exact
// arithmetic).  The state between iterations is w, y2, u2, aw, v, d, x.

// r = b - Ax (stored in w), y2 = 0, d = 0; partials of ||r||^2 and rt.r.
// With beta = 0 from the reset, the first k_step1 sees y1 = w = r.
template <typename T, typename Loc>
__global__ void k_init_r(const Loc loc, const T* __restrict__ b, const T* __restrict__ ax, T* __restrict__ w,
                         T* __restrict__ y2, T* __restrict__ d, const TfqmrScal* s, double* partial) {
  using V = typename W2<T>::type;
  const Span sp = loc.span();
  s += sp.q;
  if (s->flag != kInitRequested) return;   // partials of this problem are not read
  const uint64_t seed = s->seed;
  { const int64_t o = sp.off; b += o; ax += o; w += o; y2 += o; d += o; }
  double rr = 0, rrt = 0;
  SPAN_LOOP(i, sp) {
    const V bb = reinterpret_cast<const V*>(b)[i], aa = reinterpret_cast<const V*>(ax)[i];
    V r;
    r.x = bb.x - aa.x;
    r.y = bb.y - aa.y;
    reinterpret_cast<V*>(w)[i] = r;
    V z;
    z.x = 0;
    z.y = 0;
    reinterpret_cast<V*>(y2)[i] = z;
    reinterpret_cast<V*>(d)[i] = z;
    rr += (double)r.x * r.x + (double)r.y * r.y;
    rrt += shadow(2 * i, seed) * r.x + shadow(2 * i + 1, seed) * r.y;
  }
  span_partial(rr, rrt, partial, sp);
}

// v = ax (= A r), aw = ax when aw is given (a batch restarts some problems
// while the others keep their aw), u2 = 0; partial of rt.v
template <typename T, typename Loc>
__global__ void k_init_v(const Loc loc, const T* __restrict__ ax, T* __restrict__ v, T* __restrict__ aw,
                         T* __restrict__ u2, const TfqmrScal* s, double* partial) {
  using V = typename W2<T>::type;
  const Span sp = loc.span();
  s += sp.q;
  if (s->flag != kInitialising) return;
  const uint64_t seed = s->seed;
  { const int64_t o = sp.off; ax += o; v += o; if (aw) aw += o; u2 += o; }
  double sg = 0;
  SPAN_LOOP(i, sp) {
    const V t = reinterpret_cast<const V*>(ax)[i];
    reinterpret_cast<V*>(v)[i] = t;
    if (aw) reinterpret_cast<V*>(aw)[i] = t;
    V z;
    z.x = 0;
    z.y = 0;
    reinterpret_cast<V*>(u2)[i] = z;
    sg += shadow(2 * i, seed) * t.x + shadow(2 * i + 1, seed) * t.y;
  }
  span_partial(sg, 0.0, partial, sp);
}

// First half of an iteration, fused with the next search direction:
// y1 = w + beta y2, u1 = aw + beta u2;  y2 = y1 - alpha v;  w -= alpha u1;
// d = y1 + dcoef d;  partial ||w||^2.
template <typename T, typename Loc>
__global__ void k_step1(const Loc loc, T* __restrict__ w, T* __restrict__ y2, const T* __restrict__ v,
                        const T* __restrict__ aw, const T* __restrict__ u2, T* __restrict__ d, const TfqmrScal* s,
                        double* partial) {
  using V = typename W2<T>::type;
  const Span sp = loc.span();
  s += sp.q;
  if (s->flag) return;
  { const int64_t o = sp.off; w += o; y2 += o; v += o; aw += o; u2 += o; d += o; }
  const T alpha = (T)s->alpha, dcoef = (T)s->dcoef, beta = (T)s->beta;
  double ww = 0;
  SPAN_LOOP(i, sp) {
    V wi = reinterpret_cast<V*>(w)[i];
    const V yo = reinterpret_cast<const V*>(y2)[i];
    V y1;
    y1.x = wi.x + beta * yo.x;
    y1.y = wi.y + beta * yo.y;
    const V a = reinterpret_cast<const V*>(aw)[i], u = reinterpret_cast<const V*>(u2)[i];
    V u1;
    u1.x = a.x + beta * u.x;
    u1.y = a.y + beta * u.y;
    const V b = reinterpret_cast<const V*>(v)[i];
    V o;
    o.x = y1.x - alpha * b.x;
    o.y = y1.y - alpha * b.y;
    reinterpret_cast<V*>(y2)[i] = o;
    wi.x -= alpha * u1.x;
    wi.y -= alpha * u1.y;
    reinterpret_cast<V*>(w)[i] = wi;
    const V di = reinterpret_cast<const V*>(d)[i];
    V dn;
    dn.x = y1.x + dcoef * di.x;
    dn.y = y1.y + dcoef * di.y;
    reinterpret_cast<V*>(d)[i] = dn;
    ww += (double)wi.x * wi.x + (double)wi.y * wi.y;
  }
  span_partial(ww, 0.0, partial, sp);
}

// Second half, with the first half step's scalars folded in: (theta, tau,
// eta, dcoef) from step 1's partial sums of ||w||^2 (part_in), written to the
// problem's second scalar copy s_out by its first workgroup (s_out mirrors s
// for a stopped or frozen problem); then x += eta d;  w -= alpha u2;
// d = y2 + dcoef d;  partials of ||w||^2 and rt.w (w is final for this
// iteration).
template <typename T, typename Loc>
__global__ void k_step2(const Loc loc, T* __restrict__ x, T* __restrict__ w, const T* __restrict__ u2,
                        T* __restrict__ d, const T* __restrict__ y2, const TfqmrScal* s, TfqmrScal* s_out,
                        const double* __restrict__ part_in, double* partial) {
  using V = typename W2<T>::type;
  const Span sp = loc.span();
  s += sp.q;
  s_out += sp.q;
  TfqmrScal t = *s;
  if (t.flag) {   // uniform over the problem's workgroups
    if (sp.blk == 0 && threadIdx.x == 0) *s_out = t;
    return;
  }
  double a, unused;
  fold_partials(part_in + 2 * (sp.pidx - sp.blk), sp.nblk, a, unused);
  half_step(&t, a);
  if (sp.blk == 0 && threadIdx.x == 0) *s_out = t;
  const uint64_t seed = t.seed;
  { const int64_t o = sp.off; x += o; w += o; u2 += o; d += o; y2 += o; }
  const T alpha = (T)t.alpha, dcoef = (T)t.dcoef, eta = (T)t.eta;
  double ww = 0, rw = 0;
  SPAN_LOOP(i, sp) {
    const V di = reinterpret_cast<const V*>(d)[i];
    V xi = reinterpret_cast<V*>(x)[i];
    xi.x += eta * di.x;
    xi.y += eta * di.y;
    reinterpret_cast<V*>(x)[i] = xi;
    V wi = reinterpret_cast<V*>(w)[i];
    const V ui = reinterpret_cast<const V*>(u2)[i];
    wi.x -= alpha * ui.x;
    wi.y -= alpha * ui.y;
    reinterpret_cast<V*>(w)[i] = wi;
    const V yi = reinterpret_cast<const V*>(y2)[i];
    V dn;
    dn.x = yi.x + dcoef * di.x;
    dn.y = yi.y + dcoef * di.y;
    reinterpret_cast<V*>(d)[i] = dn;
    ww += (double)wi.x * wi.x + (double)wi.y * wi.y;
    rw += shadow(2 * i, seed) * wi.x + shadow(2 * i + 1, seed) * wi.y;
  }
  span_partial(ww, rw, partial, sp);
}

// The second half step's scalars folded in: from step 2's partial sums
// (part_in) and the scalar copy step 2 wrote (s_in), the half step, the stop
// test and beta, written back to the problem's scalars s_out by its first
// workgroup.  Then x += eta d (also on the iteration that met the bound);
// v = u1 + beta (u2 + beta v) = aw + beta (2 u2 + beta v) with u1 = A y1 =
// aw + beta u2; partial of rt.v.
template <typename T, typename Loc>
__global__ void k_v(const Loc loc, const T* __restrict__ aw, const T* __restrict__ u2, T* __restrict__ v,
                    T* __restrict__ x, const T* __restrict__ d, const TfqmrScal* s_in, TfqmrScal* s_out,
                    const double* __restrict__ part_in, double* partial) {
  using V = typename W2<T>::type;
  const Span sp = loc.span();
  s_in += sp.q;
  s_out += sp.q;
  TfqmrScal t = *s_in;
  if (t.flag) return;   // s_out already holds this state
  double a, b;
  fold_partials(part_in + 2 * (sp.pidx - sp.blk), sp.nblk, a, b);
  half_step_rho(&t, a, b);
  if (sp.blk == 0 && threadIdx.x == 0) *s_out = t;
  if (t.flag) return;   // breakdown: no update (uniform over the problem's workgroups)
  const uint64_t seed = t.seed;
  { const int64_t o = sp.off; aw += o; u2 += o; v += o; x += o; d += o; }
  const T beta = (T)t.beta, eta = (T)t.eta;
  const bool next = t.stop == 0;
  double sg = 0;
  SPAN_LOOP(i, sp) {
    V xi = reinterpret_cast<V*>(x)[i];
    const V di = reinterpret_cast<const V*>(d)[i];
    xi.x += eta * di.x;
    xi.y += eta * di.y;
    reinterpret_cast<V*>(x)[i] = xi;
    if (next) {
      const V a2 = reinterpret_cast<const V*>(aw)[i], b2 = reinterpret_cast<const V*>(u2)[i];
      const V c = reinterpret_cast<const V*>(v)[i];
      V r;
      r.x = a2.x + beta * ((T)2 * b2.x + beta * c.x);
      r.y = a2.y + beta * ((T)2 * b2.y + beta * c.y);
      reinterpret_cast<V*>(v)[i] = r;
      sg += shadow(2 * i, seed) * r.x + shadow(2 * i + 1, seed) * r.y;
    }
  }
  span_partial(sg, 0.0, partial, sp);
}

// ---------------------------------------------------------------------------
// The iteration fused into the operator launches (FP64, one problem; the
// operator side is k_apply_packed_jac's TFQMR epilogues)
// ---------------------------------------------------------------------------

// The previous iteration's deferred work, then step 1: d_b = cy y2 + cd d;
// x += eta_a d + eta_b d_b; v = aw + beta (2 u2 + beta v); y1 = w + beta y2,
// u1 = aw + beta u2; y2 = y1 - alpha v; w -= alpha u1; d = y1 + dcoef d_b;
// partials of ||w||^2 and rt.v.  After a (re)start beta = 0 and v = aw, so
// the v update rewrites the same values.
__global__ void k_step1_fused(int64_t n2, double* __restrict__ w, double* __restrict__ y2, double* __restrict__ v,
                              const double* __restrict__ aw, const double* __restrict__ u2, double* __restrict__ d,
                              double* __restrict__ x, const TfqmrScal* s, double* partial) {
  const TfqmrScal t = *s;
  if (t.flag) return;
  const double alpha = t.alpha, beta = t.beta, dcoef = t.dcoef, ea = t.eta_a, eb = t.eta_b, cy = t.cy, cd = t.cd;
  const uint64_t seed = t.seed;
  double ww = 0, sv = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 yo = reinterpret_cast<const double2*>(y2)[i];
    const double2 da = reinterpret_cast<const double2*>(d)[i];
    double2 db;
    db.x = cy * yo.x + cd * da.x;
    db.y = cy * yo.y + cd * da.y;
    double2 xi = reinterpret_cast<double2*>(x)[i];
    xi.x += ea * da.x + eb * db.x;
    xi.y += ea * da.y + eb * db.y;
    reinterpret_cast<double2*>(x)[i] = xi;
    const double2 a = reinterpret_cast<const double2*>(aw)[i], u = reinterpret_cast<const double2*>(u2)[i];
    double2 vi = reinterpret_cast<double2*>(v)[i];
    vi.x = a.x + beta * (2.0 * u.x + beta * vi.x);
    vi.y = a.y + beta * (2.0 * u.y + beta * vi.y);
    reinterpret_cast<double2*>(v)[i] = vi;
    sv += shadow(2 * i, seed) * vi.x + shadow(2 * i + 1, seed) * vi.y;
    double2 wi = reinterpret_cast<double2*>(w)[i];
    double2 y1;
    y1.x = wi.x + beta * yo.x;
    y1.y = wi.y + beta * yo.y;
    double2 o;
    o.x = y1.x - alpha * vi.x;
    o.y = y1.y - alpha * vi.y;
    reinterpret_cast<double2*>(y2)[i] = o;
    wi.x -= alpha * (a.x + beta * u.x);
    wi.y -= alpha * (a.y + beta * u.y);
    reinterpret_cast<double2*>(w)[i] = wi;
    double2 dn;
    dn.x = y1.x + dcoef * db.x;
    dn.y = y1.y + dcoef * db.y;
    reinterpret_cast<double2*>(d)[i] = dn;
    ww += wi.x * wi.x + wi.y * wi.y;
  }
  block_partial(ww, sv, partial);
}

// One workgroup of kFoldThreads: the iteration's reductions in a fixed order,
// both half steps, the stop test, beta, sigma and alpha.  pv: (||w||^2, rt.v)
// of k_step1_fused's nv blocks, interleaved; pa: ||w||^2, rt.w and rt.u2 of
// the first operator launch's na blocks and rt.aw of the second's, each an
// array of na, then the first three summed in runs of 16 blocks by the second
// launch (three arrays of ceil(na / 16)), which is what the fold reads.  sigma = rt.(aw + beta (2 u2 + beta v)) from the sums over aw,
// u2 and the stored v: the v the next k_step1_fused forms, one step from a
// directly summed rt.v.
constexpr int kFoldThreads = 1024;
__global__ void __launch_bounds__(kFoldThreads) k_tfq_fold(const double* __restrict__ pv, int nv,
                                                           const double* __restrict__ pa, int na, TfqmrScal* s) {
  if (s->flag) return;
  double v[6] = {0, 0, 0, 0, 0, 0};   // ||w||^2 and rt.v of step 1; ||w||^2, rt.w, rt.u2, rt.aw
  for (int t = threadIdx.x; t < nv; t += kFoldThreads) {
    v[0] += pv[2 * t];
    v[1] += pv[2 * t + 1];
  }
  // the first launch's three sums, already summed in runs of 16 workgroups by
  // the second launch (ng per array), then the second's rt.aw (na)
  const int ng = (na + 15) >> 4;
  const double* q = pa + 4 * (int64_t)na;
  for (int t = threadIdx.x; t < ng; t += kFoldThreads)
#pragma unroll
    for (int k = 0; k < 3; ++k) v[2 + k] += q[(int64_t)k * ng + t];
  const double* paw = pa + 3 * (int64_t)na;
  for (int t0 = threadIdx.x; t0 < na; t0 += 4 * kFoldThreads) {
    double l[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) l[u] = t0 + u * kFoldThreads < na ? paw[t0 + u * kFoldThreads] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) v[5] += l[u];
  }
  __shared__ double red[6][kFoldThreads / 64];
#pragma unroll
  for (int k = 0; k < 6; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][wid] = v[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    v[k] = 0;
    for (int q = 0; q < kFoldThreads / 64; ++q) v[k] += red[k][q];
  }
  const double a1 = v[0], sv = v[1], a2 = v[2], rw = v[3], ru = v[4], ra = v[5];
  TfqmrScal t = *s;
  half_step(&t, a1);
  t.eta_a = t.eta;
  t.cy = 1.0;
  t.cd = t.dcoef;
  half_step_rho(&t, a2, rw);
  t.eta_b = t.flag ? 0.0 : t.eta;   // a breakdown keeps x at the first half step
  if (t.stop) {
    t.flag = 1;
  } else if (!t.flag) {
    const double sg = ra + t.beta * (2.0 * ru + t.beta * sv);
    t.sigma = sg;
    if (sg == 0.0) {
      t.flag = 2;
    } else {
      t.alpha = t.rho / sg;
      t.dcoef = t.theta * t.theta * t.eta / t.alpha;
    }
  }
  *s = t;
}

// x += eta_a d + eta_b (cy y2 + cd d): the pending update of the last fused
// iteration; k_tfq_flush_scal then clears eta_a and eta_b, so that a second
// flush or the next k_step1_fused adds nothing more to x.
__global__ void k_tfq_flush(int64_t n2, double* __restrict__ x, const double* __restrict__ d,
                            const double* __restrict__ y2, const TfqmrScal* s) {
  const double ea = s->eta_a, eb = s->eta_b, cy = s->cy, cd = s->cd;
  if (ea == 0.0 && eb == 0.0) return;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x) {
    const double2 da = reinterpret_cast<const double2*>(d)[i], yo = reinterpret_cast<const double2*>(y2)[i];
    double2 xi = reinterpret_cast<double2*>(x)[i];
    xi.x += ea * da.x + eb * (cy * yo.x + cd * da.x);
    xi.y += ea * da.y + eb * (cy * yo.y + cd * da.y);
    reinterpret_cast<double2*>(x)[i] = xi;
  }
}
__global__ void k_tfq_flush_scal(TfqmrScal* s) {
  if (threadIdx.x == 0) {
    s->eta_a = 0;
    s->eta_b = 0;
  }
}

// partials of ||b - ax||^2 and ||b||^2 (FP64)
template <typename TB, typename TA, typename Loc>
__global__ void k_resid(const Loc loc, const TB* __restrict__ b, const TA* __restrict__ ax, double* partial) {
  const Span sp = loc.span();
  { const int64_t o = sp.off; b += o; ax += o; }
  double rr = 0, bb = 0;
  SPAN_LOOP(i, sp) {
    const double b0 = (double)b[2 * i], b1 = (double)b[2 * i + 1];
    const double r0 = b0 - (double)ax[2 * i], r1 = b1 - (double)ax[2 * i + 1];
    rr += r0 * r0 + r1 * r1;
    bb += b0 * b0 + b1 * b1;
  }
  span_partial(rr, bb, partial, sp);
}

// Phase 2: one block folds the partials in a fixed order and advances the
// scalars (Freund 1993; Saad, Iterative Methods 7.4.2).
enum FinalizeOp { F_INIT_R = 0, F_INIT_V, F_HALF, F_HALF_RHO, F_V, F_RESID };


// Block q finalizes problem q (one block for a single problem; the covered
// problems of a multi-size launch in order); nblk_uni is the blocks per
// problem of a one-size launch.
template <typename Loc>
__global__ void k_finalize(const Loc loc, int op, int nblk_uni, const double* __restrict__ partial, TfqmrScal* s,
                           double* out) {
  int64_t pbase = 0, nblk = 0;
  const int64_t q = loc.fin(blockIdx.x, nblk_uni, pbase, nblk);
  partial += 2 * pbase;
  if (out) out += 2 * q;
  if (s) s += q;
  if (op == F_INIT_R && s->flag != kInitRequested) return;
  if (op == F_INIT_V && s->flag != kInitialising) return;
  if (op != F_INIT_R && op != F_INIT_V && op != F_RESID && s->flag) return;   // s unused for F_RESID
  double a = 0, b = 0;
  for (int64_t t = threadIdx.x; t < nblk; t += blockDim.x) {
    a += partial[2 * t];
    b += partial[2 * t + 1];
  }
  block_sum2(a, b);
  if (threadIdx.x != 0) return;
  switch (op) {
    case F_INIT_R:
      s->rr = a;
      s->tau = sqrt(a);
      s->rho = b;
      s->theta = 0;
      s->eta = 0;
      s->dcoef = 0;
      s->stop = 0;
      s->eta_a = s->eta_b = 0;
      s->cy = 0;
      s->cd = 1;
      s->flag = (b == 0.0 && a > 0.0) ? 2 : kInitialising;
      break;
    case F_INIT_V:
      s->sigma = a;
      if (a == 0.0) { s->flag = 2; break; }
      s->alpha = s->rho / a;
      s->dcoef = 0;
      s->flag = 0;
      break;
    case F_HALF:
      half_step(s, a);
      break;
    case F_HALF_RHO:   // second half step, then rho_new = rt.w = b
      half_step_rho(s, a, b);
      break;
    case F_V:
      if (s->stop) { s->flag = 1; break; }
      s->sigma = a;
      if (a == 0.0) { s->flag = 2; break; }
      s->alpha = s->rho / a;
      s->dcoef = s->theta * s->theta * s->eta / s->alpha;
      break;
    case F_RESID:
      out[0] = a;
      out[1] = b;
      break;
  }
}

// r = b - ax in FP64, scaled into FP32: r32 = r * scale.
__global__ void k_resid_to_f32(int64_t n, const double* __restrict__ b, const double* __restrict__ ax,
                               float* __restrict__ r32, double scale) {
  GRID_LOOP(i, n) r32[i] = (float)(((double)b[i] - ax[i]) * scale);
}

// x += alpha * dx (FP64 += FP32)
__global__ void k_axpy_f32_to_f64(int64_t n, double* __restrict__ x, const float* __restrict__ dx, double alpha) {
  GRID_LOOP(i, n) x[i] += alpha * (double)dx[i];
}

template <typename T>
__global__ void k_fill(int64_t n, T* __restrict__ v, T val) {
  GRID_LOOP(i, n) v[i] = val;
}

// counter-based standard normals (splitmix64 + Box-Muller), same bits for any grid
template <typename T>
__global__ void k_randn(int64_t n, T* __restrict__ v, uint64_t seed) {
  GRID_LOOP(i, n) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u1 = ((z >> 11) + 1.0) * (1.0 / 9007199254740993.0);
    const double u2 = ((z & 0x7FFull) + 0.5) / 2048.0;
    v[i] = (T)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  }
}

// generic: out = a*x + b*y (+ c*z)
__global__ void k_lincomb3(int64_t n, double* __restrict__ out, double a, const double* __restrict__ x,
                           double b, const double* __restrict__ y, double c, const double* __restrict__ z) {
  GRID_LOOP(i, n) {
    double t = a * x[i];
    if (y) t += b * y[i];
    if (z) t += c * z[i];
    out[i] = t;
  }
}

// Projection onto the physical sector: zero the elements a tier-n ADO cannot
// hold (parity(i) + parity(j) + n odd) and Hermitianise the tier-0 block.
template <typename T>
__global__ void k_project_parity(int64_t nado, int d, T* __restrict__ x, const int8_t* __restrict__ par,
                                 TierBounds tb) {
  const int64_t dd = (int64_t)d * d;
  GRID_LOOP(idx, nado * dd) {
    const int64_t a = idx / dd;
    const int e = (int)(idx - a * dd);
    const int i = e % d, j = e / d;
    int n = 0;
    while (n + 1 < tb.ntier && tb.first[n + 1] <= a) ++n;
    if ((par[i] + par[j] + n) & 1) {
      x[2 * idx] = 0;
      x[2 * idx + 1] = 0;
    }
  }
}

template <typename T>
__global__ void k_hermitize0(int d, T* __restrict__ x) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d * d) return;
  const int i = e % d, j = e / d;
  if (i > j) return;
  const int t = j + i * d;
  const T re = (T)0.5 * (x[2 * e] + x[2 * t]);
  const T im = (T)0.5 * (x[2 * e + 1] - x[2 * t + 1]);
  x[2 * e] = re;
  x[2 * e + 1] = im;
  x[2 * t] = re;
  x[2 * t + 1] = -im;
}

// Scalar state reset on the device (no host staging of pageable memory).
// Problems p < nprob with mask[p] != 0 (every problem if mask is null) are set
// up for a (re)start: scalars cleared, flag = kInitRequested.  Problems left
// out keep their state (a frozen problem stays frozen).
__global__ void k_tfq_reset(TfqmrScal* s, double target, const int32_t* mask, int nprob, int fresh,
                            uint64_t first_seed) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob || (mask && !mask[p])) return;
  TfqmrScal* q = s + p;
  q->seed = fresh ? first_seed : q->seed + 1;
  q->rho = q->sigma = q->alpha = q->beta = 0;
  q->theta = q->tau = q->eta = q->dcoef = 0;
  q->eta_a = q->eta_b = 0;
  q->cy = 0;
  q->cd = 1;
  q->rr = 0;
  q->target = target;
  q->iter = 0;
  q->stop = 0;
  q->flag = kInitRequested;
}

// target[p] = target[p] * sqrt(rr[p] / rr_true[p]) for problems just (re)started
// (flag kInitialising): the quasi-residual bound of a scaled iteration measures
// the scaled residual, so its stop target follows the ratio at the start.
__global__ void k_tfq_scale_targets(TfqmrScal* s, const double* __restrict__ rr_true, int nprob) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob || s[p].flag != kInitialising) return;
  const double t = rr_true[2 * p];
  if (t > 0 && s[p].rr > 0) s[p].target *= sqrt(s[p].rr / t);
}

// unscaled[p] = mask[p] (0 for every problem when mask is null).
__global__ void k_tfq_set_unscaled(TfqmrScal* s, const int32_t* mask, int nprob) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nprob) s[p].unscaled = mask ? mask[p] : 0;
}

// flag[p] = kFrozen where mask[p]: finished problems drop out of every kernel.
__global__ void k_tfq_freeze(TfqmrScal* s, const int32_t* mask, int nprob) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < nprob && mask[p]) s[p].flag = kFrozen;
}

// v[p * stride + idx[k]] = val[k] for every problem p and k < m.
__global__ void k_set_pattern(double* v, int64_t stride, int nprob, SetPattern pat) {
  const int p = blockIdx.x;
  if (p >= nprob) return;
  for (int k = threadIdx.x; k < pat.m; k += blockDim.x) v[p * stride + pat.idx[k]] = pat.val[k];
}

template <typename T>
__global__ void k_set_elem(T* v, int64_t off, T val) {
  if (threadIdx.x == 0 && blockIdx.x == 0) v[off] = val;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T>
static int grid_for(int64_t n) {
  int64_t g = (n + 4 * kRedThreads - 1) / (4 * kRedThreads);
  if (g > kMaxRedBlocks) g = kMaxRedBlocks;
  if (g < 1) g = 1;
  return (int)g;
}

#define L1(kern, n, ...) hipLaunchKernelGGL(kern, dim3(grid_for<T>(n)), dim3(kRedThreads), 0, st, n, __VA_ARGS__)
#define LB(kern, n, ...)                                                                        \
  hipLaunchKernelGGL(kern, dim3(grid_for<T>(n), bt.nprob), dim3(kRedThreads), 0, st, UniLoc{n, bt.vs}, \
                     __VA_ARGS__)

static hipError_t finalize(hipStream_t st, int op, int nblk, const double* partial, TfqmrScal* s, double* out,
                           int nprob = 1) {
  hipLaunchKernelGGL(k_finalize<UniLoc>, dim3(nprob), dim3(kRedThreads), 0, st, UniLoc{0, 0}, op, nblk, partial,
                     s, out);
  return hipGetLastError();
}

// n counts reals; the vector kernels walk n/2 complex pairs of each problem
template <typename T>
hipError_t tfq_init_r(hipStream_t st, int64_t n, const T* b, const T* ax, T* w, T* y2, T* d, TfqmrScal* s,
                      double* partial, Batch bt) {
  const int64_t n2 = n / 2;
  LB((k_init_r<T, UniLoc>), n2, b, ax, w, y2, d, s, partial);
  if (hipError_t e = hipGetLastError()) return e;
  return finalize(st, F_INIT_R, grid_for<T>(n2), partial, s, nullptr, bt.nprob);
}
template <typename T>
hipError_t tfq_init_v(hipStream_t st, int64_t n, const T* ax, T* v, T* aw, T* u2, TfqmrScal* s, double* partial,
                      Batch bt) {
  const int64_t n2 = n / 2;
  LB((k_init_v<T, UniLoc>), n2, ax, v, aw, u2, s, partial);
  if (hipError_t e = hipGetLastError()) return e;
  return finalize(st, F_INIT_V, grid_for<T>(n2), partial, s, nullptr, bt.nprob);
}
template <typename T>
hipError_t tfq_step1(hipStream_t st, int64_t n, T* w, T* y2, const T* v, const T* aw, const T* u2, T* d,
                     TfqmrScal* s, double* partial, Batch bt) {
  const int64_t n2 = n / 2;
  LB((k_step1<T, UniLoc>), n2, w, y2, v, aw, u2, d, s, partial);
  return hipGetLastError();
}
template <typename T>
hipError_t tfq_step2(hipStream_t st, int64_t n, T* x, T* w, const T* u2, T* d, const T* y2, TfqmrScal* s,
                     TfqmrScal* s2, const double* partial, double* partial2, Batch bt) {
  const int64_t n2 = n / 2;
  LB((k_step2<T, UniLoc>), n2, x, w, u2, d, y2, s, s2, partial, partial2);
  return hipGetLastError();
}
template <typename T>
hipError_t tfq_v(hipStream_t st, int64_t n, const T* aw, const T* u2, T* v, T* x, const T* d, TfqmrScal* s,
                 const TfqmrScal* s2, double* partial, const double* partial2, Batch bt) {
  const int64_t n2 = n / 2;
  LB((k_v<T, UniLoc>), n2, aw, u2, v, x, d, s2, s, partial2, partial);
  if (hipError_t e = hipGetLastError()) return e;
  return finalize(st, F_V, grid_for<T>(n2), partial, s, nullptr, bt.nprob);
}
hipError_t tfq_step1_fused(hipStream_t st, int64_t n, double* w, double* y2, double* v, const double* aw,
                           const double* u2, double* d, double* x, const TfqmrScal* s, double* pv) {
  const int64_t n2 = n / 2;
  hipLaunchKernelGGL(k_step1_fused, dim3(grid_for<double>(n2)), dim3(kRedThreads), 0, st, n2, w, y2, v, aw, u2, d,
                     x, s, pv);
  return hipGetLastError();
}
hipError_t tfq_fold(hipStream_t st, int64_t n, const double* pv, const double* pa, int na, TfqmrScal* s) {
  hipLaunchKernelGGL(k_tfq_fold, dim3(1), dim3(kFoldThreads), 0, st, pv, grid_for<double>(n / 2), pa, na, s);
  return hipGetLastError();
}
hipError_t tfq_flush(hipStream_t st, int64_t n, double* x, const double* d, const double* y2, TfqmrScal* s) {
  const int64_t n2 = n / 2;
  hipLaunchKernelGGL(k_tfq_flush, dim3(grid_for<double>(n2)), dim3(kRedThreads), 0, st, n2, x, d, y2, s);
  if (hipError_t e = hipGetLastError()) return e;
  hipLaunchKernelGGL(k_tfq_flush_scal, dim3(1), dim3(64), 0, st, s);
  return hipGetLastError();
}
template <typename T>
hipError_t vec_fill(hipStream_t st, int64_t n, T* v, T val) {
  L1((k_fill<T>), n, v, val);
  return hipGetLastError();
}
template <typename T>
hipError_t vec_randn(hipStream_t st, int64_t n, T* v, uint64_t seed) {
  L1((k_randn<T>), n, v, seed);
  return hipGetLastError();
}
template <typename TB, typename TA>
hipError_t vec_resid(hipStream_t st, int64_t n, const TB* b, const TA* ax, double* out, double* partial, Batch bt) {
  using T = TB;
  const int64_t n2 = n / 2;
  LB((k_resid<TB, TA, UniLoc>), n2, b, ax, partial);
  if (hipError_t e = hipGetLastError()) return e;
  return finalize(st, F_RESID, grid_for<T>(n2), partial, nullptr, out, bt.nprob);
}
hipError_t vec_resid_to_f32(hipStream_t st, int64_t n, const double* b, const double* ax, float* r32, double scale) {
  using T = double;
  L1(k_resid_to_f32, n, b, ax, r32, scale);
  return hipGetLastError();
}
hipError_t vec_axpy_f32_to_f64(hipStream_t st, int64_t n, double* x, const float* dx, double alpha) {
  using T = double;
  L1(k_axpy_f32_to_f64, n, x, dx, alpha);
  return hipGetLastError();
}
template <typename T>
hipError_t vec_project(hipStream_t st, int64_t nado, int d, T* x, const int8_t* par, const TierBounds& tb,
                       bool hermitian) {
  L1((k_project_parity<T>), nado, d, x, par, tb);
  if (hermitian) hipLaunchKernelGGL((k_hermitize0<T>), dim3((d * d + 255) / 256), dim3(256), 0, st, d, x);
  return hipGetLastError();
}
template hipError_t vec_project<double>(hipStream_t, int64_t, int, double*, const int8_t*, const TierBounds&, bool);
template hipError_t vec_project<float>(hipStream_t, int64_t, int, float*, const int8_t*, const TierBounds&, bool);

hipError_t tfq_reset(hipStream_t st, TfqmrScal* s, double target, const int32_t* mask, int nprob, bool fresh,
                     uint64_t first_seed) {
  hipLaunchKernelGGL(k_tfq_reset, dim3((nprob + 63) / 64), dim3(64), 0, st, s, target, mask, nprob, fresh ? 1 : 0,
                     first_seed);
  return hipGetLastError();
}
hipError_t tfq_scale_targets(hipStream_t st, TfqmrScal* s, const double* rr_true, int nprob) {
  hipLaunchKernelGGL(k_tfq_scale_targets, dim3((nprob + 63) / 64), dim3(64), 0, st, s, rr_true, nprob);
  return hipGetLastError();
}
hipError_t tfq_set_unscaled(hipStream_t st, TfqmrScal* s, const int32_t* mask, int nprob) {
  hipLaunchKernelGGL(k_tfq_set_unscaled, dim3((nprob + 63) / 64), dim3(64), 0, st, s, mask, nprob);
  return hipGetLastError();
}
hipError_t tfq_freeze(hipStream_t st, TfqmrScal* s, const int32_t* mask, int nprob) {
  hipLaunchKernelGGL(k_tfq_freeze, dim3((nprob + 63) / 64), dim3(64), 0, st, s, mask, nprob);
  return hipGetLastError();
}
hipError_t vec_set_pattern(hipStream_t st, double* v, int64_t stride, int nprob, const SetPattern& pat) {
  hipLaunchKernelGGL(k_set_pattern, dim3(nprob), dim3(64), 0, st, v, stride, nprob, pat);
  return hipGetLastError();
}
hipError_t vec_set_elem(hipStream_t st, double* v, int64_t off, double val) {
  hipLaunchKernelGGL((k_set_elem<double>), dim3(1), dim3(64), 0, st, v, off, val);
  return hipGetLastError();
}

hipError_t vec_lincomb3(hipStream_t st, int64_t n, double* out, double a, const double* x, double b,
                        const double* y, double c, const double* z) {
  using T = double;
  L1(k_lincomb3, n, out, a, x, b, y, c, z);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Multi-size launches: one flat grid over every problem of every shape
// ---------------------------------------------------------------------------
int vec_blocks(int64_t n2) { return grid_for<double>(n2); }

#define LM(kern, ...) \
  hipLaunchKernelGGL(kern, dim3(m.bstart[m.ng]), dim3(kRedThreads), 0, st, MultiLoc{m}, __VA_ARGS__)

static hipError_t finalize_m(hipStream_t st, const Multi& m, int op, const double* partial, TfqmrScal* s,
                             double* out) {
  if (hipError_t e = hipGetLastError()) return e;
  hipLaunchKernelGGL(k_finalize<MultiLoc>, dim3(m.qstart[m.ng]), dim3(kRedThreads), 0, st, MultiLoc{m}, op, 0,
                     partial, s, out);
  return hipGetLastError();
}

hipError_t tfq_init_r(hipStream_t st, const Multi& m, const double* b, const double* ax, double* w, double* y2,
                      double* d, TfqmrScal* s, double* partial) {
  LM((k_init_r<double, MultiLoc>), b, ax, w, y2, d, s, partial);
  return finalize_m(st, m, F_INIT_R, partial, s, nullptr);
}
hipError_t tfq_init_v(hipStream_t st, const Multi& m, const double* ax, double* v, double* aw, double* u2,
                      TfqmrScal* s, double* partial) {
  LM((k_init_v<double, MultiLoc>), ax, v, aw, u2, s, partial);
  return finalize_m(st, m, F_INIT_V, partial, s, nullptr);
}
hipError_t tfq_step1(hipStream_t st, const Multi& m, double* w, double* y2, const double* v, const double* aw,
                     const double* u2, double* d, TfqmrScal* s, double* partial) {
  LM((k_step1<double, MultiLoc>), w, y2, v, aw, u2, d, s, partial);
  return hipGetLastError();
}
hipError_t tfq_step2(hipStream_t st, const Multi& m, double* x, double* w, const double* u2, double* d,
                     const double* y2, TfqmrScal* s, TfqmrScal* s2, const double* partial, double* partial2) {
  LM((k_step2<double, MultiLoc>), x, w, u2, d, y2, s, s2, partial, partial2);
  return hipGetLastError();
}
hipError_t tfq_v(hipStream_t st, const Multi& m, const double* aw, const double* u2, double* v, double* x,
                 const double* d, TfqmrScal* s, const TfqmrScal* s2, double* partial, const double* partial2) {
  LM((k_v<double, MultiLoc>), aw, u2, v, x, d, s2, s, partial2, partial);
  return finalize_m(st, m, F_V, partial, s, nullptr);
}
hipError_t vec_resid(hipStream_t st, const Multi& m, const double* b, const double* ax, double* out,
                     double* partial) {
  LM((k_resid<double, double, MultiLoc>), b, ax, partial);
  return finalize_m(st, m, F_RESID, partial, nullptr, out);
}

#define INST(T)                                                                                          \
  template hipError_t tfq_init_r<T>(hipStream_t, int64_t, const T*, const T*, T*, T*, T*, TfqmrScal*,     \
                                    double*, Batch);                                                    \
  template hipError_t tfq_init_v<T>(hipStream_t, int64_t, const T*, T*, T*, T*, TfqmrScal*, double*,     \
                                    Batch);                                                             \
  template hipError_t tfq_step1<T>(hipStream_t, int64_t, T*, T*, const T*, const T*, const T*, T*,       \
                                   TfqmrScal*, double*, Batch);                                         \
  template hipError_t tfq_step2<T>(hipStream_t, int64_t, T*, T*, const T*, T*, const T*, TfqmrScal*,     \
                                   TfqmrScal*, const double*, double*, Batch);                          \
  template hipError_t tfq_v<T>(hipStream_t, int64_t, const T*, const T*, T*, T*, const T*, TfqmrScal*,   \
                               const TfqmrScal*, double*, const double*, Batch);                        \
  template hipError_t vec_fill<T>(hipStream_t, int64_t, T*, T);                                           \
  template hipError_t vec_randn<T>(hipStream_t, int64_t, T*, uint64_t);
INST(double)
INST(float)
template hipError_t vec_resid<double, double>(hipStream_t, int64_t, const double*, const double*, double*,
                                              double*, Batch);
template hipError_t vec_resid<float, float>(hipStream_t, int64_t, const float*, const float*, double*, double*,
                                            Batch);

// ---------------------------------------------------------------------------
// BiCG (bicg/solvebicg.f90): real inner products of the (Re, Im) vectors
// (cdotlong, bicg/ddotlong.f90:18-31), absolute stop ||r|| <= crit.
// ---------------------------------------------------------------------------
__global__ void k_bicg_init(int64_t n2, const double* __restrict__ q, int64_t e_pair, double* __restrict__ r,
                            double* __restrict__ s, double* __restrict__ d, double* __restrict__ f, double* partial) {
  double rr = 0;
  GRID_LOOP(i, n2) {
    const double2 qi = reinterpret_cast<const double2*>(q)[i];
    double2 ri;
    ri.x = (i == e_pair ? 1.0 : 0.0) - qi.x;
    ri.y = -qi.y;
    reinterpret_cast<double2*>(r)[i] = ri;
    reinterpret_cast<double2*>(s)[i] = ri;
    reinterpret_cast<double2*>(d)[i] = ri;
    reinterpret_cast<double2*>(f)[i] = ri;
    rr += ri.x * ri.x + ri.y * ri.y;
  }
  block_partial(rr, 0.0, partial);
}

__global__ void k_bicg_fq(int64_t n2, const double* __restrict__ f, const double* __restrict__ q,
                          const BicgScal* sc, double* partial) {
  if (sc->flag) return;
  double a = 0;
  GRID_LOOP(i, n2) {
    const double2 fi = reinterpret_cast<const double2*>(f)[i], qi = reinterpret_cast<const double2*>(q)[i];
    a += fi.x * qi.x + fi.y * qi.y;
  }
  block_partial(a, 0.0, partial);
}

__global__ void k_bicg_update(int64_t n2, double* __restrict__ x, const double* __restrict__ d,
                              double* __restrict__ r, const double* __restrict__ q, double* __restrict__ s,
                              const double* __restrict__ qt, const BicgScal* sc, double* partial) {
  if (sc->flag) return;
  const double alpha = sc->alpha;
  double rr = 0, sr = 0;
  GRID_LOOP(i, n2) {
    double2 xi = reinterpret_cast<double2*>(x)[i];
    const double2 di = reinterpret_cast<const double2*>(d)[i];
    xi.x += alpha * di.x;
    xi.y += alpha * di.y;
    reinterpret_cast<double2*>(x)[i] = xi;
    double2 ri = reinterpret_cast<double2*>(r)[i];
    const double2 qi = reinterpret_cast<const double2*>(q)[i];
    ri.x -= alpha * qi.x;
    ri.y -= alpha * qi.y;
    reinterpret_cast<double2*>(r)[i] = ri;
    double2 si = reinterpret_cast<double2*>(s)[i];
    const double2 ti = reinterpret_cast<const double2*>(qt)[i];
    si.x -= alpha * ti.x;
    si.y -= alpha * ti.y;
    reinterpret_cast<double2*>(s)[i] = si;
    rr += ri.x * ri.x + ri.y * ri.y;
    sr += si.x * ri.x + si.y * ri.y;
  }
  block_partial(rr, sr, partial);
}

__global__ void k_bicg_dirs(int64_t n2, const double* __restrict__ r, const double* __restrict__ s,
                            double* __restrict__ d, double* __restrict__ f, const BicgScal* sc) {
  if (sc->flag) return;
  const double beta = sc->beta;
  GRID_LOOP(i, n2) {
    const double2 ri = reinterpret_cast<const double2*>(r)[i], si = reinterpret_cast<const double2*>(s)[i];
    double2 di = reinterpret_cast<double2*>(d)[i], fi = reinterpret_cast<double2*>(f)[i];
    di.x = ri.x + beta * di.x;
    di.y = ri.y + beta * di.y;
    fi.x = si.x + beta * fi.x;
    fi.y = si.y + beta * fi.y;
    reinterpret_cast<double2*>(d)[i] = di;
    reinterpret_cast<double2*>(f)[i] = fi;
  }
}

enum BicgOp { B_INIT = 0, B_ALPHA, B_UPDATE };

__global__ void k_bicg_finalize(int op, int nblk, const double* __restrict__ partial, BicgScal* sc, double crit) {
  if (op != B_INIT && sc->flag) return;
  double a = 0, b = 0;
  for (int t = threadIdx.x; t < nblk; t += blockDim.x) {
    a += partial[2 * t];
    b += partial[2 * t + 1];
  }
  block_sum2(a, b);
  if (threadIdx.x != 0) return;
  switch (op) {
    case B_INIT:   // solvebicg.f90:58-66: s = r, so <s, r> = ||r||^2
      sc->rr = a;
      sc->rho = a;
      sc->crit = crit;
      sc->iter = 0;
      sc->flag = sqrt(a) <= crit ? 1 : 0;
      sc->alpha = sc->beta = 0;
      break;
    case B_ALPHA:
      if (a == 0.0) { sc->flag = 2; break; }
      sc->alpha = sc->rho / a;
      break;
    case B_UPDATE:
      sc->rr = a;
      sc->iter += 1;
      if (sqrt(a) <= sc->crit) { sc->flag = 1; break; }
      if (sc->rho == 0.0) { sc->flag = 2; break; }
      sc->beta = b / sc->rho;
      sc->rho = b;
      break;
  }
}

static hipError_t bfin(hipStream_t st, int op, int nblk, const double* partial, BicgScal* sc, double crit) {
  hipLaunchKernelGGL(k_bicg_finalize, dim3(1), dim3(kRedThreads), 0, st, op, nblk, partial, sc, crit);
  return hipGetLastError();
}

hipError_t bicg_init(hipStream_t st, int64_t n, const double* q, int64_t e_off, double* r, double* s, double* d,
                     double* f, BicgScal* sc, double crit, double* partial) {
  using T = double;
  const int64_t n2 = n / 2;
  L1(k_bicg_init, n2, q, e_off / 2, r, s, d, f, partial);
  if (hipError_t e = hipGetLastError()) return e;
  return bfin(st, B_INIT, grid_for<T>(n2), partial, sc, crit);
}
hipError_t bicg_alpha(hipStream_t st, int64_t n, const double* f, const double* q, BicgScal* sc, double* partial) {
  using T = double;
  const int64_t n2 = n / 2;
  L1(k_bicg_fq, n2, f, q, sc, partial);
  if (hipError_t e = hipGetLastError()) return e;
  return bfin(st, B_ALPHA, grid_for<T>(n2), partial, sc, 0.0);
}
hipError_t bicg_update(hipStream_t st, int64_t n, double* x, const double* d, double* r, const double* q, double* s,
                       const double* qt, BicgScal* sc, double* partial) {
  using T = double;
  const int64_t n2 = n / 2;
  L1(k_bicg_update, n2, x, d, r, q, s, qt, sc, partial);
  if (hipError_t e = hipGetLastError()) return e;
  return bfin(st, B_UPDATE, grid_for<T>(n2), partial, sc, 0.0);
}
hipError_t bicg_dirs(hipStream_t st, int64_t n, const double* r, const double* s, double* d, double* f,
                     const BicgScal* sc) {
  using T = double;
  const int64_t n2 = n / 2;
  L1(k_bicg_dirs, n2, r, s, d, f, sc);
  return hipGetLastError();
}

}  // namespace heomq
