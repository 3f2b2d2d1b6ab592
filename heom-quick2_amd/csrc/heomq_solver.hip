// Vector kernels of the device-resident TFQMR steady-state solver, the RK4 and
// Chebyshev propagators.  The solver keeps every vector and every scalar on
// the GPU: each fused kernel does its vector updates and one FP64 reduction,
// and the last workgroup to finish (ticket counter) folds the per-workgroup
// partials in a fixed order and advances the TFQMR scalars, so a whole
// iteration is launch-only (no host round trip) and deterministic.
#include <hip/hip_runtime.h>

#include <cstdint>

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

// Stores this block's partials; returns true in thread 0 of the last block,
// with the ordered totals in (a, b).
__device__ inline bool finish_reduction(double& a, double& b, double* partial, unsigned* ticket) {
  block_sum2(a, b);
  __shared__ bool last;
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = b;
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  // ordered fold by one wave: lane l sums partials l, l+64, ... then a fixed tree
  double ta = 0, tb = 0;
  if (threadIdx.x < 64)
    for (unsigned t = threadIdx.x; t < gridDim.x; t += 64) {
      ta += ((volatile double*)partial)[2 * t];
      tb += ((volatile double*)partial)[2 * t + 1];
    }
  if (threadIdx.x < 64) {
    for (int off = 32; off > 0; off >>= 1) {
      ta += __shfl_down(ta, off, 64);
      tb += __shfl_down(tb, off, 64);
    }
  }
  if (threadIdx.x == 0) {
    a = ta;
    b = tb;
    *ticket = 0u;   // ready for the next launch
  }
  return threadIdx.x == 0;
}

#define GRID_LOOP(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

// r = b - Ax (stored in w and y1), ||r||^2 and rt.r; sets tau, rho, nb-based target.
template <typename T>
__global__ void k_init_r(int64_t n, const T* __restrict__ b, const T* __restrict__ ax, T* __restrict__ w,
                         T* __restrict__ y1, T* __restrict__ d, const T* __restrict__ rt, TfqmrScal* s,
                         double* partial) {
  double rr = 0, rrt = 0;
  GRID_LOOP(i, n) {
    const T r = b[i] - ax[i];
    w[i] = r;
    y1[i] = r;
    d[i] = 0;
    rr += (double)r * r;
    rrt += (double)rt[i] * r;
  }
  if (finish_reduction(rr, rrt, partial, &s->ticket)) {
    s->rr = rr;
    s->tau = sqrt(rr);
    s->rho = rrt;
    s->theta = 0;
    s->eta = 0;
    s->dcoef = 0;
    s->flag = (rrt == 0.0 && rr > 0.0) ? 2 : 0;
  }
}

// v = u1 (= A y1) and sigma = rt.v -> alpha.
template <typename T>
__global__ void k_init_v(int64_t n, const T* __restrict__ u1, T* __restrict__ v, const T* __restrict__ rt,
                         TfqmrScal* s, double* partial) {
  double sg = 0, z = 0;
  GRID_LOOP(i, n) {
    const T t = u1[i];
    v[i] = t;
    sg += (double)rt[i] * t;
  }
  if (finish_reduction(sg, z, partial, &s->ticket)) {
    s->sigma = sg;
    if (sg == 0.0) { s->flag = 2; return; }
    s->alpha = s->rho / sg;
    s->dcoef = 0;
  }
}

template <typename T>
__global__ void k_y2(int64_t n, const T* __restrict__ y1, const T* __restrict__ v, T* __restrict__ y2,
                     const TfqmrScal* s) {
  if (s->flag) return;
  const T alpha = (T)s->alpha;
  GRID_LOOP(i, n) y2[i] = y1[i] - alpha * v[i];
}

// One TFQMR half step: [x += eta d], w -= alpha u, d = y + dcoef d, ||w||^2,
// then theta, c, tau, eta, dcoef (Freund 1993; Saad, Iterative Methods 7.4.2).
template <typename T>
__global__ void k_half(int64_t n, T* __restrict__ x, int xupd, T* __restrict__ w, const T* __restrict__ u,
                       T* __restrict__ d, const T* __restrict__ yj, TfqmrScal* s, double* partial) {
  if (s->flag) return;
  const T alpha = (T)s->alpha, dcoef = (T)s->dcoef, eta = (T)s->eta;
  double ww = 0, z = 0;
  GRID_LOOP(i, n) {
    const T di = d[i];
    if (xupd) x[i] += eta * di;
    const T wi = w[i] - alpha * u[i];
    w[i] = wi;
    d[i] = yj[i] + dcoef * di;
    ww += (double)wi * wi;
  }
  if (finish_reduction(ww, z, partial, &s->ticket)) {
    const double theta = sqrt(ww) / s->tau;
    const double c = 1.0 / sqrt(1.0 + theta * theta);
    s->tau = s->tau * theta * c;
    s->theta = theta;
    s->eta = c * c * s->alpha;
    s->dcoef = theta * theta * s->eta / s->alpha;
  }
}

// x += eta d; rho_new = rt.w; beta; convergence bound tau*sqrt(2m+1) <= target.
template <typename T>
__global__ void k_rho(int64_t n, T* __restrict__ x, const T* __restrict__ d, const T* __restrict__ w,
                      const T* __restrict__ rt, TfqmrScal* s, double* partial) {
  if (s->flag) return;
  const T eta = (T)s->eta;
  double r = 0, z = 0;
  GRID_LOOP(i, n) {
    x[i] += eta * d[i];
    r += (double)rt[i] * w[i];
  }
  if (finish_reduction(r, z, partial, &s->ticket)) {
    s->iter += 1;
    if (s->tau * sqrt(2.0 * s->iter + 1.0) <= s->target) { s->flag = 1; return; }
    if (s->rho == 0.0) { s->flag = 2; return; }
    s->beta = r / s->rho;
    s->rho = r;
  }
}

template <typename T>
__global__ void k_y1(int64_t n, const T* __restrict__ w, const T* __restrict__ y2, T* __restrict__ y1,
                     const TfqmrScal* s) {
  if (s->flag) return;
  const T beta = (T)s->beta;
  GRID_LOOP(i, n) y1[i] = w[i] + beta * y2[i];
}

// v = u1 + beta (u2 + beta v); sigma = rt.v -> alpha, dcoef.
template <typename T>
__global__ void k_v(int64_t n, const T* __restrict__ u1, const T* __restrict__ u2, T* __restrict__ v,
                    const T* __restrict__ rt, TfqmrScal* s, double* partial) {
  if (s->flag) return;
  const T beta = (T)s->beta;
  double sg = 0, z = 0;
  GRID_LOOP(i, n) {
    const T t = u1[i] + beta * (u2[i] + beta * v[i]);
    v[i] = t;
    sg += (double)rt[i] * t;
  }
  if (finish_reduction(sg, z, partial, &s->ticket)) {
    s->sigma = sg;
    if (sg == 0.0) { s->flag = 2; return; }
    s->alpha = s->rho / sg;
    s->dcoef = s->theta * s->theta * s->eta / s->alpha;
  }
}

// ||b - ax||^2 and ||b||^2 into out[0], out[1] (FP64).
template <typename TB, typename TA>
__global__ void k_resid(int64_t n, const TB* __restrict__ b, const TA* __restrict__ ax, double* out,
                        unsigned* ticket, double* partial) {
  double rr = 0, bb = 0;
  GRID_LOOP(i, n) {
    const double r = (double)b[i] - (double)ax[i];
    rr += r * r;
    bb += (double)b[i] * (double)b[i];
  }
  if (finish_reduction(rr, bb, partial, ticket)) {
    out[0] = rr;
    out[1] = bb;
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
__global__ void k_tfq_reset(TfqmrScal* s, double target) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  s->rho = s->sigma = s->alpha = s->beta = 0;
  s->theta = s->tau = s->eta = s->dcoef = 0;
  s->rr = 0;
  s->target = target;
  s->iter = 0;
  s->flag = 0;
  s->ticket = 0;
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

template <typename T>
hipError_t tfq_init_r(hipStream_t st, int64_t n, const T* b, const T* ax, T* w, T* y1, T* d, const T* rt,
                      TfqmrScal* s, double* partial) {
  L1((k_init_r<T>), n, b, ax, w, y1, d, rt, s, partial);
  return hipGetLastError();
}
template <typename T>
hipError_t tfq_init_v(hipStream_t st, int64_t n, const T* u1, T* v, const T* rt, TfqmrScal* s, double* partial) {
  L1((k_init_v<T>), n, u1, v, rt, s, partial);
  return hipGetLastError();
}
template <typename T>
hipError_t tfq_y2(hipStream_t st, int64_t n, const T* y1, const T* v, T* y2, const TfqmrScal* s) {
  L1((k_y2<T>), n, y1, v, y2, s);
  return hipGetLastError();
}
template <typename T>
hipError_t tfq_half(hipStream_t st, int64_t n, T* x, int xupd, T* w, const T* u, T* d, const T* yj,
                    TfqmrScal* s, double* partial) {
  L1((k_half<T>), n, x, xupd, w, u, d, yj, s, partial);
  return hipGetLastError();
}
template <typename T>
hipError_t tfq_rho(hipStream_t st, int64_t n, T* x, const T* d, const T* w, const T* rt, TfqmrScal* s,
                   double* partial) {
  L1((k_rho<T>), n, x, d, w, rt, s, partial);
  return hipGetLastError();
}
template <typename T>
hipError_t tfq_y1(hipStream_t st, int64_t n, const T* w, const T* y2, T* y1, const TfqmrScal* s) {
  L1((k_y1<T>), n, w, y2, y1, s);
  return hipGetLastError();
}
template <typename T>
hipError_t tfq_v(hipStream_t st, int64_t n, const T* u1, const T* u2, T* v, const T* rt, TfqmrScal* s,
                 double* partial) {
  L1((k_v<T>), n, u1, u2, v, rt, s, partial);
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
hipError_t vec_resid(hipStream_t st, int64_t n, const TB* b, const TA* ax, double* out, unsigned* ticket,
                     double* partial) {
  using T = TB;
  L1((k_resid<TB, TA>), n, b, ax, out, ticket, partial);
  return hipGetLastError();
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
  const int64_t n = nado * d * d;
  L1((k_project_parity<T>), nado, d, x, par, tb);
  if (hermitian) hipLaunchKernelGGL((k_hermitize0<T>), dim3((d * d + 255) / 256), dim3(256), 0, st, d, x);
  (void)n;
  return hipGetLastError();
}
template hipError_t vec_project<double>(hipStream_t, int64_t, int, double*, const int8_t*, const TierBounds&, bool);
template hipError_t vec_project<float>(hipStream_t, int64_t, int, float*, const int8_t*, const TierBounds&, bool);

hipError_t tfq_reset(hipStream_t st, TfqmrScal* s, double target) {
  hipLaunchKernelGGL(k_tfq_reset, dim3(1), dim3(64), 0, st, s, target);
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

#define INST(T)                                                                                          \
  template hipError_t tfq_init_r<T>(hipStream_t, int64_t, const T*, const T*, T*, T*, T*, const T*,       \
                                    TfqmrScal*, double*);                                               \
  template hipError_t tfq_init_v<T>(hipStream_t, int64_t, const T*, T*, const T*, TfqmrScal*, double*);  \
  template hipError_t tfq_y2<T>(hipStream_t, int64_t, const T*, const T*, T*, const TfqmrScal*);         \
  template hipError_t tfq_half<T>(hipStream_t, int64_t, T*, int, T*, const T*, T*, const T*, TfqmrScal*, \
                                  double*);                                                             \
  template hipError_t tfq_rho<T>(hipStream_t, int64_t, T*, const T*, const T*, const T*, TfqmrScal*,     \
                                 double*);                                                              \
  template hipError_t tfq_y1<T>(hipStream_t, int64_t, const T*, const T*, T*, const TfqmrScal*);         \
  template hipError_t tfq_v<T>(hipStream_t, int64_t, const T*, const T*, T*, const T*, TfqmrScal*,       \
                               double*);                                                                \
  template hipError_t vec_fill<T>(hipStream_t, int64_t, T*, T);                                           \
  template hipError_t vec_randn<T>(hipStream_t, int64_t, T*, uint64_t);
INST(double)
INST(float)
template hipError_t vec_resid<double, double>(hipStream_t, int64_t, const double*, const double*, double*,
                                              unsigned*, double*);
template hipError_t vec_resid<float, float>(hipStream_t, int64_t, const float*, const float*, double*,
                                            unsigned*, double*);

}  // namespace heomq
