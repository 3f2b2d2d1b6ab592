// Sparse-storage drop-in: calc_ax_tfqmr_spa_omp (tfqmr_double/calc_ax_tfqmr_spa_omp.f90).
//
// The reference's sparse mode stores every ADO as a COO list of (row, column,
// value) over a pattern built once (fermion/prelude_spa.f90:532-749; module
// sparsemod: lunk_spa elements, nnz_spa / ind_spa per ADO, irow_spa /
// icol_spa).  Its apply computes the Liouvillian on that storage: neighbour
// blocks expanded from COO, the products formed, and only the pattern's
// elements of the result kept (zmat_extract_coo), i.e. y = P A P^T x with P
// the selection of the pattern; ADOs with nnz = 0 hold nothing (:225).  The
// trace of rho0 is appended as element lunk_spa + 1 (:1061-1069).
//
// Here the pattern becomes a map from COO element to dense index, and the
// apply is scatter -> the dense device apply -> gather, plus the trace: the
// arithmetic is the dense operator's, so P A P^T holds for any pattern.  The
// derivative of the sparse propagators (td/calcderiv_spa_omp.f90: rhotmp_spa
// -> rhorhs_spa, H and the lead shift at time tt, :84-98) is the same
// composition around the derivative at tt, with no trace element.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "heomq_ctx.hpp"

using namespace heomq;

namespace {

__global__ void k_spa_scatter(int64_t n, const int64_t* __restrict__ map, const double2* __restrict__ xs,
                              double2* __restrict__ dense) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    dense[map[e]] = xs[e];
}

__global__ void k_spa_gather(int64_t n, const int64_t* __restrict__ map, const double2* __restrict__ dense,
                             double2* __restrict__ ys) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    ys[e] = dense[map[e]];
}

// y[lunk] = sum of x over the tier-0 diagonal elements (one block, fixed order)
__global__ void k_spa_trace(int nd, const int64_t* __restrict__ diag, const double2* __restrict__ xs,
                            double2* __restrict__ yt) {
  __shared__ double2 part[256];
  double2 acc = make_double2(0.0, 0.0);
  for (int k = threadIdx.x; k < nd; k += 256) {
    const double2 v = xs[diag[k]];
    acc.x += v.x;
    acc.y += v.y;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      part[threadIdx.x].x += part[threadIdx.x + h].x;
      part[threadIdx.x].y += part[threadIdx.x + h].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *yt = part[0];
}

__global__ void k_spa_zero(int64_t n, double* __restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0.0;
}

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

}  // namespace

extern "C" {

int heomq_spa_pattern(heomq_ctx* c, int64_t lunk_spa, const int32_t* nnz_spa, const int64_t* ind_spa,
                      const int32_t* irow_spa, const int32_t* icol_spa) {
  if (!c || !nnz_spa || !ind_spa || (lunk_spa > 0 && (!irow_spa || !icol_spa)))
    return fail(HEOMQ_BAD_ARG, "null argument");
  if (lunk_spa < 0) return fail(HEOMQ_BAD_ARG, "sparse pattern: lunk_spa < 0");
  const int d = c->d, dd = d * d;
  std::vector<int64_t> map((size_t)lunk_spa, -1), diag;
  std::vector<uint8_t> seen(dd);
  for (int64_t a = 0; a < c->N; ++a) {
    const int32_t nz = nnz_spa[a];
    if (nz == 0) continue;
    const int64_t st = ind_spa[a] - 1;
    if (nz < 0 || nz > dd || st < 0 || st + nz > lunk_spa)
      return fail(HEOMQ_BAD_ARG, "sparse pattern: nnz_spa / ind_spa outside 1..lunk_spa");
    std::fill(seen.begin(), seen.end(), 0);
    for (int64_t e = st; e < st + nz; ++e) {
      const int32_t i = irow_spa[e] - 1, j = icol_spa[e] - 1;
      if (i < 0 || i >= d || j < 0 || j >= d) return fail(HEOMQ_BAD_ARG, "sparse pattern: irow / icol outside 1..nrho");
      if (seen[i + j * d]++) return fail(HEOMQ_BAD_ARG, "sparse pattern: an element listed twice in one ADO");
      if (map[e] >= 0) return fail(HEOMQ_BAD_ARG, "sparse pattern: two ADOs share storage");
      map[e] = a * dd + i + (int64_t)j * d;
      if (a == 0 && i == j) diag.push_back(e);
    }
  }
  for (int64_t e = 0; e < lunk_spa; ++e)
    if (map[e] < 0) return fail(HEOMQ_BAD_ARG, "sparse pattern: an element of rho_spa belongs to no ADO");
  const int ndiag = (int)diag.size();
  if (map.empty()) map.push_back(0);     // keep the device buffers non-empty
  if (diag.empty()) diag.push_back(0);
  HIPCHECK(upload(c->spa_map, map));
  HIPCHECK(upload(c->spa_diag, diag));
  c->spa_lunk = lunk_spa;
  c->spa_ndiag = ndiag;
  return HEOMQ_OK;
}

int heomq_apply_spa(heomq_ctx* c, const double* x_spa, double* y_spa, void* stream) {
  if (!c || !x_spa || !y_spa) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c->spa_lunk < 0) return fail(HEOMQ_BAD_ARG, "no sparse pattern: call heomq_spa_pattern first");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = c->spa_lunk, nd = 2 * c->N * c->d * c->d;
  HIPCHECK(c->spa_dense.alloc(2 * sizeof(double) * (size_t)nd));
  double* xd = c->spa_dense.as<double>();
  double* yd = xd + nd;
  hipLaunchKernelGGL(k_spa_zero, dim3(grid_of(nd)), dim3(256), 0, st, nd, xd);
  HIPCHECK(hipGetLastError());
  if (n > 0) {
    hipLaunchKernelGGL(k_spa_scatter, dim3(grid_of(n)), dim3(256), 0, st, n, c->spa_map.as<int64_t>(),
                       reinterpret_cast<const double2*>(x_spa), reinterpret_cast<double2*>(xd));
    HIPCHECK(hipGetLastError());
  }
  HIPCHECK(launch_apply_f64(c->fwd.op, xd, yd, -1, st));
  if (n > 0) {
    hipLaunchKernelGGL(k_spa_gather, dim3(grid_of(n)), dim3(256), 0, st, n, c->spa_map.as<int64_t>(),
                       reinterpret_cast<const double2*>(yd), reinterpret_cast<double2*>(y_spa));
    HIPCHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_spa_trace, dim3(1), dim3(256), 0, st, c->spa_ndiag, c->spa_diag.as<int64_t>(),
                     reinterpret_cast<const double2*>(x_spa), reinterpret_cast<double2*>(y_spa) + n);
  HIPCHECK(hipGetLastError());
  return HEOMQ_OK;
}

int heomq_calcderiv_spa_at(heomq_ctx* c, double tt, const double* x_spa, double* y_spa, void* stream) {
  if (!c || !x_spa || !y_spa) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c->spa_lunk < 0) return fail(HEOMQ_BAD_ARG, "no sparse pattern: call heomq_spa_pattern first");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = c->spa_lunk, nd = 2 * c->N * c->d * c->d;
  HIPCHECK(c->spa_dense.alloc(2 * sizeof(double) * (size_t)nd));
  double* xd = c->spa_dense.as<double>();
  double* yd = xd + nd;
  hipLaunchKernelGGL(k_spa_zero, dim3(grid_of(nd)), dim3(256), 0, st, nd, xd);
  HIPCHECK(hipGetLastError());
  if (n > 0) {
    hipLaunchKernelGGL(k_spa_scatter, dim3(grid_of(n)), dim3(256), 0, st, n, c->spa_map.as<int64_t>(),
                       reinterpret_cast<const double2*>(x_spa), reinterpret_cast<double2*>(xd));
    HIPCHECK(hipGetLastError());
  }
  if (int rc = heomq_calcderiv_at(c, tt, xd, yd, stream)) return rc;
  if (n > 0) {
    hipLaunchKernelGGL(k_spa_gather, dim3(grid_of(n)), dim3(256), 0, st, n, c->spa_map.as<int64_t>(),
                       reinterpret_cast<const double2*>(yd), reinterpret_cast<double2*>(y_spa));
    HIPCHECK(hipGetLastError());
  }
  return HEOMQ_OK;
}

int heomq_calcderiv_spa_tt(heomq_ctx* c, double tt, const double* x_host, double* y_host) {
  if (!c || !x_host || !y_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c->spa_lunk < 0) return fail(HEOMQ_BAD_ARG, "no sparse pattern: call heomq_spa_pattern first");
  const int64_t n = c->spa_lunk;
  if (n == 0) return HEOMQ_OK;
  HIPCHECK(c->spa_io.alloc(sizeof(double) * 2 * (size_t)(2 * n + 1)));
  double* xd = c->spa_io.as<double>();
  double* yd = xd + 2 * n;
  HIPCHECK(hipMemcpy(xd, x_host, sizeof(double) * 2 * n, hipMemcpyHostToDevice));
  if (int rc = heomq_calcderiv_spa_at(c, tt, xd, yd, nullptr)) return rc;
  HIPCHECK(hipMemcpy(y_host, yd, sizeof(double) * 2 * n, hipMemcpyDeviceToHost));
  return HEOMQ_OK;
}

int heomq_calc_ax_tfqmr_spa(heomq_ctx* c, int64_t ldim, const double* x_host, double* y_host) {
  if (!c || !x_host || !y_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c->spa_lunk < 0) return fail(HEOMQ_BAD_ARG, "no sparse pattern: call heomq_spa_pattern first");
  if (ldim != c->spa_lunk) return fail(HEOMQ_BAD_DIM, "calc_ax_tfqmr_spa_omp: error dimension, ldim != lunk_spa");
  const int64_t n = c->spa_lunk;
  HIPCHECK(c->spa_io.alloc(sizeof(double) * 2 * (size_t)(2 * n + 1)));
  double* xd = c->spa_io.as<double>();
  double* yd = xd + 2 * n;
  if (n > 0) HIPCHECK(hipMemcpy(xd, x_host, sizeof(double) * 2 * n, hipMemcpyHostToDevice));
  if (int rc = heomq_apply_spa(c, xd, yd, nullptr)) return rc;
  HIPCHECK(hipMemcpy(y_host, yd, sizeof(double) * 2 * (n + 1), hipMemcpyDeviceToHost));
  return HEOMQ_OK;
}

}  // extern "C"
