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
// Here the pattern's distinct per-ADO element lists become the "sectors" of
// the packed apply (heomq_capi.cpp, build_spa_operator): rho_spa, stored ADO
// after ADO as prelude_spa lays it out, is that layout's packed vector as it
// stands, the gather maps leave out neighbour elements outside a neighbour's
// list (P^T) and only the list's own elements are formed (P).  Device memory
// scales with lunk_spa and the number of distinct lists, not with N d^2.  The
// derivative of the sparse propagators (td/calcderiv_spa_omp.f90: rhotmp_spa
// -> rhorhs_spa, H and the lead shift at time tt, :84-98) is the same
// operator with the rates at tt; only a time-dependent system Hamiltonian
// (whose rows are not in the list tables) goes through the dense derivative
// between a scatter and a gather.
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
  // ADO after ADO, as prelude_spa.f90:741-742 stores them (ind_spa(a) =
  // ind_spa(a-1) + nnz_spa(a-1)); an empty ADO's ind_spa is not read
  std::vector<int64_t> start(c->N, 0);
  int64_t run = 0;
  for (int64_t a = 0; a < c->N; ++a) {
    if (nnz_spa[a] > 0 && ind_spa[a] - 1 != run)
      return fail(HEOMQ_BAD_ARG, "sparse pattern: ADOs not stored in order (ind_spa(a) != ind_spa(a-1) + nnz_spa(a-1))");
    start[a] = run;
    run += nnz_spa[a];
  }
  if (int rc = build_spa_operator(c, nnz_spa, start, irow_spa, icol_spa)) return rc;
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
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_apply_spa: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (c->spa_lunk < 0 || !c->spa.built) return fail(HEOMQ_BAD_ARG, "no sparse pattern: call heomq_spa_pattern first");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = c->spa_lunk;
  if (n > 0) HIPCHECK(launch_apply_packed_f64(spa_operator(c, nullptr), x_spa, y_spa, false, st));
  hipLaunchKernelGGL(k_spa_trace, dim3(1), dim3(256), 0, st, c->spa_ndiag, c->spa_diag.as<int64_t>(),
                     reinterpret_cast<const double2*>(x_spa), reinterpret_cast<double2*>(y_spa) + n);
  HIPCHECK(hipGetLastError());
  return HEOMQ_OK;
}

int heomq_calcderiv_spa_at(heomq_ctx* c, double tt, const double* x_spa, double* y_spa, void* stream) {
  if (!c || !x_spa || !y_spa) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calcderiv_spa_at: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (c->spa_lunk < 0 || !c->spa.built) return fail(HEOMQ_BAD_ARG, "no sparse pattern: call heomq_spa_pattern first");
  hipStream_t st = (hipStream_t)stream;
  const int64_t n = c->spa_lunk;
  if (n == 0) return HEOMQ_OK;
  if (c->sys_nterm == 0) {   // rates at tt (lead shift), H fixed: the list tables hold the whole operator
    DevOperator dop;
    if (int rc = operator_at(c, tt, st, dop)) return rc;
    HIPCHECK(launch_apply_packed_f64(spa_operator(c, dop.gam), x_spa, y_spa, false, st));
    return HEOMQ_OK;
  }
  const int64_t nd = 2 * c->N * c->d * c->d;
  HIPCHECK(c->spa_dense.alloc(2 * sizeof(double) * (size_t)nd));
  double* xd = c->spa_dense.as<double>();
  double* yd = xd + nd;
  hipLaunchKernelGGL(k_spa_zero, dim3(grid_of(nd)), dim3(256), 0, st, nd, xd);
  HIPCHECK(hipGetLastError());
  hipLaunchKernelGGL(k_spa_scatter, dim3(grid_of(n)), dim3(256), 0, st, n, c->spa_map.as<int64_t>(),
                     reinterpret_cast<const double2*>(x_spa), reinterpret_cast<double2*>(xd));
  HIPCHECK(hipGetLastError());
  if (int rc = heomq_calcderiv_at(c, tt, xd, yd, stream)) return rc;
  hipLaunchKernelGGL(k_spa_gather, dim3(grid_of(n)), dim3(256), 0, st, n, c->spa_map.as<int64_t>(),
                     reinterpret_cast<const double2*>(yd), reinterpret_cast<double2*>(y_spa));
  HIPCHECK(hipGetLastError());
  return HEOMQ_OK;
}

int heomq_calcderiv_spa_tt(heomq_ctx* c, double tt, const double* x_host, double* y_host) {
  if (!c || !x_host || !y_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calcderiv_spa_tt: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
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
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calc_ax_tfqmr_spa: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
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
