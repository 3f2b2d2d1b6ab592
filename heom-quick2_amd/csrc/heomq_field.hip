// Time-dependent lead energy shifts for the propagators: calcderiv_omp(tt)
// and the RK4 of td/intstp.f90 evaluate gamma_K(t) = sum_k cgamma_k +
// i dipm_k eshift(alpha_k, s_k; t) (td/calcderiv_omp.f90:51-61,144-148) with
// eshift from main/getengyshift.f90 (fieldtype 0: amps (1 - exp(-t/tchar)),
// 1: amps sin(2 pi t / tchar) + amps_dc; 0 once t >= t_off).  The system
// Hamiltonian keeps dhs = 0 (main/calcdhs.f90 without its dot-field options).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "heomq_ctx.hpp"

using namespace heomq;

namespace {

constexpr int kMaxLeadSpin = 8;

struct ShiftSet {
  double v[kMaxLeadSpin];
};

// gamma_K(t) = gamma_K(no bias) + i sum_(alpha, s) cnt[a][alpha, s] eshift(alpha, s)
__global__ void k_td_gamma(int64_t n, int k, const double2* __restrict__ base, const int8_t* __restrict__ cnt,
                           ShiftSet es, double2* __restrict__ out) {
  for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n; a += (int64_t)gridDim.x * blockDim.x) {
    double im = base[a].y;
    for (int q = 0; q < k; ++q) im += (double)cnt[a * k + q] * es.v[q];
    out[a] = make_double2(base[a].x, im);
  }
}

}  // namespace

namespace heomq {

// eshift(alpha, s; t) of main/getengyshift.f90 for every lead and spin
void field_shifts(const heomq_ctx* c, double t, double* es) {
  const int k = c->nalf * c->nspin;
  for (int q = 0; q < k; ++q) {
    double v = 0.0;
    if (t < c->field_toff) {
      if (c->field_type == 0) v = c->field_amp[q] * (1.0 - std::exp(-t / c->field_tchar[q]));
      else if (c->field_type == 1) v = c->field_amp[q] * std::sin(2.0 * M_PI * t / c->field_tchar[q]) + c->field_dc[q];
    }
    es[q] = v;
  }
}

// The forward operator at time t: the context's own when no field is set,
// else a copy whose rates are gamma_K(t), written into c->td_gam on st.
int operator_at(heomq_ctx* c, double t, hipStream_t st, DevOperator& op) {
  op = c->fwd.op;
  if (c->field_type < 0) return HEOMQ_OK;
  ShiftSet es{};
  field_shifts(c, t, es.v);
  const int k = c->nalf * c->nspin;
  const unsigned grid = (unsigned)std::min<int64_t>((c->N + 255) / 256, 2048);
  hipLaunchKernelGGL(k_td_gamma, dim3(grid), dim3(256), 0, st, c->N, k,
                     reinterpret_cast<const double2*>(c->fwd.gam.p), c->td_cnt.as<int8_t>(), es,
                     c->td_gam.as<double2>());
  HIPCHECK(hipGetLastError());
  op.gam = c->td_gam.as<double>();
  return HEOMQ_OK;
}

}  // namespace heomq

extern "C" {

int heomq_ctx_set_field(heomq_ctx* c, int fieldtype, const double* tchar, const double* amps_dc, double t_off) {
  if (!c) return fail(HEOMQ_BAD_ARG, "null argument");
  if (fieldtype != -1 && fieldtype != 0 && fieldtype != 1)
    return fail(HEOMQ_UNSUPPORTED, "fieldtype must be -1, 0 or 1 (pulsed shifts are read from the input deck)");
  const int k = c->nalf * c->nspin;
  if (k > kMaxLeadSpin) return fail(HEOMQ_UNSUPPORTED, "more than 8 lead-spin channels");
  if (fieldtype >= 0 && c->steady_model)
    return fail(HEOMQ_UNSUPPORTED, "a time-dependent field needs a model built without the steady bias (steady = 0)");
  if (fieldtype >= 0 && !tchar) return fail(HEOMQ_BAD_ARG, "tchar required");
  c->field_tchar.assign(k, 1.0);
  c->field_dc.assign(k, 0.0);
  for (int q = 0; q < k && fieldtype >= 0; ++q) {
    if (!(std::fabs(tchar[q]) > 0.0)) return fail(HEOMQ_BAD_ARG, "tchar must be nonzero");
    c->field_tchar[q] = tchar[q];
    if (amps_dc && fieldtype == 1) c->field_dc[q] = amps_dc[q] / kHbar;   // eV -> 1/fs
  }
  c->field_toff = t_off;
  c->field_type = fieldtype;
  if (fieldtype < 0) return HEOMQ_OK;
  std::vector<int8_t> cnt((size_t)c->N * k, 0);
  for (int n = 1; n <= c->L; ++n)
    for (int64_t g = c->first[n]; g < c->first[n + 1]; ++g) {
      const int32_t* K = &c->tiers[n][(g - c->first[n]) * n];
      for (int r = 0; r < n; ++r) {
        const int dr = K[r];
        cnt[(size_t)g * k + c->alf[dr] * c->nspin + c->spin[dr]] += (int8_t)(c->sgn[dr] == 0 ? 1 : -1);
      }
    }
  HIPCHECK(upload(c->td_cnt, cnt));
  HIPCHECK(c->td_gam.alloc(sizeof(double) * 2 * (size_t)c->N));
  return HEOMQ_OK;
}

int heomq_calcderiv_at(heomq_ctx* c, double tt, const double* x, double* y, void* stream) {
  if (!c || !x || !y) return fail(HEOMQ_BAD_ARG, "null argument");
  hipStream_t st = (hipStream_t)stream;
  DevOperator op;
  if (int rc = operator_at(c, tt, st, op)) return rc;
  HIPCHECK(launch_apply_f64(op, x, y, -1, st));
  return HEOMQ_OK;
}

}  // extern "C"

extern "C" int heomq_calcderiv_tt(heomq_ctx* c, double tt, const double* rhotmp_host, double* rhorhs_host) {
  if (!c || !rhotmp_host || !rhorhs_host) return fail(HEOMQ_BAD_ARG, "null argument");
  const int64_t n = 2 * c->N * c->d * c->d;
  HIPCHECK(c->ws_host.alloc(2 * sizeof(double) * (size_t)n));
  double* xd = c->ws_host.as<double>();
  double* yd = xd + n;
  HIPCHECK(hipMemcpy(xd, rhotmp_host, sizeof(double) * n, hipMemcpyHostToDevice));
  if (int rc = heomq_calcderiv_at(c, tt, xd, yd, nullptr)) return rc;
  HIPCHECK(hipMemcpy(rhorhs_host, yd, sizeof(double) * n, hipMemcpyDeviceToHost));
  return HEOMQ_OK;
}
