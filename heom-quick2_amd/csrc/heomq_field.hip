// Time-dependent terms of the propagators: calcderiv_omp(tt) and the RK4 of
// td/intstp.f90 evaluate gamma_K(t) = sum_k cgamma_k + i dipm_k
// eshift(alpha_k, s_k; t) (td/calcderiv_omp.f90:51-61,144-148) with eshift
// from main/getengyshift.f90 (fieldtype 0: amps (1 - exp(-t/tchar)), 1: amps
// sin(2 pi t / tchar) + amps_dc; 0 once t >= t_off), and the system
// Hamiltonian hs + dhs(t) (:60-68, main/calcdhs.f90) with dhs(t) a sum of
// fixed Hermitian operators times the scalar time functions of calcdhs's
// options.
#include <hip/hip_runtime.h>

#include <algorithm>
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

constexpr int kMaxSysTerms = 16;

struct SysCoef {
  double f[kMaxSysTerms];
};

// H(t) over the union row pattern: hs + sum_k f_k(t) D_k (d*d slots)
__global__ void k_td_hval(int nslot, int nterm, const double2* __restrict__ base, const double2* __restrict__ terms,
                          SysCoef cf, double2* __restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nslot; i += gridDim.x * blockDim.x) {
    double2 v = base[i];
    for (int k = 0; k < nterm; ++k) {
      const double2 t = terms[(size_t)k * nslot + i];
      v.x += cf.f[k] * t.x;
      v.y += cf.f[k] * t.y;
    }
    out[i] = v;
  }
}

// The scalar time functions of main/calcdhs.f90's options (amplitude p[0] in
// eV; times in fs, angular frequencies in 1/fs):
//   0 p0 (1 - exp(-t / p1))              dfieldtype 0 (:93-94), lbfield (:302)
//   1 p0 sin(p1 t)                       dfieldtype 1, one level (:95-96)
//   2 p0 cos(p1 t)                       dfieldtype 2 (:160-161, :263-265)
//   3 p0 for p1 < t < p2                 the gate egate on (ton_gate, toff_gate) (:112-115)
//   4 p0 for p1 <= t < p2                dfieldtype 4 (:267-270)
//   5 p0 + p1 cos(p2 t + p3)             gatetype 1 (:120-128)
//   6 p0 for t < p1                      dfieldtype 1 switched off at wdot (:140-142, :190-193)
//   7 p0 e^(-p1 t) for t < p2, -p0 e^(-p1 (t - p2)) for t < p3, p0 e^(-p1 (t - p3)) after
//                                        the Hubbard dot field (:46-58)
//   8 p0 for t > p1                      terms switched on after t = 0 (tdot12, B fields)
double sys_coef(int kind, const double* p, double t) {
  switch (kind) {
    case 0: return p[0] * (1.0 - std::exp(-t / p[1]));
    case 1: return p[0] * std::sin(p[1] * t);
    case 2: return p[0] * std::cos(p[1] * t);
    case 3: return (t > p[1] && t < p[2]) ? p[0] : 0.0;
    case 4: return (t >= p[1] && t < p[2]) ? p[0] : 0.0;
    case 5: return p[0] + p[1] * std::cos(p[2] * t + p[3]);
    case 6: return t < p[1] ? p[0] : 0.0;
    case 7:
      if (t < p[2]) return p[0] * std::exp(-p[1] * t);
      if (t < p[3]) return -p[0] * std::exp(-p[1] * (t - p[2]));
      return p[0] * std::exp(-p[1] * (t - p[3]));
    case 8: return t > p[1] ? p[0] : 0.0;
  }
  return 0.0;
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
  if (c->field_type >= 0) {
    ShiftSet es{};
    field_shifts(c, t, es.v);
    const int k = c->nalf * c->nspin;
    const unsigned grid = (unsigned)std::min<int64_t>((c->N + 255) / 256, 2048);
    hipLaunchKernelGGL(k_td_gamma, dim3(grid), dim3(256), 0, st, c->N, k,
                       reinterpret_cast<const double2*>(c->fwd.gam.p), c->td_cnt.as<int8_t>(), es,
                       c->td_gam.as<double2>());
    HIPCHECK(hipGetLastError());
    op.gam = c->td_gam.as<double>();
  }
  if (c->sys_nterm > 0) {
    SysCoef cf{};
    for (int k = 0; k < c->sys_nterm; ++k) cf.f[k] = sys_coef(c->sys_kind[k], &c->sys_par[4 * (size_t)k], t);
    const int nslot = c->d * c->d;
    hipLaunchKernelGGL(k_td_hval, dim3((nslot + 255) / 256), dim3(256), 0, st, nslot, c->sys_nterm,
                       c->sys_base.as<double2>(), c->sys_terms.as<double2>(), cf, c->sys_hval.as<double2>());
    HIPCHECK(hipGetLastError());
    op.hnnz = c->sys_hnnz.as<int32_t>();
    op.hcol = c->sys_hcol.as<int32_t>();
    op.hval = c->sys_hval.as<double>();
  }
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

int heomq_ctx_set_system_field(heomq_ctx* c, int nterm, const double* ops, const int32_t* kinds, const double* pars) {
  if (!c) return fail(HEOMQ_BAD_ARG, "null argument");
  if (nterm < 0 || nterm > kMaxSysTerms) return fail(HEOMQ_BAD_ARG, "system field: 0 to 16 terms");
  if (nterm > 0 && (!ops || !kinds || !pars)) return fail(HEOMQ_BAD_ARG, "null argument");
  const int d = c->d, dd = d * d;
  std::vector<cplx> D((size_t)nterm * dd);
  for (int k = 0; k < nterm; ++k) {
    if (kinds[k] < 0 || kinds[k] > 8) return fail(HEOMQ_BAD_ARG, "system field: time-function kind outside 0..8");
    for (int q = 0; q < 4; ++q)
      if (!std::isfinite(pars[4 * k + q])) return fail(HEOMQ_BAD_ARG, "system field: parameters must be finite");
    if (kinds[k] == 0 && !(std::fabs(pars[4 * k + 1]) > 0.0))
      return fail(HEOMQ_BAD_ARG, "system field: kind 0 needs a nonzero time constant");
    double mx = 0.0;
    for (int e = 0; e < dd; ++e) {
      D[(size_t)k * dd + e] = cplx(ops[2 * ((size_t)k * dd + e)], ops[2 * ((size_t)k * dd + e) + 1]);
      mx = std::max(mx, std::abs(D[(size_t)k * dd + e]));
    }
    for (int i = 0; i < d; ++i)
      for (int j = 0; j < d; ++j)
        if (std::abs(D[(size_t)k * dd + i + j * d] - std::conj(D[(size_t)k * dd + j + i * d])) > 1e-12 * std::max(mx, 1.0))
          return fail(HEOMQ_BAD_ARG, "system field: every operator must be Hermitian");
  }
  c->sys_nterm = 0;
  if (nterm == 0) return HEOMQ_OK;
  // hs from the operator's own row tables
  std::vector<int32_t> nnz(d), col(dd);
  std::vector<double> val(2 * (size_t)dd);
  HIPCHECK(hipMemcpy(nnz.data(), c->fwd.op.hnnz, sizeof(int32_t) * d, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(col.data(), c->fwd.op.hcol, sizeof(int32_t) * dd, hipMemcpyDeviceToHost));
  HIPCHECK(hipMemcpy(val.data(), c->fwd.op.hval, sizeof(double) * 2 * dd, hipMemcpyDeviceToHost));
  std::vector<cplx> hs((size_t)dd, cplx(0, 0));
  for (int i = 0; i < d; ++i)
    for (int p = 0; p < nnz[i]; ++p) hs[i + (size_t)col[(size_t)i * d + p] * d] = cplx(val[2 * ((size_t)i * d + p)], val[2 * ((size_t)i * d + p) + 1]);
  std::vector<int32_t> unz(d, 0), ucol(dd, 0);
  std::vector<double> base(2 * (size_t)dd, 0.0), terms(2 * (size_t)nterm * dd, 0.0);
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      bool any = hs[i + (size_t)j * d] != cplx(0, 0);
      for (int k = 0; k < nterm && !any; ++k) any = D[(size_t)k * dd + i + j * d] != cplx(0, 0);
      if (!any) continue;
      const size_t slot = (size_t)i * d + unz[i]++;
      ucol[slot] = j;
      base[2 * slot] = hs[i + (size_t)j * d].real();
      base[2 * slot + 1] = hs[i + (size_t)j * d].imag();
      for (int k = 0; k < nterm; ++k) {
        const cplx v = D[(size_t)k * dd + i + j * d] / kHbar;
        terms[2 * ((size_t)k * dd + slot)] = v.real();
        terms[2 * ((size_t)k * dd + slot) + 1] = v.imag();
      }
    }
  HIPCHECK(upload(c->sys_hnnz, unz));
  HIPCHECK(upload(c->sys_hcol, ucol));
  HIPCHECK(upload(c->sys_base, base));
  HIPCHECK(upload(c->sys_terms, terms));
  HIPCHECK(c->sys_hval.alloc(sizeof(double) * 2 * (size_t)dd));
  c->sys_kind.assign(kinds, kinds + nterm);
  c->sys_par.assign(pars, pars + 4 * (size_t)nterm);
  c->sys_nterm = nterm;
  return HEOMQ_OK;
}

int heomq_ctx_field_shifts(const heomq_ctx* c, double tt, double* eshift_ev) {
  if (!c || !eshift_ev) return fail(HEOMQ_BAD_ARG, "null argument");
  const int k = c->nalf * c->nspin;
  std::vector<double> es(kMaxLeadSpin, 0.0);
  if (c->field_type >= 0) field_shifts(c, tt, es.data());
  for (int q = 0; q < k; ++q) eshift_ev[q] = es[q] * kHbar;
  return HEOMQ_OK;
}

int heomq_calcderiv_at(heomq_ctx* c, double tt, const double* x, double* y, void* stream) {
  if (!c || !x || !y) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calcderiv_at: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
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
