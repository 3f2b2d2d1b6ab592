// Projected Liouvillian apply on symmetry-packed ADO vectors (solver path).
//
// All index work is precomputed in slot space.  For every sector the term
// i(HX - XH) is a short CSR row over the sector's packed slots, and every
// coupling carries two gather maps (output slot -> signed neighbour slot) for
// its left and right operator, so a coupling costs entry -> map -> data.
//
// Work split: ADOs below the top tier (about M couplings each) run one
// workgroup per ADO, the coupling list sliced over the lanes and the slices
// summed through LDS; top-tier ADOs (L couplings) run one thread per packed
// element.  The output is the projected operator Pi*A: off-sector elements do
// not exist and the tier-0 block is Hermitianised by a one-block fix-up.
#include "heomq_kernels.hpp"

namespace heomq {

template <typename T> struct P2;
template <> struct P2<double> { using type = double2; };
template <> struct P2<float> { using type = float2; };

__device__ inline int64_t find_ado_tab(const DevPacked& op, const int32_t* tab, int64_t t) {
  int64_t lo = tab[blockIdx.x], hi = tab[blockIdx.x + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (op.off[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <typename T, bool TRANS>
__device__ inline typename P2<T>::type ld(const typename P2<T>::type* xv, int64_t idx, int64_t zero_at) {
  typename P2<T>::type v = xv[idx];
  if (TRANS && idx == zero_at) { v.x = 0; v.y = 0; }
  return v;
}

// gamma_K x_t + hsign * i(HX - XH)_t for packed slot t of the ADO at `base`
template <typename T, bool TRANS>
__device__ inline void diag_term(const DevPacked& op, const typename P2<T>::type* xv, int64_t a, int64_t base,
                                 int s, int t, int64_t zero_at, T& ar, T& ai) {
  const int dd = op.d * op.d;
  const auto xe = ld<T, TRANS>(xv, base + t, zero_at);
  const T gr = (T)op.gam[2 * a], gi = (T)op.gam[2 * a + 1];
  ar = gr * xe.x - gi * xe.y;
  ai = gr * xe.y + gi * xe.x;
  const int r0 = op.hrow[s * (dd + 1) + t], r1 = op.hrow[s * (dd + 1) + t + 1];
  T cr = 0, ci = 0;
  for (int u = r0; u < r1; ++u) {
    const auto v = ld<T, TRANS>(xv, base + op.hq[u], zero_at);
    const T hr = (T)op.hc[2 * u], hi = (T)op.hc[2 * u + 1];
    cr += hr * v.x - hi * v.y;
    ci += hr * v.y + hi * v.x;
  }
  const T hs = (T)op.hsign;
  ar += hs * cr;
  ai += hs * ci;
}

template <typename T, bool TRANS>
__device__ inline void coupling_term(const DevEntryP& en, const DevRec& rc, const int16_t* __restrict__ maps,
                                     const typename P2<T>::type* xv, int t, int64_t zero_at, T& ar, T& ai) {
  const bool dag = en.flags & 1;
  if (en.mapL >= 0) {
    const int mv = maps[en.mapL + t];
    if (mv != 0) {
      auto v = ld<T, TRANS>(xv, (int64_t)en.base + (mv > 0 ? mv : -mv) - 1, zero_at);
      if (dag) v.y = -v.y;
      const T sg = ((mv > 0) != ((en.flags & 2) != 0)) ? (T)1 : (T)-1;
      ar += sg * ((T)rc.cLr * v.x - (T)rc.cLi * v.y);
      ai += sg * ((T)rc.cLr * v.y + (T)rc.cLi * v.x);
    }
  }
  if (en.mapR >= 0) {
    const int mv = maps[en.mapR + t];
    if (mv != 0) {
      auto v = ld<T, TRANS>(xv, (int64_t)en.base + (mv > 0 ? mv : -mv) - 1, zero_at);
      if (dag) v.y = -v.y;
      const T sg = ((mv > 0) != ((en.flags & 4) != 0)) ? (T)1 : (T)-1;
      ar += sg * ((T)rc.cRr * v.x - (T)rc.cRi * v.y);
      ai += sg * ((T)rc.cRr * v.y + (T)rc.cRi * v.x);
    }
  }
}

// Top-tier ADOs: one thread per packed element.
template <typename T, bool TRANS>
__global__ void __launch_bounds__(256)
k_apply_packed(const DevPacked op, const T* __restrict__ x, T* __restrict__ y, int64_t zero_at) {
  using V = typename P2<T>::type;
  const int64_t t = op.light_begin + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= op.total) return;   // no barrier in this kernel
  const V* __restrict__ xv = reinterpret_cast<const V*>(x);
  const int64_t a = find_ado_tab(op, op.light_block_ado, t);
  const int64_t base = op.off[a];
  const int slot = (int)(t - base);
  T ar, ai;
  diag_term<T, TRANS>(op, xv, a, base, op.sec[a], slot, zero_at, ar, ai);
  const int64_t p0 = op.rowptr[a], p1 = op.rowptr[a + 1];
  for (int64_t p = p0; p < p1; ++p) {
    const DevEntryP en = op.ent[p];
    coupling_term<T, TRANS>(en, op.rec[en.rec], op.maps, xv, slot, zero_at, ar, ai);
  }
  V o;
  o.x = ar;
  o.y = ai;
  reinterpret_cast<V*>(y)[t] = o;
}

// ADOs below the top tier: one workgroup per ADO.  EL lanes cover the ADO's
// packed elements, 256/EL slices split its coupling list (staged in LDS).
template <typename T, bool TRANS>
__global__ void __launch_bounds__(256)
k_apply_packed_heavy(const DevPacked op, const T* __restrict__ x, T* __restrict__ y, int64_t zero_at) {
  using V = typename P2<T>::type;
  constexpr int kStage = 256;
  __shared__ T red_r[256], red_i[256];
  __shared__ DevEntryP s_en[kStage];
  __shared__ DevRec s_rc[kStage];
  const int64_t a = blockIdx.x;
  const int64_t base = op.off[a];
  const int nnz = (int)(op.off[a + 1] - base);
  if (nnz == 0) return;   // uniform across the block, before any barrier
  int EL = 1;
  while (EL < nnz && EL < 256) EL <<= 1;
  const int S = 256 / EL;
  const int lane = threadIdx.x % EL, slice = threadIdx.x / EL;
  const V* __restrict__ xv = reinterpret_cast<const V*>(x);
  const int s = op.sec[a];
  const int64_t p0 = op.rowptr[a], p1 = op.rowptr[a + 1];
  const int nent = (int)(p1 - p0);
  const bool staged = nent <= kStage;
  if (staged)
    for (int k = threadIdx.x; k < nent; k += 256) {
      const DevEntryP e0 = op.ent[p0 + k];
      s_en[k] = e0;
      s_rc[k] = op.rec[e0.rec];
    }
  __syncthreads();
  for (int t0 = 0; t0 < nnz; t0 += EL) {
    const int t = t0 + lane;
    T ar = 0, ai = 0;
    if (t < nnz) {
      if (slice == 0) diag_term<T, TRANS>(op, xv, a, base, s, t, zero_at, ar, ai);
      for (int k = slice; k < nent; k += S) {
        if (staged) {
          coupling_term<T, TRANS>(s_en[k], s_rc[k], op.maps, xv, t, zero_at, ar, ai);
        } else {
          const DevEntryP en = op.ent[p0 + k];
          coupling_term<T, TRANS>(en, op.rec[en.rec], op.maps, xv, t, zero_at, ar, ai);
        }
      }
    }
    red_r[threadIdx.x] = ar;
    red_i[threadIdx.x] = ai;
    __syncthreads();
    if (slice == 0 && t < nnz) {
      T sr = 0, si = 0;
      for (int k = 0; k < S; ++k) {   // fixed order: deterministic
        sr += red_r[k * EL + lane];
        si += red_i[k * EL + lane];
      }
      V o;
      o.x = sr;
      o.y = si;
      reinterpret_cast<V*>(y)[base + t] = o;
    }
    __syncthreads();
  }
}

// Tier-0 fix-ups (one block): trace row (forward: y(d,d) = tr x0; transposed:
// z0(i,i) += w(d,d)) followed by Hermitianisation of the tier-0 block.
template <typename T>
__device__ inline void herm0(const DevPacked& op, T* y) {
  const int D = op.d, DD = D * D;
  const int n0 = (int)(op.off[1] - op.off[0]);
  const int16_t* pos = op.sec_pos + op.sec[0] * DD;
  const int16_t* el = op.sec_elems + op.sec[0] * DD;
  for (int t = threadIdx.x; t < n0; t += blockDim.x) {
    const int e = el[t];
    const int i = e % D, j = e / D;
    if (i > j) continue;
    const int q = pos[j + i * D];
    if (q < 0) continue;
    const T re = (T)0.5 * (y[2 * t] + y[2 * q]);
    const T im = (T)0.5 * (y[2 * t + 1] - y[2 * q + 1]);
    y[2 * t] = re;
    y[2 * t + 1] = im;
    y[2 * q] = re;
    y[2 * q + 1] = -im;
  }
}

template <typename T>
__global__ void k_tier0_fixup(const DevPacked op, const T* __restrict__ x, T* __restrict__ y, int rdo) {
  const int D = op.d, DD = D * D;
  const int16_t* pos = op.sec_pos + op.sec[0] * DD;
  __shared__ T tr[2];
  if (rdo) {
    if (threadIdx.x == 0) {
      if (!op.transposed) {
        T sr = 0, si = 0;
        for (int i = 0; i < D; ++i) {
          const int q = pos[i + i * D];
          if (q >= 0) { sr += x[2 * q]; si += x[2 * q + 1]; }
        }
        y[2 * op.trace_slot] = sr;
        y[2 * op.trace_slot + 1] = si;
      } else {
        tr[0] = x[2 * op.trace_slot];
        tr[1] = x[2 * op.trace_slot + 1];
      }
    }
    __syncthreads();
    if (op.transposed)
      for (int i = threadIdx.x; i < D; i += blockDim.x) {
        const int q = pos[i + i * D];
        if (q >= 0) { y[2 * q] += tr[0]; y[2 * q + 1] += tr[1]; }
      }
    __syncthreads();
  }
  herm0<T>(op, y);
}

template <typename T>
__global__ void k_herm_packed(const DevPacked op, T* __restrict__ y) {
  herm0<T>(op, y);
}

__device__ inline int64_t find_ado(const DevPacked& op, int64_t t) { return find_ado_tab(op, op.block_ado, t); }

template <typename TD, typename TP>
__global__ void k_pack(const DevPacked op, const TD* __restrict__ dense, TP* __restrict__ packed) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= op.total) return;
  const int64_t a = find_ado(op, t);
  const int DD = op.d * op.d;
  const int e = op.sec_elems[op.sec[a] * DD + (int)(t - op.off[a])];
  const int64_t src = a * DD + e;
  packed[2 * t] = (TP)dense[2 * src];
  packed[2 * t + 1] = (TP)dense[2 * src + 1];
}

template <typename TP, typename TD>
__global__ void k_unpack(const DevPacked op, const TP* __restrict__ packed, TD* __restrict__ dense) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= op.total) return;
  const int64_t a = find_ado(op, t);
  const int DD = op.d * op.d;
  const int e = op.sec_elems[op.sec[a] * DD + (int)(t - op.off[a])];
  const int64_t dst = a * DD + e;
  dense[2 * dst] = (TD)packed[2 * t];
  dense[2 * dst + 1] = (TD)packed[2 * t + 1];
}

template <typename TD>
__global__ void k_zero(int64_t n, TD* __restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = 0;
}

template <typename T>
static hipError_t launch_p(const DevPacked& op, const T* x, T* y, bool rdo, hipStream_t s) {
  if (rdo && op.trace_slot < 0) return hipErrorInvalidValue;
  if (op.total <= 0) return hipSuccess;
  const int64_t zero_at = (op.transposed && rdo) ? op.trace_slot : -1;   // ADO 0 starts at packed offset 0
  hipError_t err;
  if (op.n_heavy > 0) {
    if (op.transposed)
      hipLaunchKernelGGL((k_apply_packed_heavy<T, true>), dim3((unsigned)op.n_heavy), dim3(256), 0, s, op, x, y,
                         zero_at);
    else
      hipLaunchKernelGGL((k_apply_packed_heavy<T, false>), dim3((unsigned)op.n_heavy), dim3(256), 0, s, op, x, y,
                         zero_at);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  const int64_t light = op.total - op.light_begin;
  if (light > 0) {
    const int64_t blocks = (light + 255) / 256;
    if (op.transposed)
      hipLaunchKernelGGL((k_apply_packed<T, true>), dim3((unsigned)blocks), dim3(256), 0, s, op, x, y, zero_at);
    else
      hipLaunchKernelGGL((k_apply_packed<T, false>), dim3((unsigned)blocks), dim3(256), 0, s, op, x, y, zero_at);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  hipLaunchKernelGGL((k_tier0_fixup<T>), dim3(1), dim3(256), 0, s, op, x, y, rdo ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_apply_packed_f64(const DevPacked& op, const double* x, double* y, bool rdo, hipStream_t s) {
  return launch_p<double>(op, x, y, rdo, s);
}
hipError_t launch_apply_packed_f32(const DevPacked& op, const float* x, float* y, bool rdo, hipStream_t s) {
  return launch_p<float>(op, x, y, rdo, s);
}

template <typename TD, typename TP>
hipError_t launch_pack(const DevPacked& op, const TD* dense, TP* packed, hipStream_t s) {
  if (op.total <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_pack<TD, TP>), dim3((unsigned)((op.total + 255) / 256)), dim3(256), 0, s, op, dense, packed);
  return hipGetLastError();
}

template <typename TP, typename TD>
hipError_t launch_unpack(const DevPacked& op, const TP* packed, TD* dense, hipStream_t s) {
  const int64_t n = 2 * op.n_ado * op.d * op.d;
  hipLaunchKernelGGL((k_zero<TD>), dim3(1024), dim3(256), 0, s, n, dense);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || op.total <= 0) return e;
  hipLaunchKernelGGL((k_unpack<TP, TD>), dim3((unsigned)((op.total + 255) / 256)), dim3(256), 0, s, op, packed, dense);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_herm_packed(const DevPacked& op, T* x, hipStream_t s) {
  hipLaunchKernelGGL((k_herm_packed<T>), dim3(1), dim3(256), 0, s, op, x);
  return hipGetLastError();
}

template hipError_t launch_pack<double, double>(const DevPacked&, const double*, double*, hipStream_t);
template hipError_t launch_unpack<double, double>(const DevPacked&, const double*, double*, hipStream_t);
template hipError_t launch_pack<double, float>(const DevPacked&, const double*, float*, hipStream_t);
template hipError_t launch_unpack<float, double>(const DevPacked&, const float*, double*, hipStream_t);
template hipError_t launch_herm_packed<double>(const DevPacked&, double*, hipStream_t);
template hipError_t launch_herm_packed<float>(const DevPacked&, float*, hipStream_t);

}  // namespace heomq
