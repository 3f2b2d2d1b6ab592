// HEOM Liouvillian apply on gfx950: one table-driven gather kernel serves the
// forward operator (tfqmr_double/calc_ax_tfqmr_omp.f90, td/calcderiv_omp.f90)
// and its real transpose (bicg/calc_atrx_bicg_omp.f90 in the reference's
// BiCG convention).
//
// Mapping: the d*d elements of one output ADO are spread over min(d*d, 256)
// lanes of a 256-thread workgroup (d = 4: 16 lanes per ADO, four ADOs per
// wave64; d = 16: one ADO per workgroup; d = 64: 16 elements per lane).
// Element (i, j) lives at complex offset i + j*d (the reference's Fortran
// order), so consecutive lanes read consecutive rows: a column of a neighbour
// block is one coalesced segment.  The ADO's own block is staged in LDS for
// the H commutator; neighbour blocks are gathered straight from HBM/L2, one
// element per lane per term, because c and c^+ are signed partial
// permutations: (c X)(i, j) = s_i * X(p_i, j).
#include "heomq_kernels.hpp"

namespace heomq {

template <typename T> struct V2;
template <> struct V2<double> { using type = double2; };
template <> struct V2<float> { using type = float2; };

constexpr int kMaxOps = 12;   // 2 * norbs * nspin for norbs <= 3

template <int D, typename T, bool TRANS>
__global__ void __launch_bounds__(256)
k_apply(const DevOperator op, const T* __restrict__ x, T* __restrict__ y, int64_t trace_pos, int project) {
  using V = typename V2<T>::type;
  constexpr int E = D * D;
  constexpr int LANES = E < 256 ? E : 256;
  constexpr int EPT = E / LANES;
  constexpr int APB = 256 / LANES;
  __shared__ V xs[APB * E];
  __shared__ int8_t sop[kMaxOps * 4 * D];

  const int tid = threadIdx.x;
  for (int t = tid; t < op.nop * 4 * D; t += 256) sop[t] = op.opmap[t];
  const int slot = tid / LANES;
  const int lane = tid - slot * LANES;
  const int64_t a = (int64_t)blockIdx.x * APB + slot;
  const bool active = a < op.n_ado;
  const V* __restrict__ xv = reinterpret_cast<const V*>(x);
  if (active) {
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = lane + k * LANES;
      V v = xv[a * E + e];
      if (TRANS && a == trace_pos && e == op.trace_elem) { v.x = 0; v.y = 0; }
      xs[slot * E + e] = v;
    }
  }
  __syncthreads();
  if (!active) return;   // no barrier follows
  const V* self = &xs[slot * E];

  T ar[EPT], ai[EPT];
  const T gr = (T)op.gam[2 * a], gi = (T)op.gam[2 * a + 1];
  const T hs = (T)op.hsign;
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = lane + k * LANES;
    const int i = e % D, j = e / D;
    const V xe = self[e];
    T cr = 0, ci = 0;
    const int ni = op.hnnz[i];
    for (int t = 0; t < ni; ++t) {            // (H X)(i, j)
      const int c = op.hcol[i * D + t];
      const T hr = (T)op.hval[2 * (i * D + t)], hi = (T)op.hval[2 * (i * D + t) + 1];
      const V v = self[c + j * D];
      cr += hr * v.x - hi * v.y;
      ci += hr * v.y + hi * v.x;
    }
    const int nj = op.hnnz[j];
    for (int t = 0; t < nj; ++t) {            // (X H)(i, j) = sum_k X(i, k) conj(H(j, k))
      const int c = op.hcol[j * D + t];
      const T hr = (T)op.hval[2 * (j * D + t)], hi = -(T)op.hval[2 * (j * D + t) + 1];
      const V v = self[i + c * D];
      cr -= v.x * hr - v.y * hi;
      ci -= v.x * hi + v.y * hr;
    }
    ar[k] = gr * xe.x - gi * xe.y - hs * ci;
    ai[k] = gr * xe.y + gi * xe.x + hs * cr;
  }

  // couplings: the entry two ahead and the record one ahead are loaded while
  // the current neighbour is gathered
  const int64_t p0 = op.rowptr[a], p1 = op.rowptr[a + 1];
  DevEntry en_next{}, en_after{};
  DevRec rc_next{};
  if (p0 < p1) {
    en_next = op.ent[p0];
    rc_next = op.rec[en_next.rec];
  }
  if (p0 + 1 < p1) en_after = op.ent[p0 + 1];
  for (int64_t p = p0; p < p1; ++p) {
    const DevEntry en = en_next;
    const DevRec rc = rc_next;
    if (p + 1 < p1) {
      en_next = en_after;
      rc_next = op.rec[en_next.rec];
      if (p + 2 < p1) en_after = op.ent[p + 2];
    }
    const V* __restrict__ xn = xv + (int64_t)en.nbr * E;
    const bool dag = en.flags & 1;
    const T cLr = (en.flags & 2) ? -(T)rc.cLr : (T)rc.cLr;
    const T cLi = (en.flags & 2) ? -(T)rc.cLi : (T)rc.cLi;
    const T cRr = (en.flags & 4) ? -(T)rc.cRr : (T)rc.cRr;
    const T cRi = (en.flags & 4) ? -(T)rc.cRi : (T)rc.cRi;
    const bool zchk = TRANS && en.nbr == trace_pos;
    const int8_t* mL = &sop[(rc.opL < 0 ? 0 : rc.opL) * 4 * D];
    const int8_t* mR = &sop[(rc.opR < 0 ? 0 : rc.opR) * 4 * D];
#pragma unroll
    for (int k = 0; k < EPT; ++k) {
      const int e = lane + k * LANES;
      const int i = e % D, j = e / D;
      const int pl = rc.opL >= 0 ? mL[i] : -1;
      if (pl >= 0) {
        const int idx = dag ? (j + pl * D) : (pl + j * D);
        V v = xn[idx];
        if (dag) v.y = -v.y;
        if (zchk && idx == op.trace_elem) { v.x = 0; v.y = 0; }
        const T s = (T)mL[D + i];
        ar[k] += s * (cLr * v.x - cLi * v.y);
        ai[k] += s * (cLr * v.y + cLi * v.x);
      }
      const int pr = rc.opR >= 0 ? mR[2 * D + j] : -1;
      if (pr >= 0) {
        const int idx = dag ? (pr + i * D) : (i + pr * D);
        V v = xn[idx];
        if (dag) v.y = -v.y;
        if (zchk && idx == op.trace_elem) { v.x = 0; v.y = 0; }
        const T s = (T)mR[3 * D + j];
        ar[k] += s * (cRr * v.x - cRi * v.y);
        ai[k] += s * (cRr * v.y + cRi * v.x);
      }
    }
  }
  int tpar = 0;
  if (project) {
    int n = 0;
    while (n + 1 < op.tb.ntier && op.tb.first[n + 1] <= a) ++n;
    tpar = n & 1;
  }
  V* __restrict__ yv = reinterpret_cast<V*>(y);
#pragma unroll
  for (int k = 0; k < EPT; ++k) {
    const int e = lane + k * LANES;
    V o;
    o.x = ar[k];
    o.y = ai[k];
    if (project && ((op.parity[e % D] + op.parity[e / D] + tpar) & 1)) { o.x = 0; o.y = 0; }
    yv[a * E + e] = o;
  }
}

// y0 = (y0 + y0^H) / 2 on the tier-0 block (one workgroup).
template <typename T>
__global__ void k_hermitize_block0(T* __restrict__ y, int d) {
  for (int e = threadIdx.x; e < d * d; e += blockDim.x) {
    const int i = e % d, j = e / d;
    if (i > j) continue;
    const int t = j + i * d;
    const T re = (T)0.5 * (y[2 * e] + y[2 * t]);
    const T im = (T)0.5 * (y[2 * e + 1] - y[2 * t + 1]);
    y[2 * e] = re;
    y[2 * e + 1] = im;
    y[2 * t] = re;
    y[2 * t + 1] = -im;
  }
}

// Forward trace row: y(d,d,trace_pos) = tr x0 (calc_ax_tfqmr_omp.f90:492-495).
template <typename T>
__global__ void k_trace_fwd(const T* __restrict__ x, T* __restrict__ y, int d, int64_t pos, int elem) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  T sr = 0, si = 0;
  for (int i = 0; i < d; ++i) {
    sr += x[2 * (i + i * d)];
    si += x[2 * (i + i * d) + 1];
  }
  const int64_t o = 2 * ((int64_t)pos * d * d + elem);
  y[o] = sr;
  y[o + 1] = si;
}

// Transposed trace row: z0(i,i) += w(d,d,trace_pos).
template <typename T>
__global__ void k_trace_trn(const T* __restrict__ w, T* __restrict__ z, int d, int64_t pos, int elem) {
  const int i = threadIdx.x;
  if (i >= d || blockIdx.x != 0) return;
  const int64_t o = 2 * ((int64_t)pos * d * d + elem);
  z[2 * (i + i * d)] += w[o];
  z[2 * (i + i * d) + 1] += w[o + 1];
}

template <int D, typename T>
static hipError_t launch_d(const DevOperator& op, const T* x, T* y, int64_t trace_pos, hipStream_t s,
                           bool project) {
  constexpr int E = D * D;
  constexpr int LANES = E < 256 ? E : 256;
  constexpr int APB = 256 / LANES;
  if (op.n_ado <= 0) return hipSuccess;
  const int64_t blocks = (op.n_ado + APB - 1) / APB;
  const int pj = project ? 1 : 0;
  if (op.transposed)
    hipLaunchKernelGGL((k_apply<D, T, true>), dim3((unsigned)blocks), dim3(256), 0, s, op, x, y, trace_pos, pj);
  else
    hipLaunchKernelGGL((k_apply<D, T, false>), dim3((unsigned)blocks), dim3(256), 0, s, op, x, y, trace_pos, pj);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  if (trace_pos >= 0) {
    if (op.transposed)
      hipLaunchKernelGGL((k_trace_trn<T>), dim3(1), dim3(64), 0, s, x, y, D, trace_pos, op.trace_elem);
    else
      hipLaunchKernelGGL((k_trace_fwd<T>), dim3(1), dim3(64), 0, s, x, y, D, trace_pos, op.trace_elem);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  if (project) {
    hipLaunchKernelGGL((k_hermitize_block0<T>), dim3(1), dim3(256), 0, s, y, D);
    err = hipGetLastError();
  }
  return err;
}

template <typename T>
static hipError_t launch_any(const DevOperator& op, const T* x, T* y, int64_t trace_pos, hipStream_t s,
                             bool project) {
  if (project && !op.parity) return hipErrorInvalidValue;
  switch (op.d) {
    case 4: return launch_d<4, T>(op, x, y, trace_pos, s, project);
    case 16: return launch_d<16, T>(op, x, y, trace_pos, s, project);
    case 64: return launch_d<64, T>(op, x, y, trace_pos, s, project);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_apply_f64(const DevOperator& op, const double* x, double* y, int64_t trace_pos, hipStream_t s,
                            bool project) {
  return launch_any<double>(op, x, y, trace_pos, s, project);
}
hipError_t launch_apply_f32(const DevOperator& op, const float* x, float* y, int64_t trace_pos, hipStream_t s,
                            bool project) {
  return launch_any<float>(op, x, y, trace_pos, s, project);
}

// ---------------------------------------------------------------------------
// BiCG convention around the same operator (tier-0 block only; element (i, j)
// at i + j*d; lower triangle i > j).
// ---------------------------------------------------------------------------
// x0: real diagonal, upper = conj(lower)^T (calc_ax_bicg_omp.f90:70-77)
__global__ void k_bicg_herm_in(int d, double* __restrict__ x) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d * d) return;
  const int i = e % d, j = e / d;
  if (i == j) x[2 * e + 1] = 0.0;
  else if (i < j) {   // upper element reads the lower one, which no thread writes
    const int t = j + i * d;
    x[2 * e] = x[2 * t];
    x[2 * e + 1] = -x[2 * t + 1];
  }
}
// y0 (or w0): upper triangle zeroed, diagonal real (calc_ax_bicg_omp.f90:330-337;
// calc_atrx_bicg_omp.f90:103-110)
__global__ void k_bicg_lower(int d, double* __restrict__ y) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d * d) return;
  const int i = e % d, j = e / d;
  if (i == j) y[2 * e + 1] = 0.0;
  else if (i < j) { y[2 * e] = 0.0; y[2 * e + 1] = 0.0; }
}
// z0 folded onto its lower triangle: the transpose of k_bicg_herm_in
// (lower += conj(upper)^T, diagonal real, upper zero); each lower thread owns
// its mirror element
__global__ void k_bicg_fold(int d, double* __restrict__ z) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= d * d) return;
  const int i = e % d, j = e / d;
  if (i == j) z[2 * e + 1] = 0.0;
  else if (i > j) {
    const int t = j + i * d;
    z[2 * e] += z[2 * t];
    z[2 * e + 1] -= z[2 * t + 1];
    z[2 * t] = 0.0;
    z[2 * t + 1] = 0.0;
  }
}

hipError_t launch_herm_from_lower_f64(int d, double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_herm_in, dim3((d * d + 255) / 256), dim3(256), 0, s, d, x);
  return hipGetLastError();
}

hipError_t launch_lower_tier0_f64(int d, double* x, hipStream_t s) {
  hipLaunchKernelGGL(k_bicg_lower, dim3((d * d + 255) / 256), dim3(256), 0, s, d, x);
  return hipGetLastError();
}

hipError_t launch_apply_bicg_f64(const DevOperator& op, double* x, double* y, hipStream_t s) {
  const int d = op.d;
  hipLaunchKernelGGL(k_bicg_herm_in, dim3((d * d + 255) / 256), dim3(256), 0, s, d, x);
  if (hipError_t e = hipGetLastError()) return e;
  DevOperator o = op;
  o.trace_elem = 0;   // y(1,1) of tier 0 = tr rho0 (calc_ax_bicg_omp.f90:321-324)
  if (hipError_t e = launch_any<double>(o, x, y, 0, s, false)) return e;
  hipLaunchKernelGGL(k_bicg_lower, dim3((d * d + 255) / 256), dim3(256), 0, s, d, y);
  return hipGetLastError();
}

hipError_t launch_apply_transpose_bicg_f64(const DevOperator& op, double* w, double* z, hipStream_t s) {
  const int d = op.d;
  hipLaunchKernelGGL(k_bicg_lower, dim3((d * d + 255) / 256), dim3(256), 0, s, d, w);
  if (hipError_t e = hipGetLastError()) return e;
  DevOperator o = op;
  o.trace_elem = 0;   // w(1,1) is the trace row's dual: kept out of L^T, added to the diagonal of z0
  if (hipError_t e = launch_any<double>(o, w, z, 0, s, false)) return e;
  hipLaunchKernelGGL(k_bicg_fold, dim3((d * d + 255) / 256), dim3(256), 0, s, d, z);
  return hipGetLastError();
}

}  // namespace heomq
