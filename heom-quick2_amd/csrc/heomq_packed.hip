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
#include <algorithm>

#include "heomq_kernels.hpp"
#include "heomq_shadow.hpp"
#include "heomq_solver.hpp"

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

// Element `g` of x, from the LDS window [win_lo, win_lo + win_n) when it lies
// there (the window already holds the trace-row zero of the transposed apply).
template <typename T, bool TRANS>
__device__ inline typename P2<T>::type ldw(const typename P2<T>::type* __restrict__ xv, int64_t g, int64_t zero_at,
                                          const typename P2<T>::type* s_x, int64_t win_lo, int win_n) {
  const int64_t r = g - win_lo;
  if (s_x != nullptr && r >= 0 && r < win_n) return s_x[r];
  return ld<T, TRANS>(xv, g, zero_at);
}

// A Chebyshev term applied where the apply writes its output (no y in
// memory): with y = L p1 for element g,  p[g] += a y;  x[g] += b p[g]
// (td/chebyshev.f90's P_k = P_{k-2} + (2 / delta) L P_{k-1} over P_{k-2},
// then x += 2 J_k P_k), the same arithmetic as the two vector passes.
struct ChebEpi {
  double* p;
  double* x;
  double a, b;
};
__device__ inline void cheb_put(const ChebEpi* ep, int64_t g, double re, double im) {
  double2* pv = reinterpret_cast<double2*>(ep->p) + g;
  double2* xv = reinterpret_cast<double2*>(ep->x) + g;
  double2 q = *pv;
  q.x += ep->a * re;
  q.y += ep->a * im;
  *pv = q;
  double2 w = *xv;
  w.x += ep->b * q.x;
  w.y += ep->b * q.y;
  *xv = w;
}

// The coefficient arrays of one problem of a batch.
struct Coefs {
  const DevRec* rec;
  const double* gam;
  const double* hc_ell;
};

// gamma_K x_t + hsign * i(HX - XH)_t for packed slot t of the ADO at `base`.
// Fixed-width rows: the eight slot indices of a row come in one 16-byte load,
// the coefficients follow, and the neighbour elements (same ADO) are read from
// the LDS window when it holds them.
template <typename T, bool TRANS>
__device__ inline void diag_term(const DevPacked& op, const double* __restrict__ hc_ell,
                                 const typename P2<T>::type* __restrict__ xv, T gr, T gi,
                                 int64_t base, int s, int t, int64_t zero_at, T& ar, T& ai,
                                 const typename P2<T>::type* s_x, int64_t win_lo, int win_n, T& csr, T& csi) {
  using V = typename P2<T>::type;
  const int dd = op.d * op.d;
  const V xe = ldw<T, TRANS>(xv, base + t, zero_at, s_x, win_lo, win_n);
  ar = gr * xe.x - gi * xe.y;
  ai = gr * xe.y + gi * xe.x;
  T cr = 0, ci = 0;
  csr = 0;
  csi = 0;
  const int W = op.hell_w;
  if (W > 0) {
    const int64_t row = (int64_t)s * dd + t;
    const int4* qrow = reinterpret_cast<const int4*>(op.hq_ell + row * op.hell_wp);
    const double2* crow = reinterpret_cast<const double2*>(hc_ell) + row * W;
    for (int k0 = 0; k0 < W; k0 += 8) {
      const int4 qq = qrow[k0 >> 3];
      const int packed[4] = {qq.x, qq.y, qq.z, qq.w};
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (k0 + k >= W) break;
        const int q = (int)(int16_t)((packed[k >> 1] >> ((k & 1) * 16)) & 0xffff);
        const double2 h = crow[k0 + k];
        const V v = ldw<T, TRANS>(xv, base + q, zero_at, s_x, win_lo, win_n);
        cr += (T)h.x * v.x - (T)h.y * v.y;
        ci += (T)h.x * v.y + (T)h.y * v.x;
        if (q == t) { csr += (T)h.x; csi += (T)h.y; }
      }
    }
  } else {
    const int r0 = op.hrow[s * (dd + 1) + t], r1 = op.hrow[s * (dd + 1) + t + 1];
    for (int u = r0; u < r1; ++u) {
      const V v = ldw<T, TRANS>(xv, base + op.hq[u], zero_at, s_x, win_lo, win_n);
      const T hr = (T)op.hc[2 * u], hi = (T)op.hc[2 * u + 1];
      cr += hr * v.x - hi * v.y;
      ci += hr * v.y + hi * v.x;
      if (op.hq[u] == t) { csr += hr; csi += hi; }
    }
  }
  const T hs = (T)op.hsign;
  ar += hs * cr;
  ai += hs * ci;
}

// Left Jacobi scaling of one output element: divide by gamma + hsign * c_tt.
template <typename T>
__device__ inline void jacobi_scale(const DevPacked& op, T gr, T gi, T csr, T csi, T& ar, T& ai) {
  const T hs = (T)op.hsign;
  const T dr = gr + hs * csr, di = gi + hs * csi;
  const T n2 = dr * dr + di * di;
  if (n2 <= (T)1e-24) return;
  const T xr = (ar * dr + ai * di) / n2, xi = (ai * dr - ar * di) / n2;
  ar = xr;
  ai = xi;
}

// Coefficients (cL, cR) of a record: the first 32 bytes of DevRec.
struct CoefLR {
  double2 l, r;
};
__device__ inline CoefLR load_coef(const DevRec* rec, int k) {
  const double2* p = reinterpret_cast<const double2*>(rec + k);
  return CoefLR{p[0], p[1]};
}

// U couplings at once: every entry's paired gather-map value (one 32-bit load
// for the left and right operator), then every neighbour element, are loaded
// before any of them is used.
template <typename T, bool TRANS, int U>
__device__ inline void coupling_block(const DevEntryP* en, const CoefLR* cf, int nvalid,
                                      const int32_t* __restrict__ maps, const typename P2<T>::type* __restrict__ xv,
                                      int t, int64_t zero_at, T& ar, T& ai) {
  using V = typename P2<T>::type;
  int mL[U], mR[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int mv = (u < nvalid && en[u].map >= 0) ? maps[en[u].map + t] : 0;
    mL[u] = (int)(int16_t)(mv & 0xffff);
    mR[u] = (int)(int16_t)((uint32_t)mv >> 16);
  }
  V vL[U], vR[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int64_t b = u < nvalid ? (int64_t)en[u].base : 0;
    vL[u] = mL[u] != 0 ? ld<T, TRANS>(xv, b + (mL[u] > 0 ? mL[u] : -mL[u]) - 1, zero_at) : V{0, 0};
    vR[u] = mR[u] != 0 ? ld<T, TRANS>(xv, b + (mR[u] > 0 ? mR[u] : -mR[u]) - 1, zero_at) : V{0, 0};
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (mL[u] == 0 && mR[u] == 0) continue;
    const uint8_t fl = en[u].flags;
    const T cj = (fl & 1) ? (T)-1 : (T)1;   // X^H: conjugate the gathered element
    const T sL = ((mL[u] > 0) != ((fl & 2) != 0)) ? (T)1 : (T)-1;
    const T sR = ((mR[u] > 0) != ((fl & 4) != 0)) ? (T)1 : (T)-1;
    const T lx = vL[u].x, ly = cj * vL[u].y, rx = vR[u].x, ry = cj * vR[u].y;
    const T clr = (T)cf[u].l.x, cli = (T)cf[u].l.y, crr = (T)cf[u].r.x, cri = (T)cf[u].r.y;
    ar += sL * (clr * lx - cli * ly) + sR * (crr * rx - cri * ry);
    ai += sL * (clr * ly + cli * lx) + sR * (crr * ry + cri * rx);
  }
}

// One coupling's contribution from its gathered elements (signs in the map
// values and the entry flags, coefficients cL, cR).
template <typename T>
__device__ inline void coupling_accum(int mL, int mR, typename P2<T>::type vL, typename P2<T>::type vR, uint8_t fl,
                                      const CoefLR& cf, T& ar, T& ai) {
  if (mL == 0 && mR == 0) return;
  const T cj = (fl & 1) ? (T)-1 : (T)1;
  const T sL = ((mL > 0) != ((fl & 2) != 0)) ? (T)1 : (T)-1;
  const T sR = ((mR > 0) != ((fl & 4) != 0)) ? (T)1 : (T)-1;
  const T lx = vL.x, ly = cj * vL.y, rx = vR.x, ry = cj * vR.y;
  const T clr = (T)cf.l.x, cli = (T)cf.l.y, crr = (T)cf.r.x, cri = (T)cf.r.y;
  ar += sL * (clr * lx - cli * ly) + sR * (crr * rx - cri * ry);
  ai += sL * (clr * ly + cli * lx) + sR * (crr * ry + cri * rx);
}

// The couplings k0, k0 + S, k0 + 2S, ... of a heavy lane with the entries and
// coefficients staged in LDS, U per batch, software-pipelined: the gather-map
// values of batch b + 1 are loaded while batch b's elements are in flight, so
// each batch waits for one global load instead of two (map, then element).
template <typename T, bool TRANS, int U>
__device__ inline void staged_couplings(const DevEntryP* s_en, const CoefLR* s_cf, int nent, int k0, int S,
                                        const int32_t* __restrict__ maps, const typename P2<T>::type* __restrict__ xv,
                                        int t, int64_t zero_at, T& ar, T& ai) {
  using V = typename P2<T>::type;
  int mv[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int k = k0 + u * S;
    mv[u] = (k < nent && s_en[k].map >= 0) ? maps[s_en[k].map + t] : 0;
  }
  for (; k0 < nent; k0 += S * U) {
    V vL[U], vR[U];
    int mL[U], mR[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * S;
      const int64_t b = k < nent ? (int64_t)s_en[k].base : 0;
      mL[u] = (int)(int16_t)(mv[u] & 0xffff);
      mR[u] = (int)(int16_t)((uint32_t)mv[u] >> 16);
      vL[u] = mL[u] != 0 ? ld<T, TRANS>(xv, b + (mL[u] > 0 ? mL[u] : -mL[u]) - 1, zero_at) : V{0, 0};
      vR[u] = mR[u] != 0 ? ld<T, TRANS>(xv, b + (mR[u] > 0 ? mR[u] : -mR[u]) - 1, zero_at) : V{0, 0};
    }
    const int kn = k0 + S * U;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = kn + u * S;
      mv[u] = (k < nent && s_en[k].map >= 0) ? maps[s_en[k].map + t] : 0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int k = k0 + u * S;
      if (k < nent) coupling_accum<T>(mL[u], mR[u], vL[u], vR[u], s_en[k].flags, s_cf[k], ar, ai);
    }
  }
}

// Top-tier ADOs: one thread per packed element.  A block stages in LDS the
// headers (offset, first and last entry, sector, gamma) of the non-empty ADOs
// it covers (at most 257, checked when the tables are built) and its own 256
// elements of x (the commutator's neighbours); the couplings are then read U
// at a time (U = min(the top tier's coupling count, 2): two couplings in
// flight keep the kernel at 72 VGPRs, 7 waves per SIMD; the coefficient
// records come from L2, which holds them for every block).
// LDS carve-up of the light and heavy bodies (they share one buffer).
constexpr size_t kAlign16(size_t b) { return (b + 15) / 16 * 16; }
constexpr size_t kLightOffP0 = kAlign16(257 * 8), kLightOffP1 = kLightOffP0 + kAlign16(257 * 8),
                 kLightOffSec = kLightOffP1 + kAlign16(257 * 8), kLightOffGam = kLightOffSec + kAlign16(257 * 4),
                 kLightOffX = kLightOffGam + kAlign16(257 * 16), kLightSmem = kLightOffX + 256 * 16;
constexpr size_t kHeavyOffRi = 256 * 8, kHeavyOffEn = 2 * 256 * 8, kHeavyOffCf = kHeavyOffEn + 256 * sizeof(DevEntryP),
                 kHeavyOffX = kHeavyOffCf + 256 * sizeof(CoefLR), kHeavySmem = kHeavyOffX + 256 * 16;
constexpr size_t kApplySmem0 = kLightSmem > kHeavySmem ? kLightSmem : kHeavySmem;

template <typename T, bool TRANS, int U, bool EPI = false>
__device__ inline void light_body(const DevPacked& op, const Coefs cf_p, const T* __restrict__ x, T* __restrict__ y,
                                  int64_t zero_at, int bid, char* smem, bool scaled, const ChebEpi* ep = nullptr) {
  using V = typename P2<T>::type;
  int64_t* s_off = reinterpret_cast<int64_t*>(smem);
  int64_t* s_p0 = reinterpret_cast<int64_t*>(smem + kLightOffP0);
  int64_t* s_p1 = reinterpret_cast<int64_t*>(smem + kLightOffP1);
  int* s_sec = reinterpret_cast<int*>(smem + kLightOffSec);
  double2* s_gam = reinterpret_cast<double2*>(smem + kLightOffGam);
  V* s_x = reinterpret_cast<V*>(smem + kLightOffX);
  const V* __restrict__ xv = reinterpret_cast<const V*>(x);
  const int lo = op.light_block_ado[bid], hi = op.light_block_ado[bid + 1];
  const int cnt = hi - lo + 1;
  for (int k = threadIdx.x; k < cnt; k += 256) {
    const int64_t a = op.light_nz[lo + k];
    s_off[k] = op.off[a];
    s_p0[k] = op.rowptr[a];
    s_p1[k] = op.rowptr[a + 1];
    s_sec[k] = op.sec[a];
    s_gam[k] = reinterpret_cast<const double2*>(cf_p.gam)[a];
  }
  const int64_t win_lo = op.light_begin + (int64_t)bid * 256;
  const int win_n = (int)(op.total - win_lo < 256 ? op.total - win_lo : 256);
  if (threadIdx.x < win_n) s_x[threadIdx.x] = ld<T, TRANS>(xv, win_lo + threadIdx.x, zero_at);
  __syncthreads();
  const int64_t t = win_lo + threadIdx.x;
  if (t >= op.total) return;   // after the only barrier
  int l = 0, h = cnt - 1;
  while (l < h) {
    const int mid = (l + h + 1) >> 1;
    if (s_off[mid] <= t) l = mid; else h = mid - 1;
  }
  const int64_t base = s_off[l];
  const int slot = (int)(t - base);
  const int64_t p0 = s_p0[l], p1 = s_p1[l];
  T ar, ai, csr, csi;
  diag_term<T, TRANS>(op, cf_p.hc_ell, xv, (T)s_gam[l].x, (T)s_gam[l].y, base, s_sec[l], slot, zero_at, ar, ai,
                      s_x, win_lo, win_n, csr, csi);
  for (int64_t p = p0; p < p1; p += U) {
    const int nv = (int)(p1 - p < U ? p1 - p : U);
    DevEntryP en[U];
    CoefLR cf[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (u < nv) {
        en[u] = op.ent[p + u];
        cf[u] = load_coef(cf_p.rec, en[u].rec);
      } else {
        en[u] = DevEntryP{0, -1, -1, 0, 0, 0};
        cf[u] = CoefLR{};
      }
    }
    coupling_block<T, TRANS, U>(en, cf, nv, op.maps, xv, slot, zero_at, ar, ai);
  }
  if (scaled) jacobi_scale<T>(op, (T)s_gam[l].x, (T)s_gam[l].y, csr, csi, ar, ai);
  if (EPI) {
    cheb_put(ep, t, (double)ar, (double)ai);
    return;
  }
  V o;
  o.x = ar;
  o.y = ai;
  reinterpret_cast<V*>(y)[t] = o;
}

template <typename T>
__device__ void tier0_fixup_block(const DevPacked& op, const T* __restrict__ x, T* __restrict__ y, int rdo);

// A thread's apply output held back for a fused preconditioner (instead of
// being written to y): its value, the element's packed index (gbase + t), its
// place among the block's elements (lbase + t, the ADO's slots contiguous),
// the ADO's sector and rate.
struct StageOut {
  double2 v;
  int64_t gbase;
  int lbase, t, s;
  double gr, gi;
  bool act;
};

// ADOs below the top tier: one workgroup per ADO.  The block's 256 lanes are
// laid out as (slice, slot): S = 256 / nnz slices, each walking every S-th
// coupling of the ADO (entries and coefficients staged in LDS), two
// couplings per step; the slices are summed through LDS in a fixed order.
// Block 0 (the tier-0 ADO) then applies the trace row and Hermitianises.
template <typename T, bool TRANS, bool STAGE = false, bool EPI = false>
__device__ inline void heavy_body(const DevPacked& op, const Coefs cf_p, const T* __restrict__ x, T* __restrict__ y,
                                  int64_t zero_at, int rdo, int bid, char* smem, bool scaled,
                                  StageOut* so = nullptr, const ChebEpi* ep = nullptr) {
  using V = typename P2<T>::type;
  constexpr int kStage = 256;
  T* red_r = reinterpret_cast<T*>(smem);
  T* red_i = reinterpret_cast<T*>(smem + kHeavyOffRi);
  DevEntryP* s_en = reinterpret_cast<DevEntryP*>(smem + kHeavyOffEn);
  CoefLR* s_cf = reinterpret_cast<CoefLR*>(smem + kHeavyOffCf);
  V* s_x = reinterpret_cast<V*>(smem + kHeavyOffX);
  const int64_t a = bid;
  const int64_t base = op.off[a];
  const int nnz = (int)(op.off[a + 1] - base);
  const V* __restrict__ xv = reinterpret_cast<const V*>(x);
  const int s = op.sec[a];
  const T gr = (T)cf_p.gam[2 * a], gi = (T)cf_p.gam[2 * a + 1];
  const int64_t p0 = op.rowptr[a], p1 = op.rowptr[a + 1];
  const int nent = (int)(p1 - p0);
  const bool staged = nent <= kStage;
  if (staged)
    for (int k = threadIdx.x; k < nent; k += 256) {
      const DevEntryP e0 = op.ent[p0 + k];
      s_en[k] = e0;
      s_cf[k] = load_coef(cf_p.rec, e0.rec);
    }
  const int win_n = nnz < 256 ? nnz : 256;
  if (threadIdx.x < win_n) s_x[threadIdx.x] = ld<T, TRANS>(xv, base + threadIdx.x, zero_at);
  __syncthreads();
  constexpr int U = 2;
  for (int t0 = 0; t0 < nnz; t0 += 256) {   // nnz is uniform over the block: every thread takes the same trips
    const int w = nnz - t0 < 256 ? nnz - t0 : 256;   // slots in this pass
    const int S = 256 / w;
    const int lane = threadIdx.x % w, slice = threadIdx.x / w;
    const int t = t0 + lane;
    T ar = 0, ai = 0, csr = 0, csi = 0;
    if (slice < S) {
      if (slice == 0)
        diag_term<T, TRANS>(op, cf_p.hc_ell, xv, gr, gi, base, s, t, zero_at, ar, ai, s_x, base, win_n, csr, csi);
      if (staged) {
        staged_couplings<T, TRANS, U>(s_en, s_cf, nent, slice, S, op.maps, xv, t, zero_at, ar, ai);
      } else {   // entries from global memory (ADOs with more than 256 couplings)
        for (int k0 = slice; k0 < nent; k0 += S * U) {
          DevEntryP en[U];
          CoefLR cf[U];
          int nv = 0;
#pragma unroll
          for (int u = 0; u < U; ++u) {
            const int k = k0 + u * S;
            if (k < nent) {
              en[u] = op.ent[p0 + k];
              cf[u] = load_coef(cf_p.rec, en[u].rec);
              nv = u + 1;
            } else {
              en[u] = DevEntryP{0, -1, -1, 0, 0, 0};
              cf[u] = CoefLR{};
            }
          }
          coupling_block<T, TRANS, U>(en, cf, nv, op.maps, xv, t, zero_at, ar, ai);
        }
      }
    }
    red_r[threadIdx.x] = ar;
    red_i[threadIdx.x] = ai;
    __syncthreads();
    // pairwise tree over the S slices (fixed pairing: deterministic); S and
    // every trip count below are uniform over the block
    for (int act = S; act > 1;) {
      const int half = (act + 1) >> 1;
      if (slice < act - half) {
        red_r[threadIdx.x] += red_r[threadIdx.x + half * w];
        red_i[threadIdx.x] += red_i[threadIdx.x + half * w];
      }
      __syncthreads();
      act = half;
    }
    if (threadIdx.x < w) {   // slice 0: these lanes hold slot t's diagonal
      T sr = red_r[threadIdx.x], si = red_i[threadIdx.x];
      if (scaled) jacobi_scale<T>(op, gr, gi, csr, csi, sr, si);
      V o;
      o.x = sr;
      o.y = si;
      if (STAGE) {   // one pass (nnz <= 256): slot t = threadIdx.x
        so->v = make_double2((double)sr, (double)si);
        so->act = true;
      } else if (EPI) {
        cheb_put(ep, base + t, (double)sr, (double)si);
      } else {
        reinterpret_cast<V*>(y)[base + t] = o;
      }
    }
    __syncthreads();
  }
  if (STAGE) {
    so->gbase = base;
    so->lbase = 0;
    so->t = threadIdx.x;
    so->s = s;
    so->gr = (double)gr;
    so->gi = (double)gi;
  } else if (a == 0 && !EPI) {   // (with the epilogue no trace row and no Hermitian fix-up: host check)
    tier0_fixup_block<T>(op, x, y, rdo);   // block-uniform branch
  }
}

// ---------------------------------------------------------------------------
// Restructured apply.
//
// Heavy ADOs, grouped form.  The upper couplings of an ADO below the top tier
// come in runs that share one gather map (one c operator, one neighbour
// sector, one dagger flag; about 14 drawers each: leads x exponents).  As the
// reference does (calc_ax_tfqmr_omp.f90:289-353: accumulate the neighbours,
// then apply c once), each run is first summed in the neighbour's slot space,
// B_L[u] = sum_e wL_e x_e[u] and B_R[u] = sum_e wR_e x_e[u], one lane per
// (run, neighbour slot) pair: the lane's loads are independent and lanes read
// consecutive elements of one neighbour.  Each output slot then takes two LDS
// reads per run through the run's map.  Weights: w = s * c from the entry's
// sign flags and record, conjugated for X^H runs (sum w conj(x) = conj(sum
// conj(w) x)).
template <typename T>
struct HWgt {
  int32_t base, pad;
  T lr, li, rr, ri;
};
constexpr size_t kGOffB = 256 * sizeof(HWgt<double>);                              // accumulator pairs
constexpr size_t kGOffG = kGOffB + 2 * 16 * (size_t)kHeavyPairs;                   // run table
constexpr size_t kGSmem = kGOffG + 16 * (size_t)(kHeavyGroups + 1);

template <typename T, bool TRANS, bool EPI = false>
__device__ inline void heavy_grouped(const DevPacked& op, const Coefs cf_p, const T* __restrict__ x,
                                     T* __restrict__ y, int64_t zero_at, int rdo, int bid, char* smem, bool scaled,
                                     const ChebEpi* ep = nullptr) {
  using V = typename P2<T>::type;
  HWgt<T>* s_w = reinterpret_cast<HWgt<T>*>(smem);
  V* s_B = reinterpret_cast<V*>(smem + kGOffB);
  int4* s_g = reinterpret_cast<int4*>(smem + kGOffG);
  T* red_r = reinterpret_cast<T*>(smem);                       // phase 2: over the staged weights
  T* red_i = reinterpret_cast<T*>(smem + kHeavyOffRi);
  const int64_t a = bid;
  const int64_t base = op.off[a];
  const int nnz = (int)(op.off[a + 1] - base);
  const V* __restrict__ xv = reinterpret_cast<const V*>(x);
  const int s = op.sec[a];
  const T gr = (T)cf_p.gam[2 * a], gi = (T)cf_p.gam[2 * a + 1];
  const int64_t p0 = op.rowptr[a];
  const int nent = (int)(op.rowptr[a + 1] - p0);               // <= kHeavyStage (host check)
  const int g0 = op.hg_off[a];
  const int ng = op.hg_off[a + 1] - g0 - 1;                    // runs; the sentinel follows
  for (int k = threadIdx.x; k <= ng; k += 256) s_g[k] = op.hg[g0 + k];
  for (int k = threadIdx.x; k < nent; k += 256) {
    const DevEntryP e = op.ent[p0 + k];
    const CoefLR c = load_coef(cf_p.rec, e.rec);
    const T fl = (e.flags & 2) ? (T)-1 : (T)1, fr = (e.flags & 4) ? (T)-1 : (T)1;
    const T cj = (e.flags & 1) ? (T)-1 : (T)1;
    HWgt<T> w;
    w.base = e.base;
    w.pad = 0;
    w.lr = fl * (T)c.l.x;
    w.li = cj * fl * (T)c.l.y;
    w.rr = fr * (T)c.r.x;
    w.ri = cj * fr * (T)c.r.y;
    s_w[k] = w;
  }
  __syncthreads();
  // phase 1: the runs' accumulators in neighbour slot space
  const int P = s_g[ng].w;
  for (int l = threadIdx.x; l < P; l += 256) {
    int lo = 0, hi = ng - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((s_g[mid].w & 0x7fffffff) <= l) lo = mid; else hi = mid - 1;
    }
    const int4 G = s_g[lo];
    const int u = l - (G.w & 0x7fffffff);
    T lr = 0, li = 0, rr = 0, ri = 0;
    int e = G.x;
    for (; e + 4 <= G.y; e += 4) {   // four independent loads in flight
      V v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = ld<T, TRANS>(xv, (int64_t)s_w[e + q - p0].base + u, zero_at);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const HWgt<T>& w = s_w[e + q - p0];
        lr += w.lr * v[q].x - w.li * v[q].y;
        li += w.lr * v[q].y + w.li * v[q].x;
        rr += w.rr * v[q].x - w.ri * v[q].y;
        ri += w.rr * v[q].y + w.ri * v[q].x;
      }
    }
    for (; e < G.y; ++e) {
      const V v = ld<T, TRANS>(xv, (int64_t)s_w[e - p0].base + u, zero_at);
      const HWgt<T>& w = s_w[e - p0];
      lr += w.lr * v.x - w.li * v.y;
      li += w.lr * v.y + w.li * v.x;
      rr += w.rr * v.x - w.ri * v.y;
      ri += w.rr * v.y + w.ri * v.x;
    }
    V bl, br;
    bl.x = lr; bl.y = li;
    br.x = rr; br.y = ri;
    s_B[2 * l] = bl;
    s_B[2 * l + 1] = br;
  }
  __syncthreads();
  // phase 2: output slots, the runs sliced over the lanes, slices summed as
  // in the sliced form (fixed pairing)
  for (int t0 = 0; t0 < nnz; t0 += 256) {
    const int w = nnz - t0 < 256 ? nnz - t0 : 256;
    const int S = 256 / w;
    const int lane = threadIdx.x % w, slice = threadIdx.x / w;
    const int t = t0 + lane;
    T ar = 0, ai = 0, csr = 0, csi = 0;
    if (slice < S) {
      if (slice == 0)
        diag_term<T, TRANS>(op, cf_p.hc_ell, xv, gr, gi, base, s, t, zero_at, ar, ai, nullptr, 0, 0, csr, csi);
      for (int g = slice; g < ng; g += S) {
        const int4 G = s_g[g];
        const int ub = G.w & 0x7fffffff;
        const T cj = G.w < 0 ? (T)-1 : (T)1;   // X^H run: conj of the accumulator
        const int mv = op.maps[G.z + t];
        const int mL = (int)(int16_t)(mv & 0xffff), mR = (int)(int16_t)((uint32_t)mv >> 16);
        if (mL != 0) {
          const V b = s_B[2 * (ub + (mL > 0 ? mL : -mL) - 1)];
          const T sg = mL > 0 ? (T)1 : (T)-1;
          ar += sg * b.x;
          ai += sg * cj * b.y;
        }
        if (mR != 0) {
          const V b = s_B[2 * (ub + (mR > 0 ? mR : -mR) - 1) + 1];
          const T sg = mR > 0 ? (T)1 : (T)-1;
          ar += sg * b.x;
          ai += sg * cj * b.y;
        }
      }
    }
    red_r[threadIdx.x] = ar;
    red_i[threadIdx.x] = ai;
    __syncthreads();
    for (int act = S; act > 1;) {
      const int half = (act + 1) >> 1;
      if (slice < act - half) {
        red_r[threadIdx.x] += red_r[threadIdx.x + half * w];
        red_i[threadIdx.x] += red_i[threadIdx.x + half * w];
      }
      __syncthreads();
      act = half;
    }
    if (threadIdx.x < w) {
      T sr = red_r[threadIdx.x], si = red_i[threadIdx.x];
      if (scaled) jacobi_scale<T>(op, gr, gi, csr, csi, sr, si);
      V o;
      o.x = sr;
      o.y = si;
      if (EPI)
        cheb_put(ep, base + t, (double)sr, (double)si);
      else
        reinterpret_cast<V*>(y)[base + t] = o;
    }
    __syncthreads();
  }
  if (a == 0) tier0_fixup_block<T>(op, x, y, rdo);   // block-uniform branch
}

// Heavy ADOs of small sectors (at most 16 elements, e.g. every ADO of the
// single-impurity model): one wave per ADO, four ADOs per workgroup.  The
// wave's 64 lanes are (slice, slot) with the slot count rounded up to a power
// of two w, so 64 / w slices walk the coupling list (entries, records and
// gather maps from global memory, two couplings in flight); the slices are
// summed by shuffles in a fixed tree, with no LDS and no barrier.
template <typename T, bool TRANS, bool EPI = false>
__device__ inline void heavy_wave(const DevPacked& op, const Coefs cf_p, const T* __restrict__ x, T* __restrict__ y,
                                  int64_t zero_at, int64_t a, bool scaled, const ChebEpi* ep = nullptr) {
  using V = typename P2<T>::type;
  const int lane = threadIdx.x & 63;
  const int64_t base = op.off[a];
  const int nnz = (int)(op.off[a + 1] - base);
  if (nnz == 0) return;   // wave-uniform
  int w = 1;
  while (w < nnz) w <<= 1;   // <= kHeavyWaveMaxNnz (host check)
  const int S = 64 / w, slice = lane / w, t = lane - slice * w;
  const bool act = t < nnz;
  const V* __restrict__ xv = reinterpret_cast<const V*>(x);
  const int s = op.sec[a];
  const T gr = (T)cf_p.gam[2 * a], gi = (T)cf_p.gam[2 * a + 1];
  const int64_t p0 = op.rowptr[a];
  const int nent = (int)(op.rowptr[a + 1] - p0);
  T ar = 0, ai = 0, csr = 0, csi = 0;
  if (act) {
    if (slice == 0)
      diag_term<T, TRANS>(op, cf_p.hc_ell, xv, gr, gi, base, s, t, zero_at, ar, ai, nullptr, 0, 0, csr, csi);
    constexpr int U = 2;
    for (int k0 = slice; k0 < nent; k0 += S * U) {
      DevEntryP en[U];
      CoefLR cf[U];
      int nv = 0;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u * S;
        if (k < nent) {
          en[u] = op.ent[p0 + k];
          cf[u] = load_coef(cf_p.rec, en[u].rec);
          nv = u + 1;
        } else {
          en[u] = DevEntryP{0, -1, -1, 0, 0, 0};
          cf[u] = CoefLR{};
        }
      }
      coupling_block<T, TRANS, U>(en, cf, nv, op.maps, xv, t, zero_at, ar, ai);
    }
  }
  // lane = slice * w + t: shuffles by multiples of w fold the slices onto slice 0
  for (int off = 32; off >= w; off >>= 1) {
    ar += __shfl_down(ar, off, 64);
    ai += __shfl_down(ai, off, 64);
  }
  if (slice == 0 && act) {
    if (scaled) jacobi_scale<T>(op, gr, gi, csr, csi, ar, ai);
    if (EPI) {
      cheb_put(ep, base + t, (double)ar, (double)ai);
      return;
    }
    V o;
    o.x = ar;
    o.y = ai;
    reinterpret_cast<V*>(y)[base + t] = o;
  }
}

// Top tier, sector-batched: a block takes 256 / nnz top-tier ADOs of one
// sector, one lane per (ADO, slot).  Each lane loads its ADO's first element,
// rate and couplings from the block's own tables (one level below the block
// descriptor), then every coupling's map value and record, then the lower-tier
// elements, and sums the couplings before any barrier; only its own element
// of x goes through LDS, where the sector's commutator rows (staged once per
// block) read their neighbours.  Four levels of dependent loads in all.
constexpr size_t kLOffHc = 256 * 2, kLOffX = kLOffHc + 256 * 16, kLSmem = kLOffX + 256 * 16;

template <typename T, bool TRANS, bool STAGE = false, bool EPI = false>
__device__ inline void light_sector(const DevPacked& op, const Coefs cf_p, const T* __restrict__ x,
                                    T* __restrict__ y, int64_t zero_at, int bid, char* smem, bool scaled,
                                    StageOut* so = nullptr, const ChebEpi* ep = nullptr) {
  using V = typename P2<T>::type;
  int16_t* s_hq = reinterpret_cast<int16_t*>(smem);
  double2* s_hc = reinterpret_cast<double2*>(smem + kLOffHc);
  V* s_x = reinterpret_cast<V*>(smem + kLOffX);
  const V* __restrict__ xv = reinterpret_cast<const V*>(x);
  const int4 bd = op.lt_blk[bid];
  const int cnt = bd.y, nnz = bd.z, s = bd.w, U = op.lt_u;
  const int dd = op.d * op.d;
  const int W = op.hell_w;
  const bool hst = W > 0 && nnz * W <= 256;   // the sector's commutator rows fit the LDS table
  if (hst)
    for (int k = threadIdx.x; k < nnz * W; k += 256) {
      const int row = k / W, kk = k - row * W;
      const int64_t r0 = (int64_t)s * dd + row;
      s_hq[k] = op.hq_ell[r0 * op.hell_wp + kk];
      s_hc[k] = reinterpret_cast<const double2*>(cf_p.hc_ell)[r0 * W + kk];
    }
  const int j = threadIdx.x / nnz, t = threadIdx.x - j * nnz;   // nnz >= 1 (host check)
  const bool act = j < cnt;
  const int64_t i = (int64_t)bd.x + j;
  int64_t base = 0;
  T ar = 0, ai = 0, gr = 0, gi = 0;
  if (act) {
    base = op.lt_off[i];
    const int64_t a = op.lt_id[i];
    s_x[threadIdx.x] = ld<T, TRANS>(xv, base + t, zero_at);
    gr = (T)cf_p.gam[2 * a];
    gi = (T)cf_p.gam[2 * a + 1];
    // the lower-tier couplings, two at a time: entries, then map values and
    // records, then the four elements, then the sums
    const DevEntryP* en = op.lt_ent + i * U;
    for (int r0 = 0; r0 < U; r0 += 2) {
      DevEntryP e[2];
      int mv[2];
      CoefLR cf[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) e[q] = r0 + q < U ? en[r0 + q] : DevEntryP{0, -1, -1, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        mv[q] = e[q].map >= 0 ? op.maps[e[q].map + t] : 0;
        cf[q] = e[q].map >= 0 ? load_coef(cf_p.rec, e[q].rec) : CoefLR{};
      }
      V vl[2], vr[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int mL = (int)(int16_t)(mv[q] & 0xffff), mR = (int)(int16_t)((uint32_t)mv[q] >> 16);
        vl[q] = mL != 0 ? ld<T, TRANS>(xv, (int64_t)e[q].base + (mL > 0 ? mL : -mL) - 1, zero_at) : V{0, 0};
        vr[q] = mR != 0 ? ld<T, TRANS>(xv, (int64_t)e[q].base + (mR > 0 ? mR : -mR) - 1, zero_at) : V{0, 0};
      }
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int mL = (int)(int16_t)(mv[q] & 0xffff), mR = (int)(int16_t)((uint32_t)mv[q] >> 16);
        coupling_accum<T>(mL, mR, vl[q], vr[q], e[q].flags, cf[q], ar, ai);
      }
    }
  }
  __syncthreads();
  if (!act) return;   // after the only barrier
  T csr = 0, csi = 0;
  const V* xw = s_x + j * nnz;   // this ADO's elements
  if (hst) {
    const V xe = xw[t];
    ar += gr * xe.x - gi * xe.y;
    ai += gr * xe.y + gi * xe.x;
    T cr = 0, ci = 0;
    for (int k = 0; k < W; ++k) {
      const int q = s_hq[t * W + k];
      const double2 h = s_hc[t * W + k];
      const V v = xw[q];
      cr += (T)h.x * v.x - (T)h.y * v.y;
      ci += (T)h.x * v.y + (T)h.y * v.x;
      if (q == t) { csr += (T)h.x; csi += (T)h.y; }
    }
    const T hs = (T)op.hsign;
    ar += hs * cr;
    ai += hs * ci;
  } else {
    T dr, di;
    diag_term<T, TRANS>(op, cf_p.hc_ell, xv, gr, gi, base, s, t, zero_at, dr, di, xw, base, nnz, csr, csi);
    ar += dr;
    ai += di;
  }
  if (scaled) jacobi_scale<T>(op, gr, gi, csr, csi, ar, ai);
  if (STAGE) {
    so->v = make_double2((double)ar, (double)ai);
    so->act = true;
    so->gbase = base;
    so->lbase = j * nnz;
    so->t = t;
    so->s = s;
    so->gr = (double)gr;
    so->gi = (double)gi;
    return;
  }
  if (EPI) {
    cheb_put(ep, base + t, (double)ar, (double)ai);
    return;
  }
  V o;
  o.x = ar;
  o.y = ai;
  reinterpret_cast<V*>(y)[base + t] = o;
}

// One launch for the whole apply: blocks [0, n_heavy) take the ADOs below the
// top tier (dispatched first: their work per block is the longest), the rest
// the top tier's elements, so the two phases overlap on the device.  The LDS
// buffer allows 7 blocks per CU; the register budget is held to the same 7
// waves per SIMD (72 VGPRs).
constexpr size_t kApplySmem1 = kGSmem > kLSmem ? kGSmem : kLSmem;
constexpr size_t kApplySmem = kApplySmem0 > kApplySmem1 ? kApplySmem0 : kApplySmem1;

// Block bx of problem p of the apply (the kernels below supply the two).
template <typename T, bool TRANS, int U, bool EPI = false>
__device__ inline void apply_block(const DevPacked& op, int64_t p, int64_t bx, const T* __restrict__ x,
                                   T* __restrict__ y, int64_t zero_at, int rdo, char* smem,
                                   const ChebEpi* ep = nullptr) {
  if (op.frozen_flag && op.frozen_flag[p * op.frozen_stride] == 3) return;   // finished: block-uniform, no barrier yet
  const bool scaled = op.precond && !(op.frozen_flag && op.frozen_flag[p * op.frozen_stride + 2]);
  x += 2 * p * op.vstride;
  y += 2 * p * op.vstride;
  const Coefs cf{op.rec + p * op.rec_stride, op.gam + 2 * p * op.gam_stride, op.hc_ell + 2 * p * op.hc_stride};
  const int64_t hb = heavy_blocks(op);
  if (bx < hb) {   // block-uniform
    if (op.heavy_wave && bx > 0) {   // four ADOs of small sectors, a wave each (no barrier)
      const int64_t a = 1 + (bx - 1) * 4 + (threadIdx.x >> 6);
      if (a < op.n_heavy) heavy_wave<T, TRANS, EPI>(op, cf, x, y, zero_at, a, scaled, ep);
      return;
    }
    const int a = (int)bx;
    if (op.hg_off && op.hg_off[a + 1] > op.hg_off[a])
      heavy_grouped<T, TRANS, EPI>(op, cf, x, y, zero_at, rdo, a, smem, scaled, ep);
    else
      heavy_body<T, TRANS, false, EPI>(op, cf, x, y, zero_at, rdo, a, smem, scaled, nullptr, ep);
  } else if (op.lt_blk) {
    light_sector<T, TRANS, false, EPI>(op, cf, x, y, zero_at, (int)(bx - hb), smem, scaled, nullptr, ep);
  } else {
    light_body<T, TRANS, U, EPI>(op, cf, x, y, zero_at, (int)(bx - hb), smem, scaled, ep);
  }
}

template <typename T, bool TRANS, int U>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7)))
k_apply_packed(const DevPacked op, const T* __restrict__ x, T* __restrict__ y, int64_t zero_at, int rdo) {
  __shared__ __align__(16) char smem[kApplySmem];
  apply_block<T, TRANS, U>(op, blockIdx.y, blockIdx.x, x, y, zero_at, rdo, smem);   // grid y: problem of a batch
}

// One Chebyshev term in one launch: the forward apply of p1 (no trace row,
// no Hermitian fix-up) with each output element folded into the recurrence
// and the sum at once (ChebEpi); L p1 itself is never stored.
template <int U>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7)))
k_apply_packed_cheb(const DevPacked op, const double* __restrict__ p1, const ChebEpi ep) {
  __shared__ __align__(16) char smem[kApplySmem];
  apply_block<double, false, U, true>(op, 0, blockIdx.x, p1, nullptr, -1, 0, smem, &ep);
}

// Forward applies of problems of several shapes in one flat grid (the shapes
// come largest first, so their long heavy workgroups are dispatched first).
template <typename T, int U>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7)))
k_apply_packed_multi(const ApplyMulti am, const T* __restrict__ x, T* __restrict__ y, int rdo) {
  __shared__ __align__(16) char smem[kApplySmem];
  const int64_t b = blockIdx.x;
  int g = 0;
  while (g + 1 < am.ng && am.bstart[g + 1] <= b) ++g;   // block-uniform
  const int64_t loc = b - am.bstart[g], j = loc / am.nb[g];
  const int64_t p = am.pmap[g] ? am.pmap[g][j] : j;
  const DevPacked& op = am.ops[g];
  apply_block<T, false, U>(op, p, loc - j * am.nb[g], x + op.xoff, y + op.xoff, -1, rdo, smem);
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
__device__ void tier0_fixup_block(const DevPacked& op, const T* __restrict__ x, T* __restrict__ y, int rdo) {
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
  if (op.herm0) herm0<T>(op, y);
}

template <typename T>
__global__ void k_tier0_fixup(const DevPacked op, const T* __restrict__ x, T* __restrict__ y, int rdo) {
  tier0_fixup_block<T>(op, x, y, rdo);
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

// Block-Jacobi same-tier inverse (jacobi/calc_ax_jacobi_omp.f90:276-298 on
// tier >= 1, :95-108 on tier 0): forward half.  Thread = packed element
// (i, j) of ADO a; the transform sums over the H-groups of i and j only.
__global__ void k_jacobi_fwd(const DevPacked op, const DevJacobi jc, const double* __restrict__ r,
                             double* __restrict__ t) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= op.total) return;
  const int64_t a = find_ado(op, g);
  const int d = op.d, DD = d * d, s = op.sec[a];
  const int64_t o = op.off[a];
  const int e = op.sec_elems[s * DD + (int)(g - o)];
  const int i = e % d, j = e / d;
  const int ni = jc.gn[i], nj = jc.gn[j];
  const double2* rv = reinterpret_cast<const double2*>(r);
  const double2* uv = reinterpret_cast<const double2*>(jc.u);
  double ar = 0.0, ai = 0.0;
  for (int p = 0; p < ni; ++p) {
    const int k = jc.grp[i * jc.gmax + p];
    const double2 uki = uv[k + i * d];            // conj(U_ki) multiplies
    double br = 0.0, bi = 0.0;                    // sum_l X_kl U_lj
    for (int q = 0; q < nj; ++q) {
      const int l = jc.grp[j * jc.gmax + q];
      const double2 x = rv[o + op.sec_pos[s * DD + k + l * d]];
      const double2 ulj = uv[l + j * d];
      br += x.x * ulj.x - x.y * ulj.y;
      bi += x.x * ulj.y + x.y * ulj.x;
    }
    ar += uki.x * br + uki.y * bi;
    ai += uki.x * bi - uki.y * br;
  }
  // w = 1 / (ejac - gamma_K + i (E_i - E_j))
  const double wr = jc.ejac - op.gam[2 * a], wi = (jc.e[i] - jc.e[j]) - op.gam[2 * a + 1];
  const double den = wr * wr + wi * wi;
  t[2 * g] = (ar * wr + ai * wi) / den;
  t[2 * g + 1] = (ai * wr - ar * wi) / den;
}

// Element (i, j) of U T U^H for one ADO block of the packed vector t
// (block at o, sector s): sum_{p in group(i), q in group(j)} U_ip t_pq conj(U_jq).
__device__ inline double2 jacobi_back_elem(const DevPacked& op, const DevJacobi& jc, const double2* tv, int64_t o,
                                           int s, int i, int j) {
  const int d = op.d, DD = d * d;
  const int ni = jc.gn[i], nj = jc.gn[j];
  const double2* uv = reinterpret_cast<const double2*>(jc.u);
  double ar = 0.0, ai = 0.0;
  for (int p = 0; p < ni; ++p) {
    const int pp = jc.grp[i * jc.gmax + p];
    const double2 uip = uv[i + pp * d];
    double br = 0.0, bi = 0.0;                    // sum_q t_pq conj(U_jq)
    for (int q = 0; q < nj; ++q) {
      const int qq = jc.grp[j * jc.gmax + q];
      const double2 v = tv[o + op.sec_pos[s * DD + pp + qq * d]];
      const double2 ujq = uv[j + qq * d];
      br += v.x * ujq.x + v.y * ujq.y;
      bi += v.y * ujq.x - v.x * ujq.y;
    }
    ar += uip.x * br - uip.y * bi;
    ai += uip.x * bi + uip.y * br;
  }
  return make_double2(ar, ai);
}

// Back half: x_ij (+)= (U t U^H)_ij (accumulate: a Jacobi sweep; otherwise the
// preconditioned vector itself, whose tier-0 block is made exactly Hermitian:
// the physical iterates of the solvers keep a Hermitian rho0, and round-off
// in its anti-Hermitian part is not seen by the projected operator).
__global__ void k_jacobi_back(const DevPacked op, const DevJacobi jc, const double* __restrict__ t,
                              double* __restrict__ x, int accumulate) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= op.total) return;
  const int64_t a = find_ado(op, g);
  const int d = op.d, DD = d * d, s = op.sec[a];
  const int64_t o = op.off[a];
  const int e = op.sec_elems[s * DD + (int)(g - o)];
  const int i = e % d, j = e / d;
  const double2* tv = reinterpret_cast<const double2*>(t);
  const double2 v = jacobi_back_elem(op, jc, tv, o, s, i, j);
  double ar = v.x, ai = v.y;
  if (!accumulate && a == 0) {   // ADO 0 holds every (j, i) with its (i, j)
    const double2 m = jacobi_back_elem(op, jc, tv, o, s, j, i);
    ar = 0.5 * (ar + m.x);
    ai = 0.5 * (ai - m.y);
  }
  if (accumulate) {
    ar += x[2 * g];
    ai += x[2 * g + 1];
  }
  x[2 * g] = ar;
  x[2 * g + 1] = ai;
}

// The same-tier inverse (shift - S_K)^-1 in one kernel (calc_ax_jacobi_omp.f90:
// 276-298 restated in the eigenbasis of H): block b owns output elements
// [256 b, 256 b + 256); the whole ADOs they belong to are staged in LDS, the
// forward transform (U^H X U)_pq / w_pq of every staged element goes to a
// second LDS array, and each thread back-transforms its own element from it.
// Without accumulate the tier-0 block of the output is made exactly
// Hermitian (z_ij from the average of (i, j) and conj of (j, i)).  in and out
// must not alias: neighbouring blocks read the ADOs they share.
// sum over one slot-space list: c[r] * v[base + slot[r]] for r in [r0, r1)
__device__ inline double2 jac_list(const int32_t* row, const int16_t* slot, const double* coef, int idx,
                                   const double2* v, int base) {
  const int r0 = row[idx], r1 = row[idx + 1];
  const double2* cf = reinterpret_cast<const double2*>(coef);
  double ar = 0.0, ai = 0.0;
  for (int r = r0; r < r1; ++r) {
    const double2 c = cf[r];
    const double2 x = v[base + slot[r]];
    ar += c.x * x.x - c.y * x.y;
    ai += c.x * x.y + c.y * x.x;
  }
  return make_double2(ar, ai);
}

__global__ __launch_bounds__(256) void k_jacobi_win(const DevPacked op, const DevJacobi jc,
                                                    const double* __restrict__ in, double* __restrict__ out,
                                                    int accumulate) {
  __shared__ double2 xw[kJacobiWindow];
  __shared__ double2 tw[kJacobiWindow];
  __shared__ int64_t offs[260];
  __shared__ double2 gamw[260];
  __shared__ int secw[260];
  __shared__ uint16_t adoi[kJacobiWindow];   // window element -> index of its ADO in offs
  const int64_t e0 = (int64_t)blockIdx.x * 256;
  const int64_t a_lo = op.block_ado[blockIdx.x], a_hi = op.block_ado[blockIdx.x + 1];
  const int nado = (int)(a_hi - a_lo + 1);   // <= 257, checked on the host
  for (int k = threadIdx.x; k <= nado; k += 256) {
    offs[k] = op.off[a_lo + k];
    if (k < nado) {
      gamw[k] = make_double2(op.gam[2 * (a_lo + k)], op.gam[2 * (a_lo + k) + 1]);
      secw[k] = op.sec[a_lo + k];
    }
  }
  __syncthreads();
  const int64_t w0 = offs[0];
  const int nw = (int)(offs[nado] - w0);     // <= kJacobiWindow, checked on the host
  const double2* xin = reinterpret_cast<const double2*>(in);
  for (int k = threadIdx.x; k < nw; k += 256) xw[k] = xin[w0 + k];
  for (int ia = threadIdx.x; ia < nado; ia += 256)
    for (int k = (int)(offs[ia] - w0), k1 = (int)(offs[ia + 1] - w0); k < k1; ++k) adoi[k] = (uint16_t)ia;
  __syncthreads();
  const int R = op.d * op.d + 1;   // row length of the list tables
  // X -> X U_b
  for (int k = threadIdx.x; k < nw; k += 256) {
    const int ia = adoi[k], base = (int)(offs[ia] - w0);
    tw[k] = jac_list(jc.crow, jc.cslot, jc.ca, secw[ia] * R + (k - base), xw, base);
  }
  __syncthreads();
  // -> (U_a^H .) / (shift - gamma_K + i (E_i - E_j))
  for (int k = threadIdx.x; k < nw; k += 256) {
    const int ia = adoi[k], base = (int)(offs[ia] - w0), s = secw[ia], t = k - base;
    const double2 q = jac_list(jc.rrow, jc.rslot, jc.rb, s * R + t, tw, base);
    const double2 ga = gamw[ia];
    const double wr = jc.ejac - ga.x, wi = jc.tde[s * (R - 1) + t] - ga.y;
    const double den = wr * wr + wi * wi;
    xw[k] = make_double2((q.x * wr + q.y * wi) / den, (q.y * wr - q.x * wi) / den);
  }
  __syncthreads();
  // -> . U_b^H
  for (int k = threadIdx.x; k < nw; k += 256) {
    const int ia = adoi[k], base = (int)(offs[ia] - w0);
    tw[k] = jac_list(jc.crow, jc.cslot, jc.cc, secw[ia] * R + (k - base), xw, base);
  }
  __syncthreads();
  // -> U_a . for the block's own elements
  const int64_t g = e0 + threadIdx.x;
  if (g >= op.total) return;   // after the last barrier
  const int k = (int)(g - w0);
  const int ia = adoi[k], base = (int)(offs[ia] - w0), s = secw[ia], t = k - base;
  double2 z = jac_list(jc.rrow, jc.rslot, jc.rd, s * R + t, tw, base);
  if (!accumulate && a_lo + ia == 0) {   // tier 0: average with conj of the mirror element
    const int e = op.sec_elems[s * (R - 1) + t];
    const int i = e % op.d, j = e / op.d;
    const double2 m = jac_list(jc.rrow, jc.rslot, jc.rd, s * R + op.sec_pos[s * (R - 1) + j + i * op.d], tw, base);
    z.x = 0.5 * (z.x + m.x);
    z.y = 0.5 * (z.y - m.y);
  }
  double2* xo = reinterpret_cast<double2*>(out);
  if (accumulate) {
    const double2 w = xo[g];
    z.x += w.x;
    z.y += w.y;
  }
  xo[g] = z;
}

// The same-tier inverse with blocks aligned to ADOs: block b owns whole ADOs
// (at most 256 elements), one thread per element, so every element is
// transformed once and the four stages run on LDS.  The element's slot lists
// (JacDesc) are loaded as soon as its ADO is known, U and the eigenvalues are
// staged in LDS, and the only dependent global loads are the block bounds, the
// ADO headers and the descriptor.
__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 cmulc(double2 a, double2 b) {   // conj(a) b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// U_ki times a complex value (conj: conj(U_ki) times it); real U reads one double
template <bool RU>
struct UStore;
template <>
struct UStore<false> {
  double2 v[256];
  __device__ void put(int k, double2 u) { v[k] = u; }
  __device__ double2 mul(int k, double2 x) const { return cmul(v[k], x); }
  __device__ double2 mulc(int k, double2 x) const { return cmulc(v[k], x); }
};
template <>
struct UStore<true> {
  double v[256];
  __device__ void put(int k, double2 u) { v[k] = u.x; }
  __device__ double2 mul(int k, double2 x) const { return make_double2(v[k] * x.x, v[k] * x.y); }
  __device__ double2 mulc(int k, double2 x) const { return make_double2(v[k] * x.x, v[k] * x.y); }
};

// TIO: element type of in / out (double, or float for the FP32 inner solves of
// mixed precision); the transform itself runs in FP64 either way.
template <bool RU, typename TIO>
__global__ __launch_bounds__(256) void k_jacobi_ado(const DevPacked op, const DevJacobi jc,
                                                    const TIO* __restrict__ in, TIO* __restrict__ out,
                                                    int accumulate) {
  using VIO = typename P2<TIO>::type;
  __shared__ double2 xw[256];
  __shared__ double2 tw[256];
  __shared__ UStore<RU> us;          // U, u[k + i d] = U_ki (d <= 16)
  __shared__ double ew[16];
  __shared__ int32_t offs[257];      // ADO starts relative to the block's first element
  __shared__ double2 gamw[256];
  __shared__ int secw[256];
  const int tid = threadIdx.x, d = op.d, DD = d * d;
  const int64_t a_lo = jc.blk[2 * blockIdx.x], w0 = jc.blk[2 * blockIdx.x + 1];
  const int nado = (int)(jc.blk[2 * blockIdx.x + 2] - a_lo);   // 1..256, checked on the host
  const int nw = (int)(jc.blk[2 * blockIdx.x + 3] - w0);       // <= 256
  const VIO* xin = reinterpret_cast<const VIO*>(in);
  const bool act = tid < nw;
  double2 xv = make_double2(0.0, 0.0);
  if (act) {
    const VIO v = xin[w0 + tid];
    xv = make_double2((double)v.x, (double)v.y);
  }
  for (int k = tid; k <= nado; k += 256) {
    offs[k] = (int32_t)(op.off[a_lo + k] - w0);
    if (k < nado) {
      gamw[k] = make_double2(op.gam[2 * (a_lo + k)], op.gam[2 * (a_lo + k) + 1]);
      secw[k] = op.sec[a_lo + k];
    }
  }
  if (tid < DD) us.put(tid, reinterpret_cast<const double2*>(jc.u)[tid]);
  if (tid < d) ew[tid] = jc.e[tid];
  __syncthreads();
  // the ADO of this element: the last ia with offs[ia] <= tid
  int ia = 0;
  if (act) {
    int hi = nado - 1;
    while (ia < hi) {
      const int mid = (ia + hi + 1) >> 1;
      if (offs[mid] <= tid) ia = mid;
      else hi = mid - 1;
    }
  }
  const int base = offs[ia], s = secw[ia], t = tid - base;
  JacDesc ds{};
  if (act) ds = jc.desc[s * DD + t];
  xw[tid] = xv;
  __syncthreads();
  // X -> X U_c
  if (act) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < ds.nq) {
        const double2 p = us.mul(ds.c2[q] + ds.c * d, xw[base + ds.cslot[q]]);
        acc.x += p.x;
        acc.y += p.y;
      }
    tw[tid] = acc;
  }
  __syncthreads();
  // -> (U_r^H .) / (shift - gamma_K + i (E_r - E_c))
  if (act) {
    double2 q = make_double2(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < 4; ++p)
      if (p < ds.np) {
        const double2 v = us.mulc(ds.r2[p] + ds.r * d, tw[base + ds.rslot[p]]);
        q.x += v.x;
        q.y += v.y;
      }
    const double2 ga = gamw[ia];
    const double wr = jc.ejac - ga.x, wi = (ew[ds.r] - ew[ds.c]) - ga.y;
    const double den = wr * wr + wi * wi;
    xw[tid] = make_double2((q.x * wr + q.y * wi) / den, (q.y * wr - q.x * wi) / den);
  }
  __syncthreads();
  // -> . U_c^H
  if (act) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < ds.nq) {
        const double2 p = us.mulc(ds.c + ds.c2[q] * d, xw[base + ds.cslot[q]]);
        acc.x += p.x;
        acc.y += p.y;
      }
    tw[tid] = acc;
  }
  __syncthreads();
  // -> U_r .
  double2 z = make_double2(0.0, 0.0);
  if (act) {
#pragma unroll
    for (int p = 0; p < 4; ++p)
      if (p < ds.np) {
        const double2 v = us.mul(ds.r + ds.r2[p] * d, tw[base + ds.rslot[p]]);
        z.x += v.x;
        z.y += v.y;
      }
  }
  if (!accumulate && a_lo == 0) {   // tier 0: average with conj of the mirror element
    xw[tid] = z;
    __syncthreads();
    const int tm = (act && ia == 0) ? op.sec_pos[s * DD + ds.c + ds.r * d] : -1;
    if (tm >= 0) {
      const double2 m = xw[base + tm];
      z.x = 0.5 * (z.x + m.x);
      z.y = 0.5 * (z.y - m.y);
    }
  }
  if (!act) return;
  VIO* xo = reinterpret_cast<VIO*>(out);
  if (accumulate) {
    const VIO w = xo[w0 + tid];
    z.x += (double)w.x;
    z.y += (double)w.y;
  }
  VIO o;
  o.x = (TIO)z.x;
  o.y = (TIO)z.y;
  xo[w0 + tid] = o;
}

// The block-Jacobi-preconditioned apply in one launch, y = M^-1 (Pi A x):
// each workgroup of the apply holds whole ADOs (a heavy ADO, or a
// sector-batched run of top-tier ADOs), so it keeps its outputs in registers,
// reuses the apply's LDS for the four stages of k_jacobi_ado on them, and
// writes only M^-1 A x.  The tier-0 workgroup applies the trace row and the
// Hermitian tier-0 block to its staged outputs first, and averages each
// element with the conjugate of its mirror after the inverse, as the two
// kernels do.  Every element goes through the same arithmetic as in the
// apply followed by k_jacobi_ado.
//
// EP > 0 carries a TFQMR iteration's vector work on the outputs (one
// problem): EP 1, out = u2 = M^-1 A y2, then w -= alpha u2 and per-workgroup
// partials ||w||^2, rt.w and rt.u2 at ep.partial[k gridDim.x + blockIdx.x],
// k = 0, 1, 2; EP 2 also sums those (ep.prev) in runs of 16 workgroups into
// ep.prev_out[k ceil(gridDim.x / 16) + blockIdx.x]; EP 2, out =
// aw = M^-1 A w and the partial rt.aw at ep.partial[blockIdx.x].  Nothing
// runs once the problem's flag is set.
template <bool RU, int EP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7)))
k_apply_packed_jac(const DevPacked op, const DevJacobi jc, const double* __restrict__ x, double* __restrict__ out,
                   int rdo, const TfqEpi ep) {
  if (EP && ep.s->flag) return;   // uniform over the grid
  static_assert(2 * 256 * sizeof(double2) + sizeof(UStore<RU>) + 16 * sizeof(double) <= kApplySmem,
                "the inverse's stages fit the apply's LDS");
  __shared__ __align__(16) char smem[kApplySmem];
  const Coefs cf{op.rec, op.gam, op.hc_ell};
  const int64_t hb = op.n_heavy;
  const bool heavy = (int64_t)blockIdx.x < hb;
  StageOut so;
  so.act = false;
  so.v = make_double2(0.0, 0.0);
  so.lbase = 0;
  so.t = 0;
  if (heavy)
    heavy_body<double, false, true>(op, cf, x, nullptr, -1, rdo, (int)blockIdx.x, smem, false, &so);
  else
    light_sector<double, false, true>(op, cf, x, nullptr, -1, (int)(blockIdx.x - hb), smem, false, &so);
  __syncthreads();   // the apply's LDS is free from here
  double2* xw = reinterpret_cast<double2*>(smem);
  double2* tw = xw + 256;
  UStore<RU>* us = reinterpret_cast<UStore<RU>*>(smem + 2 * 256 * sizeof(double2));
  double* ew = reinterpret_cast<double*>(smem + 2 * 256 * sizeof(double2) + sizeof(UStore<RU>));
  const int tid = threadIdx.x, d = op.d, DD = d * d;
  const bool tier0 = heavy && blockIdx.x == 0;
  xw[tid] = so.act ? so.v : make_double2(0.0, 0.0);
  if (tid < DD) us->put(tid, reinterpret_cast<const double2*>(jc.u)[tid]);
  if (tid < d) ew[tid] = jc.e[tid];
  __syncthreads();
  if (tier0) {   // block-uniform: the tier-0 fix-ups of the apply, on the staged block
    const int16_t* pos = op.sec_pos + op.sec[0] * DD;
    const int16_t* el = op.sec_elems + op.sec[0] * DD;
    if (rdo && tid == 0) {
      double sr = 0, si = 0;
      for (int i = 0; i < d; ++i) {
        const int q = pos[i + i * d];
        if (q >= 0) { sr += x[2 * q]; si += x[2 * q + 1]; }
      }
      xw[op.trace_slot] = make_double2(sr, si);
    }
    __syncthreads();
    if (op.herm0) {
      const int n0 = (int)(op.off[1] - op.off[0]);
      for (int t = tid; t < n0; t += blockDim.x) {
        const int e = el[t];
        const int i = e % d, j = e / d;
        if (i > j) continue;
        const int q = pos[j + i * d];
        if (q < 0) continue;
        const double2 a = xw[t], b = xw[q];
        const double re = 0.5 * (a.x + b.x), im = 0.5 * (a.y - b.y);
        xw[t] = make_double2(re, im);
        xw[q] = make_double2(re, -im);
      }
    }
    __syncthreads();
  }
  const int base = so.lbase;
  JacDesc ds{};
  if (so.act) ds = jc.desc[so.s * DD + so.t];
  // X -> X U_c
  if (so.act) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < ds.nq) {
        const double2 p = us->mul(ds.c2[q] + ds.c * d, xw[base + ds.cslot[q]]);
        acc.x += p.x;
        acc.y += p.y;
      }
    tw[tid] = acc;
  }
  __syncthreads();
  // -> (U_r^H .) / (shift - gamma_K + i (E_r - E_c))
  if (so.act) {
    double2 q = make_double2(0.0, 0.0);
#pragma unroll
    for (int p = 0; p < 4; ++p)
      if (p < ds.np) {
        const double2 v = us->mulc(ds.r2[p] + ds.r * d, tw[base + ds.rslot[p]]);
        q.x += v.x;
        q.y += v.y;
      }
    const double wr = jc.ejac - so.gr, wi = (ew[ds.r] - ew[ds.c]) - so.gi;
    const double den = wr * wr + wi * wi;
    xw[tid] = make_double2((q.x * wr + q.y * wi) / den, (q.y * wr - q.x * wi) / den);
  }
  __syncthreads();
  // -> . U_c^H
  if (so.act) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < ds.nq) {
        const double2 p = us->mulc(ds.c + ds.c2[q] * d, xw[base + ds.cslot[q]]);
        acc.x += p.x;
        acc.y += p.y;
      }
    tw[tid] = acc;
  }
  __syncthreads();
  // -> U_r .
  double2 z = make_double2(0.0, 0.0);
  if (so.act) {
#pragma unroll
    for (int p = 0; p < 4; ++p)
      if (p < ds.np) {
        const double2 v = us->mul(ds.r + ds.r2[p] * d, tw[base + ds.rslot[p]]);
        z.x += v.x;
        z.y += v.y;
      }
  }
  if (tier0) {   // tier 0: average with conj of the mirror element
    xw[tid] = z;
    __syncthreads();
    const int tm = so.act ? op.sec_pos[so.s * DD + ds.c + ds.r * d] : -1;
    if (tm >= 0) {
      const double2 m = xw[tm];
      z.x = 0.5 * (z.x + m.x);
      z.y = 0.5 * (z.y - m.y);
    }
  }
  if (EP == 0) {
    if (so.act) reinterpret_cast<double2*>(out)[so.gbase + so.t] = z;
    return;
  }
  double r0 = 0, r1 = 0, r2 = 0;
  if (so.act) {
    const int64_t g = so.gbase + so.t;
    reinterpret_cast<double2*>(out)[g] = z;
    const uint64_t seed = ep.s->seed;
    const double s0 = shadow(2 * g, seed), s1 = shadow(2 * g + 1, seed);
    if (EP == 1) {
      const double alpha = ep.s->alpha;
      double2 wi = reinterpret_cast<double2*>(ep.w)[g];
      wi.x -= alpha * z.x;
      wi.y -= alpha * z.y;
      reinterpret_cast<double2*>(ep.w)[g] = wi;
      r0 = wi.x * wi.x + wi.y * wi.y;
      r1 = s0 * wi.x + s1 * wi.y;
      r2 = s0 * z.x + s1 * z.y;
    } else {
      r0 = s0 * z.x + s1 * z.y;
    }
  }
  __shared__ double red[3][4];
  for (int off = 32; off > 0; off >>= 1) {
    r0 += __shfl_down(r0, off, 64);
    if (EP == 1) {
      r1 += __shfl_down(r1, off, 64);
      r2 += __shfl_down(r2, off, 64);
    }
  }
  const int lane = tid & 63, wid = tid >> 6;
  if (lane == 0) {
    red[0][wid] = r0;
    red[1][wid] = r1;
    red[2][wid] = r2;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0, b = 0, c = 0;
    for (int k = 0; k < 4; ++k) {
      a += red[0][k];
      b += red[1][k];
      c += red[2][k];
    }
    if (EP == 1) {
      ep.partial[blockIdx.x] = a;
      ep.partial[gridDim.x + blockIdx.x] = b;
      ep.partial[2 * gridDim.x + blockIdx.x] = c;
    } else {
      ep.partial[blockIdx.x] = a;
    }
  }
  if (EP == 2) {   // the first launch's sums, 16 workgroups' worth per workgroup here, for the fold
    const int na = (int)gridDim.x, ng = (na + 15) >> 4;
    if ((int)blockIdx.x < ng && tid < 64) {   // wave 0: lanes (k, j), k = tid / 16 < 3 the sum
      const int k = tid >> 4, j = tid & 15, idx = (int)blockIdx.x * 16 + j;
      double v = (k < 3 && idx < na) ? ep.prev[(int64_t)k * na + idx] : 0.0;
      for (int off = 8; off > 0; off >>= 1) v += __shfl_down(v, off, 16);
      if (k < 3 && j == 0) ep.prev_out[(int64_t)k * ng + blockIdx.x] = v;
    }
  }
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

// ADO 0's sector holds (j, i) whenever it holds (i, j) (N_i = N_j and
// S_z,i = S_z,j are symmetric), so every off-diagonal slot has its mirror.
// Upper slots are written only by their lower mirror's thread (mode 2) or
// read only from lower slots no thread writes (mode 0).
__global__ void k_bicg_tier0_p(const DevPacked op, double* __restrict__ x, int mode) {
  const int D = op.d, DD = D * D;
  const int n0 = (int)(op.off[1] - op.off[0]);
  const int16_t* pos = op.sec_pos + op.sec[0] * DD;
  const int16_t* el = op.sec_elems + op.sec[0] * DD;
  for (int t = threadIdx.x; t < n0; t += blockDim.x) {
    const int e = el[t];
    const int i = e % D, j = e / D;
    if (i == j) {
      x[2 * t + 1] = 0.0;
      continue;
    }
    const int q = pos[j + i * D];
    if (mode == 0 && i < j) {
      x[2 * t] = q >= 0 ? x[2 * q] : 0.0;
      x[2 * t + 1] = q >= 0 ? -x[2 * q + 1] : 0.0;
    } else if (mode == 1 && i < j) {
      x[2 * t] = 0.0;
      x[2 * t + 1] = 0.0;
    } else if (mode == 2 && i > j && q >= 0) {
      x[2 * t] += x[2 * q];
      x[2 * t + 1] -= x[2 * q + 1];
      x[2 * q] = 0.0;
      x[2 * q + 1] = 0.0;
    }
  }
}

__global__ void k_flag_clear(int32_t* flag) { *flag = 0; }

__global__ void k_offsector_any(const DevPacked op, const double2* __restrict__ x, int64_t nado, int32_t* flag) {
  const int dd = op.d * op.d;
  for (int64_t a = blockIdx.x; a < nado; a += gridDim.x) {
    const int16_t* pos = op.sec_pos + (int64_t)op.sec[a] * dd;
    for (int e = threadIdx.x; e < dd; e += blockDim.x) {
      if (pos[e] >= 0) continue;
      const double2 v = x[a * dd + e];
      if (v.x != 0.0 || v.y != 0.0) atomicOr(flag, 1);
    }
  }
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
  const int64_t blocks = apply_blocks(op);
  if (blocks > 0) {
    const dim3 grid((unsigned)blocks, (unsigned)op.nprob);
    const int u = op.max_light_couplings;
    const int r = rdo ? 1 : 0;
#define HEOMQ_APPLY(UU)                                                                                  \
  do {                                                                                                   \
    if (op.transposed)                                                                                   \
      hipLaunchKernelGGL((k_apply_packed<T, true, UU>), grid, dim3(256), 0, s, op, x, y, zero_at, r);    \
    else                                                                                                 \
      hipLaunchKernelGGL((k_apply_packed<T, false, UU>), grid, dim3(256), 0, s, op, x, y, zero_at, r);   \
  } while (0)
    if (u <= 1) HEOMQ_APPLY(1);
    else HEOMQ_APPLY(2);
#undef HEOMQ_APPLY
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  if (op.n_heavy > 0) return hipSuccess;   // block 0 (the tier-0 ADO) did the tier-0 fix-up
  if (op.nprob != 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_tier0_fixup<T>), dim3(1), dim3(256), 0, s, op, x, y, rdo ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_apply_packed_f64(const DevPacked& op, const double* x, double* y, bool rdo, hipStream_t s) {
  return launch_p<double>(op, x, y, rdo, s);
}

int64_t apply_blocks(const DevPacked& op) {
  const int64_t light = op.total - op.light_begin;
  return heavy_blocks(op) + (op.lt_blk ? op.n_lt_blocks : (light > 0 ? (light + 255) / 256 : 0));
}

bool apply_cheb_fusable(const DevPacked& op) {
  return !op.transposed && op.nprob == 1 && op.herm0 == 0 && op.precond == 0 && op.n_heavy > 0;
}

hipError_t launch_apply_packed_cheb(const DevPacked& op, const double* p1, double* p, double* x, double a, double b,
                                    hipStream_t s) {
  if (!apply_cheb_fusable(op)) return hipErrorInvalidValue;
  const int64_t blocks = apply_blocks(op);
  if (blocks <= 0) return hipSuccess;
  const ChebEpi ep{p, x, a, b};
  if (op.max_light_couplings <= 1)
    hipLaunchKernelGGL((k_apply_packed_cheb<1>), dim3((unsigned)blocks), dim3(256), 0, s, op, p1, ep);
  else
    hipLaunchKernelGGL((k_apply_packed_cheb<2>), dim3((unsigned)blocks), dim3(256), 0, s, op, p1, ep);
  return hipGetLastError();
}

hipError_t launch_apply_packed_multi_f64(const ApplyMulti& am, const double* x, double* y, bool rdo, hipStream_t s) {
  if (am.ng < 1 || am.ng > kApplyShapesPerLaunch || am.bstart[am.ng] <= 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_apply_packed_multi<double, 2>), dim3((unsigned)am.bstart[am.ng]), dim3(256), 0, s, am, x, y,
                     rdo ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_apply_packed_f32(const DevPacked& op, const float* x, float* y, bool rdo, hipStream_t s) {
  return launch_p<float>(op, x, y, rdo, s);
}

hipError_t launch_bicg_tier0_packed(const DevPacked& op, double* x, int mode, hipStream_t s) {
  if (op.n_ado <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bicg_tier0_p, dim3(1), dim3(256), 0, s, op, x, mode);
  return hipGetLastError();
}

hipError_t launch_offsector_any(const DevPacked& op, const double* dense, int32_t* flag, hipStream_t s) {
  hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(1), 0, s, flag);
  if (hipError_t e = hipGetLastError()) return e;
  if (op.n_ado <= 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<int64_t>(op.n_ado, 8192);
  hipLaunchKernelGGL(k_offsector_any, dim3(grid), dim3(256), 0, s, op, reinterpret_cast<const double2*>(dense),
                     op.n_ado, flag);
  return hipGetLastError();
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

hipError_t launch_jacobi_fwd(const DevPacked& op, const DevJacobi& jc, const double* r, double* t, hipStream_t s) {
  if (op.total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jacobi_fwd, dim3((unsigned)((op.total + 255) / 256)), dim3(256), 0, s, op, jc, r, t);
  return hipGetLastError();
}

hipError_t launch_jacobi_back(const DevPacked& op, const DevJacobi& jc, const double* t, double* x, bool accumulate,
                              hipStream_t s) {
  if (op.total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jacobi_back, dim3((unsigned)((op.total + 255) / 256)), dim3(256), 0, s, op, jc, t, x,
                     accumulate ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_jacobi_win(const DevPacked& op, const DevJacobi& jc, const double* in, double* out, bool accumulate,
                             hipStream_t s) {
  if (!jc.windowed || in == out) return hipErrorInvalidValue;
  if (op.total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_jacobi_win, dim3((unsigned)((op.total + 255) / 256)), dim3(256), 0, s, op, jc, in, out,
                     accumulate ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_jacobi_ado(const DevPacked& op, const DevJacobi& jc, const double* in, double* out, bool accumulate,
                             hipStream_t s) {
  if (jc.ado_blocks <= 0 || in == out || op.nprob != 1) return hipErrorInvalidValue;
  if (jc.u_real)
    hipLaunchKernelGGL((k_jacobi_ado<true, double>), dim3((unsigned)jc.ado_blocks), dim3(256), 0, s, op, jc, in, out,
                       accumulate ? 1 : 0);
  else
    hipLaunchKernelGGL((k_jacobi_ado<false, double>), dim3((unsigned)jc.ado_blocks), dim3(256), 0, s, op, jc, in,
                       out, accumulate ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_jacobi_ado_f32(const DevPacked& op, const DevJacobi& jc, const float* in, float* out,
                                 hipStream_t s) {
  if (jc.ado_blocks <= 0 || in == out || op.nprob != 1) return hipErrorInvalidValue;
  if (jc.u_real)
    hipLaunchKernelGGL((k_jacobi_ado<true, float>), dim3((unsigned)jc.ado_blocks), dim3(256), 0, s, op, jc, in, out,
                       0);
  else
    hipLaunchKernelGGL((k_jacobi_ado<false, float>), dim3((unsigned)jc.ado_blocks), dim3(256), 0, s, op, jc, in,
                       out, 0);
  return hipGetLastError();
}

bool apply_jac_fusable(const DevPacked& op, const DevJacobi& jc) {
  return op.lt_blk != nullptr && !op.heavy_wave && op.hg_off == nullptr && !op.transposed && op.precond == 0 &&
         op.nprob == 1 && op.n_heavy > 0 && op.d <= 16 && jc.ado_blocks > 0 && jc.desc != nullptr;
}

hipError_t launch_apply_packed_jac(const DevPacked& op, const DevJacobi& jc, const double* x, double* y, bool rdo,
                                   hipStream_t s) {
  if (!apply_jac_fusable(op, jc) || (rdo && op.trace_slot < 0)) return hipErrorInvalidValue;
  const int64_t blocks = apply_jac_blocks(op);
  const TfqEpi none{};
  if (jc.u_real)
    hipLaunchKernelGGL((k_apply_packed_jac<true, 0>), dim3((unsigned)blocks), dim3(256), 0, s, op, jc, x, y,
                       rdo ? 1 : 0, none);
  else
    hipLaunchKernelGGL((k_apply_packed_jac<false, 0>), dim3((unsigned)blocks), dim3(256), 0, s, op, jc, x, y,
                       rdo ? 1 : 0, none);
  return hipGetLastError();
}

int64_t apply_jac_blocks(const DevPacked& op) { return op.n_heavy + op.n_lt_blocks; }

hipError_t launch_apply_packed_jac_tfq(const DevPacked& op, const DevJacobi& jc, const double* x, double* y, bool rdo,
                                       int kind, const TfqEpi& ep, hipStream_t s) {
  if (!apply_jac_fusable(op, jc) || (rdo && op.trace_slot < 0) || (kind != 1 && kind != 2) || !ep.s ||
      !ep.partial || (kind == 1 && !ep.w) || (kind == 2 && (!ep.prev || !ep.prev_out)))
    return hipErrorInvalidValue;
  const dim3 grid((unsigned)apply_jac_blocks(op));
  const int r = rdo ? 1 : 0;
  if (jc.u_real) {
    if (kind == 1) hipLaunchKernelGGL((k_apply_packed_jac<true, 1>), grid, dim3(256), 0, s, op, jc, x, y, r, ep);
    else hipLaunchKernelGGL((k_apply_packed_jac<true, 2>), grid, dim3(256), 0, s, op, jc, x, y, r, ep);
  } else {
    if (kind == 1) hipLaunchKernelGGL((k_apply_packed_jac<false, 1>), grid, dim3(256), 0, s, op, jc, x, y, r, ep);
    else hipLaunchKernelGGL((k_apply_packed_jac<false, 2>), grid, dim3(256), 0, s, op, jc, x, y, r, ep);
  }
  return hipGetLastError();
}

hipError_t launch_jacobi_1k(const DevPacked& op, const DevJacobi& jc, const double* in, double* out, bool accumulate,
                            hipStream_t s) {
  if (jc.ado_blocks > 0 && op.nprob == 1) return launch_jacobi_ado(op, jc, in, out, accumulate, s);
  return launch_jacobi_win(op, jc, in, out, accumulate, s);
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
