// Block-Jacobi steady-state iteration (jacobi/solve_jacobi.f90,
// calc_ax_jacobi_omp.f90, prelude_jacobi.f90) on the symmetry-packed layout.
//
// One sweep is x <- (ejac - S_K)^-1 (ejac x + C_K(x)) per ADO K, where S_K X =
// -i[H, X] + gamma_K X is the same-tier part of the Liouvillian and C_K the
// couplings to the neighbouring tiers (calc_ax_jacobi_omp.f90:130-298).  With
// A_0 the full apply without a trace row, ejac x + C(x) = (ejac - S) x + A_0 x,
// so a sweep is x <- x + (ejac - S)^-1 A_0 x: one packed apply plus the
// same-tier inverse, which is diagonal in the eigenbasis of H (the reference's
// zmatjac / dvaljac, prelude_jacobi.f90:43-82, diagonalise [H, .] instead).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "heomq_ctx.hpp"

using namespace heomq;

namespace {

using cplx = std::complex<double>;

// Cyclic complex Jacobi rotations (Golub & Van Loan 8.5; the phase of a_pq is
// removed first so each step is the real symmetric 2x2 rotation of Numerical
// Recipes 11.1).  a: n x n Hermitian, column-major, destroyed.  On return
// v's columns are orthonormal eigenvectors and ev the eigenvalues.
void herm_eig(int n, std::vector<cplx>& a, std::vector<double>& ev, std::vector<cplx>& v) {
  auto A = [&](int i, int j) -> cplx& { return a[i + (size_t)j * n]; };
  auto V = [&](int i, int j) -> cplx& { return v[i + (size_t)j * n]; };
  v.assign((size_t)n * n, cplx(0, 0));
  for (int i = 0; i < n; ++i) V(i, i) = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) (i == j ? tot : off) += std::norm(A(i, j));
    if (off == 0.0 || off <= 1e-32 * (tot + off)) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const cplx apq = A(p, q);
        const double r = std::abs(apq);
        if (r == 0.0) continue;
        const cplx ph = apq / r;
        const double theta = (A(q, q).real() - A(p, p).real()) / (2.0 * r);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        // G = diag(1, conj(ph)) [[c, s], [-s, c]] on (p, q): A <- G^H A G, V <- V G
        const cplx gqp = -s * std::conj(ph), gqq = c * std::conj(ph);
        for (int k = 0; k < n; ++k) {
          const cplx akp = A(k, p), akq = A(k, q);
          A(k, p) = akp * c + akq * gqp;
          A(k, q) = akp * s + akq * gqq;
          const cplx vkp = V(k, p), vkq = V(k, q);
          V(k, p) = vkp * c + vkq * gqp;
          V(k, q) = vkp * s + vkq * gqq;
        }
        for (int k = 0; k < n; ++k) {
          const cplx apk = A(p, k), aqk = A(q, k);
          A(p, k) = c * apk + std::conj(gqp) * aqk;
          A(q, k) = s * apk + std::conj(gqq) * aqk;
        }
      }
  }
  ev.resize(n);
  for (int i = 0; i < n; ++i) ev[i] = A(i, i).real();
}

struct JacobiHost {
  std::vector<double> u, e;   // 2*d*d (column-major U) and d
  std::vector<int16_t> grp;   // [d][gmax]
  std::vector<int8_t> gn;
  std::vector<std::vector<int>> groups;
  std::vector<int> gid;       // group of each state
  int gmax = 0, nnz = 0;
  double gap = 0.0, cbterm = 0.0, ejac = 0.0;
};

// Groups of Fock states that H connects only among themselves: equal electron
// number (its parity when H changes it) and equal S_z (unless H flips spin);
// the same quantum numbers the packed layout conserves.
int build_jacobi_host(const Model& m, JacobiHost& jh) {
  const int d = m.d;
  std::vector<int> fN(d), fS(d);
  for (int i = 0; i < d; ++i) {
    int n = 0, s2 = 0, v = i;
    for (int o = 0; o < m.norbs; ++o, v /= 4) {
      const int dig = v % 4;
      const int up = dig == 1 || dig == 3, dn = dig == 2 || dig == 3;
      n += up + dn;
      s2 += up - dn;
    }
    fN[i] = n;
    fS[i] = s2;
  }
  bool cons_n = true, cons_s = true;
  for (int i = 0; i < d; ++i)
    for (int k = 0; k < d; ++k)
      if (m.H[i + (size_t)k * d] != cplx(0, 0)) {
        if (fN[i] != fN[k]) cons_n = false;
        if (fS[i] != fS[k]) cons_s = false;
      }
  std::vector<int> key(d);
  for (int i = 0; i < d; ++i) key[i] = (cons_n ? fN[i] : (fN[i] & 1)) * 1000 + (cons_s ? fS[i] + 500 : 0);
  for (int i = 0; i < d; ++i)
    for (int k = 0; k < d; ++k)
      if (m.H[i + (size_t)k * d] != cplx(0, 0) && key[i] != key[k])
        return fail(HEOMQ_UNSUPPORTED, "H mixes electron-number parities: no block-Jacobi groups");
  std::vector<std::vector<int>> groups;
  std::vector<int> gid(d, -1);
  for (int i = 0; i < d; ++i) {
    if (gid[i] >= 0) continue;
    std::vector<int> g;
    for (int k = i; k < d; ++k)
      if (key[k] == key[i]) { gid[k] = (int)groups.size(); g.push_back(k); }
    groups.push_back(g);
  }
  jh.groups = groups;
  jh.gid = gid;
  jh.gmax = 0;
  for (auto& g : groups) jh.gmax = std::max<int>(jh.gmax, (int)g.size());
  if (jh.gmax > 64) return fail(HEOMQ_UNSUPPORTED, "block-Jacobi: an H group is larger than 64 states");
  jh.u.assign(2 * (size_t)d * d, 0.0);
  jh.e.assign(d, 0.0);
  jh.grp.assign((size_t)d * jh.gmax, (int16_t)-1);
  jh.gn.assign(d, 0);
  for (auto& g : groups) {
    const int n = (int)g.size();
    std::vector<cplx> a((size_t)n * n), v;
    std::vector<double> ev;
    for (int q = 0; q < n; ++q)
      for (int p = 0; p < n; ++p) a[p + (size_t)q * n] = m.H[g[p] + (size_t)g[q] * d];
    herm_eig(n, a, ev, v);
    // eigenvector q of the group becomes column g[q] of U
    for (int q = 0; q < n; ++q) {
      jh.e[g[q]] = ev[q];
      for (int p = 0; p < n; ++p) {
        jh.u[2 * (g[p] + (size_t)g[q] * d)] = v[p + (size_t)q * n].real();
        jh.u[2 * (g[p] + (size_t)g[q] * d) + 1] = v[p + (size_t)q * n].imag();
      }
    }
    for (int q = 0; q < n; ++q) {
      jh.gn[g[q]] = (int8_t)n;
      for (int p = 0; p < n; ++p) jh.grp[(size_t)g[q] * jh.gmax + p] = (int16_t)g[p];
    }
  }
  // prelude_jacobi.f90:84-113: ejac = max(spread of the nonzero eigenvalues of
  // [H, .], sqrt((ntier - 1) * max |cb|)); the eigenvalues of [H, .] are
  // E_i - E_j, and those within dpico = 1e-12 of zero are screened out (:62-70)
  double lo = 0.0, hi = 0.0;
  bool any = false;
  jh.nnz = 0;
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      const double lam = jh.e[i] - jh.e[j];
      if (std::fabs(lam) <= 1e-12) continue;
      ++jh.nnz;
      lo = any ? std::min(lo, lam) : lam;
      hi = any ? std::max(hi, lam) : lam;
      any = true;
    }
  jh.gap = any ? hi - lo : 0.0;
  double cbmax = 0.0;
  for (const cplx& v : m.cb) cbmax = std::max(cbmax, std::abs(v));
  jh.cbterm = std::sqrt((double)m.L * cbmax);
  jh.ejac = std::max(jh.gap, jh.cbterm);
  return HEOMQ_OK;
}

int check_shift(const heomq_ctx* c, double ejac) {
  if (!(ejac > 0.0)) return fail(HEOMQ_BAD_ARG, "block-Jacobi: ejac must be positive");
  for (size_t a = 0; 2 * a + 1 < c->gam_h.size(); ++a)
    if (std::abs(cplx(ejac - c->gam_h[2 * a], -c->gam_h[2 * a + 1])) <= 1e-12)
      return fail(HEOMQ_UNSUPPORTED, "block-Jacobi: ejac equals an ADO's gamma_K");
  return HEOMQ_OK;
}

int carve_jac(heomq_ctx* c, int64_t n, int count, std::vector<double*>& out) {
  const size_t stride = ((sizeof(double) * (size_t)n + 255) / 256) * 256;
  HIPCHECK(c->ws_jac.alloc(stride * count));
  out.clear();
  for (int k = 0; k < count; ++k) out.push_back(reinterpret_cast<double*>(static_cast<char*>(c->ws_jac.p) + k * stride));
  return HEOMQ_OK;
}

// x (packed) <- one block-Jacobi sweep; r, t: packed scratch
int sweep(heomq_ctx* c, const DevJacobi& jc, double* x, double* r, double* t, hipStream_t st) {
  DevPacked op = c->pfwd;
  op.precond = 0;
  HIPCHECK(launch_apply_packed_f64(op, x, r, false, st));
  if (jacobi_one_kernel(jc)) {
    HIPCHECK(launch_jacobi_1k(op, jc, r, x, true, st));
    return HEOMQ_OK;
  }
  HIPCHECK(launch_jacobi_fwd(op, jc, r, t, st));
  HIPCHECK(launch_jacobi_back(op, jc, t, x, true, st));
  return HEOMQ_OK;
}

// ||A x - e||_2 with the trace row of calc_ax_tfqmr_omp (the last ADO's (d, d)
// element holds tr rho0, solve_jacobi.f90:188-190), from A_0 x on the packed
// layout: ||A_0 x||^2 - |(A_0 x)_last|^2 + |tr rho0 - 1|^2.  Also returns tr rho0.
int residual(heomq_ctx* c, const double* x, double* r, const double* zero, hipStream_t st, double& res,
             cplx& tr) {
  DevPacked op = c->pfwd;
  op.precond = 0;
  HIPCHECK(launch_apply_packed_f64(op, x, r, false, st));
  double* out = c->red.as<double>();
  HIPCHECK((vec_resid<double, double>(st, 2 * c->ptotal, r, zero, out, c->partial.as<double>())));
  HIPCHECK(hipStreamSynchronize(st));
  double rr[2];
  HIPCHECK(hipMemcpy(rr, out, sizeof(rr), hipMemcpyDeviceToHost));
  const int d = c->d, dd = d * d;
  const int64_t last = c->N - 1;
  const int slot = c->sec_pos_h[(size_t)c->psec_h[last] * dd + (d - 1) + (d - 1) * d];
  double ylast[2] = {0.0, 0.0};
  if (slot >= 0) HIPCHECK(hipMemcpy(ylast, r + 2 * (c->poff_h[last] + slot), sizeof(ylast), hipMemcpyDeviceToHost));
  const int nnz0 = (int)(c->poff_h[1] - c->poff_h[0]);
  std::vector<double> x0(2 * (size_t)nnz0);
  HIPCHECK(hipMemcpy(x0.data(), x, sizeof(double) * x0.size(), hipMemcpyDeviceToHost));
  tr = 0.0;
  for (int i = 0; i < d; ++i) {
    const int sl = c->sec_pos_h[(size_t)c->psec_h[0] * dd + i + i * d];
    if (sl >= 0) tr += cplx(x0[2 * sl], x0[2 * sl + 1]);
  }
  const double r2 = rr[0] - (ylast[0] * ylast[0] + ylast[1] * ylast[1]) + std::norm(tr - 1.0);
  res = std::sqrt(std::max(r2, 0.0));
  return HEOMQ_OK;
}

}  // namespace

namespace heomq {

// nsweep block-Jacobi sweeps of x (packed) with the reference's default
// shift: the convergent iteration of solve_jacobi.f90, used to smooth a
// start before a Krylov solve.  r, t: packed scratch.
int jacobi_presmooth(heomq_ctx* c, double* x, int nsweep, double* r, double* t, hipStream_t st) {
  DevJacobi jc = c->pjac;
  jc.ejac = c->ejac_default;
  for (int k = 0; k < nsweep; ++k)
    if (int rc = sweep(c, jc, x, r, t, st)) return rc;
  return HEOMQ_OK;
}

int build_tier_jacobi(const Model& m, int S, const std::vector<int16_t>& elems, const std::vector<int16_t>& pos,
                      const std::vector<int>& nnz, const std::vector<int64_t>& off, int64_t a0, DevBuf& blk_buf,
                      DevBuf& desc_buf, DevBuf& u_buf, DevBuf& e_buf, DevJacobi& out) {
  JacobiHost jh;
  if (int rc = build_jacobi_host(m, jh)) return rc;
  const int d = m.d, dd = d * d;
  if (d > 16 || jh.gmax > 4)
    return fail(HEOMQ_UNSUPPORTED, "ADO-aligned same-tier inverse: needs d <= 16 and H groups of at most 4 states");
  std::vector<int64_t> blk;
  for (int64_t a = a0; a < m.N; ++a) {
    const int64_t sz = off[a + 1] - off[a];
    if (sz > 256) return fail(HEOMQ_UNSUPPORTED, "ADO-aligned same-tier inverse: an ADO holds more than 256 elements");
    if (blk.empty() || off[a + 1] - blk[blk.size() - 1] > 256 || a - blk[blk.size() - 2] >= 256) {
      blk.push_back(a);
      blk.push_back(off[a]);
    }
  }
  std::vector<JacDesc> desc((size_t)S * dd);
  for (int s = 0; s < S; ++s)
    for (int t = 0; t < nnz[s]; ++t) {
      const int e = elems[(size_t)s * dd + t], r = e % d, col = e / d;
      JacDesc& ds = desc[(size_t)s * dd + t];
      ds.r = (uint8_t)r;
      ds.c = (uint8_t)col;
      ds.nq = (uint8_t)jh.gn[col];
      ds.np = (uint8_t)jh.gn[r];
      for (int q = 0; q < jh.gn[col]; ++q) {
        const int c2 = jh.grp[(size_t)col * jh.gmax + q];
        ds.c2[q] = (uint8_t)c2;
        ds.cslot[q] = pos[(size_t)s * dd + r + c2 * d];
        if (ds.cslot[q] < 0) return fail(HEOMQ_UNSUPPORTED, "block-Jacobi groups do not match the packed layout");
      }
      for (int p = 0; p < jh.gn[r]; ++p) {
        const int r2 = jh.grp[(size_t)r * jh.gmax + p];
        ds.r2[p] = (uint8_t)r2;
        ds.rslot[p] = pos[(size_t)s * dd + r2 + col * d];
        if (ds.rslot[p] < 0) return fail(HEOMQ_UNSUPPORTED, "block-Jacobi groups do not match the packed layout");
      }
    }
  out = DevJacobi{};
  out.ado_blocks = (int64_t)blk.size() / 2;
  blk.push_back(m.N);
  blk.push_back(off[m.N]);
  HIPCHECK(upload(blk_buf, blk));
  HIPCHECK(upload(desc_buf, desc));
  HIPCHECK(upload(u_buf, jh.u));
  HIPCHECK(upload(e_buf, jh.e));
  out.u = u_buf.as<double>();
  out.e = e_buf.as<double>();
  out.gmax = jh.gmax;
  out.blk = blk_buf.as<int64_t>();
  out.desc = desc_buf.as<JacDesc>();
  bool real = true;
  for (size_t e = 1; e < jh.u.size(); e += 2) real = real && jh.u[e] == 0.0;
  out.u_real = real ? 1 : 0;
  out.ejac = 0.0;   // (0 - S_K)^-1: the eliminated ADO's steady equation
  return HEOMQ_OK;
}

int build_jacobi(heomq_ctx* c, const Model& m) {
  JacobiHost jh;
  if (int rc = build_jacobi_host(m, jh)) return rc;
  // every (k, l) the transform of a stored (i, j) reads must be stored too
  const int d = m.d, dd = d * d;
  for (int s = 0; s < c->nsec; ++s)
    for (int t = 0; t < c->sec_nnz_h[s]; ++t) {
      const int e = c->sec_elems_h[(size_t)s * dd + t], i = e % d, j = e / d;
      for (int p = 0; p < jh.gn[i]; ++p)
        for (int q = 0; q < jh.gn[j]; ++q) {
          const int k = jh.grp[(size_t)i * jh.gmax + p], l = jh.grp[(size_t)j * jh.gmax + q];
          if (c->sec_pos_h[(size_t)s * dd + k + l * d] < 0)
            return fail(HEOMQ_UNSUPPORTED, "block-Jacobi groups do not match the packed layout");
        }
    }
  HIPCHECK(upload(c->j_u, jh.u));
  HIPCHECK(upload(c->j_e, jh.e));
  HIPCHECK(upload(c->j_grp, jh.grp));
  HIPCHECK(upload(c->j_gn, jh.gn));
  c->pjac.u = c->j_u.as<double>();
  c->pjac.e = c->j_e.as<double>();
  c->pjac.grp = c->j_grp.as<int16_t>();
  c->pjac.gn = c->j_gn.as<int8_t>();
  c->pjac.gmax = jh.gmax;
  // slot-space list tables of the one-kernel form (per sector and slot)
  {
    const int S = c->nsec;
    std::vector<int32_t> crow((size_t)S * (dd + 1), 0), rrow((size_t)S * (dd + 1), 0);
    std::vector<int16_t> cslot, rslot;
    std::vector<double> ca, cc, rb, rd, tde((size_t)S * dd, 0.0);
    auto U = [&](int a, int b) { return cplx(jh.u[2 * (a + (size_t)b * d)], jh.u[2 * (a + (size_t)b * d) + 1]); };
    auto push = [](std::vector<double>& v, cplx z) { v.push_back(z.real()); v.push_back(z.imag()); };
    for (int s = 0; s < S; ++s) {
      for (int t = 0; t < dd; ++t) {
        crow[(size_t)s * (dd + 1) + t] = (int32_t)cslot.size();
        rrow[(size_t)s * (dd + 1) + t] = (int32_t)rslot.size();
        if (t >= c->sec_nnz_h[s]) continue;
        const int e = c->sec_elems_h[(size_t)s * dd + t], r = e % d, col = e / d;
        tde[(size_t)s * dd + t] = jh.e[r] - jh.e[col];
        for (int q = 0; q < jh.gn[col]; ++q) {   // c' in group(c)
          const int c2 = jh.grp[(size_t)col * jh.gmax + q];
          cslot.push_back(c->sec_pos_h[(size_t)s * dd + r + c2 * d]);
          push(ca, U(c2, col));
          push(cc, std::conj(U(col, c2)));
        }
        for (int p = 0; p < jh.gn[r]; ++p) {     // r' in group(r)
          const int r2 = jh.grp[(size_t)r * jh.gmax + p];
          rslot.push_back(c->sec_pos_h[(size_t)s * dd + r2 + col * d]);
          push(rb, std::conj(U(r2, r)));
          push(rd, U(r, r2));
        }
      }
      crow[(size_t)s * (dd + 1) + dd] = (int32_t)cslot.size();
      rrow[(size_t)s * (dd + 1) + dd] = (int32_t)rslot.size();
    }
    if (cslot.empty()) cslot.push_back(0), ca.assign(2, 0.0), cc.assign(2, 0.0);
    if (rslot.empty()) rslot.push_back(0), rb.assign(2, 0.0), rd.assign(2, 0.0);
    HIPCHECK(upload(c->j_crow, crow));
    HIPCHECK(upload(c->j_cslot, cslot));
    HIPCHECK(upload(c->j_ca, ca));
    HIPCHECK(upload(c->j_cc, cc));
    HIPCHECK(upload(c->j_rrow, rrow));
    HIPCHECK(upload(c->j_rslot, rslot));
    HIPCHECK(upload(c->j_rb, rb));
    HIPCHECK(upload(c->j_rd, rd));
    HIPCHECK(upload(c->j_tde, tde));
    c->pjac.crow = c->j_crow.as<int32_t>();
    c->pjac.cslot = c->j_cslot.as<int16_t>();
    c->pjac.ca = c->j_ca.as<double>();
    c->pjac.cc = c->j_cc.as<double>();
    c->pjac.rrow = c->j_rrow.as<int32_t>();
    c->pjac.rslot = c->j_rslot.as<int16_t>();
    c->pjac.rb = c->j_rb.as<double>();
    c->pjac.rd = c->j_rd.as<double>();
    c->pjac.tde = c->j_tde.as<double>();
  }
  // the one-kernel form: every 256-element block of the packed vector must lie
  // in ADOs that fit its LDS window
  {
    const int64_t total = c->poff_h[m.N];
    const int64_t nblk = (total + 255) / 256;
    bool fits = total > 0;
    int64_t a = 0;
    std::vector<int64_t> blk_ado((size_t)nblk + 1, 0);
    for (int64_t bb = 0; bb <= nblk && fits; ++bb) {
      const int64_t t = std::min<int64_t>(bb * 256, total - 1);
      while (a + 1 < m.N && c->poff_h[a + 1] <= t) ++a;
      blk_ado[bb] = a;
    }
    for (int64_t bb = 0; bb < nblk && fits; ++bb) {
      const int64_t lo = blk_ado[bb], hi = blk_ado[bb + 1];
      if (hi - lo + 1 > 257 || c->poff_h[hi + 1] - c->poff_h[lo] > kJacobiWindow) fits = false;
    }
    c->pjac.windowed = fits ? 1 : 0;
  }
  // the ADO-aligned form: runs of whole ADOs of at most 256 elements and 256
  // ADOs per block, element descriptors with groups of at most 4 states
  {
    const int S = c->nsec;
    bool ok = d <= 16 && jh.gmax <= 4;
    std::vector<int64_t> blk;
    for (int64_t a = 0; a < m.N && ok; ++a) {
      const int64_t sz = c->poff_h[a + 1] - c->poff_h[a];
      if (sz > 256) ok = false;
      else if (blk.empty() || c->poff_h[a + 1] - blk[blk.size() - 1] > 256 || a - blk[blk.size() - 2] >= 256) {
        blk.push_back(a);
        blk.push_back(c->poff_h[a]);
      }
    }
    std::vector<JacDesc> desc(ok ? (size_t)S * dd : 0);
    for (int s = 0; s < S && ok; ++s)
      for (int t = 0; t < c->sec_nnz_h[s]; ++t) {
        const int e = c->sec_elems_h[(size_t)s * dd + t], r = e % d, col = e / d;
        JacDesc& ds = desc[(size_t)s * dd + t];
        ds.r = (uint8_t)r;
        ds.c = (uint8_t)col;
        ds.nq = (uint8_t)jh.gn[col];
        ds.np = (uint8_t)jh.gn[r];
        for (int q = 0; q < jh.gn[col]; ++q) {
          const int c2 = jh.grp[(size_t)col * jh.gmax + q];
          ds.c2[q] = (uint8_t)c2;
          ds.cslot[q] = c->sec_pos_h[(size_t)s * dd + r + c2 * d];
        }
        for (int p = 0; p < jh.gn[r]; ++p) {
          const int r2 = jh.grp[(size_t)r * jh.gmax + p];
          ds.r2[p] = (uint8_t)r2;
          ds.rslot[p] = c->sec_pos_h[(size_t)s * dd + r2 + col * d];
        }
      }
    if (ok && !blk.empty() && c->poff_h[m.N] > 0) {
      const int64_t nb = (int64_t)blk.size() / 2;
      blk.push_back(m.N);
      blk.push_back(c->poff_h[m.N]);
      HIPCHECK(upload(c->j_blk, blk));
      HIPCHECK(upload(c->j_desc, desc));
      c->pjac.blk = c->j_blk.as<int64_t>();
      c->pjac.desc = c->j_desc.as<JacDesc>();
      c->pjac.ado_blocks = nb;
      bool real = true;
      for (size_t e = 1; e < jh.u.size(); e += 2) real = real && jh.u[e] == 0.0;
      c->pjac.u_real = real ? 1 : 0;
    } else {
      c->pjac.ado_blocks = 0;
    }
  }
  c->pjac.ejac = jh.ejac;
  c->ejac_default = jh.ejac;
  c->ejac_gap = jh.gap;
  c->ejac_cb = jh.cbterm;
  c->jac_nnz = jh.nnz;
  build_gamma(m, m.gdraw, c->gam_h, false);
  // The preconditioner's denominators are shift - gamma_K + i dE: re-check
  // the shift against the rates just rebuilt.  A user shift that the new
  // parameters make singular falls back to the default; a singular default
  // is refused, as the reference refuses its 1/(i lambda) branch.
  auto singular = [&](double v) {
    for (size_t a = 0; 2 * a + 1 < c->gam_h.size(); ++a)
      if (std::abs(cplx(v - c->gam_h[2 * a], -c->gam_h[2 * a + 1])) <= 1e-12) return true;
    return false;
  };
  if (c->prec_shift_user > 0.0 && singular(c->prec_shift_user)) c->prec_shift_user = 0.0;
  c->prec_shift = c->prec_shift_user > 0.0 ? c->prec_shift_user : kPrecShiftFactor * jh.ejac;
  if (singular(c->prec_shift))
    return fail(HEOMQ_UNSUPPORTED, "block-Jacobi preconditioner: the default shift equals an ADO's gamma_K");
  return HEOMQ_OK;
}

}  // namespace heomq

extern "C" {

int heomq_model_jacobi_info(const heomq_model* h, double* info, double* evals, double* evecs) {
  if (!h || !info) return fail(HEOMQ_BAD_ARG, "null argument");
  JacobiHost jh;
  if (int rc = build_jacobi_host(h->m, jh)) return rc;
  info[0] = jh.ejac;
  info[1] = jh.gap;
  info[2] = jh.cbterm;
  info[3] = jh.nnz;
  info[4] = jh.gmax;
  if (evals) std::copy(jh.e.begin(), jh.e.end(), evals);
  if (evecs) std::copy(jh.u.begin(), jh.u.end(), evecs);
  return HEOMQ_OK;
}

int heomq_jacobi_info(const heomq_ctx* c, double* info) {
  if (!c || !info) return fail(HEOMQ_BAD_ARG, "null argument");
  info[0] = c->ejac_default;
  info[1] = c->ejac_gap;
  info[2] = c->ejac_cb;
  info[3] = c->jac_nnz;
  info[4] = c->pjac.gmax;
  return HEOMQ_OK;
}

int heomq_jacobi_step(heomq_ctx* c, const double* x, double* y, double ejac, void* stream) {
  if (!c || !x || !y) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_jacobi_step: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (ejac <= 0.0) ejac = c->ejac_default;
  if (int rc = check_shift(c, ejac)) return rc;
  hipStream_t st = (hipStream_t)stream;
  std::vector<double*> v;
  if (int rc = carve_jac(c, 2 * c->ptotal, 3, v)) return rc;
  DevJacobi jc = c->pjac;
  jc.ejac = ejac;
  HIPCHECK((launch_pack<double, double>(c->pfwd, x, v[0], st)));
  if (int rc = sweep(c, jc, v[0], v[1], v[2], st)) return rc;
  HIPCHECK((launch_unpack<double, double>(c->pfwd, v[0], y, st)));
  return HEOMQ_OK;
}

int heomq_calc_ax_jacobi(heomq_ctx* c, int32_t dim0, const double* x_host, double* y_host) {
  if (!c || !x_host || !y_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_calc_ax_jacobi: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (dim0 != c->d) return fail(HEOMQ_BAD_DIM, "error! dim0 != nrho in calc_ax_jacobi_omp");
  const int64_t n = 2 * c->N * c->d * c->d;
  HIPCHECK(c->ws_host.alloc(2 * sizeof(double) * (size_t)n));
  double* xd = c->ws_host.as<double>();
  double* yd = xd + n;
  HIPCHECK(hipMemcpy(xd, x_host, sizeof(double) * n, hipMemcpyHostToDevice));
  if (int rc = heomq_jacobi_step(c, xd, yd, 0.0, nullptr)) return rc;
  HIPCHECK(hipMemcpy(y_host, yd, sizeof(double) * n, hipMemcpyDeviceToHost));
  return HEOMQ_OK;
}

// solve_jacobi.f90 on the device: rho0 made Hermitian from its lower triangle
// and required to have unit trace (:70-85), then sweeps until the trace-row
// residual ||A x - e||_2, checked after sweeps 1, 51, 101, ... (:183-229),
// is <= crit, or maxiter sweeps.  x (dense, in/out) holds the last iterate.
int heomq_solve_jacobi(heomq_ctx* c, double* x, double crit, int maxiter, double ejac, double* stats,
                       void* stream) {
  if (!c || !x) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_solve_jacobi: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (ejac <= 0.0) ejac = c->ejac_default;
  if (int rc = check_shift(c, ejac)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t n = 2 * c->ptotal;
  std::vector<double*> v;
  if (int rc = carve_jac(c, n, 4, v)) return rc;
  double *xs = v[0], *r = v[1], *t = v[2], *zero = v[3];
  DevJacobi jc = c->pjac;
  jc.ejac = ejac;
  HIPCHECK(launch_herm_from_lower_f64(c->d, x, st));
  HIPCHECK((launch_pack<double, double>(c->pfwd, x, xs, st)));
  HIPCHECK(vec_fill<double>(st, n, zero, 0.0));
  double res = 0.0, best = 0.0;
  cplx tr;
  if (int rc = residual(c, xs, r, zero, st, res, tr)) return rc;
  if (std::abs(tr - 1.0) > 1e-12) {
    char msg[96];
    std::snprintf(msg, sizeof msg, "error! rho is not normalized (trace %.15g%+.3gi)", tr.real(), tr.imag());
    return fail(HEOMQ_BAD_ARG, msg);
  }
  best = res;
  int iter = 0, applies = 1, status = HEOMQ_NOT_CONVERGED;
  if (res <= crit) {
    status = HEOMQ_OK;
  } else {
    for (;;) {
      if (int rc = sweep(c, jc, xs, r, t, st)) return rc;
      ++applies;
      if (iter % 50 == 0) {
        if (int rc = residual(c, xs, r, zero, st, res, tr)) return rc;
        ++applies;
        best = std::min(best, res);
        if (res <= crit) { status = HEOMQ_OK; break; }
      }
      if (iter >= maxiter) break;
      ++iter;
    }
  }
  HIPCHECK((launch_unpack<double, double>(c->pfwd, xs, x, st)));
  HIPCHECK(hipStreamSynchronize(st));
  if (stats) {
    std::fill(stats, stats + HEOMQ_STATS_LEN, 0.0);
    stats[0] = iter;
    stats[1] = applies;
    stats[2] = res;
    stats[3] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats[4] = ejac;
    stats[5] = best;
  }
  if (status != HEOMQ_OK) return fail(HEOMQ_NOT_CONVERGED, "Jacobi iteration did not reach the criterion");
  return HEOMQ_OK;
}

int heomq_solve_jacobi_host(heomq_ctx* c, double* rho_host, double crit, int maxiter, double* stats) {
  if (!c || !rho_host) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_solve_jacobi_host: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  const int64_t n = 2 * c->N * c->d * c->d;
  HIPCHECK(c->ws_host.alloc(sizeof(double) * (size_t)n));
  double* x = c->ws_host.as<double>();
  HIPCHECK(hipMemcpy(x, rho_host, sizeof(double) * n, hipMemcpyHostToDevice));
  const int rc = heomq_solve_jacobi(c, x, crit, maxiter, 0.0, stats, nullptr);
  if (rc != HEOMQ_OK && rc != HEOMQ_NOT_CONVERGED) return rc;
  HIPCHECK(hipMemcpy(rho_host, x, sizeof(double) * n, hipMemcpyDeviceToHost));
  return rc;
}

}  // extern "C"
