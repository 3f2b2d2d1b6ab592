// Host-side HEOM model builder: the preparation phase that the reference runs
// once before its solvers (bath decomposition, fermion operators, H, the ADO
// hierarchy and its tier-index coupling table), restated in C++ so the device
// context can be built without the Fortran program.
#include "heomq_model.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace heomq {

std::string& last_error() {
  static thread_local std::string s;
  return s;
}

int fail(int code, const std::string& msg) {
  last_error() = msg;
  return code;
}

int Model::tier_of(int64_t a) const {
  int n = 0;
  while (n + 1 < (int)first.size() && first[n + 1] <= a) ++n;
  return n;
}

// ---------------------------------------------------------------------------
// Pade spectrum decomposition (main/zpout_psd.f90:34-114).  The two symmetric
// tridiagonal matrices have zero diagonal; their eigenvalues are found by
// Sturm-sequence bisection (each to full double precision).
// ---------------------------------------------------------------------------
static int sturm_below(const std::vector<double>& off, double x) {
  const int n = (int)off.size() + 1;
  int cnt = 0;
  double q = -x;
  if (q < 0) ++cnt;
  for (int i = 1; i < n; ++i) {
    const double qq = (q == 0.0) ? 1e-300 : q;
    q = -x - off[i - 1] * off[i - 1] / qq;
    if (q < 0) ++cnt;
  }
  return cnt;
}

static double kth_eigenvalue(const std::vector<double>& off, int k) {
  double bound = 0.0;
  for (double v : off) bound = std::max(bound, 2.0 * std::fabs(v));
  double lo = -bound - 1.0, hi = bound + 1.0;
  for (int it = 0; it < 400 && hi - lo > 0.0; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (sturm_below(off, mid) > k) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

void pade_psd(int n, std::vector<double>& eta, std::vector<double>& xi) {
  const int dim0 = 2 * n, dim1 = 2 * n - 1;
  std::vector<double> off0(dim0 - 1), off1(std::max(dim1 - 1, 0));
  for (int i = 1; i < dim0; ++i) off0[i - 1] = 1.0 / std::sqrt((2.0 * i - 1.0) * (2.0 * (i + 1) - 1.0));
  for (int i = 1; i < dim1; ++i) off1[i - 1] = 1.0 / std::sqrt((2.0 * i + 1.0) * (2.0 * (i + 1) + 1.0));
  std::vector<double> x0(n), x1(std::max(n - 1, 0));
  for (int k = 0; k < n; ++k) {
    const double e = kth_eigenvalue(off0, k);   // the n negative eigenvalues, ascending
    x0[k] = 4.0 / (e * e);
  }
  for (int k = 0; k < n - 1; ++k) {
    const double e = kth_eigenvalue(off1, k);
    x1[k] = 4.0 / (e * e);
  }
  eta.assign(n, 1.0);
  xi.assign(n, 0.0);
  for (int j = 0; j < n; ++j) {
    for (int k = 0; k < j; ++k) eta[j] *= (x1[k] - x0[j]) / (x0[k] - x0[j]);
    for (int k = j + 1; k < n; ++k) eta[j] *= (x1[k - 1] - x0[j]) / (x0[k] - x0[j]);
    eta[j] *= 0.5 * n * (2.0 * n + 1.0);
    xi[j] = std::sqrt(x0[j]);
  }
}

cplx fermi_pade(cplx z, const std::vector<double>& eta, const std::vector<double>& xi) {
  cplx out(0.5, 0.0);                       // main/fermifunc.f90:18-23
  for (size_t k = 0; k < eta.size(); ++k)
    out -= eta[k] / (z + cplx(0, xi[k])) + eta[k] / (z - cplx(0, xi[k]));
  return out;
}

// ---------------------------------------------------------------------------
// Fermion operators in the occupation basis (main/calcams.f90:79-103 and the
// swaps of main/buildoperator.f90:168-226, main/swapspin.f90, swapms.f90).
// Orbital 1 is the fastest base-4 digit; digit 0 empty, 1 up, 2 down, 3 both.
// ---------------------------------------------------------------------------
namespace {
struct Occ { int up, dn; };

std::vector<Occ> decode(int idx0, int d, int norbs) {
  std::vector<Occ> o(norbs);
  for (int m = 0; m < norbs; ++m) {
    int len0 = 1;
    for (int t = 0; t < m; ++t) len0 *= 4;
    const int dig = (idx0 / len0) % 4;
    o[m] = Occ{dig == 1 || dig == 3, dig == 2 || dig == 3};
  }
  (void)d;
  return o;
}

int encode(const std::vector<Occ>& o) {
  int idx = 0, len0 = 1;
  for (size_t m = 0; m < o.size(); ++m) {
    const int dig = o[m].up && o[m].dn ? 3 : o[m].up ? 1 : o[m].dn ? 2 : 0;
    idx += dig * len0;
    len0 *= 4;
  }
  return idx;
}

struct Nz { int r, c, s; };

// swap the spins of orbital m (swapspin.f90): sign -1 per index whose orbital m is doubly occupied
Nz swap_spin(const Nz& e, int d, int norbs, int m) {
  Nz out{0, 0, e.s};
  for (int w = 0; w < 2; ++w) {
    auto o = decode(w == 0 ? e.r : e.c, d, norbs);
    if (o[m].up && o[m].dn) out.s = -out.s;
    std::swap(o[m].up, o[m].dn);
    (w == 0 ? out.r : out.c) = encode(o);
  }
  return out;
}

// swap orbitals m-1 and m for one spin (swapms.f90)
Nz swap_orb(const Nz& e, int d, int norbs, int m, int s) {
  Nz out{0, 0, e.s};
  for (int w = 0; w < 2; ++w) {
    auto o = decode(w == 0 ? e.r : e.c, d, norbs);
    int& a = s == 0 ? o[m - 1].up : o[m - 1].dn;
    int& b = s == 0 ? o[m].up : o[m].dn;
    if (a && b) out.s = -out.s;
    std::swap(a, b);
    (w == 0 ? out.r : out.c) = encode(o);
  }
  return out;
}
}  // namespace

static void build_operators(Model& m) {
  const int d = m.d, norbs = m.norbs;
  m.ams.assign((size_t)norbs * 2 * d * d, 0.0);
  std::vector<Nz> c1;
  // c_{1,up}: |empty><up| + |down><both| on orbital 1 for every other configuration
  for (int base = 0; base < d; base += 4) {
    c1.push_back(Nz{base + 0, base + 1, 1});
    c1.push_back(Nz{base + 2, base + 3, 1});
  }
  std::vector<std::vector<Nz>> cur(2);
  cur[0] = c1;
  for (auto& e : c1) cur[1].push_back(swap_spin(e, d, norbs, 0));
  auto store = [&](int orb, int s, const std::vector<Nz>& ent) {
    for (auto& e : ent) m.ams[((size_t)(orb * 2 + s) * d + e.c) * d + e.r] = e.s;
  };
  store(0, 0, cur[0]);
  store(0, 1, cur[1]);
  int s = 1;
  for (int orb = 1; orb < norbs; ++orb) {
    std::vector<Nz> nw;
    for (auto& e : cur[s]) nw.push_back(swap_orb(e, d, norbs, orb, s));
    cur[s] = nw;
    store(orb, s, nw);
    s = 1 - s;
    nw.clear();
    for (auto& e : cur[1 - s]) nw.push_back(swap_spin(e, d, norbs, orb));
    cur[s] = nw;
    store(orb, s, nw);
  }
  std::sort(c1.begin(), c1.end(), [](const Nz& a, const Nz& b) { return a.r < b.r; });
}

// ---------------------------------------------------------------------------
// Hamiltonian (main/calchs.f90:368-387 para1, :514-567 para4), divided by hbar.
// ---------------------------------------------------------------------------
static void build_hamiltonian(Model& m) {
  const int d = m.d;
  m.H.assign((size_t)d * d, cplx(0, 0));
  auto C = [&](int orb, int s) {
    std::vector<double> c(d * d);
    for (int j = 0; j < d; ++j)
      for (int i = 0; i < d; ++i) c[i + j * d] = m.amsv(orb, s, i, j);
    return c;
  };
  auto mul = [&](const std::vector<double>& a, const std::vector<double>& b, bool ta) {
    std::vector<double> r(d * d, 0.0);
    for (int j = 0; j < d; ++j)
      for (int k = 0; k < d; ++k) {
        const double bk = b[k + j * d];
        if (bk == 0.0) continue;
        for (int i = 0; i < d; ++i) r[i + j * d] += (ta ? a[k + i * d] : a[i + k * d]) * bk;
      }
    return r;
  };
  auto add = [&](const std::vector<double>& a, double f) {
    for (int t = 0; t < d * d; ++t) m.H[t] += f * a[t];
  };
  const double* hp = m.p.hpar;
  if (m.p.ham == HEOMQ_HAM_SIAM) {
    auto cu = C(0, 0), cdn = C(0, 1);
    auto nu = mul(cu, cu, true), nd = mul(cdn, cdn, true);
    add(nu, hp[0]);
    add(nd, hp[1]);
    add(mul(nu, nd, false), hp[2]);
    auto sf = mul(cu, cdn, true);   // c_up^+ c_down
    for (int j = 0; j < d; ++j)
      for (int i = 0; i < d; ++i) m.H[i + j * d] += hp[3] * (sf[i + j * d] + sf[j + i * d]);
  } else {
    auto c1u = C(0, 0), c1d = C(0, 1), c2u = C(1, 0), c2d = C(1, 1);
    auto n1u = mul(c1u, c1u, true), n1d = mul(c1d, c1d, true);
    auto n2u = mul(c2u, c2u, true), n2d = mul(c2d, c2d, true);
    add(n1u, hp[0]); add(n1d, hp[1]); add(n2u, hp[2]); add(n2d, hp[3]);
    add(mul(n1u, n1d, false), hp[4]);
    add(mul(n2u, n2d, false), hp[5]);
    std::vector<double> n1(d * d), n2(d * d);
    for (int t = 0; t < d * d; ++t) { n1[t] = n1u[t] + n1d[t]; n2[t] = n2u[t] + n2d[t]; }
    add(mul(n1, n2, false), hp[6]);
    auto h1 = mul(c1u, c2u, true), h2 = mul(c2u, c1u, true), h3 = mul(c1d, c2d, true), h4 = mul(c2d, c1d, true);
    for (int t = 0; t < d * d; ++t) m.H[t] += hp[7] * (h1[t] + h2[t] + h3[t] + h4[t]);
  }
  for (auto& h : m.H) h /= kHbar;
}

// ---------------------------------------------------------------------------
// Bath coefficients (evaluatepara.f90:1897-1982; cgamma :2210-2227; bias
// eleadinfty :2431 with getengyshift(t -> inf) = amplitude).
// ---------------------------------------------------------------------------
// lscale (main/evaluatepara.f90:2927-2990): dbsqrt_k = |cb_k|^(1/2) of the
// unscaled cb, then cb and cd divided by it (by nothing where |cb| <= dpico).
static void apply_lscale(Model& m) {
  m.lscale = m.p.lscale != 0;
  m.dbsqrt.assign(m.M, 1.0);
  if (!m.lscale) return;
  for (int k = 0; k < m.M; ++k) {
    const double a = std::abs(m.cb[k]);
    m.dbsqrt[k] = std::sqrt(a);
    const double inv = a > 1e-12 ? 1.0 / m.dbsqrt[k] : 0.0;
    m.cb[k] *= inv;
    m.cd[k] *= inv;
  }
}

static void build_bath(Model& m) {
  std::vector<double> eta, xi;
  pade_psd(m.p.nmats, eta, xi);
  const int M = m.M;
  m.cb.assign(M, 0);
  m.cd.assign(M, 0);
  m.cgamma.assign(M, 0);
  m.gdraw.assign(M, 0);
  m.eleadinfty.assign((size_t)m.nalf * m.nspin, 0.0);
  for (int a = 0; a < m.nalf; ++a)
    for (int s = 0; s < m.nspin; ++s) m.eleadinfty[a * m.nspin + s] = m.p.engyshift[a][s] / kHbar;
  const cplx I(0, 1);
  const double h2 = kHbar * kHbar;
  for (int k = 0; k < M; ++k) {
    const int a = m.alf[k], orb = m.orb[k], sg = m.sgn[k], j = m.cor[k];
    const double W = m.p.bandwidth[a], T = m.p.temperature[a];
    const double zxip = m.p.linewidth[a][orb] * W * W * 0.5;
    const double dpm = sg == 0 ? 1.0 : -1.0;
    if (j == 0) {   // Drude pole of the Lorentzian band
      m.cgamma[k] = cplx(-W, 0.0) / kHbar;
      const cplx f = fermi_pade(cplx(0.0, W / T), eta, xi);
      m.cb[k] = zxip / W * f / h2;
      m.cd[k] = std::conj(zxip / W * f) / h2;
    } else {        // Pade poles
      const int jm = j - 1;
      const cplx pole = sg == 0 ? cplx(0, xi[jm]) : cplx(0, -xi[jm]);
      const cplx pole_b = -pole;
      const double coef = -eta[jm];
      m.cgamma[k] = dpm * I * T * pole / kHbar;
      m.cb[k] = I * 2.0 * coef * zxip * T / ((pole * T) * (pole * T) + W * W) / h2;
      m.cd[k] = std::conj(I * 2.0 * coef * zxip * T) / std::conj((pole_b * T) * (pole_b * T) + W * W) / h2;
    }
    const double dipm = sg == 0 ? 1.0 : -1.0;
    m.gdraw[k] = m.cgamma[k];
    if (m.steady) m.gdraw[k] += I * dipm * m.eleadinfty[a * m.nspin + m.spin[k]];
  }
  apply_lscale(m);
}

// ---------------------------------------------------------------------------
// Hierarchy (ball2drawer.f90:84-253, findoprindex.f90, sortindextable3.f90):
// for each tier the sorted drawer tuples whose operator multiset is not
// lexicographically above its sign-flipped complement, in lex order.
// ---------------------------------------------------------------------------
namespace {
int parity_strict(const int* v, int n) {
  int inv = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) inv += v[i] > v[j];
  return (inv & 1) ? -1 : 1;
}

struct Lookup {
  const Model* m;
  // binary search of a sorted tuple in tier n; returns local index or -1
  int64_t find(int n, const int* K) const {
    const auto& T = m->tiers[n];
    int64_t lo = 0, hi = (int64_t)T.size() / n - 1;
    while (lo <= hi) {
      const int64_t mid = (lo + hi) / 2;
      const int32_t* r = &T[mid * n];
      int cmp = 0;
      for (int t = 0; t < n && !cmp; ++t) cmp = (r[t] < K[t]) ? -1 : (r[t] > K[t]) ? 1 : 0;
      if (cmp == 0) return mid;
      if (cmp < 0) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
  }
  // findindex.f90:51-76: direct record, else reversed sign-flipped complement
  bool locate(int n, const int* Ksorted, int64_t& idx, int& itype, int& fact) const {
    if (n == 0) { idx = 0; itype = 0; fact = 1; return true; }
    int64_t loc = find(n, Ksorted);
    if (loc >= 0) { idx = m->first[n] + loc; itype = 0; fact = 1; return true; }
    int comp[32];
    for (int t = 0; t < n; ++t) comp[t] = m->bar[Ksorted[n - 1 - t]];
    fact = parity_strict(comp, n);
    std::sort(comp, comp + n);
    loc = find(n, comp);
    if (loc < 0) return false;
    idx = m->first[n] + loc;
    itype = 1;
    return true;
  }
};
}  // namespace

static bool opset_recorded(const Model& m, const int* K, int n) {
  int ops[32], comp[32];
  const int half = m.nopr / 2;
  for (int t = 0; t < n; ++t) {
    ops[t] = K[t] / m.ncor;
    comp[t] = ops[t] < half ? ops[t] + half : ops[t] - half;
  }
  std::sort(comp, comp + n);
  for (int t = 0; t < n; ++t) {
    if (ops[t] < comp[t]) return true;
    if (ops[t] > comp[t]) return false;
  }
  return true;
}

static void build_hierarchy(Model& m) {
  m.tiers.assign(m.L + 1, {});
  m.first.assign(m.L + 2, 0);
  m.first[0] = 0;
  m.first[1] = 1;
  const bool rep = m.p.lsdraw != 0;
  for (int n = 1; n <= m.L; ++n) {
    std::vector<int> K(n);
    for (int t = 0; t < n; ++t) K[t] = rep ? 0 : t;
    auto& out = m.tiers[n];
    if (!rep && n > m.M) { m.first[n + 1] = m.first[n]; continue; }
    while (true) {
      if (opset_recorded(m, K.data(), n)) out.insert(out.end(), K.begin(), K.end());
      int t = n - 1;   // odometer in lex order
      while (t >= 0) {
        const int maxv = rep ? m.M - 1 : m.M - n + t;
        if (K[t] < maxv) break;
        --t;
      }
      if (t < 0) break;
      ++K[t];
      for (int u = t + 1; u < n; ++u) K[u] = rep ? K[t] : K[u - 1] + 1;
    }
    m.first[n + 1] = m.first[n] + (int64_t)out.size() / n;
  }
  m.N = m.first[m.L + 1];
}

// ---------------------------------------------------------------------------
// Coupling table (buildcoefindex.f90:239-362 with findlowertier.f90:153-200,
// findhighertier.f90:50-107).  Upper entries in the loop order isgn, ispin,
// iorbs, ialf, imats, which is the drawer order: entry e is drawer e.
// ---------------------------------------------------------------------------
static int build_coupling(Model& m) {
  Lookup lk{&m};
  m.lower.clear();
  m.lower_off.assign(m.N + 1, 0);
  m.n_upper_ados = m.first[m.L];
  m.upper.assign((size_t)m.n_upper_ados * m.M, UpperEntry{-1, 0, 0, 0});
  for (int n = 0; n <= m.L; ++n) {
    const auto& T = m.tiers[n];
    const int64_t cnt = m.first[n + 1] - m.first[n];
    for (int64_t a = 0; a < cnt; ++a) {
      const int64_t g = m.first[n] + a;
      const int* K = n ? &T[a * n] : nullptr;
      m.lower_off[g] = (int64_t)m.lower.size();
      int buf[33];
      for (int r = 0; r < n; ++r) {
        int q = 0;
        for (int t = 0; t < n; ++t) if (t != r) buf[q++] = K[t];
        int64_t idx; int itype, fact;
        if (!lk.locate(n - 1, buf, idx, itype, fact))
          return fail(HEOMQ_BAD_ARG, "lower-tier ADO not found");
        int ff = (r % 2 == 0) ? 1 : -1;                  // buildcoefindex.f90:267-271
        int fr = ((n - 1 - r) % 2 == 0) ? 1 : -1;        // :272-276
        if (itype) { ff *= fact; fr *= fact; }
        m.lower.push_back(LowerEntry{(int32_t)idx, (int8_t)itype, (int8_t)ff, (int8_t)fr, K[r]});
      }
      if (n < m.L) {
        for (int e = 0; e < m.M; ++e) {
          bool present = false;
          int less = 0, greater = 0;
          for (int t = 0; t < n; ++t) {
            present |= K[t] == e;
            less += K[t] < e;
            greater += K[t] > e;
          }
          if (present && !m.p.lsdraw) continue;   // Pauli: no unique-index ADO
          int q = 0, ins = 0;
          for (int t = 0; t < n; ++t) {
            if (!ins && e < K[t]) { buf[q++] = e; ins = 1; }
            buf[q++] = K[t];
          }
          if (!ins) buf[q++] = e;
          int64_t idx; int itype, fact;
          if (!lk.locate(n + 1, buf, idx, itype, fact))
            return fail(HEOMQ_BAD_ARG, "higher-tier ADO not found");
          const int ffact = (less & 1) ? -1 : 1;          // findhighertier.f90:52-65
          const int rfact = (greater & 1) ? -1 : 1;       // :69-82
          m.upper[(size_t)g * m.M + e] = UpperEntry{(int32_t)idx, (int8_t)itype,
                                                    (int8_t)(ffact * fact), (int8_t)(rfact * fact)};
        }
      }
    }
  }
  m.lower_off[m.N] = (int64_t)m.lower.size();
  return HEOMQ_OK;
}

static void init_drawers(Model& m) {
  m.sgn.clear(); m.spin.clear(); m.orb.clear(); m.alf.clear(); m.cor.clear(); m.bar.clear();
  for (int sg = 0; sg < 2; ++sg)
    for (int s = 0; s < m.nspin; ++s)
      for (int o = 0; o < m.norbs; ++o)
        for (int a = 0; a < m.nalf; ++a)
          for (int j = 0; j < m.ncor; ++j) {
            m.sgn.push_back(sg); m.spin.push_back(s); m.orb.push_back(o);
            m.alf.push_back(a); m.cor.push_back(j);
          }
  const int half = m.M / 2;
  for (int k = 0; k < m.M; ++k) m.bar.push_back(k < half ? k + half : k - half);
}

static int check_params(const heomq_params& p) {
  if (p.nspin != 2) return fail(HEOMQ_UNSUPPORTED, "only nspin = 2 is supported");
  if (p.norbs < 1 || p.norbs > 3) return fail(HEOMQ_UNSUPPORTED, "norbs must be 1..3");
  if (p.nalf < 1 || p.nalf > 4) return fail(HEOMQ_BAD_ARG, "nalf must be 1..4");
  if (p.nmats < 1 || p.nmats > 64) return fail(HEOMQ_BAD_ARG, "nmats must be 1..64");
  if (p.ntop < 1 || p.ntop > 8) return fail(HEOMQ_BAD_ARG, "ntop must be 1..8");
  if (p.ham == HEOMQ_HAM_DQD && p.norbs != 2) return fail(HEOMQ_BAD_ARG, "two-impurity model needs norbs = 2");
  if (p.ham != HEOMQ_HAM_SIAM && p.ham != HEOMQ_HAM_DQD) return fail(HEOMQ_BAD_ARG, "unknown Hamiltonian kind");
  if (p.ham == HEOMQ_HAM_SIAM && p.norbs != 1) return fail(HEOMQ_BAD_ARG, "SIAM needs norbs = 1");
  if (p.lscale != 0 && p.lscale != 1) return fail(HEOMQ_BAD_ARG, "lscale must be 0 or 1");
  for (int a = 0; a < p.nalf; ++a)
    if (!(p.bandwidth[a] > 0) || !(p.temperature[a] > 0)) return fail(HEOMQ_BAD_ARG, "bandwidth and temperature must be > 0");
  return HEOMQ_OK;
}

int build_model(const heomq_params& p, Model& m) {
  if (int rc = check_params(p)) return rc;
  m = Model{};
  m.p = p;
  m.from_params = true;
  m.norbs = p.norbs; m.nspin = p.nspin; m.nalf = p.nalf;
  m.ncor = 1 + p.nmats;
  m.nopr = 2 * p.norbs * p.nspin * p.nalf;
  m.M = m.nopr * m.ncor;
  m.d = 1;
  for (int t = 0; t < p.norbs; ++t) m.d *= 2 * p.nspin;
  m.L = p.ntop;
  m.steady = p.steady != 0;
  init_drawers(m);
  build_bath(m);
  build_operators(m);
  build_hamiltonian(m);
  build_hierarchy(m);
  if (m.N > (int64_t)1 << 28) return fail(HEOMQ_UNSUPPORTED, "hierarchy too large (N > 2^28 ADOs)");
  return build_coupling(m);
}

int set_params(Model& m, const heomq_params& p) {
  if (int rc = check_params(p)) return rc;
  if (p.norbs != m.p.norbs || p.nalf != m.p.nalf || p.nmats != m.p.nmats || p.ntop != m.p.ntop ||
      p.lsdraw != m.p.lsdraw || p.ham != m.p.ham || p.lscale != m.p.lscale)
    return fail(HEOMQ_BAD_ARG, "set_params: hierarchy shape (or the lscale representation) differs");
  m.p = p;
  m.steady = p.steady != 0;
  build_bath(m);
  build_hamiltonian(m);
  return HEOMQ_OK;
}

// ---------------------------------------------------------------------------
// Drop-in loader: the reference's module arrays (module/matmod.f90), 1-based.
// ---------------------------------------------------------------------------
int model_from_reference(
    int32_t nrho, int64_t nunk, int32_t ntier, int32_t norbs, int32_t nspin, int32_t nalf,
    int32_t ncor, int32_t igroundsteady, int64_t ntable, int64_t lall, const int32_t* indextable,
    const int64_t* nfirst,
    const int64_t* nlast, const int64_t* ifirst, const int64_t* indexcoef,
    const int8_t* itypecoef, const int8_t* ifactfcoef, const int8_t* ifactrcoef,
    const int64_t* index_coef_ref, const int32_t* indexdraw, const double* cb,
    const double* cd, const double* cgamma, const double* dipm, const int32_t* ilead,
    const int32_t* mspin, const int32_t* morbs, const int32_t* mpm, const double* eleadinfty,
    const double* hs, const double* amsall, const double* dbsqrt, Model& m) {
  if (nspin != 2) return fail(HEOMQ_UNSUPPORTED, "only nspin = 2 is supported");
  int dd = 1;
  for (int t = 0; t < norbs; ++t) dd *= 4;
  if (nrho != dd) return fail(HEOMQ_BAD_DIM, "nrho != (2*nspin)**norbs");
  if (ntable < 1 || lall < 0) return fail(HEOMQ_BAD_ARG, "ntable < 1 or lall < 0");
  m = Model{};
  m.norbs = norbs; m.nspin = nspin; m.nalf = nalf; m.ncor = ncor;
  m.nopr = 2 * norbs * nspin * nalf;
  m.M = m.nopr * ncor;
  m.d = nrho;
  m.L = ntier - 1;
  m.N = nunk;
  m.steady = igroundsteady != 0;
  m.p.norbs = norbs; m.p.nspin = nspin; m.p.nalf = nalf; m.p.nmats = ncor - 1;
  m.p.ntop = m.L; m.p.lsdraw = 1; m.p.steady = m.steady;
  init_drawers(m);
  const int M = m.M;
  // drawer attributes given by the reference must match its own ordering
  for (int k = 0; k < M; ++k) {
    if (mpm[k] - 1 != m.sgn[k] || mspin[k] - 1 != m.spin[k] || morbs[k] - 1 != m.orb[k] ||
        ilead[k] - 1 != m.alf[k])
      return fail(HEOMQ_BAD_ARG, "drawer attributes do not follow the reference ordering");
  }
  m.cb.resize(M); m.cd.resize(M); m.cgamma.resize(M); m.gdraw.resize(M);
  m.dbsqrt.assign(M, 1.0);
  m.lscale = dbsqrt != nullptr;
  m.p.lscale = m.lscale ? 1 : 0;
  m.eleadinfty.assign((size_t)nalf * nspin, 0.0);
  for (int a = 0; a < nalf; ++a)
    for (int s = 0; s < nspin; ++s) m.eleadinfty[a * nspin + s] = eleadinfty[a + nalf * s];
  for (int k = 0; k < M; ++k) {
    const size_t idx = m.orb[k] + (size_t)norbs * (m.spin[k] + (size_t)nspin * (m.cor[k] + (size_t)ncor * (m.alf[k] + (size_t)nalf * m.sgn[k])));
    m.cb[k] = cplx(cb[2 * idx], cb[2 * idx + 1]);
    m.cd[k] = cplx(cd[2 * idx], cd[2 * idx + 1]);
    m.cgamma[k] = cplx(cgamma[2 * k], cgamma[2 * k + 1]);
    m.gdraw[k] = m.cgamma[k];
    if (m.steady) m.gdraw[k] += cplx(0, 1) * dipm[k] * m.eleadinfty[m.alf[k] * nspin + m.spin[k]];
    // lscale: the reference has already divided cb and cd by dbsqrt
    // (evaluatepara.f90:2976-2990); dbsqrt scales the upper couplings
    m.dbsqrt[k] = dbsqrt ? dbsqrt[idx] : 1.0;
    if (dbsqrt && !(m.dbsqrt[k] >= 0.0)) return fail(HEOMQ_BAD_ARG, "dbsqrt must be >= 0");
  }
  const int d = nrho;
  m.H.assign((size_t)d * d, 0);
  for (int j = 0; j < d; ++j)          // zhemm 'u': only the upper triangle is referenced
    for (int i = 0; i <= j; ++i) {
      const cplx h(hs[2 * (i + j * d)], hs[2 * (i + j * d) + 1]);
      m.H[i + j * d] = h;
      m.H[j + i * d] = std::conj(h);
    }
  m.ams.assign((size_t)norbs * nspin * d * d, 0.0);
  for (int s = 0; s < nspin; ++s)
    for (int o = 0; o < norbs; ++o)
      for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i)
          m.ams[((size_t)(o * nspin + s) * d + j) * d + i] = amsall[i + (size_t)d * (j + (size_t)d * (o + (size_t)norbs * s))];
  // tiers
  m.first.assign(m.L + 2, 0);
  m.tiers.assign(m.L + 1, {});
  for (int it = 1; it <= ntier; ++it) {
    const int n = it - 1;
    m.first[n] = nfirst[it - 1] - 1;
    const int64_t cnt = nlast[it - 1] - nfirst[it - 1] + 1;
    if (cnt < (n == 0 ? 1 : 0) || m.first[n] != (n == 0 ? 0 : nlast[it - 2]) || nlast[it - 1] > nunk)
      return fail(HEOMQ_BAD_ARG, "nfirst/nlast do not tile 1..nunk tier by tier");
    if (n > 0) {
      // tier n's records occupy indextable(ifirst(it) .. ifirst(it) + cnt*n - 1)
      if (ifirst[it - 1] < 1 || ifirst[it - 1] - 1 + cnt * n > ntable)
        return fail(HEOMQ_BAD_ARG, "ifirst: tier " + std::to_string(n) + " records run past indextable(ntable)");
      m.tiers[n].resize((size_t)cnt * n);
      for (int64_t a = 0; a < cnt; ++a)
        for (int r = 0; r < n; ++r) {
          const int32_t k = indextable[ifirst[it - 1] - 1 + a * n + r] - 1;
          // drawer tuples are sorted, ids in 1..nvar (fermion/sortindextable3.f90)
          if (k < 0 || k >= M || (r > 0 && k < m.tiers[n][a * n + r - 1]))
            return fail(HEOMQ_BAD_ARG, "indextable: drawer id out of range or tuple not sorted (tier " +
                                           std::to_string(n) + ", ADO " + std::to_string(a + 1) + ")");
          m.tiers[n][a * n + r] = k;
        }
    }
  }
  m.first[m.L + 1] = nunk;
  if (nlast[ntier - 1] != nunk) return fail(HEOMQ_BAD_ARG, "nlast(ntier) != nunk");
  auto bad_entry = [&](int32_t nb, int8_t it, bool upper) {
    return nb < (upper ? -1 : 0) || nb >= nunk || (it != 0 && it != 1);
  };
  // coupling
  m.lower.clear();
  m.lower_off.assign(m.N + 1, 0);
  m.n_upper_ados = m.first[m.L];
  m.upper.assign((size_t)m.n_upper_ados * M, UpperEntry{-1, 0, 0, 0});
  const int half = M / 2;
  for (int e = 0; e < M; ++e) {        // tier 0 (calc_ax_tfqmr_omp.f90:108-149 via indexdraw)
    const int pe = e < half ? e : e - half;
    if (m.L >= 1 && (indexdraw[pe] < 1 || indexdraw[pe] > nlast[1] - nfirst[1] + 1))
      return fail(HEOMQ_BAD_ARG, "indexdraw out of tier-1 range");
    const int32_t nb = (int32_t)(nfirst[1] - 1 + indexdraw[pe] - 1);
    m.upper[e] = UpperEntry{nb, (int8_t)(e < half ? 0 : 1), 1, 1};
  }
  m.lower_off[0] = 0;
  for (int n = 1; n <= m.L; ++n) {
    for (int64_t g = m.first[n]; g < m.first[n + 1]; ++g) {
      m.lower_off[g] = (int64_t)m.lower.size();
      int64_t lnl = index_coef_ref[g] - 1;   // 0-based position into the coupling stream
      if (lnl < 0) return fail(HEOMQ_BAD_ARG, "index_coef_ref < 1");
      // n lower entries, then M upper ones below the top tier
      if (lnl + n + (n < m.L ? M : 0) > lall)
        return fail(HEOMQ_BAD_ARG, "index_coef_ref: ADO " + std::to_string(g + 1) +
                                       "'s couplings run past indexcoef(lall)");
      for (int r = 0; r < n; ++r, ++lnl) {
        const int32_t nb = (int32_t)(indexcoef[lnl] - 1);
        if (bad_entry(nb, itypecoef[lnl], false)) return fail(HEOMQ_BAD_ARG, "indexcoef: bad lower-tier entry");
        m.lower.push_back(LowerEntry{nb, itypecoef[lnl], ifactfcoef[lnl], ifactrcoef[lnl],
                                     m.tiers[n][(g - m.first[n]) * n + r]});
      }
      if (n < m.L)
        for (int e = 0; e < M; ++e, ++lnl) {
          const int32_t nb = (int32_t)(indexcoef[lnl] - 1);
          if (nb == -1) continue;   // absent (indexcoef = 0): its type and signs are not read
          if (bad_entry(nb, itypecoef[lnl], true)) return fail(HEOMQ_BAD_ARG, "indexcoef: bad upper-tier entry");
          m.upper[(size_t)g * M + e] = UpperEntry{nb, itypecoef[lnl], ifactfcoef[lnl], ifactrcoef[lnl]};
        }
    }
  }
  m.lower_off[m.N] = (int64_t)m.lower.size();
  return HEOMQ_OK;
}

}  // namespace heomq

namespace heomq {

namespace {
// splitmix64 and a Box-Muller normal from it: the mixing coefficients of the
// pattern builder, a pure function of (seed, ADO, coupling, slot)
uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double normal_of(uint64_t key) {
  const double u1 = ((mix64(key) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((mix64(key ^ 0xA5A5A5A5A5A5A5A5ull) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}
}  // namespace

// The sparse-mode pattern of fermion/prelude_spa.f90:532-749.  Tier 0 holds
// the nonzeros of E = sum_{k <= 10} (a H)^k / k!, a = 0.31415926
// (utility/calc_exp_hmat.f90, H already / hbar; prelude_spa.f90:231-238
// requires 10 and 11 terms to give the same sparsity), with E as its values.
// An ADO K of tier n >= 1 sums, over its lower neighbours J (K without drawer
// k), m1 E (op_k X) + m2 E (X op_k) with X = rho_J (or rho_J^H when J is
// stored as K's partner's neighbour) and op_k = c^+ for sigma = +, c for
// sigma = - (:640-686); m1, m2 are random complex normals that keep the two
// products from cancelling (:676-683).  K's pattern is the elements of the sum
// above dsmall in column-major order (zmat_dns2coo), and the sum is K's value
// for the next tier.  A drawer whose lead-orbital coupling vanishes empties
// the ADO (:650-653, :733-737).
int spa_pattern_build(const Model& m, double dsmall, uint64_t seed, std::vector<int32_t>& nnz,
                      std::vector<int32_t>& irow, std::vector<int32_t>& icol) {
  const int d = m.d, dd = d * d;
  auto expH = [&](int kmax) {
    std::vector<cplx> out((size_t)dd, cplx(0, 0)), pw((size_t)dd, cplx(0, 0)), nx((size_t)dd);
    for (int i = 0; i < d; ++i) out[i + (size_t)i * d] = pw[i + (size_t)i * d] = cplx(1, 0);
    double f = 1.0;
    for (int k = 1; k <= kmax; ++k) {
      f /= k;
      for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i) {
          cplx s(0, 0);
          for (int q = 0; q < d; ++q) s += pw[i + (size_t)q * d] * m.H[q + (size_t)j * d];
          nx[i + (size_t)j * d] = 0.31415926 * s;
        }
      pw.swap(nx);
      for (size_t t = 0; t < (size_t)dd; ++t) out[t] += f * pw[t];
    }
    return out;
  };
  const std::vector<cplx> E = expH(10), E11 = expH(11);
  for (size_t t = 0; t < (size_t)dd; ++t)
    if ((std::abs(E[t]) > dsmall) != (std::abs(E11[t]) > dsmall))
      return fail(HEOMQ_UNSUPPORTED, "sparse pattern: the sparsity of exp(a H) differs between 10 and 11 terms");
  // columns of E as (row, value) lists
  std::vector<std::vector<std::pair<int, cplx>>> Ecol(d);
  for (int k = 0; k < d; ++k)
    for (int i = 0; i < d; ++i)
      if (E[i + (size_t)k * d] != cplx(0, 0)) Ecol[k].push_back({i, E[i + (size_t)k * d]});
  nnz.assign(m.N, 0);
  irow.clear();
  icol.clear();
  std::vector<int64_t> start(m.N + 1, 0);
  std::vector<cplx> val;   // every stored value, ADO after ADO
  auto emit = [&](int64_t g, const std::vector<cplx>& Z, bool empty) {
    start[g] = (int64_t)irow.size();
    if (!empty)
      for (int j = 0; j < d; ++j)
        for (int i = 0; i < d; ++i) {
          const cplx z = Z[i + (size_t)j * d];
          if (std::abs(z) > dsmall) {
            irow.push_back(i + 1);
            icol.push_back(j + 1);
            val.push_back(z);
          }
        }
    nnz[g] = (int32_t)((int64_t)irow.size() - start[g]);
  };
  emit(0, E, false);
  // op_k (r, c) as row lists (for X op) and column lists (for op X)
  auto opv = [&](int k, int r, int c) { return m.sgn[k] == 0 ? m.amsv(m.orb[k], m.spin[k], c, r)
                                                             : m.amsv(m.orb[k], m.spin[k], r, c); };
  std::vector<cplx> Z((size_t)dd), YL((size_t)dd), YR((size_t)dd);
  for (int n = 1; n <= m.L; ++n)
    for (int64_t g = m.first[n]; g < m.first[n + 1]; ++g) {
      std::fill(Z.begin(), Z.end(), cplx(0, 0));
      bool empty = false;
      for (int64_t p = m.lower_off[g]; p < m.lower_off[g + 1]; ++p) {
        const LowerEntry& le = m.lower[p];
        const int k = le.drawer;
        if (m.cb[k] == cplx(0, 0) && m.cd[k] == cplx(0, 0)) { empty = true; continue; }
        std::fill(YL.begin(), YL.end(), cplx(0, 0));
        std::fill(YR.begin(), YR.end(), cplx(0, 0));
        const int64_t J = le.nbr;
        for (int64_t e = start[J]; e < start[J] + nnz[J]; ++e) {
          int a = irow[e] - 1, b = icol[e] - 1;
          cplx x = val[e];
          if (le.itype) { std::swap(a, b); x = std::conj(x); }   // X^H
          for (int r = 0; r < d; ++r) {
            const double o = opv(k, r, a);                        // (op X)(r, b) += op(r, a) X(a, b)
            if (o != 0.0) YL[r + (size_t)b * d] += o * x;
          }
          for (int c = 0; c < d; ++c) {
            const double o = opv(k, b, c);                        // (X op)(a, c) += X(a, b) op(b, c)
            if (o != 0.0) YR[a + (size_t)c * d] += x * o;
          }
        }
        const uint64_t key = seed * 0x100000001B3ull + (uint64_t)g * 64 + (uint64_t)(p - m.lower_off[g]) * 4;
        cplx m1, m2;
        for (uint64_t t = 0;; t += 0x10000) {
          m1 = cplx(normal_of(key + t), normal_of(key + t + 1));
          m2 = cplx(normal_of(key + t + 2), normal_of(key + t + 3));
          if (std::abs(m1) > 1e-6 && std::abs(m2) > 1e-6) break;
        }
        for (int j = 0; j < d; ++j)
          for (int q = 0; q < d; ++q) {
            const cplx yl = YL[q + (size_t)j * d], yr = YR[q + (size_t)j * d];
            if (yl == cplx(0, 0) && yr == cplx(0, 0)) continue;
            const cplx y = m1 * yl + m2 * yr;
            for (const auto& ev : Ecol[q]) Z[ev.first + (size_t)j * d] += ev.second * y;
          }
      }
      emit(g, Z, empty);
    }
  return HEOMQ_OK;
}

}  // namespace heomq
