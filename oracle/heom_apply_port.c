/*
 * CPU restatement of the reference's dense Liouvillian apply -- TEST AND
 * BASELINE INFRASTRUCTURE ONLY (the checker's CPU port, never the product).
 *
 * Follows tfqmr_double/calc_ax_tfqmr_omp.f90 on the reference's own module
 * tables (module/matmod.f90 layout, 1-based values): an OpenMP loop over ADOs
 * with a dynamic schedule (:190-200), per ADO the same-tier term
 * -i(H X - X H) + gamma_K X (:180-203, zhemm there, dense d^3 loops here),
 * the lower-tier couplings (:221-271) and the upper-tier couplings
 * (:275-358) through the +-1 partial permutations of the c operators
 * (utility/zamsmm1_omp.f90), tier 0's upper couplings located through
 * indexdraw (:108-149), and the trace row (:492-495).
 *
 * Conventions (the same as the numpy oracle, oracle/heom_oracle.py): drawer
 * k (0-based) has sign sgn[k] (0: sigma = +), spin, orbital; lower coupling
 * through drawer k applies op = c^+ (sigma = +) or c (sigma = -):
 * y += -i ff cb_k op X~ + i fr cd_k X~ op; upper coupling through drawer e
 * applies op' = c (sigma = +) or c^+: y += -i ff op' X~ + i fr X~ op';
 * X~ = x_nb or x_nb^H (itype 1).  Matrices are column-major, element (i, j)
 * of ADO a at i + j*d + a*d*d.
 */
#include <complex.h>
#include <stdint.h>
#include <stdlib.h>
#include <math.h>
#include <string.h>

typedef double _Complex cplx;

typedef struct {
  int32_t nrho, ntier, nvar, norbs, nspin;
  int64_t nunk;
  const int32_t* indextable;      /* module/matmod.f90:4 (1-based drawer ids, tier 0's empty record first) */
  const int64_t* nfirst;          /* :13, 1-based */
  const int64_t* ifirst;          /* :12, 1-based position of each tier in indextable */
  const int64_t* indexcoef;       /* :7, 1-based neighbour ADO (0: absent) */
  const int8_t* itypecoef;        /* :8 */
  const int8_t* ifactfcoef;       /* :9 */
  const int8_t* ifactrcoef;       /* :9 */
  const int64_t* index_coef_ref;  /* :7, 1-based start of each ADO's entries */
  const int32_t* indexdraw;       /* :16, tier-1 position of each sigma = + drawer (1-based) */
  const double* cb;               /* per drawer, complex interleaved */
  const double* cd;
  const double* gdraw;            /* per drawer gamma incl. steady bias, complex interleaved */
  const double* hs;               /* d*d complex, column-major, already / hbar */
  const double* ams;              /* [(m*nspin + s)*d*d + i + j*d], real */
  const int32_t* sgn;             /* per drawer */
  const int32_t* spin;
  const int32_t* orb;
  const double* dbsqrt;           /* lscale: |cb_k|^(1/2) per drawer on every upper coupling (:120-122,
                                     :325-328); null without lscale */
} heom_port_tables;

/* one c operator as its +-1 entries (row, column, sign) */
typedef struct {
  int n;
  int row[64], col[64];
  double sg[64];
} sparse_op;

static void build_op(const heom_port_tables* t, int m, int s, int dagger, sparse_op* op) {
  const int d = t->nrho;
  const double* a = t->ams + (size_t)(m * t->nspin + s) * d * d;
  op->n = 0;
  for (int j = 0; j < d; ++j)
    for (int i = 0; i < d; ++i) {
      const double v = a[i + j * d];
      if (v != 0.0 && op->n < 64) {
        op->row[op->n] = dagger ? j : i;
        op->col[op->n] = dagger ? i : j;
        op->sg[op->n] = v;
        op->n++;
      }
    }
}

/* y += a * (op X) + b * (X op), X = x or x^H */
static void couple(int d, const sparse_op* op, const cplx* xn, int dag, cplx a, cplx b, cplx* y) {
  for (int e = 0; e < op->n; ++e) {
    const int r = op->row[e], c = op->col[e];
    const double s = op->sg[e];
    for (int j = 0; j < d; ++j) {   /* (op X)(r, j) += s X(c, j) */
      const cplx v = dag ? conj(xn[j + c * d]) : xn[c + j * d];
      y[r + j * d] += a * s * v;
    }
    for (int i = 0; i < d; ++i) {   /* (X op)(i, c) += s X(i, r) */
      const cplx v = dag ? conj(xn[r + i * d]) : xn[i + r * d];
      y[i + c * d] += b * s * v;
    }
  }
}

static int tier_of(const heom_port_tables* t, int64_t a) {
  int n = 0;
  while (n + 1 < t->ntier && t->nfirst[n + 1] - 1 <= a) ++n;
  return n;
}

/* y = L x; trace_mode 0: none, 1: y(d,d) of the last ADO = tr x0 (the
 * reference), 2: y(d,d) of ADO 0 = tr x0. */
void heom_port_apply(const heom_port_tables* t, const double* xr, double* yr, int trace_mode) {
  const int d = t->nrho, M = t->nvar, L = t->ntier - 1, half = M / 2;
  const cplx* x = (const cplx*)xr;
  cplx* y = (cplx*)yr;
  const cplx* H = (const cplx*)t->hs;
  const cplx* cb = (const cplx*)t->cb;
  const cplx* cd = (const cplx*)t->cd;
  const cplx* gd = (const cplx*)t->gdraw;
  sparse_op* lop = (sparse_op*)malloc(sizeof(sparse_op) * (size_t)M);   /* lower-coupling operator of each drawer */
  sparse_op* uop = (sparse_op*)malloc(sizeof(sparse_op) * (size_t)M);   /* upper-coupling operator */
  for (int k = 0; k < M; ++k) {
    build_op(t, t->orb[k], t->spin[k], t->sgn[k] == 0, &lop[k]);
    build_op(t, t->orb[k], t->spin[k], t->sgn[k] != 0, &uop[k]);
  }
  const int64_t n_ado = t->nunk;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t a = 0; a < n_ado; ++a) {
    const int n = tier_of(t, a);
    const cplx* X = x + a * d * d;
    cplx* Y = y + a * d * d;
    const int32_t* K = n > 0 ? t->indextable + (t->ifirst[n] - 1) + (a - (t->nfirst[n] - 1)) * n : 0;
    cplx gam = 0;
    for (int r = 0; r < n; ++r) gam += gd[K[r] - 1];
    /* same tier: -i (H X - X H) + gamma X (zhemm on the full H) */
    for (int j = 0; j < d; ++j)
      for (int i = 0; i < d; ++i) {
        cplx acc = 0;
        for (int k = 0; k < d; ++k) acc += H[i + k * d] * X[k + j * d] - X[i + k * d] * H[k + j * d];
        Y[i + j * d] = -I * acc + gam * X[i + j * d];
      }
    if (n > 0) {
      int64_t p = t->index_coef_ref[a] - 1;
      for (int r = 0; r < n; ++r, ++p) {   /* lower tier: drawer K[r] removed */
        const int k = K[r] - 1;
        const int64_t nb = t->indexcoef[p] - 1;
        couple(d, &lop[k], x + nb * d * d, t->itypecoef[p], -I * t->ifactfcoef[p] * cb[k],
               I * t->ifactrcoef[p] * cd[k], Y);
      }
      if (n < L)
        for (int e = 0; e < M; ++e, ++p) {   /* upper tier: drawer e added */
          if (t->indexcoef[p] <= 0) continue;
          const int64_t nb = t->indexcoef[p] - 1;
          const double sc = t->dbsqrt ? t->dbsqrt[e] : 1.0;
          couple(d, &uop[e], x + nb * d * d, t->itypecoef[p], -I * sc * t->ifactfcoef[p], I * sc * t->ifactrcoef[p],
                 Y);
        }
    } else if (L > 0) {
      for (int e = 0; e < M; ++e) {   /* tier 0: the tier-1 ADO of drawer e, or of its partner */
        const int pe = e < half ? e : e - half;
        const int64_t nb = (t->nfirst[1] - 1) + t->indexdraw[pe] - 1;
        const double sc = t->dbsqrt ? t->dbsqrt[e] : 1.0;
        couple(d, &uop[e], x + nb * d * d, e < half ? 0 : 1, -I * sc, I * sc, Y);
      }
    }
  }
  if (trace_mode == 1 || trace_mode == 2) {
    cplx tr = 0;
    for (int i = 0; i < d; ++i) tr += x[i + i * d];
    const int64_t pos = trace_mode == 1 ? n_ado - 1 : 0;
    y[pos * d * d + d * d - 1] = tr;
  }
  free(lop);
  free(uop);
}

int heom_port_threads(void);
#ifdef _OPENMP
#include <omp.h>
int heom_port_threads(void) { return omp_get_max_threads(); }
void heom_port_set_threads(int n) { omp_set_num_threads(n); }
#else
int heom_port_threads(void) { return 1; }
void heom_port_set_threads(int n) { (void)n; }
#endif

/*
 * TFQMR iterations (Freund, SIAM J. Sci. Comput. 14 (1993) 470) on A x = b
 * over the real (Re, Im) vector of n = 2 N d^2 doubles, with the work the
 * reference does per iteration of its reverse-communication loop
 * (tfqmr_double/solve_tfqmr.f90:194-280 around dutfx.f90:196-357): two
 * applies, nine axpby passes, two dot products and two norms.  par = 0 runs
 * the vector passes on one thread, as dutfx.f90, daxpby.f90, ddot_long.f90
 * and dnrm2_long.f90 do (they have no OpenMP); par = 1 spreads them over the
 * OpenMP threads.  init = 1 starts the recurrences (x = 0, the shadow
 * vector +-1 from a fixed hash -- the reference draws it with DRANDN once per
 * solve -- and one apply); init = 0 resumes them from work and state.  Then
 * nit iterations advance x.  work holds 8 vectors of n doubles, state the
 * scalars (tau, theta, eta, rho, m).  Returns the quasi-residual bound
 * tau sqrt(2 m + 1) after the last iteration.
 */
static double vdot(int64_t n, const double* a, const double* b, int par) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) if (par)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}
/* z = a x + b y (daxpby) */
static void vaxpby(int64_t n, double* z, double a, const double* x, double b, const double* y, int par) {
#pragma omp parallel for if (par)
  for (int64_t i = 0; i < n; ++i) z[i] = a * x[i] + b * y[i];
}

double heom_port_tfqmr(const heom_port_tables* t, const double* b, double* x, double* work, double* state,
                       int nit, int par, int trace_mode, int init) {
  const int64_t n = 2 * t->nunk * (int64_t)t->nrho * t->nrho;
  double *w = work, *y1 = work + n, *y2 = work + 2 * n, *u1 = work + 3 * n, *u2 = work + 4 * n;
  double *v = work + 5 * n, *dv = work + 6 * n, *rt = work + 7 * n;
  if (init) {
    for (int64_t i = 0; i < n; ++i) {
      uint64_t z = 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
      z = (z ^ (z >> 31)) * 0xBF58476D1CE4E5B9ull;
      rt[i] = ((z >> 40) & 1) ? 1.0 : -1.0;
      x[i] = 0.0;
      w[i] = b[i];
      y1[i] = b[i];
      dv[i] = 0.0;
    }
    heom_port_apply(t, y1, v, trace_mode);
    memcpy(u1, v, sizeof(double) * (size_t)n);
    state[0] = sqrt(vdot(n, w, w, par));
    state[1] = state[2] = 0.0;
    state[3] = vdot(n, rt, w, par);
    state[4] = 0.0;
  }
  double tau = state[0], theta = state[1], eta = state[2], rho = state[3];
  int m = (int)state[4];
  for (int it = 0; it < nit; ++it) {
    const double sigma = vdot(n, rt, v, par);
    const double alpha = rho / sigma;
    vaxpby(n, y2, 1.0, y1, -alpha, v, par);
    heom_port_apply(t, y2, u2, trace_mode);
    for (int j = 0; j < 2; ++j) {
      const double* yj = j ? y2 : y1;
      const double* uj = j ? u2 : u1;
      vaxpby(n, w, 1.0, w, -alpha, uj, par);
      vaxpby(n, dv, 1.0, yj, theta * theta * eta / alpha, dv, par);
      theta = sqrt(vdot(n, w, w, par)) / tau;
      const double c = 1.0 / sqrt(1.0 + theta * theta);
      tau = tau * theta * c;
      eta = c * c * alpha;
      vaxpby(n, x, 1.0, x, eta, dv, par);
    }
    ++m;
    const double rho_n = vdot(n, rt, w, par);
    const double beta = rho_n / rho;
    rho = rho_n;
    vaxpby(n, y1, 1.0, w, beta, y2, par);
    heom_port_apply(t, y1, u1, trace_mode);
#pragma omp parallel for if (par)
    for (int64_t i = 0; i < n; ++i) v[i] = u1[i] + beta * (u2[i] + beta * v[i]);
  }
  state[0] = tau;
  state[1] = theta;
  state[2] = eta;
  state[3] = rho;
  state[4] = m;
  return tau * sqrt(2.0 * m + 1.0);
}
