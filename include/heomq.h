/*
 * heomq.h -- C ABI of the MI355X-native HEOM Liouvillian engine.
 *
 * One data-parallel hot path of HEOM-QUICK2 (Fortran/OpenMP, MKL) is rebuilt
 * here for one gfx950 GPU: the action of the fermionic HEOM Liouvillian on the
 * auxiliary-density-operator (ADO) hierarchy, its real transpose, and the
 * steady-state / propagation loops that consume them.  Every entry point takes
 * plain pointers and sizes; device pointers are HIP device addresses, streams
 * are hipStream_t passed as void*.  No torch types appear here.
 *
 * Vector layout: an ADO vector is complex (re, im interleaved) in the
 * reference's Fortran order x(i, j, a) -- element (i, j) of ADO a at complex
 * offset i + j*d + a*d*d -- which is also the real (Re, Im)-interleaved TFQMR
 * vector of tfqmr_double/solve_tfqmr.f90:104-113.  FP64 entry points take
 * double*, FP32 ones float*.
 *
 * Error behaviour mirrors the reference's fatal checks (e.g. dim0 != nrho,
 * tfqmr_double/calc_ax_tfqmr_omp.f90:44-48) as status codes instead of STOP;
 * heomq_last_error() gives the message.  Calls on one context are not
 * reentrant (the reference applies are not either: `save` counters and shared
 * module scratch, calc_ax_tfqmr_omp.f90:34-35,61).
 */
#ifndef HEOMQ_H
#define HEOMQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  NOT_CONVERGED / BREAKDOWN follow DUTFX's INFO(1) meanings
 * (tfqmr_double/dutfx.f90:137-147: 4 = no convergence, 8 = breakdown). */
#define HEOMQ_OK             0
#define HEOMQ_BAD_DIM        1
#define HEOMQ_UNSUPPORTED    2
#define HEOMQ_HIP_ERROR      3
#define HEOMQ_NOT_CONVERGED  4
#define HEOMQ_BREAKDOWN      8
#define HEOMQ_STOPPED       16
#define HEOMQ_BAD_ARG       32

/* Length (doubles) of the stats array every solver entry point fills.  The
 * solvers write exactly these slots (unused ones are set to 0) and never
 * beyond them: heomq_solve_steady, heomq_solve_bicg, heomq_solve_jacobi and
 * their _host variants take HEOMQ_STATS_LEN; heomq_solve_steady_mixed takes
 * HEOMQ_STATS_MIXED_LEN; heomq_chebyshev takes HEOMQ_STATS_CHEBY_LEN. */
#define HEOMQ_STATS_LEN        8
#define HEOMQ_STATS_MIXED_LEN 64
#define HEOMQ_STATS_CHEBY_LEN  3

/* Trace (normalisation) row placement of the steady-state operator. */
#define HEOMQ_TRACE_NONE  0   /* plain L x (td/calcderiv_omp.f90)                         */
#define HEOMQ_TRACE_LAST  1   /* y(d,d,last ADO) = tr x0 (calc_ax_tfqmr_omp.f90:492-495)  */
#define HEOMQ_TRACE_RDO   2   /* y(d,d,tier 0) = tr x0 (a dependent row for every tier)   */

/* Hamiltonian kinds: SIAM ($para1, main/calchs.f90:257-387) and the
 * two-impurity model ($para4, main/calchs.f90:463-567). */
#define HEOMQ_HAM_SIAM 0
#define HEOMQ_HAM_DQD  1

typedef struct heomq_params {
  int32_t norbs, nspin, nalf;   /* nspin must be 2 */
  int32_t nmats;                /* Pade terms Nk; one Drude pole is added */
  int32_t ntop;                 /* top tier L (reference ntier = L + 1)   */
  int32_t lsdraw;               /* 1: keep ADOs with repeated drawers     */
  int32_t ham;                  /* HEOMQ_HAM_*                            */
  int32_t steady;               /* 1: bias enters gamma (igroundsteady=1) */
  int32_t lscale;               /* 1: ADOs stored scaled by prod_k |cb_k|^(-1/2) ($jobinfo
                                   lscale, main/evaluatepara.f90:2927-2990): cb, cd divided by
                                   |cb|^(1/2), every upper coupling times |cb_e|^(1/2); the
                                   vectors of every entry point are then the scaled ones      */
  int32_t reserved;             /* 0                                       */
  double bandwidth[4];          /* W per lead                              */
  double linewidth[4][4];       /* Gamma[lead][orbital]                    */
  double temperature[4];        /* k_B T per lead                          */
  double engyshift[4][2];       /* bias per [lead][spin]                   */
  /* SIAM: eup, edown, uu, tupdn.  DQD: e1up, e1down, e2up, e2down, uu1, uu2,
   * u12, t12.  Energies in meV; divided by hbar inside.                    */
  double hpar[16];
} heomq_params;

typedef struct heomq_model heomq_model;   /* host-side tables   */
typedef struct heomq_ctx   heomq_ctx;     /* device-resident    */

const char* heomq_last_error(void);
int heomq_version(void);

/* ---------------------------------------------------------------------------
 * Host model: bath decomposition, fermion operators, H, hierarchy and the
 * tier-index (coupling) table.  Replaces the preparation phase the reference
 * runs before its solvers: main/evaluatepara.f90:1157-2232 (counts, drawers,
 * cb/cd/cgamma), main/zpout_psd.f90 (Pade), main/calchs.f90 (H),
 * main/calcams.f90 + buildoperator.f90 (c operators), main/ball2drawer.f90 +
 * fermion/sortindextable3.f90 (hierarchy), main/buildcoefindex.f90 with
 * findlowertier/findhighertier/findindex (coupling table).
 * ------------------------------------------------------------------------- */
int heomq_model_build(const heomq_params* p, heomq_model** out);
void heomq_model_free(heomq_model* m);
/* info[0..7] = d, M, L, N (ADOs), ncor, nopr, n_lower_entries, n_upper_entries;
 * first[0..L+1] = first ADO of each tier (0-based, first[L+1] = N). */
int heomq_model_info(const heomq_model* m, int64_t* info, int64_t* first);
/* Per-drawer coefficients (complex, interleaved): cb, cd, gamma (incl. bias). */
int heomq_model_coefs(const heomq_model* m, double* cb, double* cd, double* gam);
/* H/hbar (d*d complex, column-major) and c_{m,s} (norbs*nspin*d*d real). */
int heomq_model_operators(const heomq_model* m, double* H, double* ams);
/* Drawer tuples of tier n (N_n * n int32, 0-based drawer ids). */
int heomq_model_tier(const heomq_model* m, int tier, int32_t* tuples);
/* Coupling table, 0-based: lower[a*n + r] and upper[(a - 0)*M + e] rows of
 * (nbr, itype, ffront, frear) as int32 quadruples; absent upper entries have
 * nbr = -1.  lower covers tiers >= 1 (ordered by ADO), upper tiers < L. */
int heomq_model_coupling(const heomq_model* m, int32_t* lower, int32_t* upper);
/* Symmetry-packed layout the device solvers use, built on the host (no
 * device needed; the same construction a context runs, with its consistency
 * checks).  info[0] = packed complex elements, [1] = sectors, [2] = N
 * conserved, [3] = S_z conserved, [4] = dense complex elements, [5] = ADOs
 * below the top tier, [6] = first top-tier packed element, [7] = top-tier
 * blocks of 256 elements, [8] = most non-empty ADOs one such block spans,
 * [9] = ADOs with no allowed element, [10] = most couplings of a top-tier
 * ADO, [11] = gather-map entries.  No counterpart in the reference. */
int heomq_model_packed_layout(const heomq_model* m, int64_t* info);
/* Replace physical parameters that leave the hierarchy shape unchanged
 * (energies, Gamma, T, bias).  Used by sweeps and finite differences. */
int heomq_model_set_params(heomq_model* m, const heomq_params* p);

/* ---------------------------------------------------------------------------
 * Drop-in loader from the reference's own module arrays (1-based Fortran
 * values, as built by the reference's preparation phase; module/matmod.f90).
 * Called once by the ISO_C_BINDING shim (fortran/heomq_iface.f90) with
 * c_loc of:  indextable, nfirst/nlast, ifirst (per tier), indexcoef (int64),
 * itypecoef/ifactfcoef/ifactrcoef (int8), index_coef_ref (int64),
 * indexdraw (position of each drawer in tier 1), cb/cd (complex,
 * (norbs,nspin,ncor,nalf,nsgn)), cgamma(nvar), dipm(nvar), ilead/mspin/morbs/
 * mpm (nvar), eleadinfty(nalf,nspin), hs(nrho,nrho), amsall(nrho,nrho,norbs,
 * nspin).  Only the default-flag subset plus lscale is accepted (psdjob, no
 * offcor, lscreen, lsimple, lfilter, ltrun_der, lhf, lhb): the shim rejects
 * the rest with HEOMQ_UNSUPPORTED as the reference would need them.
 * ntable and lall are the lengths of indextable and of the coupling stream
 * (indexcoef, itypecoef, ifactfcoef, ifactrcoef), common block info3 of
 * include/common:41-47 (indextable(ntable), main/main.f90:537;
 * indexcoef(lall), main/buildcoefindex.f90:608): every position the tables
 * name is checked against them, and one outside is HEOMQ_BAD_ARG.
 * dbsqrt: with the reference's lscale, its dbsqrt(norbs,nspin,ncor,nalf,nsgn)
 * (main/evaluatepara.f90:2931-2960; cb and cd are then the scaled ones it
 * holds), which scales every upper coupling; NULL without lscale.
 * ------------------------------------------------------------------------- */
int heomq_model_from_reference(
    int32_t nrho, int64_t nunk, int32_t ntier, int32_t norbs, int32_t nspin,
    int32_t nalf, int32_t ncor, int32_t igroundsteady, int64_t ntable, int64_t lall,
    const int32_t* indextable, const int64_t* nfirst, const int64_t* nlast,
    const int64_t* ifirst, const int64_t* indexcoef, const int8_t* itypecoef,
    const int8_t* ifactfcoef, const int8_t* ifactrcoef,
    const int64_t* index_coef_ref, const int32_t* indexdraw,
    const double* cb, const double* cd, const double* cgamma,
    const double* dipm, const int32_t* ilead, const int32_t* mspin,
    const int32_t* morbs, const int32_t* mpm, const double* eleadinfty,
    const double* hs, const double* amsall, const double* dbsqrt, heomq_model** out);

/* ---------------------------------------------------------------------------
 * Device context: uploads the tables once (heomq_ctx_load in the reference's
 * terms) and owns the solver workspace.  dtype: 0 = FP64, 1 = FP32 apply.
 * ------------------------------------------------------------------------- */
int heomq_ctx_create(const heomq_model* m, int device, heomq_ctx** out);
void heomq_ctx_free(heomq_ctx* c);
/* Re-upload coefficients after heomq_model_set_params (same shape). */
int heomq_ctx_refresh(heomq_ctx* c, const heomq_model* m);
/* Solver iterations are replayed from captured HIP graphs (one launch per
 * eight TFQMR iterations; solves requested on the null stream run on a
 * context-owned stream).  enable = 0 launches the same kernels one by one. */
int heomq_ctx_set_graphs(heomq_ctx* c, int enable);
/* Derivative terminator of the top tier ($jobinfo ltrun_der;
 * tfqmr_double/calc_ax_tfqmr_omp.f90:362-478 with fermion/prelude_trun.f90):
 * every top-tier ADO K receives the upper couplings of the tier-(L+1) ADOs
 * K + j, each set to the steady solution of its own equation without upper
 * couplings, (i [H, .] - gamma_{K+j})^-1 (its lower-coupling terms).  The
 * context then holds the packed tables of the hierarchy one tier higher
 * (same parameters; a model from heomq_model_from_reference is refused) and
 * its forward packed applies (heomq_apply_packed*, the preconditioned and
 * allocation-free ones) and the steady TFQMR solve (trace row RDO) include
 * the terminator; every other apply, solver and propagator of the context
 * returns HEOMQ_UNSUPPORTED while it is set.  enable 0 removes it.  The
 * steady state with the terminator equals, on tiers 0..L, the steady state of
 * the hierarchy one tier higher (a Schur complement). */
int heomq_ctx_set_trun_der(heomq_ctx* c, int enable);
/* Block-Jacobi-preconditioned steady solves with b = e_trace start with two
 * sweeps of the reference's block-Jacobi iteration (jacobi/solve_jacobi.f90,
 * its default shift; the trace is kept) applied to the initial guess
 * (enable = 1, the default); 0 starts TFQMR from the guess as given.  No
 * counterpart in the reference. */
int heomq_ctx_set_presmooth(heomq_ctx* c, int enable);
/* Block-Jacobi-preconditioned packed applies (the steady solve's iteration,
 * heomq_apply_packed_precond) run as one fused launch, the inverse applied to
 * each workgroup's ADOs before anything is written, when the forward apply
 * form has the sector-batched top tier and sliced heavy rows (enable = 1, the
 * default; the results are the same bits); 0 always runs the apply, then the
 * inverse.  The same switch fuses each Chebyshev term of the packed
 * propagators (the apply's output folded into the recurrence and the sum,
 * one launch, the same bits).  No counterpart in the reference. */
int heomq_ctx_set_precond_fusion(heomq_ctx* c, int enable);
/* Left preconditioning of the forward steady-state TFQMR on the packed
 * layout: 0 none; 1 Jacobi (diagonal) scaling, default right-hand side only;
 * 2 block Jacobi, (shift - S_K)^-1 per ADO with S_K the same-tier operator
 * (the sweep operator of jacobi/calc_ax_jacobi_omp.f90 used as M^-1); -1
 * automatic: 1 when H is diagonal in the occupation basis (the single-
 * impurity model), 2 otherwise.  The convergence test always uses the
 * unpreconditioned residual ||b - A x||.  No counterpart in the reference,
 * whose TFQMR is unpreconditioned. */
int heomq_ctx_set_precond(heomq_ctx* c, int mode);
/* Shift of the block-Jacobi preconditioner (> 0), or 0 for the default:
 * 0.05 x the block-Jacobi iteration's default ejac. */
int heomq_ctx_set_precond_shift(heomq_ctx* c, double shift);
/* Seed of TFQMR's +-1 shadow vector at the first start of every solve
 * (restarts take the following seeds); default 0.  The reference draws its
 * shadow vector with DRANDN, seed 1 (tfqmr_double/dutfx.f90:178). */
int heomq_ctx_set_shadow_seed(heomq_ctx* c, uint64_t seed);
/* Autotuner of the packed apply: times each of its forms (bit 0: ADOs below
 * the top tier with their couplings grouped by gather map, summed in the
 * neighbour's slot space before the map; bit 1: top-tier blocks batched by
 * symmetry sector; bit 2: one wave per ADO below the top tier, four per
 * workgroup, when every such ADO but the tier-0 one holds at most 16
 * elements; 0: one workgroup per ADO with the couplings sliced over the
 * lanes, top tier in packed-element order) over `reps` launches on this GPU,
 * for the forward and the transposed operator, and keeps the fastest.  The
 * forward forms are timed on the operator the steady solve iterates with:
 * where that is block-Jacobi preconditioned (the default when H couples Fock
 * states), the apply and its inverse, one fused launch in the forms that
 * allow it (heomq_ctx_set_precond_fusion).
 * forms[0..1] (may be NULL) receive the choices, times_us[0..15] (may be
 * NULL) the times, form f of operator w at [8 w + f] (forms the layout does
 * not take are left untouched).  Every form computes the same operator.  No
 * counterpart in the reference. */
int heomq_ctx_autotune(heomq_ctx* c, int reps, int32_t* forms, double* times_us);
/* Set the packed apply's form per operator (bits as heomq_ctx_autotune; a
 * form the layout does not take falls back per part), and read the forms in
 * use (forms[0] forward, forms[1] transposed).  No counterpart in the
 * reference. */
int heomq_ctx_set_apply_form(heomq_ctx* c, int forward, int transposed);
int heomq_ctx_get_apply_form(const heomq_ctx* c, int32_t* forms);
/* Number of complex elements N*d*d of one ADO vector. */
int64_t heomq_ctx_vector_len(const heomq_ctx* c);

/* y = L x on device buffers (FP64), the drop-in for
 * tfqmr_double/calc_ax_tfqmr_omp.f90 (trace = HEOMQ_TRACE_LAST) and
 * td/calcderiv_omp.f90 (HEOMQ_TRACE_NONE). */
int heomq_apply(heomq_ctx* c, const double* x, double* y, int trace, void* stream);
/* y = L^T x, the real transpose (Re<w, Lx> = Re<L^T w, x>); the drop-in for
 * bicg/calc_atrx_bicg_omp.f90 up to its unknown convention. */
int heomq_apply_transpose(heomq_ctx* c, const double* x, double* y, int trace, void* stream);
/* FP32 (complex64) apply, same layout with float. */
int heomq_apply_f32(heomq_ctx* c, const float* x, float* y, int trace, void* stream);
/* y = dL/dtheta x with L's coefficients replaced by their derivatives
 * (dH, dcb, dcd, dgamma per drawer; complex interleaved; dH d*d). */
int heomq_apply_dparam(heomq_ctx* c, const double* x, double* y, const double* dH,
                       const double* dcb, const double* dcd, const double* dgam,
                       void* stream);

/* Projected apply in the solvers' symmetry-packed layout (H and the bath
 * conserve N and S_z, so element (i, j) of ADO K is kept only if N_i - N_j
 * and 2(S_z,i - S_z,j) equal the charge / spin carried by K's drawers; the
 * tier-0 block is Hermitianised).  Dense in, dense out (off-sector elements of
 * y are zero); trace is HEOMQ_TRACE_RDO or HEOMQ_TRACE_NONE.  No counterpart
 * in the reference: it is the operator the device solvers iterate with. */
int heomq_apply_packed(heomq_ctx* c, const double* x, double* y, int trace, int transpose, void* stream);
/* The same projected operator on packed device vectors (ptotal complex
 * elements each, see heomq_packed_info), and the dense <-> packed maps. */
int heomq_apply_packed_raw(heomq_ctx* c, const double* xp, double* yp, int trace, int transpose, void* stream);
int heomq_apply_packed_raw_f32(heomq_ctx* c, const float* xp, float* yp, int trace, int transpose, void* stream);
/* yp = M^-1 (Pi A xp) on packed vectors, M^-1 the block-Jacobi inverse
 * (shift - S_K)^-1 per ADO with the preconditioner shift (tier-0 output made
 * Hermitian): the operator the block-Jacobi-preconditioned TFQMR iterates
 * with.  form 1: apply, then the one-kernel inverse the solver uses (blocks
 * of whole ADOs when d <= 16 and H couples groups of <= 4 states, else the
 * LDS-window kernel; HEOMQ_UNSUPPORTED when neither fits the layout); 2:
 * apply, then the two-kernel transform; 3: apply, then the LDS-window kernel
 * (when every ADO range a 256-element block touches fits its LDS); 4: one
 * fused launch, the inverse applied in each apply workgroup to its whole
 * ADOs (apply forms 2 and 6 without the wave form, d <= 16, the ADO-aligned
 * inverse; the same bits as form 1), HEOMQ_UNSUPPORTED otherwise.  All forms
 * compute the same values.  No counterpart in the reference. */
int heomq_apply_packed_precond(heomq_ctx* c, const double* xp, double* yp, int trace, int form, void* stream);
/* The same operator as heomq_apply_packed_precond form 1, with no allocation:
 * A x is staged in a static device array of 2^21 complex FP64 elements
 * (32 MiB per GPU, held until the process ends) instead of a context
 * workspace; a packed vector longer than that is HEOMQ_BAD_DIM.  Calls that
 * overlap on one GPU share the array and race.  No counterpart in the
 * reference: for callers that cannot allocate (HIP graph capture, serving
 * loops). */
int heomq_apply_packed_precond_noalloc(heomq_ctx* c, const double* xp, double* yp, int trace, void* stream);
int heomq_pack(heomq_ctx* c, const double* dense, double* packed, void* stream);
int heomq_unpack(heomq_ctx* c, const double* packed, double* dense, void* stream);
/* info[0] = packed complex elements, [1] = sectors, [2] = N conserved,
 * [3] = S_z conserved, [4] = dense complex elements N*d*d. */
int heomq_packed_info(const heomq_ctx* c, int64_t* info);

/* Host-pointer apply: the calc_ax_tfqmr_omp(dim0, rcgtmp, rcgrhs) drop-in
 * (copies in, applies, copies out).  dim0 must equal d. */
int heomq_calc_ax_tfqmr(heomq_ctx* c, int32_t dim0, const double* x_host,
                        double* y_host);
/* Host-pointer transpose (calc_atrx_bicg_omp(dim0, rcgtmp, rcgrhs) position). */
int heomq_calc_atrx(heomq_ctx* c, int32_t dim0, const double* x_host, double* y_host);
/* BiCG convention on device buffers: y = A_B x, the operator of
 * bicg/calc_ax_bicg_omp.f90 (x's tier 0 is rebuilt in place from its lower
 * triangle, y(1,1) of tier 0 = tr rho0, y's tier-0 upper triangle zero and
 * diagonal real), and z = A_B^T w, the operator of bicg/calc_atrx_bicg_omp.f90
 * (w's tier 0 is reduced in place to its lower triangle with a real diagonal).
 * Both inputs are modified as the reference's are (intent inout). */
int heomq_apply_bicg(heomq_ctx* c, double* x, double* y, void* stream);
int heomq_apply_transpose_bicg(heomq_ctx* c, double* w, double* z, void* stream);
/* Host-pointer drop-ins: calc_ax_bicg_omp(dim0, rcgtmp, rcgrhs) and
 * calc_atrx_bicg_omp(dim0, rcgtmp, rcgrhs) (calc_ax_bicg_omp.f90:16-17,
 * calc_atrx_bicg_omp.f90:35-36: both arguments inout). */
int heomq_calc_ax_bicg(heomq_ctx* c, int32_t dim0, double* x_host, double* y_host);
int heomq_calc_atrx_bicg(heomq_ctx* c, int32_t dim0, double* x_host, double* y_host);
/* calcderiv_omp(tt): rhorhs = L(rhotmp) for fixed H (host pointers). */
int heomq_calcderiv(heomq_ctx* c, const double* rhotmp_host, double* rhorhs_host);

/* ---------------------------------------------------------------------------
 * Steady state: solve A x = e_trace on the device with TFQMR (Freund 1993),
 * the solver-level drop-in for tfqmr_double/solve_tfqmr.f90 + dutfx.f90.
 * x (device, in/out): initial guess in, solution out.  Converged when
 * ||b - A x||_2 <= tol * ||b||_2 on the (Re, Im) vector.  transpose = 1 solves
 * with A^T (the adjoint solve).  b may be NULL (then b = e_trace).
 * stats (HEOMQ_STATS_LEN): [0] iterations, [1] applies, [2] final relative
 * residual, [3] seconds, [4] the left preconditioner the iteration ended
 * with (0 none, 1 Jacobi scaling, 2 block Jacobi), [5] length of the real
 * vector iterated on, [6] 1 when block-Jacobi iterations ran with their
 * vector work carried by the fused preconditioned launches (FP64 forward
 * solve, fusable apply form, fusion on: heomq_ctx_set_precond_fusion; the
 * same iterates in exact arithmetic, reductions in another order), [7] 0.
 * Returns HEOMQ_NOT_CONVERGED if maxiter hit.
 * ------------------------------------------------------------------------- */
int heomq_solve_steady(heomq_ctx* c, double* x, const double* b, double tol,
                       int maxiter, int trace, int transpose, double* stats,
                       void* stream);
/* Mixed precision: FP32 TFQMR inner solves (tolerance inner_tol) inside FP64
 * iterative refinement; residuals, updates, dots and norms in FP64.
 * stats (HEOMQ_STATS_MIXED_LEN): [0] FP32 inner iterations, [1] applies,
 * [2] final relative residual, [3] seconds, [4] refinement steps, [5] 0,
 * [6] iterations of the FP64 finish, [7] 0, [8..47] relative FP64 residual
 * at the start of refinement steps 0..39, [48..63] inner iterations of
 * refinement steps 0..15 (0 past the steps taken). */
int heomq_solve_steady_mixed(heomq_ctx* c, double* x, const double* b, double tol,
                             double inner_tol, int maxiter, int trace,
                             double* stats, void* stream);
/* Steady state by BiCG in the reference's convention (bicg/solvebicg.f90):
 * x (device, dense, in/out) holds the initial guess, b = e_(1,1) of tier 0,
 * stop when ||b - A_B x||_2 <= crit (absolute, as solvebicg.f90:124-126);
 * x's rho0 is returned Hermitian.  stats (HEOMQ_STATS_LEN) as
 * heomq_solve_steady, [2] the absolute residual, [4] 1 when the iteration ran
 * on the symmetry-packed layout (an x with nothing outside the sectors; the
 * iterates are the dense ones to rounding), 0 on the dense one. */
int heomq_solve_bicg(heomq_ctx* c, double* x, double crit, int maxiter, double* stats, void* stream);
/* The same on a host array (the solvebicg drop-in of the Fortran layer). */
int heomq_solve_bicg_host(heomq_ctx* c, double* rho_host, double crit, int maxiter, double* stats);
/*
 * Time-dependent lead energy shift of the propagators (main/getengyshift.f90,
 * td/calcderiv_omp.f90:51-61,144-148): gamma_K(t) = sum_k cgamma_k +
 * i dipm_k eshift(alpha_k, s_k; t) with eshift = amps (1 - exp(-t/tchar))
 * (fieldtype 0) or amps sin(2 pi t / tchar) + amps_dc (fieldtype 1), zero
 * for t >= t_off; amps are the model's lead shifts (engyshift).  Needs a
 * model built with steady = 0 (no static bias in gamma).  tchar, amps_dc:
 * nalf * nspin entries, [alpha * nspin + s], fs and eV (amps_dc may be NULL);
 * fieldtype -1 removes the field.  Pulsed shifts (fieldtype 2) are
 * HEOMQ_UNSUPPORTED.
 */
int heomq_ctx_set_field(heomq_ctx* c, int fieldtype, const double* tchar, const double* amps_dc, double t_off);
/*
 * Time-dependent system Hamiltonian of the propagators (main/calcdhs.f90;
 * td/calcderiv_omp.f90:60-68 uses hs + dhs(tt)): dhs(t) = sum_k f_k(t) D_k
 * with nterm <= 16 fixed Hermitian operators D_k (ops: nterm * d * d complex,
 * column-major, the energy unit eV) and f_k from calcdhs's options, kinds[k]
 * with pars[4k..4k+3] (p0 in eV, times in fs, angular frequencies in 1/fs):
 * 0 p0 (1 - e^(-t/p1)); 1 p0 sin(p1 t); 2 p0 cos(p1 t); 3 p0 on p1 < t < p2;
 * 4 p0 on p1 <= t < p2; 5 p0 + p1 cos(p2 t + p3); 6 p0 for t < p1; 7 the
 * Hubbard dot field p0 e^(-p1 t) | -p0 e^(-p1 (t - p2)) | p0 e^(-p1 (t - p3))
 * split at p2 and p3; 8 p0 for t > p1.  nterm = 0 removes it.  The options
 * that feed back the state (pha) or project with cproj are not covered.
 * Non-Hermitian operators are HEOMQ_BAD_ARG.
 */
int heomq_ctx_set_system_field(heomq_ctx* c, int nterm, const double* ops, const int32_t* kinds, const double* pars);
/* eshift(alpha, s; tt) of the lead field in eV, nalf * nspin values
 * [alpha * nspin + s] (main/getengyshift.f90 as main/main.f90:1453-1458
 * prints it); zeros when no field is set.  Host only. */
int heomq_ctx_field_shifts(const heomq_ctx* c, double tt, double* eshift_ev);
/* y = L(tt) x (dense device vectors), the derivative calcderiv_omp(tt) forms. */
int heomq_calcderiv_at(heomq_ctx* c, double tt, const double* x, double* y, void* stream);
/* calcderiv_omp(tt) on host arrays (the Fortran drop-in's backend). */
int heomq_calcderiv_tt(heomq_ctx* c, double tt, const double* rhotmp_host, double* rhorhs_host);
/* RK4 of td/intstp.f90 from time t0 (the derivative at t, t + dt/2 twice, t + dt). */
int heomq_rk4_at(heomq_ctx* c, double* x, double t0, double dt, int nstep, void* stream);

/*
 * Sparse-storage drop-in (tfqmr_double/calc_ax_tfqmr_spa_omp.f90): ADOs
 * stored as COO lists over a fixed pattern (module sparsemod, built by
 * fermion/prelude_spa.f90:532-749).  y = P A P^T x with P the pattern's
 * selection, and y(lunk_spa + 1) = tr rho0 (:1061-1069).
 */
/* Build the pattern as fermion/prelude_spa.f90:532-749 does (replaces its
 * tier loop; the reference writes the result into module sparsemod): tier 0
 * holds the nonzeros of exp(a H) (a = 0.31415926, 10 Taylor terms,
 * utility/calc_exp_hmat.f90); an ADO of tier n >= 1 holds the elements above
 * dsmall of sum_J m1 E (op X_J) + m2 E (X_J op) over its lower neighbours J
 * (random complex m1, m2 drawn from `seed`: they only keep the two products
 * from cancelling).  dsmall: the reference's dsmall_spa (1e-32, or 1e-12 with
 * lproj).  Two calls: with irow_spa / icol_spa NULL, or cap < the pattern's
 * length, only *lunk_spa (and nnz_spa / ind_spa when given, N each) are set
 * and HEOMQ_BAD_DIM is returned when arrays were given but too short; then
 * with arrays of lunk_spa entries.  ind_spa(a) = ind_spa(a-1) + nnz_spa(a-1),
 * 1-based, as :741-742. */
int heomq_spa_build_pattern(const heomq_model* m, double dsmall, uint64_t seed, int64_t cap, int64_t* lunk_spa,
                            int32_t* nnz_spa, int64_t* ind_spa, int32_t* irow_spa, int32_t* icol_spa);
/* Set the pattern from the reference's arrays: nnz_spa and ind_spa (1-based
 * start in rho_spa) per ADO (N each), irow_spa / icol_spa (1-based, lunk_spa
 * each); ADOs are stored in order (ind_spa(a) = ind_spa(a-1) + nnz_spa(a-1)
 * for every non-empty ADO, as prelude_spa.f90:741-742 lays them out).
 * Rejected with HEOMQ_BAD_ARG: indices out of range, overlapping or
 * out-of-order ADOs, an element listed twice, elements of no ADO.  The apply
 * then runs on rho_spa as it is: device tables scale with lunk_spa and the
 * number of distinct per-ADO element lists (at most 65535). */
int heomq_spa_pattern(heomq_ctx* c, int64_t lunk_spa, const int32_t* nnz_spa, const int64_t* ind_spa,
                      const int32_t* irow_spa, const int32_t* icol_spa);
/* The steady state on the registered pattern (tfqmr_double/solve_tfqmr.f90
 * with lsparse = T, which iterates calc_ax_tfqmr_spa_omp): TFQMR on P A P^T
 * over x_spa (lunk_spa complex, in: a trace-one guess, out: the state), the
 * trace row on element (d, d) of tier 0 (the pattern must hold it), tier 0
 * kept Hermitian.  Preconditioned as heomq_solve_steady (block Jacobi needs
 * every ADO's list closed under H's groups; else HEOMQ_UNSUPPORTED unless
 * heomq_ctx_set_precond chose none or Jacobi scaling).  stats
 * (HEOMQ_STATS_LEN): [0] iterations, [1] applies, [2] relative residual,
 * [3] seconds, [4] preconditioner used (0 none, 1 Jacobi scaling, 2 block
 * Jacobi), [5] reals per vector. */
int heomq_spa_solve_steady(heomq_ctx* c, double* x_spa, double tol, int maxiter, double* stats, void* stream);
/* Device vectors: x_spa (lunk_spa complex), y_spa (lunk_spa + 1 complex). */
int heomq_apply_spa(heomq_ctx* c, const double* x_spa, double* y_spa, void* stream);
/* calc_ax_tfqmr_spa_omp(ldim, rcgtmp, rcgrhs) on host arrays; ldim /= lunk_spa
 * is HEOMQ_BAD_DIM, as calc_ax_tfqmr_spa_omp.f90:61-65 stops. */
int heomq_calc_ax_tfqmr_spa(heomq_ctx* c, int64_t ldim, const double* x_host, double* y_host);
/* The sparse propagators' derivative (td/calcderiv_spa_omp.f90:1-11,84-98):
 * y_spa = P L(tt) P^T x_spa, lunk_spa complex each (no trace element), with
 * the lead field of heomq_ctx_set_field at tt.  Device vectors. */
int heomq_calcderiv_spa_at(heomq_ctx* c, double tt, const double* x_spa, double* y_spa, void* stream);
/* calcderiv_spa_omp(tt) on host arrays (rhotmp_spa -> rhorhs_spa of module
 * sparsemod; the Fortran drop-in's backend). */
int heomq_calcderiv_spa_tt(heomq_ctx* c, double tt, const double* x_host, double* y_host);

/*
 * Block-Jacobi iteration (jacobi/solve_jacobi.f90, calc_ax_jacobi_omp.f90,
 * prelude_jacobi.f90; J. Chem. Phys. 147, 044105 (2017), Eq. 5).  One sweep
 * replaces every ADO K by (ejac - S_K)^-1 (ejac x_K + C_K(x)), S_K X =
 * -i[H, X] + gamma_K X the same-tier operator and C_K the tier couplings,
 * computed on the symmetry-packed layout with S_K diagonalised through the
 * eigenvectors of H.  ejac <= 0 selects the reference's default shift
 * (prelude_jacobi.f90:84-113): max(spread of the nonzero eigenvalues of
 * [H, .], sqrt((ntier - 1) * max |cb|)).
 */
/* info[0] default ejac, [1] eigenvalue spread of [H, .], [2] sqrt((ntier-1)
 * max|cb|), [3] nonzero eigenvalues of [H, .] (nnzjac), [4] largest group of
 * Fock states H couples.  evals (d) and evecs (d*d complex, column-major) may
 * be NULL.  Host only: needs no GPU.  Replaces prelude_jacobi.f90:43-113. */
int heomq_model_jacobi_info(const heomq_model* m, double* info, double* evals, double* evecs);
/* The same five numbers for a context's current parameters. */
int heomq_jacobi_info(const heomq_ctx* c, double* info);
/* y = one sweep of x (dense device vectors, N*d*d complex each): the operator
 * calc_ax_jacobi_omp(dim0, rcgtmp, rcgrhs) applies (calc_ax_jacobi_omp.f90:1-324). */
int heomq_jacobi_step(heomq_ctx* c, const double* x, double* y, double ejac, void* stream);
/* calc_ax_jacobi_omp on host arrays (the Fortran drop-in), default ejac. */
int heomq_calc_ax_jacobi(heomq_ctx* c, int32_t dim0, const double* x_host, double* y_host);
/* solve_jacobi on the device: rho0 of x is made Hermitian from its lower
 * triangle and must have unit trace (else HEOMQ_BAD_ARG, as
 * solve_jacobi.f90:70-85 stops); sweeps until ||A x - e||_2 (the
 * calc_ax_tfqmr_omp trace row, last ADO), checked after sweeps 1, 51, 101, ...
 * (:183-229), is <= crit or maxiter sweeps are done.  x returns the last
 * iterate.  stats (HEOMQ_STATS_LEN): sweeps, applies, final residual, seconds, ejac used,
 * smallest residual seen, [6..7] 0. */
int heomq_solve_jacobi(heomq_ctx* c, double* x, double crit, int maxiter, double ejac, double* stats,
                       void* stream);
/* The same on a host array (solve_jacobi_gpu of the Fortran layer). */
int heomq_solve_jacobi_host(heomq_ctx* c, double* rho_host, double crit, int maxiter, double* stats);
/* Host-pointer steady state (the solve_tfqmr drop-in of the Fortran shim):
 * rho_host (d, d, N) complex holds the initial guess on entry (a trace-one
 * tier-0 block, Hermitian as tfqmr_double/solve_tfqmr.f90:97-103 makes it)
 * and the solution on exit;
 * solved with the tier-0 trace row on the symmetry-packed layout.  stats
 * (HEOMQ_STATS_LEN) as heomq_solve_steady. */
int heomq_solve_steady_host(heomq_ctx* c, double* rho_host, double tol, int maxiter,
                            double* stats);

/* ---------------------------------------------------------------------------
 * Batched steady states.  The reference solves one parameter point per run
 * (a shell loop over generated inputs, tools/make_input_didv/makeinput.f90);
 * here nprob problems of one hierarchy shape (same norbs, nalf, nmats, ntop,
 * lsdraw, Hamiltonian kind; energies, Gamma, T and bias free) share every
 * structural table and are solved together, one launch per solver step for
 * the whole batch.  No counterpart in the reference.
 * ------------------------------------------------------------------------- */
typedef struct heomq_batch heomq_batch;
int heomq_batch_create(const heomq_params* params, int nprob, int device, heomq_batch** out);
void heomq_batch_free(heomq_batch* b);
/* info[0] = problems, [1] = ADOs per problem, [2] = packed elements per
 * problem, [3] = d, [4] = commutator row width. */
int heomq_batch_info(const heomq_batch* b, int64_t* info);
/* y_p = Pi A_p x_p for every problem at once; problem p's packed vector starts
 * at complex element p * info[2]. */
int heomq_batch_apply_packed(heomq_batch* b, const double* xp, double* yp, int trace, void* stream);
/* TFQMR steady state of every problem (tier-0 trace row, projected operator,
 * trace-one start), to ||b - A x|| <= tol ||b||.  Host outputs: rdo (nprob *
 * d * d complex, column-major tier-0 blocks), current (nprob * nalf * nspin,
 * td/calcj.f90 units), occ (nprob * norbs * nspin), stats (nprob * 4:
 * iterations, relative residual, status, restarts).  Any output may be NULL. */
int heomq_batch_solve(heomq_batch* b, double tol, int maxiter, double* rdo, double* current, double* occ,
                      double* stats, void* stream);

/* ---------------------------------------------------------------------------
 * Sweeps: nprob problems of up to 16 hierarchy shapes (Pade terms, top tier;
 * the same orbitals, leads and d for all) solved together with ONE launch per
 * solver step (the apply: one launch per 10 shapes).  The problems are
 * grouped by shape (each shape's structural tables shared, as a batch), and
 * every packed apply and every TFQMR vector kernel runs over a flat grid that
 * covers every problem of every shape, largest shape first; each shape's
 * apply form is the one heomq_ctx_autotune picks for it.  Each problem keeps the blocks-per-problem count of a solve
 * on its own, so its reductions fold the same partial sums in the same order.
 * The reference's counterpart is one process per point
 * (tools/make_input_ivcurve/ivcurve.csh); no entry point of its own.
 * ------------------------------------------------------------------------- */
typedef struct heomq_sweep heomq_sweep;
int heomq_sweep_create(const heomq_params* params, int nprob, int device, heomq_sweep** out);
void heomq_sweep_free(heomq_sweep* w);
/* info[0] = problems, [1] = shapes, [2] = packed complex elements of all
 * problems, [3] = d, [4] = workgroups of one apply launch, [5] = workgroups of
 * one vector-kernel launch. */
int heomq_sweep_info(const heomq_sweep* w, int64_t* info);
/* Per problem (caller's order; any output may be NULL): its shape (shapes
 * numbered largest first), the complex offset and length of its packed vector
 * in the sweep's flat layout. */
int heomq_sweep_problems(const heomq_sweep* w, int32_t* shape, int64_t* off, int64_t* len);
/* y = Pi A x for every problem at once on flat-layout device vectors (trace
 * row HEOMQ_TRACE_RDO), on `stream`; one launch for up to 10 shapes. */
int heomq_sweep_apply_packed(heomq_sweep* w, const double* xp, double* yp, int trace, void* stream);
/* TFQMR steady state of every problem, as heomq_batch_solve; outputs in the
 * caller's problem order. */
int heomq_sweep_solve(heomq_sweep* w, double tol, int maxiter, double* rdo, double* current, double* occ,
                      double* stats, void* stream);

/* ---------------------------------------------------------------------------
 * Propagation (td/intstp.f90 RK4 and td/chebyshev.f90 Chebyshev), fixed H.
 * x (device) is advanced in place by nstep steps of dt.  Chebyshev's stats
 * (HEOMQ_STATS_CHEBY_LEN, may be NULL): [0] terms per step, [1] delta, [2] 1 when
 * the steps ran on the packed layout, 0 on the dense one.
 * A dense x with no nonzero element outside the symmetry-packed sectors (every
 * physical state) is propagated on the packed layout and unpacked (heomq_rk4_at
 * likewise, unless a time-dependent system Hamiltonian is set); any other x
 * takes the dense apply.
 * ------------------------------------------------------------------------- */
int heomq_rk4(heomq_ctx* c, double* x, double dt, int nstep, void* stream);
int heomq_chebyshev(heomq_ctx* c, double* x, double dt, int nstep, double* stats,
                    void* stream);
/* The same on a symmetry-packed vector xp (heomq_packed_info()[0] complex
 * elements, device), with the operator A restricted to the sectors (no trace
 * row, no projection): td/intstp.f90:20-100 from time t0 (lead-field rates
 * gamma_K(t) at every stage), td/chebyshev.f90:126-200.  HEOMQ_UNSUPPORTED
 * when a time-dependent system Hamiltonian is set. */
int heomq_rk4_packed(heomq_ctx* c, double* xp, double t0, double dt, int nstep, void* stream);
int heomq_chebyshev_packed(heomq_ctx* c, double* xp, double dt, int nstep, double* stats,
                           void* stream);

/* Observables from a device state: I[alpha*nspin + s] (td/calcj.f90, times
 * jconst) and occupation occ[m*nspin + s] (main/calcocc.f90). */
int heomq_observables(heomq_ctx* c, const double* x, double* current, double* occ,
                      void* stream);

#ifdef __cplusplus
}
#endif
#endif /* HEOMQ_H */
