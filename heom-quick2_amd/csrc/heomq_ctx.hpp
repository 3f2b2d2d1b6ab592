// Internal definitions shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "heomq.h"
#include "heomq_kernels.hpp"
#include "heomq_model.hpp"
#include "heomq_solver.hpp"

struct heomq_model {
  heomq::Model m;
};

#define HIPCHECK(expr)                                                                \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return heomq::fail(HEOMQ_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    release();
    bytes = std::max<size_t>(b, 16);
    return hipMalloc(&p, bytes);
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

template <typename T>
inline hipError_t upload(DevBuf& b, const std::vector<T>& v) {
  hipError_t e = b.alloc(v.size() * sizeof(T));
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

inline int op_low(const heomq::Model& m, int k) {   // lower-tier operator: c^+ for sigma = +, c for sigma = -
  return 2 * (m.orb[k] * m.nspin + m.spin[k]) + (m.sgn[k] == 0 ? 1 : 0);
}
inline int op_up(const heomq::Model& m, int k) {    // upper-tier operator: c for sigma = +, c^+ for sigma = -
  return 2 * (m.orb[k] * m.nspin + m.spin[k]) + (m.sgn[k] == 0 ? 0 : 1);
}

struct OperatorTables {
  DevBuf rowptr, ent, gam;
  heomq::DevOperator op;
};

// Sparse-storage operator (calc_ax_tfqmr_spa_omp): the packed apply's tables
// over "sectors" that are the distinct per-ADO element lists of a COO pattern,
// so y = P A P^T x runs directly on rho_spa (no dense scratch).
struct SpaOperator {
  DevBuf off, sec, elems, pos, block, lblock, lnz, ent, maps, hgoff, hg, ltblk, ltados, ltoff, ltent;
  DevBuf hrow, hq, hc, hq_ell, hc_ell;
  DevBuf jblk, jdesc, ju, je;        // block-Jacobi tables of the pattern's lists (steady solves)
  heomq::DevPacked op;               // structural fields; coefficients come from the context's operator
  heomq::DevJacobi jac;
  bool has_jac = false;              // every list closed under H's groups, every ADO <= 256 elements
  int trace_slot = -1;               // slot of element (d, d) in ADO 0's list, -1 if absent
  int nsec = 0;
  bool built = false;
  std::vector<int16_t> elems_h, pos_h;   // host copies (H tables are rebuilt when H changes)
  std::vector<int> nnz_h;
};

// Derivative terminator of the top tier (ltrun_der): the packed tables of the
// hierarchy one tier higher.  Its tiers 0..L are the context's own ADOs at
// the same packed offsets; its top tier L + 1 holds the ADOs the terminator
// eliminates in every apply (their steady equation without upper couplings).
struct TrunOperator {
  DevBuf off, sec, elems, pos, block, lblock, lnz, rowptr, ent, maps, hgoff, hg, ltblk, ltados, ltoff, ltent;
  DevBuf gam, hrow, hq, hc, hq_ell, hc_ell, jblk, jdesc, ju, je;
  DevBuf xext, text;                 // the extended iterate, the eliminated tier's right-hand sides
  heomq::DevPacked op;               // every table of the extended hierarchy (records: the context's)
  heomq::DevJacobi jac;              // (-S_V)^-1 on the eliminated tier (shift 0)
  int64_t total = 0, vis = 0;        // packed elements of the extended / the context's hierarchy
  bool built = false;
};

struct heomq_ctx {
  int device = 0;
  heomq::Model model;                // the model the tables were built from (sparse layouts are built later)
  int d = 0, M = 0, L = 0, nspin = 2, norbs = 1, nalf = 2;
  int64_t N = 0, n1first = 1, n1count = 0;
  std::vector<int64_t> first;
  std::vector<int> sgn, spin, orb, alf;
  // tier-1 lookup of drawers with sign + (for observables): ADO index
  std::vector<int64_t> tier1_of;     // per drawer: ADO index or -1
  std::vector<int> tier1_dag;        // 1 if stored as its Hermitian partner
  std::vector<int> tier1_fact;
  std::vector<double> ams;           // host copy for observables
  std::vector<double> dbsqrt;        // lscale: |cb_k|^(1/2) (the stored tier-1 ADO is rho_k / dbsqrt_k); else 1
  bool lscale = false;
  std::vector<std::vector<int32_t>> tiers;   // drawer tuples (for per-ADO rates)
  // shared device tables
  DevBuf rec, rec_d, opmap, hnnz, hcol, hval, hnnz_d, hcol_d, hval_d;
  OperatorTables fwd, trn, dpar;
  // host-side entry lists (kept for re-upload of coefficients)
  std::vector<int32_t> lower_drawers;
  // solver workspace
  DevBuf ws64, ws64b, ws32, ws_io, scal, partial, red;
  DevBuf partial_apply;     // per-workgroup sums of the fused TFQMR iteration's operator launches
  DevBuf scal2, partial2;   // TFQMR's mid-iteration scalar copy and second partial-sum buffer
  DevBuf prop_flag;                  // off-sector test of a dense state (propagators)
  DevBuf ws_host;                    // dense state for the host-pointer solve
  // host copies of the sector tables (the packed H table is rebuilt from them)
  std::vector<int16_t> sec_elems_h, sec_pos_h;
  std::vector<int> sec_nnz_h, fock_n, fock_s;
  std::vector<int64_t> poff_h;       // packed offset of each ADO (N + 1)
  std::vector<uint16_t> psec_h;      // sector of each ADO
  DevBuf parity;                     // electron-number parity of each Fock state
  // symmetry-packed layout (solver path)
  DevBuf p_off, p_sec, p_elems, p_pos, p_block, p_lblock, p_entf, p_entt, p_hrow, p_hq, p_hc, p_maps;
  DevBuf p_hq_ell, p_hc_ell, p_lnz;
  DevBuf p_hgoff_f, p_hgoff_t, p_hg_f, p_hg_t, p_ltblk, p_ltados, p_ltoff, p_ltent_f, p_ltent_t;   // restructured apply
  heomq::DevPacked pfwd, ptrn;
  int apply_form[2] = {2, 2};        // packed apply form of the forward / transposed operator
  bool heavy_wave_ok = false;        // every heavy ADO but ADO 0 holds at most kHeavyWaveMaxNnz elements
  int64_t ptotal = 0;
  int nsec = 0;
  bool cons_n = true, cons_s = true;
  void* pinned = nullptr;            // small pinned host buffer for scalar read-back
  double* scratch_dev = nullptr;     // this device's address of the static scratch (no-alloc entry point)
  heomq::TierBounds tb{};
  // Solves asked for on the null stream run on this non-blocking stream (the
  // null stream cannot be captured into a graph).
  hipStream_t solve_stream = nullptr;
  // Captured TFQMR iteration chunks.  Every pointer a chunk's kernels take
  // follows from the key (iterate, right-hand side, workspace base, length)
  // and the context's tables, so a key match means the graph is valid; the
  // cache is dropped whenever the tables are re-uploaded.
  struct ChunkGraph {
    const void* x;
    const void* b;
    const void* work;
    int64_t n, tpos;
    int mode, trans, bytes, stream_id;
    hipStream_t stream;
    hipGraphExec_t exec;
  };
  std::vector<ChunkGraph> graphs;
  bool use_graphs = true;            // false: launch the chunk's kernels one by one
  // block-Jacobi-preconditioned applies as one fused launch where the apply
  // form allows it (else the apply, then the inverse)
  bool fuse_precond = true;
  // left preconditioning of forward steady-state solves: -1 automatic (Jacobi
  // scaling when H is diagonal in the occupation basis, block Jacobi
  // otherwise), 0 none, 1 Jacobi scaling, 2 block Jacobi
  int precond = -1;
  bool h_diagonal = false;
  // block-Jacobi iteration (jacobi/): eigensystem of H per conserved group of
  // Fock states, the default shift, per-ADO rates on the host (for the shift's
  // degeneracy check) and the iteration's own workspace
  DevBuf j_u, j_e, j_grp, j_gn, ws_jac;
  DevBuf j_crow, j_cslot, j_ca, j_cc, j_rrow, j_rslot, j_rb, j_rd, j_tde;   // slot-space list tables
  DevBuf j_blk, j_desc;                                                        // ADO-aligned form
  // time-dependent lead energy shift eshift(alpha, s; t) of the propagators
  // (main/getengyshift.f90): field_type -1 none, 0 amps (1 - exp(-t / tchar)),
  // 1 amps sin(2 pi t / tchar) + amps_dc; zero for t >= t_off.  Amplitudes
  // are the model's lead shifts (1/fs); td_cnt[a][alpha*nspin + s] sums
  // dipm over ADO a's drawers of lead alpha and spin s; td_gam holds
  // gamma_K(t) for the apply
  int field_type = -1;
  std::vector<double> field_amp, field_tchar, field_dc;
  double field_toff = 1e99;
  bool steady_model = true;
  DevBuf td_cnt, td_gam;
  // time-dependent system Hamiltonian of the propagators (main/calcdhs.f90):
  // H(t) = hs + sum_k f_k(t) D_k with fixed Hermitian D_k (eV) and the time
  // functions of calcdhs's options (sys_kind, 4 parameters each); the row
  // pattern of H(t) is the union of hs's and the D_k's (sys_hnnz, sys_hcol,
  // d slots per row), sys_base holds hs and sys_terms the D_k / hbar on it,
  // sys_hval the values at the last evaluated t
  int presmooth = 1;        // block-Jacobi sweeps before preconditioned steady solves (heomq_ctx_set_presmooth)
  int sys_nterm = 0;
  std::vector<int> sys_kind;
  std::vector<double> sys_par;
  DevBuf sys_hnnz, sys_hcol, sys_base, sys_terms, sys_hval;
  // sparse-storage drop-in (calc_ax_tfqmr_spa_omp): COO element -> dense
  // index, positions of the tier-0 diagonal, dense scratch; spa_lunk < 0: no pattern
  DevBuf spa_map, spa_diag, spa_dense, spa_io;
  int64_t spa_lunk = -1;
  int spa_ndiag = 0;
  SpaOperator spa;
  TrunOperator trun;
  heomq::DevJacobi pjac;
  double ejac_default = 0.0, ejac_gap = 0.0, ejac_cb = 0.0;
  // shift of the block-Jacobi preconditioner of TFQMR: (shift - S_K)^-1 per
  // ADO; prec_shift_user 0 selects kPrecShiftFactor * ejac_default
  double prec_shift = 0.0, prec_shift_user = 0.0;
  // seed of TFQMR's hashed +-1 shadow vector at a solve's first start (the
  // reference draws it from DRANDN with seed 1, tfqmr_double/dutfx.f90:178)
  uint64_t shadow_seed = 0;
  int jac_nnz = 0;
  std::vector<double> gam_h;
  void drop_graphs() {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
  }
  ~heomq_ctx() {
    drop_graphs();
    if (solve_stream) (void)hipStreamDestroy(solve_stream);
    if (pinned) (void)hipHostFree(pinned);
  }
};

// Currents I[alpha*nspin + s] and occupations from the dense tier-0/1 blocks
// h (complex interleaved, first[2] * d * d elements), host side.
void observables_host(const heomq_ctx* c, const std::vector<double>& h, double* current, double* occ);

namespace heomq {
// Coefficient records of the forward and transposed operators (6 M of them)
// and the per-ADO diagonal rates gamma_K (conjugated for the transpose).
void build_records(const Model& m, std::vector<DevRec>& rec);
void build_gamma(const Model& m, const std::vector<cplx>& gdraw, std::vector<double>& gam, bool conj);
// Eigensystem of H for the block-Jacobi iteration (uploaded into c->pjac) and
// the default shift of jacobi/prelude_jacobi.f90; run whenever H or the bath
// coefficients change.
int build_jacobi(heomq_ctx* c, const Model& m);
int jacobi_presmooth(heomq_ctx* c, double* x, int nsweep, double* r, double* t, hipStream_t st);
// block-Jacobi sweeps (default shift) applied to the start of a
// block-Jacobi-preconditioned steady solve with b = e_trace
constexpr int kPresmoothSweeps = 2;
// default block-Jacobi preconditioner shift, as a multiple of the iteration's ejac
constexpr double kPrecShiftFactor = 0.05;
// the static device scratch of heomq_apply_packed_precond_noalloc: complex FP64
// elements (32 MiB) and the devices whose addresses are cached
constexpr int64_t kNoAllocElems = int64_t(1) << 21;
constexpr int kMaxDevices = 64;
int scratch_address(int device, double** out);
// The sparse-storage operator of a validated COO pattern: start[a] is ADO a's
// first element (0-based) in rho_spa, irow / icol 1-based per element.
int build_spa_operator(heomq_ctx* c, const int32_t* nnz_spa, const std::vector<int64_t>& start,
                       const int32_t* irow, const int32_t* icol);
// Its H tables again (after H changed).
int build_spa_h(heomq_ctx* c);
// The derivative terminator's tables from the context's model (built from
// parameters: the hierarchy one tier higher is built the same way).
int build_trun(heomq_ctx* c);
// The same-tier inverse (-S_K)^-1 (shift 0) of the ADOs [a0, N) of a packed
// layout (sectors S, elems, pos, nnz; offsets off), ADO-aligned blocks; the
// eigensystem comes from m.H.
int build_tier_jacobi(const Model& m, int S, const std::vector<int16_t>& elems, const std::vector<int16_t>& pos,
                      const std::vector<int>& nnz, const std::vector<int64_t>& off, int64_t a0, DevBuf& blk_buf,
                      DevBuf& desc_buf, DevBuf& u_buf, DevBuf& e_buf, DevJacobi& out);
// y = (A + T) x on the context's packed layout with T the derivative
// terminator; precond: the apply's Jacobi scaling flag; herm0 as DevPacked.
hipError_t trun_apply(heomq_ctx* c, const double* x, double* y, bool rdo, int precond, int herm0, hipStream_t st);
// The sparse operator with the context's current coefficients (rates gam).
heomq::DevPacked spa_operator(const heomq_ctx* c, const double* gam);
// The forward operator at time t (the context's own when no time-dependent
// field is set; else gamma_K(t) written into the context's td_gam on st).
int operator_at(heomq_ctx* c, double t, hipStream_t st, DevOperator& op);
void field_shifts(const heomq_ctx* c, double t, double* es);
}  // namespace heomq
