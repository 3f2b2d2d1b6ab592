// Device-side data layout and kernel launchers of the HEOM apply engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace heomq {

// One coupling of an output ADO to a neighbour ADO (8 bytes).  The term added
// to the output block Y is
//     Y += sL * cL * (OpL * Xt) + sR * cR * (Xt * OpR),   Xt = X or X^H,
// with (cL, cR, OpL, OpR) taken from the record `rec`.  Every term of the
// forward Liouvillian and of its real transpose has this form.
struct DevEntry {
  int32_t nbr;
  uint16_t rec;
  uint8_t flags;   // bit0: Xt = X^H, bit1: sL = -1, bit2: sR = -1
  uint8_t pad;
};

struct DevRec {
  double cLr, cLi, cRr, cRi;
  int32_t opL, opR;   // operator ids: 2*(orb*nspin + spin) + (0: c, 1: c^+); -1 none
  int32_t pad0, pad1;
};

// First ADO of each tier (for tier-dependent projections inside kernels).
struct TierBounds {
  int64_t first[12];   // first[ntier] = N
  int ntier;
};

// Device tables of one operator (forward or transposed).
struct DevOperator {
  int64_t n_ado = 0;
  const int64_t* rowptr = nullptr;   // n_ado + 1
  const DevEntry* ent = nullptr;
  const DevRec* rec = nullptr;
  const double* gam = nullptr;       // per-ADO diagonal rate (complex)
  const int8_t* opmap = nullptr;     // [nop][4][d]: pL, sL, pR, sR
  const int32_t* hnnz = nullptr;     // H in row-CSR: nnz per row [d]
  const int32_t* hcol = nullptr;     // [d][d]
  const double* hval = nullptr;      // [d][d] complex
  int nop = 0;
  int d = 0;
  double hsign = -1.0;               // coefficient of i in hsign*i*(HX - XH)
  bool transposed = false;
  const int8_t* parity = nullptr;    // electron-number parity of each Fock state [d]
  TierBounds tb{};
  int trace_elem = 0;                // element (i + j*d) of the trace-row ADO the trace row replaces:
                                     // d*d - 1 (TFQMR convention) or 0 (BiCG convention)
};

// ---------------------------------------------------------------------------
// Symmetry-packed layout (solver path).  H and the bath couplings conserve the
// electron number N and S_z, so element (i, j) of ADO K can be nonzero only if
// N_i - N_j = dN_K and 2(S_z,i - S_z,j) = d2S_K, with (dN_K, d2S_K) the sums of
// q = +-1 (sigma = +/-) and q*(+-1 for spin up/down) over K's drawers.  ADOs
// with the same (dN, d2S) share one "sector": an ordered list of allowed
// elements.  A packed vector stores, ADO after ADO, only those elements.
// ---------------------------------------------------------------------------
struct DevEntryP {       // packed coupling (16 bytes)
  int32_t base;          // neighbour's first packed element
  int32_t map;           // offset of the paired gather map (output slot -> two signed neighbour
                         // slots, left and right operator, as int16 halves); -1: no term
  int32_t ncls;          // the neighbour's sector (host-side table building only)
  uint16_t rec;          // coefficient record (cL, cR)
  uint8_t flags;         // bit0 X^H, bit1 sL = -1, bit2 sR = -1
  uint8_t pad;
};

// LDS limits of the restructured packed apply (checked when its tables are built)
constexpr int kHeavyStage = 256;   // entries of one heavy row staged at once
constexpr int kHeavyPairs = 296;   // (group, neighbour slot) accumulator pairs of one heavy ADO
constexpr int kHeavyGroups = 47;   // entry runs of one heavy ADO
constexpr int kLightMaxU = 8;      // couplings of one top-tier ADO

struct DevPacked {
  int64_t n_ado = 0, total = 0;      // ADOs, packed complex elements
  const int64_t* off = nullptr;      // packed offset of each ADO (n_ado + 1)
  const uint16_t* sec = nullptr;     // sector of each ADO
  const int16_t* sec_elems = nullptr;   // [S][d*d]: element index i + j*d of slot t
  const int16_t* sec_pos = nullptr;     // [S][d*d]: slot of element i + j*d, -1 if forbidden
  const int32_t* block_ado = nullptr;   // ADO of packed element 256*b (one per 256-thread block, + sentinel)
  const int64_t* rowptr = nullptr;
  const DevEntryP* ent = nullptr;
  const DevRec* rec = nullptr;
  const double* gam = nullptr;
  const int8_t* opmap = nullptr;
  const int32_t* hnnz = nullptr;
  const int32_t* hcol = nullptr;
  const double* hval = nullptr;
  int nop = 0, d = 0;
  double hsign = -1.0;
  bool transposed = false;
  int32_t trace_slot = -1;           // packed index (within ADO 0) of element (d-1, d-1)
  // split of the work: ADOs below n_heavy (the tiers under the top one, with
  // ~M couplings each) run one workgroup per ADO with the coupling list sliced
  // over the lanes; the top tier (few couplings) runs one thread per element
  // starting at packed element light_begin, with its own block table.
  int64_t n_heavy = 0;
  int64_t light_begin = 0;
  const int32_t* light_nz = nullptr;         // top-tier ADOs that hold >= 1 packed element, in order
  const int32_t* light_block_ado = nullptr;  // per light block: index into light_nz of the ADO holding its
                                             // first element (+ one entry for the end); a block spans at most
                                             // 257 of these ADOs
  int max_light_couplings = 0;       // largest coupling count of a top-tier ADO (<= 8)
  // Grouped form of the ADOs below the top tier (the reference's
  // accumulate-then-apply order, calc_ax_tfqmr_omp.f90:289-353): the row of
  // heavy ADO a is sorted so that entries sharing a gather map (one c
  // operator, one neighbour sector, one dagger flag) are contiguous runs;
  // hg[hg_off[a] .. hg_off[a + 1] - 1] lists them as {first entry, end entry,
  // map, first accumulator pair | dagger << 31}, closed by a sentinel whose
  // last field is the ADO's accumulator pair count.  hg_off[a + 1] ==
  // hg_off[a]: the ADO takes the sliced form.  Null: no grouped form.
  const int32_t* hg_off = nullptr;
  const int4* hg = nullptr;
  // Sector-batched top tier: light block b takes the top-tier ADOs i =
  // lt_blk[b].x .. + lt_blk[b].y - 1 of a list sorted by sector, all of
  // sector lt_blk[b].w with lt_blk[b].z elements each; lt_off[i] is ADO i's
  // first packed element, lt_id[i] the ADO, lt_ent[i * lt_u + r] its
  // coupling r (map -1 past its last), lt_u the most couplings of a top-tier
  // ADO.  Null: the top tier runs in packed-element order (light_block_ado).
  const int4* lt_blk = nullptr;
  const int32_t* lt_off = nullptr;
  const int32_t* lt_id = nullptr;
  const DevEntryP* lt_ent = nullptr;
  int n_lt_blocks = 0;
  int lt_u = 0;
  // slot-space tables: i(HX - XH) per sector as CSR over packed slots, and the
  // gather maps (value (q + 1) * sign, 0 = no term) referenced by entries
  // the same rows padded to a fixed width hell_w (slot -1 / coefficient 0 as
  // padding) when every row is short: index loads then depend only on (s, t)
  int hell_w = 0;                    // 0: use the CSR form
  int hell_wp = 0;                   // hell_w rounded up to 8 (index rows are read 8 at a time)
  const int16_t* hq_ell = nullptr;   // [S][d*d][hell_wp]
  const double* hc_ell = nullptr;    // [S][d*d][hell_w] complex
  const int32_t* hrow = nullptr;     // [S][d*d + 1] row starts into hq / hc
  const int16_t* hq = nullptr;
  const double* hc = nullptr;        // complex coefficients of i(HX - XH)
  const int32_t* maps = nullptr;     // paired gather maps (int16 left | int16 right << 16)
  // batch of problems sharing every structural table (grid y = problem):
  // problem p's vectors start p * vstride complex elements in, its records,
  // ADO rates and commutator coefficients p * rec_stride / gam_stride /
  // hc_stride entries in (all 0 for one problem)
  int64_t vstride = 0, rec_stride = 0, gam_stride = 0, hc_stride = 0;
  int nprob = 1;
  // batches: problem p is skipped when frozen_flag[p * frozen_stride] == 3
  // (the solver's state word of a finished problem), and is not scaled when
  // frozen_flag[p * frozen_stride + 2] != 0 (its unscaled-iteration word);
  // null for neither
  const int32_t* frozen_flag = nullptr;
  int frozen_stride = 0;
  // 1: left Jacobi scaling, output element t divided by the operator's own
  // diagonal gamma_K + hsign * c_tt (c_tt: x_t's coefficient in i(HX - XH)_t);
  // elements whose diagonal vanishes (tier-0 populations) and the trace row
  // are left unscaled
  int precond = 0;
  // 1: the tier-0 output block is Hermitianised (the solvers' projected
  // operator); 0: the plain restriction of A to the sector (propagation)
  int herm0 = 1;
  // 1: ADOs 1 .. n_heavy - 1 run one wave each, four per workgroup (sectors
  // of at most 16 elements; ADO 0 keeps its own workgroup for the tier-0
  // fix-up)
  int heavy_wave = 0;
  // multi-shape launches: reals before this shape's first problem in x and y
  int64_t xoff = 0;
};

// Workgroups of the heavy part of a packed apply launch.
__host__ __device__ inline int64_t heavy_blocks(const DevPacked& op) {
  if (!op.heavy_wave) return op.n_heavy;
  return op.n_heavy > 0 ? 1 + (op.n_heavy - 1 + 3) / 4 : 0;
}
constexpr int kHeavyWaveMaxNnz = 16;

// Forward applies of problems of up to kApplyShapesPerLaunch shapes in one
// launch.  The shapes' operators travel in the kernel argument itself (so
// their table pointers are known to be global memory); shape g's blocks are
// bstart[g] .. bstart[g + 1] - 1 of the flat grid, nb[g] = apply_blocks(ops[g])
// per covered problem, the j-th covered problem being pmap[g][j] (j when
// null), and its vectors start ops[g].xoff reals into x and y.
constexpr int kApplyShapesPerLaunch = 10;
struct ApplyMulti {
  DevPacked ops[kApplyShapesPerLaunch];
  const int32_t* pmap[kApplyShapesPerLaunch] = {};
  int32_t bstart[kApplyShapesPerLaunch + 1] = {};
  int32_t nb[kApplyShapesPerLaunch] = {};
  int ng = 0;
};
static_assert(sizeof(ApplyMulti) <= 4000, "the multi-shape apply's kernel argument must stay under 4 KB");
// workgroups of one problem's packed apply
int64_t apply_blocks(const DevPacked& op);
// One Chebyshev term (td/chebyshev.f90) in one launch: y = L p1 (forward, no
// trace row, no Hermitian fix-up, one problem) is not stored; each element
// updates p += a y, then x += b p.  The same arithmetic as the apply followed
// by the two vector passes.
bool apply_cheb_fusable(const DevPacked& op);
hipError_t launch_apply_packed_cheb(const DevPacked& op, const double* p1, double* p, double* x, double a, double b,
                                    hipStream_t s);
// every shape needs a heavy part (top tier >= 1), whose block 0 does the
// tier-0 fix-up, and its trace slot
hipError_t launch_apply_packed_multi_f64(const ApplyMulti& am, const double* x, double* y, bool rdo, hipStream_t s);

// Projected apply on packed vectors: y = Pi A x with the trace row on the
// tier-0 block (rdo = true) or none, tier-0 output Hermitianised (herm0).
hipError_t launch_apply_packed_f64(const DevPacked& op, const double* x, double* y, bool rdo, hipStream_t s);
hipError_t launch_apply_packed_f32(const DevPacked& op, const float* x, float* y, bool rdo, hipStream_t s);
// dense (N*d*d complex) <-> packed conversions; unpack writes zeros off-sector.
template <typename TD, typename TP>
hipError_t launch_pack(const DevPacked& op, const TD* dense, TP* packed, hipStream_t s);
template <typename TP, typename TD>
hipError_t launch_unpack(const DevPacked& op, const TP* packed, TD* dense, hipStream_t s);
// flag = 1 when a dense (N*d*d complex) vector holds a nonzero element outside
// the packed sectors, else 0 (flag is written first).
hipError_t launch_offsector_any(const DevPacked& op, const double* dense, int32_t* flag, hipStream_t s);
// The BiCG convention on the tier-0 block of a packed vector (ADO 0's slots),
// one block: mode 0 rebuilds the upper triangle from the lower one with a real
// diagonal (calc_ax_bicg_omp.f90:70-77), mode 1 zeroes the upper triangle and
// the diagonal's imaginary part (:330-337), mode 2 folds the upper triangle
// onto the lower one (lower += conj(upper)^T, upper 0, diagonal real).
hipError_t launch_bicg_tier0_packed(const DevPacked& op, double* x, int mode, hipStream_t s);
// Hermitianise the tier-0 block of a packed vector in place.
template <typename T>
hipError_t launch_herm_packed(const DevPacked& op, T* x, hipStream_t s);

// y = A x.  trace_pos >= 0: apply the trace row at element (d-1, d-1) of ADO
// trace_pos (forward: overwrite with tr x0; transposed: treat that element
// of x as zero and add it to the diagonal of output block 0).
// project = true: the output is projected onto the physical sector (elements
// with parity(i) + parity(j) + tier odd are zeroed, the tier-0 block is
// Hermitianised); the solvers iterate with this projected operator.
hipError_t launch_apply_f64(const DevOperator& op, const double* x, double* y,
                            int64_t trace_pos, hipStream_t s, bool project = false);
hipError_t launch_apply_f32(const DevOperator& op, const float* x, float* y,
                            int64_t trace_pos, hipStream_t s, bool project = false);

// BiCG convention (bicg/calc_ax_bicg_omp.f90, calc_atrx_bicg_omp.f90).  Forward:
// x's tier 0 is rebuilt in place from its lower triangle, y = L x, y(1,1) of
// tier 0 = tr rho0, then y's tier-0 upper triangle is zeroed and its diagonal
// made real.  Transposed (op = the transposed operator): w's tier 0 is reduced
// in place to its lower triangle with a real diagonal, z = L^T w with w(1,1)
// moved onto the diagonal of z0, and z0 is folded back onto its lower triangle.
hipError_t launch_apply_bicg_f64(const DevOperator& op, double* x, double* y, hipStream_t s);
hipError_t launch_apply_transpose_bicg_f64(const DevOperator& op, double* w, double* z, hipStream_t s);
// rho0 of a vector rebuilt from its lower triangle (real diagonal, upper = conj lower^T).
hipError_t launch_herm_from_lower_f64(int d, double* x, hipStream_t s);
// rho0 of a vector reduced to its lower triangle with a real diagonal.
hipError_t launch_lower_tier0_f64(int d, double* x, hipStream_t s);

// Same-tier block inverse of the block-Jacobi iteration, in the eigenbasis of
// H.  H is block-diagonal over groups of Fock states sharing the quantum
// numbers the packed layout conserves; the eigenvectors (columns of u) of
// group g are supported on g, so element (i, j) of U^H X U only needs the
// elements (k, l) of X with k in group(i) and l in group(j), all of which the
// packed layout stores whenever it stores (i, j).
// One element (r, c) of a sector in the ADO-aligned block-Jacobi kernel: the
// slots of (r, c') for the nq states c' of group(c) and of (r', c) for the np
// states r' of group(r), with those states themselves (entries past nq / np
// are padding).
struct alignas(16) JacDesc {
  int16_t cslot[4];
  int16_t rslot[4];
  uint8_t c2[4];
  uint8_t r2[4];
  uint8_t r, c, nq, np;
  uint8_t pad[4];
};
static_assert(sizeof(JacDesc) == 32, "JacDesc is two 16-byte loads");

struct DevJacobi {
  const double* u = nullptr;        // d*d complex, column-major: u[k + i*d] = U_ki
  const double* e = nullptr;        // d eigenvalues of H
  const int16_t* grp = nullptr;     // [d][gmax]: the states of group(i)
  const int8_t* gn = nullptr;       // group size of each state
  int gmax = 0;
  double ejac = 0.0;                // the iteration's shift (> 0)
  // 1: every 256-element block of the packed layout lies in ADOs spanning at
  // most kJacobiWindow elements and 257 ADOs, so the one-kernel form applies
  int windowed = 0;
  // slot-space tables of the one-kernel form, per sector s and slot t of
  // element (r, c): the column list [crow[s][t], crow[s][t+1]) runs over
  // c' in group(c) with the slot of (r, c') and the coefficients U_{c'c}
  // (X -> X U) and conj(U_{cc'}) (X -> X U^H); the row list runs over r' in
  // group(r) with the slot of (r', c) and conj(U_{r'r}) (X -> U^H X) and
  // U_{rr'} (X -> U X); tde[s][t] = E_r - E_c
  const int32_t* crow = nullptr;    // [S][d*d + 1]
  const int16_t* cslot = nullptr;
  const double* ca = nullptr;       // complex
  const double* cc = nullptr;       // complex
  const int32_t* rrow = nullptr;    // [S][d*d + 1]
  const int16_t* rslot = nullptr;
  const double* rb = nullptr;       // complex
  const double* rd = nullptr;       // complex
  const double* tde = nullptr;      // [S][d*d]
  // ADO-aligned one-kernel form (d <= 16, groups of <= 4 states): block b owns
  // ADOs [blk[2b], blk[2b+2]), whose packed elements [blk[2b+1], blk[2b+3])
  // number at most 256; 0 blocks when the layout does not allow it
  int64_t ado_blocks = 0;
  int u_real = 0;                   // 1: every U_ki is real (a real H), the kernel reads U as doubles
  const int64_t* blk = nullptr;     // [ado_blocks + 1][2]: first ADO, first packed element
  const JacDesc* desc = nullptr;    // [S][d*d]
};

// t = (U^H r U)_ij / (ejac - gamma_K + i (E_i - E_j)) on the packed layout
hipError_t launch_jacobi_fwd(const DevPacked& op, const DevJacobi& jc, const double* r, double* t, hipStream_t s);
// x += U t U^H (accumulate) or x = U t U^H on the packed layout
hipError_t launch_jacobi_back(const DevPacked& op, const DevJacobi& jc, const double* t, double* x, bool accumulate,
                              hipStream_t s);
// The whole same-tier inverse in one kernel (jc.windowed): out (+)= U [(U^H in U) / w] U^H
// per ADO; in and out must differ.  Without accumulate the tier-0 block of out
// is made exactly Hermitian.
hipError_t launch_jacobi_win(const DevPacked& op, const DevJacobi& jc, const double* in, double* out, bool accumulate,
                             hipStream_t s);
// The same with one block per run of whole ADOs (jc.ado_blocks > 0); in and out must differ.
hipError_t launch_jacobi_ado(const DevPacked& op, const DevJacobi& jc, const double* in, double* out, bool accumulate,
                             hipStream_t s);
// The ADO-aligned form on FP32 vectors (the transform in FP64): the
// preconditioner of the FP32 inner solves of mixed precision.
hipError_t launch_jacobi_ado_f32(const DevPacked& op, const DevJacobi& jc, const float* in, float* out,
                                 hipStream_t s);
// The one-kernel form the layout allows (ADO-aligned, else the window form).
inline bool jacobi_one_kernel(const DevJacobi& jc) { return jc.ado_blocks > 0 || jc.windowed; }
hipError_t launch_jacobi_1k(const DevPacked& op, const DevJacobi& jc, const double* in, double* out, bool accumulate,
                            hipStream_t s);
constexpr int kJacobiWindow = 384;    // LDS window of the one-kernel form (complex elements); d <= 16 needs <= 328
// y = M^-1 (Pi A x), the block-Jacobi-preconditioned forward apply, in one
// launch (no A x in HBM).  Needs the sector-batched top tier and sliced heavy
// rows (apply forms 2 and 6 without the wave form), one problem, no Jacobi
// scaling, d <= 16 and the ADO-aligned inverse's tables.
bool apply_jac_fusable(const DevPacked& op, const DevJacobi& jc);
hipError_t launch_apply_packed_jac(const DevPacked& op, const DevJacobi& jc, const double* x, double* y, bool rdo,
                                   hipStream_t s);

// A TFQMR iteration's vector work carried by the preconditioned launch (kind
// 1: out = u2, w -= alpha u2, partials ||w||^2, rt.w, rt.u2 per workgroup,
// three arrays of the workgroup count; kind 2: out = aw, partial rt.aw per
// workgroup, and kind 1's arrays summed in runs of 16 workgroups into
// prev_out, three arrays of ceil(count / 16)); s is the problem's device
// scalars.  apply_jac_blocks is the launch's workgroup count.
struct TfqmrScal;
struct TfqEpi {
  const TfqmrScal* s = nullptr;
  double* w = nullptr;
  double* partial = nullptr;
  const double* prev = nullptr;   // kind 2: kind 1's three arrays, summed in runs of 16 into prev_out
  double* prev_out = nullptr;
};
int64_t apply_jac_blocks(const DevPacked& op);
hipError_t launch_apply_packed_jac_tfq(const DevPacked& op, const DevJacobi& jc, const double* x, double* y, bool rdo,
                                       int kind, const TfqEpi& ep, hipStream_t s);

}  // namespace heomq
