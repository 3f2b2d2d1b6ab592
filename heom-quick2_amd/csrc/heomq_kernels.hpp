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
};

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

}  // namespace heomq
