// C ABI of the engine: device context (tables uploaded once), applies,
// device-resident TFQMR steady-state solver, FP32-in-FP64 iterative
// refinement, RK4 / Chebyshev propagation and observables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstdint>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <memory>
#include <vector>

#include "heomq.h"
#include "heomq_kernels.hpp"
#include "heomq_model.hpp"
#include "heomq_solver.hpp"
#include "heomq_ctx.hpp"

using namespace heomq;

namespace heomq {

void build_records(const Model& m, std::vector<DevRec>& rec) {
  const int M = m.M;
  rec.assign(6 * (size_t)M, DevRec{});
  auto set = [&](DevRec& r, std::complex<double> cL, std::complex<double> cR, int oL, int oR) {
    r.cLr = cL.real(); r.cLi = cL.imag(); r.cRr = cR.real(); r.cRi = cR.imag();
    r.opL = oL; r.opR = oR; r.pad0 = r.pad1 = 0;
  };
  const std::complex<double> I(0, 1);
  for (int k = 0; k < M; ++k) {
    // forward lower (calc_ax_tfqmr_omp.f90:247-258): -i*ff*cb op(c) X + i*fr*cd X op(c)
    set(rec[k], -I * m.cb[k], I * m.cd[k], op_low(m, k), op_low(m, k));
    // forward upper (:278-279, :342-353): -i*ff op'(c) X + i*fr X op'(c), times
    // |cb_k|^(1/2) with lscale (:325-328)
    const double sc = m.dbsqrt.empty() ? 1.0 : m.dbsqrt[k];
    set(rec[M + k], -I * sc, I * sc, op_up(m, k), op_up(m, k));
    for (int kind = 0; kind < 2; ++kind) {
      const DevRec& f = rec[kind * M + k];
      const std::complex<double> cL(f.cLr, f.cLi), cR(f.cRr, f.cRi);
      // transpose of a direct term: conj(c) * Op^+ (the operator id with the dagger bit flipped)
      set(rec[(2 + kind) * M + k], std::conj(cL), std::conj(cR), f.opL ^ 1, f.opR ^ 1);
      // transpose of a term on X^H: a*Op*X^H -> a * W^H * Op, b*X^H*Op -> b * Op * W^H
      set(rec[(4 + kind) * M + k], cR, cL, f.opR, f.opL);
    }
  }
}

void build_opmap(const Model& m, std::vector<int8_t>& map) {
  const int d = m.d, nop = 2 * m.norbs * m.nspin;
  map.assign((size_t)nop * 4 * d, 0);
  for (int orb = 0; orb < m.norbs; ++orb)
    for (int s = 0; s < m.nspin; ++s)
      for (int t = 0; t < 2; ++t) {
        const int o = 2 * (orb * m.nspin + s) + t;
        int8_t* pL = &map[(size_t)o * 4 * d];
        int8_t* sL = pL + d;
        int8_t* pR = pL + 2 * d;
        int8_t* sR = pL + 3 * d;
        for (int i = 0; i < d; ++i) { pL[i] = -1; sL[i] = 0; pR[i] = -1; sR[i] = 0; }
        for (int r = 0; r < d; ++r)
          for (int c = 0; c < d; ++c) {
            const double v = t == 0 ? m.amsv(orb, s, r, c) : m.amsv(orb, s, c, r);   // op(r, c)
            if (v == 0.0) continue;
            pL[r] = (int8_t)c;                    // (op X)(r, j) = v * X(c, j)
            sL[r] = (int8_t)(v > 0 ? 1 : -1);
            pR[c] = (int8_t)r;                    // (X op)(i, c) = v * X(i, r)
            sR[c] = (int8_t)(v > 0 ? 1 : -1);
          }
      }
}

void build_hcsr(const std::vector<cplx>& H, int d, std::vector<int32_t>& nnz, std::vector<int32_t>& col,
                std::vector<double>& val) {
  nnz.assign(d, 0);
  col.assign((size_t)d * d, 0);
  val.assign((size_t)2 * d * d, 0.0);
  for (int i = 0; i < d; ++i)
    for (int k = 0; k < d; ++k) {
      const cplx h = H[i + (size_t)k * d];
      if (h == cplx(0, 0)) continue;
      const int t = nnz[i]++;
      col[(size_t)i * d + t] = k;
      val[2 * ((size_t)i * d + t)] = h.real();
      val[2 * ((size_t)i * d + t) + 1] = h.imag();
    }
}

// per-ADO diagonal rates gamma_K = sum_r gdraw[K_r]
void build_gamma(const Model& m, const std::vector<cplx>& gdraw, std::vector<double>& gam, bool conj) {
  gam.assign(2 * (size_t)m.N, 0.0);
  for (int n = 1; n <= m.L; ++n)
    for (int64_t g = m.first[n]; g < m.first[n + 1]; ++g) {
      cplx s(0, 0);
      const int32_t* K = &m.tiers[n][(g - m.first[n]) * n];
      for (int r = 0; r < n; ++r) s += gdraw[K[r]];
      gam[2 * g] = s.real();
      gam[2 * g + 1] = conj ? -s.imag() : s.imag();
    }
}

}  // namespace heomq

namespace {

int build_entries(const Model& m, std::vector<int64_t>& rowf, std::vector<DevEntry>& entf,
                  std::vector<int64_t>& rowt, std::vector<DevEntry>& entt) {
  const int M = m.M;
  rowf.assign(m.N + 1, 0);
  entf.clear();
  entf.reserve(m.lower.size() + m.upper.size());
  auto flags = [](int itype, int ff, int fr) {
    return (uint8_t)((itype ? 1 : 0) | (ff < 0 ? 2 : 0) | (fr < 0 ? 4 : 0));
  };
  for (int64_t g = 0; g < m.N; ++g) {
    rowf[g] = (int64_t)entf.size();
    for (int64_t p = m.lower_off[g]; p < m.lower_off[g + 1]; ++p) {
      const LowerEntry& e = m.lower[p];
      entf.push_back(DevEntry{e.nbr, (uint16_t)e.drawer, flags(e.itype, e.ff, e.fr), 0});
    }
    if (g < m.n_upper_ados)
      for (int k = 0; k < M; ++k) {
        const UpperEntry& e = m.upper[(size_t)g * M + k];
        if (e.nbr < 0) continue;
        entf.push_back(DevEntry{e.nbr, (uint16_t)(M + k), flags(e.itype, e.ff, e.fr), 0});
      }
  }
  rowf[m.N] = (int64_t)entf.size();
  if (6 * (int64_t)M > 65535) return fail(HEOMQ_UNSUPPORTED, "too many drawers for the record table");
  // reverse every forward entry: K <- J becomes J <- K
  rowt.assign(m.N + 1, 0);
  for (const DevEntry& e : entf) rowt[e.nbr + 1]++;
  for (int64_t g = 0; g < m.N; ++g) rowt[g + 1] += rowt[g];
  entt.assign(entf.size(), DevEntry{});
  std::vector<int64_t> fill(rowt.begin(), rowt.end() - 1);
  for (int64_t g = 0; g < m.N; ++g)
    for (int64_t p = rowf[g]; p < rowf[g + 1]; ++p) {
      const DevEntry& e = entf[p];
      const int kind = e.rec >= M ? 1 : 0;
      const int k = e.rec - kind * M;
      const bool dag = e.flags & 1;
      const bool nL = e.flags & 2, nR = e.flags & 4;
      DevEntry t;
      t.nbr = (int32_t)g;
      t.rec = (uint16_t)((dag ? 4 + kind : 2 + kind) * M + k);
      t.flags = dag ? (uint8_t)(1 | (nR ? 2 : 0) | (nL ? 4 : 0)) : (uint8_t)((nL ? 2 : 0) | (nR ? 4 : 0));
      t.pad = 0;
      entt[fill[e.nbr]++] = t;
    }
  return HEOMQ_OK;
}

int upload_coefs(heomq_ctx* c, const Model& m) {
  std::vector<DevRec> rec;
  build_records(m, rec);
  HIPCHECK(upload(c->rec, rec));
  std::vector<double> gam;
  build_gamma(m, m.gdraw, gam, false);
  HIPCHECK(upload(c->fwd.gam, gam));
  build_gamma(m, m.gdraw, gam, true);
  HIPCHECK(upload(c->trn.gam, gam));
  std::vector<int32_t> nnz, col;
  std::vector<double> val;
  build_hcsr(m.H, m.d, nnz, col, val);
  HIPCHECK(upload(c->hnnz, nnz));
  HIPCHECK(upload(c->hcol, col));
  HIPCHECK(upload(c->hval, val));
  return HEOMQ_OK;
}

void fill_op(heomq_ctx* c, OperatorTables& t, const DevBuf& rec, bool transposed, double hsign) {
  DevOperator& o = t.op;
  o.n_ado = c->N;
  o.rowptr = t.rowptr.as<int64_t>();
  o.ent = t.ent.as<DevEntry>();
  o.rec = rec.as<DevRec>();
  o.gam = t.gam.as<double>();
  o.opmap = c->opmap.as<int8_t>();
  o.hnnz = c->hnnz.as<int32_t>();
  o.hcol = c->hcol.as<int32_t>();
  o.hval = c->hval.as<double>();
  o.nop = 2 * c->norbs * c->nspin;
  o.d = c->d;
  o.hsign = hsign;
  o.transposed = transposed;
  o.parity = c->parity.as<int8_t>();
  o.tb = c->tb;
  o.trace_elem = c->d * c->d - 1;
}

// i(HX - XH) per sector as a CSR over its slots (row t of sector s: the slots
// q and coefficients v with (i(HX - XH))_t = sum v * X_q, elements outside the
// sector dropped), and the same rows padded to a fixed width when every row
// is short and the padded table stays under ell_cap bytes.
struct HTables {
  std::vector<int32_t> hrow;
  std::vector<int16_t> hq, hq_ell;
  std::vector<double> hc, hc_ell;
  int ew = 0, ewp = 0;
};

void build_h_tables(const std::vector<cplx>& H, int d, int S, const std::vector<int16_t>& elems,
                    const std::vector<int16_t>& pos, const std::vector<int>& nnz, size_t ell_cap, HTables& ht) {
  const int dd = d * d;
  auto& hrow = ht.hrow;
  auto& hq = ht.hq;
  auto& hc = ht.hc;
  hrow.assign((size_t)S * (dd + 1), 0);
  hq.clear();
  hc.clear();
  for (int s = 0; s < S; ++s) {
    const int nz = nnz[s];
    for (int t = 0; t < nz; ++t) {
      hrow[(size_t)s * (dd + 1) + t] = (int32_t)hq.size();
      const int e = elems[(size_t)s * dd + t];
      const int i = e % d, j = e / d;
      std::vector<std::pair<int, cplx>> terms;
      auto addterm = [&](int q, cplx v) {
        if (q < 0 || v == cplx(0, 0)) return;
        for (auto& tm : terms)
          if (tm.first == q) { tm.second += v; return; }
        terms.push_back({q, v});
      };
      for (int k = 0; k < d; ++k) {
        const cplx hik = H[i + (size_t)k * d], hkj = H[k + (size_t)j * d];
        if (hik != cplx(0, 0)) addterm(pos[(size_t)s * dd + k + j * d], cplx(0, 1) * hik);
        if (hkj != cplx(0, 0)) addterm(pos[(size_t)s * dd + i + k * d], -cplx(0, 1) * hkj);
      }
      for (auto& tm : terms) {
        if (tm.second == cplx(0, 0)) continue;
        hq.push_back((int16_t)tm.first);
        hc.push_back(tm.second.real());
        hc.push_back(tm.second.imag());
      }
    }
    for (int t = nz; t <= dd; ++t) hrow[(size_t)s * (dd + 1) + t] = (int32_t)hq.size();
  }
  // fixed-width copy when all rows are short
  int width = 0;
  for (int s = 0; s < S; ++s)
    for (int t = 0; t < dd; ++t)
      width = std::max(width, hrow[(size_t)s * (dd + 1) + t + 1] - hrow[(size_t)s * (dd + 1) + t]);
  const int kMaxEll = 16;
  int ew = width <= kMaxEll ? std::max(width, 1) : 0;
  if (ew > 0 && (size_t)S * dd * ew * 16 > ell_cap) ew = 0;
  const int ewp = (ew + 7) / 8 * 8;
  if (ew > 0) {
    ht.hq_ell.assign((size_t)S * dd * ewp, 0);
    ht.hc_ell.assign((size_t)S * dd * ew * 2, 0.0);
    for (int s = 0; s < S; ++s)
      for (int t = 0; t < dd; ++t) {
        const int r0 = hrow[(size_t)s * (dd + 1) + t], r1 = hrow[(size_t)s * (dd + 1) + t + 1];
        for (int u = r0; u < r1; ++u) {
          const size_t at = ((size_t)s * dd + t) * ew + (u - r0);
          ht.hq_ell[((size_t)s * dd + t) * ewp + (u - r0)] = hq[u];
          ht.hc_ell[2 * at] = hc[2 * (size_t)u];
          ht.hc_ell[2 * at + 1] = hc[2 * (size_t)u + 1];
        }
      }
  } else {
    ht.hq_ell.assign(1, 0);
    ht.hc_ell.assign(2, 0.0);
  }
  if (hq.empty()) { hq.push_back(0); hc.push_back(0); hc.push_back(0); }
  ht.ew = ew;
  ht.ewp = ewp;
}

// The solver layout's H tables.  Built at context creation and again whenever
// H changes; H must keep the N / S_z conservation the sector layout was built
// for.
int build_packed_h(heomq_ctx* c, const Model& m) {
  const int d = m.d;
  c->h_diagonal = true;
  for (int i = 0; i < d; ++i)
    for (int k = 0; k < d; ++k)
      if (i != k && m.H[i + (size_t)k * d] != cplx(0, 0)) c->h_diagonal = false;
  for (int i = 0; i < d; ++i)
    for (int k = 0; k < d; ++k)
      if (m.H[i + (size_t)k * d] != cplx(0, 0)) {
        if ((c->cons_n && c->fock_n[i] != c->fock_n[k]) || (c->cons_s && c->fock_s[i] != c->fock_s[k]))
          return fail(HEOMQ_UNSUPPORTED,
                      "H no longer conserves the quantum numbers the packed layout uses: create a new context");
      }
  HTables ht;
  build_h_tables(m.H, d, c->nsec, c->sec_elems_h, c->sec_pos_h, c->sec_nnz_h, SIZE_MAX, ht);
  HIPCHECK(upload(c->p_hrow, ht.hrow));
  HIPCHECK(upload(c->p_hq, ht.hq));
  HIPCHECK(upload(c->p_hc, ht.hc));
  HIPCHECK(upload(c->p_hq_ell, ht.hq_ell));
  HIPCHECK(upload(c->p_hc_ell, ht.hc_ell));
  for (int w = 0; w < 2; ++w) {
    DevPacked& p = w == 0 ? c->pfwd : c->ptrn;
    p.hrow = c->p_hrow.as<int32_t>();
    p.hq = c->p_hq.as<int16_t>();
    p.hc = c->p_hc.as<double>();
    p.hell_w = ht.ew;
    p.hell_wp = ht.ewp;
    p.hq_ell = c->p_hq_ell.as<int16_t>();
    p.hc_ell = c->p_hc_ell.as<double>();
  }
  return HEOMQ_OK;
}

// Symmetry-packed layout: sectors, packed offsets, packed coupling entries.
// Host-side symmetry-packed layout (sectors, offsets, block tables, packed
// entries and gather maps); built without a device, uploaded by build_packed.
// A "sector" is an ordered list of the elements an ADO stores; the solver
// layout takes the (N, S_z) selection rule's lists, the sparse-storage layout
// the per-ADO lists of a COO pattern (heomq_spa_pattern).
struct PackedHost {
  bool cons_n = true, cons_s = true;
  std::vector<int> fN, fS;           // electron number and 2 S_z of each Fock state
  int S = 0;                         // sectors
  std::vector<int16_t> elems, pos;   // [S][d*d]
  std::vector<int> nnz;              // elements per sector
  std::vector<uint16_t> sec;         // sector of each ADO
  std::vector<int64_t> off;          // N + 1
  int64_t total = 0;
  std::vector<int32_t> block, lblock, lnz;
  int64_t n_heavy = 0, light_begin = 0;
  std::vector<DevEntryP> pentf, pentt;
  std::vector<int32_t> maps;
  int tslot = -1;
  int max_light[2] = {0, 0};
  // grouped heavy rows (per operator) and sector-batched top-tier blocks
  std::vector<int32_t> hg_off[2];
  std::vector<int4> hg[2];
  std::vector<int4> lt_blk;
  std::vector<int32_t> lt_ados, lt_off;
  std::vector<DevEntryP> lt_ent[2];
  int lt_u = 0;
};

constexpr int kMaxSectors = 65535;

// Sort each heavy ADO's packed entries into runs of one gather map (and one
// dagger flag) and list the runs; an ADO whose row or accumulator pairs do
// not fit the kernel's LDS keeps the sliced form (no runs listed).
void group_heavy_rows(const std::vector<int64_t>& rows, std::vector<DevEntryP>& ent, const std::vector<int>& nnz,
                      int64_t n_heavy, std::vector<int32_t>& off, std::vector<int4>& grp) {
  off.assign((size_t)n_heavy + 1, 0);
  grp.clear();
  for (int64_t g = 0; g < n_heavy; ++g) {
    const int64_t p0 = rows[g], p1 = rows[g + 1];
    auto key = [](const DevEntryP& e) { return std::make_pair((uint32_t)e.map, e.flags & 1); };
    std::stable_sort(ent.begin() + p0, ent.begin() + p1,
                     [&](const DevEntryP& a, const DevEntryP& b) { return key(a) < key(b); });
    std::vector<int4> mine;
    int P = 0;
    for (int64_t p = p0; p < p1;) {
      int64_t q = p;
      while (q < p1 && key(ent[q]) == key(ent[p])) ++q;
      if (ent[p].map >= 0) {
        mine.push_back(make_int4((int)p, (int)q, ent[p].map, P | ((ent[p].flags & 1) ? (int)0x80000000 : 0)));
        P += nnz[ent[p].ncls];
      }
      p = q;
    }
    const bool fits = p1 - p0 <= kHeavyStage && P <= kHeavyPairs && (int)mine.size() <= kHeavyGroups && !mine.empty();
    if (fits) {
      grp.insert(grp.end(), mine.begin(), mine.end());
      grp.push_back(make_int4(0, 0, -1, P));
    }
    off[g + 1] = (int32_t)grp.size();
  }
  if (grp.empty()) grp.push_back(make_int4(0, 0, -1, 0));
}

// Electron number and 2 S_z of every Fock state (occupation basis, 4 states
// per orbital: empty, up, down, both).
void fock_numbers(const Model& m, std::vector<int>& fN, std::vector<int>& fS) {
  fN.assign(m.d, 0);
  fS.assign(m.d, 0);
  for (int i = 0; i < m.d; ++i) {
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
}

// The solver layout's sectors: ADOs with equal (dN, 2 dS_z) share one list of
// the elements the selection rule allows, column-major over (i, j).
int symmetry_sectors(const Model& m, PackedHost& ph) {
  const int d = m.d, dd = d * d;
  fock_numbers(m, ph.fN, ph.fS);
  const auto& fN = ph.fN;
  const auto& fS = ph.fS;
  ph.cons_n = ph.cons_s = true;
  for (int i = 0; i < d; ++i)
    for (int k = 0; k < d; ++k)
      if (m.H[i + (size_t)k * d] != cplx(0, 0)) {
        if (fN[i] != fN[k]) ph.cons_n = false;
        if (fS[i] != fS[k]) ph.cons_s = false;
      }
  // per-ADO (dN, d2S)
  std::vector<std::pair<int, int>> key(m.N, {0, 0});
  for (int n = 1; n <= m.L; ++n)
    for (int64_t g = m.first[n]; g < m.first[n + 1]; ++g) {
      const int32_t* K = &m.tiers[n][(g - m.first[n]) * n];
      int dN = 0, dS = 0;
      for (int r = 0; r < n; ++r) {
        const int q = m.sgn[K[r]] == 0 ? 1 : -1;
        dN += q;
        dS += q * (m.spin[K[r]] == 0 ? 1 : -1);
      }
      if (!ph.cons_n) dN = ((dN % 2) + 2) % 2;
      if (!ph.cons_s) dS = 0;
      key[g] = {dN, dS};
    }
  std::vector<std::pair<int, int>> keys;
  ph.sec.assign(m.N, 0);
  for (int64_t g = 0; g < m.N; ++g) {
    auto it = std::find(keys.begin(), keys.end(), key[g]);
    if (it == keys.end()) {
      if (keys.size() >= 255) return fail(HEOMQ_UNSUPPORTED, "too many symmetry sectors");
      keys.push_back(key[g]);
      it = keys.end() - 1;
    }
    ph.sec[g] = (uint16_t)(it - keys.begin());
  }
  const int S = (int)keys.size();
  ph.S = S;
  ph.elems.assign((size_t)S * dd, -1);
  ph.pos.assign((size_t)S * dd, -1);
  ph.nnz.assign(S, 0);
  for (int s = 0; s < S; ++s)
    for (int j = 0; j < d; ++j)
      for (int i = 0; i < d; ++i) {
        const int dn = fN[i] - fN[j], ds = fS[i] - fS[j];
        const bool okn = ph.cons_n ? dn == keys[s].first : (((dn - keys[s].first) % 2) == 0);
        const bool oks = ph.cons_s ? ds == keys[s].second : true;
        if (!okn || !oks) continue;
        ph.pos[(size_t)s * dd + i + j * d] = (int16_t)ph.nnz[s];
        ph.elems[(size_t)s * dd + ph.nnz[s]] = (int16_t)(i + j * d);
        ph.nnz[s]++;
      }
  return HEOMQ_OK;
}

// Everything that follows from the sectors (ph.S, sec, elems, pos, nnz):
// packed offsets, block tables, gather maps, packed entries, the grouped heavy
// rows and the sector-batched top tier.  both = false builds the forward
// operator's entries only.
int finish_packed_host(const Model& m, const std::vector<int64_t>& rowf, const std::vector<DevEntry>& entf,
                       const std::vector<int64_t>& rowt, const std::vector<DevEntry>& entt, PackedHost& ph,
                       bool both) {
  std::vector<DevRec> recs;
  build_records(m, recs);
  std::vector<int8_t> omap;
  build_opmap(m, omap);
  const int d = m.d, dd = d * d;
  const int S = ph.S;
  const auto& sec = ph.sec;
  const auto& elems = ph.elems;
  const auto& pos = ph.pos;
  const auto& nnz = ph.nnz;
  std::vector<int64_t> off(m.N + 1, 0);
  for (int64_t g = 0; g < m.N; ++g) off[g + 1] = off[g] + nnz[sec[g]];
  const int64_t total = off[m.N];
  if (total >= ((int64_t)1 << 31)) return fail(HEOMQ_UNSUPPORTED, "packed vector too long");
  const int64_t nblk = (total + 255) / 256;
  std::vector<int32_t> block((size_t)nblk + 1, 0);
  {
    int64_t a = 0;
    for (int64_t b = 0; b <= nblk; ++b) {
      const int64_t t = std::min<int64_t>(b * 256, std::max<int64_t>(total - 1, 0));
      while (a + 1 < m.N && off[a + 1] <= t) ++a;
      block[b] = (int32_t)a;
    }
  }
  // heavy / light split at the top tier
  const int64_t n_heavy = m.L >= 1 ? m.first[m.L] : m.N;
  const int64_t light_begin = off[n_heavy];
  const int64_t nlb = (total - light_begin + 255) / 256;
  // top-tier ADOs with at least one allowed element (an ADO whose drawers carry
  // a charge no pair of Fock states can match holds none)
  std::vector<int32_t> lnz;
  for (int64_t a = n_heavy; a < m.N; ++a)
    if (off[a + 1] > off[a]) lnz.push_back((int32_t)a);
  std::vector<int32_t> lblock((size_t)std::max<int64_t>(nlb, 0) + 1, 0);
  {
    size_t q = 0;
    for (int64_t b = 0; b <= nlb && !lnz.empty(); ++b) {
      const int64_t t = std::min<int64_t>(light_begin + b * 256, total - 1);
      while (q + 1 < lnz.size() && off[lnz[q + 1]] <= t) ++q;
      lblock[b] = (int32_t)q;
    }
    // what the light kernel assumes: every block's element range lies in the
    // ADOs lnz[lo .. hi] and there are at most 257 of them (its LDS tables)
    for (int64_t b = 0; b < nlb; ++b) {
      const int32_t lo = lblock[b], hi = lblock[b + 1];
      const int64_t t0 = light_begin + b * 256, t1 = std::min<int64_t>(t0 + 255, total - 1);
      if (hi < lo || hi - lo + 1 > 257 || hi >= (int32_t)lnz.size() || off[lnz[lo]] > t0 ||
          off[lnz[hi] + 1] <= t1)
        return fail(HEOMQ_UNSUPPORTED, "packed layout: light block table inconsistent");
    }
  }
  if (lnz.empty()) lnz.push_back(0);
  // paired gather maps: output slot t of sector s -> the signed slots in the
  // neighbour (sector s2) of its left-operator and right-operator terms
  std::vector<int32_t> maps;
  std::map<std::tuple<int, int, int, int, int>, int32_t> mapidx;
  auto side_value = [&](int s, int s2, int o, int dag, int side, int t) -> int {
    if (o < 0) return 0;
    const int8_t* pL = &omap[(size_t)o * 4 * d];
    const int8_t* sL = pL + d;
    const int8_t* pR = pL + 2 * d;
    const int8_t* sR = pL + 3 * d;
    const int e = elems[(size_t)s * dd + t];
    const int i = e % d, j = e / d;
    if (side == 0) {
      const int pl = pL[i];
      if (pl < 0) return 0;
      const int q = pos[(size_t)s2 * dd + (dag ? (j + pl * d) : (pl + j * d))];
      return q >= 0 ? (q + 1) * sL[i] : 0;
    }
    const int pr = pR[j];
    if (pr < 0) return 0;
    const int q = pos[(size_t)s2 * dd + (dag ? (pr + i * d) : (i + pr * d))];
    return q >= 0 ? (q + 1) * sR[j] : 0;
  };
  auto getmap = [&](int s, int s2, int oL, int oR, int dag) -> int32_t {
    const auto key = std::make_tuple(s, s2, oL, oR, dag);
    auto it = mapidx.find(key);
    if (it != mapidx.end()) return it->second;
    const int32_t at = (int32_t)maps.size();
    bool any = false;
    for (int t = 0; t < nnz[s]; ++t) {
      const int vl = side_value(s, s2, oL, dag, 0, t), vr = side_value(s, s2, oR, dag, 1, t);
      any |= (vl != 0 || vr != 0);
      maps.push_back((int32_t)((uint32_t)(uint16_t)(int16_t)vl | ((uint32_t)(uint16_t)(int16_t)vr << 16)));
    }
    const int32_t res = any ? at : -1;
    if (!any) maps.resize(at);
    if ((int64_t)maps.size() >= ((int64_t)1 << 31) - 1) return -2;
    mapidx[key] = res;
    return res;
  };
  bool too_many = false;
  auto conv = [&](const std::vector<int64_t>& rows, const std::vector<DevEntry>& e) {
    std::vector<DevEntryP> out(e.size());
    for (int64_t g = 0; g < m.N; ++g)
      for (int64_t p = rows[g]; p < rows[g + 1]; ++p) {
        const DevRec& r = recs[e[p].rec];
        const int dag = e[p].flags & 1;
        const int s = sec[g], s2 = sec[e[p].nbr];
        int32_t mp = getmap(s, s2, r.opL, r.opR, dag);
        if (mp == -2) { too_many = true; mp = -1; }
        out[p] = DevEntryP{(int32_t)off[e[p].nbr], mp, s2, e[p].rec, e[p].flags, 0};
      }
    return out;
  };
  ph.pentf = conv(rowf, entf);
  if (both) ph.pentt = conv(rowt, entt);
  if (too_many) return fail(HEOMQ_UNSUPPORTED, "packed layout: gather maps exceed 2^31 entries");
  group_heavy_rows(rowf, ph.pentf, nnz, n_heavy, ph.hg_off[0], ph.hg[0]);
  if (both) group_heavy_rows(rowt, ph.pentt, nnz, n_heavy, ph.hg_off[1], ph.hg[1]);
  if (maps.empty()) maps.push_back(0);
  ph.off = std::move(off);
  ph.total = total;
  ph.block = std::move(block);
  ph.lblock = std::move(lblock);
  ph.lnz = std::move(lnz);
  ph.n_heavy = n_heavy;
  ph.light_begin = light_begin;
  ph.maps = std::move(maps);
  ph.tslot = ph.pos[(size_t)ph.sec[0] * dd + (d - 1) + (d - 1) * d];
  for (int w = 0; w < (both ? 2 : 1); ++w) {
    const auto& rows = w == 0 ? rowf : rowt;
    int mx = 0;
    for (int64_t g = n_heavy; g < m.N; ++g) mx = std::max<int>(mx, (int)(rows[g + 1] - rows[g]));
    ph.max_light[w] = mx;
  }
  // sector-batched top-tier blocks: the non-empty top-tier ADOs of one sector
  // (in ADO order), 256 / nnz of them per block; each with its first packed
  // element and its couplings padded to lt_u entries (map -1)
  ph.lt_u = std::max(1, std::max(ph.max_light[0], ph.max_light[1]));
  bool lt_ok = ph.lt_u <= kLightMaxU && n_heavy < m.N;
  for (int s = 0; s < S && lt_ok; ++s)
    if (ph.nnz[s] > 256) lt_ok = false;
  if (lt_ok) {
    std::vector<std::vector<int64_t>> by_sec(S);
    for (int64_t a = n_heavy; a < m.N; ++a) by_sec[ph.sec[a]].push_back(a);
    for (int s = 0; s < S; ++s) {
      if (ph.nnz[s] == 0) continue;
      const int chunk = 256 / ph.nnz[s];
      int in_block = 0;
      for (int64_t a : by_sec[s]) {
        if (in_block == 0) ph.lt_blk.push_back(make_int4((int)ph.lt_ados.size(), 0, ph.nnz[s], s));
        ph.lt_ados.push_back((int32_t)a);
        ph.lt_off.push_back((int32_t)ph.off[a]);
        ph.lt_blk.back().y += 1;
        if (++in_block == chunk) in_block = 0;
      }
    }
    for (int w = 0; w < (both ? 2 : 1); ++w) {
      const auto& rows = w == 0 ? rowf : rowt;
      const auto& pent = w == 0 ? ph.pentf : ph.pentt;
      auto& out = ph.lt_ent[w];
      out.assign(ph.lt_ados.size() * (size_t)ph.lt_u, DevEntryP{0, -1, -1, 0, 0, 0});
      for (size_t i = 0; i < ph.lt_ados.size(); ++i) {
        const int64_t a = ph.lt_ados[i];
        for (int64_t p = rows[a]; p < rows[a + 1]; ++p) out[i * ph.lt_u + (p - rows[a])] = pent[p];
      }
    }
  }
  return HEOMQ_OK;
}

int build_packed_host(const Model& m, const std::vector<int64_t>& rowf, const std::vector<DevEntry>& entf,
                      const std::vector<int64_t>& rowt, const std::vector<DevEntry>& entt, PackedHost& ph) {
  if (int rc = symmetry_sectors(m, ph)) return rc;
  return finish_packed_host(m, rowf, entf, rowt, entt, ph, true);
}
void set_apply_form(heomq_ctx* c, int w, int form);

int build_packed(heomq_ctx* c, const Model& m, const std::vector<int64_t>& rowf, const std::vector<DevEntry>& entf,
                 const std::vector<int64_t>& rowt, const std::vector<DevEntry>& entt) {
  PackedHost ph;
  if (int rc = build_packed_host(m, rowf, entf, rowt, entt, ph)) return rc;
  c->cons_n = ph.cons_n;
  c->cons_s = ph.cons_s;
  c->sec_elems_h = ph.elems;
  c->sec_pos_h = ph.pos;
  c->sec_nnz_h = ph.nnz;
  c->fock_n = ph.fN;
  c->fock_s = ph.fS;
  c->poff_h = ph.off;
  c->psec_h = ph.sec;
  HIPCHECK(upload(c->p_lblock, ph.lblock));
  HIPCHECK(upload(c->p_lnz, ph.lnz));
  HIPCHECK(upload(c->p_off, ph.off));
  HIPCHECK(upload(c->p_sec, ph.sec));
  HIPCHECK(upload(c->p_elems, ph.elems));
  HIPCHECK(upload(c->p_pos, ph.pos));
  HIPCHECK(upload(c->p_block, ph.block));
  HIPCHECK(upload(c->p_entf, ph.pentf));
  HIPCHECK(upload(c->p_entt, ph.pentt));
  HIPCHECK(upload(c->p_maps, ph.maps));
  HIPCHECK(upload(c->p_hgoff_f, ph.hg_off[0]));
  HIPCHECK(upload(c->p_hgoff_t, ph.hg_off[1]));
  HIPCHECK(upload(c->p_hg_f, ph.hg[0]));
  HIPCHECK(upload(c->p_hg_t, ph.hg[1]));
  if (!ph.lt_blk.empty()) {
    HIPCHECK(upload(c->p_ltblk, ph.lt_blk));
    HIPCHECK(upload(c->p_ltados, ph.lt_ados));
    HIPCHECK(upload(c->p_ltoff, ph.lt_off));
    HIPCHECK(upload(c->p_ltent_f, ph.lt_ent[0]));
    HIPCHECK(upload(c->p_ltent_t, ph.lt_ent[1]));
  }
  c->nsec = ph.S;
  if (int rc = build_packed_h(c, m)) return rc;
  c->ptotal = ph.total;
  c->heavy_wave_ok = ph.n_heavy > 1;
  for (int64_t a = 1; a < ph.n_heavy; ++a)
    if (ph.off[a + 1] - ph.off[a] > kHeavyWaveMaxNnz) c->heavy_wave_ok = false;
  for (int w = 0; w < 2; ++w) {
    DevPacked& p = w == 0 ? c->pfwd : c->ptrn;
    p.n_ado = m.N;
    p.total = ph.total;
    p.off = c->p_off.as<int64_t>();
    p.sec = c->p_sec.as<uint16_t>();
    p.sec_elems = c->p_elems.as<int16_t>();
    p.sec_pos = c->p_pos.as<int16_t>();
    p.block_ado = c->p_block.as<int32_t>();
    p.rowptr = (w == 0 ? c->fwd.rowptr : c->trn.rowptr).as<int64_t>();
    p.ent = (w == 0 ? c->p_entf : c->p_entt).as<DevEntryP>();
    p.trace_slot = ph.tslot;
    p.maps = c->p_maps.as<int32_t>();
    p.n_heavy = ph.n_heavy;
    p.light_begin = ph.light_begin;
    p.light_block_ado = c->p_lblock.as<int32_t>();
    p.light_nz = c->p_lnz.as<int32_t>();
    p.max_light_couplings = ph.max_light[w];
    p.n_lt_blocks = (int)ph.lt_blk.size();
    p.lt_u = ph.lt_u;
    // the default form: the sector-batched top tier with the sliced heavy
    // rows (the fastest on both test workloads' forward applies); the
    // autotuner (heomq_ctx_autotune) picks per operator, HEOMQ_APPLY_FORM
    // forces one
    const char* form = std::getenv("HEOMQ_APPLY_FORM");
    set_apply_form(c, w, form ? std::atoi(form) : 2);
  }
  return HEOMQ_OK;
}

// Form of the packed apply of operator w (0 forward, 1 transposed): bit 0
// the grouped heavy rows, bit 1 the sector-batched top tier (0: the sliced
// heavy and packed-order light forms), bit 2 one wave per heavy ADO (small
// sectors only).  Every form computes the same operator (to rounding); a form
// whose tables were not built falls back per half.
void set_apply_form(heomq_ctx* c, int w, int form) {
  DevPacked& p = w == 0 ? c->pfwd : c->ptrn;
  const bool wave = (form & 4) != 0 && c->heavy_wave_ok;
  p.heavy_wave = wave ? 1 : 0;
  const bool grouped = (form & 1) != 0, sector = (form & 2) != 0 && c->p_ltblk.p != nullptr;
  p.hg_off = grouped ? (w == 0 ? c->p_hgoff_f : c->p_hgoff_t).as<int32_t>() : nullptr;
  p.hg = grouped ? (w == 0 ? c->p_hg_f : c->p_hg_t).as<int4>() : nullptr;
  p.lt_blk = sector ? c->p_ltblk.as<int4>() : nullptr;
  p.lt_off = sector ? c->p_ltoff.as<int32_t>() : nullptr;
  p.lt_id = sector ? c->p_ltados.as<int32_t>() : nullptr;
  p.lt_ent = sector ? (w == 0 ? c->p_ltent_f : c->p_ltent_t).as<DevEntryP>() : nullptr;
  c->apply_form[w] = (grouped ? 1 : 0) | (sector ? 2 : 0) | (wave ? 4 : 0);
}

// coefficient pointers of the packed operators follow the dense ones
void sync_packed(heomq_ctx* c) {
  for (int w = 0; w < 2; ++w) {
    DevPacked& p = w == 0 ? c->pfwd : c->ptrn;
    const DevOperator& o = w == 0 ? c->fwd.op : c->trn.op;
    p.rec = o.rec;
    p.gam = o.gam;
    p.opmap = o.opmap;
    p.hnnz = o.hnnz;
    p.hcol = o.hcol;
    p.hval = o.hval;
    p.nop = o.nop;
    p.d = o.d;
    p.hsign = o.hsign;
    p.transposed = o.transposed;
  }
}

int64_t trace_pos(const heomq_ctx* c, int trace) {
  if (trace == HEOMQ_TRACE_LAST) return c->N - 1;
  if (trace == HEOMQ_TRACE_RDO) return 0;
  return -1;
}

}  // namespace

namespace heomq {

int build_spa_h(heomq_ctx* c) {
  SpaOperator& so = c->spa;
  HTables ht;
  // the fixed-width rows are kept while their table stays under 256 MiB
  build_h_tables(c->model.H, c->d, so.nsec, so.elems_h, so.pos_h, so.nnz_h, (size_t)256 << 20, ht);
  HIPCHECK(upload(so.hrow, ht.hrow));
  HIPCHECK(upload(so.hq, ht.hq));
  HIPCHECK(upload(so.hc, ht.hc));
  HIPCHECK(upload(so.hq_ell, ht.hq_ell));
  HIPCHECK(upload(so.hc_ell, ht.hc_ell));
  DevPacked& p = so.op;
  p.hrow = so.hrow.as<int32_t>();
  p.hq = so.hq.as<int16_t>();
  p.hc = so.hc.as<double>();
  p.hell_w = ht.ew;
  p.hell_wp = ht.ewp;
  p.hq_ell = so.hq_ell.as<int16_t>();
  p.hc_ell = so.hc_ell.as<double>();
  return HEOMQ_OK;
}

// The pattern's sectors are its distinct per-ADO element lists (in storage
// order, so rho_spa is the packed vector of this layout as it stands); every
// table of the packed apply then follows as for the symmetry sectors, with
// neighbour elements outside a neighbour's list absent from the gather maps
// (P^T: they are zero) and output elements outside the list never formed (P).
int build_spa_operator(heomq_ctx* c, const int32_t* nnz_spa, const std::vector<int64_t>& start,
                       const int32_t* irow, const int32_t* icol) {
  const Model& m = c->model;
  const int d = m.d, dd = d * d;
  PackedHost ph;
  std::map<std::vector<int16_t>, int> cls;
  std::vector<const std::vector<int16_t>*> lists;
  ph.sec.assign(m.N, 0);
  std::vector<int16_t> key;
  for (int64_t a = 0; a < m.N; ++a) {
    const int nz = nnz_spa[a];
    key.resize(nz);
    for (int e = 0; e < nz; ++e)
      key[e] = (int16_t)((irow[start[a] + e] - 1) + (icol[start[a] + e] - 1) * d);
    auto it = cls.find(key);
    if (it == cls.end()) {
      if ((int)lists.size() >= kMaxSectors)
        return fail(HEOMQ_UNSUPPORTED, "sparse pattern: more than 65535 distinct per-ADO element lists");
      it = cls.emplace(key, (int)lists.size()).first;
      lists.push_back(&it->first);
    }
    ph.sec[a] = (uint16_t)it->second;
  }
  const int S = (int)lists.size();
  if ((size_t)S * dd * 4 > ((size_t)1 << 30))
    return fail(HEOMQ_UNSUPPORTED, "sparse pattern: per-list tables over 1 GiB (too many distinct element lists)");
  ph.S = S;
  ph.elems.assign((size_t)S * dd, -1);
  ph.pos.assign((size_t)S * dd, -1);
  ph.nnz.assign(S, 0);
  for (int s = 0; s < S; ++s) {
    const auto& l = *lists[s];
    ph.nnz[s] = (int)l.size();
    for (size_t t = 0; t < l.size(); ++t) {
      ph.elems[(size_t)s * dd + t] = l[t];
      ph.pos[(size_t)s * dd + l[t]] = (int16_t)t;
    }
  }
  std::vector<int64_t> rowf, rowt;
  std::vector<DevEntry> entf, entt;
  if (int rc = build_entries(m, rowf, entf, rowt, entt)) return rc;
  if (int rc = finish_packed_host(m, rowf, entf, rowt, entt, ph, false)) return rc;
  SpaOperator& so = c->spa;
  so.built = false;
  HIPCHECK(upload(so.off, ph.off));
  HIPCHECK(upload(so.sec, ph.sec));
  HIPCHECK(upload(so.elems, ph.elems));
  HIPCHECK(upload(so.pos, ph.pos));
  HIPCHECK(upload(so.block, ph.block));
  HIPCHECK(upload(so.lblock, ph.lblock));
  HIPCHECK(upload(so.lnz, ph.lnz));
  HIPCHECK(upload(so.ent, ph.pentf));
  HIPCHECK(upload(so.maps, ph.maps));
  HIPCHECK(upload(so.hgoff, ph.hg_off[0]));
  HIPCHECK(upload(so.hg, ph.hg[0]));
  // the top tier batched by element list when the blocks stay mostly full
  // (a pattern with few distinct lists), else in rho_spa order
  int64_t top = ph.total - ph.light_begin;
  const bool batched = !ph.lt_blk.empty() && top * 2 >= (int64_t)ph.lt_blk.size() * 256;
  if (batched) {
    HIPCHECK(upload(so.ltblk, ph.lt_blk));
    HIPCHECK(upload(so.ltados, ph.lt_ados));
    HIPCHECK(upload(so.ltoff, ph.lt_off));
    HIPCHECK(upload(so.ltent, ph.lt_ent[0]));
  }
  so.nsec = S;
  so.elems_h = std::move(ph.elems);
  so.pos_h = std::move(ph.pos);
  so.nnz_h = std::move(ph.nnz);
  DevPacked p = c->pfwd;
  p.n_ado = m.N;
  p.total = ph.total;
  p.off = so.off.as<int64_t>();
  p.sec = so.sec.as<uint16_t>();
  p.sec_elems = so.elems.as<int16_t>();
  p.sec_pos = so.pos.as<int16_t>();
  p.block_ado = so.block.as<int32_t>();
  p.rowptr = c->fwd.rowptr.as<int64_t>();
  p.ent = so.ent.as<DevEntryP>();
  p.trace_slot = -1;
  p.maps = so.maps.as<int32_t>();
  p.n_heavy = ph.n_heavy;
  p.light_begin = ph.light_begin;
  p.light_block_ado = so.lblock.as<int32_t>();
  p.light_nz = so.lnz.as<int32_t>();
  p.max_light_couplings = ph.max_light[0];
  p.hg_off = so.hgoff.as<int32_t>();
  p.hg = so.hg.as<int4>();
  p.n_lt_blocks = batched ? (int)ph.lt_blk.size() : 0;
  p.lt_u = ph.lt_u;
  p.lt_blk = batched ? so.ltblk.as<int4>() : nullptr;
  p.lt_off = batched ? so.ltoff.as<int32_t>() : nullptr;
  p.lt_id = batched ? so.ltados.as<int32_t>() : nullptr;
  p.lt_ent = batched ? so.ltent.as<DevEntryP>() : nullptr;
  p.precond = 0;
  p.herm0 = 0;
  p.nprob = 1;
  p.vstride = p.rec_stride = p.gam_stride = p.hc_stride = 0;
  p.frozen_flag = nullptr;
  p.heavy_wave = ph.n_heavy > 1 ? 1 : 0;   // the pattern's heavy ADOs all small: one wave each
  for (int64_t a = 1; a < ph.n_heavy && p.heavy_wave; ++a)
    if (ph.off[a + 1] - ph.off[a] > kHeavyWaveMaxNnz) p.heavy_wave = 0;
  so.op = p;
  if (int rc = build_spa_h(c)) return rc;
  so.trace_slot = so.pos_h[(size_t)ph.sec[0] * dd + (d - 1) + (size_t)(d - 1) * d];
  // the preconditioner of sparse steady solves, where the lists allow it
  {
    const std::string keep = last_error();
    so.has_jac = build_tier_jacobi(m, S, so.elems_h, so.pos_h, so.nnz_h, ph.off, 0, so.jblk, so.jdesc, so.ju, so.je,
                                   so.jac) == HEOMQ_OK;
    last_error() = keep;
  }
  so.built = true;
  return HEOMQ_OK;
}

DevPacked spa_operator(const heomq_ctx* c, const double* gam) {
  DevPacked p = c->spa.op;
  const DevPacked& f = c->pfwd;
  p.rec = f.rec;
  p.gam = gam ? gam : f.gam;
  p.opmap = f.opmap;
  p.hnnz = f.hnnz;
  p.hcol = f.hcol;
  p.hval = f.hval;
  p.nop = f.nop;
  p.d = f.d;
  p.hsign = f.hsign;
  p.transposed = false;
  return p;
}

}  // namespace heomq

namespace heomq {

// The derivative terminator's tables (calc_ax_tfqmr_omp.f90:362-478): the
// hierarchy one tier higher, built from the same parameters.  Its ADOs of
// tiers 0..L are the context's (same tuples, same order, same packed
// elements, checked here), so the context's packed vectors are the prefix of
// its packed vectors; the apply eliminates its top tier in every call.
int build_trun(heomq_ctx* c) {
  TrunOperator& tr = c->trun;
  tr.built = false;
  const Model& m = c->model;
  if (m.L < 1) return fail(HEOMQ_UNSUPPORTED, "derivative terminator: needs a top tier L >= 1");
  if (!m.from_params || m.p.ntop != m.L)
    return fail(HEOMQ_UNSUPPORTED, "derivative terminator: needs a model built from parameters (heomq_model_build)");
  heomq_params px = m.p;
  px.ntop = m.L + 1;
  Model mx;
  if (int rc = build_model(px, mx)) return rc;
  for (int n = 0; n <= m.L; ++n)
    if (mx.first[n] != m.first[n] || mx.tiers[n] != m.tiers[n])
      return fail(HEOMQ_UNSUPPORTED, "derivative terminator: the extended hierarchy's lower tiers differ");
  if (mx.first[m.L + 1] != m.N) return fail(HEOMQ_UNSUPPORTED, "derivative terminator: tier sizes differ");
  PackedHost ph;
  if (int rc = symmetry_sectors(mx, ph)) return rc;
  std::vector<int64_t> rowf, rowt;
  std::vector<DevEntry> entf, entt;
  if (int rc = build_entries(mx, rowf, entf, rowt, entt)) return rc;
  if (int rc = finish_packed_host(mx, rowf, entf, rowt, entt, ph, false)) return rc;
  const int dd = m.d * m.d;
  for (int64_t a = 0; a < m.N; ++a) {
    const int s = ph.sec[a], s0 = c->psec_h[a];
    if (ph.off[a] != c->poff_h[a] || ph.nnz[s] != c->sec_nnz_h[s0] ||
        !std::equal(ph.elems.begin() + (size_t)s * dd, ph.elems.begin() + (size_t)s * dd + ph.nnz[s],
                    c->sec_elems_h.begin() + (size_t)s0 * dd))
      return fail(HEOMQ_UNSUPPORTED, "derivative terminator: the extended packed layout's prefix differs");
  }
  HIPCHECK(upload(tr.off, ph.off));
  HIPCHECK(upload(tr.sec, ph.sec));
  HIPCHECK(upload(tr.elems, ph.elems));
  HIPCHECK(upload(tr.pos, ph.pos));
  HIPCHECK(upload(tr.block, ph.block));
  HIPCHECK(upload(tr.lblock, ph.lblock));
  HIPCHECK(upload(tr.lnz, ph.lnz));
  HIPCHECK(upload(tr.rowptr, rowf));
  HIPCHECK(upload(tr.ent, ph.pentf));
  HIPCHECK(upload(tr.maps, ph.maps));
  HIPCHECK(upload(tr.hgoff, ph.hg_off[0]));
  HIPCHECK(upload(tr.hg, ph.hg[0]));
  const bool batched = !ph.lt_blk.empty();
  if (batched) {
    HIPCHECK(upload(tr.ltblk, ph.lt_blk));
    HIPCHECK(upload(tr.ltados, ph.lt_ados));
    HIPCHECK(upload(tr.ltoff, ph.lt_off));
    HIPCHECK(upload(tr.ltent, ph.lt_ent[0]));
  }
  std::vector<double> gam;
  build_gamma(mx, mx.gdraw, gam, false);
  HIPCHECK(upload(tr.gam, gam));
  HTables ht;
  build_h_tables(mx.H, mx.d, ph.S, ph.elems, ph.pos, ph.nnz, SIZE_MAX, ht);
  HIPCHECK(upload(tr.hrow, ht.hrow));
  HIPCHECK(upload(tr.hq, ht.hq));
  HIPCHECK(upload(tr.hc, ht.hc));
  HIPCHECK(upload(tr.hq_ell, ht.hq_ell));
  HIPCHECK(upload(tr.hc_ell, ht.hc_ell));
  if (int rc = build_tier_jacobi(mx, ph.S, ph.elems, ph.pos, ph.nnz, ph.off, m.N, tr.jblk, tr.jdesc, tr.ju, tr.je,
                                 tr.jac))
    return rc;
  // coefficient records: the extended hierarchy has the context's drawers and bath
  DevPacked p = c->pfwd;
  p.n_ado = mx.N;
  p.total = ph.total;
  p.off = tr.off.as<int64_t>();
  p.sec = tr.sec.as<uint16_t>();
  p.sec_elems = tr.elems.as<int16_t>();
  p.sec_pos = tr.pos.as<int16_t>();
  p.block_ado = tr.block.as<int32_t>();
  p.rowptr = tr.rowptr.as<int64_t>();
  p.ent = tr.ent.as<DevEntryP>();
  p.gam = tr.gam.as<double>();
  p.trace_slot = ph.tslot;
  p.maps = tr.maps.as<int32_t>();
  p.n_heavy = ph.n_heavy;
  p.light_begin = ph.light_begin;
  p.light_block_ado = tr.lblock.as<int32_t>();
  p.light_nz = tr.lnz.as<int32_t>();
  p.max_light_couplings = ph.max_light[0];
  p.hg_off = tr.hgoff.as<int32_t>();
  p.hg = tr.hg.as<int4>();
  p.n_lt_blocks = batched ? (int)ph.lt_blk.size() : 0;
  p.lt_u = ph.lt_u;
  p.lt_blk = batched ? tr.ltblk.as<int4>() : nullptr;
  p.lt_off = batched ? tr.ltoff.as<int32_t>() : nullptr;
  p.lt_id = batched ? tr.ltados.as<int32_t>() : nullptr;
  p.lt_ent = batched ? tr.ltent.as<DevEntryP>() : nullptr;
  p.hrow = tr.hrow.as<int32_t>();
  p.hq = tr.hq.as<int16_t>();
  p.hc = tr.hc.as<double>();
  p.hell_w = ht.ew;
  p.hell_wp = ht.ewp;
  p.hq_ell = tr.hq_ell.as<int16_t>();
  p.hc_ell = tr.hc_ell.as<double>();
  p.precond = 0;
  p.herm0 = 0;
  p.nprob = 1;
  p.vstride = p.rec_stride = p.gam_stride = p.hc_stride = 0;
  p.frozen_flag = nullptr;
  if (p.trace_slot != c->pfwd.trace_slot) return fail(HEOMQ_UNSUPPORTED, "derivative terminator: trace slot differs");
  tr.op = p;
  HIPCHECK(tr.xext.alloc(sizeof(double) * 2 * (size_t)ph.total));
  HIPCHECK(tr.text.alloc(sizeof(double) * 2 * (size_t)ph.total));
  tr.total = ph.total;
  tr.vis = c->ptotal;
  tr.built = true;
  return HEOMQ_OK;
}

}  // namespace heomq


extern "C" {

const char* heomq_last_error(void) { return last_error().c_str(); }
int heomq_version(void) { return 10000; }

int heomq_model_build(const heomq_params* p, heomq_model** out) {
  if (!p || !out) return fail(HEOMQ_BAD_ARG, "null argument");
  auto* h = new heomq_model;
  const int rc = build_model(*p, h->m);
  if (rc) { delete h; return rc; }
  *out = h;
  return HEOMQ_OK;
}

void heomq_model_free(heomq_model* m) { delete m; }

int heomq_model_info(const heomq_model* h, int64_t* info, int64_t* first) {
  if (!h) return fail(HEOMQ_BAD_ARG, "null model");
  const Model& m = h->m;
  if (info) {
    info[0] = m.d; info[1] = m.M; info[2] = m.L; info[3] = m.N;
    info[4] = m.ncor; info[5] = m.nopr;
    info[6] = (int64_t)m.lower.size();
    int64_t nu = 0;
    for (const auto& e : m.upper) nu += e.nbr >= 0;
    info[7] = nu;
  }
  if (first) for (int n = 0; n <= m.L + 1; ++n) first[n] = m.first[n];
  return HEOMQ_OK;
}

int heomq_model_coefs(const heomq_model* h, double* cb, double* cd, double* gam) {
  if (!h) return fail(HEOMQ_BAD_ARG, "null model");
  const Model& m = h->m;
  for (int k = 0; k < m.M; ++k) {
    if (cb) { cb[2 * k] = m.cb[k].real(); cb[2 * k + 1] = m.cb[k].imag(); }
    if (cd) { cd[2 * k] = m.cd[k].real(); cd[2 * k + 1] = m.cd[k].imag(); }
    if (gam) { gam[2 * k] = m.gdraw[k].real(); gam[2 * k + 1] = m.gdraw[k].imag(); }
  }
  return HEOMQ_OK;
}

int heomq_model_operators(const heomq_model* h, double* H, double* ams) {
  if (!h) return fail(HEOMQ_BAD_ARG, "null model");
  const Model& m = h->m;
  if (H) for (size_t t = 0; t < m.H.size(); ++t) { H[2 * t] = m.H[t].real(); H[2 * t + 1] = m.H[t].imag(); }
  if (ams) std::memcpy(ams, m.ams.data(), m.ams.size() * sizeof(double));
  return HEOMQ_OK;
}

int heomq_model_tier(const heomq_model* h, int tier, int32_t* tuples) {
  if (!h || tier < 0 || tier > h->m.L) return fail(HEOMQ_BAD_ARG, "bad tier");
  const auto& t = h->m.tiers[tier];
  if (tuples && !t.empty()) std::memcpy(tuples, t.data(), t.size() * sizeof(int32_t));
  return HEOMQ_OK;
}

int heomq_model_coupling(const heomq_model* h, int32_t* lower, int32_t* upper) {
  if (!h) return fail(HEOMQ_BAD_ARG, "null model");
  const Model& m = h->m;
  if (lower)
    for (size_t p = 0; p < m.lower.size(); ++p) {
      lower[4 * p] = m.lower[p].nbr; lower[4 * p + 1] = m.lower[p].itype;
      lower[4 * p + 2] = m.lower[p].ff; lower[4 * p + 3] = m.lower[p].fr;
    }
  if (upper)
    for (size_t p = 0; p < m.upper.size(); ++p) {
      upper[4 * p] = m.upper[p].nbr; upper[4 * p + 1] = m.upper[p].itype;
      upper[4 * p + 2] = m.upper[p].ff; upper[4 * p + 3] = m.upper[p].fr;
    }
  return HEOMQ_OK;
}

int heomq_model_packed_layout(const heomq_model* h, int64_t* info) {
  if (!h || !info) return fail(HEOMQ_BAD_ARG, "null argument");
  const Model& m = h->m;
  std::vector<int64_t> rowf, rowt;
  std::vector<DevEntry> entf, entt;
  if (int rc = build_entries(m, rowf, entf, rowt, entt)) return rc;
  PackedHost ph;
  if (int rc = build_packed_host(m, rowf, entf, rowt, entt, ph)) return rc;
  int64_t empty = 0, worst = 0;
  for (int64_t a = 0; a < m.N; ++a) empty += ph.off[a + 1] == ph.off[a];
  for (size_t b = 0; b + 1 < ph.lblock.size(); ++b) worst = std::max<int64_t>(worst, ph.lblock[b + 1] - ph.lblock[b] + 1);
  info[0] = ph.total;
  info[1] = ph.S;
  info[2] = ph.cons_n;
  info[3] = ph.cons_s;
  info[4] = m.N * (int64_t)m.d * m.d;
  info[5] = ph.n_heavy;
  info[6] = ph.light_begin;
  info[7] = (int64_t)ph.lblock.size() - 1;
  info[8] = worst;
  info[9] = empty;
  info[10] = ph.max_light[0];
  info[11] = (int64_t)ph.maps.size();
  return HEOMQ_OK;
}

int heomq_spa_build_pattern(const heomq_model* h, double dsmall, uint64_t seed, int64_t cap, int64_t* lunk_spa,
                            int32_t* nnz_spa, int64_t* ind_spa, int32_t* irow_spa, int32_t* icol_spa) {
  if (!h || !lunk_spa) return fail(HEOMQ_BAD_ARG, "null argument");
  if (!(dsmall >= 0.0)) return fail(HEOMQ_BAD_ARG, "dsmall must be >= 0");
  std::vector<int32_t> nnz, ir, ic;
  if (int rc = spa_pattern_build(h->m, dsmall, seed, nnz, ir, ic)) return rc;
  *lunk_spa = (int64_t)ir.size();
  if (nnz_spa) std::memcpy(nnz_spa, nnz.data(), nnz.size() * sizeof(int32_t));
  if (ind_spa) {
    int64_t run = 1;
    for (size_t a = 0; a < nnz.size(); ++a) { ind_spa[a] = run; run += nnz[a]; }
  }
  if (!irow_spa || !icol_spa) return HEOMQ_OK;
  if (cap < (int64_t)ir.size()) return fail(HEOMQ_BAD_DIM, "sparse pattern: irow_spa / icol_spa shorter than lunk_spa");
  std::memcpy(irow_spa, ir.data(), ir.size() * sizeof(int32_t));
  std::memcpy(icol_spa, ic.data(), ic.size() * sizeof(int32_t));
  return HEOMQ_OK;
}

int heomq_model_set_params(heomq_model* h, const heomq_params* p) {
  if (!h || !p) return fail(HEOMQ_BAD_ARG, "null argument");
  return set_params(h->m, *p);
}

int heomq_model_from_reference(
    int32_t nrho, int64_t nunk, int32_t ntier, int32_t norbs, int32_t nspin, int32_t nalf,
    int32_t ncor, int32_t igroundsteady, int64_t ntable, int64_t lall, const int32_t* indextable,
    const int64_t* nfirst, const int64_t* nlast, const int64_t* ifirst, const int64_t* indexcoef, const int8_t* itypecoef,
    const int8_t* ifactfcoef, const int8_t* ifactrcoef, const int64_t* index_coef_ref,
    const int32_t* indexdraw, const double* cb, const double* cd, const double* cgamma,
    const double* dipm, const int32_t* ilead, const int32_t* mspin, const int32_t* morbs,
    const int32_t* mpm, const double* eleadinfty, const double* hs, const double* amsall, const double* dbsqrt,
    heomq_model** out) {
  if (!out) return fail(HEOMQ_BAD_ARG, "null output");
  auto* h = new heomq_model;
  const int rc = model_from_reference(nrho, nunk, ntier, norbs, nspin, nalf, ncor, igroundsteady, ntable,
                                      lall, indextable, nfirst, nlast, ifirst, indexcoef, itypecoef,
                                      ifactfcoef, ifactrcoef, index_coef_ref, indexdraw, cb, cd, cgamma,
                                      dipm, ilead, mspin, morbs, mpm, eleadinfty, hs, amsall, dbsqrt,
                                      h->m);
  if (rc) { delete h; return rc; }
  *out = h;
  return HEOMQ_OK;
}

int heomq_ctx_create(const heomq_model* h, int device, heomq_ctx** out) {
  if (!h || !out) return fail(HEOMQ_BAD_ARG, "null argument");
  const Model& m = h->m;
  if (m.d != 4 && m.d != 16 && m.d != 64) return fail(HEOMQ_UNSUPPORTED, "d must be 4, 16 or 64");
  HIPCHECK(hipSetDevice(device));
  auto c = std::make_unique<heomq_ctx>();
  c->device = device;
  c->model = m;
  c->d = m.d; c->M = m.M; c->L = m.L; c->N = m.N;
  c->nspin = m.nspin; c->norbs = m.norbs; c->nalf = m.nalf;
  c->first = m.first;
  c->sgn = m.sgn; c->spin = m.spin; c->orb = m.orb; c->alf = m.alf;
  c->ams = m.ams;
  c->dbsqrt = m.dbsqrt.empty() ? std::vector<double>(m.M, 1.0) : m.dbsqrt;
  c->lscale = m.lscale;
  c->tiers = m.tiers;
  c->steady_model = m.steady;
  c->field_amp = m.eleadinfty;
  // tier-1 position of each sigma=+ drawer (td/calcj.f90 reads these ADOs)
  c->tier1_of.assign(m.M, -1);
  c->tier1_dag.assign(m.M, 0);
  c->tier1_fact.assign(m.M, 1);
  if (m.L >= 1)
    for (int k = 0; k < m.M; ++k) {
      const UpperEntry& e = m.upper[k];   // tier-0 upper entries are the tier-1 lookups
      c->tier1_of[k] = e.nbr;
      c->tier1_dag[k] = e.itype;
      c->tier1_fact[k] = e.ff;
    }
  std::vector<int64_t> rowf, rowt;
  std::vector<DevEntry> entf, entt;
  if (int rc = build_entries(m, rowf, entf, rowt, entt)) return rc;
  HIPCHECK(upload(c->fwd.rowptr, rowf));
  HIPCHECK(upload(c->fwd.ent, entf));
  HIPCHECK(upload(c->trn.rowptr, rowt));
  HIPCHECK(upload(c->trn.ent, entt));
  std::vector<int8_t> map;
  build_opmap(m, map);
  HIPCHECK(upload(c->opmap, map));
  if (int rc = upload_coefs(c.get(), m)) return rc;
  {
    std::vector<int8_t> par(m.d);
    for (int i = 0; i < m.d; ++i) {
      int ne = 0, v = i;
      for (int o = 0; o < m.norbs; ++o, v /= 4) ne += (v % 4 == 1) + (v % 4 == 2) + 2 * (v % 4 == 3);
      par[i] = (int8_t)(ne & 1);
    }
    HIPCHECK(upload(c->parity, par));
    c->tb.ntier = m.L + 1;
    for (int n = 0; n <= m.L + 1 && n < 12; ++n) c->tb.first[n] = m.first[n];
  }
  fill_op(c.get(), c->fwd, c->rec, false, -1.0);
  fill_op(c.get(), c->trn, c->rec, true, +1.0);
  // derivative operator shares the forward entries; records/H/gamma set per call
  c->dpar.op = c->fwd.op;
  if (int rc = build_packed(c.get(), m, rowf, entf, rowt, entt)) return rc;
  sync_packed(c.get());
  if (int rc = build_jacobi(c.get(), m)) return rc;
  HIPCHECK(hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault));
  HIPCHECK(hipStreamCreateWithFlags(&c->solve_stream, hipStreamNonBlocking));
  if (int rc = scratch_address(device, &c->scratch_dev)) return rc;   // with `device` current
  HIPCHECK(c->scal.alloc(sizeof(TfqmrScal) * 2));
  HIPCHECK(hipMemset(c->scal.p, 0, sizeof(TfqmrScal) * 2));
  HIPCHECK(c->partial.alloc(sizeof(double) * 2 * kMaxRedBlocks));
  HIPCHECK(c->scal2.alloc(sizeof(TfqmrScal)));
  HIPCHECK(hipMemset(c->scal2.p, 0, sizeof(TfqmrScal)));
  HIPCHECK(c->partial2.alloc(sizeof(double) * 2 * kMaxRedBlocks));
  HIPCHECK(c->red.alloc(sizeof(double) * 4 + sizeof(unsigned) * 4));
  HIPCHECK(hipMemset(c->red.p, 0, c->red.bytes));
  HIPCHECK(hipDeviceSynchronize());
  *out = c.release();
  return HEOMQ_OK;
}

void heomq_ctx_free(heomq_ctx* c) { delete c; }

// The autotuner of the packed apply: every form of each operator is timed
// with HIP events over `reps` back-to-back launches on a trace-one state (the
// operator's own tables, this GPU) and the fastest is kept.
int heomq_ctx_autotune(heomq_ctx* c, int reps, int32_t* forms, double* times_us) {
  if (!c || reps < 1) return fail(HEOMQ_BAD_ARG, "null context or reps < 1");
  if (c->ptotal <= 0) return HEOMQ_OK;
  HIPCHECK(hipSetDevice(c->device));
  c->drop_graphs();   // captured chunks hold the operator's tables by value
  DevBuf xb, yb, tb;
  const size_t bytes = sizeof(double) * 2 * (size_t)c->ptotal;
  HIPCHECK(xb.alloc(bytes));
  HIPCHECK(yb.alloc(bytes));
  hipStream_t st = c->solve_stream;
  HIPCHECK(vec_fill<double>(st, 2 * c->ptotal, xb.as<double>(), 1.0));
  // the forward form is timed on the operator the steady solve iterates
  // with: under block-Jacobi preconditioning (the default where H couples
  // Fock states) the apply and its inverse, as one fused launch where the
  // form allows it
  const int want = c->precond >= 0 ? c->precond : (c->h_diagonal ? 1 : 2);
  DevJacobi jc = c->pjac;
  jc.ejac = c->prec_shift > 0 ? c->prec_shift : 1.0;   // timing only: any positive shift
  const bool prec_fwd = want == 2 && jacobi_one_kernel(jc) && !c->trun.built;
  if (prec_fwd) HIPCHECK(tb.alloc(bytes));
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  int rc = HEOMQ_OK;
  for (int w = 0; w < 2 && rc == HEOMQ_OK; ++w) {
    double best = 1e300;
    int best_f = 2;
    for (int f = 0; f < 8; ++f) {
      if ((f & 4) && !c->heavy_wave_ok) continue;
      set_apply_form(c, w, f);
      const DevPacked& op = w == 0 ? c->pfwd : c->ptrn;
      auto one = [&]() -> hipError_t {
        if (w == 1 || !prec_fwd) return launch_apply_packed_f64(op, xb.as<double>(), yb.as<double>(), true, st);
        if (c->fuse_precond && apply_jac_fusable(op, jc))
          return launch_apply_packed_jac(op, jc, xb.as<double>(), yb.as<double>(), true, st);
        if (hipError_t e = launch_apply_packed_f64(op, xb.as<double>(), tb.as<double>(), true, st)) return e;
        return launch_jacobi_1k(op, jc, tb.as<double>(), yb.as<double>(), false, st);
      };
      hipError_t e = hipSuccess;
      for (int k = 0; k < 2 && e == hipSuccess; ++k) e = one();
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (int k = 0; k < reps && e == hipSuccess; ++k) e = one();
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(HEOMQ_HIP_ERROR, std::string("autotune: ") + hipGetErrorString(e));
        break;
      }
      const double us = 1e3 * ms / reps;
      if (times_us) times_us[8 * w + f] = us;
      if (us < best) { best = us; best_f = f; }
    }
    set_apply_form(c, w, best_f);
    if (forms) forms[w] = c->apply_form[w];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int heomq_ctx_refresh(heomq_ctx* c, const heomq_model* h) {
  if (!c || !h) return fail(HEOMQ_BAD_ARG, "null argument");
  if (h->m.N != c->N || h->m.M != c->M) return fail(HEOMQ_BAD_ARG, "model shape differs from context");
  HIPCHECK(hipSetDevice(c->device));
  c->drop_graphs();   // captured chunks hold table pointers and values' addresses
  if (int rc = build_packed_h(c, h->m)) return rc;   // checks conservation before anything changes
  if (int rc = upload_coefs(c, h->m)) return rc;
  fill_op(c, c->fwd, c->rec, false, -1.0);
  fill_op(c, c->trn, c->rec, true, +1.0);
  sync_packed(c);
  if (int rc = build_jacobi(c, h->m)) return rc;
  c->model = h->m;
  if (c->spa.built)
    if (int rc = build_spa_h(c)) return rc;
  if (c->trun.built)
    if (int rc = build_trun(c)) return rc;
  c->steady_model = h->m.steady;
  c->field_amp = h->m.eleadinfty;
  HIPCHECK(hipDeviceSynchronize());
  return HEOMQ_OK;
}

int heomq_ctx_set_apply_form(heomq_ctx* c, int forward, int transposed) {
  if (!c) return fail(HEOMQ_BAD_ARG, "null argument");
  if (forward < 0 || forward > 7 || transposed < 0 || transposed > 7) return fail(HEOMQ_BAD_ARG, "forms are 0..7");
  c->drop_graphs();   // captured chunks hold the operator by value
  set_apply_form(c, 0, forward);
  set_apply_form(c, 1, transposed);
  return HEOMQ_OK;
}

int heomq_ctx_get_apply_form(const heomq_ctx* c, int32_t* forms) {
  if (!c || !forms) return fail(HEOMQ_BAD_ARG, "null argument");
  forms[0] = c->apply_form[0];
  forms[1] = c->apply_form[1];
  return HEOMQ_OK;
}

int heomq_ctx_set_trun_der(heomq_ctx* c, int enable) {
  if (!c) return fail(HEOMQ_BAD_ARG, "null argument");
  HIPCHECK(hipSetDevice(c->device));
  c->drop_graphs();   // captured chunks hold the apply's kernels
  if (!enable) {
    c->trun.built = false;
    return HEOMQ_OK;
  }
  if (int rc = build_trun(c)) return rc;
  HIPCHECK(hipDeviceSynchronize());
  return HEOMQ_OK;
}

int64_t heomq_ctx_vector_len(const heomq_ctx* c) { return c ? c->N * c->d * c->d : -1; }

int heomq_apply(heomq_ctx* c, const double* x, double* y, int trace, void* stream) {
  if (!c || !x || !y) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_apply: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  HIPCHECK(launch_apply_f64(c->fwd.op, x, y, trace_pos(c, trace), (hipStream_t)stream));
  return HEOMQ_OK;
}

int heomq_apply_transpose(heomq_ctx* c, const double* x, double* y, int trace, void* stream) {
  if (!c || !x || !y) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_apply_transpose: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  HIPCHECK(launch_apply_f64(c->trn.op, x, y, trace_pos(c, trace), (hipStream_t)stream));
  return HEOMQ_OK;
}

int heomq_apply_f32(heomq_ctx* c, const float* x, float* y, int trace, void* stream) {
  if (!c || !x || !y) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_apply_f32: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  HIPCHECK(launch_apply_f32(c->fwd.op, x, y, trace_pos(c, trace), (hipStream_t)stream));
  return HEOMQ_OK;
}

int heomq_apply_dparam(heomq_ctx* c, const double* x, double* y, const double* dH, const double* dcb,
                       const double* dcd, const double* dgam, void* stream) {
  if (!c || !x || !y || !dH || !dcb || !dcd || !dgam) return fail(HEOMQ_BAD_ARG, "null argument");
  if (c && c->trun.built) return fail(HEOMQ_UNSUPPORTED, "heomq_apply_dparam: not with the derivative terminator (it runs in the packed forward apply and the steady TFQMR solve)");
  if (c->lscale)   // the scaling factors |cb|^(1/2) are not affine in the parameters
    return fail(HEOMQ_UNSUPPORTED, "derivative apply: lscale models are not supported");
  const int M = c->M, d = c->d;
  // records: lower-tier kind with derivative coefficients; upper-tier terms
  // (coefficients -i, +i) do not depend on parameters and are switched off
  std::vector<DevRec> rec(6 * (size_t)M, DevRec{});
  const std::complex<double> I(0, 1);
  std::vector<cplx> gd(M);
  for (int k = 0; k < M; ++k) {
    const cplx b(dcb[2 * k], dcb[2 * k + 1]), dd(dcd[2 * k], dcd[2 * k + 1]);
    const cplx cL = -I * b, cR = I * dd;
    const int o = 2 * (c->orb[k] * c->nspin + c->spin[k]) + (c->sgn[k] == 0 ? 1 : 0);
    rec[k] = DevRec{cL.real(), cL.imag(), cR.real(), cR.imag(), o, o, 0, 0};
    rec[M + k] = DevRec{0, 0, 0, 0, -1, -1, 0, 0};
    gd[k] = cplx(dgam[2 * k], dgam[2 * k + 1]);
  }
  HIPCHECK(upload(c->rec_d, rec));
  std::vector<double> gam(2 * (size_t)c->N, 0.0);
  for (int n = 1; n <= c->L; ++n)
    for (int64_t g = c->first[n]; g < c->first[n + 1]; ++g) {
      cplx s(0, 0);
      const int32_t* K = &c->tiers[n][(g - c->first[n]) * n];
      for (int r = 0; r < n; ++r) s += gd[K[r]];
      gam[2 * g] = s.real();
      gam[2 * g + 1] = s.imag();
    }
  HIPCHECK(upload(c->dpar.gam, gam));
  std::vector<cplx> Hd((size_t)d * d);
  for (size_t t = 0; t < Hd.size(); ++t) Hd[t] = cplx(dH[2 * t], dH[2 * t + 1]);
  std::vector<int32_t> nnz, col;
  std::vector<double> val;
  build_hcsr(Hd, d, nnz, col, val);
  HIPCHECK(upload(c->hnnz_d, nnz));
  HIPCHECK(upload(c->hcol_d, col));
  HIPCHECK(upload(c->hval_d, val));
  DevOperator op = c->fwd.op;
  op.rec = c->rec_d.as<DevRec>();
  op.gam = c->dpar.gam.as<double>();
  op.hnnz = c->hnnz_d.as<int32_t>();
  op.hcol = c->hcol_d.as<int32_t>();
  op.hval = c->hval_d.as<double>();
  HIPCHECK(launch_apply_f64(op, x, y, -1, (hipStream_t)stream));
  HIPCHECK(hipStreamSynchronize((hipStream_t)stream));   // tables above are reused by the next call
  return HEOMQ_OK;
}

}  // extern "C"
