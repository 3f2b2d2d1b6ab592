// Host-side HEOM model: bath coefficients, fermion operators, Hamiltonian,
// ADO hierarchy and coupling table.  The device context (heomq_device.hip)
// turns this into GPU tables.
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "heomq.h"

namespace heomq {

using cplx = std::complex<double>;

constexpr double kHbar = 0.658211928;    // include/sizes:23
constexpr double kJconst = 1.602177e5;   // include/sizes:27

struct LowerEntry {   // one lower-tier coupling of an ADO (buildcoefindex.f90:239-285)
  int32_t nbr;        // 0-based ADO index
  int8_t itype;       // 1: neighbour stored as its Hermitian partner (use x^H)
  int8_t ff, fr;      // front / rear signs
  int32_t drawer;     // drawer removed (K[r])
};

struct UpperEntry {   // one upper-tier coupling (buildcoefindex.f90:289-362)
  int32_t nbr;        // -1: absent (drawer already present, unique-index hierarchy)
  int8_t itype, ff, fr;
};

struct Model {
  heomq_params p{};
  bool from_params = false;                  // built by build_model (p describes it); false: loaded tables
  int d = 0, M = 0, L = 0, ncor = 0, nopr = 0, norbs = 0, nspin = 0, nalf = 0;
  int64_t N = 0;
  bool steady = true;
  // drawer attributes (evaluatepara.f90:1701-1772)
  std::vector<int> sgn, spin, orb, alf, cor, bar;
  std::vector<cplx> cb, cd, cgamma, gdraw;   // gdraw = cgamma + i*dipm*eleadinfty
  // lscale: |cb_k|^(1/2) of the unscaled cb (1 without lscale); cb, cd are
  // then the scaled ones and every upper coupling carries this factor
  std::vector<double> dbsqrt;
  bool lscale = false;
  std::vector<double> eleadinfty;           // [alf*nspin + s]
  std::vector<cplx> H;                       // d*d column-major, already / hbar
  std::vector<double> ams;                   // [(m*nspin + s)*d*d + i + j*d]
  std::vector<int64_t> first;                // L+2
  std::vector<std::vector<int32_t>> tiers;   // tier n: N_n * n sorted tuples
  std::vector<LowerEntry> lower;             // ADOs of tiers >= 1 in order, n each
  std::vector<int64_t> lower_off;            // per ADO start into lower (N+1)
  std::vector<UpperEntry> upper;             // ADOs of tiers < L, M each
  int64_t n_upper_ados = 0;

  int tier_of(int64_t a) const;
  double amsv(int m, int s, int i, int j) const { return ams[((size_t)(m * nspin + s) * d + j) * d + i]; }
};

std::string& last_error();
int fail(int code, const std::string& msg);

// Pade decomposition (main/zpout_psd.f90:34-114): eta, xi.
void pade_psd(int n, std::vector<double>& eta, std::vector<double>& xi);
cplx fermi_pade(cplx z, const std::vector<double>& eta, const std::vector<double>& xi);

int build_model(const heomq_params& p, Model& m);
int set_params(Model& m, const heomq_params& p);   // coefficients only
int model_from_reference(
    int32_t nrho, int64_t nunk, int32_t ntier, int32_t norbs, int32_t nspin,
    int32_t nalf, int32_t ncor, int32_t igroundsteady, int64_t ntable, int64_t lall,
    const int32_t* indextable,
    const int64_t* nfirst, const int64_t* nlast, const int64_t* ifirst,
    const int64_t* indexcoef, const int8_t* itypecoef, const int8_t* ifactfcoef,
    const int8_t* ifactrcoef, const int64_t* index_coef_ref, const int32_t* indexdraw,
    const double* cb, const double* cd, const double* cgamma, const double* dipm,
    const int32_t* ilead, const int32_t* mspin, const int32_t* morbs, const int32_t* mpm,
    const double* eleadinfty, const double* hs, const double* amsall, const double* dbsqrt, Model& m);

// The sparse-mode pattern of fermion/prelude_spa.f90:532-749 (module
// sparsemod layout: nnz per ADO, 1-based row / column per element, ADO after
// ADO); seed fixes the mixing coefficients.
int spa_pattern_build(const Model& m, double dsmall, uint64_t seed, std::vector<int32_t>& nnz,
                      std::vector<int32_t>& irow, std::vector<int32_t>& icol);

}  // namespace heomq
