// General chemistry networks: get_rates and get_derivatives
// (src/m_chemistry.f90:672-804) for an arbitrary reaction list, one thread
// per cell, on column-major arrays (cell index fastest).  Every rate is
// formed with the reference's operations in its order; exp and pow are the
// device library's.
#pragma once

#include "afgpu.h"
#include "afgpu_common.h"

namespace afg {

// UC_boltzmann_const / UC_elec_volt and electron_eV_to_K
// (src/m_units_constants.f90:15-39, m_chemistry.f90:686-687)
constexpr double UC_boltzmann_const = 1.380649e-23;
constexpr double CHEM_KB_EV = UC_boltzmann_const / UC_elem_charge;
constexpr double CHEM_EV_TO_K = 2 * UC_elem_charge / (3 * UC_boltzmann_const);

struct ChemArgs {
  const afgpu_reaction_t *r;
  int n_reactions, n_species;
  DTable fld;       // chemtbl_fld (one column per tabulated reaction)
  DTable energy;    // td_tbl's mean-energy column (cols null: absent)
  double tg;        // gas_temperature
  int need_te;
};

__device__ inline double chem_sq(double x) { return x * x; }

// The rate coefficient of reaction R at field fld (Td) and electron
// temperature te (m_chemistry.f90:699-767)
__device__ inline double chem_rate(const afgpu_reaction_t &R, const ChemArgs &a, double fld, double te, int ix,
                                   double w) {
  const double c0 = R.rate_factor;
  const double *c = R.rate_data;
  const double tg = a.tg;
  switch (R.rate_type) {
    case AFGPU_RATE_TABULATED_FIELD:
      return c0 * lt_at(a.fld, R.lookup_col, ix, w);
    case AFGPU_RATE_CONSTANT:
      return c0 * c[0];
    case AFGPU_RATE_LINEAR:
      return c0 * c[0] * (fld - c[1]);
    case AFGPU_RATE_EXP_V1:
      return c0 * c[0] * exp(-chem_sq(c[1] / (c[2] + fld)));
    case AFGPU_RATE_EXP_V2:
      return c0 * c[0] * exp(-chem_sq(fld / c[1]));
    case AFGPU_RATE_K1:
      return c0 * c[0] * pow(300 / te, c[1]);
    case AFGPU_RATE_K3:
      return c0 * (c[0] * chem_sq(CHEM_KB_EV * te + c[1]) - c[2]) * c[3];
    case AFGPU_RATE_K4:
      return c0 * c[0] * pow(tg / 300, c[1]) * exp(-c[2] / tg);
    case AFGPU_RATE_K5:
      return c0 * c[0] * exp(-c[1] / tg);
    case AFGPU_RATE_K6:
      return c0 * c[0] * pow(tg, c[1]);
    case AFGPU_RATE_K7:
      return c0 * c[0] * pow(tg / c[1], c[2]);
    case AFGPU_RATE_K8:
      return c0 * c[0] * pow(300 / tg, c[1]);
    case AFGPU_RATE_K9:
      return c0 * c[0] * exp(-c[1] * tg);
    case AFGPU_RATE_K10:
      return c0 * pow(10.0, c[0] + c[1] * (tg - 300));
    case AFGPU_RATE_K11:
      return c0 * c[0] * pow(300 / tg, c[1]) * exp(-c[2] / tg);
    case AFGPU_RATE_K12:
      return c0 * c[0] * pow(tg, c[1]) * exp(-c[2] / tg);
    case AFGPU_RATE_K13:
      return c0 * c[0] * exp(-pow(c[1] / (c[2] + fld), c[3]));
    case AFGPU_RATE_K14:
      return c0 * c[0] * exp(-pow(fld / c[1], c[2]));
    case AFGPU_RATE_K15:
      return c0 * c[0] * exp(-pow(c[1] / (UC_boltzmann_const * (tg + fld / c[2])), c[3]));
    case AFGPU_RATE_K16:
      return c0 * c[0] * pow(300 / te, c[1]) * exp(-c[2] / tg) * exp(c[3] * (te - tg) / (te * tg));
    default:
      return 0;  // (rejected when the network is created)
  }
}

// mode 0: rates[n][i] = get_rates; 1: get_derivatives from rates (rates
// become the reaction rates); 2: both, without the rates array.  Dynamic
// LDS: the derivatives of the workgroup's BS cells ([n_species][BS]) and,
// with STAGE, their densities (read from HBM once per cell and species).
// RLDS: the reaction records are copied to LDS first (after the species
// blocks), else read from memory.
template <int BS, bool STAGE, bool RLDS>
__global__ void __launch_bounds__(BS) k_chem(ChemArgs a, int n_cells, const double *__restrict__ fields,
                                             const double *__restrict__ dens, double *rates, double *derivs,
                                             int mode) {
  extern __shared__ double lds[];
  double *sd = lds;
  double *sn = lds + (size_t)a.n_species * BS;
  afgpu_reaction_t *sr = (afgpu_reaction_t *)(lds + (size_t)(STAGE ? 2 : 1) * a.n_species * BS);
  const int t = threadIdx.x;
  if (RLDS) {
    constexpr int W = sizeof(afgpu_reaction_t) / sizeof(int);
    const int *src = (const int *)a.r;
    int *dst = (int *)sr;
    for (int q = t; q < a.n_reactions * W; q += BS) dst[q] = src[q];
    __syncthreads();
  }
  const int i = blockIdx.x * BS + t;
  const bool in = i < n_cells;
  const size_t n = (size_t)n_cells;
  if (mode != 0)
    for (int s = 0; s < a.n_species; s++) {
      sd[s * BS + t] = 0.0;
      if (STAGE) sn[s * BS + t] = in ? dens[s * n + i] : 0.0;
    }
  double fld = 0, te = 0, w = 1;
  int ix = 1;
  if (in && mode != 1) {
    fld = fields[i];
    lt_loc(a.fld, fld, ix, w);
    if (a.need_te) {
      int ie;
      double we;
      lt_loc(a.energy, fld, ie, we);
      te = CHEM_EV_TO_K * lt_at(a.energy, 0, ie, we);
    }
  }
  if (in)
    for (int r = 0; r < a.n_reactions; r++) {
      const afgpu_reaction_t &R = RLDS ? sr[r] : a.r[r];
      double k = (mode == 1) ? rates[r * n + i] : chem_rate(R, a, fld, te, ix, w);
      if (mode == 0) {
        rates[r * n + i] = k;
        continue;
      }
      double prod = 1.0;
      for (int q = 0; q < R.n_in; q++) prod = prod * (STAGE ? sn[R.ix_in[q] * BS + t] : dens[R.ix_in[q] * n + i]);
      k = k * prod;
      if (mode == 1) rates[r * n + i] = k;
      for (int q = 0; q < R.n_in; q++) {
        double &d = sd[R.ix_in[q] * BS + t];
        d = d - k;
      }
      for (int q = 0; q < R.n_out; q++) {
        double &d = sd[R.ix_out[q] * BS + t];
        d = d + k * R.multiplicity_out[q];
      }
    }
  if (in && mode != 0)
    for (int s = 0; s < a.n_species; s++) derivs[s * n + i] = sd[s * BS + t];
}

}  // namespace afg
