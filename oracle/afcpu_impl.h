// Internal declarations of the CPU oracle (test infrastructure only).
//
// The oracle is a C++/OpenMP restatement of afivo-streamer's per-time-step
// hot path, written box by box in the reference's own loop structure and
// operation order so that it can serve as the FP64 parity oracle for the
// HIP path and as the timed CPU baseline.  Only tests/, the smoke check and
// the cpu_baseline leg of bench.py may load it.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "afgpu.h"

struct afcpu_chem;  // a reaction list (afcpu_chem_create)

namespace afcpu {

constexpr double UC_eps0 = 8.8541878128e-12;       // src/m_units_constants.f90:12
constexpr double UC_elem_charge = 1.602176634e-19;  // src/m_units_constants.f90:15
constexpr double SI_to_Townsend = 1e21;             // src/m_gas.f90:39

enum BcType { BC_DIRICHLET = 0, BC_NEUMANN = 1, BC_DIRICHLET_COPY = 2 };
enum RbType { RB_NONE = 0, RB_MG = 1, RB_INTERP = 2, RB_INTERP_LIM = 3 };
enum LimType { LIM_MC = 0, LIM_GMINMOD43 = 1 };

struct Box {
  int lvl = 0;
  int ix[3] = {0, 0, 0};
  int parent = 0;
  int children[8] = {0};
  int nb[6] = {0};
  int nbmat[27] = {0};
  double r_min[3] = {0, 0, 0};
  double dr[3] = {0, 0, 0};
};

struct VarMethods {
  int bc_type[6];
  double bc_val[6];
  int rb;
  int limiter;
  bool ok;
};

struct Table {
  int n = 0;
  double x_min = 0, inv_fac = 0;
  std::vector<double> cols;  // [col][row]
  double col(int c, double x) const;
  void loc(double x, int &ix, double &w) const;
  double at(int c, int ix, double w) const;
};

struct Operator {
  double lambda2 = 0;
  bool have_factors = false;
  std::vector<double> vinv, v, eig;
};

struct Ctx {
  int D = 2, nc = 8, coord = AFGPU_XYZ;
  int ng = 10;
  size_t ncell = 100, nface = 81;
  int n_boxes = 0, highest_lvl = 0;
  std::vector<Box> boxes;                 // index id-1
  std::vector<std::vector<int>> ids, leaves, parents;  // index lvl (0 unused)
  std::vector<std::vector<double>> cc;    // [var][box*ncell]
  std::vector<std::vector<double>> fc;    // [var][box*D*nface]
  afgpu_params_t p{};
  bool have_params = false;
  Table td, chem;
  const ::afcpu_chem *net = nullptr;  // a reaction list in place of the standard model
  // species: 0-2 in the pools of the electrons and ions, further ones of a
  // reaction list in xs[(k - 3) * 3 + state] (gen_sp: the general update)
  int n_sp = 3;
  bool gen_sp = false;
  std::vector<int> sp_charge{-1, 1, -1};
  std::vector<std::vector<double>> xs;
  double *SP(int k, int state, int id) {
    return k < 3 ? C(3 * k + state, id) : &xs[(size_t)(k - 3) * 3 + state][(size_t)(id - 1) * ncell];
  }
  double voltage = 0;
  Operator ops[1 + AFGPU_MAX_MODES];
  // time-loop state
  int it = 0;
  double time = 0, global_time = 0, global_dt = 0, frac_rejected = 0, dt = 0;
  bool started = false;
  int photoi_cycles = 0;  // FMG cycles of the Helmholtz modes in the last photoionization solve

  double *C(int iv, int id) { return &cc[iv][(size_t)(id - 1) * ncell]; }
  double *F(int ifc, int id) { return &fc[ifc][(size_t)(id - 1) * D * nface]; }
  const Box &B(int id) const { return boxes[id - 1]; }
  bool has_children(int id) const { return boxes[id - 1].children[0] > 0; }
  long idx(int i, int j, int k) const { return i + (long)ng * (j + (long)ng * k); }
  long fidx(int i, int j, int k, int d) const {
    return (i - 1) + (long)(nc + 1) * ((j - 1) + (long)(nc + 1) * (k - 1)) + (long)d * nface;
  }
  int n_states() const { return p.time_integrator == AFGPU_HEUNS_METHOD ? 2 : 1; }
};

void set_error(const std::string &msg);

}  // namespace afcpu
