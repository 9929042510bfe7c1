// CPU oracle: C++/OpenMP restatement of afivo-streamer's per-time-step hot
// path (test infrastructure; loaded only by tests/, the smoke check and the
// cpu_baseline leg of bench.py).
//
// Every routine follows the reference routine cited next to it, box by box
// in the reference's OpenMP loop structure, with the reference's
// floating-point operation order (compiled with -ffp-contract=off like
// gfortran -O2 on x86-64, which does not contract).  The only deliberate
// departure is the level-1 solve: hypre PFMG (tolerance 1e-6) is replaced by
// an exact solve whose factors are supplied by afcpu_set_coarse_factors
// (parity unpinned against hypre).
//
// Pins: the afivo and lookup-table routines restated here are checked bit
// for bit against the reference's own routines, built from its sources
// (oracle/ref, tests/test_oracle_ref.py, through the afcpu_test_* entry
// points at the end of this file); the rest against closed forms and
// invariants (tests/test_oracle_pins.py and the other CPU tests).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "afcpu_impl.h"

struct afcpu_chem {
  int n_species = 0, n_reactions = 0;
  std::vector<afgpu_reaction_t> r;
  afcpu::Table fld, energy;
  double tg = 300;
  bool need_te = false;
};


namespace afcpu {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

// ---- constant tables (afivo/src/m_af_types.f90) ---------------------------
static const int child_dix2[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
static const int child_dix3[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
static const int child_adj_nb2[4][2] = {{1, 3}, {2, 4}, {1, 2}, {3, 4}};
static const int child_adj_nb3[6][4] = {{1, 3, 5, 7}, {2, 4, 6, 8}, {1, 2, 5, 6},
                                        {3, 4, 7, 8}, {1, 2, 3, 4}, {5, 6, 7, 8}};
static const int edge_dim3[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
static const int edge_dir3[12][3] = {{0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1},
                                     {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
                                     {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0}};
static const int edge_min_ix3[12][3] = {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1},
                                        {0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {1, 0, 1},
                                        {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}};

// neighbor_mat(-1:1, -1:1[, -1:1]) in column-major order (3^D entries)
static inline int nbmat_index(int di, int dj, int dk) { return (di + 1) + 3 * (dj + 1) + 9 * (dk + 1); }
static inline int nbmat_index2(int di, int dj) { return (di + 1) + 3 * (dj + 1); }

static inline void tangential(int D, int d, int &t0, int &t1) {
  if (D == 2) {
    t0 = 1 - d;
    t1 = -1;
  } else {
    t0 = (d == 0) ? 1 : 0;
    t1 = (d == 2) ? 1 : 2;
  }
}

// ---- lookup table (src/lookup_table_fortran/m_lookup_table.f90:330-406) ----
void Table::loc(double x, int &ix, double &w) const {
  double frac = (x - x_min) * inv_fac;
  if (frac <= 0) {
    ix = 1;
    w = 1;
  } else if (frac >= n - 1) {
    ix = n - 1;
    w = 0;
  } else {
    ix = (int)std::ceil(frac);
    w = ix - frac;
  }
}
double Table::at(int c, int ix, double w) const {
  const double *y = &cols[(size_t)c * n];
  return w * y[ix - 1] + (1 - w) * y[ix];
}
double Table::col(int c, double x) const {
  int ix;
  double w;
  loc(x, ix, w);
  return at(c, ix, w);
}

// ---- variable methods (src/streamer.f90:84-86, src/m_field.f90:392,
//      src/m_photoi.f90:115, afivo/src/m_af_multigrid.f90:100-103) ----------
static VarMethods methods(const Ctx &c, int iv) {
  VarMethods m{};
  m.ok = true;
  m.limiter = (c.D < 3) ? LIM_MC : LIM_GMINMOD43;  // m_af_core.f90:396-404
  for (int nb = 0; nb < 2 * c.D; nb++) {
    m.bc_type[nb] = BC_NEUMANN;
    m.bc_val[nb] = 0.0;
  }
  const int last = c.D - 1;
  if (iv == AFGPU_CC_PHI) {  // field_bc_homogeneous, src/m_field.f90:590-610
    m.rb = RB_MG;
    m.bc_type[2 * last] = BC_DIRICHLET;
    m.bc_val[2 * last] = 0.0;
    m.bc_type[2 * last + 1] = BC_DIRICHLET;
    m.bc_val[2 * last + 1] = c.voltage;
  } else if (iv >= AFGPU_CC_HELMH && iv < AFGPU_CC_HELMH + AFGPU_MAX_MODES) {
    m.rb = RB_MG;  // photoi_helmh_bc, src/m_photoi_helmh.f90:210-228
    m.bc_type[2 * last] = BC_DIRICHLET;
    m.bc_type[2 * last + 1] = BC_DIRICHLET;
  } else if (iv == AFGPU_CC_EFIELD) {
    m.rb = RB_INTERP;  // af_bc_neumann_zero + af_gc_interp
  } else if (iv == AFGPU_CC_PHOTO) {
    m.rb = RB_INTERP;
    m.bc_type[2 * last] = BC_DIRICHLET;
    m.bc_type[2 * last + 1] = BC_DIRICHLET;
  } else if (iv == AFGPU_CC_RHS) {
    m.rb = RB_INTERP;  // af_bc_neumann_zero + af_gc_interp (src/streamer.f90:103-108)
  } else if (iv >= 0 && iv < 9) {
    m.rb = RB_INTERP_LIM;  // bc_species + af_gc_interp_lim
    if (c.p.species_bc == AFGPU_SPECIES_BC_DIRICHLET_ZERO) {
      m.bc_type[2 * last] = BC_DIRICHLET_COPY;  // src/m_streamer.f90:503-518
      m.bc_type[2 * last + 1] = BC_DIRICHLET_COPY;
    }
  } else {
    m.ok = false;
  }
  return m;
}

// ---- limiters (afivo/src/m_af_limiters.f90) ------------------------------
static inline double limiter_gminmod(double a, double b, double theta) {
  if (a * b > 0) {
    double x = std::fabs(theta * a), y = std::fabs(theta * b), z = std::fabs(0.5 * (a + b));
    double m = std::min(std::min(x, y), z);
    return std::copysign(m, a);
  }
  return 0.0;
}
static inline double limiter_apply(double a, double b, int lim) {
  if (lim == LIM_MC) return limiter_gminmod(a, b, 2.0);
  return limiter_gminmod(a, b, 4 / 3.0);
}
static inline double limiter_koren(double a, double b) {  // m_af_limiters.f90:72-95
  const double third = 1 / 3.0;
  double aa = a * a, ab = a * b;
  if (ab <= 0) return 0;
  if (aa <= 0.25 * ab) return 2 * a;
  if (aa <= 2.5 * ab) return third * (b + 2 * a);
  return 2 * b;
}

// ---- ghost cells (afivo/src/m_af_ghostcell.f90) -------------------------
static void bc_to_gc(Ctx &c, int id, int nb, int iv, const VarMethods &m) {  // :173-279
  const Box &bx = c.B(id);
  double *u = c.C(iv, id);
  const int d = nb / 2;
  const bool low = (nb % 2) == 0;
  double c0, c1, c2;
  switch (m.bc_type[nb]) {
    case BC_DIRICHLET: c0 = 2; c1 = -1; c2 = 0; break;
    case BC_NEUMANN: c0 = bx.dr[d] * (low ? -1 : 1); c1 = 1; c2 = 0; break;
    default: c0 = 1; c1 = 0; c2 = 0; break;
  }
  const double bv = m.bc_val[nb];
  const int g = low ? 0 : c.nc + 1, x1 = low ? 1 : c.nc, x2 = low ? 2 : c.nc - 1;
  int t0, t1;
  tangential(c.D, d, t0, t1);
  const int bmax = (c.D == 3) ? c.nc : 1;
  for (int b = 1; b <= bmax; b++)
    for (int a = 1; a <= c.nc; a++) {
      int p[3] = {0, 0, 0};
      p[t0] = a;
      if (t1 >= 0) p[t1] = b;
      p[d] = g;
      long ig = c.idx(p[0], p[1], p[2]);
      p[d] = x1;
      long i1 = c.idx(p[0], p[1], p[2]);
      p[d] = x2;
      long i2 = c.idx(p[0], p[1], p[2]);
      u[ig] = c0 * bv + c1 * u[i1] + c2 * u[i2];
    }
}

static void copy_face_from_nb(Ctx &c, int id, int nb, int nb_id, int iv) {  // copy_from_nb :654
  double *u = c.C(iv, id);
  const double *w = c.C(iv, nb_id);
  const int d = nb / 2;
  const bool low = (nb % 2) == 0;
  const int g = low ? 0 : c.nc + 1, gs = low ? c.nc : 1;
  int t0, t1;
  tangential(c.D, d, t0, t1);
  const int bmax = (c.D == 3) ? c.nc : 1;
  for (int b = 1; b <= bmax; b++)
    for (int a = 1; a <= c.nc; a++) {
      int p[3] = {0, 0, 0};
      p[t0] = a;
      if (t1 >= 0) p[t1] = b;
      p[d] = g;
      long ig = c.idx(p[0], p[1], p[2]);
      p[d] = gs;
      u[ig] = w[c.idx(p[0], p[1], p[2])];
    }
}

// mg_sides_rb, afivo/src/m_af_multigrid.f90:294-461
static void mg_sides_rb(Ctx &c, int id, int nb, int iv) {
  const int nc = c.nc, hnc = nc / 2;
  const Box &bx = c.B(id);
  const int p_id = bx.parent;
  const int p_nb_id = c.B(p_id).nb[nb];
  const double *P = c.C(iv, p_nb_id);
  double *u = c.C(iv, id);
  int co[3] = {0, 0, 0};
  for (int dd = 0; dd < c.D; dd++) co[dd] = ((bx.ix[dd] - 1) & 1) * hnc;
  const int d = nb / 2;
  const bool low = (nb % 2) == 0;
  const int cn = low ? nc : 1;
  int t0, t1;
  tangential(c.D, d, t0, t1);
  const int ix = low ? 1 : nc, di = low ? 1 : -1;
  if (c.D == 2) {
    double tmp[34], gc[34];
    for (int q = 0; q <= hnc + 1; q++) {
      int p[3] = {0, 0, 0};
      p[d] = cn;
      p[t0] = co[t0] + q;
      tmp[q] = P[c.idx(p[0], p[1], p[2])];
    }
    for (int i = 1; i <= hnc; i++) {
      double grad = 0.125 * (tmp[i + 1] - tmp[i - 1]);
      gc[2 * i - 1] = tmp[i] - grad;
      gc[2 * i] = tmp[i] + grad;
    }
    for (int a = 1; a <= nc; a++) {
      int p[3] = {0, 0, 0};
      p[t0] = a;
      p[d] = ix - di;
      long ig = c.idx(p[0], p[1], p[2]);
      p[d] = ix;
      long i1 = c.idx(p[0], p[1], p[2]);
      p[d] = ix + di;
      long i2 = c.idx(p[0], p[1], p[2]);
      u[ig] = 0.5 * gc[a] + 0.75 * u[i1] - 0.25 * u[i2];
    }
  } else {
    const int W = hnc + 2;
    double tmp[34 * 34];
    static thread_local std::vector<double> gcv;
    gcv.assign((size_t)(nc + 1) * (nc + 1), 0.0);
    for (int q1 = 0; q1 <= hnc + 1; q1++)
      for (int q0 = 0; q0 <= hnc + 1; q0++) {
        int p[3] = {0, 0, 0};
        p[d] = cn;
        p[t0] = co[t0] + q0;
        p[t1] = co[t1] + q1;
        tmp[q0 + W * q1] = P[c.idx(p[0], p[1], p[2])];
      }
    auto T = [&](int a, int b) { return tmp[a + W * b]; };
    auto G = [&](int a, int b) -> double & { return gcv[a + (nc + 1) * b]; };
    for (int j = 1; j <= hnc; j++)
      for (int i = 1; i <= hnc; i++) {
        double g1 = 0.125 * (T(i + 1, j) - T(i - 1, j));
        double g2 = 0.125 * (T(i, j + 1) - T(i, j - 1));
        G(2 * i - 1, 2 * j - 1) = T(i, j) - g1 - g2;
        G(2 * i, 2 * j - 1) = T(i, j) + g1 - g2;
        G(2 * i - 1, 2 * j) = T(i, j) - g1 + g2;
        G(2 * i, 2 * j) = T(i, j) + g1 + g2;
      }
    for (int b = 1; b <= nc; b++)
      for (int a = 1; a <= nc; a++) {
        int p[3] = {0, 0, 0};
        p[t0] = a;
        p[t1] = b;
        p[d] = ix - di;
        long ig = c.idx(p[0], p[1], p[2]);
        p[d] = ix;
        long i1 = c.idx(p[0], p[1], p[2]);
        p[d] = ix + di;
        long i2 = c.idx(p[0], p[1], p[2]);
        u[ig] = 0.5 * G(a, b) + 0.75 * u[i1] - 0.25 * u[i2];
      }
  }
}

// af_gc_interp (:394-498) and af_gc_interp_lim (:503-605)
static void gc_interp(Ctx &c, int id, int nb, int iv, bool limited) {
  const int nc = c.nc, hnc = nc / 2;
  const Box &bx = c.B(id);
  const int p_nb_id = c.B(bx.parent).nb[nb];
  const double *P = c.C(iv, p_nb_id);
  double *u = c.C(iv, id);
  int off[3] = {0, 0, 0};
  for (int dd = 0; dd < c.D; dd++) off[dd] = ((bx.ix[dd] - 1) & 1) * hnc;
  const int d = nb / 2;
  const bool low = (nb % 2) == 0;
  const int gix = low ? 0 : nc + 1, ix_f = low ? 1 : nc, ix_c = low ? nc : 1;
  int t0, t1;
  tangential(c.D, d, t0, t1);
  const double third = 1 / 3.0, sixth = 1 / 6.0;
  if (c.D == 2) {
    for (int a = 1; a <= nc; a++) {
      int a1 = off[t0] + ((a + 1) >> 1);
      int a2 = a1 + 1 - 2 * (a & 1);
      int p[3] = {0, 0, 0};
      p[d] = ix_c;
      p[t0] = a1;
      double c1 = P[c.idx(p[0], p[1], p[2])];
      p[t0] = a2;
      double c2 = P[c.idx(p[0], p[1], p[2])];
      p[t0] = a;
      p[d] = ix_f;
      double f = u[c.idx(p[0], p[1], p[2])];
      p[d] = gix;
      double v = 0.5 * c1 + sixth * c2 + third * f;
      if (limited && v > 2 * c1) v = 2 * c1;
      u[c.idx(p[0], p[1], p[2])] = v;
    }
  } else {
    for (int b = 1; b <= nc; b++) {
      int b1 = off[t1] + ((b + 1) >> 1);
      int b2 = b1 + 1 - 2 * (b & 1);
      for (int a = 1; a <= nc; a++) {
        int a1 = off[t0] + ((a + 1) >> 1);
        int a2 = a1 + 1 - 2 * (a & 1);
        int p[3];
        p[d] = ix_c;
        p[t0] = a1; p[t1] = b1;
        double c1 = P[c.idx(p[0], p[1], p[2])];
        p[t0] = a2; p[t1] = b1;
        double ca = P[c.idx(p[0], p[1], p[2])];
        p[t0] = a1; p[t1] = b2;
        double cb = P[c.idx(p[0], p[1], p[2])];
        p[t0] = a; p[t1] = b; p[d] = ix_f;
        double f = u[c.idx(p[0], p[1], p[2])];
        p[d] = gix;
        // The z-face case sums the two off-diagonal coarse values in the
        // opposite order (m_af_ghostcell.f90:485-496, 592-603).
        double c2 = (d == 2) ? cb : ca, c3 = (d == 2) ? ca : cb;
        double v = third * c1 + sixth * c2 + sixth * c3 + third * f;
        if (limited && v > 2 * c1) v = 2 * c1;
        u[c.idx(p[0], p[1], p[2])] = v;
      }
    }
  }
}

// af_gc_box_corner (:125-170) with af_edge_gc_extrap (:885-924) and
// af_corner_gc_extrap (:860-879)
static void gc_box_corner(Ctx &c, int id, int iv) {
  const int nc = c.nc;
  const Box &bx = c.B(id);
  double *u = c.C(iv, id);
  if (c.D == 2) {
    for (int n = 0; n < 4; n++) {
      int dnb[2] = {2 * child_dix2[n][0] - 1, 2 * child_dix2[n][1] - 1};
      int nb_id = bx.nbmat[nbmat_index2(dnb[0], dnb[1])];
      int lo[2] = {child_dix2[n][0] * (nc + 1), child_dix2[n][1] * (nc + 1)};
      if (nb_id > 0) {
        u[c.idx(lo[0], lo[1], 0)] = c.C(iv, nb_id)[c.idx(lo[0] - dnb[0] * nc, lo[1] - dnb[1] * nc, 0)];
      } else {
        int di = 1 - 2 * (lo[0] & 1), dj = 1 - 2 * (lo[1] & 1);
        u[c.idx(lo[0], lo[1], 0)] = u[c.idx(lo[0] + di, lo[1], 0)] + u[c.idx(lo[0], lo[1] + dj, 0)] -
                                    u[c.idx(lo[0] + di, lo[1] + dj, 0)];
      }
    }
    return;
  }
  for (int n = 0; n < 12; n++) {
    const int dim = edge_dim3[n];
    const int *dir = edge_dir3[n];
    int nb_id = bx.nbmat[nbmat_index(dir[0], dir[1], dir[2])];
    int lo[3];
    for (int dd = 0; dd < 3; dd++) lo[dd] = edge_min_ix3[n][dd] * (nc + 1);
    lo[dim] = 1;
    if (nb_id > 0) {
      const double *w = c.C(iv, nb_id);
      for (int q = 1; q <= nc; q++) {
        int p[3] = {lo[0], lo[1], lo[2]};
        p[dim] = q;
        u[c.idx(p[0], p[1], p[2])] = w[c.idx(p[0] - dir[0] * nc, p[1] - dir[1] * nc, p[2] - dir[2] * nc)];
      }
    } else {
      int o0 = (dim + 1) % 3, o1 = (dim + 2) % 3;
      int di[3];
      for (int dd = 0; dd < 3; dd++) di[dd] = 1 - 2 * (lo[dd] & 1);
      di[dim] = 0;
      for (int q = 1; q <= nc; q++) {
        int x[3] = {lo[0], lo[1], lo[2]};
        x[dim] = q;
        int ia[3] = {x[0], x[1], x[2]}, ib[3] = {x[0], x[1], x[2]};
        ia[o0] += di[o0];
        ib[o1] += di[o1];
        int ic[3] = {x[0] + di[0], x[1] + di[1], x[2] + di[2]};
        u[c.idx(x[0], x[1], x[2])] = u[c.idx(ia[0], ia[1], ia[2])] + u[c.idx(ib[0], ib[1], ib[2])] -
                                     u[c.idx(ic[0], ic[1], ic[2])];
      }
    }
  }
  for (int n = 0; n < 8; n++) {
    int dnb[3], lo[3];
    for (int dd = 0; dd < 3; dd++) {
      dnb[dd] = 2 * child_dix3[n][dd] - 1;
      lo[dd] = child_dix3[n][dd] * (nc + 1);
    }
    int nb_id = bx.nbmat[nbmat_index(dnb[0], dnb[1], dnb[2])];
    if (nb_id > 0) {
      u[c.idx(lo[0], lo[1], lo[2])] =
          c.C(iv, nb_id)[c.idx(lo[0] - dnb[0] * nc, lo[1] - dnb[1] * nc, lo[2] - dnb[2] * nc)];
    } else {
      int di = 1 - 2 * (lo[0] & 1), dj = 1 - 2 * (lo[1] & 1), dk = 1 - 2 * (lo[2] & 1);
      int x = lo[0], y = lo[1], z = lo[2];
      u[c.idx(x, y, z)] = u[c.idx(x, y + dj, z + dk)] + u[c.idx(x + di, y, z + dk)] +
                          u[c.idx(x + di, y + dj, z)] - 2 * u[c.idx(x + di, y + dj, z + dk)];
    }
  }
}

// af_gc_box, afivo/src/m_af_ghostcell.f90:64-120
static void gc_box(Ctx &c, int id, int iv, bool corners) {
  VarMethods m = methods(c, iv);
  const Box &bx = c.B(id);
  for (int nb = 0; nb < 2 * c.D; nb++) {
    int nb_id = bx.nb[nb];
    if (nb_id > 0) {
      copy_face_from_nb(c, id, nb, nb_id, iv);
    } else if (nb_id == 0) {
      if (m.rb == RB_MG)
        mg_sides_rb(c, id, nb, iv);
      else
        gc_interp(c, id, nb, iv, m.rb == RB_INTERP_LIM);
    } else {
      bc_to_gc(c, id, nb, iv, m);
    }
  }
  if (corners) gc_box_corner(c, id, iv);
}

static void gc_lvl(Ctx &c, int lvl, int iv, bool corners) {  // af_gc_lvl :49
  const std::vector<int> &ids = c.ids[lvl];
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < ids.size(); i++) gc_box(c, ids[i], iv, corners);
}

static void gc_tree(Ctx &c, int iv) {  // af_gc_tree :24
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) gc_lvl(c, lvl, iv, true);
}

// ---- restriction (afivo/src/m_af_restrict.f90:62-136) ---------------------
static void restrict_box(Ctx &c, int id_c, int iv_from, int iv_to, bool use_geom) {
  const int nc = c.nc, hnc = nc / 2;
  const Box &bc_ = c.B(id_c);
  const int id_p = bc_.parent;
  const Box &bp = c.B(id_p);
  const double *f = c.C(iv_from, id_c);
  double *u = c.C(iv_to, id_p);
  int off[3] = {0, 0, 0};
  for (int d = 0; d < c.D; d++) off[d] = ((bc_.ix[d] - 1) & 1) * hnc;
  if (c.D == 2) {
    const bool cyl = (c.coord == AFGPU_CYL) && use_geom;
    for (int j = 1; j <= hnc; j++) {
      int jc = off[1] + j, jf = 2 * j - 1;
      for (int i = 1; i <= hnc; i++) {
        int ic = off[0] + i, i_f = 2 * i - 1;
        if (cyl) {
          double r = bp.r_min[0] + (ic - 0.5) * bp.dr[0];
          double tmp = 0.25 * bp.dr[0] / r;
          double w1 = 1 - tmp, w2 = 1 + tmp;
          u[c.idx(ic, jc, 0)] =
              0.25 * (w1 * (f[c.idx(i_f, jf, 0)] + f[c.idx(i_f, jf + 1, 0)]) +
                      w2 * (f[c.idx(i_f + 1, jf, 0)] + f[c.idx(i_f + 1, jf + 1, 0)]));
        } else {
          double s = f[c.idx(i_f, jf, 0)];
          s = s + f[c.idx(i_f + 1, jf, 0)];
          s = s + f[c.idx(i_f, jf + 1, 0)];
          s = s + f[c.idx(i_f + 1, jf + 1, 0)];
          u[c.idx(ic, jc, 0)] = 0.25 * s;
        }
      }
    }
  } else {
    for (int k = 1; k <= hnc; k++) {
      int kc = off[2] + k, kf = 2 * k - 1;
      for (int j = 1; j <= hnc; j++) {
        int jc = off[1] + j, jf = 2 * j - 1;
        for (int i = 1; i <= hnc; i++) {
          int ic = off[0] + i, i_f = 2 * i - 1;
          double s = f[c.idx(i_f, jf, kf)];
          s = s + f[c.idx(i_f + 1, jf, kf)];
          s = s + f[c.idx(i_f, jf + 1, kf)];
          s = s + f[c.idx(i_f + 1, jf + 1, kf)];
          s = s + f[c.idx(i_f, jf, kf + 1)];
          s = s + f[c.idx(i_f + 1, jf, kf + 1)];
          s = s + f[c.idx(i_f, jf + 1, kf + 1)];
          s = s + f[c.idx(i_f + 1, jf + 1, kf + 1)];
          u[c.idx(ic, jc, kc)] = 0.125 * s;
        }
      }
    }
  }
}

// ---- stencils (afivo/src/m_af_multigrid.f90:1246-1264, m_af_stencil.f90) --
struct Stencil {
  double c[7];
};
static Stencil lpl_stencil(const Ctx &c, const Box &bx, double lambda2) {
  Stencil s{};
  for (int d = 0; d < c.D; d++) {
    double inv_dr2 = 1 / (bx.dr[d] * bx.dr[d]);
    s.c[1 + 2 * d] = inv_dr2;
    s.c[2 + 2 * d] = inv_dr2;
  }
  double sum = 0;
  for (int q = 1; q <= 2 * c.D; q++) sum = sum + s.c[q];
  s.c[0] = -sum - lambda2;
  return s;
}
static void cyl_factors(const Ctx &c, const Box &bx, double *rf1, double *rf2) {
  // af_cyl_flux_factors, afivo/src/m_af_types.f90:1199-1211
  for (int i = 1; i <= c.nc; i++) {
    double r = bx.r_min[0] + (i - 0.5) * bx.dr[0];
    double inv_r = 1 / r;
    rf1[i] = (r - 0.5 * bx.dr[0]) * inv_r;
    rf2[i] = (r + 0.5 * bx.dr[0]) * inv_r;
  }
}

// stencil_gsrb_357, afivo/src/m_af_stencil.f90:838-998
static void gsrb_box(Ctx &c, int id, int redblack, int iv, int i_rhs, double lambda2) {
  const Box &bx = c.B(id);
  const int nc = c.nc;
  Stencil s = lpl_stencil(c, bx, lambda2);
  double *u = c.C(iv, id);
  const double *r = c.C(i_rhs, id);
  if (c.D == 2) {
    const int ng = c.ng;
    if (c.coord == AFGPU_CYL) {
      double rf1[33], rf2[33], cc1[33], cc2[33], cc3[33], inv1[33];
      cyl_factors(c, bx, rf1, rf2);
      for (int i = 1; i <= nc; i++) {
        cc2[i] = rf1[i] * s.c[1];
        cc3[i] = rf2[i] * s.c[2];
        cc1[i] = s.c[0] - (cc2[i] - s.c[1]) - (cc3[i] - s.c[2]);
        inv1[i] = 1 / cc1[i];
      }
      for (int j = 1; j <= nc; j++) {
        int i0 = 2 - ((redblack ^ j) & 1);
        for (int i = i0; i <= nc; i += 2) {
          long x = i + ng * j;
          u[x] = (r[x] - cc2[i] * u[x - 1] - cc3[i] * u[x + 1] - s.c[3] * u[x - ng] -
                  s.c[4] * u[x + ng]) *
                 inv1[i];
        }
      }
    } else {
      double inv_c1 = 1 / s.c[0];
      for (int j = 1; j <= nc; j++) {
        int i0 = 2 - ((redblack ^ j) & 1);
        for (int i = i0; i <= nc; i += 2) {
          long x = i + ng * j;
          u[x] = (r[x] - s.c[1] * u[x - 1] - s.c[2] * u[x + 1] - s.c[3] * u[x - ng] -
                  s.c[4] * u[x + ng]) *
                 inv_c1;
        }
      }
    }
  } else {
    const long ng = c.ng, ng2 = ng * ng;
    double inv_c1 = 1 / s.c[0];
    for (int k = 1; k <= nc; k++)
      for (int j = 1; j <= nc; j++) {
        int i0 = 2 - ((redblack ^ (k + j)) & 1);
        for (int i = i0; i <= nc; i += 2) {
          long x = c.idx(i, j, k);
          u[x] = (r[x] - s.c[1] * u[x - 1] - s.c[2] * u[x + 1] - s.c[3] * u[x - ng] -
                  s.c[4] * u[x + ng] - s.c[5] * u[x - ng2] - s.c[6] * u[x + ng2]) *
                 inv_c1;
        }
      }
  }
}

// stencil_apply_357 (afivo/src/m_af_stencil.f90:367-496) into i_out
static void apply_box(Ctx &c, int id, int iv, int i_out, double lambda2) {
  const Box &bx = c.B(id);
  const int nc = c.nc;
  Stencil s = lpl_stencil(c, bx, lambda2);
  const double *u = c.C(iv, id);
  double *o = c.C(i_out, id);
  if (c.D == 2) {
    const int ng = c.ng;
    if (c.coord == AFGPU_CYL) {
      double rf1[33], rf2[33], cc1[33], cc2[33], cc3[33];
      cyl_factors(c, bx, rf1, rf2);
      for (int i = 1; i <= nc; i++) {
        cc2[i] = rf1[i] * s.c[1];
        cc3[i] = rf2[i] * s.c[2];
        cc1[i] = s.c[0] - (cc2[i] - s.c[1]) - (cc3[i] - s.c[2]);
      }
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          long x = i + ng * j;
          o[x] = cc1[i] * u[x] + cc2[i] * u[x - 1] + cc3[i] * u[x + 1] + s.c[3] * u[x - ng] +
                 s.c[4] * u[x + ng];
        }
    } else {
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          long x = i + ng * j;
          o[x] = s.c[0] * u[x] + s.c[1] * u[x - 1] + s.c[2] * u[x + 1] + s.c[3] * u[x - ng] +
                 s.c[4] * u[x + ng];
        }
    }
  } else {
    const long ng = c.ng, ng2 = ng * ng;
    for (int k = 1; k <= nc; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          long x = c.idx(i, j, k);
          o[x] = s.c[0] * u[x] + s.c[1] * u[x - 1] + s.c[2] * u[x + 1] + s.c[3] * u[x - ng] +
                 s.c[4] * u[x + ng] + s.c[5] * u[x - ng2] + s.c[6] * u[x + ng2];
        }
  }
}

// residual_box, afivo/src/m_af_multigrid.f90:801-810
static void residual_box(Ctx &c, int id, int i_phi, double lambda2) {
  apply_box(c, id, i_phi, AFGPU_CC_TMP, lambda2);
  double *t = c.C(AFGPU_CC_TMP, id);
  const double *r = c.C(AFGPU_CC_RHS, id);
  const int nc = c.nc;
  const int kmax = (c.D == 3) ? nc : 0, kmin = (c.D == 3) ? 1 : 0;
  for (int k = kmin; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) {
        long x = c.idx(i, j, k);
        t[x] = r[x] - t[x];
      }
}

// stencil_prolong_248 with add (afivo/src/m_af_stencil.f90:695-815)
static void prolong_add(Ctx &c, int id_p, int id_c, int iv, int iv_to) {
  const int nc = c.nc, hnc = nc / 2;
  const Box &bc_ = c.B(id_c);
  const double *P = c.C(iv, id_p);
  double *u = c.C(iv_to, id_c);
  int off[3] = {0, 0, 0};
  for (int d = 0; d < c.D; d++) off[d] = ((bc_.ix[d] - 1) & 1) * hnc;
  if (c.D == 2) {
    const double w[4] = {9 / 16.0, 3 / 16.0, 3 / 16.0, 1 / 16.0};
    for (int j = 1; j <= nc; j++) {
      int j1 = off[1] + ((j + 1) >> 1), j2 = j1 + 1 - 2 * (j & 1);
      for (int i = 1; i <= nc; i++) {
        int i1 = off[0] + ((i + 1) >> 1), i2 = i1 + 1 - 2 * (i & 1);
        long x = c.idx(i, j, 0);
        u[x] = u[x] + w[0] * P[c.idx(i1, j1, 0)] + w[1] * P[c.idx(i2, j1, 0)] +
               w[2] * P[c.idx(i1, j2, 0)] + w[3] * P[c.idx(i2, j2, 0)];
      }
    }
  } else {
    const double w[8] = {27 / 64.0, 9 / 64.0, 9 / 64.0, 3 / 64.0, 9 / 64.0, 3 / 64.0, 3 / 64.0, 1 / 64.0};
    for (int k = 1; k <= nc; k++) {
      int k1 = off[2] + ((k + 1) >> 1), k2 = k1 + 1 - 2 * (k & 1);
      for (int j = 1; j <= nc; j++) {
        int j1 = off[1] + ((j + 1) >> 1), j2 = j1 + 1 - 2 * (j & 1);
        for (int i = 1; i <= nc; i++) {
          int i1 = off[0] + ((i + 1) >> 1), i2 = i1 + 1 - 2 * (i & 1);
          long x = c.idx(i, j, k);
          u[x] = u[x] + w[0] * P[c.idx(i1, j1, k1)] + w[1] * P[c.idx(i2, j1, k1)] +
                 w[2] * P[c.idx(i1, j2, k1)] + w[3] * P[c.idx(i2, j2, k1)] +
                 w[4] * P[c.idx(i1, j1, k2)] + w[5] * P[c.idx(i2, j1, k2)] +
                 w[6] * P[c.idx(i1, j2, k2)] + w[7] * P[c.idx(i2, j2, k2)];
        }
      }
    }
  }
}

// ---- whole-box helpers (afivo/src/m_af_utils.f90:385-576) ---------------
static void box_copy(Ctx &c, int id, int from, int to) {
  std::memcpy(c.C(to, id), c.C(from, id), c.ncell * sizeof(double));
}
static void box_clear(Ctx &c, int id, int iv) { std::memset(c.C(iv, id), 0, c.ncell * sizeof(double)); }

// ---- multigrid (afivo/src/m_af_multigrid.f90) ---------------------------
struct MG {
  int op;
  int i_phi;
  double lambda2;
};
static MG mg_of(const Ctx &c, int op) {
  MG m;
  m.op = op;
  m.i_phi = (op == AFGPU_OP_POISSON) ? AFGPU_CC_PHI : AFGPU_CC_HELMH + (op - AFGPU_OP_HELMH);
  m.lambda2 = c.ops[op].lambda2;
  return m;
}

// gsrb_boxes :648-687
static void gsrb_boxes(Ctx &c, int lvl, const MG &mg, bool down) {
  const int n_cycle = down ? c.p.mg_n_cycle_down : c.p.mg_n_cycle_up;
  const std::vector<int> &ids = c.ids[lvl];
  for (int n = 1; n <= 2 * n_cycle; n++) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < ids.size(); i++) gsrb_box(c, ids[i], n, mg.i_phi, AFGPU_CC_RHS, mg.lambda2);
    bool use_corners = (!down && n == 2 * n_cycle);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < ids.size(); i++) gc_box(c, ids[i], mg.i_phi, use_corners);
  }
}

// update_coarse :691-738
static void update_coarse(Ctx &c, int lvl, const MG &mg) {
  const std::vector<int> &ids = c.ids[lvl];
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < ids.size(); i++) {
    int id = ids[i];
    std::vector<double> save(c.C(AFGPU_CC_TMP, id), c.C(AFGPU_CC_TMP, id) + c.ncell);
    residual_box(c, id, mg.i_phi, mg.lambda2);
    restrict_box(c, id, AFGPU_CC_TMP, AFGPU_CC_TMP, true);
    restrict_box(c, id, mg.i_phi, mg.i_phi, false);
    // Only the interior of tmp is saved and restored in the reference; the
    // ghost cells of tmp are not touched by residual_box, so a full copy is
    // equivalent.
    std::memcpy(c.C(AFGPU_CC_TMP, id), save.data(), c.ncell * sizeof(double));
  }
  gc_lvl(c, lvl - 1, mg.i_phi, true);
  const std::vector<int> &par = c.parents[lvl - 1];
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < par.size(); i++) {
    int id = par[i];
    apply_box(c, id, mg.i_phi, AFGPU_CC_RHS, mg.lambda2);
    double *rh = c.C(AFGPU_CC_RHS, id);
    const double *t = c.C(AFGPU_CC_TMP, id);
    for (size_t q = 0; q < c.ncell; q++) rh[q] = rh[q] + t[q];
    box_copy(c, id, mg.i_phi, AFGPU_CC_TMP);
  }
}

// set_coarse_phi_rhs :742-776
static void set_coarse_phi_rhs(Ctx &c, int lvl, const MG &mg) {
  if (lvl == c.highest_lvl) gc_lvl(c, lvl, mg.i_phi, true);
  const std::vector<int> &ids = c.ids[lvl];
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < ids.size(); i++) {
    int id = ids[i];
    residual_box(c, id, mg.i_phi, mg.lambda2);
    restrict_box(c, id, AFGPU_CC_TMP, AFGPU_CC_TMP, true);
    restrict_box(c, id, mg.i_phi, mg.i_phi, false);
  }
  gc_lvl(c, lvl - 1, mg.i_phi, true);
  const std::vector<int> &par = c.parents[lvl - 1];
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < par.size(); i++) {
    int id = par[i];
    apply_box(c, id, mg.i_phi, AFGPU_CC_RHS, mg.lambda2);
    double *rh = c.C(AFGPU_CC_RHS, id);
    const double *t = c.C(AFGPU_CC_TMP, id);
    for (size_t q = 0; q < c.ncell; q++) rh[q] = rh[q] + t[q];
  }
}

// init_phi_rhs :779-799
static void init_phi_rhs(Ctx &c, const MG &mg) {
  for (int lvl = c.highest_lvl; lvl >= 2; lvl--) {
    const std::vector<int> &ids = c.ids[lvl];
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < ids.size(); i++) {
      box_clear(c, ids[i], mg.i_phi);
      restrict_box(c, ids[i], AFGPU_CC_RHS, AFGPU_CC_RHS, true);
    }
  }
}

// correct_children :624-646
static void correct_children(Ctx &c, int lvl_p, const MG &mg) {
  const std::vector<int> &par = c.parents[lvl_p];
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < par.size(); i++) {
    int id = par[i];
    double *t = c.C(AFGPU_CC_TMP, id);
    const double *u = c.C(mg.i_phi, id);
    for (size_t q = 0; q < c.ncell; q++) t[q] = u[q] - t[q];
    const Box &bx = c.B(id);
    for (int ic = 0; ic < (1 << c.D); ic++) {
      int c_id = bx.children[ic];
      if (c_id == 0) continue;
      prolong_add(c, id, c_id, AFGPU_CC_TMP, mg.i_phi);
    }
  }
}

// Exact level-1 solve (replaces solve_coarse_grid/hypre :266-291)
static void solve_coarse_grid(Ctx &c, const MG &mg) {
  const std::vector<int> &ids = c.ids[1];
  const int nc = c.nc, D = c.D;
  const Operator &op = c.ops[mg.op];
  const int id = ids[0];
  const Box &bx = c.B(id);
  VarMethods m = methods(c, mg.i_phi);
  Stencil s = lpl_stencil(c, bx, mg.lambda2);
  double rf1[33], rf2[33];
  if (c.coord == AFGPU_CYL) cyl_factors(c, bx, rf1, rf2);
  const int n3 = (D == 3) ? nc : 1;
  const size_t N = (size_t)nc * nc * n3;
  std::vector<double> b(N), y(N);
  auto lin = [&](int i, int j, int k) { return (size_t)(i - 1) + (size_t)nc * ((j - 1) + (size_t)nc * (k - 1)); };
  const int kmin = 1, kmax = n3;
  const double *rhs = c.C(AFGPU_CC_RHS, id);
  for (int k = kmin; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) b[lin(i, j, k)] = rhs[c.idx(i, j, D == 3 ? k : 0)];
  // Fold boundary values into the rhs (coarse_solver_set_rhs_phi,
  // afivo/src/m_coarse_solver.f90:286-338, coefficients :442-491).
  for (int nb = 0; nb < 2 * D; nb++) {
    if (bx.nb[nb] >= 0) continue;
    const int d = nb / 2;
    const bool low = (nb % 2) == 0;
    const int ib = low ? 1 : nc;
    for (int k = kmin; k <= kmax; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          int p[3] = {i, j, k};
          if (p[d] != ib) continue;
          double cnb = s.c[1 + nb];
          if (c.coord == AFGPU_CYL && d == 0) cnb = (low ? rf1[i] : rf2[i]) * s.c[1 + nb];
          double f;
          if (m.bc_type[nb] == BC_DIRICHLET)
            f = -2 * cnb;
          else
            f = -(cnb * bx.dr[d]) * (double)(low ? -1 : 1);
          size_t q = lin(i, j, k);
          b[q] = b[q] + f * m.bc_val[nb];
        }
  }
  auto transform = [&](const std::vector<double> &M, int d, const std::vector<double> &in,
                       std::vector<double> &out) {
    const double *Md = &M[(size_t)d * nc * nc];
    for (int k = kmin; k <= kmax; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          int p[3] = {i, j, k};
          int row = p[d];
          double sum = 0.0;
          for (int mm = 1; mm <= nc; mm++) {
            int q[3] = {i, j, k};
            q[d] = mm;
            sum = sum + Md[(row - 1) * nc + (mm - 1)] * in[lin(q[0], q[1], q[2])];
          }
          out[lin(i, j, k)] = sum;
        }
  };
  for (int d = 0; d < D; d++) {
    transform(op.vinv, d, b, y);
    b.swap(y);
  }
  for (int k = kmin; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) {
        double den = op.eig[i - 1] + op.eig[nc + j - 1];
        if (D == 3) den = den + op.eig[2 * nc + k - 1];
        den = den - mg.lambda2;
        b[lin(i, j, k)] = b[lin(i, j, k)] / den;
      }
  for (int d = 0; d < D; d++) {
    transform(op.v, d, b, y);
    b.swap(y);
  }
  double *u = c.C(mg.i_phi, id);
  for (int k = kmin; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) u[c.idx(i, j, D == 3 ? k : 0)] = b[lin(i, j, k)];
  gc_lvl(c, 1, mg.i_phi, true);
}

// mg_fas_vcycle :185-264
static void vcycle(Ctx &c, const MG &mg, bool set_residual, int max_lvl) {
  for (int lvl = max_lvl; lvl >= 2; lvl--) {
    gsrb_boxes(c, lvl, mg, true);
    update_coarse(c, lvl, mg);
  }
  solve_coarse_grid(c, mg);
  for (int lvl = 2; lvl <= max_lvl; lvl++) {
    correct_children(c, lvl - 1, mg);
    gc_lvl(c, lvl, mg.i_phi, true);
    gsrb_boxes(c, lvl, mg, false);
  }
  if (set_residual) {
    for (int lvl = 1; lvl <= max_lvl; lvl++) {
      const std::vector<int> &ids = c.ids[lvl];
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < ids.size(); i++) residual_box(c, ids[i], mg.i_phi, mg.lambda2);
    }
  }
}

// mg_fas_fmg :137-180
static void fmg(Ctx &c, const MG &mg, bool set_residual, bool have_guess) {
  if (have_guess) {
    for (int lvl = c.highest_lvl; lvl >= 2; lvl--) set_coarse_phi_rhs(c, lvl, mg);
  } else {
    init_phi_rhs(c, mg);
  }
  int lvl = 1;
  for (int id : c.ids[lvl]) box_copy(c, id, mg.i_phi, AFGPU_CC_TMP);
  vcycle(c, mg, set_residual && lvl == c.highest_lvl, lvl);
  for (lvl = 2; lvl <= c.highest_lvl; lvl++) {
    const std::vector<int> &ids = c.ids[lvl];
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < ids.size(); i++) box_copy(c, ids[i], mg.i_phi, AFGPU_CC_TMP);
    correct_children(c, lvl - 1, mg);
    gc_lvl(c, lvl, mg.i_phi, true);
    vcycle(c, mg, set_residual && lvl == c.highest_lvl, lvl);
  }
}

// af_tree_maxabs_cc, afivo/src/m_af_utils.f90:694-786 (leaves, interior)
static double tree_maxabs(Ctx &c, int iv) {
  double result = -std::numeric_limits<double>::max() / 10;
  const int nc = c.nc;
  const int kmin = (c.D == 3) ? 1 : 0, kmax = (c.D == 3) ? nc : 0;
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
    const std::vector<int> &lv = c.leaves[lvl];
    double lm = result;
#pragma omp parallel for schedule(static) reduction(max : lm)
    for (size_t n = 0; n < lv.size(); n++) {
      const double *u = c.C(iv, lv[n]);
      double bm = 0;
      for (int k = kmin; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) bm = std::max(bm, std::fabs(u[c.idx(i, j, k)]));
      lm = std::max(lm, bm);
    }
    result = std::max(result, lm);
  }
  return result;
}

// ---- field (src/m_field.f90) ----------------------------------------------
static void field_set_rhs(Ctx &c, int s_in) {  // :406-444
  const double fac = -UC_elem_charge / UC_eps0;
  const std::vector<int> &charge = c.sp_charge;
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
    const std::vector<int> &lv = c.leaves[lvl];
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < lv.size(); i++) {
      int id = lv[i];
      double *rh = c.C(AFGPU_CC_RHS, id);
      for (size_t q = 0; q < c.ncell; q++) rh[q] = 0.0;
      for (int n = 0; n < c.n_sp; n++) {
        if (charge[n] == 0) continue;
        const double *u = c.SP(n, s_in, id);
        double qf = charge[n] * fac;
        for (size_t q = 0; q < c.ncell; q++) rh[q] = rh[q] + qf * u[q];
      }
    }
  }
}

static double min_dr(const Ctx &c) {  // af_min_dr: finest level
  const Box &bx = c.B(c.ids[c.highest_lvl][0]);
  double m = bx.dr[0];
  for (int d = 1; d < c.D; d++) m = std::min(m, bx.dr[d]);
  return m;
}

// mg_compute_phi_gradient + mg_box_field_norm (afivo/src/m_af_multigrid.f90:1857-2045)
static void field_from_potential(Ctx &c) {
  const int nc = c.nc, D = c.D;
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
    const std::vector<int> &ids = c.ids[lvl];
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < ids.size(); n++) {
      int id = ids[n];
      const Box &bx = c.B(id);
      const double *u = c.C(AFGPU_CC_PHI, id);
      double *f = c.F(AFGPU_FC_FIELD, id);
      double inv_dr[3];
      for (int d = 0; d < D; d++) inv_dr[d] = -1.0 / bx.dr[d];
      const int k0 = (D == 3) ? 1 : 0;
      const int kn = (D == 3) ? nc : 0;
      // x faces
      for (int k = k0; k <= kn; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc + 1; i++)
            f[c.fidx(i, j, D == 3 ? k : 1, 0)] = inv_dr[0] * (u[c.idx(i, j, k)] - u[c.idx(i - 1, j, k)]);
      for (int k = k0; k <= kn; k++)
        for (int j = 1; j <= nc + 1; j++)
          for (int i = 1; i <= nc; i++)
            f[c.fidx(i, j, D == 3 ? k : 1, 1)] = inv_dr[1] * (u[c.idx(i, j, k)] - u[c.idx(i, j - 1, k)]);
      if (D == 3)
        for (int k = 1; k <= nc + 1; k++)
          for (int j = 1; j <= nc; j++)
            for (int i = 1; i <= nc; i++)
              f[c.fidx(i, j, k, 2)] = inv_dr[2] * (u[c.idx(i, j, k)] - u[c.idx(i, j, k - 1)]);
      double *e = c.C(AFGPU_CC_EFIELD, id);
      for (int k = k0; k <= kn; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            int kk = (D == 3) ? k : 1;
            double a = f[c.fidx(i, j, kk, 0)] + f[c.fidx(i + 1, j, kk, 0)];
            double b = f[c.fidx(i, j, kk, 1)] + f[c.fidx(i, j + 1, kk, 1)];
            double s = a * a + b * b;
            if (D == 3) {
              double z = f[c.fidx(i, j, k, 2)] + f[c.fidx(i, j, k + 1, 2)];
              s = s + z * z;
            }
            e[c.idx(i, j, k)] = 0.5 * std::sqrt(s);
          }
    }
  }
  gc_tree(c, AFGPU_CC_EFIELD);
}

// field_compute, src/m_field.f90:448-528
static int field_compute(Ctx &c, int s_in, bool have_guess, int *n_cycles, double *res_out) {
  field_set_rhs(c, s_in);
  double max_rhs = tree_maxabs(c, AFGPU_CC_RHS);
  const double conv_fac = 1e-10, min_residual = 1e-6, max_residual = 1e8;
  double threshold = std::max(std::max(min_residual, max_rhs * c.p.mg_max_rel_residual),
                              conv_fac * std::fabs(c.voltage) / (c.p.domain_len_last * min_dr(c)));
  MG mg = mg_of(c, AFGPU_OP_POISSON);
  std::vector<double> residuals(101, 0.0);
  int count = 0;
  double last = 0;
  if (!have_guess) {
    int i;
    for (i = 1; i <= 100; i++) {
      fmg(c, mg, true, true);
      residuals[i] = tree_maxabs(c, AFGPU_CC_TMP);
      count++;
      last = residuals[i];
      if (residuals[i] < threshold) break;
      if (i > 2) {
        double mn = std::min(std::min(residuals[i - 2], residuals[i - 1]), residuals[i]);
        double mx = std::max(std::max(residuals[i - 2], residuals[i - 1]), residuals[i]);
        double ratio = mn / mx;
        if (ratio < 2.0 && ratio > 0.5 && residuals[i] < max_residual) break;
      }
    }
    if (i == 101) {
      set_error("No convergence in initial field computation");
      return AFGPU_ERR_NO_CONVERGENCE;
    }
  }
  for (int i = 1; i <= c.p.mg_num_vcycles; i++) {
    vcycle(c, mg, true, c.highest_lvl);
    last = tree_maxabs(c, AFGPU_CC_TMP);
    count++;
    if (last < threshold) break;
  }
  field_from_potential(c);
  if (n_cycles) *n_cycles = count;
  if (res_out) *res_out = last;
  return AFGPU_OK;
}

// ---- drift-diffusion flux ---------------------------------------------------
// af_restrict_ref_boundary, afivo/src/m_af_restrict.f90:140-161
static void restrict_ref_boundary(Ctx &c, int iv) {
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
    const std::vector<int> &lv = c.leaves[lvl];
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < lv.size(); i++) {
      int id = lv[i];
      const Box &bx = c.B(id);
      bool at_rb = false;
      for (int nb = 0; nb < 2 * c.D; nb++)
        if (bx.nb[nb] == 0) at_rb = true;
      if (bx.parent > 0 && at_rb) restrict_box(c, id, iv, iv, true);
    }
  }
}

// af_gc2_box (afivo/src/m_af_ghostcell.f90:672-744) into g[(nc+4)^D],
// offset 2 (index -1 maps to 0).
static void gc2_box(Ctx &c, int id, int iv, std::vector<double> &g) {
  const int nc = c.nc, D = c.D, ng4 = nc + 4;
  const Box &bx = c.B(id);
  double *u = c.C(iv, id);
  VarMethods m = methods(c, iv);
  auto G = [&](int i, int j, int k) -> double & {
    return g[(size_t)(i + 1) + ng4 * ((size_t)(j + 1) + ng4 * (size_t)(D == 3 ? k + 1 : 0))];
  };
  const int kmin = (D == 3) ? 0 : 0, kmax = (D == 3) ? nc + 1 : 0;
  for (int k = kmin; k <= kmax; k++)
    for (int j = 0; j <= nc + 1; j++)
      for (int i = 0; i <= nc + 1; i++) G(i, j, k) = u[c.idx(i, j, k)];
  for (int nb = 0; nb < 2 * D; nb++) {
    const int d = nb / 2;
    const bool low = (nb % 2) == 0;
    int t0, t1;
    tangential(D, d, t0, t1);
    const int bmax = (D == 3) ? nc : 1;
    const int nb_id = bx.nb[nb];
    if (nb_id > 0) {
      const double *w = c.C(iv, nb_id);
      const int dnb = low ? -1 : 1;
      for (int layer = 1; layer <= 2; layer++) {
        int gi = low ? 1 - layer : nc + layer;
        for (int b = 1; b <= bmax; b++)
          for (int a = 1; a <= nc; a++) {
            int p[3] = {0, 0, 0};
            p[t0] = a;
            if (t1 >= 0) p[t1] = b;
            p[d] = gi;
            int q[3] = {p[0], p[1], p[2]};
            q[d] = gi - dnb * nc;
            G(p[0], p[1], p[2]) = w[c.idx(q[0], q[1], q[2])];
          }
      }
    } else if (nb_id == 0) {
      // gc2_prolong_rb, afivo/src/m_af_ghostcell.f90:753-856
      const int hnc = nc / 2;
      const int p_nb_id = c.B(bx.parent).nb[nb];
      const double *P = c.C(iv, p_nb_id);
      int off[3] = {0, 0, 0};
      for (int dd = 0; dd < D; dd++) off[dd] = ((bx.ix[dd] - 1) & 1) * hnc;
      int lo[3] = {1, 1, 1}, hi[3] = {nc, nc, nc};
      if (D == 2) lo[2] = hi[2] = 0;
      lo[d] = low ? -1 : nc + 1;
      hi[d] = low ? 0 : nc + 2;
      int lo_c[3], hi_c[3];
      for (int dd = 0; dd < 3; dd++) {
        lo_c[dd] = off[dd] + (lo[dd] + 1) / 2;
        hi_c[dd] = off[dd] + (hi[dd] + 1) / 2;
      }
      const int dix = low ? -1 : 1;
      lo_c[d] -= dix * nc;
      hi_c[d] -= dix * nc;
      if (D == 2) {
        for (int j = lo_c[1]; j <= hi_c[1]; j++) {
          int jf = lo[1] + 2 * (j - lo_c[1]);
          for (int i = lo_c[0]; i <= hi_c[0]; i++) {
            int i_f = lo[0] + 2 * (i - lo_c[0]);
            double f0 = P[c.idx(i, j, 0)];
            double sa0 = P[c.idx(i, j, 0)] - P[c.idx(i - 1, j, 0)];
            double sa1 = P[c.idx(i, j, 0)] - P[c.idx(i, j - 1, 0)];
            double sb0 = P[c.idx(i + 1, j, 0)] - P[c.idx(i, j, 0)];
            double sb1 = P[c.idx(i, j + 1, 0)] - P[c.idx(i, j, 0)];
            double f1 = 0.25 * limiter_apply(sa0, sb0, m.limiter);
            double f2 = 0.25 * limiter_apply(sa1, sb1, m.limiter);
            G(i_f, jf, 0) = f0 - f1 - f2;
            G(i_f, jf + 1, 0) = f0 - f1 + f2;
            G(i_f + 1, jf, 0) = f0 + f1 - f2;
            G(i_f + 1, jf + 1, 0) = f0 + f1 + f2;
          }
        }
      } else {
        for (int k = lo_c[2]; k <= hi_c[2]; k++) {
          int kf = lo[2] + 2 * (k - lo_c[2]);
          for (int j = lo_c[1]; j <= hi_c[1]; j++) {
            int jf = lo[1] + 2 * (j - lo_c[1]);
            for (int i = lo_c[0]; i <= hi_c[0]; i++) {
              int i_f = lo[0] + 2 * (i - lo_c[0]);
              double f0 = P[c.idx(i, j, k)];
              double sa0 = P[c.idx(i, j, k)] - P[c.idx(i - 1, j, k)];
              double sa1 = P[c.idx(i, j, k)] - P[c.idx(i, j - 1, k)];
              double sa2 = P[c.idx(i, j, k)] - P[c.idx(i, j, k - 1)];
              double sb0 = P[c.idx(i + 1, j, k)] - P[c.idx(i, j, k)];
              double sb1 = P[c.idx(i, j + 1, k)] - P[c.idx(i, j, k)];
              double sb2 = P[c.idx(i, j, k + 1)] - P[c.idx(i, j, k)];
              double f1 = 0.25 * limiter_apply(sa0, sb0, m.limiter);
              double f2 = 0.25 * limiter_apply(sa1, sb1, m.limiter);
              double f3 = 0.25 * limiter_apply(sa2, sb2, m.limiter);
              G(i_f, jf, kf) = f0 - f1 - f2 - f3;
              G(i_f, jf, kf + 1) = f0 - f1 - f2 + f3;
              G(i_f, jf + 1, kf) = f0 - f1 + f2 - f3;
              G(i_f, jf + 1, kf + 1) = f0 - f1 + f2 + f3;
              G(i_f + 1, jf, kf) = f0 + f1 - f2 - f3;
              G(i_f + 1, jf, kf + 1) = f0 + f1 - f2 + f3;
              G(i_f + 1, jf + 1, kf) = f0 + f1 + f2 - f3;
              G(i_f + 1, jf + 1, kf + 1) = f0 + f1 + f2 + f3;
            }
          }
        }
      }
    } else {
      // bc_to_gc2, afivo/src/m_af_ghostcell.f90:282-375
      double c0, c1, c2;
      switch (m.bc_type[nb]) {
        case BC_DIRICHLET: c0 = 2; c1 = -1; c2 = c0; break;
        case BC_NEUMANN: c0 = bx.dr[d] * (low ? -1 : 1); c1 = 1; c2 = 3 * c0; break;
        default: c0 = 1; c1 = 0; c2 = c0; break;
      }
      // The reference's 3D high-y branch uses c0 for the second layer (:361).
      if (D == 3 && nb == 3) c2 = c0;
      const double bv = m.bc_val[nb];
      for (int b = 1; b <= bmax; b++)
        for (int a = 1; a <= nc; a++) {
          int p[3] = {0, 0, 0};
          p[t0] = a;
          if (t1 >= 0) p[t1] = b;
          int q0[3] = {p[0], p[1], p[2]}, q1[3] = {p[0], p[1], p[2]}, r1[3] = {p[0], p[1], p[2]},
              r2[3] = {p[0], p[1], p[2]};
          q0[d] = low ? 0 : nc + 1;
          q1[d] = low ? -1 : nc + 2;
          r1[d] = low ? 1 : nc;
          r2[d] = low ? 2 : nc - 1;
          G(q0[0], q0[1], q0[2]) = c0 * bv + c1 * G(r1[0], r1[1], r1[2]);
          G(q1[0], q1[1], q1[2]) = c2 * bv + c1 * G(r2[0], r2[1], r2[2]);
        }
    }
  }
  for (int k = kmin; k <= kmax; k++)
    for (int j = 0; j <= nc + 1; j++)
      for (int i = 0; i <= nc + 1; i++) u[c.idx(i, j, k)] = G(i, j, k);
}

// flux_upwind_box (afivo/src/m_af_flux_schemes.f90:715-848) with the
// electron flux of src/m_fluid.f90:102-227.
static void flux_upwind_box(Ctx &c, int id, int s_deriv, double *dt_lim) {
  const int nc = c.nc, D = c.D, ng4 = nc + 4;
  static thread_local std::vector<double> g, cfl;
  g.assign((size_t)ng4 * ng4 * (D == 3 ? ng4 : 1), 0.0);
  cfl.assign((size_t)nc * nc * (D == 3 ? nc : 1), 0.0);
  const int iv = AFGPU_CC_ELECTRON + s_deriv;
  gc2_box(c, id, iv, g);
  const Box &bx = c.B(id);
  const double *ne = c.C(iv, id);
  const double *E = c.C(AFGPU_CC_EFIELD, id);
  const double *Ef = c.F(AFGPU_FC_FIELD, id);
  double *flux = c.F(AFGPU_FC_FLUX, id);
  const double N_inv = 1 / c.p.gas_number_density;
  double drt_min = 1e100;
  auto G = [&](int i, int j, int k) {
    return g[(size_t)(i + 1) + ng4 * ((size_t)(j + 1) + ng4 * (size_t)(D == 3 ? k + 1 : 0))];
  };
  double line[36], ecc[36], necc[36], ex[36], u[36], v[36], dc[36], sg[36];
  for (int d = 0; d < D; d++) {
    int t0, t1;
    tangential(D, d, t0, t1);
    const double inv_dx = 1 / bx.dr[d];
    const int bmax = (D == 3) ? nc : 1;
    for (int b = 1; b <= bmax; b++)
      for (int a = 1; a <= nc; a++) {
        for (int q = -1; q <= nc + 2; q++) {
          int p[3] = {0, 0, 0};
          p[t0] = a;
          if (t1 >= 0) p[t1] = b;
          p[d] = q;
          line[q + 1] = G(p[0], p[1], p[2]);
          if (q >= 0 && q <= nc + 1) {
            long x = c.idx(p[0], p[1], p[2]);
            ecc[q] = E[x];
            necc[q] = ne[x];
          }
          if (q >= 1 && q <= nc + 1) {
            int fp[3] = {p[0], p[1], D == 3 ? p[2] : 1};
            ex[q] = Ef[c.fidx(fp[0], fp[1], fp[2], d)];
          }
        }
        auto L = [&](int q) { return line[q + 1]; };
        // reconstruct_upwind_1d (m_af_flux_schemes.f90:282-303)
        for (int f = 1; f <= nc + 1; f++) {
          bool pos = (-1 * ex[f] > 0);
          if (pos) {
            double aa = L(f) - L(f - 1), bb = L(f - 1) - L(f - 2);
            u[f] = L(f - 1) + 0.5 * limiter_koren(aa, bb);
          } else {
            double bb = L(f) - L(f - 1), aa = L(f + 1) - L(f);
            u[f] = L(f) - 0.5 * limiter_koren(bb, aa);
          }
        }
        double sig_max = -1e300;
        for (int f = 1; f <= nc + 1; f++) {
          double tmp_fc = 0.5 * (ecc[f - 1] + ecc[f]) * SI_to_Townsend * N_inv;
          int lix;
          double lw;
          c.td.loc(tmp_fc, lix, lw);
          double mu = c.td.at(0, lix, lw) * N_inv;
          dc[f] = c.td.at(1, lix, lw) * N_inv;
          v[f] = -mu * ex[f];
          double fl = v[f] * u[f] - dc[f] * inv_dx * (necc[f] - necc[f - 1]);
          sg[f] = mu * u[f];
          sig_max = std::max(sig_max, sg[f]);
          int fp[3] = {0, 0, 0};
          fp[t0] = a;
          if (t1 >= 0) fp[t1] = b;
          fp[d] = f;
          flux[c.fidx(fp[0], fp[1], D == 3 ? fp[2] : 1, d)] = fl;
        }
        for (int q = 1; q <= nc; q++) {
          double cl = 1.0 * std::max(std::fabs(v[q + 1]), std::fabs(v[q])) * inv_dx +
                      2 * std::max(dc[q + 1], dc[q]) * (inv_dx * inv_dx);
          int p[3] = {0, 0, 0};
          p[t0] = a;
          if (t1 >= 0) p[t1] = b;
          p[d] = q;
          size_t ci = (size_t)(p[0] - 1) + nc * ((size_t)(p[1] - 1) + nc * (size_t)(D == 3 ? p[2] - 1 : 0));
          cfl[ci] = cfl[ci] + cl;
        }
        double drt = UC_eps0 / (UC_elem_charge * std::max(sig_max, 1e-100));
        drt_min = std::min(drt_min, drt);
      }
  }
  double cmax = -std::numeric_limits<double>::infinity();
  for (double x : cfl) cmax = std::max(cmax, x);
  dt_lim[0] = 1 / cmax;
  dt_lim[1] = drt_min;
}

// af_consistent_fluxes + flux_from_children, afivo/src/m_af_core.f90:1257-1402
static void consistent_fluxes(Ctx &c) {
  const int nc = c.nc, nch = nc / 2, D = c.D;
  for (int lvl = 1; lvl <= c.highest_lvl - 1; lvl++) {
    const std::vector<int> &par = c.parents[lvl];
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < par.size(); n++) {
      int id = par[n];
      const Box &bx = c.B(id);
      for (int nb = 0; nb < 2 * D; nb++) {
        int nb_id = bx.nb[nb];
        if (nb_id <= 0 || c.has_children(nb_id)) continue;
        const int d = nb / 2;
        const bool low = (nb % 2) == 0;
        const int fi = low ? 1 : nc + 1, fi_nb = low ? nc + 1 : 1;
        double *fn = c.F(AFGPU_FC_FLUX, nb_id);
        const Box &bn = c.B(nb_id);
        for (int ic = 0; ic < (1 << (D - 1)); ic++) {
          int i_ch = (D == 2) ? child_adj_nb2[nb][ic] - 1 : child_adj_nb3[nb][ic] - 1;
          int c_id = bx.children[i_ch];
          const double *fc_ = c.F(AFGPU_FC_FLUX, c_id);
          int ioff[3] = {0, 0, 0};
          for (int dd = 0; dd < D; dd++)
            ioff[dd] = nch * ((D == 2) ? child_dix2[i_ch][dd] : child_dix3[i_ch][dd]);
          if (D == 2) {
            int t0 = 1 - d;
            for (int t = 1; t <= nch; t++) {
              int pn[2], pa[2], pb[2];
              pn[d] = fi_nb;
              pn[t0] = ioff[t0] + t;
              pa[d] = fi;
              pa[t0] = 2 * t - 1;
              pb[d] = fi;
              pb[t0] = 2 * t;
              double A = fc_[c.fidx(pa[0], pa[1], 1, d)], Bv = fc_[c.fidx(pb[0], pb[1], 1, d)];
              double val;
              if (d == 1 && bn.lvl > 0 && c.coord == AFGPU_CYL) {
                double r = bn.r_min[0] + ((ioff[0] + t) - 0.5) * bn.dr[0];
                double tmp = 0.25 * bn.dr[0] / r;
                double w1 = 1 - tmp, w2 = 1 + tmp;
                val = 0.5 * (w1 * A + w2 * Bv);
              } else {
                val = 0.5 * (A + Bv);
              }
              fn[c.fidx(pn[0], pn[1], 1, d)] = val;
            }
          } else {
            int t0, t1;
            tangential(3, d, t0, t1);
            for (int q = 1; q <= nch; q++)
              for (int p = 1; p <= nch; p++) {
                int pn[3], s00[3], s10[3], s01[3], s11[3];
                pn[d] = fi_nb;
                pn[t0] = ioff[t0] + p;
                pn[t1] = ioff[t1] + q;
                s00[d] = s10[d] = s01[d] = s11[d] = fi;
                s00[t0] = 2 * p - 1; s00[t1] = 2 * q - 1;
                s10[t0] = 2 * p;     s10[t1] = 2 * q - 1;
                s01[t0] = 2 * p - 1; s01[t1] = 2 * q;
                s11[t0] = 2 * p;     s11[t1] = 2 * q;
                double val = 0.25 * (fc_[c.fidx(s00[0], s00[1], s00[2], d)] +
                                     fc_[c.fidx(s10[0], s10[1], s10[2], d)] +
                                     fc_[c.fidx(s01[0], s01[1], s01[2], d)] +
                                     fc_[c.fidx(s11[0], s11[1], s11[2], d)]);
                fn[c.fidx(pn[0], pn[1], pn[2], d)] = val;
              }
          }
        }
      }
    }
  }
}

// flux_upwind_tree, afivo/src/m_af_flux_schemes.f90:666-712
static void flux_upwind_tree(Ctx &c, int s_deriv, double *dt_lim) {
  restrict_ref_boundary(c, AFGPU_CC_ELECTRON + s_deriv);
  double l0 = 1e100, l1 = 1e100;
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
    const std::vector<int> &lv = c.leaves[lvl];
#pragma omp parallel for schedule(static) reduction(min : l0, l1)
    for (size_t i = 0; i < lv.size(); i++) {
      double my[2];
      flux_upwind_box(c, lv[i], s_deriv, my);
      l0 = std::min(l0, my[0]);
      l1 = std::min(l1, my[1]);
    }
  }
  consistent_fluxes(c);
  dt_lim[0] = l0;
  dt_lim[1] = l1;
}

// The rate coefficient of reaction R at field f (Td) and electron
// temperature te, get_rates (src/m_chemistry.f90:699-767)
static double chem_rate_cell(const afcpu_chem &c, const afgpu_reaction_t &R, double f, double te) {
  const double kB = 1.380649e-23, eV = 1.602176634e-19;  // src/m_units_constants.f90:15-39
  const double c0 = R.rate_factor;
  const double *k = R.rate_data;
  const double tg = c.tg;
  switch (R.rate_type) {
    case AFGPU_RATE_TABULATED_FIELD: return c0 * c.fld.col(R.lookup_col, f);
    case AFGPU_RATE_CONSTANT: return c0 * k[0];
    case AFGPU_RATE_LINEAR: return c0 * k[0] * (f - k[1]);
    case AFGPU_RATE_EXP_V1: {
      const double q = k[1] / (k[2] + f);
      return c0 * k[0] * std::exp(-(q * q));
    }
    case AFGPU_RATE_EXP_V2: {
      const double q = f / k[1];
      return c0 * k[0] * std::exp(-(q * q));
    }
    case AFGPU_RATE_K1: return c0 * k[0] * std::pow(300 / te, k[1]);
    case AFGPU_RATE_K3: {
      const double q = (kB / eV) * te + k[1];
      return c0 * (k[0] * (q * q) - k[2]) * k[3];
    }
    case AFGPU_RATE_K4: return c0 * k[0] * std::pow(tg / 300, k[1]) * std::exp(-k[2] / tg);
    case AFGPU_RATE_K5: return c0 * k[0] * std::exp(-k[1] / tg);
    case AFGPU_RATE_K6: return c0 * k[0] * std::pow(tg, k[1]);
    case AFGPU_RATE_K7: return c0 * k[0] * std::pow(tg / k[1], k[2]);
    case AFGPU_RATE_K8: return c0 * k[0] * std::pow(300 / tg, k[1]);
    case AFGPU_RATE_K9: return c0 * k[0] * std::exp(-k[1] * tg);
    case AFGPU_RATE_K10: return c0 * std::pow(10.0, k[0] + k[1] * (tg - 300));
    case AFGPU_RATE_K11: return c0 * k[0] * std::pow(300 / tg, k[1]) * std::exp(-k[2] / tg);
    case AFGPU_RATE_K12: return c0 * k[0] * std::pow(tg, k[1]) * std::exp(-k[2] / tg);
    case AFGPU_RATE_K13: return c0 * k[0] * std::exp(-std::pow(k[1] / (k[2] + f), k[3]));
    case AFGPU_RATE_K14: return c0 * k[0] * std::exp(-std::pow(f / k[1], k[2]));
    case AFGPU_RATE_K15: return c0 * k[0] * std::exp(-std::pow(k[1] / (kB * (tg + f / k[2])), k[3]));
    case AFGPU_RATE_K16:
      return c0 * k[0] * std::pow(300 / te, k[1]) * std::exp(-k[2] / tg) * std::exp(k[3] * (te - tg) / (te * tg));
  }
  return 0;
}
static double chem_te(const afcpu_chem &c, double f) {
  const double kB = 1.380649e-23, eV = 1.602176634e-19;
  const double electron_eV_to_K = 2 * eV / (3 * kB);
  return c.need_te ? electron_eV_to_K * c.energy.col(0, f) : 0.0;
}

// add_source_terms (src/m_fluid.f90:298-466) for a reaction list over
// c.n_sp species on one box, after the states' combination of
// flux_update_densities; returns the box's chemistry dt limit.
static double update_species_box(Ctx &c, int id, double dt, int s_deriv, int n_prev, const int *s_prev,
                                 const double *w_prev, int s_out, bool last_step) {
  const int nc = c.nc, D = c.D, ns = c.n_sp;
  const int kmin = (D == 3) ? 1 : 0, kmax = (D == 3) ? nc : 0;
  const double eps = 1e-100;
  for (int sp = 0; sp < ns; sp++) {
    double *out = c.SP(sp, s_out, id);
    for (int k = kmin; k <= kmax; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          long x = c.idx(i, j, k);
          double tmp = 0.0;
          for (int m = 0; m < n_prev; m++) tmp = tmp + w_prev[m] * c.SP(sp, s_prev[m], id)[x];
          out[x] = tmp;
        }
  }
  const double Ninv = 1 / c.p.gas_number_density;
  const double *E = c.C(AFGPU_CC_EFIELD, id);
  const double *photo = c.C(AFGPU_CC_PHOTO, id);
  const afcpu_chem &net = *c.net;
  double box_min = 1e100;
  std::vector<double> der((size_t)ns * nc * nc * (D == 3 ? nc : 1)), dens(ns), dd(ns);
  size_t cell = 0;
  for (int k = kmin; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++, cell++) {
        long x = c.idx(i, j, k);
        const double field = SI_to_Townsend * Ninv * E[x];
        for (int sp = 0; sp < ns; sp++) {
          dens[sp] = std::max(c.SP(sp, s_deriv, id)[x], 0.0);
          dd[sp] = 0.0;
        }
        const double te = chem_te(net, field);
        for (const afgpu_reaction_t &R : net.r) {
          double kr = chem_rate_cell(net, R, field, te);
          double prod = 1.0;
          for (int q = 0; q < R.n_in; q++) prod = prod * dens[R.ix_in[q]];
          kr = kr * prod;
          for (int q = 0; q < R.n_in; q++) dd[R.ix_in[q]] = dd[R.ix_in[q]] - kr;
          for (int q = 0; q < R.n_out; q++) dd[R.ix_out[q]] = dd[R.ix_out[q]] + kr * R.multiplicity_out[q];
        }
        if (last_step)
          for (int sp = 0; sp < ns; sp++) {
            double t;
            if (c.p.dt_chemistry_nmin > 0)
              t = (dens[sp] + c.p.dt_chemistry_nmin) / std::max(std::fabs(dd[sp]), eps);
            else if (c.p.dt_chemistry_limit_loss)
              t = std::max(dens[sp], eps) / std::max(-dd[sp], eps);
            else
              t = 1e100;
            box_min = std::min(box_min, t);
          }
        if (c.p.photoi_enabled) {
          dd[0] = dd[0] + photo[x];
          dd[c.p.photoi_species] = dd[c.p.photoi_species] + photo[x];
        }
        for (int sp = 0; sp < ns; sp++) der[(size_t)ns * cell + sp] = dd[sp];
      }
  for (int sp = 0; sp < ns; sp++) {
    double *out = c.SP(sp, s_out, id);
    cell = 0;
    for (int k = kmin; k <= kmax; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++, cell++) {
          long x = c.idx(i, j, k);
          out[x] = out[x] + dt * der[(size_t)ns * cell + sp];
        }
  }
  return box_min;
}

// The conservative flux divergence of the electrons on one box
// (flux_update_densities, afivo/src/m_af_flux_schemes.f90:320-436).
static void electron_flux_divergence(Ctx &c, int id, const double *dt_dr, int s_out) {
  const int nc = c.nc, D = c.D;
  const Box &bx = c.B(id);
  double *ne = c.C(AFGPU_CC_ELECTRON + s_out, id);
  const double *fl = c.F(AFGPU_FC_FLUX, id);
  if (D == 2) {
    double rf1[33], rf2[33];
    const bool cyl = (c.coord == AFGPU_CYL);
    if (cyl) cyl_factors(c, bx, rf1, rf2);
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) {
        long x = c.idx(i, j, 0);
        if (cyl)
          ne[x] = ne[x] + dt_dr[0] * (rf1[i] * fl[c.fidx(i, j, 1, 0)] - rf2[i] * fl[c.fidx(i + 1, j, 1, 0)]) +
                  dt_dr[1] * (fl[c.fidx(i, j, 1, 1)] - fl[c.fidx(i, j + 1, 1, 1)]);
        else
          ne[x] = ne[x] + dt_dr[0] * (fl[c.fidx(i, j, 1, 0)] - fl[c.fidx(i + 1, j, 1, 0)]) +
                  dt_dr[1] * (fl[c.fidx(i, j, 1, 1)] - fl[c.fidx(i, j + 1, 1, 1)]);
      }
  } else {
    for (int k = 1; k <= nc; k++)
      for (int j = 1; j <= nc; j++)
        for (int i = 1; i <= nc; i++) {
          long x = c.idx(i, j, k);
          ne[x] = ne[x] + dt_dr[0] * (fl[c.fidx(i, j, k, 0)] - fl[c.fidx(i + 1, j, k, 0)]) +
                  dt_dr[1] * (fl[c.fidx(i, j, k, 1)] - fl[c.fidx(i, j + 1, k, 1)]) +
                  dt_dr[2] * (fl[c.fidx(i, j, k, 2)] - fl[c.fidx(i, j, k + 1, 2)]);
        }
  }
}

// flux_update_densities (afivo/src/m_af_flux_schemes.f90:320-436) with
// add_source_terms (src/m_fluid.f90:298-466), get_rates / get_derivatives
// for the standard model (src/m_chemistry.f90:220-258, 672-804).
static void flux_update_densities(Ctx &c, double dt, int s_deriv, int n_prev, const int *s_prev,
                                  const double *w_prev, int s_out, bool last_step, double *dt_lim) {
  const int nc = c.nc, D = c.D;
  const int kmin = (D == 3) ? 1 : 0, kmax = (D == 3) ? nc : 0;
  double chem_min = 1e100;
  const double eps = 1e-100;
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
    const std::vector<int> &lv = c.leaves[lvl];
#pragma omp parallel for schedule(static) reduction(min : chem_min)
    for (size_t n = 0; n < lv.size(); n++) {
      int id = lv[n];
      const Box &bx = c.B(id);
      double dt_dr[3];
      for (int d = 0; d < D; d++) dt_dr[d] = dt / bx.dr[d];
      if (c.gen_sp) {
        chem_min = std::min(chem_min, update_species_box(c, id, dt, s_deriv, n_prev, s_prev, w_prev, s_out, last_step));
        electron_flux_divergence(c, id, dt_dr, s_out);
        continue;
      }
      for (int sp = 0; sp < 3; sp++) {
        const int iv = AFGPU_CC_SPECIES(sp);
        double *out = c.C(iv + s_out, id);
        for (int k = kmin; k <= kmax; k++)
          for (int j = 1; j <= nc; j++)
            for (int i = 1; i <= nc; i++) {
              long x = c.idx(i, j, k);
              double tmp = 0.0;
              for (int m = 0; m < n_prev; m++) tmp = tmp + w_prev[m] * c.C(iv + s_prev[m], id)[x];
              out[x] = tmp;
            }
      }
      // add_source_terms
      const double Ninv = 1 / c.p.gas_number_density;
      const double *E = c.C(AFGPU_CC_EFIELD, id);
      const double *photo = c.C(AFGPU_CC_PHOTO, id);
      double box_min = 1e100;
      std::vector<double> der(3 * nc * nc * (D == 3 ? nc : 1));
      size_t cell = 0;
      for (int k = kmin; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++, cell++) {
            long x = c.idx(i, j, k);
            double field = SI_to_Townsend * Ninv * E[x];
            double dens[3];
            for (int sp = 0; sp < 3; sp++) dens[sp] = std::max(c.C(AFGPU_CC_SPECIES(sp) + s_deriv, id)[x], 0.0);
            double dd[3] = {0.0, 0.0, 0.0};
            if (c.net) {
              // get_rates and get_derivatives of the list (m_chemistry.f90:672-804)
              const double te = chem_te(*c.net, field);
              for (const afgpu_reaction_t &R : c.net->r) {
                double kr = chem_rate_cell(*c.net, R, field, te);
                double prod = 1.0;
                for (int q = 0; q < R.n_in; q++) prod = prod * dens[R.ix_in[q]];
                kr = kr * prod;
                for (int q = 0; q < R.n_in; q++) dd[R.ix_in[q]] = dd[R.ix_in[q]] - kr;
                for (int q = 0; q < R.n_out; q++) dd[R.ix_out[q]] = dd[R.ix_out[q]] + kr * R.multiplicity_out[q];
              }
            } else {
              double k_ion = 1.0 * c.chem.col(0, field);
              double k_att = 1.0 * c.chem.col(1, field);
              double r1 = k_ion * dens[0];
              double r2 = k_att * dens[0];
              double de = 0.0, dp = 0.0, dm = 0.0;
              de = de - r1;
              de = de + r1 * 2;
              dp = dp + r1 * 1;
              de = de - r2;
              dm = dm + r2 * 1;
              dd[0] = de;
              dd[1] = dp;
              dd[2] = dm;
            }
            if (last_step) {
              for (int sp = 0; sp < 3; sp++) {
                double t;
                if (c.p.dt_chemistry_nmin > 0)
                  t = (dens[sp] + c.p.dt_chemistry_nmin) / std::max(std::fabs(dd[sp]), eps);
                else if (c.p.dt_chemistry_limit_loss)
                  t = std::max(dens[sp], eps) / std::max(-dd[sp], eps);
                else
                  t = 1e100;
                box_min = std::min(box_min, t);
              }
            }
            if (c.p.photoi_enabled) {
              dd[0] = dd[0] + photo[x];
              dd[c.p.photoi_species] = dd[c.p.photoi_species] + photo[x];
            }
            for (int sp = 0; sp < 3; sp++) der[3 * cell + sp] = dd[sp];
          }
      for (int sp = 0; sp < 3; sp++) {
        double *out = c.C(AFGPU_CC_SPECIES(sp) + s_out, id);
        cell = 0;
        for (int k = kmin; k <= kmax; k++)
          for (int j = 1; j <= nc; j++)
            for (int i = 1; i <= nc; i++, cell++) {
              long x = c.idx(i, j, k);
              out[x] = out[x] + dt * der[3 * cell + sp];
            }
      }
      chem_min = std::min(chem_min, box_min);
      // conservative flux divergence for the electrons
      double *ne = c.C(AFGPU_CC_ELECTRON + s_out, id);
      const double *fl = c.F(AFGPU_FC_FLUX, id);
      if (D == 2) {
        double rf1[33], rf2[33];
        const bool cyl = (c.coord == AFGPU_CYL);
        if (cyl) cyl_factors(c, bx, rf1, rf2);
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            long x = c.idx(i, j, 0);
            if (cyl)
              ne[x] = ne[x] + dt_dr[0] * (rf1[i] * fl[c.fidx(i, j, 1, 0)] - rf2[i] * fl[c.fidx(i + 1, j, 1, 0)]) +
                      dt_dr[1] * (fl[c.fidx(i, j, 1, 1)] - fl[c.fidx(i, j + 1, 1, 1)]);
            else
              ne[x] = ne[x] + dt_dr[0] * (fl[c.fidx(i, j, 1, 0)] - fl[c.fidx(i + 1, j, 1, 0)]) +
                      dt_dr[1] * (fl[c.fidx(i, j, 1, 1)] - fl[c.fidx(i, j + 1, 1, 1)]);
          }
      } else {
        for (int k = 1; k <= nc; k++)
          for (int j = 1; j <= nc; j++)
            for (int i = 1; i <= nc; i++) {
              long x = c.idx(i, j, k);
              ne[x] = ne[x] + dt_dr[0] * (fl[c.fidx(i, j, k, 0)] - fl[c.fidx(i + 1, j, k, 0)]) +
                      dt_dr[1] * (fl[c.fidx(i, j, k, 1)] - fl[c.fidx(i, j + 1, k, 1)]) +
                      dt_dr[2] * (fl[c.fidx(i, j, k, 2)] - fl[c.fidx(i, j, k + 1, 2)]);
            }
      }
    }
  }
  dt_lim[0] = chem_min;
  dt_lim[1] = 1e100;
}

// forward_euler, src/m_fluid.f90:21-99
static int forward_euler(Ctx &c, double dt, int s_deriv, int n_prev, const int *s_prev,
                         const double *w_prev, int s_out, int i_step, int n_steps, double *dt_lim) {
  bool last_step = (i_step == n_steps);
  if (i_step > 1) {
    int rc = field_compute(c, s_deriv, true, nullptr, nullptr);
    if (rc) return rc;
  }
  double lim[4];
  flux_upwind_tree(c, s_deriv, lim);
  flux_update_densities(c, dt, s_deriv, n_prev, s_prev, w_prev, s_out, last_step, lim + 2);
  lim[0] = lim[0] * c.p.dt_cfl_number;
  double mv = std::min(std::min(std::min(lim[0], lim[1]), lim[2]), lim[3]);
  *dt_lim = std::min(c.p.dt_max, mv);
  return AFGPU_OK;
}

// af_advance, afivo/src/m_af_advance.f90:121-214
static int advance(Ctx &c, double dt, double *dt_lim) {
  *dt_lim = 1e100;
  int rc;
  if (c.p.time_integrator == AFGPU_FORWARD_EULER) {
    int sp[1] = {0};
    double w[1] = {1.0};
    rc = forward_euler(c, dt, 0, 1, sp, w, 0, 1, 1, dt_lim);
  } else {
    int sp1[1] = {0};
    double w1[1] = {1.0};
    rc = forward_euler(c, dt, 0, 1, sp1, w1, 1, 1, 2, dt_lim);
    if (rc) return rc;
    int sp2[2] = {0, 1};
    double w2[2] = {0.5, 0.5};
    rc = forward_euler(c, 0.5 * dt, 1, 2, sp2, w2, 0, 2, 2, dt_lim);
  }
  c.time = c.time + dt;
  return rc;
}

// photoi_set_src (src/m_photoi.f90:153-213, 259-291) + photoi_helmh_compute
// (src/m_photoi_helmh.f90:162-204)
static int photoi_set_src(Ctx &c) {
  const int nc = c.nc, D = c.D;
  const int kmin = (D == 3) ? 1 : 0, kmax = (D == 3) ? nc : 0;
  const double coeff = c.p.photoi_eta_quench;
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
    const std::vector<int> &lv = c.leaves[lvl];
#pragma omp parallel for schedule(static)
    for (size_t n = 0; n < lv.size(); n++) {
      int id = lv[n];
      const double *E = c.C(AFGPU_CC_EFIELD, id);
      const double *ne = c.C(AFGPU_CC_ELECTRON, id);
      double *rh = c.C(AFGPU_CC_RHS, id);
      for (int k = kmin; k <= kmax; k++)
        for (int j = 1; j <= nc; j++)
          for (int i = 1; i <= nc; i++) {
            long x = c.idx(i, j, k);
            double fld = E[x];
            double Td = fld * SI_to_Townsend / c.p.gas_number_density;
            int lix;
            double lw;
            c.td.loc(Td, lix, lw);
            double alpha = c.td.at(2, lix, lw);
            double mob = c.td.at(0, lix, lw);
            double tmp = fld * mob * alpha * ne[x] * coeff;
            if (tmp < 0) tmp = 0;
            rh[x] = tmp;
          }
    }
  }
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++)
    for (int id : c.ids[lvl]) box_clear(c, id, AFGPU_CC_PHOTO);
  double max_rhs = tree_maxabs(c, AFGPU_CC_RHS);
  max_rhs = std::max(max_rhs, std::sqrt(std::numeric_limits<double>::epsilon()));
  c.photoi_cycles = 0;
  for (int n = 0; n < c.p.photoi_n_modes; n++) {
    MG mg = mg_of(c, AFGPU_OP_HELMH + n);
    for (int i = 1; i <= c.p.photoi_max_fmg_cycles; i++) {
      c.photoi_cycles++;
      fmg(c, mg, true, true);
      double residu = tree_maxabs(c, AFGPU_CC_TMP);
      if (residu / max_rhs < c.p.photoi_max_rel_residual) break;
    }
    const double cf = c.p.photoi_coeff[n];
    for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
      const std::vector<int> &lv = c.leaves[lvl];
#pragma omp parallel for schedule(static)
      for (size_t q = 0; q < lv.size(); q++) {
        double *ph = c.C(AFGPU_CC_PHOTO, lv[q]);
        const double *psi = c.C(mg.i_phi, lv[q]);
        for (size_t x = 0; x < c.ncell; x++) ph[x] = ph[x] - cf * psi[x];
      }
    }
  }
  return AFGPU_OK;
}

static void copy_states(Ctx &c, bool backup) {  // src/streamer.f90:659-688
  const int ns = c.n_states();
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
    const std::vector<int> &ids = c.ids[lvl];
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < ids.size(); i++) {
      for (int sp = 0; sp < c.n_sp; sp++) {
        double *a = c.SP(sp, 0, ids[i]), *b = c.SP(sp, ns, ids[i]);
        if (backup)
          std::copy(a, a + c.ncell, b);
        else
          std::copy(b, b + c.ncell, a);
      }
      if (backup)
        box_copy(c, ids[i], AFGPU_CC_PHI, AFGPU_CC_PHI_2);
      else
        box_copy(c, ids[i], AFGPU_CC_PHI_2, AFGPU_CC_PHI);
    }
  }
  if (!backup) field_from_potential(c);
}

// One iteration of the time loop, src/streamer.f90:181-365.
static int step(Ctx &c, afgpu_step_info_t *info) {
  if (!c.started) {
    c.global_dt = c.p.dt_min;  // src/streamer.f90:459
    c.dt = c.global_dt;
    c.started = true;
  }
  c.it = c.it + 1;
  int photoi_done = 0;
  if (c.p.photoi_enabled && c.p.photoi_per_steps > 0 && c.it % c.p.photoi_per_steps == 0) {
    int rc = photoi_set_src(c);
    if (rc) return rc;
    photoi_done = 1;
  }
  double dt_lim = std::numeric_limits<double>::max();
  double dt = c.dt;
  int n;
  const int max_attempts = 10;
  int n_rej = 0;
  for (n = 1; n <= max_attempts; n++) {
    copy_states(c, true);
    double dt_lim_step;
    int rc = advance(c, dt, &dt_lim_step);
    if (rc) return rc;
    dt_lim = std::min(dt_lim, dt_lim_step);
    if (dt <= dt_lim_step) break;
    n_rej++;
    dt = c.p.dt_safety_factor * dt_lim_step;
    c.time = c.global_time;
    copy_states(c, false);
  }
  c.frac_rejected = 0.99 * c.frac_rejected;
  if (n > 1) c.frac_rejected = c.frac_rejected + 0.01;
  if (n == max_attempts + 1) {
    set_error("All time steps were rejected");
    return AFGPU_ERR_ALL_REJECTED;
  }
  const double dt_used = dt;
  int n_cyc = 0;
  double res = 0;
  int rc = field_compute(c, 0, true, &n_cyc, &res);
  if (rc) return rc;
  double tmp = c.p.dt_max_growth_factor;
  if (c.frac_rejected > 0.1) tmp = 1.0;
  dt = std::min(tmp * c.global_dt, c.p.dt_safety_factor * std::min(dt_lim, c.p.dt_max));
  c.global_dt = dt;
  c.global_time = c.time;
  c.dt = dt;
  if (info) {
    info->it = c.it;
    info->time = c.time;
    info->dt_used = dt_used;
    info->dt_next = dt;
    info->dt_lim = dt_lim;
    info->n_rejected = n_rej;
    info->n_vcycles = n_cyc;
    info->residual = res;
    double mr = tree_maxabs(c, AFGPU_CC_RHS);
    info->rel_residual = mr > 0 ? res / mr : 0;
    info->photoi_done = photoi_done;
    info->photoi_cycles = photoi_done ? c.photoi_cycles : 0;
    long long cells = 0;
    for (int lvl = 1; lvl <= c.highest_lvl; lvl++) cells += (long long)c.leaves[lvl].size();
    for (int d = 0; d < c.D; d++) cells *= c.nc;
    info->leaf_cells = cells;
  }
  if (c.global_dt < c.p.dt_min) {
    set_error("dt too small");
    return AFGPU_ERR_DT_TOO_SMALL;
  }
  return AFGPU_OK;
}

static int sync_tree(Ctx &c, int n_boxes, const int *lvl, const int *ix, const int *parent,
                     const int *children, const int *neighbors, const int *nbmat, const double *r_min,
                     const double *dr) {
  const int D = c.D, nch = 1 << D, nnb = 2 * D, nm = (D == 3) ? 27 : 9;
  std::vector<Box> nb(n_boxes);
  int hl = 0;
  for (int b = 0; b < n_boxes; b++) {
    Box &x = nb[b];
    x.lvl = lvl[b];
    for (int d = 0; d < D; d++) {
      x.ix[d] = ix[b * D + d];
      x.r_min[d] = r_min[b * D + d];
      x.dr[d] = dr[b * D + d];
    }
    x.parent = parent[b];
    for (int q = 0; q < nch; q++) x.children[q] = children[b * nch + q];
    for (int q = 0; q < nnb; q++) x.nb[q] = neighbors[b * nnb + q];
    for (int q = 0; q < nm; q++) x.nbmat[q] = nbmat[b * nm + q];
    hl = std::max(hl, x.lvl);
  }
  c.boxes = nb;
  c.n_boxes = n_boxes;
  c.highest_lvl = hl;
  c.ids.assign(hl + 1, {});
  c.leaves.assign(hl + 1, {});
  c.parents.assign(hl + 1, {});
  for (int b = 0; b < n_boxes; b++) {
    int l = nb[b].lvl;
    if (l <= 0) continue;
    c.ids[l].push_back(b + 1);
    if (nb[b].children[0] > 0)
      c.parents[l].push_back(b + 1);
    else
      c.leaves[l].push_back(b + 1);
  }
  for (int iv = 0; iv < AFGPU_N_CC; iv++) c.cc[iv].resize((size_t)n_boxes * c.ncell, 0.0);
  for (int f = 0; f < AFGPU_N_FC; f++) c.fc[f].resize((size_t)n_boxes * D * c.nface, 0.0);
  for (auto &x : c.xs) x.resize((size_t)n_boxes * c.ncell, 0.0);
  if (hl < 1 || c.ids[1].size() != 1) {
    set_error("level 1 must consist of exactly one box");
    return AFGPU_ERR_UNSUPPORTED;
  }
  return AFGPU_OK;
}

}  // namespace afcpu

using namespace afcpu;

struct afcpu_ctx {
  Ctx c;
};
typedef struct afcpu_ctx afcpu_t;

#define CHECK_H(h)                          \
  do {                                      \
    if (!(h)) {                             \
      set_error("null handle");             \
      return AFGPU_ERR_ARG;                 \
    }                                       \
  } while (0)

namespace afcpu {

// ---- refinement flags -------------------------------------------------------
// af_rm_ref, af_keep_ref, af_do_ref, af_derefine, af_refine
// (afivo/src/m_af_types.f90:23-35)
constexpr int REF_RM = -1, REF_KEEP = 0, REF_DO = 1, REF_DEREFINE = -2, REF_REFINE = 2;

// GM_dist_vec_line / GM_dist_line (src/m_geometry.f90:23-51): distance of r
// to the segment r0-r1; the length as the square root of the sum of squares
static double dist_line(int D, const double *r, const double *r0, const double *r1) {
  double len2 = 0, frac = 0;
  for (int d = 0; d < D; d++) len2 = len2 + (r1[d] - r0[d]) * (r1[d] - r0[d]);
  for (int d = 0; d < D; d++) frac = frac + (r[d] - r0[d]) * (r1[d] - r0[d]);
  double v[3];
  for (int d = 0; d < D; d++) {
    if (frac <= 0)
      v[d] = r[d] - r0[d];
    else if (frac >= len2)
      v[d] = r[d] - r1[d];
    else
      v[d] = r[d] - (r0[d] + frac / len2 * (r1[d] - r0[d]));
  }
  double s2 = 0;
  for (int d = 0; d < D; d++) s2 = s2 + v[d] * v[d];
  return std::sqrt(s2);
}

// default_refinement (src/m_refine.f90:198-298): the flags of the cells of
// box id, stored i fastest.  The gas density is constant; |E| and n_e are
// the box's own values.
static void cell_flags(const Ctx &c, const afgpu_refine_t &rp, int id, std::vector<int> &cf) {
  const Box &b = c.B(id);
  const int D = c.D, nc = c.nc;
  double min_dx = b.dr[0], max_dx = b.dr[0];
  for (int d = 1; d < D; d++) {
    min_dx = std::min(min_dx, b.dr[d]);
    max_dx = std::max(max_dx, b.dr[d]);
  }
  const double gas = c.p.gas_number_density;
  const double *E = &c.cc[AFGPU_CC_EFIELD][(size_t)(id - 1) * c.ncell];
  const double *ne = &c.cc[AFGPU_CC_ELECTRON][(size_t)(id - 1) * c.ncell];
  const int kmax = (D == 3) ? nc : 1;
  int q = 0;
  for (int k = 1; k <= kmax; k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++, q++) {
        const long x = c.idx(i, j, D == 3 ? k : 0);
        const double fld = E[x] * SI_to_Townsend / gas;
        double alpha;
        if (rp.use_alpha_effective) {
          alpha = (c.td.col(2, rp.refine_adx_fac * fld) - c.td.col(3, rp.refine_adx_fac * fld)) * gas /
                  rp.refine_adx_fac;
          alpha = std::max(0.0, alpha);
        } else {
          alpha = c.td.col(2, rp.refine_adx_fac * fld) * gas / rp.refine_adx_fac;
        }
        const double adx = max_dx * alpha;
        int f;
        if (adx > rp.refine_adx && ne[x] > rp.refine_min_dens)
          f = REF_DO;
        else if (adx < 0.125 * rp.refine_adx && max_dx < rp.derefine_dx)
          f = REF_RM;
        else
          f = REF_KEEP;
        if (rp.global_time < rp.refine_init_time) {
          const int ijk[3] = {i, j, k};
          double r[3];
          for (int d = 0; d < D; d++) r[d] = b.r_min[d] + (ijk[d] - 0.5) * b.dr[d];
          for (int n = 0; n < rp.n_seeds; n++) {
            const double dist = dist_line(D, r, rp.seed_r0[n], rp.seed_r1[n]);
            if (dist - rp.seed_width[n] < 2 * max_dx && max_dx > rp.refine_init_fac * rp.seed_width[n]) f = REF_DO;
          }
        }
        cf[q] = f;
      }
  double rmin[3], rmax[3];
  for (int d = 0; d < D; d++) {
    rmin[d] = b.r_min[d];
    rmax[d] = b.r_min[d] + b.dr[d] * nc;
  }
  for (int n = 0; n < rp.n_regions; n++) {
    bool all = true;
    for (int d = 0; d < D; d++) all = all && rmax[d] >= rp.regions_rmin[n][d] && rmin[d] <= rp.regions_rmax[n][d];
    if (rp.global_time <= rp.regions_tstop[n] && max_dx > rp.regions_dr[n] && all) {
      const int h = nc / 2 - 1;  // the centre cell (nc/2, nc/2[, nc/2])
      cf[h + nc * (h + (D == 3 ? nc * h : 0))] = REF_DO;
    }
  }
  for (int n = 0; n < rp.n_limits; n++) {
    bool all = true;
    for (int d = 0; d < D; d++) all = all && rmin[d] >= rp.limits_rmin[n][d] && rmax[d] <= rp.limits_rmax[n][d];
    if (max_dx < 2 * rp.limits_dr[n] && all)
      for (int &f : cf)
        if (f == REF_DO) f = REF_KEEP;
  }
  if (max_dx > rp.refine_max_dx) {
    for (int &f : cf) f = REF_DO;
  } else if (min_dx < 2 * rp.refine_min_dx) {
    for (int &f : cf)
      if (f == REF_DO) f = REF_KEEP;
  }
}

// cell_to_ref_flags (afivo/src/m_af_core.f90:1091-1145)
static void cell_to_ref_flags(const Ctx &c, const std::vector<int> &cf, int buffer, int id, std::vector<int> &flags) {
  const int D = c.D, nc = c.nc;
  bool any_do = false, any_keep = false;
  for (int f : cf) {
    any_do = any_do || f == REF_DO;
    any_keep = any_keep || f == REF_KEEP;
  }
  int &fl = flags[id - 1];
  if (any_do)
    fl = REF_DO;
  else if (any_keep)
    fl = std::max(fl, REF_KEEP);
  else
    fl = std::max(fl, REF_RM);
  if (buffer <= 0) return;
  const Box &b = c.B(id);
  const int kmin = (D == 3) ? -1 : 0, kmax = (D == 3) ? 1 : 0;
  for (int dk = kmin; dk <= kmax; dk++)
    for (int dj = -1; dj <= 1; dj++)
      for (int di = -1; di <= 1; di++) {
        if (di == 0 && dj == 0 && dk == 0) continue;
        const int nb_id = (D == 3) ? b.nbmat[nbmat_index(di, dj, dk)] : b.nbmat[nbmat_index2(di, dj)];
        if (nb_id <= 0) continue;
        const int dir[3] = {di, dj, dk};
        int lo[3] = {1, 1, 1}, hi[3] = {nc, nc, (D == 3) ? nc : 1};
        for (int d = 0; d < D; d++) {
          if (dir[d] == 1) lo[d] = nc - buffer + 1;
          if (dir[d] == -1) hi[d] = buffer;
        }
        bool hit = false;
        for (int k = lo[2]; k <= hi[2] && !hit; k++)
          for (int j = lo[1]; j <= hi[1] && !hit; j++)
            for (int i = lo[0]; i <= hi[0] && !hit; i++)
              hit = cf[(i - 1) + nc * ((j - 1) + nc * (k - 1))] == REF_DO;
        if (hit) flags[nb_id - 1] = REF_DO;
      }
}

// consistent_ref_flags (afivo/src/m_af_core.f90:929-1013) with
// ensure_two_one_balance (:1016-1057) and handle_derefinement_flags
// (:1059-1088)
static void refine_flags(const Ctx &c, const afgpu_refine_t &rp, std::vector<int> &flags) {
  const int unset = std::numeric_limits<int>::min();
  flags.assign(c.n_boxes, unset);
  std::vector<int> cf((size_t)c.nc * c.nc * (c.D == 3 ? c.nc : 1));
  const int nch = 1 << c.D;
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++)
    for (int id : c.leaves[lvl]) {
      cell_flags(c, rp, id, cf);
      cell_to_ref_flags(c, cf, rp.buffer_width, id, flags);
      const int p = c.B(id).parent;
      if (c.B(id).lvl > 1 && p > 0) {
        // the parent, once, from its first child that is itself a leaf
        const Box &bp = c.B(p);
        int ich = 0;
        while (ich < nch && bp.children[ich] != id) ich++;
        int i_ch = 0;
        while (i_ch < ich && c.has_children(bp.children[i_ch])) i_ch++;
        if (i_ch == ich) {
          cell_flags(c, rp, p, cf);
          cell_to_ref_flags(c, cf, rp.buffer_width, p, flags);
        }
      }
    }
  for (int &f : flags)
    if (f == unset) f = REF_KEEP;
  if (rp.max_lvl >= 1 && rp.max_lvl <= c.highest_lvl)
    for (int id : c.ids[rp.max_lvl])
      if (flags[id - 1] == REF_DO) flags[id - 1] = REF_KEEP;
  // ensure_two_one_balance
  for (int lvl = c.highest_lvl; lvl >= 1; lvl--)
    for (int id : c.leaves[lvl]) {
      int &f = flags[id - 1];
      const Box &b = c.B(id);
      if (f == REF_DO || f == REF_REFINE) {
        f = REF_REFINE;
        for (int nb = 0; nb < 2 * c.D; nb++)
          if (b.nb[nb] == 0) flags[c.B(b.parent).nb[nb] - 1] = REF_REFINE;
      } else if (f == REF_RM) {
        for (int nb = 0; nb < 2 * c.D; nb++) {
          const int nb_id = b.nb[nb];
          if (nb_id > 0 && (c.has_children(nb_id) || flags[nb_id - 1] > REF_KEEP)) {
            f = REF_KEEP;
            break;
          }
        }
      }
    }
  // handle_derefinement_flags
  for (int lvl = c.highest_lvl - 1; lvl >= 1; lvl--)
    for (int id : c.parents[lvl]) {
      const Box &b = c.B(id);
      bool all_refined = true, all_rm = true;
      for (int q = 0; q < nch; q++) {
        all_refined = all_refined && c.has_children(b.children[q]);
        all_rm = all_rm && flags[b.children[q] - 1] == REF_RM;
      }
      if (all_refined) continue;
      if (all_rm && flags[id - 1] <= REF_KEEP) {
        flags[id - 1] = REF_DEREFINE;
      } else {
        flags[id - 1] = REF_KEEP;
        for (int q = 0; q < nch; q++) {
          int &fc = flags[b.children[q] - 1];
          if (fc != REF_DEREFINE) fc = std::max(fc, REF_KEEP);
        }
      }
    }
}

// ---- data of boxes removed or added by refinement ---------------------------
// The variables with methods, in the order the streamer sets them
// (photoi_initialize, field_initialize, mg_init, src/streamer.f90:84-86 and
// 103-108): photo (with photoionization), |E|, phi, the densities, rhs.
static std::vector<int> auto_vars(const Ctx &c) {
  std::vector<int> v;
  if (c.p.photoi_enabled) v.push_back(AFGPU_CC_PHOTO);
  v.push_back(AFGPU_CC_EFIELD);
  v.push_back(AFGPU_CC_PHI);
  for (int k = 0; k < 3; k++) v.push_back(AFGPU_CC_SPECIES(k));
  v.push_back(AFGPU_CC_RHS);
  return v;
}
// the densities and rhs use af_prolong_limit (the streamer's default
// prolongation, src/m_streamer.f90:402-417); the others af_prolong_linear
static bool prolongs_limited(int iv) { return iv < 9 || iv == AFGPU_CC_RHS; }

// af_prolong_limit (afivo/src/m_af_prolong.f90:311-420), not adding
static void prolong_limit(Ctx &c, int id_c, int iv) {
  const int nc = c.nc, hnc = nc / 2, D = c.D;
  const Box &bc_ = c.B(id_c);
  const double *P = c.C(iv, bc_.parent);
  double *u = c.C(iv, id_c);
  const int lim = methods(c, iv).limiter;
  int off[3] = {0, 0, 0};
  for (int d = 0; d < D; d++) off[d] = ((bc_.ix[d] - 1) & 1) * hnc;
  const int kmax = (D == 3) ? nc : 1;
  for (int k = (D == 3) ? 1 : 0; k <= (D == 3 ? kmax : 0); k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) u[c.idx(i, j, k)] = 0;
  const int st[3] = {1, c.ng, c.ng * c.ng};
  for (int k = 1; k <= (D == 3 ? hnc : 1); k++)
    for (int j = 1; j <= hnc; j++)
      for (int i = 1; i <= hnc; i++) {
        const int ic = off[0] + i, jc = off[1] + j, kc = (D == 3) ? off[2] + k : 0;
        const long x = c.idx(ic, jc, kc);
        const double f0 = P[x];
        double f[3];
        for (int d = 0; d < D; d++) {
          const double a = P[x] - P[x - st[d]], b = P[x + st[d]] - P[x];
          f[d] = 0.25 * limiter_apply(a, b, lim);
        }
        const int i_f = 2 * i - 1, jf = 2 * j - 1, kf = (D == 3) ? 2 * k - 1 : 0;
        for (int q = 0; q < (1 << D); q++) {
          const int di = q & 1, dj = (q >> 1) & 1, dk = (q >> 2) & 1;
          double v = f0;
          v = di ? v + f[0] : v - f[0];
          v = dj ? v + f[1] : v - f[1];
          if (D == 3) v = dk ? v + f[2] : v - f[2];
          double &y = u[c.idx(i_f + di, jf + dj, (D == 3) ? kf + dk : 0)];
          y = v + y;
        }
      }
}

// af_prolong_linear (afivo/src/m_af_prolong.f90:531-679), not adding
static void prolong_linear(Ctx &c, int id_c, int iv) {
  const int nc = c.nc, hnc = nc / 2, D = c.D;
  const Box &bc_ = c.B(id_c);
  const double *P = c.C(iv, bc_.parent);
  double *u = c.C(iv, id_c);
  int off[3] = {0, 0, 0};
  for (int d = 0; d < D; d++) off[d] = ((bc_.ix[d] - 1) & 1) * hnc;
  for (int k = (D == 3) ? 1 : 0; k <= (D == 3 ? nc : 0); k++)
    for (int j = 1; j <= nc; j++)
      for (int i = 1; i <= nc; i++) u[c.idx(i, j, k)] = 0;
  if (D == 2) {
    const double f1 = 1 / 16.0, f3 = 3 / 16.0, f9 = 9 / 16.0;
    for (int j = 1; j <= hnc; j++)
      for (int i = 1; i <= hnc; i++) {
        const int ic = off[0] + i, jc = off[1] + j, i_f = 2 * i - 1, jf = 2 * j - 1;
        auto p = [&](int a, int b) { return P[c.idx(ic + a, jc + b, 0)]; };
        const double f0 = f9 * p(0, 0), flx = f3 * p(-1, 0), fhx = f3 * p(1, 0), fly = f3 * p(0, -1),
                     fhy = f3 * p(0, 1), fll = f1 * p(-1, -1), fhl = f1 * p(1, -1), flh = f1 * p(-1, 1),
                     fhh = f1 * p(1, 1);
        double *y = &u[c.idx(i_f, jf, 0)];
        y[0] = f0 + flx + fly + fll + y[0];
        y[1] = f0 + fhx + fly + fhl + y[1];
        y[c.ng] = f0 + flx + fhy + flh + y[c.ng];
        y[c.ng + 1] = f0 + fhx + fhy + fhh + y[c.ng + 1];
      }
    return;
  }
  const double f1 = 1 / 64.0, f3 = 3 / 64.0, f9 = 9 / 64.0, f27 = 27 / 64.0;
  for (int k = 1; k <= hnc; k++)
    for (int j = 1; j <= hnc; j++)
      for (int i = 1; i <= hnc; i++) {
        const int ic = off[0] + i, jc = off[1] + j, kc = off[2] + k;
        const int i_f = 2 * i - 1, jf = 2 * j - 1, kf = 2 * k - 1;
        auto p = [&](int a, int b, int e) { return P[c.idx(ic + a, jc + b, kc + e)]; };
        const double f000 = f27 * p(0, 0, 0);
        const double f00l = f9 * p(0, 0, -1), f0l0 = f9 * p(0, -1, 0), f0ll = f3 * p(0, -1, -1),
                     fl00 = f9 * p(-1, 0, 0), fl0l = f3 * p(-1, 0, -1), fll0 = f3 * p(-1, -1, 0),
                     flll = f1 * p(-1, -1, -1);
        const double f00h = f9 * p(0, 0, 1), f0h0 = f9 * p(0, 1, 0), f0hh = f3 * p(0, 1, 1),
                     fh00 = f9 * p(1, 0, 0), fh0h = f3 * p(1, 0, 1), fhh0 = f3 * p(1, 1, 0),
                     fhhh = f1 * p(1, 1, 1);
        const double fl0h = f3 * p(-1, 0, 1), fh0l = f3 * p(1, 0, -1), flh0 = f3 * p(-1, 1, 0),
                     fhl0 = f3 * p(1, -1, 0), f0lh = f3 * p(0, -1, 1), f0hl = f3 * p(0, 1, -1);
        const double fllh = f1 * p(-1, -1, 1), flhl = f1 * p(-1, 1, -1), fhll = f1 * p(1, -1, -1);
        const double fhhl = f1 * p(1, 1, -1), fhlh = f1 * p(1, -1, 1), flhh = f1 * p(-1, 1, 1);
        auto Y = [&](int a, int b, int e) -> double & { return u[c.idx(i_f + a, jf + b, kf + e)]; };
        Y(0, 0, 0) = f000 + fl00 + f0l0 + f00l + fll0 + fl0l + f0ll + flll + Y(0, 0, 0);
        Y(1, 0, 0) = f000 + fh00 + f0l0 + f00l + fhl0 + fh0l + f0ll + fhll + Y(1, 0, 0);
        Y(0, 1, 0) = f000 + fl00 + f0h0 + f00l + flh0 + fl0l + f0hl + flhl + Y(0, 1, 0);
        Y(1, 1, 0) = f000 + fh00 + f0h0 + f00l + fhh0 + fh0l + f0hl + fhhl + Y(1, 1, 0);
        Y(0, 0, 1) = f000 + fl00 + f0l0 + f00h + fll0 + fl0h + f0lh + fllh + Y(0, 0, 1);
        Y(1, 0, 1) = f000 + fh00 + f0l0 + f00h + fhl0 + fh0h + f0lh + fhlh + Y(1, 0, 1);
        Y(0, 1, 1) = f000 + fl00 + f0h0 + f00h + flh0 + fl0h + f0hh + flhh + Y(0, 1, 1);
        Y(1, 1, 1) = f000 + fh00 + f0h0 + f00h + fhh0 + fh0h + f0hh + fhhh + Y(1, 1, 1);
      }
}

// auto_restrict (afivo/src/m_af_core.f90:825-840): every child into box id
// Species of a reaction list beyond the three pools: the routines address
// pools by index, so each such species is handled in the electron pool's
// slot (same methods), its storage swapped in and out around the call.
template <typename F>
static void for_extra_species(Ctx &c, F f) {
  for (int k = 3; k < c.n_sp; k++) {
    std::vector<double> &x = c.xs[(size_t)(k - 3) * 3];
    std::swap(c.cc[AFGPU_CC_ELECTRON], x);
    f(AFGPU_CC_ELECTRON);
    std::swap(c.cc[AFGPU_CC_ELECTRON], x);
  }
}

static void restrict_boxes(Ctx &c, const std::vector<int> &ids) {
  const std::vector<int> vars = auto_vars(c);
  const int nch = 1 << c.D;
  auto one = [&](int iv) {
    for (int id : ids)
      for (int q = 0; q < nch; q++) restrict_box(c, c.B(id).children[q], iv, iv, true);
  };
  for (int iv : vars) one(iv);
  for_extra_species(c, one);
}

// auto_prolong (afivo/src/m_af_core.f90:843-881): the new boxes level by
// level, each from its parent, then their ghost cells (af_gc_box, corners)
static void prolong_boxes(Ctx &c, const std::vector<int> &ids) {
  const std::vector<int> vars = auto_vars(c);
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
    std::vector<int> l;
    for (int id : ids)
      if (c.B(id).lvl == lvl) l.push_back(id);
    for (int id : l)
      for (int iv : vars) {
        if (prolongs_limited(iv))
          prolong_limit(c, id, iv);
        else
          prolong_linear(c, id, iv);
      }
    for (int id : l)
      for (int iv : vars) gc_box(c, id, iv, true);
    for_extra_species(c, [&](int iv) {
      for (int id : l) prolong_limit(c, id, iv);
      for (int id : l) gc_box(c, id, iv, true);
    });
  }
}

}  // namespace afcpu

extern "C" {

const char *afcpu_last_error(void) { return g_err.c_str(); }

int afcpu_create(int ndim, int n_cell, int coord, int precision, afcpu_t **out) {
  if ((ndim != 2 && ndim != 3) || n_cell < 4 || n_cell > 32 || (n_cell % 2) || !out ||
      (coord != AFGPU_XYZ && coord != AFGPU_CYL) || (coord == AFGPU_CYL && ndim != 2)) {
    set_error("afcpu_create: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  if (precision != AFGPU_PREC_FP64) {
    set_error("the oracle computes in FP64 only");
    return AFGPU_ERR_UNSUPPORTED;
  }
  afcpu_t *h = new afcpu_t();
  Ctx &c = h->c;
  c.D = ndim;
  c.nc = n_cell;
  c.coord = coord;
  c.ng = n_cell + 2;
  c.ncell = 1;
  c.nface = 1;
  for (int d = 0; d < ndim; d++) {
    c.ncell *= c.ng;
    c.nface *= (n_cell + 1);
  }
  c.cc.resize(AFGPU_N_CC);
  c.fc.resize(AFGPU_N_FC);
  *out = h;
  return AFGPU_OK;
}

int afcpu_destroy(afcpu_t *h) {
  delete h;
  return AFGPU_OK;
}

int afcpu_sync_tree(afcpu_t *h, int n_boxes, const int *lvl, const int *ix, const int *parent,
                    const int *children, const int *neighbors, const int *neighbor_mat,
                    const double *r_min, const double *dr) {
  CHECK_H(h);
  if (n_boxes < 1 || !lvl || !ix || !parent || !children || !neighbors || !neighbor_mat || !r_min || !dr) {
    set_error("afcpu_sync_tree: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  return sync_tree(h->c, n_boxes, lvl, ix, parent, children, neighbors, neighbor_mat, r_min, dr);
}

static int xfer(afcpu_t *h, bool cc, int iv, int n_ids, const int *ids, double *data, bool put) {
  CHECK_H(h);
  Ctx &c = h->c;
  if ((cc && (iv < 0 || iv >= AFGPU_N_CC)) || (!cc && (iv < 0 || iv >= AFGPU_N_FC)) || n_ids < 0 ||
      (n_ids > 0 && (!ids || !data))) {
    set_error("transfer: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  size_t blk = cc ? c.ncell : c.D * c.nface;
  for (int n = 0; n < n_ids; n++) {
    if (ids[n] < 1 || ids[n] > c.n_boxes) {
      set_error("transfer: box id out of range");
      return AFGPU_ERR_ARG;
    }
    double *dst = cc ? c.C(iv, ids[n]) : c.F(iv, ids[n]);
    if (put)
      std::memcpy(dst, data + n * blk, blk * sizeof(double));
    else
      std::memcpy(data + n * blk, dst, blk * sizeof(double));
  }
  return AFGPU_OK;
}

int afcpu_put_cc(afcpu_t *h, int iv, int n, const int *ids, const double *d) {
  return xfer(h, true, iv, n, ids, const_cast<double *>(d), true);
}
int afcpu_get_cc(afcpu_t *h, int iv, int n, const int *ids, double *d) { return xfer(h, true, iv, n, ids, d, false); }
int afcpu_put_fc(afcpu_t *h, int iv, int n, const int *ids, const double *d) {
  return xfer(h, false, iv, n, ids, const_cast<double *>(d), true);
}
int afcpu_get_fc(afcpu_t *h, int iv, int n, const int *ids, double *d) { return xfer(h, false, iv, n, ids, d, false); }

int afcpu_set_params(afcpu_t *h, const afgpu_params_t *p) {
  CHECK_H(h);
  if (!p || p->photoi_n_modes < 0 || p->photoi_n_modes > AFGPU_MAX_MODES || p->photoi_species < 0 ||
      p->photoi_species >= 16) {
    set_error("afcpu_set_params: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  h->c.p = *p;
  h->c.have_params = true;
  h->c.ops[AFGPU_OP_POISSON].lambda2 = 0.0;
  for (int n = 0; n < p->photoi_n_modes; n++)
    h->c.ops[AFGPU_OP_HELMH + n].lambda2 = p->photoi_lambda[n] * p->photoi_lambda[n];
  return AFGPU_OK;
}

int afcpu_set_tables(afcpu_t *h, int n_rows, double td_x_min, double td_inv_fac, const double *td_cols,
                     double chem_x_min, double chem_inv_fac, const double *chem_cols) {
  CHECK_H(h);
  if (n_rows < 2 || !td_cols || !chem_cols) {
    set_error("afcpu_set_tables: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  Ctx &c = h->c;
  c.td.n = n_rows;
  c.td.x_min = td_x_min;
  c.td.inv_fac = td_inv_fac;
  c.td.cols.assign(td_cols, td_cols + 4 * (size_t)n_rows);
  c.chem.n = n_rows;
  c.chem.x_min = chem_x_min;
  c.chem.inv_fac = chem_inv_fac;
  c.chem.cols.assign(chem_cols, chem_cols + 2 * (size_t)n_rows);
  return AFGPU_OK;
}

int afcpu_set_voltage(afcpu_t *h, double v) {
  CHECK_H(h);
  h->c.voltage = v;
  return AFGPU_OK;
}

int afcpu_set_coarse_factors(afcpu_t *h, int op, const double *vinv, const double *v, const double *eig) {
  CHECK_H(h);
  if (op < 0 || op > AFGPU_MAX_MODES || !vinv || !v || !eig) {
    set_error("afcpu_set_coarse_factors: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  Ctx &c = h->c;
  size_t nn = (size_t)c.D * c.nc * c.nc;
  c.ops[op].vinv.assign(vinv, vinv + nn);
  c.ops[op].v.assign(v, v + nn);
  c.ops[op].eig.assign(eig, eig + (size_t)c.D * c.nc);
  c.ops[op].have_factors = true;
  return AFGPU_OK;
}

static int ready(afcpu_t *h, int op) {
  CHECK_H(h);
  if (h->c.highest_lvl < 1 || !h->c.have_params) {
    set_error("tree or parameters not set");
    return AFGPU_ERR_STATE;
  }
  if (op >= 0 && !h->c.ops[op].have_factors) {
    set_error("coarse factors not set for operator");
    return AFGPU_ERR_STATE;
  }
  return AFGPU_OK;
}

// The species receiving the photoionization source exists in the model.
static int check_photoi_species(const Ctx &c) {
  if (c.p.photoi_enabled && c.p.photoi_species >= c.n_sp) {
    set_error("photoionization species not among the species of the model");
    return AFGPU_ERR_ARG;
  }
  return AFGPU_OK;
}

int afcpu_mg_fas_vcycle(afcpu_t *h, int op, int set_residual, int max_lvl) {
  int rc = ready(h, op);
  if (rc) return rc;
  Ctx &c = h->c;
  if (max_lvl <= 0 || max_lvl > c.highest_lvl) max_lvl = c.highest_lvl;
  vcycle(c, mg_of(c, op), set_residual != 0, max_lvl);
  return AFGPU_OK;
}

int afcpu_mg_fas_fmg(afcpu_t *h, int op, int set_residual, int have_guess) {
  int rc = ready(h, op);
  if (rc) return rc;
  fmg(h->c, mg_of(h->c, op), set_residual != 0, have_guess != 0);
  return AFGPU_OK;
}

int afcpu_tree_maxabs_cc(afcpu_t *h, int iv, double *out) {
  int rc = ready(h, -1);
  if (rc) return rc;
  if (iv < 0 || iv >= AFGPU_N_CC || !out) return AFGPU_ERR_ARG;
  *out = tree_maxabs(h->c, iv);
  return AFGPU_OK;
}

}  // extern "C"

// Species k of a reaction list (iv = AFGPU_VAR_SPECIES(k)): the pool index
// of species 0-2, or -1 - (its pool in xs) beyond them; n_sp when none.
static int species_var(const Ctx &c, int iv) {
  const int k = iv - AFGPU_VAR_SPECIES(0);
  if (k < 0 || k >= c.n_sp) return AFGPU_N_CC;
  return k < 3 ? AFGPU_CC_SPECIES(k) : -1 - (k - 3) * 3;
}
// Runs f(iv') with the variable of iv at a density index: a species beyond
// the three pools is swapped into the electrons' third state for the call
// (the densities' boundary methods) and back.
template <typename F>
static int with_density_var(Ctx &c, int iv, F f) {
  const int sv = species_var(c, iv);
  if (sv == AFGPU_N_CC) return f(iv);
  if (sv >= 0) return f(sv);
  const int spare = AFGPU_CC_ELECTRON + 2;
  std::swap(c.cc[spare], c.xs[(size_t)(-1 - sv)]);
  const int rc = f(spare);
  std::swap(c.cc[spare], c.xs[(size_t)(-1 - sv)]);
  return rc;
}

static int restrict_tree_var(Ctx &c, int iv) {
  if (iv < 0 || iv >= AFGPU_N_CC) return AFGPU_ERR_ARG;
  for (int lvl = c.highest_lvl - 1; lvl >= 1; lvl--) {
    const std::vector<int> &par = c.parents[lvl];
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < par.size(); i++) {
      const Box &bx = c.B(par[i]);
      for (int q = 0; q < (1 << c.D); q++)
        if (bx.children[q] > 0) restrict_box(c, bx.children[q], iv, iv, true);
    }
  }
  return AFGPU_OK;
}

extern "C" {

int afcpu_gc_tree(afcpu_t *h, int iv) {
  int rc = ready(h, -1);
  if (rc) return rc;
  return with_density_var(h->c, iv, [&](int v) {
    if (v < 0 || v >= AFGPU_N_CC || !methods(h->c, v).ok) return (int)AFGPU_ERR_ARG;
    gc_tree(h->c, v);
    return (int)AFGPU_OK;
  });
}

int afcpu_restrict_tree(afcpu_t *h, int iv) {
  int rc = ready(h, -1);
  if (rc) return rc;
  return with_density_var(h->c, iv, [&](int v) { return restrict_tree_var(h->c, v); });
}


int afcpu_field_compute(afcpu_t *h, int s_in, int have_guess, int *n_cycles, double *residual) {
  int rc = ready(h, AFGPU_OP_POISSON);
  if (rc) return rc;
  return field_compute(h->c, s_in, have_guess != 0, n_cycles, residual);
}

int afcpu_field_from_potential(afcpu_t *h) {
  int rc = ready(h, -1);
  if (rc) return rc;
  field_from_potential(h->c);
  return AFGPU_OK;
}

int afcpu_forward_euler(afcpu_t *h, double dt, int s_deriv, int n_prev, const int *s_prev, const double *w_prev,
                        int s_out, int i_step, int n_steps, double *dt_lim) {
  int rc = ready(h, -1);
  if (rc) return rc;
  if (!dt_lim || n_prev < 1 || !s_prev || !w_prev) return AFGPU_ERR_ARG;
  if ((rc = check_photoi_species(h->c))) return rc;
  return forward_euler(h->c, dt, s_deriv, n_prev, s_prev, w_prev, s_out, i_step, n_steps, dt_lim);
}

int afcpu_advance(afcpu_t *h, double dt, double *dt_lim) {
  int rc = ready(h, -1);
  if (rc) return rc;
  if (!dt_lim) return AFGPU_ERR_ARG;
  if ((rc = check_photoi_species(h->c))) return rc;
  return advance(h->c, dt, dt_lim);
}

int afcpu_photoi_set_src(afcpu_t *h) {
  int rc = ready(h, -1);
  if (rc) return rc;
  for (int n = 0; n < h->c.p.photoi_n_modes; n++)
    if (!h->c.ops[AFGPU_OP_HELMH + n].have_factors) {
      set_error("coarse factors not set for a Helmholtz mode");
      return AFGPU_ERR_STATE;
    }
  return photoi_set_src(h->c);
}

int afcpu_refine_flags(afcpu_t *h, const afgpu_refine_t *rp, int n_boxes, int *ref_flags) {
  int rc = ready(h, -1);
  if (rc) return rc;
  if (!rp || !ref_flags || n_boxes != h->c.n_boxes || rp->buffer_width < 0 || rp->buffer_width > h->c.nc ||
      rp->n_seeds < 0 || rp->n_seeds > AFGPU_REF_MAX_ITEMS || rp->n_regions < 0 ||
      rp->n_regions > AFGPU_REF_MAX_ITEMS || rp->n_limits < 0 || rp->n_limits > AFGPU_REF_MAX_ITEMS) {
    set_error("afcpu_refine_flags: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  if (h->c.td.n < 2) {
    set_error("afcpu_refine_flags: tables not set");
    return AFGPU_ERR_STATE;
  }
  std::vector<int> f;
  refine_flags(h->c, *rp, f);
  std::copy(f.begin(), f.end(), ref_flags);
  return AFGPU_OK;
}

static int check_ids(afcpu_t *h, int n, const int *ids, bool parents) {
  if (n < 0 || (n > 0 && !ids)) return AFGPU_ERR_ARG;
  for (int q = 0; q < n; q++) {
    const int id = ids[q];
    if (id < 1 || id > h->c.n_boxes || h->c.B(id).lvl < 1) return AFGPU_ERR_ARG;
    if (parents ? !h->c.has_children(id) : h->c.B(id).parent < 1) return AFGPU_ERR_ARG;
  }
  return AFGPU_OK;
}
int afcpu_restrict_boxes(afcpu_t *h, int n_ids, const int *ids) {
  int rc = ready(h, -1);
  if (rc) return rc;
  if (check_ids(h, n_ids, ids, true)) {
    set_error("afcpu_restrict_boxes: invalid ids");
    return AFGPU_ERR_ARG;
  }
  restrict_boxes(h->c, std::vector<int>(ids, ids + n_ids));
  return AFGPU_OK;
}
int afcpu_prolong_boxes(afcpu_t *h, int n_ids, const int *ids) {
  int rc = ready(h, -1);
  if (rc) return rc;
  if (check_ids(h, n_ids, ids, false)) {
    set_error("afcpu_prolong_boxes: invalid ids");
    return AFGPU_ERR_ARG;
  }
  prolong_boxes(h->c, std::vector<int>(ids, ids + n_ids));
  return AFGPU_OK;
}

// output_regression_log's statistics (src/m_output.f90:797-810) as the
// reference computes them with one thread: af_tree_sum_cc
// (afivo/src/m_af_utils.f90:966-1026: levels in order, their leaves in
// order, each box's interior summed with i fastest, sum_2pr_box in 2D
// cylindrical coordinates), af_tree_max_cc (:758-769, box_max_cc and
// reduce_max) and af_total_volume (m_af_types.f90:805-825).
int afcpu_regression_values(afcpu_t *h, int n_vars, const int *ivs, double *out) {
  int rc = ready(h, -1);
  if (rc) return rc;
  if (n_vars < 1 || n_vars > AFGPU_DIAG_MAX_VARS || !ivs || !out) return AFGPU_ERR_ARG;
  Ctx &c = h->c;
  for (int v = 0; v < n_vars; v++) {
    const int k = ivs[v] - AFGPU_VAR_SPECIES(0);
    if (!(k >= 0 && k < c.n_sp) && (ivs[v] < 0 || ivs[v] >= AFGPU_N_CC)) return AFGPU_ERR_ARG;
  }
  auto var = [&](int iv, int id) {
    const int k = iv - AFGPU_VAR_SPECIES(0);
    return (k >= 0 && k < c.n_sp) ? c.SP(k, 0, id) : c.C(iv, id);
  };
  const int D = c.D, nc = c.nc;
  const int kmin = (D == 3) ? 1 : 0, kmax = (D == 3) ? nc : 0;
  const double pi = std::acos(-1.0), two_pi = 2 * std::acos(-1.0);
  const Box &b1 = c.B(c.ids[1][0]);
  double box_len[3];
  for (int d = 0; d < D; d++) box_len[d] = nc * b1.dr[d];
  double vol;
  if (D == 2 && c.coord == AFGPU_CYL) {
    vol = 0.0;
    const double r0 = b1.r_min[0], r1 = r0 + box_len[0];
    vol = vol + pi * (r1 * r1 - r0 * r0) * box_len[D - 1];
  } else {
    double p = 1;
    for (int d = 0; d < D; d++) p = p * box_len[d];
    vol = (double)c.ids[1].size() * p;
  }
  for (int v = 0; v < n_vars; v++) {
    const int iv = ivs[v];
    for (int pow = 1; pow <= 2; pow++) {
      double my_sum = 0;
      for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
        if (c.ids[lvl].empty()) continue;
        const Box &bl = c.B(c.ids[lvl][0]);
        double fac = 1;
        for (int d = 0; d < D; d++) fac = fac * bl.dr[d];
        for (int id : c.leaves[lvl]) {
          const double *u = var(iv, id);
          const Box &bx = c.B(id);
          double tmp = 0;
          for (int k = kmin; k <= kmax; k++)
            for (int j = 1; j <= nc; j++)
              for (int i = 1; i <= nc; i++) {
                const double x = u[c.idx(i, j, k)];
                const double xp = (pow == 1) ? x : x * x;
                if (D == 2 && c.coord == AFGPU_CYL)
                  tmp = tmp + xp * (bx.r_min[0] + (i - 0.5) * bx.dr[0]);
                else
                  tmp = tmp + xp;
              }
          if (D == 2 && c.coord == AFGPU_CYL) tmp = tmp * two_pi;
          my_sum = my_sum + fac * tmp;
        }
      }
      out[3 * v + pow - 1] = my_sum / vol;
    }
    double mx = -std::numeric_limits<double>::max() / 10;
    for (int lvl = 1; lvl <= c.highest_lvl; lvl++)
      for (int id : c.leaves[lvl]) {
        const double *u = var(iv, id);
        double bm = -std::numeric_limits<double>::max();
        for (int k = kmin; k <= kmax; k++)
          for (int j = 1; j <= nc; j++)
            for (int i = 1; i <= nc; i++) bm = std::max(bm, u[c.idx(i, j, k)]);
        mx = std::max(bm, mx);
      }
    out[3 * v + 2] = mx;
  }
  return AFGPU_OK;
}

// Box data in af_write_tree's order (afivo/src/m_af_output.f90:90-114):
// per box, the variables in turn, box%cc(DTIMES(:), iv) for a cell variable
// or a species, box%fc(DTIMES(:), :, ifc) for AFGPU_VAR_FACE(ifc)
static int box_io(afcpu_t *h, int n_vars, const int *ivs, int n_ids, const int *ids, double *data, bool put) {
  CHECK_H(h);
  Ctx &c = h->c;
  if (n_vars < 1 || n_vars > AFGPU_DIAG_MAX_VARS || !ivs || n_ids < 0 || (n_ids > 0 && (!ids || !data)))
    return AFGPU_ERR_ARG;
  for (int q = 0; q < n_ids; q++)
    if (ids[q] < 1 || ids[q] > c.n_boxes) return AFGPU_ERR_ARG;
  // each variable's box data: pool, words per box
  std::vector<std::vector<double> *> pool(n_vars);
  std::vector<size_t> len(n_vars);
  size_t total = 0;
  for (int v = 0; v < n_vars; v++) {
    const int f = ivs[v] - AFGPU_VAR_FACE(0);
    const int sv = species_var(c, ivs[v]);
    if (f >= 0 && f < AFGPU_N_FC) {
      pool[v] = &c.fc[f];
      len[v] = (size_t)c.D * c.nface;
    } else if (sv != AFGPU_N_CC) {
      pool[v] = sv >= 0 ? &c.cc[sv] : &c.xs[(size_t)(-1 - sv)];
      len[v] = c.ncell;
    } else if (ivs[v] >= 0 && ivs[v] < AFGPU_N_CC) {
      pool[v] = &c.cc[ivs[v]];
      len[v] = c.ncell;
    } else {
      return AFGPU_ERR_ARG;
    }
    total += len[v];
  }
  for (int q = 0; q < n_ids; q++) {
    double *b = data + (size_t)q * total;
    for (int v = 0; v < n_vars; v++) {
      double *u = pool[v]->data() + (size_t)(ids[q] - 1) * len[v];
      if (put)
        std::memcpy(u, b, len[v] * sizeof(double));
      else
        std::memcpy(b, u, len[v] * sizeof(double));
      b += len[v];
    }
  }
  return AFGPU_OK;
}
int afcpu_get_boxes(afcpu_t *h, int n_vars, const int *ivs, int n_ids, const int *ids, double *data) {
  return box_io(h, n_vars, ivs, n_ids, ids, data, false);
}
int afcpu_put_boxes(afcpu_t *h, int n_vars, const int *ivs, int n_ids, const int *ids, const double *data) {
  return box_io(h, n_vars, ivs, n_ids, ids, const_cast<double *>(data), true);
}

int afcpu_copy_current_state(afcpu_t *h) {
  int rc = ready(h, -1);
  if (rc) return rc;
  copy_states(h->c, true);
  return AFGPU_OK;
}

int afcpu_restore_previous_state(afcpu_t *h) {
  int rc = ready(h, -1);
  if (rc) return rc;
  copy_states(h->c, false);
  return AFGPU_OK;
}

int afcpu_step(afcpu_t *h, afgpu_step_info_t *info) {
  int rc = ready(h, AFGPU_OP_POISSON);
  if (rc) return rc;
  if ((rc = check_photoi_species(h->c))) return rc;
  return step(h->c, info);
}

int afcpu_get_time_state(afcpu_t *h, int *it, double *time, double *global_dt) {
  CHECK_H(h);
  Ctx &c = h->c;
  if (it) *it = c.it;
  if (time) *time = c.time;
  if (global_dt) *global_dt = c.started ? c.global_dt : c.p.dt_min;
  return AFGPU_OK;
}

int afcpu_set_time_state(afcpu_t *h, int it, double time, double global_dt) {
  CHECK_H(h);
  if (it < 0 || !(global_dt > 0)) return AFGPU_ERR_ARG;
  Ctx &c = h->c;
  c.it = it;
  c.time = c.global_time = time;
  c.global_dt = c.dt = global_dt;
  c.started = true;
  return AFGPU_OK;
}

int afcpu_limit_step_dt(afcpu_t *h, double dt_cap) {  // src/streamer.f90:222-229
  CHECK_H(h);
  Ctx &c = h->c;
  if (!(dt_cap >= 0) || !c.have_params) return AFGPU_ERR_ARG;
  if (!c.started) {
    c.global_dt = c.p.dt_min;
    c.dt = c.global_dt;
    c.started = true;
  }
  c.dt = std::min(c.dt, dt_cap);
  return AFGPU_OK;
}

long long afcpu_num_leaf_cells(afcpu_t *h) {
  if (!h) return 0;
  long long n = 0;
  for (int lvl = 1; lvl <= h->c.highest_lvl; lvl++) n += (long long)h->c.leaves[lvl].size();
  for (int d = 0; d < h->c.D; d++) n *= h->c.nc;
  return n;
}

int afcpu_num_threads(void) { return omp_get_max_threads(); }

// ---- general chemistry networks ---------------------------------------------
// get_rates (src/m_chemistry.f90:672-768) and get_derivatives (:770-804):
// reaction by reaction over all cells, as the reference's array statements.

int afcpu_chem_create(int n_species, int n_reactions, const afgpu_reaction_t *reactions, int table_n,
                      double table_x_min, double table_x_max, int n_fld_cols, const double *fld_cols,
                      const double *energy_col, double gas_temperature, afcpu_chem **out) {
  if (!out || n_species < 1 || n_species > AFGPU_CHEM_MAX_SPECIES || n_reactions < 0 ||
      n_reactions > AFGPU_CHEM_MAX_REACTIONS || (n_reactions > 0 && !reactions) || table_n < 2 ||
      !(table_x_max > table_x_min) || n_fld_cols < 0 || (n_fld_cols > 0 && !fld_cols))
    return AFGPU_ERR_ARG;
  afcpu_chem *c = new afcpu_chem;
  c->n_species = n_species;
  c->n_reactions = n_reactions;
  c->r.assign(reactions, reactions + n_reactions);
  const double inv_fac = (table_n - 1) / (table_x_max - table_x_min);  // LT_create, linear spacing
  c->fld.n = c->energy.n = table_n;
  c->fld.x_min = c->energy.x_min = table_x_min;
  c->fld.inv_fac = c->energy.inv_fac = inv_fac;
  c->fld.cols.assign(fld_cols, fld_cols + (size_t)n_fld_cols * table_n);
  if (energy_col) c->energy.cols.assign(energy_col, energy_col + table_n);
  for (const auto &R : c->r) {
    const int t = R.rate_type;
    c->need_te = c->need_te || t == AFGPU_RATE_K1 || t == AFGPU_RATE_K3 || t == AFGPU_RATE_K16;
    if (t == AFGPU_RATE_TABULATED_ENERGY || t == 7 || t < 0 || t > AFGPU_RATE_K16) {
      delete c;
      return AFGPU_ERR_UNSUPPORTED;
    }
    bool ok = R.n_in >= 1 && R.n_in <= AFGPU_CHEM_MAX_IN && R.n_out >= 0 && R.n_out <= AFGPU_CHEM_MAX_OUT &&
              (t != AFGPU_RATE_TABULATED_FIELD || (R.lookup_col >= 0 && R.lookup_col < n_fld_cols));
    for (int q = 0; ok && q < R.n_in; q++) ok = R.ix_in[q] >= 0 && R.ix_in[q] < n_species;
    for (int q = 0; ok && q < R.n_out; q++) ok = R.ix_out[q] >= 0 && R.ix_out[q] < n_species;
    if (!ok) {
      delete c;
      return AFGPU_ERR_ARG;
    }
  }
  if (c->need_te && !energy_col) {
    delete c;
    return AFGPU_ERR_ARG;
  }
  c->tg = gas_temperature;
  *out = c;
  return AFGPU_OK;
}

int afcpu_chem_destroy(afcpu_chem *c) {
  delete c;
  return AFGPU_OK;
}

int afcpu_chem_rates(afcpu_chem *c, int n_cells, const double *fields, double *rates) {
  if (!c || n_cells < 0) return AFGPU_ERR_ARG;
  const size_t n = (size_t)n_cells;
  // (reaction by reaction over all cells, as the reference's array statements)
  for (int r = 0; r < c->n_reactions; r++)
    for (size_t i = 0; i < n; i++)
      rates[(size_t)r * n + i] = chem_rate_cell(*c, c->r[r], fields[i], chem_te(*c, fields[i]));
  return AFGPU_OK;
}

int afcpu_chem_derivatives(afcpu_chem *c, int n_cells, const double *dens, double *rates, double *derivs) {
  if (!c || n_cells < 0) return AFGPU_ERR_ARG;
  const size_t n = (size_t)n_cells;
  for (size_t x = 0; x < (size_t)c->n_species * n; x++) derivs[x] = 0.0;
  for (int r = 0; r < c->n_reactions; r++) {
    const afgpu_reaction_t &R = c->r[r];
    double *k = rates + (size_t)r * n;
    for (size_t i = 0; i < n; i++) {
      double prod = 1.0;
      for (int q = 0; q < R.n_in; q++) prod = prod * dens[(size_t)R.ix_in[q] * n + i];
      k[i] = k[i] * prod;
    }
    for (int q = 0; q < R.n_in; q++) {
      double *d = derivs + (size_t)R.ix_in[q] * n;
      for (size_t i = 0; i < n; i++) d[i] = d[i] - k[i];
    }
    for (int q = 0; q < R.n_out; q++) {
      double *d = derivs + (size_t)R.ix_out[q] * n;
      for (size_t i = 0; i < n; i++) d[i] = d[i] + k[i] * R.multiplicity_out[q];
    }
  }
  return AFGPU_OK;
}

int afcpu_chem_source(afcpu_chem *c, int n_cells, const double *fields, const double *dens, double *derivs) {
  if (!c || n_cells < 0) return AFGPU_ERR_ARG;
  std::vector<double> rates((size_t)c->n_reactions * n_cells);
  int rc = afcpu_chem_rates(c, n_cells, fields, rates.data());
  if (rc) return rc;
  return afcpu_chem_derivatives(c, n_cells, dens, rates.data(), derivs);
}

}  // extern "C"

extern "C" int afcpu_set_chemistry(afcpu_t *h, const afcpu_chem *chem, const int *charges) {
  CHECK_H(h);
  Ctx &c = h->c;
  const int n = chem ? chem->n_species : 3;
  if (chem && (n < 3 || n > 16 || (n > 3 && !charges) || (charges && charges[0] != -1))) {
    set_error("afcpu_set_chemistry: 3 to 16 species, the electrons first, charges for more than three");
    return AFGPU_ERR_ARG;
  }
  std::vector<int> q{-1, 1, -1};
  if (chem && charges) q.assign(charges, charges + n);
  const bool gen = chem && (n > 3 || q != std::vector<int>{-1, 1, -1});
  if (gen && c.p.photoi_enabled && (c.p.photoi_species < 0 || c.p.photoi_species >= n)) return AFGPU_ERR_ARG;
  const size_t want = gen ? (size_t)(n - 3) * 3 : 0;
  if (c.xs.size() != want) c.xs.assign(want, std::vector<double>((size_t)c.n_boxes * c.ncell, 0.0));
  c.net = chem;
  c.n_sp = gen ? n : 3;
  c.gen_sp = gen;
  c.sp_charge = gen ? q : std::vector<int>{-1, 1, -1};
  return AFGPU_OK;
}

static int species_xfer(afcpu_t *h, int k, int state, int n_ids, const int *ids, double *data, bool put) {
  CHECK_H(h);
  Ctx &c = h->c;
  if (k < 0 || k >= c.n_sp || state < 0 || state > 2 || n_ids < 0 || (n_ids > 0 && (!ids || !data)))
    return AFGPU_ERR_ARG;
  for (int q = 0; q < n_ids; q++)
    if (ids[q] < 1 || ids[q] > c.n_boxes) return AFGPU_ERR_ARG;
  for (int q = 0; q < n_ids; q++) {
    double *u = c.SP(k, state, ids[q]);
    double *b = data + (size_t)q * c.ncell;
    if (put)
      std::copy(b, b + c.ncell, u);
    else
      std::copy(u, u + c.ncell, b);
  }
  return AFGPU_OK;
}
extern "C" int afcpu_get_species(afcpu_t *h, int k, int state, int n_ids, const int *ids, double *data) {
  return species_xfer(h, k, state, n_ids, ids, data, false);
}
extern "C" int afcpu_put_species(afcpu_t *h, int k, int state, int n_ids, const int *ids, const double *data) {
  return species_xfer(h, k, state, n_ids, ids, const_cast<double *>(data), true);
}

// ---- single routines, for the pins against the reference's own routines -
// (the fixture drivers under oracle/ref call the same afivo routines, built
// from the reference's sources, on the same inputs)
namespace afcpu {
static int test_ids(const Ctx &c, int n, const int *ids) {
  if (n < 0 || (n > 0 && !ids)) return AFGPU_ERR_ARG;
  for (int q = 0; q < n; q++)
    if (ids[q] < 1 || ids[q] > c.n_boxes || c.B(ids[q]).lvl < 1) return AFGPU_ERR_ARG;
  return AFGPU_OK;
}
}  // namespace afcpu

extern "C" {

// kind 0: af_limiter_koren, 1: af_limiter_mc, 2: af_limiter_gminmod43
// (afivo/src/m_af_limiters.f90:72-151)
int afcpu_test_limiter(int kind, int n, const double *a, const double *b, double *out) {
  if (kind < 0 || kind > 2 || n < 0 || (n > 0 && (!a || !b || !out))) return AFGPU_ERR_ARG;
  for (int q = 0; q < n; q++)
    out[q] = kind == 0 ? limiter_koren(a[q], b[q]) : limiter_apply(a[q], b[q], kind == 1 ? LIM_MC : LIM_GMINMOD43);
  return AFGPU_OK;
}

// LT_get_loc / LT_get_col (src/lookup_table_fortran/m_lookup_table.f90:330-406)
// on one column of n_rows values; low_ix is the reference's 1-based row
int afcpu_test_table(int n_rows, double x_min, double inv_fac, const double *col, int n, const double *x,
                     double *y, int *low_ix, double *low_frac) {
  if (n_rows < 2 || !col || n < 0 || (n > 0 && (!x || !y))) return AFGPU_ERR_ARG;
  Table t;
  t.n = n_rows;
  t.x_min = x_min;
  t.inv_fac = inv_fac;
  t.cols.assign(col, col + n_rows);
  for (int q = 0; q < n; q++) {
    int ix;
    double w;
    t.loc(x[q], ix, w);
    y[q] = t.at(0, ix, w);
    if (low_ix) low_ix[q] = ix;
    if (low_frac) low_frac[q] = w;
  }
  return AFGPU_OK;
}

// af_stencil_gsrb_box with the multigrid's constant Laplacian stencil
// (redblack 1 or 2), for each listed box in turn
int afcpu_test_gsrb_boxes(afcpu_t *h, int n, const int *ids, int redblack, int iv, int i_rhs, double lambda2) {
  CHECK_H(h);
  Ctx &c = h->c;
  if (test_ids(c, n, ids) || iv < 0 || iv >= AFGPU_N_CC || i_rhs < 0 || i_rhs >= AFGPU_N_CC) return AFGPU_ERR_ARG;
  for (int q = 0; q < n; q++) gsrb_box(c, ids[q], redblack, iv, i_rhs, lambda2);
  return AFGPU_OK;
}

// af_stencil_apply_box with the multigrid's constant Laplacian stencil
int afcpu_test_apply_boxes(afcpu_t *h, int n, const int *ids, int iv, int i_out, double lambda2) {
  CHECK_H(h);
  Ctx &c = h->c;
  if (test_ids(c, n, ids) || iv < 0 || iv >= AFGPU_N_CC || i_out < 0 || i_out >= AFGPU_N_CC) return AFGPU_ERR_ARG;
  for (int q = 0; q < n; q++) apply_box(c, ids[q], iv, i_out, lambda2);
  return AFGPU_OK;
}

// mode 0: stencil_prolong_248 adding iv of the parent to iv_to of the box;
// 1: af_prolong_limit of iv (the variable's limiter); 2: af_prolong_linear
// of iv; for each listed box (with a parent) in turn
int afcpu_test_prolong_boxes(afcpu_t *h, int n, const int *ids, int mode, int iv, int iv_to) {
  CHECK_H(h);
  Ctx &c = h->c;
  if (test_ids(c, n, ids) || mode < 0 || mode > 2 || iv < 0 || iv >= AFGPU_N_CC || iv_to < 0 || iv_to >= AFGPU_N_CC)
    return AFGPU_ERR_ARG;
  for (int q = 0; q < n; q++)
    if (c.B(ids[q]).parent < 1) return AFGPU_ERR_ARG;
  for (int q = 0; q < n; q++) {
    const int id = ids[q];
    if (mode == 0)
      prolong_add(c, c.B(id).parent, id, iv, iv_to);
    else if (mode == 1)
      prolong_limit(c, id, iv);
    else
      prolong_linear(c, id, iv);
  }
  return AFGPU_OK;
}

// af_gc2_box of iv for each listed box in turn: its (n_cell+4)^D halo into
// out (box after box); each call writes its first layer back into the box
int afcpu_test_gc2_boxes(afcpu_t *h, int n, const int *ids, int iv, double *out) {
  CHECK_H(h);
  Ctx &c = h->c;
  if (test_ids(c, n, ids) || iv < 0 || iv >= AFGPU_N_CC || !methods(c, iv).ok || (n > 0 && !out))
    return AFGPU_ERR_ARG;
  const size_t ng4 = (size_t)c.nc + 4, m = c.D == 3 ? ng4 * ng4 * ng4 : ng4 * ng4;
  std::vector<double> g;
  for (int q = 0; q < n; q++) {
    g.assign(m, 0.0);
    gc2_box(c, ids[q], iv, g);
    std::copy(g.begin(), g.end(), out + q * m);
  }
  return AFGPU_OK;
}

// af_consistent_fluxes of the electron flux
int afcpu_test_consistent_fluxes(afcpu_t *h) {
  CHECK_H(h);
  consistent_fluxes(h->c);
  return AFGPU_OK;
}

// flux_upwind_tree of the electrons (state s_deriv) with dt_lim[0] the CFL
// limit and dt_lim[1] the dielectric relaxation time
int afcpu_test_flux_upwind_tree(afcpu_t *h, int s_deriv, double *dt_lim) {
  int rc = ready(h, -1);
  if (rc) return rc;
  if (!dt_lim || s_deriv < 0 || s_deriv > 2 || h->c.td.n < 2) return AFGPU_ERR_ARG;
  flux_upwind_tree(h->c, s_deriv, dt_lim);
  return AFGPU_OK;
}

// The flux divergence of flux_update_densities on the electrons (state 0,
// in place: one previous state of weight 1, no sources), every leaf
int afcpu_test_flux_divergence(afcpu_t *h, double dt) {
  int rc = ready(h, -1);
  if (rc) return rc;
  Ctx &c = h->c;
  for (int lvl = 1; lvl <= c.highest_lvl; lvl++)
    for (int id : c.leaves[lvl]) {
      double dt_dr[3];
      for (int d = 0; d < c.D; d++) dt_dr[d] = dt / c.B(id).dr[d];
      electron_flux_divergence(c, id, dt_dr, 0);
    }
  return AFGPU_OK;
}

}  // extern "C"
