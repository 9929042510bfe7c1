// Multigrid and ghost-cell kernels (afivo/src/m_af_multigrid.f90,
// m_af_stencil.f90, m_af_ghostcell.f90, m_af_restrict.f90).
//
// Work decomposition: a launch covers a box list; each box is handled by
// `bpb` workgroups ("blocks per box", > 1 on the coarse levels that have few
// boxes, so that those launches still spread over the CUs).  Every cell
// value is computed with the reference's operation order, so the FP64
// instantiation reproduces the CPU path bit for bit (hipcc is run with
// -ffp-contract=off).  The same templates instantiated with T = float form
// the inner cycle of the mixed-precision solver.
#pragma once

#include "afgpu_common.h"

namespace afg {

// Box of work item w (box w / bpb, part w % bpb) and the first cell index /
// stride of this thread.
struct Part {
  int id;
  int q0;
  int stride;
};
template <int BS>
__device__ inline Part part_of(const int *__restrict__ ids, int bpb, int w) {
  Part p;
  const int slot = w / bpb;
  p.id = ids[slot];
  p.q0 = (w - slot * bpb) * BS + ltid<BS>();
  p.stride = bpb * BS;
  return p;
}

// Constant 3/5/7-point stencil of a box (mg_box_lpl_stencil,
// afivo/src/m_af_multigrid.f90:1246-1264).
template <int D, typename T>
struct Sten {
  T c[7];
};
template <int D, typename T>
__device__ inline Sten<D, T> lpl_stencil(const DBox &bx, T lambda2) {
  Sten<D, T> s;
  for (int d = 0; d < D; d++) {
    T dr = (T)bx.dr[d];
    T inv_dr2 = T(1) / (dr * dr);
    s.c[1 + 2 * d] = inv_dr2;
    s.c[2 + 2 * d] = inv_dr2;
  }
  T sum = T(0);
  for (int q = 1; q <= 2 * D; q++) sum = sum + s.c[q];
  s.c[0] = -sum - lambda2;
  return s;
}

// Per-i cylindrical coefficients (af_cyl_flux_factors m_af_types.f90:1199
// and the cc_cyl construction of m_af_stencil.f90:388-394, 871-877).
template <typename T>
__device__ inline void cyl_coeffs(const DBox &bx, int i, const T *c, T &c1, T &c2, T &c3) {
  T dr = (T)bx.dr[0];
  T r = (T)bx.r_min[0] + (T(i) - T(0.5)) * dr;
  T inv_r = T(1) / r;
  T rf1 = (r - T(0.5) * dr) * inv_r;
  T rf2 = (r + T(0.5) * dr) * inv_r;
  c2 = rf1 * c[1];
  c3 = rf2 * c[2];
  c1 = c[0] - (c2 - c[1]) - (c3 - c[2]);
}

// L phi at interior cell (i,j,k) (stencil_apply_357, m_af_stencil.f90:367-496)
// of a box stored with layout L (afivo's by default: LDS copies, FP64 pools).
template <int D, int NC, typename T, typename L = LayStd<D, NC>>
__device__ inline T apply_cell(const T *u, const Sten<D, T> &s, bool cyl, const DBox &bx, int i, int j, int k) {
  using G = Geo<D, NC>;
  if constexpr (L::SPLIT) {
    const T uc = u[L::in_at(i, j, k)], xm = u[L::nb_at(i, j, k, 0, -1)], xp = u[L::nb_at(i, j, k, 0, 1)];
    const T ym = u[L::nb_at(i, j, k, 1, -1)], yp = u[L::nb_at(i, j, k, 1, 1)];
    if (D == 2) {
      if (cyl) {
        T c1, c2, c3;
        cyl_coeffs<T>(bx, i, s.c, c1, c2, c3);
        return c1 * uc + c2 * xm + c3 * xp + s.c[3] * ym + s.c[4] * yp;
      }
      return s.c[0] * uc + s.c[1] * xm + s.c[2] * xp + s.c[3] * ym + s.c[4] * yp;
    }
    return s.c[0] * uc + s.c[1] * xm + s.c[2] * xp + s.c[3] * ym + s.c[4] * yp + s.c[5] * u[L::nb_at(i, j, k, 2, -1)] +
           s.c[6] * u[L::nb_at(i, j, k, 2, 1)];
  }
  const int x = G::idx(i, j, k);
  if (D == 2) {
    if (cyl) {
      T c1, c2, c3;
      cyl_coeffs<T>(bx, i, s.c, c1, c2, c3);
      return c1 * u[x] + c2 * u[x - 1] + c3 * u[x + 1] + s.c[3] * u[x - G::NG] + s.c[4] * u[x + G::NG];
    }
    return s.c[0] * u[x] + s.c[1] * u[x - 1] + s.c[2] * u[x + 1] + s.c[3] * u[x - G::NG] +
           s.c[4] * u[x + G::NG];
  } else {
    constexpr int NG2 = G::NG * G::NG;
    return s.c[0] * u[x] + s.c[1] * u[x - 1] + s.c[2] * u[x + 1] + s.c[3] * u[x - G::NG] +
           s.c[4] * u[x + G::NG] + s.c[5] * u[x - NG2] + s.c[6] * u[x + NG2];
  }
}

template <int D, int NC>
__device__ inline void interior_ijk(int q, int &i, int &j, int &k) {
  i = q % NC + 1;
  j = (q / NC) % NC + 1;
  k = (D == 3) ? q / (NC * NC) + 1 : 0;
}

// ---- ghost cells --------------------------------------------------------------
// Value of the ghost cell of face nb at tangential position (a, b) of box
// `id`, as af_gc_box (afivo/src/m_af_ghostcell.f90:64-120) sets it: copy of
// the same-level neighbour (copy_from_nb :654), physical boundary (bc_to_gc
// :173-279), or refinement boundary (mg_sides_rb, m_af_multigrid.f90:294-461,
// or af_gc_interp / af_gc_interp_lim, m_af_ghostcell.f90:394-605).  x1 and x2
// are the box's first and second interior cells along the face normal.
// `coarse` (default: var) holds the coarser level's values when the box's
// own level is kept in a different buffer.
// The part of a ghost value of a face without a same-level neighbour
// (nb_id <= 0) that does not depend on the box's own cells: the boundary
// value's term, or the coarse interpolation at a refinement boundary (and
// the coarse cell the limiter compares with).  ghost_finish() adds the
// terms of the box's cells x1, x2 in the order ghost_value() always used.
template <typename T>
struct GhostPart {
  T p, c1;
};
// ghost_coarse in two phases: the coarse values it reads (none on a
// physical boundary), then the part formed from them, so that a kernel can
// have the loads of several ghosts in flight before it forms any.
template <typename T>
struct CoarseVals {
  T v[5];
};
template <int D, int NC, typename T>
__device__ inline CoarseVals<T> ghost_coarse_load(const DBox &bx, const T *var, const VarBc &m, int nb, int a, int b,
                                                  const T *coarse = nullptr) {
  using L = Lay<D, NC, T>;
  constexpr int H = NC / 2;
  const int d = nb / 2;
  const bool low = (nb % 2) == 0;
  int t0, t1;
  tangential(D, d, t0, t1);
  CoarseVals<T> cv = {{T(0), T(0), T(0), T(0), T(0)}};
  if (bx.nb[nb] < 0) return cv;
  const T *P = (coarse ? coarse : var) + (size_t)(bx.pnb[nb] - 1) * L::STRIDE;
  int co[3] = {0, 0, 0};
  for (int dd = 0; dd < D; dd++) co[dd] = ((bx.ix[dd] - 1) & 1) * H;
  // coarse cell at tangential (qa, qb) on the coarse neighbour's face cn
  auto at = [&](int cn, int qa, int qb) {
    int q[3] = {0, 0, 0};
    q[d] = cn;
    q[t0] = qa;
    if (D == 3) q[t1] = qb;
    return P[L::at(q[0], q[1], q[2])];
  };
  const int cn = low ? NC : 1;
  if (m.rb == RB_MG) {
    const int i = (a + 1) >> 1, j = (b + 1) >> 1;
    if (D == 2) {
      cv.v[0] = at(cn, co[t0] + i + 1, 0);
      cv.v[1] = at(cn, co[t0] + i - 1, 0);
      cv.v[2] = at(cn, co[t0] + i, 0);
    } else {
      cv.v[0] = at(cn, co[t0] + i + 1, co[t1] + j);
      cv.v[1] = at(cn, co[t0] + i - 1, co[t1] + j);
      cv.v[2] = at(cn, co[t0] + i, co[t1] + j + 1);
      cv.v[3] = at(cn, co[t0] + i, co[t1] + j - 1);
      cv.v[4] = at(cn, co[t0] + i, co[t1] + j);
    }
    return cv;
  }
  const int a1 = co[t0] + ((a + 1) >> 1), a2 = a1 + 1 - 2 * (a & 1);
  if (D == 2) {
    cv.v[0] = at(cn, a1, 0);
    cv.v[1] = at(cn, a2, 0);
  } else {
    const int b1 = co[t1] + ((b + 1) >> 1), b2 = b1 + 1 - 2 * (b & 1);
    cv.v[0] = at(cn, a1, b1);
    cv.v[1] = at(cn, a2, b1);
    cv.v[2] = at(cn, a1, b2);
  }
  return cv;
}
template <int D, int NC, typename T>
__device__ inline GhostPart<T> ghost_coarse_form(const DBox &bx, const VarBc &m, int nb, int a, int b,
                                                 const CoarseVals<T> &cv) {
  const int d = nb / 2;
  const bool low = (nb % 2) == 0;
  GhostPart<T> g;
  g.c1 = T(0);
  if (bx.nb[nb] < 0) {
    T c0;
    if (m.type[nb] == BC_DIRICHLET)
      c0 = 2;
    else if (m.type[nb] == BC_NEUMANN)
      c0 = (T)bx.dr[d] * T(low ? -1 : 1);
    else
      c0 = 1;
    g.p = c0 * (T)bc_value(m, nb, bx.run);
    return g;
  }
  if (m.rb == RB_MG) {
    if (D == 2) {
      const T grad = T(0.125) * (cv.v[0] - cv.v[1]);
      g.p = (a & 1) ? cv.v[2] - grad : cv.v[2] + grad;
    } else {
      const T g1 = T(0.125) * (cv.v[0] - cv.v[1]);
      const T g2 = T(0.125) * (cv.v[2] - cv.v[3]);
      const T tc = cv.v[4];
      const T h = (a & 1) ? tc - g1 : tc + g1;
      g.p = (b & 1) ? h - g2 : h + g2;
    }
    return g;
  }
  const T third = T(1 / 3.0), sixth = T(1 / 6.0);
  if (D == 2) {
    g.p = T(0.5) * cv.v[0] + sixth * cv.v[1];
  } else {
    // the z-face case sums the off-diagonal coarse values in the other order
    const T c2 = (d == 2) ? cv.v[2] : cv.v[1], c3 = (d == 2) ? cv.v[1] : cv.v[2];
    g.p = third * cv.v[0] + sixth * c2 + sixth * c3;
  }
  g.c1 = cv.v[0];
  return g;
}
template <int D, int NC, typename T>
__device__ inline GhostPart<T> ghost_coarse(const DBox &bx, const T *var, const VarBc &m, int nb, int a, int b,
                                            const T *coarse = nullptr) {
  return ghost_coarse_form<D, NC, T>(bx, m, nb, a, b, ghost_coarse_load<D, NC, T>(bx, var, m, nb, a, b, coarse));
}

template <int D, int NC, typename T>
__device__ inline T ghost_finish(const DBox &bx, const VarBc &m, int nb, const GhostPart<T> &g, T x1, T x2) {
  if (bx.nb[nb] < 0) {
    T c1, c2;
    if (m.type[nb] == BC_DIRICHLET) {
      c1 = -1; c2 = 0;
    } else if (m.type[nb] == BC_NEUMANN) {
      c1 = 1; c2 = 0;
    } else {
      c1 = 0; c2 = 0;
    }
    return g.p + c1 * x1 + c2 * x2;
  }
  if (m.rb == RB_MG) return T(0.5) * g.p + T(0.75) * x1 - T(0.25) * x2;
  T v = g.p + T(1 / 3.0) * x1;
  if (m.rb == RB_INTERP_LIM && v > 2 * g.c1) v = 2 * g.c1;
  return v;
}

template <int D, int NC, typename T>
__device__ inline T ghost_value(const DBox *__restrict__ boxes, const DBox &bx, const T *var, const VarBc &m, int nb,
                                int a, int b, T x1, T x2, const T *coarse = nullptr) {
  using L = Lay<D, NC, T>;
  const int d = nb / 2;
  const bool low = (nb % 2) == 0;
  int t0, t1;
  tangential(D, d, t0, t1);
  const int nb_id = bx.nb[nb];
  if (nb_id > 0) {
    const T *w = var + (size_t)(nb_id - 1) * L::STRIDE;
    int q[3] = {0, 0, 0};
    q[t0] = a;
    if (D == 3) q[t1] = b;
    q[d] = low ? NC : 1;
    return w[L::at(q[0], q[1], q[2])];
  }
  return ghost_finish<D, NC, T>(bx, m, nb, ghost_coarse<D, NC, T>(bx, var, m, nb, a, b, coarse), x1, x2);
}

// Position of face cell (nb, a, b) at normal index p.
template <int D>
__device__ inline void face_pos(int nb, int a, int b, int pn, int *p) {
  const int d = nb / 2;
  int t0, t1;
  tangential(D, d, t0, t1);
  p[0] = p[1] = p[2] = 0;
  p[t0] = a;
  if (D == 3) p[t1] = b;
  p[d] = pn;
}

// Corner (2D) / edge + corner (3D) ghost cells (af_gc_box_corner :125-170,
// af_edge_gc_extrap :885-924, af_corner_gc_extrap :860-879).
__constant__ int c_edge_dim3[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
__constant__ int c_edge_dir3[12][3] = {{0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1},
                                       {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
                                       {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0}};

template <int D, int NC, typename T>
__device__ inline void gc_edge_cell(const DBox &bx, T *var, int id, int n, int qq) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  T *u = var + (size_t)(id - 1) * L::STRIDE;
  const int dim = c_edge_dim3[n];
  const int *dir = c_edge_dir3[n];
  int lo[3];
  for (int dd = 0; dd < 3; dd++) lo[dd] = (dir[dd] > 0) ? NC + 1 : 0;
  lo[dim] = qq;
  const int nb_id = bx.nbmat[nbmat_at<D>(dir[0], dir[1], dir[2])];
  if (nb_id > 0) {
    const T *w = var + (size_t)(nb_id - 1) * L::STRIDE;
    u[L::at(lo[0], lo[1], lo[2])] = w[L::at(lo[0] - dir[0] * NC, lo[1] - dir[1] * NC, lo[2] - dir[2] * NC)];
  } else {
    const int o0 = (dim + 1) % 3, o1 = (dim + 2) % 3;
    int di[3];
    for (int dd = 0; dd < 3; dd++) di[dd] = 1 - 2 * (lo[dd] & 1);
    di[dim] = 0;
    int ia[3] = {lo[0], lo[1], lo[2]}, ib[3] = {lo[0], lo[1], lo[2]};
    ia[o0] += di[o0];
    ib[o1] += di[o1];
    u[L::at(lo[0], lo[1], lo[2])] = u[L::at(ia[0], ia[1], ia[2])] + u[L::at(ib[0], ib[1], ib[2])] -
                                     u[L::at(lo[0] + di[0], lo[1] + di[1], lo[2] + di[2])];
  }
}

template <int D, int NC, typename T>
__device__ inline void gc_corner_cell(const DBox &bx, T *var, int id, int n) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  T *u = var + (size_t)(id - 1) * L::STRIDE;
  const int cx = n & 1, cy = (n >> 1) & 1, cz = (n >> 2) & 1;
  const int dnb[3] = {2 * cx - 1, 2 * cy - 1, (D == 3) ? 2 * cz - 1 : 0};
  const int lo[3] = {cx * (NC + 1), cy * (NC + 1), (D == 3) ? cz * (NC + 1) : 0};
  const int nb_id = bx.nbmat[nbmat_at<D>(dnb[0], dnb[1], dnb[2])];
  if (nb_id > 0) {
    const T *w = var + (size_t)(nb_id - 1) * L::STRIDE;
    u[L::at(lo[0], lo[1], lo[2])] = w[L::at(lo[0] - dnb[0] * NC, lo[1] - dnb[1] * NC, lo[2] - dnb[2] * NC)];
  } else if (D == 2) {
    const int di = 1 - 2 * (lo[0] & 1), dj = 1 - 2 * (lo[1] & 1);
    u[L::at(lo[0], lo[1], 0)] =
        u[L::at(lo[0] + di, lo[1], 0)] + u[L::at(lo[0], lo[1] + dj, 0)] - u[L::at(lo[0] + di, lo[1] + dj, 0)];
  } else {
    const int di = 1 - 2 * (lo[0] & 1), dj = 1 - 2 * (lo[1] & 1), dk = 1 - 2 * (lo[2] & 1);
    const int x = lo[0], y = lo[1], z = lo[2];
    u[L::at(x, y, z)] = u[L::at(x, y + dj, z + dk)] + u[L::at(x + di, y, z + dk)] +
                         u[L::at(x + di, y + dj, z)] - 2 * u[L::at(x + di, y + dj, z + dk)];
  }
}

// af_gc_box for all boxes of a list.  Corners need the face ghosts of the
// same box, so with corners the box must be handled by one workgroup
// (bpb = 1); face-only fills may be split over several workgroups.
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_gc(const int *__restrict__ ids, const DBox *__restrict__ boxes, T *var,
                                                VarBc m, int corners, int bpb, int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const int id = pt.id;
    const DBox &bx = boxes[id - 1];
    T *u = var + (size_t)(id - 1) * L::STRIDE;
    constexpr int NFC = 2 * D * G::NTAN;
    for (int q = pt.q0; q < NFC; q += pt.stride) {
      const int nb = q / G::NTAN, r = q % G::NTAN;
      const int a = r % NC + 1, b = (D == 3) ? r / NC + 1 : 0;
      const bool low = (nb % 2) == 0;
      int p[3];
      face_pos<D>(nb, a, b, low ? 1 : NC, p);
      const T x1 = u[L::at(p[0], p[1], p[2])];
      face_pos<D>(nb, a, b, low ? 2 : NC - 1, p);
      const T x2 = u[L::at(p[0], p[1], p[2])];
      const T g = ghost_value<D, NC, T>(boxes, bx, var, m, nb, a, b, x1, x2);
      face_pos<D>(nb, a, b, low ? 0 : NC + 1, p);
      u[L::at(p[0], p[1], p[2])] = g;
    }
    if (!corners) continue;
    gsync<BS, BB>();
    if (D == 3) {
      for (int q = ltid<BS>(); q < 12 * NC; q += BS) gc_edge_cell<D, NC, T>(bx, var, id, q / NC, q % NC + 1);
      gsync<BS, BB>();
    }
    for (int q = ltid<BS>(); q < (1 << D); q += BS) gc_corner_cell<D, NC, T>(bx, var, id, q);
  }
}

// Red-black Gauss-Seidel half-sweep (stencil_gsrb_357,
// m_af_stencil.f90:838-998) fused with the ghost-cell fill that the
// reference performs before it (gsrb_boxes, m_af_multigrid.f90:648-687):
// a neighbour outside the box is taken from ghost_value() with the same
// inputs af_gc_box would use, so the updated cells are bit-identical while
// no separate ghost-fill pass is needed between half-sweeps.  (The stored
// ghost layer is refreshed once after the last half-sweep.)  stored: the
// first half-sweep of a V-cycle's finest level reads the stored layer as
// the reference does, also where it is out of date (the neighbours of boxes
// just added or removed by refinement).
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_gsrb(const int *__restrict__ ids, const DBox *__restrict__ boxes, T *phi,
                                                  const T *__restrict__ rhs, int redblack, T lambda2, int cyl, VarBc m,
                                                  int stored, int bpb, int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const int id = pt.id;
    const DBox &bx = boxes[id - 1];
    const Sten<D, T> s = lpl_stencil<D, T>(bx, lambda2);
    T *u = phi + (size_t)(id - 1) * L::STRIDE;
    const T *r = rhs + (size_t)(id - 1) * L::STRIDE;
    constexpr int HALF = G::NINT / 2;
    for (int q = pt.q0; q < HALF; q += pt.stride) {
      const int mm = q % (NC / 2);
      const int j = (q / (NC / 2)) % NC + 1;
      const int k = (D == 3) ? q / ((NC / 2) * NC) + 1 : 0;
      const int i0 = 2 - ((redblack ^ (D == 3 ? k + j : j)) & 1);
      const int i = i0 + 2 * mm;
      const int x = L::in_at(i, j, k);
      const int pos[3] = {i, j, k};
      T nbv[6];
      for (int nb = 0; nb < 2 * D; nb++) {
        const int d = nb / 2;
        const bool low = (nb % 2) == 0;
        const bool edge = low ? pos[d] == 1 : pos[d] == NC;
        if (!edge || (stored && !L::SPLIT)) {
          nbv[nb] = u[L::nb_at(i, j, k, d, low ? -1 : 1)];
        } else {
          int t0, t1;
          tangential(D, d, t0, t1);
          const T x2 = u[L::nb_at(i, j, k, d, low ? 1 : -1)];
          nbv[nb] = ghost_value<D, NC, T>(boxes, bx, phi, m, nb, pos[t0], (D == 3) ? pos[t1] : 0, u[x], x2);
        }
      }
      if (D == 2) {
        if (cyl) {
          T c1, c2, c3;
          cyl_coeffs<T>(bx, i, s.c, c1, c2, c3);
          T inv1 = T(1) / c1;
          u[x] = (r[x] - c2 * nbv[0] - c3 * nbv[1] - s.c[3] * nbv[2] - s.c[4] * nbv[3]) * inv1;
        } else {
          T inv_c1 = T(1) / s.c[0];
          u[x] = (r[x] - s.c[1] * nbv[0] - s.c[2] * nbv[1] - s.c[3] * nbv[2] - s.c[4] * nbv[3]) * inv_c1;
        }
      } else {
        T inv_c1 = T(1) / s.c[0];
        u[x] = (r[x] - s.c[1] * nbv[0] - s.c[2] * nbv[1] - s.c[3] * nbv[2] - s.c[4] * nbv[3] - s.c[5] * nbv[4] -
                s.c[6] * nbv[5]) *
               inv_c1;
      }
    }
  }
}

// tmp = rhs - L phi on interior cells (residual_box, m_af_multigrid.f90:801-810);
// for leaf boxes the max |tmp| is also reduced into *leafmax
// (af_tree_maxabs_cc, afivo/src/m_af_utils.f90:773).
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_residual(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                      const T *__restrict__ phi, const T *__restrict__ rhs, T *tmp,
                                                      T lambda2, int cyl, unsigned long long *leafmax, int bpb,
                                                      int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  __shared__ double sh[BS / 64];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const int id = pt.id;
    const DBox &bx = boxes[id - 1];
    const Sten<D, T> s = lpl_stencil<D, T>(bx, lambda2);
    const size_t base = (size_t)(id - 1) * L::STRIDE;
    double mx = 0;
    for (int q = pt.q0; q < G::NINT; q += pt.stride) {
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      const int x = L::in_at(i, j, k);
      T lp = apply_cell<D, NC, T, L>(phi + base, s, cyl != 0, bx, i, j, k);
      const T rv = rhs[base + x] - lp;
      tmp[base + x] = rv;
      mx = fmax(mx, (double)fabs(rv));
    }
    if (leafmax && bx.children[0] == 0) {
      mx = group_max<BS, BB>(mx, sh);
      if (ltid<BS>() == 0) atomic_max_nonneg(run_slot(leafmax, bx), mx);
    }
  }
}

// Positions in layout L of the 2^D children cells of a parent cell:
// child c is (i_f + (c & 1), jf + ((c >> 1) & 1), kf + ((c >> 2) & 1)),
// with i_f, jf (and kf) odd.  (In the colour-split layout the first cell
// of each pair of the x direction is at an even lexicographic index, so both
// share one position within their colour blocks.)
template <int D, int NC, typename L>
__host__ __device__ inline void child_pos(int i_f, int jf, int kf, int *pos) {
  constexpr int nch = (D == 3) ? 8 : 4;
  if constexpr (L::SPLIT) {
    const int h0 = ((i_f - 1) + NC * ((jf - 1) + NC * (D == 3 ? kf - 1 : 0))) >> 1;
    const int col0 = (i_f + jf + kf) & 1;
    for (int c = 0; c < nch; c++) {
      const int di = c & 1, dj = (c >> 1) & 1, dk = (c >> 2) & 1;
      pos[c] = ((col0 + di + dj + dk) & 1) * L::HALF + h0 + dj * (NC / 2) + dk * (NC * NC / 2);
    }
  } else {
    for (int c = 0; c < nch; c++) pos[c] = L::in_at(i_f + (c & 1), jf + ((c >> 1) & 1), kf + ((c >> 2) & 1));
  }
}

// Restriction of one parent cell from its 2^D children values
// (af_restrict_box, afivo/src/m_af_restrict.f90:62-136).
template <int D, int NC, typename T, typename L = LayStd<D, NC>>
__device__ inline T restrict_cell(const T *f, int i_f, int jf, int kf, bool cyl_geom, const DBox &bp, int ic) {
  int p[8];
  child_pos<D, NC, L>(i_f, jf, kf, p);
  if (D == 2) {
    if (cyl_geom) {
      T r = (T)bp.r_min[0] + (T(ic) - T(0.5)) * (T)bp.dr[0];
      T tmp = T(0.25) * (T)bp.dr[0] / r;
      T w1 = T(1) - tmp, w2 = T(1) + tmp;
      return T(0.25) * (w1 * (f[p[0]] + f[p[2]]) + w2 * (f[p[1]] + f[p[3]]));
    }
    T s = f[p[0]];
    s = s + f[p[1]];
    s = s + f[p[2]];
    s = s + f[p[3]];
    return T(0.25) * s;
  } else {
    T s = f[p[0]];
    for (int c = 1; c < 8; c++) s = s + f[p[c]];
    return T(0.125) * s;
  }
}

// Restrict variable `src` of child boxes into `dst` of their parents.
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_restrict(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                      const T *src, T *dst, int use_geom, int cyl, int bpb, int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  constexpr int H = NC / 2;
  constexpr int NQ = (D == 3) ? H * H * H : H * H;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const int id = pt.id;
    const DBox &bc = boxes[id - 1];
    const DBox &bp = boxes[bc.parent - 1];
    int off[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) off[d] = ((bc.ix[d] - 1) & 1) * H;
    const T *f = src + (size_t)(id - 1) * L::STRIDE;
    T *u = dst + (size_t)(bc.parent - 1) * L::STRIDE;
    for (int q = pt.q0; q < NQ; q += pt.stride) {
      const int i = q % H + 1, j = (q / H) % H + 1, k = (D == 3) ? q / (H * H) + 1 : 0;
      const int ic = off[0] + i, jc = off[1] + j, kc = (D == 3) ? off[2] + k : 0;
      u[L::at(ic, jc, kc)] =
          restrict_cell<D, NC, T, L>(f, 2 * i - 1, 2 * j - 1, (D == 3) ? 2 * k - 1 : 0, cyl && use_geom, bp, ic);
    }
  }
}

// update_coarse, first half (m_af_multigrid.f90:703-716): per child box the
// residual rhs - L phi is restricted (with geometry) into the parent's tmp
// and phi (without geometry) into the parent's phi.  The residual is formed
// on the fly, which leaves the child's tmp untouched exactly as the
// reference's save/restore of tmp does.  With store_residual the residual is
// also kept in the child's tmp (set_coarse_phi_rhs, :742-776).
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_restrict_residual(const int *__restrict__ ids,
                                                               const DBox *__restrict__ boxes, T *phi,
                                                               const T *__restrict__ rhs, T *tmp, T lambda2, int cyl,
                                                               int store_residual, int bpb, int nw) {
  using L = Lay<D, NC, T>;
  constexpr int H = NC / 2;
  constexpr int NQ = (D == 3) ? H * H * H : H * H;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const int id = pt.id;
    const DBox &bc = boxes[id - 1];
    const DBox &bp = boxes[bc.parent - 1];
    const Sten<D, T> s = lpl_stencil<D, T>(bc, lambda2);
    const size_t cb = (size_t)(id - 1) * L::STRIDE, pb = (size_t)(bc.parent - 1) * L::STRIDE;
    int off[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) off[d] = ((bc.ix[d] - 1) & 1) * H;
    for (int q = pt.q0; q < NQ; q += pt.stride) {
      const int i = q % H + 1, j = (q / H) % H + 1, k = (D == 3) ? q / (H * H) + 1 : 0;
      const int i_f = 2 * i - 1, jf = 2 * j - 1, kf = (D == 3) ? 2 * k - 1 : 0;
      // residuals of the 2^D children cells, in the summation order of
      // af_restrict_box (i fastest, then j, then k)
      T rs[8];
      int cp[8];
      child_pos<D, NC, L>(i_f, jf, kf, cp);
      const int nch = (D == 3) ? 8 : 4;
      for (int c = 0; c < nch; c++) {
        const int ii = i_f + (c & 1), jj = jf + ((c >> 1) & 1), kk = kf + ((c >> 2) & 1);
        T lp = apply_cell<D, NC, T, L>(phi + cb, s, cyl != 0, bc, ii, jj, kk);
        rs[c] = rhs[cb + cp[c]] - lp;
        if (store_residual) tmp[cb + cp[c]] = rs[c];
      }
      const int ic = off[0] + i, jc = off[1] + j, kc = (D == 3) ? off[2] + k : 0;
      T rv;
      if (D == 2) {
        if (cyl) {
          T r = (T)bp.r_min[0] + (T(ic) - T(0.5)) * (T)bp.dr[0];
          T t = T(0.25) * (T)bp.dr[0] / r;
          T w1 = T(1) - t, w2 = T(1) + t;
          rv = T(0.25) * (w1 * (rs[0] + rs[2]) + w2 * (rs[1] + rs[3]));
        } else {
          T sm = rs[0];
          sm = sm + rs[1];
          sm = sm + rs[2];
          sm = sm + rs[3];
          rv = T(0.25) * sm;
        }
      } else {
        T sm = rs[0];
        for (int c = 1; c < 8; c++) sm = sm + rs[c];
        rv = T(0.125) * sm;
      }
      tmp[pb + L::in_at(ic, jc, kc)] = rv;
      phi[pb + L::in_at(ic, jc, kc)] = restrict_cell<D, NC, T, L>(phi + cb, i_f, jf, kf, false, bp, ic);
    }
  }
}

// update_coarse / set_coarse_phi_rhs, second half (m_af_multigrid.f90:726-736):
// on parents, rhs = L phi (interior) + tmp (all cells); optionally tmp = phi.
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_coarse_rhs(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                        const T *__restrict__ phi, T *rhs, T *tmp, T lambda2, int cyl,
                                                        int copy_phi, int bpb, int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const int id = pt.id;
    const DBox &bx = boxes[id - 1];
    const Sten<D, T> s = lpl_stencil<D, T>(bx, lambda2);
    const size_t base = (size_t)(id - 1) * L::STRIDE;
    for (int x = pt.q0; x < G::NCELL; x += pt.stride) {
      int i, j, k;
      bool inner;
      if constexpr (L::SPLIT) {
        inner = x < G::NINT;
        if (inner) colour_cell<D, NC>(x >= L::HALF, x >= L::HALF ? x - L::HALF : x, i, j, k);
      } else {
        i = x % G::NG, j = (x / G::NG) % G::NG, k = (D == 3) ? x / (G::NG * G::NG) : 0;
        inner = i >= 1 && i <= NC && j >= 1 && j <= NC && (D == 2 || (k >= 1 && k <= NC));
      }
      T rv = inner ? apply_cell<D, NC, T, L>(phi + base, s, cyl != 0, bx, i, j, k) : rhs[base + x];
      rhs[base + x] = rv + tmp[base + x];
      if (copy_phi) tmp[base + x] = phi[base + x];
    }
  }
}

// Whole-box elementwise operations (afivo/src/m_af_utils.f90:385-576).
enum BoxOp { OP_COPY = 0, OP_CLEAR = 1, OP_SUB_FROM = 2, OP_AXPY_NEG = 3 };
// OP_COPY: b = a; OP_CLEAR: b = 0; OP_SUB_FROM: b = a - b; OP_AXPY_NEG: b = b - s*a
// A box's NCELL words are a whole number of 16-byte words (every box size
// here), so the operations run on 16-byte words: each lane moves 16 bytes
// per access.
template <typename T>
struct alignas(16) Word16 {
  T v[16 / sizeof(T)];
};
template <int NCELL, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_boxop(const int *__restrict__ ids, const T *a, T *b, int op, T s, int bpb,
                                                   int nw) {
  using V = Word16<T>;
  constexpr int NL = 16 / sizeof(T);
  static_assert((NCELL * sizeof(T)) % 16 == 0, "boxes of whole 16-byte words");
  constexpr int NV = NCELL / NL;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const V *av = reinterpret_cast<const V *>(a + (size_t)(pt.id - 1) * NCELL);
    V *bv = reinterpret_cast<V *>(b + (size_t)(pt.id - 1) * NCELL);
    for (int x = pt.q0; x < NV; x += pt.stride) {
      V r;
      switch (op) {
        case OP_COPY: r = av[x]; break;
        case OP_CLEAR:
          for (int l = 0; l < NL; l++) r.v[l] = T(0);
          break;
        case OP_SUB_FROM: {
          const V p = av[x], q = bv[x];
          for (int l = 0; l < NL; l++) r.v[l] = p.v[l] - q.v[l];
          break;
        }
        default: {
          const V p = av[x], q = bv[x];
          for (int l = 0; l < NL; l++) r.v[l] = q.v[l] - s * p.v[l];
          break;
        }
      }
      bv[x] = r;
    }
  }
}

// phi(child) += P(tmp(parent)): stencil_prolong_248 with add
// (afivo/src/m_af_stencil.f90:695-815), weights of
// mg_box_prolong_linear_stencil (m_af_multigrid.f90:1267-1284).
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_prolong_add(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                         const T *__restrict__ src, T *dst, int bpb, int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  constexpr int H = NC / 2;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const int id = pt.id;
    const DBox &bc = boxes[id - 1];
    const T *P = src + (size_t)(bc.parent - 1) * L::STRIDE;
    T *u = dst + (size_t)(id - 1) * L::STRIDE;
    int off[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) off[d] = ((bc.ix[d] - 1) & 1) * H;
    for (int q = pt.q0; q < G::NINT; q += pt.stride) {
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      const int i1 = off[0] + ((i + 1) >> 1), i2 = i1 + 1 - 2 * (i & 1);
      const int j1 = off[1] + ((j + 1) >> 1), j2 = j1 + 1 - 2 * (j & 1);
      const int x = L::at(i, j, k);
      if (D == 2) {
        const T w0 = T(9 / 16.0), w1 = T(3 / 16.0), w2 = T(3 / 16.0), w3 = T(1 / 16.0);
        u[x] = u[x] + w0 * P[L::at(i1, j1, 0)] + w1 * P[L::at(i2, j1, 0)] + w2 * P[L::at(i1, j2, 0)] +
               w3 * P[L::at(i2, j2, 0)];
      } else {
        const int k1 = off[2] + ((k + 1) >> 1), k2 = k1 + 1 - 2 * (k & 1);
        const T w0 = T(27 / 64.0), w1 = T(9 / 64.0), w3 = T(3 / 64.0), w7 = T(1 / 64.0);
        u[x] = u[x] + w0 * P[L::at(i1, j1, k1)] + w1 * P[L::at(i2, j1, k1)] + w1 * P[L::at(i1, j2, k1)] +
               w3 * P[L::at(i2, j2, k1)] + w1 * P[L::at(i1, j1, k2)] + w3 * P[L::at(i2, j1, k2)] +
               w3 * P[L::at(i1, j2, k2)] + w7 * P[L::at(i2, j2, k2)];
      }
    }
  }
}

// k_prolong_add for pools in the colour-split layout: the parent cells the
// box's cells interpolate from (its octant and one cell around it) are
// staged in LDS once, then every cell is formed from LDS, colour block by
// colour block (same weights and order as k_prolong_add).
template <int D, int NC, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_prolong_add_split(const int *__restrict__ ids,
                                                               const DBox *__restrict__ boxes,
                                                               const float *__restrict__ src, float *dst, int bpb,
                                                               int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, float>;
  constexpr int H = NC / 2, P = H + 2;
  constexpr int NP = (D == 3) ? P * P * P : P * P;
  __shared__ float s_all[BB][NP];
  float *sp = s_all[lgrp<BS, BB>()];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const DBox &bc = boxes[pt.id - 1];
    const float *Pb = src + (size_t)(bc.parent - 1) * L::STRIDE;
    float *u = dst + (size_t)(pt.id - 1) * L::STRIDE;
    int off[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) off[d] = ((bc.ix[d] - 1) & 1) * H;
    // parent cells off[d] .. off[d] + H + 1 along each axis
    for (int q = ltid<BS>(); q < NP; q += BS) {
      const int a = q % P, b = (q / P) % P, c = (D == 3) ? q / (P * P) : 0;
      sp[q] = Pb[L::at(off[0] + a, off[1] + b, (D == 3) ? off[2] + c : 0)];
    }
    gsync<BS, BB>();
    auto S = [&](int a, int b, int c) { return sp[a + P * (b + P * c)]; };
    for (int x = pt.q0; x < G::NINT; x += pt.stride) {
      int i, j, k;
      colour_cell<D, NC>(x >= L::HALF, x >= L::HALF ? x - L::HALF : x, i, j, k);
      const int i1 = (i + 1) >> 1, i2 = i1 + 1 - 2 * (i & 1);
      const int j1 = (j + 1) >> 1, j2 = j1 + 1 - 2 * (j & 1);
      if (D == 2) {
        const float w0 = 9 / 16.0f, w1 = 3 / 16.0f, w2 = 3 / 16.0f, w3 = 1 / 16.0f;
        u[x] = u[x] + w0 * S(i1, j1, 0) + w1 * S(i2, j1, 0) + w2 * S(i1, j2, 0) + w3 * S(i2, j2, 0);
      } else {
        const int k1 = (k + 1) >> 1, k2 = k1 + 1 - 2 * (k & 1);
        const float w0 = 27 / 64.0f, w1 = 9 / 64.0f, w3 = 3 / 64.0f, w7 = 1 / 64.0f;
        u[x] = u[x] + w0 * S(i1, j1, k1) + w1 * S(i2, j1, k1) + w1 * S(i1, j2, k1) + w3 * S(i2, j2, k1) +
               w1 * S(i1, j1, k2) + w3 * S(i2, j1, k2) + w3 * S(i1, j2, k2) + w7 * S(i2, j2, k2);
      }
    }
    gsync<BS, BB>();  // (the group's LDS is reused by its next item)
  }
}

// max |u| over interior cells of a box list (af_tree_maxabs_cc,
// afivo/src/m_af_utils.f90:773).
template <int D, int NC, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_maxabs(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                    const double *__restrict__ u, unsigned long long *out, int bpb,
                                                    int nw) {
  using G = Geo<D, NC>;
  __shared__ double sh[BS / 64];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const size_t base = (size_t)(pt.id - 1) * G::NCELL;
    double m = 0;
    for (int q = pt.q0; q < G::NINT; q += pt.stride) {
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      m = fmax(m, fabs(u[base + G::idx(i, j, k)]));
    }
    m = group_max<BS, BB>(m, sh);
    if (ltid<BS>() == 0) atomic_max_nonneg(run_slot(out, boxes[pt.id - 1]), m);
  }
}

// Exact level-1 solve by fast diagonalisation (replaces hypre PFMG,
// afivo/src/m_coarse_solver.f90:421): fold boundary values into the rhs as
// coarse_solver_set_rhs_phi (:286-338) does, transform with V^-1 along each
// axis, divide by the eigenvalue sums, transform back with V.  One
// workgroup of CBS threads per level-1 box (one per run of a forest); the
// box lives in LDS.
// TS is the storage type of rhs and phi; the transforms always run in FP64.
template <int D, int NC, int CBS, typename TS>
__global__ void __launch_bounds__(CBS) k_coarse_solve(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                      const TS *rhs,
                                                      TS *phi, VarBc m, double lambda2, int cyl,
                                                      const double *__restrict__ vinv,
                                                      const double *__restrict__ v,
                                                      const double *__restrict__ eig) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, TS>;
  constexpr int N = G::NINT;
  constexpr int PER = (N + CBS - 1) / CBS;
  __shared__ double buf[N];
  const int id = ids[blockIdx.x];
  if (id <= 0) return;  // run masked out
  const DBox &bx = boxes[id - 1];
  const Sten<D, double> s = lpl_stencil<D, double>(bx, lambda2);
  const TS *r = rhs + (size_t)(id - 1) * L::STRIDE;
  auto lin = [](int i, int j, int k) { return (i - 1) + NC * ((j - 1) + NC * (D == 3 ? k - 1 : 0)); };
  for (int q = threadIdx.x; q < N; q += CBS) {
    int i, j, k;
    interior_ijk<D, NC>(q, i, j, k);
    double bval = (double)r[L::in_at(i, j, k)];
    int p[3] = {i, j, (D == 3) ? k : 1};
    for (int nb = 0; nb < 2 * D; nb++) {
      if (bx.nb[nb] >= 0) continue;
      const int d = nb / 2;
      const bool low = (nb % 2) == 0;
      if (p[d] != (low ? 1 : NC)) continue;
      double cnb = s.c[1 + nb];
      if (cyl && d == 0) {
        double dr = bx.dr[0];
        double rr = bx.r_min[0] + (i - 0.5) * dr;
        double inv_r = 1 / rr;
        double rf = low ? (rr - 0.5 * dr) * inv_r : (rr + 0.5 * dr) * inv_r;
        cnb = rf * s.c[1 + nb];
      }
      double f;
      if (m.type[nb] == BC_DIRICHLET)
        f = -2 * cnb;
      else
        f = -(cnb * bx.dr[d]) * (double)(low ? -1 : 1);
      bval = bval + f * bc_value(m, nb, bx.run);
    }
    buf[lin(i, j, k)] = bval;
  }
  __syncthreads();
  for (int pass = 0; pass < 2 * D; pass++) {
    const int d = pass % D;
    const double *M = ((pass < D) ? vinv : v) + (size_t)d * NC * NC;
    if (pass == D) {
      for (int q = threadIdx.x; q < N; q += CBS) {
        int i, j, k;
        interior_ijk<D, NC>(q, i, j, k);
        double den = eig[i - 1] + eig[NC + j - 1];
        if (D == 3) den = den + eig[2 * NC + k - 1];
        den = den - lambda2;
        buf[q] = buf[q] / den;
      }
      __syncthreads();
    }
    double out[PER];
    for (int t = 0; t < PER; t++) {
      const int q = threadIdx.x + t * CBS;
      if (q >= N) break;
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      int p[3] = {i, j, (D == 3) ? k : 1};
      const int row = p[d];
      double sum = 0.0;
      for (int mm = 1; mm <= NC; mm++) {
        int qq[3] = {p[0], p[1], p[2]};
        qq[d] = mm;
        sum = sum + M[(row - 1) * NC + (mm - 1)] * buf[lin(qq[0], qq[1], qq[2])];
      }
      out[t] = sum;
    }
    __syncthreads();
    for (int t = 0; t < PER; t++) {
      const int q = threadIdx.x + t * CBS;
      if (q < N) buf[q] = out[t];
    }
    __syncthreads();
  }
  TS *u = phi + (size_t)(id - 1) * L::STRIDE;
  for (int q = threadIdx.x; q < N; q += CBS) {
    int i, j, k;
    interior_ijk<D, NC>(q, i, j, k);
    u[L::in_at(i, j, k)] = (TS)buf[q];
  }
}

// ---- mixed precision: the FP64 side of the defect correction ---------------
// The inner multigrid cycle works on the correction e (single precision,
// homogeneous boundary conditions) of the FP64 potential; these two kernels
// close each cycle in FP64.

// FP64 residual rhs - L phi on the interior cells of a box list, with every
// neighbour outside the box taken from ghost_value() (so the stored ghost
// layer of phi need not be current), stored in single precision as the rhs
// of the next correction cycle; max |residual| over leaf boxes is reduced
// into *leafmax.
template <int D, int NC, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_mp_residual(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                         const double *__restrict__ phi, const double *__restrict__ rhs,
                                                         float *r32, double lambda2, int cyl, VarBc m,
                                                         unsigned long long *leafmax, int bpb, int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, float>;
  __shared__ double sh[BS / 64];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const Part pt = part_of<BS>(ids, bpb, w);
    if (pt.id <= 0) continue;  // run masked out
    const int id = pt.id;
    const DBox &bx = boxes[id - 1];
    const Sten<D, double> s = lpl_stencil<D, double>(bx, lambda2);
    const size_t base = (size_t)(id - 1) * G::NCELL;
    const double *u = phi + base;
    float *r = r32 + (size_t)(id - 1) * L::STRIDE;
    constexpr int st[3] = {1, G::NG, G::NG * G::NG};
    double mx = 0;
    for (int q = pt.q0; q < G::NINT; q += pt.stride) {
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      const int x = G::idx(i, j, k);
      const int pos[3] = {i, j, k};
      const double uc = u[x];
      double nbv[6];
      for (int nb = 0; nb < 2 * D; nb++) {
        const int d = nb / 2;
        const bool low = (nb % 2) == 0;
        const bool edge = low ? pos[d] == 1 : pos[d] == NC;
        if (!edge) {
          nbv[nb] = u[low ? x - st[d] : x + st[d]];
        } else {
          int t0, t1;
          tangential(D, d, t0, t1);
          const double x2 = u[low ? x + st[d] : x - st[d]];
          nbv[nb] = ghost_value<D, NC, double>(boxes, bx, phi, m, nb, pos[t0], (D == 3) ? pos[t1] : 0, uc, x2);
        }
      }
      double lp;
      if (D == 2) {
        if (cyl) {
          double c1, c2, c3;
          cyl_coeffs<double>(bx, i, s.c, c1, c2, c3);
          lp = c1 * uc + c2 * nbv[0] + c3 * nbv[1] + s.c[3] * nbv[2] + s.c[4] * nbv[3];
        } else {
          lp = s.c[0] * uc + s.c[1] * nbv[0] + s.c[2] * nbv[1] + s.c[3] * nbv[2] + s.c[4] * nbv[3];
        }
      } else {
        lp = s.c[0] * uc + s.c[1] * nbv[0] + s.c[2] * nbv[1] + s.c[3] * nbv[2] + s.c[4] * nbv[3] +
             s.c[5] * nbv[4] + s.c[6] * nbv[5];
      }
      const double rv = rhs[base + x] - lp;
      r[L::in_at(i, j, k)] = (float)rv;
      mx = fmax(mx, fabs(rv));
    }
    if (bx.children[0] == 0) {
      mx = group_max<BS, BB>(mx, sh);
      if (ltid<BS>() == 0) atomic_max_nonneg(run_slot(leafmax, bx), mx);
    }
  }
}

// ---- LDS-staged forms (one workgroup per box) ---------------------------------
// The box interior is loaded once, coalesced, into LDS; the face ghosts that
// the pass needs are formed by ghost_value() into the LDS ghost layer (all
// threads of an iteration on the same face, so without divergence); the
// stencil then reads LDS only.  Same arithmetic, in the same order, as the
// direct kernels above.

template <int D, int NC, int BS, typename T, int BB = 1>
__device__ inline void lds_load_interior(T *s, const T *u) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  for (int q = ltid<BS>(); q < G::NINT; q += BS) {
    int i, j, k;
    interior_ijk<D, NC>(q, i, j, k);
    s[G::idx(i, j, k)] = u[L::in_at(i, j, k)];
  }
}

// Staging of a box and its face ghosts into LDS in two memory rounds: the
// interior and the same-level neighbour copies are loaded together (into
// registers, then LDS); after a barrier, the ghosts that depend on the box's
// own cells (physical boundary, refinement boundary) are formed from LDS.
// colour >= 0: only the ghosts next to interior cells of that GSRB colour,
// and in the colour-split FP32 layout only the interior cells of the other
// colour, the only ones a half-sweep of that colour reads (the few cells of
// the updated colour that those ghosts depend on are read from the pool).
// stored: the face ghosts are the box's own stored ghost layer (the
// reference's smoother reads that layer as it stands before its first
// half-sweep; FP64 layout only).
// Ends with a barrier.
template <int D, int NC, int BS, typename T, int BB = 1>
__device__ inline void lds_stage(T *s, const DBox *__restrict__ boxes, const DBox &bx, int id, const T *var,
                                 const VarBc &m, int colour, const T *coarse = nullptr, bool stored = false) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  stored = stored && !L::SPLIT;
  constexpr int NI = (G::NINT + BS - 1) / BS;
  constexpr int NFC = 2 * D * G::NTAN;
  constexpr int NF = (NFC + BS - 1) / BS;
  const T *u = var + (size_t)(id - 1) * L::STRIDE;
  const bool one = L::SPLIT && colour >= 0;
  const int n_load = one ? G::NINT / 2 : G::NINT;
  // interior cell (i, j, k) of load q and its position in the pool: in the
  // split layout the loads run over the colour blocks in order
  auto cell = [&](int q, int &i, int &j, int &k) {
    if constexpr (L::SPLIT) {
      constexpr int HALF = G::NINT / 2;
      const int x = q + (one ? (1 - colour) * HALF : 0);
      colour_cell<D, NC>(x >= HALF, x >= HALF ? x - HALF : x, i, j, k);
      return x;
    } else {
      interior_ijk<D, NC>(q, i, j, k);
      return L::in_at(i, j, k);
    }
  };
  // with a colour, the physical and refinement-boundary ghosts' parts that
  // do not depend on the box (the coarse values) are loaded in the first
  // round too, and so are the updated colour's cells they need in the split
  // layout: the second round then works in LDS only
  const bool pre = colour >= 0;
  T vi[NI], vf[NF], x1g[NF];
  GhostPart<T> gp[NF];
  int gi[NF];
  for (int t = 0; t < NI; t++) {
    const int q = ltid<BS>() + t * BS;
    if (q < n_load) {
      int i, j, k;
      vi[t] = u[cell(q, i, j, k)];
    }
  }
  for (int t = 0; t < NF; t++) {
    gi[t] = -1;
    const int q = ltid<BS>() + t * BS;
    if (q >= NFC) continue;
    const int nb = q / G::NTAN, r = q % G::NTAN;
    const int nb_id = bx.nb[nb];
    if (nb_id <= 0 && !pre && !stored) continue;
    const int a = r % NC + 1, b = (D == 3) ? r / NC + 1 : 0;
    const bool low = (nb % 2) == 0;
    int p[3];
    face_pos<D>(nb, a, b, low ? 1 : NC, p);
    if (colour >= 0 && ((p[0] + p[1] + p[2] + colour) & 1)) continue;
    if (stored) {
      p[nb / 2] = low ? 0 : NC + 1;
      gi[t] = G::idx(p[0], p[1], p[2]);
      vf[t] = u[L::at(p[0], p[1], p[2])];
      continue;
    }
    if (nb_id <= 0) {
      gp[t] = ghost_coarse<D, NC, T>(bx, var, m, nb, a, b, coarse);
      if (one) x1g[t] = u[L::in_at(p[0], p[1], p[2])];
      continue;
    }
    int d = nb / 2;
    p[d] = low ? 0 : NC + 1;
    gi[t] = G::idx(p[0], p[1], p[2]);
    p[d] = low ? NC : 1;  // the neighbour's cell across the face
    vf[t] = var[(size_t)(nb_id - 1) * L::STRIDE + L::in_at(p[0], p[1], p[2])];
  }
  for (int t = 0; t < NI; t++) {
    const int q = ltid<BS>() + t * BS;
    if (q < n_load) {
      int i, j, k;
      cell(q, i, j, k);
      s[G::idx(i, j, k)] = vi[t];
    }
  }
  for (int t = 0; t < NF; t++)
    if (gi[t] >= 0) s[gi[t]] = vf[t];
  gsync<BS, BB>();
  bool any = false;
  for (int nb = 0; nb < 2 * D; nb++) any = any || bx.nb[nb] <= 0;
  if (any && !stored) {
    if (pre) {
      for (int t = 0; t < NF; t++) {
        const int q = ltid<BS>() + t * BS;
        if (q >= NFC) continue;
        const int nb = q / G::NTAN, r = q % G::NTAN;
        if (bx.nb[nb] > 0) continue;
        const int a = r % NC + 1, b = (D == 3) ? r / NC + 1 : 0;
        const bool low = (nb % 2) == 0;
        int p[3];
        face_pos<D>(nb, a, b, low ? 1 : NC, p);
        if ((p[0] + p[1] + p[2] + colour) & 1) continue;
        const T x1 = one ? x1g[t] : s[G::idx(p[0], p[1], p[2])];
        face_pos<D>(nb, a, b, low ? 2 : NC - 1, p);
        const T x2 = s[G::idx(p[0], p[1], p[2])];
        face_pos<D>(nb, a, b, low ? 0 : NC + 1, p);
        s[G::idx(p[0], p[1], p[2])] = ghost_finish<D, NC, T>(bx, m, nb, gp[t], x1, x2);
      }
    } else {
      for (int q = ltid<BS>(); q < NFC; q += BS) {
        const int nb = q / G::NTAN, r = q % G::NTAN;
        if (bx.nb[nb] > 0) continue;
        const int a = r % NC + 1, b = (D == 3) ? r / NC + 1 : 0;
        const bool low = (nb % 2) == 0;
        int p[3];
        face_pos<D>(nb, a, b, low ? 1 : NC, p);
        const T x1 = s[G::idx(p[0], p[1], p[2])];
        face_pos<D>(nb, a, b, low ? 2 : NC - 1, p);
        const T x2 = s[G::idx(p[0], p[1], p[2])];
        const T g = ghost_value<D, NC, T>(boxes, bx, var, m, nb, a, b, x1, x2, coarse);
        face_pos<D>(nb, a, b, low ? 0 : NC + 1, p);
        s[G::idx(p[0], p[1], p[2])] = g;
      }
    }
    gsync<BS, BB>();
  }
}

// k_gsrb with the box staged in LDS (levels with one workgroup per box).
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_gsrb_lds(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                      T *phi, const T *__restrict__ rhs, int redblack, T lambda2,
                                                      int cyl, VarBc m, int stored, int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  __shared__ T s_all[BB][G::NCELL];
  T *s = s_all[lgrp<BS, BB>()];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const DBox &bx = boxes[id - 1];
    const Sten<D, T> st = lpl_stencil<D, T>(bx, lambda2);
    T *u = phi + (size_t)(id - 1) * L::STRIDE;
    const T *r = rhs + (size_t)(id - 1) * L::STRIDE;
    constexpr int HALF = G::NINT / 2;
    constexpr int NH = (HALF + BS - 1) / BS;
    constexpr int sx = 1, sy = G::NG, sz = G::NG * G::NG;
    // updated cell t of this thread, and its rhs (loaded with the box)
    auto cell = [&](int q, int &i, int &j, int &k) {
      const int mm = q % (NC / 2);
      j = (q / (NC / 2)) % NC + 1;
      k = (D == 3) ? q / ((NC / 2) * NC) + 1 : 0;
      const int i0 = 2 - ((redblack ^ (D == 3 ? k + j : j)) & 1);
      i = i0 + 2 * mm;
    };
    T rr[NH];
    for (int t = 0; t < NH; t++) {
      const int q = ltid<BS>() + t * BS;
      if (q < HALF) {
        int i, j, k;
        cell(q, i, j, k);
        rr[t] = r[L::SPLIT ? (redblack & 1) * HALF + q : L::in_at(i, j, k)];
      }
    }
    // the updated cells satisfy (i + j + k + redblack) even
    lds_stage<D, NC, BS, T, BB>(s, boxes, bx, id, phi, m, redblack & 1, nullptr, stored != 0);
    for (int t = 0; t < NH; t++) {
      const int q = ltid<BS>() + t * BS;
      if (q >= HALF) break;
      int i, j, k;
      cell(q, i, j, k);
      const int x = G::idx(i, j, k);
      const T rx = rr[t];
      if (D == 2) {
        if (cyl) {
          T c1, c2, c3;
          cyl_coeffs<T>(bx, i, st.c, c1, c2, c3);
          T inv1 = T(1) / c1;
          s[x] = (rx - c2 * s[x - sx] - c3 * s[x + sx] - st.c[3] * s[x - sy] - st.c[4] * s[x + sy]) * inv1;
        } else {
          T inv_c1 = T(1) / st.c[0];
          s[x] = (rx - st.c[1] * s[x - sx] - st.c[2] * s[x + sx] - st.c[3] * s[x - sy] - st.c[4] * s[x + sy]) *
                 inv_c1;
        }
      } else {
        T inv_c1 = T(1) / st.c[0];
        s[x] = (rx - st.c[1] * s[x - sx] - st.c[2] * s[x + sx] - st.c[3] * s[x - sy] - st.c[4] * s[x + sy] -
                st.c[5] * s[x - sz] - st.c[6] * s[x + sz]) *
               inv_c1;
      }
    }
    gsync<BS, BB>();
    if (L::SPLIT) {
      // the updated colour's block
      for (int q = ltid<BS>(); q < HALF; q += BS) {
        int i, j, k;
        cell(q, i, j, k);
        u[(redblack & 1) * HALF + q] = s[G::idx(i, j, k)];
      }
    } else {
      // Whole interior rows back (the other colour unchanged): full-line
      // stores instead of every other word, so no line has to be fetched to
      // merge a partial write.  Neighbouring boxes only read this box's
      // other colour during the launch, which keeps its value.
      for (int q = ltid<BS>(); q < G::NINT; q += BS) {
        int i, j, k;
        interior_ijk<D, NC>(q, i, j, k);
        const int x = G::idx(i, j, k);
        u[x] = s[x];
      }
    }
    gsync<BS, BB>();  // (the group's LDS is reused by its next item)
  }
}

// The coarse values ghost_coarse_load reads for ghost (a, b) of a 3D face at
// a refinement boundary (mg_sides_rb, afivo/src/m_af_multigrid.f90:294-461;
// af_gc_interp(_lim), afivo/src/m_af_ghostcell.f90:394-605), in its order,
// from a PP x PP patch of the coarse neighbour's face layer: patch cell
// pa + PP pb holds the coarse cell at tangential (co[t0] + pa, co[t1] + pb),
// co the box's offset within its parent.
template <int NC>
__device__ inline CoarseVals<float> coarse_patch_vals(const float *cp, const VarBc &m, int a, int b) {
  constexpr int PP = NC / 2 + 2;
  auto at = [&](int qa, int qb) { return cp[qa + PP * qb]; };
  CoarseVals<float> cv = {{0.0f, 0.0f, 0.0f, 0.0f, 0.0f}};
  if (m.rb == RB_MG) {
    const int i = (a + 1) >> 1, j = (b + 1) >> 1;
    cv.v[0] = at(i + 1, j);
    cv.v[1] = at(i - 1, j);
    cv.v[2] = at(i, j + 1);
    cv.v[3] = at(i, j - 1);
    cv.v[4] = at(i, j);
    return cv;
  }
  const int a1 = (a + 1) >> 1, a2 = a1 + 1 - 2 * (a & 1);
  const int b1 = (b + 1) >> 1, b2 = b1 + 1 - 2 * (b & 1);
  cv.v[0] = at(a1, b1);
  cv.v[1] = at(a2, b1);
  cv.v[2] = at(a1, b2);
  return cv;
}

// k_gsrb_lds for 3D boxes in the colour-split single-precision layout.  A
// half-sweep reads only the other colour's interior, the face ghosts next to
// the updated cells and the updated colour's rhs; the updated colour's old
// values are never read (but for the cells next to faces without a
// same-level neighbour, whose ghosts use them).  Everything a box reads goes
// straight from memory into LDS (global_load_lds, no registers held): the
// other colour's block and the rhs as they lie in the pool (16-byte words),
// and for each face either the same-level neighbour's cells across it (one
// value per ghost, already the ghost) or the box's own cells next to it and,
// at a refinement boundary, a patch of the coarse neighbour's face layer.
// After one barrier those faces' ghosts are formed, after a second each
// thread updates four consecutive cells of the colour block per pass and
// stores them as one 16-byte word.  The stencil and ghost expressions and
// their operands are those of k_gsrb_lds, so the results are identical.
// With few registers and 25 KB of LDS, six boxes fit on a CU: all 1536
// boxes of a 256-CU finest level run at once.
constexpr int GSRB_SPLIT_NT = 256;
template <int D, int NC, int NT>
__global__ void __launch_bounds__(NT) k_gsrb_split(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                   float *phi, const float *__restrict__ rhs, int redblack,
                                                   float lambda2, VarBc m, int nw) {
  static_assert(D == 3, "the colour blocks of 3D boxes");
  using G = Geo<D, NC>;
  using L = Lay<D, NC, float>;
  constexpr int HALF = L::HALF;
  constexpr int NV = HALF / 4;               // 16-byte words of a colour block
  constexpr int NW = NV / NT;                // per thread
  constexpr int NE = G::NTAN / 2;            // ghosts of a face next to updated cells
  constexpr int NGH = 2 * D * NE;            // of the box
  constexpr int NGT = (NGH + NT - 1) / NT;   // per thread
  constexpr int PP = NC / 2 + 2;             // coarse patch of a face: its coarse cells and a ring
  static_assert(NV % NT == 0 && (L::STRIDE * sizeof(float)) % 16 == 0, "16-byte words of whole boxes");
  static_assert(NE % 64 == 0 && NT % 64 == 0 && PP * PP <= NE, "a face's ghosts in whole wavefronts");
  __shared__ float4 s_other4[NV];
  __shared__ float4 s_rhs4[NV];
  __shared__ float s_gh[2 * D][NE];  // the ghosts: face nb, item e
  __shared__ float s_x1[2 * D][NE];  // the box's cells next to a face without a same-level neighbour
  __shared__ float s_cp[2 * D][NE];  // the coarse patch of a face at a refinement boundary
  const float *s_other = reinterpret_cast<const float *>(s_other4);
  const int t = threadIdx.x, wv = __builtin_amdgcn_readfirstlane(t / 64);  // (the wavefront)
  const int uc = redblack & 1, oc = 1 - uc;  // updated cells: (i + j + k + redblack) even
  // ghost item e of face nb: tangential (a, b), next to an updated cell
  auto item_ab = [&](int nb, int e, int &a, int &b) {
    const int pd = (nb % 2) == 0 ? 1 : NC;
    b = e / (NC / 2) + 1;
    a = 2 * (e % (NC / 2)) + 1 + ((uc + pd + b + 1) & 1);  // (pd + a + b + uc) even
  };
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const DBox &bx = boxes[id - 1];
    const Sten<D, float> st = lpl_stencil<D, float>(bx, lambda2);
    float *u = phi + (size_t)(id - 1) * L::STRIDE;
    const float4 *ob = reinterpret_cast<const float4 *>(u + oc * HALF);
    const float4 *rb = reinterpret_cast<const float4 *>(rhs + (size_t)(id - 1) * L::STRIDE + uc * HALF);
    for (int q = 0; q < NW; q++) {
      __builtin_amdgcn_global_load_lds((const void *)(ob + t + q * NT),
                                       (__attribute__((address_space(3))) void *)(s_other4 + 64 * wv + q * NT), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void *)(rb + t + q * NT),
                                       (__attribute__((address_space(3))) void *)(s_rhs4 + 64 * wv + q * NT), 16, 0, 0);
    }
    int nbid[2 * D];
    for (int nb = 0; nb < 2 * D; nb++) nbid[nb] = bx.nb[nb];
    // the faces' values, one per lane: a wavefront's items lie on one face
    // (its face, its first item and the neighbour there are uniform), and
    // item e of a wavefront lands at e of the face's arrays
    for (int q = 0; q < NGT; q++) {
      const int g0 = 64 * wv + q * NT;
      if (g0 >= NGH) continue;
      const int nb = g0 / NE, e0 = g0 % NE, e = e0 + (t & 63);
      const int nbv = __builtin_amdgcn_readfirstlane(nbid[nb]);
      const bool low = (nb % 2) == 0;
      int a, b, p[3];
      item_ab(nb, e, a, b);
      if (nbv > 0) {  // the same-level neighbour's cell across the face: the ghost
        face_pos<D>(nb, a, b, low ? NC : 1, p);
        const float *src = phi + (size_t)(nbv - 1) * L::STRIDE + L::in_at(p[0], p[1], p[2]);
        __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)&s_gh[nb][e0],
                                         4, 0, 0);
      } else {
        face_pos<D>(nb, a, b, low ? 1 : NC, p);
        __builtin_amdgcn_global_load_lds((const void *)(u + L::in_at(p[0], p[1], p[2])),
                                         (__attribute__((address_space(3))) void *)&s_x1[nb][e0], 4, 0, 0);
        if (nbv == 0 && e < PP * PP) {  // refinement boundary: coarse patch cell e
          const int d = nb / 2;
          int t0, t1;
          tangential(D, d, t0, t1);
          int qc[3];
          qc[d] = low ? NC : 1;
          qc[t0] = ((bx.ix[t0] - 1) & 1) * (NC / 2) + e % PP;
          qc[t1] = ((bx.ix[t1] - 1) & 1) * (NC / 2) + e / PP;
          const float *P = phi + (size_t)(__builtin_amdgcn_readfirstlane(bx.pnb[nb]) - 1) * L::STRIDE;
          __builtin_amdgcn_global_load_lds((const void *)(P + L::at(qc[0], qc[1], qc[2])),
                                           (__attribute__((address_space(3))) void *)&s_cp[nb][e0], 4, 0, 0);
        }
      }
    }
    __syncthreads();
    // the ghosts of faces without a same-level neighbour (the box's second
    // cell along the normal is of the other colour: in the staged block)
    for (int q = 0; q < NGT; q++) {
      const int g0 = 64 * wv + q * NT;
      if (g0 >= NGH) continue;
      const int nb = g0 / NE, e = g0 % NE + (t & 63);
      const int nbv = __builtin_amdgcn_readfirstlane(nbid[nb]);
      if (nbv > 0) continue;
      const bool low = (nb % 2) == 0;
      int a, b, p[3];
      item_ab(nb, e, a, b);
      face_pos<D>(nb, a, b, low ? 2 : NC - 1, p);
      const float x2 = s_other[L::in_at(p[0], p[1], p[2]) - oc * HALF];
      CoarseVals<float> cv = {{0.0f, 0.0f, 0.0f, 0.0f, 0.0f}};
      if (nbv == 0) cv = coarse_patch_vals<NC>(s_cp[nb], m, a, b);
      s_gh[nb][e] =
          ghost_finish<D, NC, float>(bx, m, nb, ghost_coarse_form<D, NC, float>(bx, m, nb, a, b, cv), s_x1[nb][e], x2);
    }
    __syncthreads();
    const float inv_c1 = 1.0f / st.c[0];
    float4 *ub = reinterpret_cast<float4 *>(u + uc * HALF);
    for (int q = 0; q < NW; q++) {
      // four cells: half of a row (j, k) of the colour block, at
      // i = i0 + 2 (m0 + c); their other-colour neighbours in the row are
      // the five values from B on, in the rows and planes beside it four
      // consecutive values (one 16-byte word), unless they are ghosts
      const int h0 = 4 * (t + q * NT);
      const int row = h0 / (NC / 2), m0 = h0 % (NC / 2);
      const int j = row % NC + 1, k = row / NC + 1;
      const int i0 = 1 + ((1 + j + k + uc) & 1);
      const int R = row * (NC / 2);
      const int B = R + m0 + i0 - 2;
      // (one 16-byte word at R + m0 and one value before or after it:
      // consecutive lanes read consecutive words, free of bank conflicts)
      const bool lo_gh = i0 == 1 && m0 == 0, hi_gh = i0 == 2 && m0 == NC / 2 - 4;
      const float4 xc = s_other4[(R + m0) / 4];
      // (a face's ghosts are stored every other cell: cell r of the face
      // at r / 2)
      const float xe = (i0 == 1) ? (lo_gh ? s_gh[0][row >> 1] : s_other[max(B, 0)])
                                 : (hi_gh ? s_gh[1][row >> 1] : s_other[min(B + 4, HALF - 1)]);
      float xw[5];
      if (i0 == 1) {
        xw[0] = xe; xw[1] = xc.x; xw[2] = xc.y; xw[3] = xc.z; xw[4] = xc.w;
      } else {
        xw[0] = xc.x; xw[1] = xc.y; xw[2] = xc.z; xw[3] = xc.w; xw[4] = xe;
      }
      float ym[4], yp[4], zm[4], zp[4];
      auto four = [&](float *v, bool ghost, int nb, int r0, int off) {
        if (ghost) {
          for (int c = 0; c < 4; c++) v[c] = s_gh[nb][(r0 + i0 - 1 + 2 * (m0 + c)) >> 1];
        } else {
          const float4 w = s_other4[min(max(R + off + m0, 0), HALF - 4) / 4];
          v[0] = w.x;
          v[1] = w.y;
          v[2] = w.z;
          v[3] = w.w;
        }
      };
      four(ym, j == 1, 2, NC * (k - 1), -NC / 2);
      four(yp, j == NC, 3, NC * (k - 1), NC / 2);
      four(zm, k == 1, 4, NC * (j - 1), -NC * NC / 2);
      four(zp, k == NC, 5, NC * (j - 1), NC * NC / 2);
      const float4 rv = s_rhs4[t + q * NT];
      const float rr[4] = {rv.x, rv.y, rv.z, rv.w};
      float res[4];
      for (int c = 0; c < 4; c++)
        res[c] = (rr[c] - st.c[1] * xw[c] - st.c[2] * xw[c + 1] - st.c[3] * ym[c] - st.c[4] * yp[c] -
                  st.c[5] * zm[c] - st.c[6] * zp[c]) *
                 inv_c1;
      ub[t + q * NT] = make_float4(res[0], res[1], res[2], res[3]);
    }
    __syncthreads();  // (the LDS is reused by the next item)
  }
}

// k_gsrb_lds for 2D boxes of 8^2 cells in the colour-split single-precision
// layout (the ensembles' half-sweep).  A wavefront takes eight boxes, a lane
// one row of the box's updated colour (four cells, one 16-byte word).  All
// that the eight boxes read goes straight from memory into the wavefront's
// part of LDS (global_load_lds, one round, no registers held): each row's
// other-colour word and rhs word, the row's x ghost (the neighbour's cell
// across the face), the y ghosts (the neighbour's facing row of the other
// colour, one word per face), and for a face without a same-level neighbour
// the box's own cells next to it and the coarse values its ghosts need.
// Those ghosts are then formed in LDS, and each lane updates its row.  The
// ghost and stencil expressions and their operands are those of
// k_gsrb_lds, so the results are identical.  Wavefronts never wait for
// each other.
#define LDS_PTR(p) ((__attribute__((address_space(3))) void *)(p))
template <int NC>
__global__ void __launch_bounds__(256) k_gsrb_split2d(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                      float *phi, const float *__restrict__ rhs, int redblack,
                                                      float lambda2, int cyl, VarBc m, int n_boxes, int nw) {
  constexpr int D = 2;
  using L = Lay<D, NC, float>;
  constexpr int HALF = L::HALF;
  static_assert(NC == 8 && HALF == 32 && (L::STRIDE * sizeof(float)) % 16 == 0, "8^2 boxes: a row is one word");
  constexpr int NBW = 8;                       // boxes per wavefront
  __shared__ float4 s_o[4][64];                // other colour, row r of box bl at lane 8 bl + r
  __shared__ float4 s_r[4][64];                // rhs of the updated colour, the same
  __shared__ float4 s_y[4][64];                // y ghosts of box bl: face 2 at 8 bl, face 3 at 8 bl + 4
  __shared__ float s_gx[4][64];                // x ghost of row r
  __shared__ float s_x1[2][4][64];             // the box's cell next to a face: x item, y item
  __shared__ float s_cv[2][3][4][64];          // coarse values of an item at a refinement boundary
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x & 63;
  const int bl = lane >> 3, r = lane & 7;
  const int uc = redblack & 1, oc = 1 - uc;    // updated cells: (i + j + redblack) even
  for (int w = xcd_item(); w < nw; w += gridDim.x) {
    const int b = (4 * w + wv) * NBW + bl;
    const int id = b < n_boxes ? ids[b] : 0;
    const bool on = id > 0;  // (past the end, or a run masked out)
    const DBox &bx = boxes[on ? id - 1 : 0];
    float *u = phi + (size_t)(on ? id - 1 : 0) * L::STRIDE;
    const int j = r + 1, i0 = 1 + ((1 + j + uc) & 1);
    // x item: the ghost of row j on the side of its first or last cell
    const int nbx = (i0 == 1) ? 0 : 1;
    // y item: face 2 (r < 4) or 3, the cell i = i0y + 2 mm of the face row
    const int fy = r >> 2, mm = r & 3, nby = 2 + fy, jy = fy ? NC : 1;
    const int iy = 1 + ((1 + jy + uc) & 1) + 2 * mm;
    int nbvx = 0, nbvy = 0;
    if (on) {
      nbvx = bx.nb[nbx];
      nbvy = bx.nb[nby];
      __builtin_amdgcn_global_load_lds((const void *)(u + oc * HALF + 4 * r), LDS_PTR(&s_o[wv][0]), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void *)(rhs + (size_t)(id - 1) * L::STRIDE + uc * HALF + 4 * r),
                                       LDS_PTR(&s_r[wv][0]), 16, 0, 0);
    }
    // (a DMA's LDS address is its wavefront's base plus the lane's slot:
    // every lane's destination below is its own slot of a 64-entry array)
    if (on && nbvx > 0) {  // the neighbour's cell across the x face
      const int p = (nbx == 0) ? NC : 1;
      __builtin_amdgcn_global_load_lds((const void *)(phi + (size_t)(nbvx - 1) * L::STRIDE + L::in_at(p, j, 0)),
                                       LDS_PTR(&s_gx[wv][0]), 4, 0, 0);
    }
    if (on && nbvx <= 0)
      __builtin_amdgcn_global_load_lds((const void *)(u + L::in_at(nbx == 0 ? 1 : NC, j, 0)), LDS_PTR(&s_x1[0][wv][0]), 4,
                                       0, 0);
    if (on && nbvy > 0 && mm == 0) {  // the neighbour's facing row of the other colour
      const int row = fy ? 1 : NC;
      __builtin_amdgcn_global_load_lds(
          (const void *)(phi + (size_t)(nbvy - 1) * L::STRIDE + oc * HALF + 4 * (row - 1)), LDS_PTR(&s_y[wv][0]), 16, 0,
          0);
    }
    if (on && nbvy <= 0)
      __builtin_amdgcn_global_load_lds((const void *)(u + L::in_at(iy, jy, 0)), LDS_PTR(&s_x1[1][wv][0]), 4, 0, 0);
    // coarse values (ghost_coarse_load's, in its order) of the items at a
    // refinement boundary
    for (int it = 0; it < 2; it++) {
      const int nb = it ? nby : nbx, nbv = it ? nbvy : nbvx, a = it ? iy : j;
      if (!(on && nbv == 0)) continue;
      const int d = nb / 2, t0 = 1 - d;
      const bool low = (nb % 2) == 0;
      const float *P = phi + (size_t)(bx.pnb[nb] - 1) * L::STRIDE;
      const int co = ((bx.ix[t0] - 1) & 1) * (NC / 2), cn = low ? NC : 1;
      auto at = [&](int qa) {
        int q[3] = {0, 0, 0};
        q[d] = cn;
        q[t0] = qa;
        return P + L::at(q[0], q[1], q[2]);
      };
      int pos[3];
      int n = 3;
      if (m.rb == RB_MG) {
        const int i = (a + 1) >> 1;
        pos[0] = co + i + 1;
        pos[1] = co + i - 1;
        pos[2] = co + i;
      } else {
        const int a1 = co + ((a + 1) >> 1);
        pos[0] = a1;
        pos[1] = a1 + 1 - 2 * (a & 1);
        pos[2] = a1;
        n = 2;
      }
      for (int v = 0; v < 3; v++)
        if (v < n) __builtin_amdgcn_global_load_lds((const void *)at(pos[v]), LDS_PTR(&s_cv[it][v][wv][0]), 4, 0, 0);
    }
    // every lane's DMAs complete before any lane reads LDS: a wavefront's
    // fence orders no DMA, and the compiler does not see the DMAs' LDS
    // writes, so the wait is explicit and keeps the reads below it
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    gsync<64, 2>();
    // the ghosts of faces without a same-level neighbour (the box's second
    // cell from the face is of the other colour: in s_o)
    if (on && nbvx <= 0) {
      const float4 o = s_o[wv][lane];
      const float x2 = (nbx == 0) ? o.x : o.w;  // cell 2 or NC - 1 of the row
      CoarseVals<float> cv = {{0.0f, 0.0f, 0.0f, 0.0f, 0.0f}};
      if (nbvx == 0)
        for (int v = 0; v < 3; v++) cv.v[v] = s_cv[0][v][wv][lane];
      s_gx[wv][lane] = ghost_finish<D, NC, float>(bx, m, nbx, ghost_coarse_form<D, NC, float>(bx, m, nbx, j, 0, cv),
                                                  s_x1[0][wv][lane], x2);
    }
    if (on && nbvy <= 0) {
      const float4 o = s_o[wv][8 * bl + (fy ? NC - 2 : 1)];  // row 2 or NC - 1
      const float x2 = (mm == 0) ? o.x : (mm == 1) ? o.y : (mm == 2) ? o.z : o.w;
      CoarseVals<float> cv = {{0.0f, 0.0f, 0.0f, 0.0f, 0.0f}};
      if (nbvy == 0)
        for (int v = 0; v < 3; v++) cv.v[v] = s_cv[1][v][wv][lane];
      reinterpret_cast<float *>(&s_y[wv][8 * bl + 4 * fy])[mm] = ghost_finish<D, NC, float>(
          bx, m, nby, ghost_coarse_form<D, NC, float>(bx, m, nby, iy, 0, cv), s_x1[1][wv][lane], x2);
    }
    gsync<64, 2>();
    if (on) {
      // four cells of row j at i = i0 + 2 c; the x neighbours are the row's
      // other-colour word and its x ghost, the y neighbours the words of
      // rows j -+ 1 or the y ghosts
      const Sten<D, float> st = lpl_stencil<D, float>(bx, lambda2);
      const float4 xc = s_o[wv][lane];
      const float xe = s_gx[wv][lane];
      float xw[5];
      if (i0 == 1) {
        xw[0] = xe; xw[1] = xc.x; xw[2] = xc.y; xw[3] = xc.z; xw[4] = xc.w;
      } else {
        xw[0] = xc.x; xw[1] = xc.y; xw[2] = xc.z; xw[3] = xc.w; xw[4] = xe;
      }
      const float4 ym4 = (j == 1) ? s_y[wv][8 * bl] : s_o[wv][lane - 1];
      const float4 yp4 = (j == NC) ? s_y[wv][8 * bl + 4] : s_o[wv][lane + 1];
      const float ym[4] = {ym4.x, ym4.y, ym4.z, ym4.w}, yp[4] = {yp4.x, yp4.y, yp4.z, yp4.w};
      const float4 rv = s_r[wv][lane];
      const float rr[4] = {rv.x, rv.y, rv.z, rv.w};
      float res[4];
      if (cyl) {
        for (int c = 0; c < 4; c++) {
          float c1, c2, c3;
          cyl_coeffs<float>(bx, i0 + 2 * c, st.c, c1, c2, c3);
          const float inv1 = 1.0f / c1;
          res[c] = (rr[c] - c2 * xw[c] - c3 * xw[c + 1] - st.c[3] * ym[c] - st.c[4] * yp[c]) * inv1;
        }
      } else {
        const float inv_c1 = 1.0f / st.c[0];
        for (int c = 0; c < 4; c++)
          res[c] = (rr[c] - st.c[1] * xw[c] - st.c[2] * xw[c + 1] - st.c[3] * ym[c] - st.c[4] * yp[c]) * inv_c1;
      }
      reinterpret_cast<float4 *>(u + uc * HALF)[r] = make_float4(res[0], res[1], res[2], res[3]);
    }
    gsync<64, 2>();  // (the wavefront's LDS is reused by its next boxes)
  }
}

// k_mp_residual with the box staged in LDS.
template <int D, int NC, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_mp_residual_lds(const int *__restrict__ ids,
                                                             const DBox *__restrict__ boxes,
                                                             const double *__restrict__ phi,
                                                             const double *__restrict__ rhs, float *r32, double lambda2,
                                                             int cyl, VarBc m, unsigned long long *leafmax, int nw) {
  using G = Geo<D, NC>;
  __shared__ double s_all[BB][G::NCELL];
  double *s = s_all[lgrp<BS, BB>()];
  __shared__ double sh[BS / 64];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const DBox &bx = boxes[id - 1];
    const Sten<D, double> st = lpl_stencil<D, double>(bx, lambda2);
    const size_t base = (size_t)(id - 1) * G::NCELL;
    constexpr int NI = (G::NINT + BS - 1) / BS;
    double rr[NI];
    for (int t = 0; t < NI; t++) {
      const int q = ltid<BS>() + t * BS;
      if (q < G::NINT) {
        int i, j, k;
        interior_ijk<D, NC>(q, i, j, k);
        rr[t] = rhs[base + G::idx(i, j, k)];
      }
    }
    lds_stage<D, NC, BS, double, BB>(s, boxes, bx, id, phi, m, -1);
    double mx = 0;
    for (int t = 0; t < NI; t++) {
      const int q = ltid<BS>() + t * BS;
      if (q >= G::NINT) break;
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      const double lp = apply_cell<D, NC, double>(s, st, cyl != 0, bx, i, j, k);
      const double rv = rr[t] - lp;
      r32[(size_t)(id - 1) * Lay<D, NC, float>::STRIDE + Lay<D, NC, float>::in_at(i, j, k)] = (float)rv;
      mx = fmax(mx, fabs(rv));
    }
    if (bx.children[0] == 0) {
      mx = group_max<BS, BB>(mx, sh);
      if (ltid<BS>() == 0) atomic_max_nonneg(run_slot(leafmax, bx), mx);
    }
    gsync<BS, BB>();  // (the group's LDS is reused by its next item)
  }
}

// k_restrict_residual with the child box staged in LDS and its face ghosts
// formed in-kernel, so no ghost fill has to precede it.
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_restrict_residual_lds(const int *__restrict__ ids,
                                                                   const DBox *__restrict__ boxes, const T *child,
                                                                   T *phi, const T *__restrict__ rhs, T *tmp, T lambda2,
                                                                   int cyl, int store_residual, VarBc m, int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  constexpr int H = NC / 2;
  constexpr int NQ = (D == 3) ? H * H * H : H * H;
  __shared__ T s_all[BB][G::NCELL];
  T *s = s_all[lgrp<BS, BB>()];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const DBox &bc = boxes[id - 1];
    const DBox &bp = boxes[bc.parent - 1];
    const Sten<D, T> st = lpl_stencil<D, T>(bc, lambda2);
    const size_t cb = (size_t)(id - 1) * L::STRIDE, pb = (size_t)(bc.parent - 1) * L::STRIDE;
    lds_stage<D, NC, BS, T, BB>(s, boxes, bc, id, child, m, -1, (const T *)phi);
    int off[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) off[d] = ((bc.ix[d] - 1) & 1) * H;
    for (int q = ltid<BS>(); q < NQ; q += BS) {
      const int i = q % H + 1, j = (q / H) % H + 1, k = (D == 3) ? q / (H * H) + 1 : 0;
      const int i_f = 2 * i - 1, jf = 2 * j - 1, kf = (D == 3) ? 2 * k - 1 : 0;
      T rs[8];
      int cp[8];
      child_pos<D, NC, L>(i_f, jf, kf, cp);
      const int nch = (D == 3) ? 8 : 4;
      for (int c = 0; c < nch; c++) {
        const int ii = i_f + (c & 1), jj = jf + ((c >> 1) & 1), kk = kf + ((c >> 2) & 1);
        T lp = apply_cell<D, NC, T>(s, st, cyl != 0, bc, ii, jj, kk);
        rs[c] = rhs[cb + cp[c]] - lp;
        if (store_residual) tmp[cb + cp[c]] = rs[c];
      }
      const int ic = off[0] + i, jc = off[1] + j, kc = (D == 3) ? off[2] + k : 0;
      T rv;
      if (D == 2) {
        if (cyl) {
          T r = (T)bp.r_min[0] + (T(ic) - T(0.5)) * (T)bp.dr[0];
          T t = T(0.25) * (T)bp.dr[0] / r;
          T w1 = T(1) - t, w2 = T(1) + t;
          rv = T(0.25) * (w1 * (rs[0] + rs[2]) + w2 * (rs[1] + rs[3]));
        } else {
          T sm = rs[0];
          sm = sm + rs[1];
          sm = sm + rs[2];
          sm = sm + rs[3];
          rv = T(0.25) * sm;
        }
      } else {
        T sm = rs[0];
        for (int c = 1; c < 8; c++) sm = sm + rs[c];
        rv = T(0.125) * sm;
      }
      tmp[pb + L::in_at(ic, jc, kc)] = rv;
      phi[pb + L::in_at(ic, jc, kc)] = restrict_cell<D, NC, T>(s, i_f, jf, kf, false, bp, ic);
    }
    gsync<BS, BB>();  // (the group's LDS is reused by its next item)
  }
}

// Block smoother (single-precision inner cycle only): all 2 n_cycle
// half-sweeps of one smoothing phase on a box in LDS, in one launch.  The
// values of same-level neighbour boxes are taken from `in` at the start and
// held for the whole phase (the box couples to its neighbours as in a
// block-Jacobi sweep), which lets every workgroup read `in` while writing
// `out`; physical and refinement-boundary ghosts depend on the box's own
// cells and are formed again before every half-sweep.  Inside the box the
// half-sweeps are exactly those of k_gsrb.
template <int D, int NC, int BS, typename T>
__global__ void __launch_bounds__(BS) k_smooth_block(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                     const T *in, T *out, const T *__restrict__ rhs,
                                                     const T *coarse, int n_half, T lambda2, int cyl, VarBc m) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  __shared__ T s[G::NCELL];
  const int id = ids[blockIdx.x];
  if (id <= 0) return;  // run masked out
  const DBox &bx = boxes[id - 1];
  const Sten<D, T> st = lpl_stencil<D, T>(bx, lambda2);
  const size_t base = (size_t)(id - 1) * L::STRIDE;
  const T *r = rhs + base;
  lds_load_interior<D, NC, BS, T>(s, in + base);
  constexpr int NFC = 2 * D * G::NTAN;
  // same-level neighbour copies, once
  for (int q = threadIdx.x; q < NFC; q += BS) {
    const int nb = q / G::NTAN;
    if (bx.nb[nb] <= 0) continue;
    const int rr = q % G::NTAN;
    const int a = rr % NC + 1, b = (D == 3) ? rr / NC + 1 : 0;
    const bool low = (nb % 2) == 0;
    int p[3];
    face_pos<D>(nb, a, b, low ? 0 : NC + 1, p);
    s[G::idx(p[0], p[1], p[2])] = ghost_value<D, NC, T>(boxes, bx, in, m, nb, a, b, T(0), T(0));
  }
  constexpr int HALF = G::NINT / 2;
  constexpr int sx = 1, sy = G::NG, sz = G::NG * G::NG;
  for (int n = 1; n <= n_half; n++) {
    __syncthreads();
    // ghosts that depend on the box's own cells, for the colour updated now
    for (int q = threadIdx.x; q < NFC; q += BS) {
      const int nb = q / G::NTAN;
      if (bx.nb[nb] > 0) continue;
      const int rr = q % G::NTAN;
      const int a = rr % NC + 1, b = (D == 3) ? rr / NC + 1 : 0;
      const bool low = (nb % 2) == 0;
      int p[3];
      face_pos<D>(nb, a, b, low ? 1 : NC, p);
      if ((p[0] + p[1] + p[2] + n) & 1) continue;
      const T x1 = s[G::idx(p[0], p[1], p[2])];
      face_pos<D>(nb, a, b, low ? 2 : NC - 1, p);
      const T x2 = s[G::idx(p[0], p[1], p[2])];
      const T g = ghost_value<D, NC, T>(boxes, bx, in, m, nb, a, b, x1, x2, coarse);
      face_pos<D>(nb, a, b, low ? 0 : NC + 1, p);
      s[G::idx(p[0], p[1], p[2])] = g;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < HALF; q += BS) {
      const int mm = q % (NC / 2);
      const int j = (q / (NC / 2)) % NC + 1;
      const int k = (D == 3) ? q / ((NC / 2) * NC) + 1 : 0;
      const int i0 = 2 - ((n ^ (D == 3 ? k + j : j)) & 1);
      const int i = i0 + 2 * mm;
      const int x = G::idx(i, j, k);
      const T rx = r[L::in_at(i, j, k)];
      T v;
      if (D == 2) {
        if (cyl) {
          T c1, c2, c3;
          cyl_coeffs<T>(bx, i, st.c, c1, c2, c3);
          v = (rx - c2 * s[x - sx] - c3 * s[x + sx] - st.c[3] * s[x - sy] - st.c[4] * s[x + sy]) / c1;
        } else {
          v = (rx - st.c[1] * s[x - sx] - st.c[2] * s[x + sx] - st.c[3] * s[x - sy] - st.c[4] * s[x + sy]) /
              st.c[0];
        }
      } else {
        v = (rx - st.c[1] * s[x - sx] - st.c[2] * s[x + sx] - st.c[3] * s[x - sy] - st.c[4] * s[x + sy] -
             st.c[5] * s[x - sz] - st.c[6] * s[x + sz]) /
            st.c[0];
      }
      s[x] = v;
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < G::NINT; q += BS) {
    int i, j, k;
    interior_ijk<D, NC>(q, i, j, k);
    out[base + L::in_at(i, j, k)] = s[G::idx(i, j, k)];
  }
}

// phi += e on the interior of leaf boxes and e = 0 on the whole leaf box,
// ghost layer included: the refinement-boundary ghosts of finer boxes read
// a coarse leaf's ghost cells during the next cycle's first sweeps.  (Refined
// cells get the restriction of the finer phi instead, and e there is
// overwritten before the next cycle reads it.)
template <int D, int NC, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_mp_correct_leaves(const int *__restrict__ ids, double *phi, float *e,
                                                               int nw) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, float>;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    double *u = phi + (size_t)(id - 1) * G::NCELL;
    float *c = e + (size_t)(id - 1) * L::STRIDE;
    // interior: the pair of cells 2s, 2s + 1 (lexicographic), one of each
    // colour, both at s in their colour's block
    for (int s = ltid<BS>(); s < L::HALF; s += BS) {
      int i, j, k;
      interior_ijk<D, NC>(2 * s, i, j, k);
      const int y = G::idx(i, j, k), c0 = (i + j + k) & 1;
      const int e0 = c0 * L::HALF + s, e1 = (1 - c0) * L::HALF + s;
      u[y] = u[y] + (double)c[e0];
      u[y + 1] = u[y + 1] + (double)c[e1];
      c[e0] = 0.0f;
      c[e1] = 0.0f;
    }
    for (int x = G::NINT + ltid<BS>(); x < G::NCELL; x += BS) c[x] = 0.0f;  // the ghosts
  }
}

// af_gc_box with the box staged in LDS: face ghosts as lds_stage forms them,
// then (with corners) the edge and corner ghosts of gc_edge_cell /
// gc_corner_cell from the LDS values, then only the ghost layer is written.
// One workgroup per box.
template <int D, int NC, int BS, typename T, int BB>
__device__ inline void gc_lds_box(T *s, const DBox *__restrict__ boxes, int id, T *var, const VarBc &m,
                                  int corners) {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, T>;
  const DBox &bx = boxes[id - 1];
  T *u = var + (size_t)(id - 1) * L::STRIDE;
  lds_stage<D, NC, BS, T, BB>(s, boxes, bx, id, var, m, -1);
  if (corners) {
    if (D == 3) {
      for (int q = ltid<BS>(); q < 12 * NC; q += BS) {
        const int n = q / NC, qq = q % NC + 1;
        const int dim = c_edge_dim3[n];
        const int *dir = c_edge_dir3[n];
        int lo[3];
        for (int dd = 0; dd < 3; dd++) lo[dd] = (dir[dd] > 0) ? NC + 1 : 0;
        lo[dim] = qq;
        const int nb_id = bx.nbmat[nbmat_at<D>(dir[0], dir[1], dir[2])];
        T v;
        if (nb_id > 0) {
          const T *w = var + (size_t)(nb_id - 1) * L::STRIDE;
          v = w[L::in_at(lo[0] - dir[0] * NC, lo[1] - dir[1] * NC, lo[2] - dir[2] * NC)];
        } else {
          const int o0 = (dim + 1) % 3, o1 = (dim + 2) % 3;
          int di[3];
          for (int dd = 0; dd < 3; dd++) di[dd] = 1 - 2 * (lo[dd] & 1);
          di[dim] = 0;
          int ia[3] = {lo[0], lo[1], lo[2]}, ib[3] = {lo[0], lo[1], lo[2]};
          ia[o0] += di[o0];
          ib[o1] += di[o1];
          v = s[G::idx(ia[0], ia[1], ia[2])] + s[G::idx(ib[0], ib[1], ib[2])] -
              s[G::idx(lo[0] + di[0], lo[1] + di[1], lo[2] + di[2])];
        }
        s[G::idx(lo[0], lo[1], lo[2])] = v;
      }
      gsync<BS, BB>();
    }
    for (int n = ltid<BS>(); n < (1 << D); n += BS) {
      const int cx = n & 1, cy = (n >> 1) & 1, cz = (n >> 2) & 1;
      const int dnb[3] = {2 * cx - 1, 2 * cy - 1, (D == 3) ? 2 * cz - 1 : 0};
      const int lo[3] = {cx * (NC + 1), cy * (NC + 1), (D == 3) ? cz * (NC + 1) : 0};
      const int nb_id = bx.nbmat[nbmat_at<D>(dnb[0], dnb[1], dnb[2])];
      T v;
      if (nb_id > 0) {
        const T *w = var + (size_t)(nb_id - 1) * L::STRIDE;
        v = w[L::in_at(lo[0] - dnb[0] * NC, lo[1] - dnb[1] * NC, lo[2] - dnb[2] * NC)];
      } else if (D == 2) {
        const int di = 1 - 2 * (lo[0] & 1), dj = 1 - 2 * (lo[1] & 1);
        v = s[G::idx(lo[0] + di, lo[1], 0)] + s[G::idx(lo[0], lo[1] + dj, 0)] - s[G::idx(lo[0] + di, lo[1] + dj, 0)];
      } else {
        const int di = 1 - 2 * (lo[0] & 1), dj = 1 - 2 * (lo[1] & 1), dk = 1 - 2 * (lo[2] & 1);
        const int x = lo[0], y = lo[1], z = lo[2];
        v = s[G::idx(x, y + dj, z + dk)] + s[G::idx(x + di, y, z + dk)] + s[G::idx(x + di, y + dj, z)] -
            2 * s[G::idx(x + di, y + dj, z + dk)];
      }
      s[G::idx(lo[0], lo[1], lo[2])] = v;
    }
    gsync<BS, BB>();
  }
  // the ghost layer back: faces, and with corners the edges and corners
  for (int x = ltid<BS>(); x < G::NCELL; x += BS) {
    const int i = x % G::NG, j = (x / G::NG) % G::NG, k = (D == 3) ? x / (G::NG * G::NG) : 0;
    const int out = (i == 0 || i == NC + 1) + (j == 0 || j == NC + 1) + (D == 3 && (k == 0 || k == NC + 1));
    if (out == 1 || (corners && out > 1)) u[L::at(i, j, k)] = s[x];
  }
  gsync<BS, BB>();  // (the group's LDS is reused by its next item)
}
template <int D, int NC, int BS, typename T, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_gc_lds(const int *__restrict__ ids, const DBox *__restrict__ boxes, T *var,
                                                    VarBc m, int corners, int nw) {
  __shared__ T s_all[BB][Geo<D, NC>::NCELL];
  T *s = s_all[lgrp<BS, BB>()];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    gc_lds_box<D, NC, BS, T, BB>(s, boxes, id, var, m, corners);
  }
}

// The same for several FP64 variables in one launch: work item w is box
// w / vars.n, variable w % vars.n.
constexpr int GC_MAX_VARS = 8;
struct GcVars {
  double *v[GC_MAX_VARS];
  VarBc m[GC_MAX_VARS];
  int n;
};
template <int D, int NC, int BS>
__global__ void __launch_bounds__(BS) k_gc_lds_vars(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                    GcVars vars, int corners, int nw) {
  __shared__ double s[Geo<D, NC>::NCELL];
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int id = ids[w / vars.n];
    if (id <= 0) continue;
    gc_lds_box<D, NC, BS, double, 1>(s, boxes, id, vars.v[w % vars.n], vars.m[w % vars.n], corners);
  }
}

}  // namespace afg
