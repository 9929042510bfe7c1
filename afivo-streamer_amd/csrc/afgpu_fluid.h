// Drift-diffusion, source and field kernels:
//   flux_upwind_box + af_gc2_box + flux_upwind   (k_flux)
//   af_consistent_fluxes                          (k_consistent_fluxes)
//   flux_update_densities + add_source_terms      (k_update)
//   field_set_rhs, mg_compute_phi_gradient, photoi source/accumulation.
#pragma once

#include <limits>

#include "afgpu_chem.h"
#include "afgpu_mg.h"

namespace afg {

// Two-layer halo value of n_e for face nb, layer (0 = adjacent, 1 = outer),
// tangential (a, b): af_gc2_box (afivo/src/m_af_ghostcell.f90:672-744) with
// gc2_prolong_rb (:753-856) and bc_to_gc2 (:282-375).
template <int D, int NC>
__device__ inline double gc2_value(const DBox *__restrict__ boxes, int id, const double *var, const VarBc &m, int nb,
                                   int layer, int a, int b) {
  using G = Geo<D, NC>;
  constexpr int H = NC / 2;
  const DBox &bx = boxes[id - 1];
  const double *u = var + (size_t)(id - 1) * G::NCELL;
  const int d = nb / 2;
  const bool low = (nb % 2) == 0;
  int t0, t1;
  tangential(D, d, t0, t1);
  int p[3] = {0, 0, 0};
  p[t0] = a;
  if (D == 3) p[t1] = b;
  const int nb_id = bx.nb[nb];
  if (nb_id > 0) {
    const double *w = var + (size_t)(nb_id - 1) * G::NCELL;
    p[d] = low ? NC - layer : 1 + layer;
    return w[G::idx(p[0], p[1], p[2])];
  }
  if (nb_id < 0) {
    double c0, c1, c2;
    if (m.type[nb] == BC_DIRICHLET) {
      c0 = 2; c1 = -1; c2 = c0;
    } else if (m.type[nb] == BC_NEUMANN) {
      c0 = bx.dr[d] * (low ? -1 : 1); c1 = 1; c2 = 3 * c0;
    } else {
      c0 = 1; c1 = 0; c2 = c0;
    }
    if (D == 3 && nb == 3) c2 = c0;  // the reference's high-y branch (:361)
    const double bv = bc_value(m, nb, bx.run);
    p[d] = low ? 1 + layer : NC - layer;
    const double x = u[G::idx(p[0], p[1], p[2])];
    return (layer == 0) ? c0 * bv + c1 * x : c2 * bv + c1 * x;
  }
  // gc2_prolong_rb: limited linear prolongation from the parent's neighbour
  const double *P = var + (size_t)(bx.pnb[nb] - 1) * G::NCELL;
  int off[3] = {0, 0, 0};
  for (int dd = 0; dd < D; dd++) off[dd] = ((bx.ix[dd] - 1) & 1) * H;
  // fine index along every dimension, and the coarse cell containing it
  int fine[3] = {p[0], p[1], p[2]};
  fine[d] = low ? -1 + layer : NC + 1 + layer;
  int cc[3] = {0, 0, 0};
  bool lower[3] = {false, false, false};
  for (int dd = 0; dd < D; dd++) {
    if (dd == d) {
      // normal direction: the low face's fine cells are -1 (lower) and 0,
      // the high face's are nc+1 (lower) and nc+2
      cc[dd] = low ? NC : 1;
      lower[dd] = low ? (layer == 1) : (layer == 0);
    } else {
      cc[dd] = off[dd] + ((fine[dd] + 1) >> 1);
      lower[dd] = (fine[dd] & 1) != 0;
    }
  }
  const int c = G::idx(cc[0], cc[1], cc[2]);
  constexpr int st[3] = {1, G::NG, G::NG * G::NG};
  const double f0 = P[c];
  double v = f0;
  for (int dd = 0; dd < D; dd++) {
    const double sa = P[c] - P[c - st[dd]];
    const double sb = P[c + st[dd]] - P[c];
    const double fd = 0.25 * lim_apply<double>(sa, sb, m.limiter);
    v = lower[dd] ? v - fd : v + fd;
  }
  return v;
}

// One leaf box: n_e halo, face fluxes, CFL sum and conductivity maxima.
// Launched per refinement level, lowest level first (the reference's level
// loop makes a coarse leaf's written-back ghost layer visible to the finer
// level's halo, m_af_flux_schemes.f90:696-707).
// TF: arithmetic type of the stencil (double: the reference's arithmetic;
// float: the mixed-precision form's FP32 flux, accumulated in FP64 by
// k_update).  The n_e halo is formed in FP64 and its first layer written
// back at FP64 precision in either case.
template <int D, int NC, int BS, typename TF, int BB = 1>
__global__ void __launch_bounds__(BS * BB)
    k_flux(const int *__restrict__ ids, const DBox *__restrict__ boxes, double *ne, const double *__restrict__ efield,
           const double *__restrict__ efc, double *flux, VarBc m, DTable td, double N_inv_d,
           unsigned long long *cfl_max, unsigned long long *sig_max, int nw) {
  using G = Geo<D, NC>;
  constexpr int NT = G::NTAN;
  __shared__ TF halo_all[BB][2 * D][2][NT];
  TF(&halo)[2 * D][2][NT] = halo_all[lgrp<BS, BB>()];
  // the mobility and diffusion columns of the transport table (3D, single
  // precision, up to TABMAX rows), converted as lt_at_t converts them
  constexpr int TABMAX = (D == 3 && sizeof(TF) < sizeof(double)) ? 1024 : 1;
  __shared__ TF tab_all[BB][2 * TABMAX];
  const bool tab_lds = TABMAX > 1 && td.n <= TABMAX;
  if (tab_lds)  // (read after the first item's halo barrier)
    for (int q = ltid<BS>(); q < 2 * td.n; q += BS) tab_all[lgrp<BS, BB>()][(q / td.n) * TABMAX + q % td.n] = (TF)td.cols[q];
  __shared__ double sh[BS / 64];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const DBox &bx = boxes[id - 1];
    const size_t base = (size_t)(id - 1) * G::NCELL;
    const size_t fbase = (size_t)(id - 1) * D * G::NFACE;
    const TF N_inv = (TF)N_inv_d;

    // the halo; its first layer also goes back into the box
    // (m_af_ghostcell.f90:740-742): no thread of any box reads a ghost cell
    // of this level during the launch
    for (int q = ltid<BS>(); q < 2 * D * 2 * NT; q += BS) {
      const int nb = q / (2 * NT), layer = (q / NT) % 2, r = q % NT;
      const int a = r % NC + 1, b = (D == 3) ? r / NC + 1 : 0;
      const double hv = gc2_value<D, NC>(boxes, id, ne, m, nb, layer, a, b);
      halo[nb][layer][r] = (TF)hv;
      if (layer == 0) {
        const int d = nb / 2;
        int t0, t1;
        tangential(D, d, t0, t1);
        int p[3] = {0, 0, 0};
        p[t0] = a;
        if (D == 3) p[t1] = b;
        p[d] = (nb % 2) ? NC + 1 : 0;
        ne[base + G::idx(p[0], p[1], p[2])] = hv;
      }
    }
    gsync<BS, BB>();

    auto line = [&](int d, int r, const int *p, int q) -> TF {
      if (q >= 1 && q <= NC) {
        int pp[3] = {p[0], p[1], p[2]};
        pp[d] = q;
        return (TF)ne[base + G::idx(pp[0], pp[1], pp[2])];
      }
      if (q <= 0) return halo[2 * d][-q][r];
      return halo[2 * d + 1][q - NC - 1][r];
    };

    TF smax = -std::numeric_limits<TF>::max();
    TF cmax = 0;
    if constexpr (D == 3) {
      const TF *tab = tab_lds ? tab_all[lgrp<BS, BB>()] : nullptr;
      auto coef = [&](int c, int ix, TF w) {
        return tab_lds ? w * tab[c * TABMAX + ix - 1] + (1 - w) * tab[c * TABMAX + ix] : lt_at_t<TF>(td, c, ix, w);
      };
      // one face: the upwind value, velocity, diffusion coefficient, flux
      auto face = [&](TF Lm2, TF Lm1, TF L0, TF Lp1, TF ex, TF e_lo, TF e_hi, TF inv_dx, TF &v, TF &dc) -> TF {
        TF u;
        if (-1 * ex > 0) {
          u = Lm1 + TF(0.5) * lim_koren<TF>(L0 - Lm1, Lm1 - Lm2);
        } else {
          u = L0 - TF(0.5) * lim_koren<TF>(L0 - Lm1, Lp1 - L0);
        }
        const TF tmp_fc = TF(0.5) * (e_lo + e_hi) * (TF)SI_to_Townsend * N_inv;
        int lix;
        TF lw;
        lt_loc_t<TF>(td, tmp_fc, lix, lw);
        const TF mu = coef(0, lix, lw) * N_inv;
        dc = coef(1, lix, lw) * N_inv;
        v = -mu * ex;
        smax = fmax(smax, mu * u);
        return v * u - dc * inv_dx * (L0 - Lm1);
      };
      // The cell's CFL term of one direction from its lower and upper faces
      // (m_fluid.f90:194-195, m_af_flux_schemes.f90:779-848); the terms are
      // summed over the directions in the order x, y, z.
      auto cfl_term = [](TF av_hi, TF av_lo, TF dc_hi, TF dc_lo, TF inv_dx) -> TF {
        return TF(1) * fmax(av_hi, av_lo) * inv_dx + TF(2) * fmax(dc_hi, dc_lo) * (inv_dx * inv_dx);
      };
      if constexpr (BS == NC * NC) {
        // Face fluxes (flux_upwind, src/m_fluid.f90:102-209), one thread per
        // grid column (i, j), walking up the box in z.  Plane k is staged in
        // LDS (n_e with its two halo layers in x and y, |E| with its ghost
        // ring), so the four upwind values and two field values of an x or
        // y face come from LDS; the column's own z neighbours stay in
        // registers.  Every face is formed once: the thread forms the lower
        // x, y and z faces of its cell, the first wavefront the box's upper
        // x and y faces of the plane, and every thread its column's upper z
        // face at the end.  A cell takes its upper faces' |v| and D for the
        // CFL term from the threads that formed them (LDS), or, in z, from
        // the next plane.  Global loads and stores are rows of cells or
        // faces in x, the contiguous direction.
        static_assert(BB == 1 && BS >= 12 * NC, "one box of NC x NC columns per workgroup");
        constexpr int NP = NC + 4, NE = NC + 2;
        __shared__ TF pn[NP * NP];            // n_e of the plane: (i, j) at (j + 1) NP + i + 1, i, j = -1 .. NC + 2
        __shared__ TF pe[NE * NE];            // |E| of the plane: (i, j) at j NE + i, i, j = 0 .. NC + 1
        __shared__ TF fv[2][BS + NC], fd[2][BS + NC];  // |v| and D of the plane's x and y faces (lower, upper boundary)
        constexpr int st2 = G::NG * G::NG;
        const int t = ltid<BS>();
        const int i = t % NC + 1, j = t / NC + 1;
        const TF inv_dx[3] = {TF(1) / (TF)bx.dr[0], TF(1) / (TF)bx.dr[1], TF(1) / (TF)bx.dr[2]};
        const int rz = (i - 1) + NC * (j - 1);
        const size_t col = base + G::idx(i, j, 0);
        // the column's n_e at k - 2 .. k + 1 and |E| at k - 1, k
        TF n0 = halo[4][1][rz], n1 = halo[4][0][rz], n2 = (TF)ne[col + st2], n3 = (TF)ne[col + 2 * st2];
        TF em = (TF)efield[col], e0 = (TF)efield[col + st2];
        TF pend = 0, avz_prev = 0, dz_prev = 0;  // cell k - 1: its x + y CFL sum, its lower z face
        // loads of plane k + 1 are made on plane k (the faces' field and the
        // |E| ghost ring), so no plane waits for memory
        const int rs = (t - 8 * NC) / NC, ra = (t - 8 * NC) % NC + 1;  // |E| ring: side, position
        const int rq = (rs & 1) ? NC + 1 : 0;
        const int ri = (rs < 2) ? rq : ra, rj = (rs < 2) ? ra : rq;
        const bool ring = t >= 8 * NC && t < 12 * NC;
        TF exn = (TF)efc[fbase + G::fidx(i, j, 1, 0)], eyn = (TF)efc[fbase + G::fidx(i, j, 1, 1)],
           ezn = (TF)efc[fbase + G::fidx(i, j, 1, 2)];
        TF ern = ring ? (TF)efield[base + G::idx(ri, rj, 1)] : TF(0);
#pragma unroll 1
        for (int k = 1; k <= NC; k++) {
          const size_t c0 = col + (size_t)k * st2;
          const TF ex = exn, ey = eyn, ez = ezn, er = ern;
          if (k < NC) {
            exn = (TF)efc[fbase + G::fidx(i, j, k + 1, 0)];
            eyn = (TF)efc[fbase + G::fidx(i, j, k + 1, 1)];
            ezn = (TF)efc[fbase + G::fidx(i, j, k + 1, 2)];
            if (ring) ern = (TF)efield[base + G::idx(ri, rj, k + 1)];
          }
          const TF n4 = (k + 2 <= NC) ? (TF)ne[c0 + 2 * st2] : halo[5][k + 1 - NC][rz];
          const TF ep = (TF)efield[c0 + st2];
          pn[(j + 1) * NP + i + 1] = n2;
          pe[j * NE + i] = e0;
          if (t < 8 * NC) {
            // halo of n_e: s = direction (x, y) x side (low, high) x layer
            const int s = t / NC, a = t % NC + 1;
            const int layer = s & 1, side = (s >> 1) & 1, dd = s >> 2;
            const TF hv = halo[2 * dd + side][layer][(a - 1) + NC * (k - 1)];
            const int q = side ? NC + 1 + layer : -layer;
            if (dd == 0)
              pn[(a + 1) * NP + q + 1] = hv;
            else
              pn[(q + 1) * NP + a + 1] = hv;
          } else if (ring) {
            // ghost ring of |E|: x low, x high, y low, y high
            pe[rj * NE + ri] = er;
          }
          gsync<BS, BB>();
          TF vx, dx, vy, dy;
          {
            const int o = (j + 1) * NP + i + 1;
            flux[fbase + G::fidx(i, j, k, 0)] = (double)face(pn[o - 2], pn[o - 1], pn[o], pn[o + 1], ex,
                                                              pe[j * NE + i - 1], pe[j * NE + i], inv_dx[0], vx, dx);
            flux[fbase + G::fidx(i, j, k, 1)] =
                (double)face(pn[o - 2 * NP], pn[o - NP], pn[o], pn[o + NP], ey, pe[(j - 1) * NE + i], pe[j * NE + i],
                             inv_dx[1], vy, dy);
          }
          fv[0][t] = fabs(vx);
          fd[0][t] = dx;
          fv[1][t] = fabs(vy);
          fd[1][t] = dy;
          if (t < 2 * NC) {
            // the box's upper x face of row j = a (dd = 0) or y face of column i = a (dd = 1)
            const int dd = t / NC, a = t % NC + 1;
            const int ii = dd ? a : NC + 1, jj = dd ? NC + 1 : a;
            const int o = (jj + 1) * NP + ii + 1, sp = dd ? NP : 1, se = dd ? NE : 1;
            TF vb, db;
            flux[fbase + G::fidx(ii, jj, k, dd)] =
                (double)face(pn[o - 2 * sp], pn[o - sp], pn[o], pn[o + sp], (TF)efc[fbase + G::fidx(ii, jj, k, dd)],
                             pe[jj * NE + ii - se], pe[jj * NE + ii], inv_dx[dd], vb, db);
            fv[dd][BS + a - 1] = fabs(vb);
            fd[dd][BS + a - 1] = db;
          }
          gsync<BS, BB>();
          const int hx = (i < NC) ? t + 1 : BS + j - 1, hy = (j < NC) ? t + NC : BS + i - 1;
          TF c = 0;
          c = c + cfl_term(fv[0][hx], fabs(vx), fd[0][hx], dx, inv_dx[0]);
          c = c + cfl_term(fv[1][hy], fabs(vy), fd[1][hy], dy, inv_dx[1]);
          // z face k: the lower face of cell k, the upper one of cell k - 1
          TF vz, dz;
          flux[fbase + G::fidx(i, j, k, 2)] = (double)face(n0, n1, n2, n3, ez, em, e0, inv_dx[2], vz, dz);
          if (k > 1) cmax = fmax(cmax, pend + cfl_term(fabs(vz), avz_prev, dz, dz_prev, inv_dx[2]));
          pend = c;
          avz_prev = fabs(vz);
          dz_prev = dz;
          n0 = n1;
          n1 = n2;
          n2 = n3;
          n3 = n4;
          em = e0;
          e0 = ep;
        }
        // the box's upper z face and the top cell's CFL sum
        TF vz, dz;
        flux[fbase + G::fidx(i, j, NC + 1, 2)] =
            (double)face(n0, n1, n2, n3, (TF)efc[fbase + G::fidx(i, j, NC + 1, 2)], em, e0, inv_dx[2], vz, dz);
        cmax = fmax(cmax, pend + cfl_term(fabs(vz), avz_prev, dz, dz_prev, inv_dx[2]));
      } else {
        // Boxes with fewer columns than threads (8^3): face fluxes
        // (flux_upwind, src/m_fluid.f90:102-209) cell by cell, x fastest
        // over the threads, so that the loads of every direction
        // are coalesced and a cell's n_e and field are read while still in
        // cache by all three directions.  For each direction a cell forms its
        // lower and its upper face: an interior face is formed by both of its
        // cells, from the same values, and written by the cell above it; the
        // box's last face by the cell below it.  The two faces' velocities and
        // diffusion coefficients give the cell's CFL term (m_fluid.f90:194-195,
        // m_af_flux_schemes.f90:779-848), summed over the directions in the
        // order x, y, z.
        for (int q = ltid<BS>(); q < G::NINT; q += BS) {
          const int p[3] = {q % NC + 1, (q / NC) % NC + 1, q / (NC * NC) + 1};
          TF cfl = 0;
          for (int d = 0; d < D; d++) {
            const TF inv_dx = TF(1) / (TF)bx.dr[d];
            int t0, t1;
            tangential(D, d, t0, t1);
            const int r = (p[t0] - 1) + NC * (p[t1] - 1);  // the cell's position on a face of direction d
            constexpr int st[3] = {1, G::NG, G::NG * G::NG};
            const int cd = p[d];
            const size_t c0 = base + G::idx(p[0], p[1], p[2]);
            TF nl[5];  // n_e at cd - 2 .. cd + 2 along d
            for (int s2 = -2; s2 <= 2; s2++) {
              const int pos = cd + s2;
              nl[s2 + 2] = (pos >= 1 && pos <= NC) ? (TF)ne[c0 + s2 * st[d]]
                           : (pos <= 0)              ? halo[2 * d][-pos][r]
                                                     : halo[2 * d + 1][pos - NC - 1][r];
            }
            const TF e_m = (TF)efield[c0 - st[d]], e_0 = (TF)efield[c0], e_p = (TF)efield[c0 + st[d]];
            int fp[3] = {p[0], p[1], p[2]};
            const size_t f_lo = fbase + G::fidx(fp[0], fp[1], fp[2], d);
            fp[d] = cd + 1;
            const size_t f_hi = fbase + G::fidx(fp[0], fp[1], fp[2], d);
            TF v_lo, dc_lo, v_hi, dc_hi;
            const TF fl_lo = face(nl[0], nl[1], nl[2], nl[3], (TF)efc[f_lo], e_m, e_0, inv_dx, v_lo, dc_lo);
            const TF fl_hi = face(nl[1], nl[2], nl[3], nl[4], (TF)efc[f_hi], e_0, e_p, inv_dx, v_hi, dc_hi);
            flux[f_lo] = (double)fl_lo;
            if (cd == NC) flux[f_hi] = (double)fl_hi;
            cfl = cfl + cfl_term(fabs(v_hi), fabs(v_lo), dc_hi, dc_lo, inv_dx);
          }
          cmax = fmax(cmax, cfl);
        }
      }
      double smax_d = group_max<BS, BB>((double)smax, sh);
      if (ltid<BS>() == 0) atomic_max_nonneg(run_slot(sig_max, bx), fmax(smax_d, 1e-100));
    } else {
      // 2D boxes have too few lines per direction for a thread each: one
      // face per thread (flux_upwind, src/m_fluid.f90:102-209), then the
      // CFL sum per cell from its faces
      for (int q = ltid<BS>(); q < D * (NC + 1) * NT; q += BS) {
        const int d = q / ((NC + 1) * NT);
        // x faces: consecutive threads take consecutive faces of one grid row
        // (contiguous in memory); y/z faces: consecutive rows of one face plane
        const int qd = q % ((NC + 1) * NT);
        const int f = (d == 0) ? qd % (NC + 1) + 1 : qd / NT + 1;
        const int r = (d == 0) ? qd / (NC + 1) : qd % NT;
        int t0, t1;
        tangential(D, d, t0, t1);
        int p[3] = {0, 0, 0};
        p[t0] = r % NC + 1;
        if (D == 3) p[t1] = r / NC + 1;
        const TF inv_dx = TF(1) / (TF)bx.dr[d];
        int fp[3] = {p[0], p[1], (D == 3) ? p[2] : 1};
        fp[d] = f;
        const TF ex = (TF)efc[fbase + G::fidx(fp[0], fp[1], fp[2], d)];
        const TF Lm2 = line(d, r, p, f - 2), Lm1 = line(d, r, p, f - 1), L0 = line(d, r, p, f),
                 Lp1 = line(d, r, p, f + 1);
        TF u;
        if (-1 * ex > 0) {
          u = Lm1 + TF(0.5) * lim_koren<TF>(L0 - Lm1, Lm1 - Lm2);
        } else {
          u = L0 - TF(0.5) * lim_koren<TF>(L0 - Lm1, Lp1 - L0);
        }
        int e0[3] = {p[0], p[1], p[2]}, e1[3] = {p[0], p[1], p[2]};
        e0[d] = f - 1;
        e1[d] = f;
        const TF tmp_fc = TF(0.5) * ((TF)efield[base + G::idx(e0[0], e0[1], e0[2])] +
                                     (TF)efield[base + G::idx(e1[0], e1[1], e1[2])]) *
                          (TF)SI_to_Townsend * N_inv;
        int lix;
        TF lw;
        lt_loc_t<TF>(td, tmp_fc, lix, lw);
        const TF mu = lt_at_t<TF>(td, 0, lix, lw) * N_inv;
        const TF dc = lt_at_t<TF>(td, 1, lix, lw) * N_inv;
        const TF v = -mu * ex;
        flux[fbase + G::fidx(fp[0], fp[1], fp[2], d)] = (double)(v * u - dc * inv_dx * (L0 - Lm1));
        smax = fmax(smax, mu * u);
      }
      double smax_d = group_max<BS, BB>((double)smax, sh);
      if (ltid<BS>() == 0) atomic_max_nonneg(run_slot(sig_max, bx), fmax(smax_d, 1e-100));

      // CFL sum per cell (m_fluid.f90:194-195 summed over dimensions,
      // m_af_flux_schemes.f90:779-848)
      for (int q = ltid<BS>(); q < G::NINT; q += BS) {
        int i, j, k;
        interior_ijk<D, NC>(q, i, j, k);
        TF cfl = 0;
        for (int d = 0; d < D; d++) {
          const TF inv_dx = TF(1) / (TF)bx.dr[d];
          TF vv[2], dd[2];
          for (int s = 0; s < 2; s++) {
            int c[3] = {i, j, k};
            int fp[3] = {i, j, (D == 3) ? k : 1};
            fp[d] += s;  // faces c and c+1
            const TF ex = (TF)efc[fbase + G::fidx(fp[0], fp[1], fp[2], d)];
            int e0[3] = {c[0], c[1], c[2]};
            e0[d] = c[d] - 1 + s;
            int e1[3] = {c[0], c[1], c[2]};
            e1[d] = c[d] + s;
            const TF tmp_fc = TF(0.5) * ((TF)efield[base + G::idx(e0[0], e0[1], e0[2])] +
                                         (TF)efield[base + G::idx(e1[0], e1[1], e1[2])]) *
                              (TF)SI_to_Townsend * N_inv;
            int lix;
            TF lw;
            lt_loc_t<TF>(td, tmp_fc, lix, lw);
            const TF mu = lt_at_t<TF>(td, 0, lix, lw) * N_inv;
            dd[s] = lt_at_t<TF>(td, 1, lix, lw) * N_inv;
            vv[s] = -mu * ex;
          }
          const TF cl = TF(1) * fmax(fabs(vv[1]), fabs(vv[0])) * inv_dx + TF(2) * fmax(dd[1], dd[0]) * (inv_dx * inv_dx);
          cfl = cfl + cl;
        }
        cmax = fmax(cmax, cfl);
      }
    }
    const double cmax_d = group_max<BS, BB>((double)cmax, sh);
    if (ltid<BS>() == 0) atomic_max_nonneg(run_slot(cfl_max, bx), cmax_d);
    gsync<BS, BB>();  // (the group's halo is reused by its next item)
  }
}

// Coarse-leaf faces next to a refined neighbour get the (cylindrically
// weighted) average of the fine fluxes: af_consistent_fluxes +
// flux_from_children (afivo/src/m_af_core.f90:1257-1402).
template <int D, int NC, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_consistent_fluxes(const int *__restrict__ ids,
                                                               const DBox *__restrict__ boxes, double *flux, int cyl,
                                                               int nw) {
  using G = Geo<D, NC>;
  constexpr int H = NC / 2;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const DBox &bx = boxes[id - 1];
    for (int q = ltid<BS>(); q < 2 * D * G::NTAN; q += BS) {
      const int nb = q / G::NTAN, r = q % G::NTAN;
      const int nb_id = bx.nb[nb];
      if (nb_id <= 0 || boxes[nb_id - 1].children[0] > 0) continue;
      const int d = nb / 2;
      const bool low = (nb % 2) == 0;
      int t0, t1;
      tangential(D, d, t0, t1);
      const int a = r % NC + 1, b = (D == 3) ? r / NC + 1 : 0;
      int cd[3] = {0, 0, 0};
      cd[d] = low ? 0 : 1;
      cd[t0] = (a > H) ? 1 : 0;
      if (D == 3) cd[t1] = (b > H) ? 1 : 0;
      const int i_ch = cd[0] + 2 * cd[1] + 4 * cd[2];
      const int c_id = bx.children[i_ch];
      const double *fc = flux + (size_t)(c_id - 1) * D * G::NFACE;
      double *fn = flux + (size_t)(nb_id - 1) * D * G::NFACE;
      const int fi = low ? 1 : NC + 1, fi_nb = low ? NC + 1 : 1;
      const int pa = a - cd[t0] * H;  // index on the child's half
      if (D == 2) {
        int pn[2], s0[2], s1[2];
        pn[d] = fi_nb;
        pn[t0] = a;
        s0[d] = s1[d] = fi;
        s0[t0] = 2 * pa - 1;
        s1[t0] = 2 * pa;
        const double A = fc[G::fidx(s0[0], s0[1], 1, d)], B = fc[G::fidx(s1[0], s1[1], 1, d)];
        double val;
        if (d == 1 && cyl) {
          const DBox &bn = boxes[nb_id - 1];
          double rr = bn.r_min[0] + (a - 0.5) * bn.dr[0];
          double tmp = 0.25 * bn.dr[0] / rr;
          double w1 = 1 - tmp, w2 = 1 + tmp;
          val = 0.5 * (w1 * A + w2 * B);
        } else {
          val = 0.5 * (A + B);
        }
        fn[G::fidx(pn[0], pn[1], 1, d)] = val;
      } else {
        const int pb = b - cd[t1] * H;
        int pn[3], s00[3], s10[3], s01[3], s11[3];
        pn[d] = fi_nb;
        pn[t0] = a;
        pn[t1] = b;
        s00[d] = s10[d] = s01[d] = s11[d] = fi;
        s00[t0] = 2 * pa - 1; s00[t1] = 2 * pb - 1;
        s10[t0] = 2 * pa;     s10[t1] = 2 * pb - 1;
        s01[t0] = 2 * pa - 1; s01[t1] = 2 * pb;
        s11[t0] = 2 * pa;     s11[t1] = 2 * pb;
        fn[G::fidx(pn[0], pn[1], pn[2], d)] =
            0.25 * (fc[G::fidx(s00[0], s00[1], s00[2], d)] + fc[G::fidx(s10[0], s10[1], s10[2], d)] +
                    fc[G::fidx(s01[0], s01[1], s01[2], d)] + fc[G::fidx(s11[0], s11[1], s11[2], d)]);
      }
    }
  }
}

struct UpdateArgs {
  double *sp[3][3];   // species k, state s
  const double *efield;
  const double *photo;
  const double *flux;
  DTable chem;
  double dt, N_inv;
  // forest of runs: run r advances by dt_scale * dt_run[r] (dt unused)
  const double *dt_run;
  double dt_scale;
  int s_deriv, s_out, n_prev;
  int s_prev[4];
  double w_prev[4];
  int last_step;
  int chem_nmin_on;
  double chem_nmin;
  int chem_limit_loss;
  int photoi_enabled, photoi_species;
  int cyl;
  // a reaction list over the three species (afgpu_set_chemistry), else the
  // standard model from the chem table
  int use_net;
  ChemArgs net;
};

// get_rates and get_derivatives of a reaction list at one cell, for the
// three species of the pools (m_chemistry.f90:672-804)
__device__ inline void net_derivs(const ChemArgs &a, double field, const double *dens, double *dd) {
  int ix;
  double w, te = 0;
  lt_loc(a.fld, field, ix, w);
  if (a.need_te) {
    int ie;
    double we;
    lt_loc(a.energy, field, ie, we);
    te = CHEM_EV_TO_K * lt_at(a.energy, 0, ie, we);
  }
  for (int r = 0; r < a.n_reactions; r++) {
    const afgpu_reaction_t &R = a.r[r];
    double k = chem_rate(R, a, field, te, ix, w);
    double prod = 1.0;
    for (int q = 0; q < R.n_in; q++) {
      const int s = R.ix_in[q];
      prod = prod * (s == 0 ? dens[0] : s == 1 ? dens[1] : dens[2]);
    }
    k = k * prod;
    for (int q = 0; q < R.n_in; q++) {
      const int s = R.ix_in[q];
      if (s == 0) dd[0] = dd[0] - k;
      else if (s == 1) dd[1] = dd[1] - k;
      else dd[2] = dd[2] - k;
    }
    for (int q = 0; q < R.n_out; q++) {
      const int s = R.ix_out[q];
      const double v = k * R.multiplicity_out[q];
      if (s == 0) dd[0] = dd[0] + v;
      else if (s == 1) dd[1] = dd[1] + v;
      else dd[2] = dd[2] + v;
    }
  }
}

// flux_update_densities (afivo/src/m_af_flux_schemes.f90:320-436) with
// add_source_terms (src/m_fluid.f90:298-466, src/m_chemistry.f90:672-804)
// over all leaves: the standard model, or a reaction list over the three
// species (A.use_net).
// NET: the reaction list (A.use_net) compiled in; the standard model's
// instantiation holds none of its code or registers.
template <int D, int NC, int BS, int BB = 1, bool NET = false>
__global__ void __launch_bounds__(BS * BB) k_update(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                    UpdateArgs A, unsigned long long *chem_min, int nw) {
  using G = Geo<D, NC>;
  __shared__ double sh[BS / 64];
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const DBox &bx = boxes[id - 1];
    const size_t base = (size_t)(id - 1) * G::NCELL;
    const size_t fbase = (size_t)(id - 1) * D * G::NFACE;
    const double dt = A.dt_run ? A.dt_scale * A.dt_run[bx.run] : A.dt;
    double dt_dr[3];
    for (int d = 0; d < D; d++) dt_dr[d] = dt / bx.dr[d];
    const double eps = 1e-100;
    double bmin = 1e100;
    for (int q = ltid<BS>(); q < G::NINT; q += BS) {
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      const size_t x = base + G::idx(i, j, k);
      double out[3];
      for (int sp = 0; sp < 3; sp++) {
        double tmp = 0.0;
        for (int m = 0; m < A.n_prev; m++) tmp = tmp + A.w_prev[m] * A.sp[sp][A.s_prev[m]][x];
        out[sp] = tmp;
      }
      double dens[3];
      for (int sp = 0; sp < 3; sp++) {
        const double cur = (A.s_deriv == A.s_out) ? out[sp] : A.sp[sp][A.s_deriv][x];
        dens[sp] = fmax(cur, 0.0);
      }
      const double field = SI_to_Townsend * A.N_inv * A.efield[x];
      double dd[3] = {0.0, 0.0, 0.0};
      if constexpr (NET) {
        net_derivs(A.net, field, dens, dd);
      } else {
        int lix;
        double lw;
        lt_loc(A.chem, field, lix, lw);
        const double k_ion = 1.0 * lt_at(A.chem, 0, lix, lw);
        const double k_att = 1.0 * lt_at(A.chem, 1, lix, lw);
        const double r1 = k_ion * dens[0];
        const double r2 = k_att * dens[0];
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
      if (A.last_step) {
        for (int sp = 0; sp < 3; sp++) {
          double t;
          if (A.chem_nmin_on)
            t = (dens[sp] + A.chem_nmin) / fmax(fabs(dd[sp]), eps);
          else if (A.chem_limit_loss)
            t = fmax(dens[sp], eps) / fmax(-dd[sp], eps);
          else
            t = 1e100;
          bmin = fmin(bmin, t);
        }
      }
      if (A.photoi_enabled) {
        const double ph = A.photo[x];
        dd[0] = dd[0] + ph;
        dd[A.photoi_species] = dd[A.photoi_species] + ph;
      }
      for (int sp = 0; sp < 3; sp++) out[sp] = out[sp] + dt * dd[sp];
      // conservative flux divergence of the electrons
      const double *fl = A.flux + fbase;
      double ne = out[0];
      if (D == 2) {
        if (A.cyl) {
          const double dr = bx.dr[0];
          const double r = bx.r_min[0] + (i - 0.5) * dr;
          const double inv_r = 1 / r;
          const double rf1 = (r - 0.5 * dr) * inv_r, rf2 = (r + 0.5 * dr) * inv_r;
          ne = ne + dt_dr[0] * (rf1 * fl[G::fidx(i, j, 1, 0)] - rf2 * fl[G::fidx(i + 1, j, 1, 0)]) +
               dt_dr[1] * (fl[G::fidx(i, j, 1, 1)] - fl[G::fidx(i, j + 1, 1, 1)]);
        } else {
          ne = ne + dt_dr[0] * (fl[G::fidx(i, j, 1, 0)] - fl[G::fidx(i + 1, j, 1, 0)]) +
               dt_dr[1] * (fl[G::fidx(i, j, 1, 1)] - fl[G::fidx(i, j + 1, 1, 1)]);
        }
      } else {
        ne = ne + dt_dr[0] * (fl[G::fidx(i, j, k, 0)] - fl[G::fidx(i + 1, j, k, 0)]) +
             dt_dr[1] * (fl[G::fidx(i, j, k, 1)] - fl[G::fidx(i, j + 1, k, 1)]) +
             dt_dr[2] * (fl[G::fidx(i, j, k, 2)] - fl[G::fidx(i, j, k + 1, 2)]);
      }
      out[0] = ne;
      for (int sp = 0; sp < 3; sp++) A.sp[sp][A.s_out][x] = out[sp];
    }
    if (A.last_step) {
      bmin = group_min<BS, BB>(bmin, sh);
      if (ltid<BS>() == 0) atomic_min_nonneg(run_slot(chem_min, bx), bmin);
    }
  }
}

// field_set_rhs (src/m_field.f90:406-444) on leaves, all cells.
template <int NCELL, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_set_rhs(const int *__restrict__ ids, const double *n0, const double *n1,
                                                     const double *n2, double *rhs, int nw) {
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const size_t base = (size_t)(id - 1) * NCELL;
    const double fac = -UC_elem_charge / UC_eps0;
    const double q0 = -1 * fac, q1 = 1 * fac, q2 = -1 * fac;
    for (int x = ltid<BS>(); x < NCELL; x += BS) {
      double r = 0.0;
      r = r + q0 * n0[base + x];
      r = r + q1 * n1[base + x];
      r = r + q2 * n2[base + x];
      rhs[base + x] = r;
    }
  }
}

// More species than the standard three (a reaction list, afgpu_set_chemistry):
// pool pointers and charges of every species.
constexpr int SP_MAX = 16;
struct SpeciesArgs {
  double *sp[SP_MAX][3];  // species k, state s
  int charge[SP_MAX];
  int n;
};

// field_set_rhs (src/m_field.f90:406-444) with every charged species, in
// species order.
template <int NCELL, int BS>
__global__ void __launch_bounds__(BS) k_set_rhs_n(const int *__restrict__ ids, SpeciesArgs S, int s_in, double *rhs,
                                                  int nw) {
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int id = ids[w];
    if (id <= 0) continue;
    const size_t base = (size_t)(id - 1) * NCELL;
    const double fac = -UC_elem_charge / UC_eps0;
    for (int x = threadIdx.x; x < NCELL; x += BS) {
      double r = 0.0;
      for (int k = 0; k < S.n; k++)
        if (S.charge[k] != 0) r = r + (S.charge[k] * fac) * S.sp[k][s_in][base + x];
      rhs[base + x] = r;
    }
  }
}

// k_update for a reaction list over S.n species: the densities and
// derivatives of the workgroup's cells in LDS ([S.n][BS] each, then the
// reaction records); the
// electrons (species 0) also take the flux divergence.  One workgroup per
// leaf box.
template <int D, int NC, int BS>
__global__ void __launch_bounds__(BS) k_update_n(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                 UpdateArgs A, SpeciesArgs S, unsigned long long *chem_min, int nw) {
  using G = Geo<D, NC>;
  extern __shared__ double lds[];
  double *sdn = lds, *sdd = lds + (size_t)S.n * BS;
  afgpu_reaction_t *sr = (afgpu_reaction_t *)(lds + 2 * (size_t)S.n * BS);  // the records, copied once
  __shared__ double sh[BS / 64];
  const int t = threadIdx.x;
  {
    constexpr int W = sizeof(afgpu_reaction_t) / sizeof(int);
    const int *src = (const int *)A.net.r;
    int *dst = (int *)sr;
    for (int q = t; q < A.net.n_reactions * W; q += BS) dst[q] = src[q];
    __syncthreads();
  }
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int id = ids[w];
    if (id <= 0) continue;
    const DBox &bx = boxes[id - 1];
    const size_t base = (size_t)(id - 1) * G::NCELL;
    const size_t fbase = (size_t)(id - 1) * D * G::NFACE;
    const double dt = A.dt_run ? A.dt_scale * A.dt_run[bx.run] : A.dt;
    double dt_dr[3];
    for (int d = 0; d < D; d++) dt_dr[d] = dt / bx.dr[d];
    const double eps = 1e-100;
    double bmin = 1e100;
    for (int q = t; q < G::NINT; q += BS) {
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      const size_t x = base + G::idx(i, j, k);
      for (int s = 0; s < S.n; s++) {
        double tmp = 0.0;
        for (int m = 0; m < A.n_prev; m++) tmp = tmp + A.w_prev[m] * S.sp[s][A.s_prev[m]][x];
        const double cur = (A.s_deriv == A.s_out) ? tmp : S.sp[s][A.s_deriv][x];
        sdn[s * BS + t] = fmax(cur, 0.0);
        sdd[s * BS + t] = 0.0;
        S.sp[s][A.s_out][x] = tmp;  // (the output state's old value is no longer read)
      }
      const double field = SI_to_Townsend * A.N_inv * A.efield[x];
      const ChemArgs &a = A.net;
      int ix;
      double wl, te = 0;
      lt_loc(a.fld, field, ix, wl);
      if (a.need_te) {
        int ie;
        double we;
        lt_loc(a.energy, field, ie, we);
        te = CHEM_EV_TO_K * lt_at(a.energy, 0, ie, we);
      }
      for (int r = 0; r < a.n_reactions; r++) {
        const afgpu_reaction_t &R = sr[r];
        double kr = chem_rate(R, a, field, te, ix, wl);
        double prod = 1.0;
        for (int qq = 0; qq < R.n_in; qq++) prod = prod * sdn[R.ix_in[qq] * BS + t];
        kr = kr * prod;
        for (int qq = 0; qq < R.n_in; qq++) {
          double &d = sdd[R.ix_in[qq] * BS + t];
          d = d - kr;
        }
        for (int qq = 0; qq < R.n_out; qq++) {
          double &d = sdd[R.ix_out[qq] * BS + t];
          d = d + kr * R.multiplicity_out[qq];
        }
      }
      if (A.last_step)
        for (int s = 0; s < S.n; s++) {
          const double dn = sdn[s * BS + t], dd = sdd[s * BS + t];
          double tl;
          if (A.chem_nmin_on)
            tl = (dn + A.chem_nmin) / fmax(fabs(dd), eps);
          else if (A.chem_limit_loss)
            tl = fmax(dn, eps) / fmax(-dd, eps);
          else
            tl = 1e100;
          bmin = fmin(bmin, tl);
        }
      if (A.photoi_enabled) {
        const double ph = A.photo[x];
        sdd[0 * BS + t] = sdd[0 * BS + t] + ph;
        sdd[A.photoi_species * BS + t] = sdd[A.photoi_species * BS + t] + ph;
      }
      for (int s = 0; s < S.n; s++) S.sp[s][A.s_out][x] = S.sp[s][A.s_out][x] + dt * sdd[s * BS + t];
      // conservative flux divergence of the electrons
      const double *fl = A.flux + fbase;
      double ne = S.sp[0][A.s_out][x];
      if (D == 2) {
        if (A.cyl) {
          const double dr = bx.dr[0];
          const double r = bx.r_min[0] + (i - 0.5) * dr;
          const double inv_r = 1 / r;
          const double rf1 = (r - 0.5 * dr) * inv_r, rf2 = (r + 0.5 * dr) * inv_r;
          ne = ne + dt_dr[0] * (rf1 * fl[G::fidx(i, j, 1, 0)] - rf2 * fl[G::fidx(i + 1, j, 1, 0)]) +
               dt_dr[1] * (fl[G::fidx(i, j, 1, 1)] - fl[G::fidx(i, j + 1, 1, 1)]);
        } else {
          ne = ne + dt_dr[0] * (fl[G::fidx(i, j, 1, 0)] - fl[G::fidx(i + 1, j, 1, 0)]) +
               dt_dr[1] * (fl[G::fidx(i, j, 1, 1)] - fl[G::fidx(i, j + 1, 1, 1)]);
        }
      } else {
        ne = ne + dt_dr[0] * (fl[G::fidx(i, j, k, 0)] - fl[G::fidx(i + 1, j, k, 0)]) +
             dt_dr[1] * (fl[G::fidx(i, j, k, 1)] - fl[G::fidx(i, j + 1, k, 1)]) +
             dt_dr[2] * (fl[G::fidx(i, j, k, 2)] - fl[G::fidx(i, j, k + 1, 2)]);
      }
      S.sp[0][A.s_out][x] = ne;
    }
    if (A.last_step) {
      bmin = block_min<BS>(bmin, sh);
      if (t == 0) atomic_min_nonneg(run_slot(chem_min, bx), bmin);
    }
  }
}

// E = -grad phi on all faces and |E| at cell centres
// (mg_compute_phi_gradient / mg_box_lpl_gradient / mg_box_field_norm,
// afivo/src/m_af_multigrid.f90:1857-2045).
template <int D, int NC, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_gradient(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                      const double *__restrict__ phi, double *efc, double *enorm,
                                                      int nw) {
  using G = Geo<D, NC>;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const DBox &bx = boxes[id - 1];
    const double *u = phi + (size_t)(id - 1) * G::NCELL;
    double *f = efc + (size_t)(id - 1) * D * G::NFACE;
    double inv_dr[3];
    for (int d = 0; d < D; d++) inv_dr[d] = -1.0 / bx.dr[d];
    constexpr int NFD = (NC + 1) * G::NTAN;
    constexpr int st[3] = {1, G::NG, G::NG * G::NG};
    // consecutive threads take consecutive x positions (contiguous cells
    // and faces) in every direction
    for (int q = ltid<BS>(); q < D * NFD; q += BS) {
      const int d = q / NFD, r = q % NFD;
      const int ni = (d == 0) ? NC + 1 : NC, nj = (d == 1) ? NC + 1 : NC;
      const int p[3] = {r % ni + 1, (r / ni) % nj + 1, (D == 3) ? r / (ni * nj) + 1 : 1};
      const int x = G::idx(p[0], p[1], D == 3 ? p[2] : 0);
      f[G::fidx(p[0], p[1], p[2], d)] = inv_dr[d] * (u[x] - u[x - st[d]]);
    }
    double *e = enorm + (size_t)(id - 1) * G::NCELL;
    for (int q = ltid<BS>(); q < G::NINT; q += BS) {
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      const int x = G::idx(i, j, k);
      double s = 0;
      for (int d = 0; d < D; d++) {
        const double lo = inv_dr[d] * (u[x] - u[x - st[d]]);
        const double hi = inv_dr[d] * (u[x + st[d]] - u[x]);
        const double a = lo + hi;
        s = (d == 0) ? a * a : s + a * a;
      }
      e[x] = 0.5 * sqrt(s);
    }
  }
}

// Photon production on leaves (photoionization_rate_from_alpha,
// src/m_photoi.f90:259-291).
template <int D, int NC, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_photoi_src(const int *__restrict__ ids, const double *__restrict__ efield,
                                                        const double *__restrict__ ne, double *rhs, DTable td, double N,
                                                        double coeff, int nw) {
  using G = Geo<D, NC>;
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const size_t base = (size_t)(id - 1) * G::NCELL;
    for (int q = ltid<BS>(); q < G::NINT; q += BS) {
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      const size_t x = base + G::idx(i, j, k);
      const double fld = efield[x];
      const double Td = fld * SI_to_Townsend / N;
      int lix;
      double lw;
      lt_loc(td, Td, lix, lw);
      const double alpha = lt_at(td, 2, lix, lw);
      const double mob = lt_at(td, 0, lix, lw);
      double tmp = fld * mob * alpha * ne[x] * coeff;
      if (tmp < 0) tmp = 0;
      rhs[x] = tmp;
    }
  }
}

// The photoionization source from the Helmholtz modes' solutions
// (photoi_helmh_compute, src/m_photoi_helmh.f90:162-204): on whole boxes,
// zero on refined boxes and on leaves the modes' terms -c_n phi_n added
// mode by mode, in the order and with the operations of the reference's
// clear followed by one update per mode.
constexpr int PHOTO_MAX_MODES = 3;
struct PhotoModes {
  const double *phi[PHOTO_MAX_MODES];
  double coeff[PHOTO_MAX_MODES];
  int n;
};
template <int NCELL, int BS, int BB = 1>
__global__ void __launch_bounds__(BS * BB) k_photo_sum(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                       PhotoModes pm, double *photo, int nw) {
  for (int w = item0<BS, BB>(); w < nw; w += item_stride<BB>()) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const bool leaf = boxes[id - 1].children[0] <= 0;
    const size_t base = (size_t)(id - 1) * NCELL;
    for (int x = ltid<BS>(); x < NCELL; x += BS) {
      double v = 0.0;
      if (leaf)
        for (int n = 0; n < pm.n; n++) v = v - pm.coeff[n] * pm.phi[n][base + x];
      photo[base + x] = v;
    }
  }
}

}  // namespace afg
