// Device part of the refinement flags: the streamer's default_refinement
// (src/m_refine.f90:198-298) on every cell of a listed box, and
// cell_to_ref_flags (afivo/src/m_af_core.f90:1091-1145) turning the cell
// flags into the box's flag and, within the buffer width of a side, into
// refinement of the neighbour there.  Every contribution is a maximum (the
// reference keeps the largest value a box receives, over its own cells,
// its neighbours' buffers and all threads, m_af_core.f90:982-986), so the
// boxes are evaluated in parallel with atomic maxima on the flags.
#pragma once

#include "afgpu_common.h"

namespace afg {

// af_rm_ref, af_keep_ref, af_do_ref (afivo/src/m_af_types.f90:23-29)
constexpr int REF_RM = -1, REF_KEEP = 0, REF_DO = 1;

// GM_dist_vec_line / GM_dist_line (src/m_geometry.f90:23-51): distance of r
// to the segment r0-r1, the length as the square root of the sum of squares
__device__ inline double ref_dist_line(int D, const double *r, const double *r0, const double *r1) {
  double len2 = 0, frac = 0;
  for (int d = 0; d < D; d++) len2 = len2 + (r1[d] - r0[d]) * (r1[d] - r0[d]);
  for (int d = 0; d < D; d++) frac = frac + (r[d] - r0[d]) * (r1[d] - r0[d]);
  double s2 = 0;
  for (int d = 0; d < D; d++) {
    double v;
    if (frac <= 0)
      v = r[d] - r0[d];
    else if (frac >= len2)
      v = r[d] - r1[d];
    else
      v = r[d] - (r0[d] + frac / len2 * (r1[d] - r0[d]));
    s2 = s2 + v * v;
  }
  return sqrt(s2);
}

// OR over one group of BS threads (BB = 1: the workgroup)
template <int BS>
__device__ inline unsigned block_or(unsigned v, unsigned *sh) {
  for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off);
  if (BS == 64) return v;
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned r = 0;
  for (int w = 0; w < BS / 64; w++) r |= sh[w];
  __syncthreads();
  return r;
}

// One workgroup per listed box.  Cell flag bits gathered over the box: bit
// nbmat_at(di, dj, dk) for a cell flagged for refinement within the buffer
// of side (di, dj, dk), bit 30 for any cell flagged for refinement, bit 31
// for any cell to keep.  The gas density is constant (gas).
template <int D, int NC, int BS>
__global__ void __launch_bounds__(BS) k_refine_cells(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                     const double *__restrict__ efield,
                                                     const double *__restrict__ ne, afgpu_refine_t rp, DTable td,
                                                     double gas, int *flags, int nw) {
  using G = Geo<D, NC>;
  __shared__ unsigned sh[BS / 64];
  constexpr unsigned ANY_DO = 1u << 30, ANY_KEEP = 1u << 31;
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int id = ids[w];
    if (id <= 0) continue;  // run masked out
    const DBox &bx = boxes[id - 1];
    double min_dx = bx.dr[0], max_dx = bx.dr[0];
    for (int d = 1; d < D; d++) {
      min_dx = fmin(min_dx, bx.dr[d]);
      max_dx = fmax(max_dx, bx.dr[d]);
    }
    // the rules of the whole box (m_refine.f90:274-296)
    double rmin[3], rmax[3];
    for (int d = 0; d < D; d++) {
      rmin[d] = bx.r_min[d];
      rmax[d] = bx.r_min[d] + bx.dr[d] * NC;
    }
    bool region_do = false, limit_keep = false;
    for (int n = 0; n < rp.n_regions; n++) {
      bool all = true;
      for (int d = 0; d < D; d++) all = all && rmax[d] >= rp.regions_rmin[n][d] && rmin[d] <= rp.regions_rmax[n][d];
      region_do = region_do || (rp.global_time <= rp.regions_tstop[n] && max_dx > rp.regions_dr[n] && all);
    }
    for (int n = 0; n < rp.n_limits; n++) {
      bool all = true;
      for (int d = 0; d < D; d++) all = all && rmin[d] >= rp.limits_rmin[n][d] && rmax[d] <= rp.limits_rmax[n][d];
      limit_keep = limit_keep || (max_dx < 2 * rp.limits_dr[n] && all);
    }
    const bool all_do = max_dx > rp.refine_max_dx;
    const bool min_keep = !all_do && min_dx < 2 * rp.refine_min_dx;
    const bool seeds = rp.global_time < rp.refine_init_time;
    const int buf = rp.buffer_width;
    const double *E = efield + (size_t)(id - 1) * G::NCELL;
    const double *n_e = ne + (size_t)(id - 1) * G::NCELL;
    unsigned m = 0;
    for (int q = threadIdx.x; q < G::NINT; q += BS) {
      int i, j, k;
      interior_ijk<D, NC>(q, i, j, k);
      const int x = G::idx(i, j, k);
      const double fld = E[x] * SI_to_Townsend / gas;
      int lix;
      double lw;
      lt_loc(td, rp.refine_adx_fac * fld, lix, lw);
      double alpha;
      if (rp.use_alpha_effective) {
        alpha = (lt_at(td, 2, lix, lw) - lt_at(td, 3, lix, lw)) * gas / rp.refine_adx_fac;
        alpha = fmax(0.0, alpha);
      } else {
        alpha = lt_at(td, 2, lix, lw) * gas / rp.refine_adx_fac;
      }
      const double adx = max_dx * alpha;
      int f;
      if (adx > rp.refine_adx && n_e[x] > rp.refine_min_dens)
        f = REF_DO;
      else if (adx < 0.125 * rp.refine_adx && max_dx < rp.derefine_dx)
        f = REF_RM;
      else
        f = REF_KEEP;
      if (seeds) {
        const int ijk[3] = {i, j, k};
        double r[3];
        for (int d = 0; d < D; d++) r[d] = bx.r_min[d] + (ijk[d] - 0.5) * bx.dr[d];
        for (int n = 0; n < rp.n_seeds; n++) {
          const double dist = ref_dist_line(D, r, rp.seed_r0[n], rp.seed_r1[n]);
          if (dist - rp.seed_width[n] < 2 * max_dx && max_dx > rp.refine_init_fac * rp.seed_width[n]) f = REF_DO;
        }
      }
      if (region_do && i == NC / 2 && j == NC / 2 && (D == 2 || k == NC / 2)) f = REF_DO;
      if (limit_keep && f == REF_DO) f = REF_KEEP;
      if (all_do)
        f = REF_DO;
      else if (min_keep && f == REF_DO)
        f = REF_KEEP;
      if (f == REF_KEEP) m |= ANY_KEEP;
      if (f != REF_DO) continue;
      m |= ANY_DO;
      if (buf <= 0) continue;
      // the sides whose buffer holds this cell: per axis the low side
      // (index <= buf), the high side (index > NC - buf) or neither
      const int ijk[3] = {i, j, k};
      int lo_ok[3], hi_ok[3];
      for (int d = 0; d < 3; d++) {
        lo_ok[d] = d < D && ijk[d] <= buf;
        hi_ok[d] = d < D && ijk[d] >= NC - buf + 1;
      }
      const int kr = (D == 3) ? 1 : 0;
      for (int dk = -kr; dk <= kr; dk++)
        for (int dj = -1; dj <= 1; dj++)
          for (int di = -1; di <= 1; di++) {
            const int dir[3] = {di, dj, dk};
            bool in = !(di == 0 && dj == 0 && dk == 0);
            for (int d = 0; d < D; d++) in = in && (dir[d] == 0 || (dir[d] < 0 ? lo_ok[d] : hi_ok[d]));
            if (in) m |= 1u << nbmat_at<D>(di, dj, dk);
          }
    }
    m = block_or<BS>(m, sh);
    if (threadIdx.x == 0) {
      atomicMax(&flags[id - 1], (m & ANY_DO) ? REF_DO : (m & ANY_KEEP) ? REF_KEEP : REF_RM);
      if (buf > 0)
        for (int b = 0; b < ((D == 3) ? 27 : 9); b++)
          if ((m >> b) & 1u) {
            const int nb = bx.nbmat[b];
            if (nb > 0) atomicMax(&flags[nb - 1], REF_DO);
          }
    }
  }
}

// A new box's values from its parent (auto_prolong, afivo/src/m_af_core.f90:
// 843-881): af_prolong_limit (limiter >= 0: the slopes limited with the
// variable's limiter, m_af_prolong.f90:311-420) or af_prolong_linear
// (m_af_prolong.f90:531-679), not adding: every interior cell is written
// once, as the reference's zero plus its sum.  One workgroup per new box.
// The variables of one call: pool pointer and limiter (-1: linear).
constexpr int NEW_MAX_VARS = 8;
struct NewVars {
  double *v[NEW_MAX_VARS];
  int limiter[NEW_MAX_VARS];
  int n;
};

// Work item w is box w / vars.n, variable w % vars.n.
template <int D, int NC, int BS>
__global__ void __launch_bounds__(BS) k_prolong_new(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                    NewVars vars, int nw) {
  using G = Geo<D, NC>;
  constexpr int H = NC / 2;
  constexpr int NQ = (D == 3) ? H * H * H : H * H;
  constexpr int st[3] = {1, G::NG, G::NG * G::NG};
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int id = ids[w / vars.n];
    if (id <= 0) continue;
    double *var = vars.v[w % vars.n];
    const int limiter = vars.limiter[w % vars.n];
    const bool LIMITED = limiter >= 0;
    const DBox &bc = boxes[id - 1];
    const double *P = var + (size_t)(bc.parent - 1) * G::NCELL;
    double *u = var + (size_t)(id - 1) * G::NCELL;
    int off[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) off[d] = ((bc.ix[d] - 1) & 1) * H;
    for (int q = threadIdx.x; q < NQ; q += BS) {
      const int i = q % H + 1, j = (q / H) % H + 1, k = (D == 3) ? q / (H * H) + 1 : 0;
      const int x = G::idx(off[0] + i, off[1] + j, (D == 3) ? off[2] + k : 0);
      const int y0 = G::idx(2 * i - 1, 2 * j - 1, (D == 3) ? 2 * k - 1 : 0);
      if (LIMITED) {
        const double f0 = P[x];
        double f[3];
        for (int d = 0; d < D; d++) f[d] = 0.25 * lim_apply<double>(P[x] - P[x - st[d]], P[x + st[d]] - P[x], limiter);
        for (int c = 0; c < (1 << D); c++) {
          const int di = c & 1, dj = (c >> 1) & 1, dk = (c >> 2) & 1;
          double v = f0;
          v = di ? v + f[0] : v - f[0];
          v = dj ? v + f[1] : v - f[1];
          if (D == 3) v = dk ? v + f[2] : v - f[2];
          u[y0 + di * st[0] + dj * st[1] + dk * st[2]] = v + 0.0;
        }
      } else if (D == 2) {
        const double f1 = 1 / 16.0, f3 = 3 / 16.0, f9 = 9 / 16.0;
        auto p = [&](int a, int b) { return P[x + a + b * st[1]]; };
        const double f0 = f9 * p(0, 0), flx = f3 * p(-1, 0), fhx = f3 * p(1, 0), fly = f3 * p(0, -1),
                     fhy = f3 * p(0, 1), fll = f1 * p(-1, -1), fhl = f1 * p(1, -1), flh = f1 * p(-1, 1),
                     fhh = f1 * p(1, 1);
        u[y0] = f0 + flx + fly + fll + 0.0;
        u[y0 + 1] = f0 + fhx + fly + fhl + 0.0;
        u[y0 + st[1]] = f0 + flx + fhy + flh + 0.0;
        u[y0 + st[1] + 1] = f0 + fhx + fhy + fhh + 0.0;
      } else {
        const double f1 = 1 / 64.0, f3 = 3 / 64.0, f9 = 9 / 64.0, f27 = 27 / 64.0;
        auto p = [&](int a, int b, int e) { return P[x + a + b * st[1] + e * st[2]]; };
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
        auto Y = [&](int a, int b, int e) -> double & { return u[y0 + a + b * st[1] + e * st[2]]; };
        Y(0, 0, 0) = f000 + fl00 + f0l0 + f00l + fll0 + fl0l + f0ll + flll + 0.0;
        Y(1, 0, 0) = f000 + fh00 + f0l0 + f00l + fhl0 + fh0l + f0ll + fhll + 0.0;
        Y(0, 1, 0) = f000 + fl00 + f0h0 + f00l + flh0 + fl0l + f0hl + flhl + 0.0;
        Y(1, 1, 0) = f000 + fh00 + f0h0 + f00l + fhh0 + fh0l + f0hl + fhhl + 0.0;
        Y(0, 0, 1) = f000 + fl00 + f0l0 + f00h + fll0 + fl0h + f0lh + fllh + 0.0;
        Y(1, 0, 1) = f000 + fh00 + f0l0 + f00h + fhl0 + fh0h + f0lh + fhlh + 0.0;
        Y(0, 1, 1) = f000 + fl00 + f0h0 + f00h + flh0 + fl0h + f0hh + flhh + 0.0;
        Y(1, 1, 1) = f000 + fh00 + f0h0 + f00h + fhh0 + fh0h + f0hh + fhhh + 0.0;
      }
    }
  }
}

__global__ void k_fill_int(int *p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace afg
