// Diagnostics and box I/O from the device pool: the statistics of the
// regression log (src/m_output.f90:788-842) and the box data of
// checkpoints (afivo/src/m_af_output.f90:90-114, 269-295).
#pragma once

#include <cfloat>

#include "afgpu_common.h"

namespace afg {

constexpr int DIAG_MAX_VARS = 8;

// Pool pointers of the variables of one call (read or written).
struct VarPtrs {
  double *v[DIAG_MAX_VARS];
  int n;
};

template <int BS>
__device__ inline double block_sum(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < BS / 64; w++) r += sh[w];
  __syncthreads();
  return r;
}

// One leaf box's terms of af_tree_sum_cc (afivo/src/m_af_utils.f90:
// 966-1026) for powers 1 and 2 (the level's cell volume times the sum of
// the interior values, each times 2 pi r in 2D cylindrical coordinates as
// sum_2pr_box does), and its box_max_cc maximum (m_af_utils.f90:758-769),
// for every variable: part[(w * n + v) * 3 + {0, 1, 2}].  One workgroup
// per box; each variable's block is read once.
template <int D, int NC, int BS>
__global__ void __launch_bounds__(BS) k_diag_box(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                 VarPtrs vars, int cyl, double *part, int nw) {
  using G = Geo<D, NC>;
  __shared__ double sh[BS / 64];
  const double two_pi = 2 * 3.141592653589793;  // 2 acos(-1)
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int id = ids[w];
    if (id <= 0) continue;  // (the fold skips these entries)
    const DBox &bx = boxes[id - 1];
    double fac = 1;
    for (int d = 0; d < D; d++) fac *= bx.dr[d];
    for (int v = 0; v < vars.n; v++) {
      const double *u = vars.v[v] + (size_t)(id - 1) * G::NCELL;
      double s1 = 0, s2 = 0, mx = -DBL_MAX / 10;
      for (int q = threadIdx.x; q < G::NINT; q += BS) {
        int i, j, k;
        interior_ijk<D, NC>(q, i, j, k);
        const double x = u[G::idx(i, j, k)];
        if (cyl) {
          const double r = bx.r_min[0] + (i - 0.5) * bx.dr[0];
          s1 += x * r;
          s2 += x * x * r;
        } else {
          s1 += x;
          s2 += x * x;
        }
        mx = fmax(mx, x);
      }
      s1 = block_sum<BS>(s1, sh);
      s2 = block_sum<BS>(s2, sh);
      mx = block_max<BS>(mx, sh);
      if (threadIdx.x == 0) {
        double *o = part + ((size_t)w * vars.n + v) * 3;
        o[0] = fac * (cyl ? s1 * two_pi : s1);
        o[1] = fac * (cyl ? s2 * two_pi : s2);
        o[2] = mx;
      }
    }
  }
}

// The box terms summed (and maximised) per run: workgroup r takes the
// entries of run r in a fixed order, so the result does not vary between
// calls; every variable in one pass over the entries.
// out[(r * nv + v) * 3 + {0, 1, 2}].
template <int BS>
__global__ void __launch_bounds__(BS) k_diag_fold(const int *__restrict__ ids, const DBox *__restrict__ boxes,
                                                  const double *__restrict__ part, int nw, int nv, int one_run,
                                                  double *out) {
  __shared__ double sh[BS / 64];
  const int r = blockIdx.x;
  double s1[DIAG_MAX_VARS], s2[DIAG_MAX_VARS], mx[DIAG_MAX_VARS];
  for (int v = 0; v < DIAG_MAX_VARS; v++) {
    s1[v] = 0;
    s2[v] = 0;
    mx[v] = -DBL_MAX / 10;
  }
  for (int w = threadIdx.x; w < nw; w += BS) {
    const int id = ids[w];
    if (id <= 0 || (!one_run && boxes[id - 1].run != r)) continue;
    const double *p = part + (size_t)w * nv * 3;
    for (int v = 0; v < DIAG_MAX_VARS; v++)
      if (v < nv) {
        s1[v] += p[3 * v];
        s2[v] += p[3 * v + 1];
        mx[v] = fmax(mx[v], p[3 * v + 2]);
      }
  }
  for (int v = 0; v < nv; v++) {
    const double a = block_sum<BS>(s1[v], sh);
    const double b = block_sum<BS>(s2[v], sh);
    const double m = block_max<BS>(mx[v], sh);
    if (threadIdx.x == 0) {
      double *o = out + ((size_t)r * nv + v) * 3;
      o[0] = a;
      o[1] = b;
      o[2] = m;
    }
  }
}

// Box-major staging of whole boxes (ghost cells included): item w is box
// w / n, variable w % n; variable v has len[v] words per box in its pool
// and starts at off[v] within a box's total words in buf.  to_pool = 0
// gathers the pools into buf, 1 scatters buf into the pools.
struct BoxIoVars {
  double *v[DIAG_MAX_VARS];
  int len[DIAG_MAX_VARS];
  int off[DIAG_MAX_VARS];
  int n, total;
};
__global__ void __launch_bounds__(256) k_box_io(const int *__restrict__ ids, BoxIoVars vars, double *buf, int to_pool,
                                                int nw) {
  for (int w = blockIdx.x; w < nw; w += gridDim.x) {
    const int box = w / vars.n, v = w % vars.n, n = vars.len[v];
    double *pool = vars.v[v] + (size_t)(ids[box] - 1) * n;
    double *b = buf + (size_t)box * vars.total + vars.off[v];
    if (to_pool)
      for (int q = threadIdx.x; q < n; q += 256) pool[q] = b[q];
    else
      for (int q = threadIdx.x; q < n; q += 256) b[q] = pool[q];
  }
}

}  // namespace afg
