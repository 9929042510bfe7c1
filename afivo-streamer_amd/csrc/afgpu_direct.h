// Exact solve on a uniformly refined level (mixed-precision form only).
//
// When every level up to lc covers the whole domain, the level-lc operator
// (afivo's constant-coefficient 3/5/7-point stencil with the physical
// boundary conditions folded in, m_af_stencil.f90:367-496,
// m_af_ghostcell.f90:173-279) is a Kronecker sum of 1-D operators, exactly
// like the level-1 box handled by k_coarse_solve.  The mixed-precision
// inner cycle solves its correction equation exactly on that level with the
// same fast diagonalisation, applied to the whole level at once: the
// level's boxes are gathered into one contiguous N^D grid (x fastest), each
// axis is transformed with the 1-D eigenvector matrix (small dense
// products, tiled through LDS), divided by the eigenvalue sums, and
// transformed back.  All transforms run in FP64.
#pragma once

#include "afgpu_common.h"

namespace afg {

// Interior cells of the level's boxes <-> the contiguous grid.  `map` lists
// the box id of every box position (bx + nb*(by + nb*bz)), one workgroup
// per box; in a forest, run r's positions and grid follow those of run r-1.
template <int D, int NC, int BS, typename T>
__global__ void __launch_bounds__(BS) k_lc_gather(const int *__restrict__ map, int nb, const T *__restrict__ src,
                                                  double *__restrict__ A) {
  using G = Geo<D, NC>;
  const int b = blockIdx.x;
  const int id = map[b];
  if (id <= 0) return;  // run masked out
  const int per = (D == 3) ? nb * nb * nb : nb * nb;  // boxes of one run's level
  const int run = b / per, b0 = b % per;
  const int bx = b0 % nb, by = (b0 / nb) % nb, bz = (D == 3) ? b0 / (nb * nb) : 0;
  const int N = nb * NC;
  A += (size_t)run * ((D == 3) ? (size_t)N * N * N : (size_t)N * N);
  const T *u = src + (size_t)(id - 1) * Lay<D, NC, T>::STRIDE;
  for (int q = threadIdx.x; q < G::NINT; q += BS) {
    const int i = q % NC, j = (q / NC) % NC, k = (D == 3) ? q / (NC * NC) : 0;
    const size_t g = (size_t)(bx * NC + i) + (size_t)N * ((by * NC + j) + (size_t)N * (bz * NC + k));
    A[g] = (double)u[Lay<D, NC, T>::in_at(i + 1, j + 1, (D == 3) ? k + 1 : 0)];
  }
}

template <int D, int NC, int BS, typename T>
__global__ void __launch_bounds__(BS) k_lc_scatter(const int *__restrict__ map, int nb, const double *__restrict__ A,
                                                   T *__restrict__ dst) {
  using G = Geo<D, NC>;
  const int b = blockIdx.x;
  const int id = map[b];
  if (id <= 0) return;  // run masked out
  const int per = (D == 3) ? nb * nb * nb : nb * nb;
  const int run = b / per, b0 = b % per;
  const int bx = b0 % nb, by = (b0 / nb) % nb, bz = (D == 3) ? b0 / (nb * nb) : 0;
  const int N = nb * NC;
  A += (size_t)run * ((D == 3) ? (size_t)N * N * N : (size_t)N * N);
  T *u = dst + (size_t)(id - 1) * Lay<D, NC, T>::STRIDE;
  for (int q = threadIdx.x; q < G::NINT; q += BS) {
    const int i = q % NC, j = (q / NC) % NC, k = (D == 3) ? q / (NC * NC) : 0;
    const size_t g = (size_t)(bx * NC + i) + (size_t)N * ((by * NC + j) + (size_t)N * (bz * NC + k));
    u[Lay<D, NC, T>::in_at(i + 1, j + 1, (D == 3) ? k + 1 : 0)] = (T)A[g];
  }
}

// Transform along axis a >= 1 (y, or z in 3D):
//   out[.., i, .., x] = sum_m M[i][m] in[.., m, .., x]
// A workgroup takes TC consecutive x columns at one fixed value of the
// remaining coordinate; the N x TC input tile is staged in LDS, lanes run
// over the columns (coalesced) and the thread groups over output rows.
// With `div`, the result is divided by the eigenvalue sum of its cell
// (the last forward transform).
template <int D, int N>
__global__ void __launch_bounds__(256) k_lc_lines(const double *__restrict__ in, double *__restrict__ out,
                                                  const double *__restrict__ M, int axis, int div,
                                                  const double *__restrict__ eig, double lambda2) {
  constexpr int TC = (N < 64) ? N : (N == 64 ? 64 : 32);
  constexpr int G = 256 / TC;
  constexpr int RPT = (N + G - 1) / G;
  __shared__ double tile[N * TC];
  const int nxc = N / TC;
  const int xc = blockIdx.x % nxc, o = blockIdx.x / nxc;  // o: the other coordinate
  const int x0 = xc * TC;
  const size_t sa = (axis == 1) ? (size_t)N : (size_t)N * N;            // stride of the axis
  // stride of o (2D: the run of a forest)
  const size_t so = (D == 2) ? (size_t)N * N : ((axis == 1) ? (size_t)N * N : (size_t)N);
  const size_t base = (size_t)o * so + x0;
  for (int q = threadIdx.x; q < N * TC; q += 256) {
    const int m = q / TC, c = q % TC;
    tile[q] = in[base + (size_t)m * sa + c];
  }
  __syncthreads();
  const int c = threadIdx.x % TC, g = threadIdx.x / TC;
  double acc[RPT];
  for (int r = 0; r < RPT; r++) acc[r] = 0.0;
  for (int m = 0; m < N; m++) {
    const double t = tile[m * TC + c];
    for (int r = 0; r < RPT; r++) {
      const int i = g + r * G;
      if (i < N) acc[r] = acc[r] + M[i * N + m] * t;
    }
  }
  for (int r = 0; r < RPT; r++) {
    const int i = g + r * G;
    if (i >= N) break;
    double v = acc[r];
    if (div) {
      const int x = x0 + c;
      int y, z;
      if (D == 2) {
        y = i;
        z = 0;
      } else if (axis == 1) {
        y = i;
        z = o;
      } else {
        y = o;
        z = i;
      }
      double den = eig[x] + eig[N + y];
      if (D == 3) den = den + eig[2 * N + z];
      den = den - lambda2;
      v = v / den;
    }
    out[base + (size_t)i * sa + c] = v;
  }
}

// Transform along x (contiguous lines): out[r][i] = sum_m in[r][m] MT[m][i]
// with MT the transposed matrix.  A workgroup takes TR consecutive rows,
// staged in LDS (padded rows); lanes run over i (coalesced stores and
// matrix loads), thread groups over (i chunk, row subset).
template <int D, int N>
__global__ void __launch_bounds__(256) k_lc_rows(const double *__restrict__ in, double *__restrict__ out,
                                                 const double *__restrict__ MT, int nrows) {
  constexpr int TI = (N < 64) ? N : 64;
  constexpr int NIC = N / TI;
  constexpr int G = 256 / TI;
  constexpr int RG = G / NIC;
  constexpr int RPT = 16;
  constexpr int TR = RG * RPT;
  constexpr int LD = N + 1;
  __shared__ double tile[TR * LD];
  const int r0 = blockIdx.x * TR;
  for (int q = threadIdx.x; q < TR * N; q += 256) {
    const int rr = q / N, m = q % N;
    tile[rr * LD + m] = (r0 + rr < nrows) ? in[(size_t)(r0 + rr) * N + m] : 0.0;
  }
  __syncthreads();
  const int lane = threadIdx.x % TI, g = threadIdx.x / TI;
  const int ic = g % NIC, rg = g / NIC;
  const int i = ic * TI + lane;
  double acc[RPT];
  for (int r = 0; r < RPT; r++) acc[r] = 0.0;
  for (int m = 0; m < N; m++) {
    const double w = MT[m * N + i];
    for (int r = 0; r < RPT; r++) acc[r] = acc[r] + tile[(rg * RPT + r) * LD + m] * w;
  }
  for (int r = 0; r < RPT; r++) {
    const int row = r0 + rg * RPT + r;
    if (row < nrows) out[(size_t)row * N + i] = acc[r];
  }
}

// The same transforms for N = 64 and 128 as a tiled product
//   out(i, col) = sum_m M[i][m] X(m, col)
// over 64 x 64 output tiles (i0.., col0..): chunks of 16 m are staged in LDS
// for both M and X (the next chunk is prefetched into registers while the
// current one is used), and each thread accumulates a 4 x 4 block.
// ROWS: X(m, col) = in[col*N + m] (the x transform, col = row of the grid);
// else X(m, col) = in[o*so + m*sa + x0 + col] (the y / z transforms).  With
// `div` (y/z forms only) the result is divided by the eigenvalue sum.
template <int D, int N, bool ROWS>
__global__ void __launch_bounds__(256) k_lc_gemm(const double *__restrict__ in, double *__restrict__ out,
                                                 const double *__restrict__ M, int axis, int div,
                                                 const double *__restrict__ eig, double lambda2) {
  static_assert(N == 64 || N == 128, "tiled transform for N = 64, 128");
  constexpr int T = 64, KC = 16, KP = KC + 1, NIT = N / T;
  __shared__ double Ms[T * KP];
  __shared__ double Xs[KC * T];
  const int it = blockIdx.x % NIT, rest = blockIdx.x / NIT;
  const int i0 = it * T;
  size_t sm, sc, base;
  int o = 0, x0 = 0;
  if (ROWS) {
    sm = 1;
    sc = N;
    base = (size_t)rest * T * N;
  } else {
    constexpr int nxc = N / T;
    const int xc = rest % nxc;
    o = rest / nxc;
    x0 = xc * T;
    sm = (axis == 1) ? (size_t)N : (size_t)N * N;
    const size_t so = (D == 2) ? (size_t)N * N : ((axis == 1) ? (size_t)N * N : (size_t)N);
    sc = 1;
    base = (size_t)o * so + x0;
  }
  const int tid = threadIdx.x;
  const int ti = ROWS ? (tid % 16) : (tid / 16);
  const int tc = ROWS ? (tid / 16) : (tid % 16);
  // global -> register staging of one chunk: 4 M and 4 X values per thread
  double pm[4], px[4];
  auto fetch = [&](int k0) {
    for (int u = 0; u < 4; u++) {
      const int q = tid + 256 * u;
      const int i = q / KC, kk = q % KC;
      pm[u] = M[(size_t)(i0 + i) * N + k0 + kk];
      int kx, c;
      if (ROWS) {
        kx = q % KC;
        c = q / KC;
      } else {
        c = q % T;
        kx = q / T;
      }
      px[u] = in[base + (size_t)(k0 + kx) * sm + (size_t)c * sc];
    }
  };
  auto stash = [&]() {
    for (int u = 0; u < 4; u++) {
      const int q = tid + 256 * u;
      Ms[(q / KC) * KP + q % KC] = pm[u];
      if (ROWS)
        Xs[(q % KC) * T + q / KC] = px[u];
      else
        Xs[(q / T) * T + q % T] = px[u];
    }
  };
  double acc[4][4];
  for (int r = 0; r < 4; r++)
    for (int q = 0; q < 4; q++) acc[r][q] = 0.0;
  fetch(0);
  for (int k0 = 0; k0 < N; k0 += KC) {
    stash();
    __syncthreads();
    if (k0 + KC < N) fetch(k0 + KC);
    for (int kk = 0; kk < KC; kk++) {
      double a[4], b[4];
      for (int r = 0; r < 4; r++) a[r] = Ms[(ti + 16 * r) * KP + kk];
      for (int q = 0; q < 4; q++) b[q] = Xs[kk * T + tc + 16 * q];
      for (int r = 0; r < 4; r++)
        for (int q = 0; q < 4; q++) acc[r][q] = acc[r][q] + a[r] * b[q];
    }
    __syncthreads();
  }
  for (int r = 0; r < 4; r++)
    for (int q = 0; q < 4; q++) {
      const int i = i0 + ti + 16 * r, c = tc + 16 * q;
      double v = acc[r][q];
      if (!ROWS && div) {
        const int x = x0 + c;
        int y, z;
        if (D == 2) {
          y = i;
          z = 0;
        } else if (axis == 1) {
          y = i;
          z = o;
        } else {
          y = o;
          z = i;
        }
        double den = eig[x] + eig[N + y];
        if (D == 3) den = den + eig[2 * N + z];
        den = den - lambda2;
        v = v / den;
      }
      out[base + (size_t)i * sm + (size_t)c * sc] = v;
    }
}

// The same transforms on the FP64 matrix cores (v_mfma_f64_16x16x4f64):
//   D[p][q] = sum_k A[p][k] B[k][q]
// y / z forms: p = output index i, q = column x, A = M, B = the input lines;
// x form (ROWS): p = grid row r, q = output index i, A = the input rows,
// B[k][i] = M[i][k].  A workgroup (4 waves) computes a 64 x 64 tile of D,
// each wave a 32 x 32 block as 2 x 2 MFMA tiles; K runs in chunks of 16
// staged in LDS.  Lane layout of the instruction: A operand A[l % 16][l / 16],
// B operand B[l / 16][l % 16], result register j of lane l D[4 j + l / 16][l % 16].
typedef double lc_double4 __attribute__((ext_vector_type(4)));
typedef float lc_float4 __attribute__((ext_vector_type(4)));

// Offset in the box pool (of values of type TS) of point (x, y, r) of the
// uniform level, stored as boxes of NC^D cells: r is z in 3D and the run of
// a forest in 2D; -1 when the box is masked out.
template <int D, int NC, typename TS>
__device__ inline long long lc_pool_offset(const int *__restrict__ map, int nb, int x, int y, int r) {
  using L = Lay<D, NC, TS>;
  int b, ck = 0;
  if (D == 3) {
    b = x / NC + nb * (y / NC + nb * (r / NC));
    ck = r % NC + 1;
  } else {
    b = x / NC + nb * (y / NC) + r * nb * nb;
  }
  const int id = map[b];
  return id > 0 ? (long long)(id - 1) * L::STRIDE + L::in_at(x % NC + 1, y % NC + 1, ck) : -1;
}

// IO (x form only): 0 grid to grid; 1 the input rows read from the box
// pool `pin` (the gather, fused); 2 the output rows written to the box pool
// `pout` (the scatter, fused).  Box map and boxes per axis as k_lc_gather.
// WAVES wavefronts share a 64 x 64 output tile: two rows of WAVES / 2, each
// wave 32 x (128 / WAVES) (at 8, two waves per SIMD in every workgroup, for
// the few workgroups a pass has: (N / 64)^2 N).  Every output element is
// one wave's sum over K in the same order whatever WAVES is.
//
// FOLD (N = 128, an axis whose 1-D operator is symmetric under reversal,
// m -> N - 1 - m): its eigenvectors are even or odd, so a transform needs
// half the products.  1, forward: M is N x N/2, the even modes' rows first
// (M[i][m], m < N/2, of the even eigenvectors), then the odd ones'; the
// even half of the output (the first 64-row tile) takes u_m + u_{N-1-m},
// the odd half u_m - u_{N-1-m}, over K = N/2.  2, backward: M is N/2 x N,
// M[m][i] for m < N/2 over the modes in the same order; the sums over the
// even modes (E) and the odd ones (O) give u_m = E + O and
// u_{N-1-m} = E - O, one 64-row output tile writing both halves.
template <int D, int N, bool ROWS, int IO = 0, int NC = 16, typename TS = float, int WAVES = 8, int FOLD = 0>
__global__ void __launch_bounds__(64 * WAVES) k_lc_mfma(const double *__restrict__ in, double *__restrict__ out,
                                                        const double *__restrict__ M, int axis, int div,
                                                        const double *__restrict__ eig, double lambda2,
                                                        const int *__restrict__ map, int nb,
                                                        const TS *__restrict__ pin, TS *__restrict__ pout) {
  static_assert(IO == 0 || ROWS, "box-pool input and output in the x form only");
  static_assert(N == 64 || N == 128, "matrix-core transform for N = 64, 128");
  static_assert(WAVES == 4 || WAVES == 8, "4 or 8 wavefronts per 64 x 64 tile");
  static_assert(FOLD == 0 || N == 128, "the folded form has one 64-row tile per parity");
  constexpr int BT = 64 * WAVES;
  constexpr int T = 64, KC = 32, AP = KC + 1, BP = T + 1, NT = N / T;
  constexpr int NTO = (FOLD == 2) ? 1 : NT;    // output tiles
  constexpr int KN = (FOLD == 1) ? N / 2 : N;  // K extent
  constexpr int MS = (FOLD == 1) ? N / 2 : N;  // row stride of M
  constexpr int NS = (FOLD == 2) ? 2 : 1;      // accumulator sets (E, O)
  constexpr int PER = T * KC / BT;  // values of each operand tile per thread
  constexpr int WQ = WAVES / 2, QW = T / WQ, NTB = QW / 16;  // waves along q, their width, 16-wide tiles
  __shared__ double As[T * AP];   // As[p][kk]
  __shared__ double Bs[KC * BP];  // Bs[kk][q]
  // blocks: ROWS: (output tiles) x (row tiles); else (output tiles) x
  // (column tiles) x (the remaining coordinate)
  const int t0 = blockIdx.x % NTO, rest = blockIdx.x / NTO;
  const double sg = (t0 == 0) ? 1.0 : -1.0;  // the fold of the forward form: even, odd half
  size_t sm = 1, base = 0;
  int o = 0, x0 = 0, p0 = 0, q0 = 0;
  if (ROWS) {
    p0 = rest * T;  // grid rows
    q0 = t0 * T;    // output index i
  } else {
    p0 = t0 * T;  // output index i
    x0 = (rest % NT) * T;
    o = rest / NT;
    q0 = 0;
    sm = (axis == 1) ? (size_t)N : (size_t)N * N;
    const size_t so = (D == 2) ? (size_t)N * N : ((axis == 1) ? (size_t)N * N : (size_t)N);
    base = (size_t)o * so + x0;
  }
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wp = (w / WQ) * 32, wq = (w % WQ) * QW;
  // one K chunk of both operands, global -> registers (consecutive threads
  // take consecutive addresses) and registers -> LDS
  double pa[PER], pb[PER];
  auto fetch = [&](int k0) {
    for (int u = 0; u < PER; u++) {
      const int q = tid + BT * u;
      const int pp = q / KC, kk = q % KC;
      const int k = k0 + kk;
      if (IO == 1) {
        const int row = p0 + pp;
        const long long off = lc_pool_offset<D, NC, TS>(map, nb, k, row % N, row / N);
        double a = off >= 0 ? (double)pin[off] : 0.0;
        if (FOLD == 1) {
          const long long off2 = lc_pool_offset<D, NC, TS>(map, nb, N - 1 - k, row % N, row / N);
          a = a + sg * (off2 >= 0 ? (double)pin[off2] : 0.0);
        }
        pa[u] = a;
      } else if (ROWS) {
        const double *r = in + (size_t)(p0 + pp) * N;
        pa[u] = (FOLD == 1) ? r[k] + sg * r[N - 1 - k] : r[k];
      } else {
        pa[u] = M[(size_t)(p0 + pp) * MS + k];
      }
      if (ROWS) {
        pb[u] = M[(size_t)(q0 + pp) * MS + k];
      } else {
        const int kb = q / T, qq = q % T;
        const double *c = in + base + qq;
        pb[u] = (FOLD == 1) ? c[(size_t)(k0 + kb) * sm] + sg * c[(size_t)(N - 1 - k0 - kb) * sm]
                            : c[(size_t)(k0 + kb) * sm];
      }
    }
  };
  auto stash = [&]() {
    for (int u = 0; u < PER; u++) {
      const int q = tid + BT * u;
      As[(q / KC) * AP + q % KC] = pa[u];
      if (ROWS)
        Bs[(q % KC) * BP + q / KC] = pb[u];
      else
        Bs[(q / T) * BP + q % T] = pb[u];
    }
  };
  lc_double4 acc[NS][2][NTB];
  for (int e = 0; e < NS; e++)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < NTB; b++) acc[e][a][b] = lc_double4{0.0, 0.0, 0.0, 0.0};
  fetch(0);
  for (int k0 = 0; k0 < KN; k0 += KC) {
    stash();
    __syncthreads();
    if (k0 + KC < KN) fetch(k0 + KC);
    const int e = (NS == 2 && k0 >= N / 2) ? 1 : 0;  // (the backward form: the odd modes' sum)
    for (int ks = 0; ks < KC; ks += 4) {
      const int kk = ks + (lane >> 4);
      double a[2], b[NTB];
      for (int t = 0; t < 2; t++) a[t] = As[(wp + 16 * t + (lane & 15)) * AP + kk];
      for (int t = 0; t < NTB; t++) b[t] = Bs[kk * BP + wq + 16 * t + (lane & 15)];
      for (int ta = 0; ta < 2; ta++)
        for (int tb = 0; tb < NTB; tb++) {
          if (NS == 2 && e == 1)
            acc[NS - 1][ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta], b[tb], acc[NS - 1][ta][tb], 0, 0, 0);
          else
            acc[0][ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ta], b[tb], acc[0][ta][tb], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  for (int ta = 0; ta < 2; ta++)
    for (int tb = 0; tb < NTB; tb++)
      for (int j = 0; j < 4; j++) {
        const int pp = p0 + wp + 16 * ta + 4 * j + (lane >> 4);
        const int qq = q0 + wq + 16 * tb + (lane & 15);
        if (FOLD == 2) {
          // u_m = E + O, u_{N-1-m} = E - O; m is qq (ROWS) or pp (lines)
          const double vE = acc[0][ta][tb][j], vO = acc[NS - 1][ta][tb][j];
          const double lo = vE + vO, hi = vE - vO;
          if (ROWS && IO == 2) {
            long long off = lc_pool_offset<D, NC, TS>(map, nb, qq, pp % N, pp / N);
            if (off >= 0) pout[off] = (TS)lo;
            off = lc_pool_offset<D, NC, TS>(map, nb, N - 1 - qq, pp % N, pp / N);
            if (off >= 0) pout[off] = (TS)hi;
          } else if (ROWS) {
            out[(size_t)pp * N + qq] = lo;
            out[(size_t)pp * N + N - 1 - qq] = hi;
          } else {
            out[base + (size_t)pp * sm + qq] = lo;
            out[base + (size_t)(N - 1 - pp) * sm + qq] = hi;
          }
          continue;
        }
        double v = acc[0][ta][tb][j];
        if (ROWS && IO == 2) {
          const long long off = lc_pool_offset<D, NC, TS>(map, nb, qq, pp % N, pp / N);
          if (off >= 0) pout[off] = (TS)v;
        } else if (ROWS) {
          out[(size_t)pp * N + qq] = v;
        } else {
          if (div) {
            const int x = x0 + qq;
            int y, z;
            if (D == 2) {
              y = pp;
              z = 0;
            } else if (axis == 1) {
              y = pp;
              z = o;
            } else {
              y = o;
              z = pp;
            }
            double den = eig[x] + eig[N + y];
            if (D == 3) den = den + eig[2 * N + z];
            den = den - lambda2;
            v = v / den;
          }
          out[base + (size_t)pp * sm + qq] = v;
        }
      }
}

// k_lc_mfma in single precision (the mixed form's inner cycle, option
// AFGPU_OPT_DIRECT_F32): grids, matrices and eigenvalues in FP32, products
// on the FP32 matrix cores (v_mfma_f32_16x16x4f32: the operands as in the
// FP64 form, result register j of lane l holding row 4 (l / 16) + j,
// column l % 16 of the 16 x 16 block).
template <int D, int N, bool ROWS, int IO = 0, int NC = 16>
__global__ void __launch_bounds__(256) k_lc_mfma32(const float *__restrict__ in, float *__restrict__ out,
                                                   const float *__restrict__ M, int axis, int div,
                                                   const float *__restrict__ eig, float lambda2,
                                                   const int *__restrict__ map, int nb, const float *__restrict__ pin,
                                                   float *__restrict__ pout) {
  static_assert(IO == 0 || ROWS, "box-pool input and output in the x form only");
  static_assert(N == 64 || N == 128, "matrix-core transform for N = 64, 128");
  constexpr int T = 64, KC = 32, AP = KC + 1, BP = T + 1, NT = N / T;
  constexpr int PER = T * KC / 256;  // values of each operand tile per thread
  __shared__ float As[T * AP];   // As[p][kk]
  __shared__ float Bs[KC * BP];  // Bs[kk][q]
  // blocks: ROWS: (output tiles) x (row tiles); else (output tiles) x
  // (column tiles) x (the remaining coordinate)
  const int t0 = blockIdx.x % NT, rest = blockIdx.x / NT;
  size_t sm = 1, base = 0;
  int o = 0, x0 = 0, p0 = 0, q0 = 0;
  if (ROWS) {
    p0 = rest * T;  // grid rows
    q0 = t0 * T;    // output index i
  } else {
    p0 = t0 * T;  // output index i
    x0 = (rest % NT) * T;
    o = rest / NT;
    q0 = 0;
    sm = (axis == 1) ? (size_t)N : (size_t)N * N;
    const size_t so = (D == 2) ? (size_t)N * N : ((axis == 1) ? (size_t)N * N : (size_t)N);
    base = (size_t)o * so + x0;
  }
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wp = (w >> 1) * 32, wq = (w & 1) * 32;
  // one K chunk of both operands, global -> registers (consecutive threads
  // take consecutive addresses) and registers -> LDS
  float pa[PER], pb[PER];
  auto fetch = [&](int k0) {
    for (int u = 0; u < PER; u++) {
      const int q = tid + 256 * u;
      const int pp = q / KC, kk = q % KC;
      if (IO == 1) {
        const int row = p0 + pp;
        const long long off = lc_pool_offset<D, NC, float>(map, nb, k0 + kk, row % N, row / N);
        pa[u] = off >= 0 ? pin[off] : 0.0f;
      } else {
        pa[u] = ROWS ? in[(size_t)(p0 + pp) * N + k0 + kk] : M[(size_t)(p0 + pp) * N + k0 + kk];
      }
      if (ROWS) {
        pb[u] = M[(size_t)(q0 + pp) * N + k0 + kk];
      } else {
        const int kb = q / T, qq = q % T;
        pb[u] = in[base + (size_t)(k0 + kb) * sm + qq];
      }
    }
  };
  auto stash = [&]() {
    for (int u = 0; u < PER; u++) {
      const int q = tid + 256 * u;
      As[(q / KC) * AP + q % KC] = pa[u];
      if (ROWS)
        Bs[(q % KC) * BP + q / KC] = pb[u];
      else
        Bs[(q / T) * BP + q % T] = pb[u];
    }
  };
  lc_float4 acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = lc_float4{0.0f, 0.0f, 0.0f, 0.0f};
  fetch(0);
  for (int k0 = 0; k0 < N; k0 += KC) {
    stash();
    __syncthreads();
    if (k0 + KC < N) fetch(k0 + KC);
    for (int ks = 0; ks < KC; ks += 4) {
      const int kk = ks + (lane >> 4);
      float a[2], b[2];
      for (int t = 0; t < 2; t++) {
        a[t] = As[(wp + 16 * t + (lane & 15)) * AP + kk];
        b[t] = Bs[kk * BP + wq + 16 * t + (lane & 15)];
      }
      for (int ta = 0; ta < 2; ta++)
        for (int tb = 0; tb < 2; tb++)
          acc[ta][tb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[ta], b[tb], acc[ta][tb], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int ta = 0; ta < 2; ta++)
    for (int tb = 0; tb < 2; tb++)
      for (int j = 0; j < 4; j++) {
        const int pp = p0 + wp + 16 * ta + 4 * (lane >> 4) + j;
        const int qq = q0 + wq + 16 * tb + (lane & 15);
        float v = acc[ta][tb][j];
        if (ROWS && IO == 2) {
          const long long off = lc_pool_offset<D, NC, float>(map, nb, qq, pp % N, pp / N);
          if (off >= 0) pout[off] = v;
        } else if (ROWS) {
          out[(size_t)pp * N + qq] = v;
        } else {
          if (div) {
            const int x = x0 + qq;
            int y, z;
            if (D == 2) {
              y = pp;
              z = 0;
            } else if (axis == 1) {
              y = pp;
              z = o;
            } else {
              y = o;
              z = pp;
            }
            float den = eig[x] + eig[N + y];
            if (D == 3) den = den + eig[2 * N + z];
            den = den - lambda2;
            v = v / den;
          }
          out[base + (size_t)pp * sm + qq] = v;
        }
      }
}

// The last forward transform (z in 3D, y in 2D), the division by the
// eigenvalue sums and the first backward transform along the same lines, in
// one pass: a workgroup (N/16 waves) holds all N points of a 64-column tile
// of lines in LDS, forms D1 = Mf B on the matrix cores, divides, keeps D1 in
// LDS and forms D2 = Mb D1 (each product with the K order of k_lc_mfma).
// One read and one write of the grid instead of two of each.
template <int D, int N>
__global__ void __launch_bounds__(4 * N) k_lc_mfma_fused(const double *__restrict__ in, double *__restrict__ out,
                                                         const double *__restrict__ Mf, const double *__restrict__ Mb,
                                                         const double *__restrict__ eig, double lambda2) {
  static_assert(N == 64 || N == 128, "matrix-core transform for N = 64, 128");
  constexpr int T = 64, KC = 32, AP = KC + 1, BP = T + 1, NT = N / T, NTH = 4 * N;
  __shared__ double Bs[N * BP];   // the lines Bs[k][q], then D1
  __shared__ double As[N * AP];   // one K chunk of the matrix, As[p][kk]
  const int xt = blockIdx.x % NT, o = blockIdx.x / NT;  // o: y in 3D, the run in 2D
  const int x0 = xt * T;
  const size_t sm = (D == 3) ? (size_t)N * N : (size_t)N;  // stride along the lines
  const size_t so = (D == 3) ? (size_t)N : (size_t)N * N;  // stride of o
  const size_t base = (size_t)o * so + x0;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wp = (w >> 1) * 32, wq = (w & 1) * 32;
  for (int q = tid; q < N * T; q += NTH) Bs[(q / T) * BP + q % T] = in[base + (size_t)(q / T) * sm + q % T];
  lc_double4 acc[2][2];
  auto product = [&](const double *__restrict__ M) {
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) acc[a][b] = lc_double4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < N; k0 += KC) {
      __syncthreads();  // (the chunk and, before the first, the lines in LDS)
      for (int q = tid; q < N * KC; q += NTH) As[(q / KC) * AP + q % KC] = M[(size_t)(q / KC) * N + k0 + q % KC];
      __syncthreads();
      for (int ks = 0; ks < KC; ks += 4) {
        const int kk = ks + (lane >> 4);
        double av[2], bv[2];
        for (int t = 0; t < 2; t++) {
          av[t] = As[(wp + 16 * t + (lane & 15)) * AP + kk];
          bv[t] = Bs[(k0 + kk) * BP + wq + 16 * t + (lane & 15)];
        }
        for (int ta = 0; ta < 2; ta++)
          for (int tb = 0; tb < 2; tb++)
            acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ta], bv[tb], acc[ta][tb], 0, 0, 0);
      }
    }
  };
  product(Mf);
  __syncthreads();  // every wave is done with the lines
  for (int ta = 0; ta < 2; ta++)
    for (int tb = 0; tb < 2; tb++)
      for (int j = 0; j < 4; j++) {
        const int pp = wp + 16 * ta + 4 * j + (lane >> 4);  // position along the lines
        const int qq = wq + 16 * tb + (lane & 15);
        double den = eig[x0 + qq] + eig[N + ((D == 3) ? o : pp)];
        if (D == 3) den = den + eig[2 * N + pp];
        den = den - lambda2;
        Bs[pp * BP + qq] = acc[ta][tb][j] / den;
      }
  product(Mb);
  for (int ta = 0; ta < 2; ta++)
    for (int tb = 0; tb < 2; tb++)
      for (int j = 0; j < 4; j++) {
        const int pp = wp + 16 * ta + 4 * j + (lane >> 4);
        const int qq = wq + 16 * tb + (lane & 15);
        out[base + (size_t)pp * sm + qq] = acc[ta][tb][j];
      }
}

}  // namespace afg
