// Shared device-side definitions of the afivo-streamer hot path for gfx950.
//
// Data layout in HBM: one pool per cell-centred variable,
//   pool[(id-1) * NG^D + i + NG*(j + NG*k)],  NG = NC + 2,
// i.e. exactly afivo's box%cc(DTIMES(0:nc+1), iv) per box
// (afivo/src/m_af_types.f90:302), so a box can be copied to/from the Fortran
// array without repacking.  Face variables use box%fc's layout, D blocks of
// (NC+1)^D per box.  Topology lives in an array of DBox records.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace afg {

// A box's record.  What a smoothing half-sweep reads (neighbours, level,
// position, coordinates) comes first and the record is aligned to a
// 128-byte cache line, so those fields cost one line per box: with 8^2
// boxes a line per box is a third of the half-sweep's own bytes.
struct alignas(128) DBox {
  int lvl;
  int run;  // index of the independent run (tree of a forest) the box belongs to
  int nb[6];
  int pnb[6];  // faces on a refinement boundary (nb 0): the parent's neighbour there; else 0
  int ix[3];
  int parent;
  double r_min[3];
  double dr[3];
  int children[8];
  int nbmat[27];
};
static_assert(offsetof(DBox, dr) + 3 * sizeof(double) <= 128, "the fields a half-sweep reads share one line");

enum BcType { BC_DIRICHLET = 0, BC_NEUMANN = 1, BC_DIRICHLET_COPY = 2 };
enum RbType { RB_NONE = 0, RB_MG = 1, RB_INTERP = 2, RB_INTERP_LIM = 3 };
enum LimType { LIM_MC = 0, LIM_GMINMOD43 = 1 };

// Boundary methods of one variable (bc per face + refinement-boundary rule).
// In a forest of independent runs, the value of face `run_face` may differ
// per run: it is then run_val[run] (a per-run applied voltage).
struct VarBc {
  int type[6];
  double val[6];
  int rb;
  int limiter;
  int run_face;
  const double *run_val;
};
__device__ inline double bc_value(const VarBc &m, int nb, int run) {
  return (m.run_val && nb == m.run_face) ? m.run_val[run] : m.val[nb];
}

// Reduction slots: each run of a forest has RED_STRIDE slots, each kept in
// RED_STRIPES copies that the workgroups address by their index, so that
// the one atomic per box of a launch over 10^5 boxes does not serialise on
// one address; k_fold_stripes combines the copies.  Layout
// [run][stripe][slot].  Workgroups are dealt to the 8 XCDs in turn, so the
// stripe is taken from the index within the XCD (blockIdx / 8).
constexpr int RED_STRIDE = 16;
constexpr int RED_STRIPES = 32;
__device__ inline unsigned long long *run_slot(unsigned long long *base, const DBox &bx) {
  return base + ((size_t)bx.run * RED_STRIPES + ((blockIdx.x >> 3) & (RED_STRIPES - 1))) * RED_STRIDE;
}

// Lookup table (src/lookup_table_fortran/m_lookup_table.f90): columns [col][row].
struct DTable {
  const double *cols;
  int n;
  double x_min, inv_fac;
};

constexpr double UC_eps0 = 8.8541878128e-12;
constexpr double UC_elem_charge = 1.602176634e-19;
constexpr double SI_to_Townsend = 1e21;

template <int D, int NC>
struct Geo {
  static constexpr int NG = NC + 2;
  static constexpr int NCELL = (D == 3) ? NG * NG * NG : NG * NG;
  static constexpr int NF = NC + 1;
  static constexpr int NFACE = (D == 3) ? NF * NF * NF : NF * NF;  // per dimension
  static constexpr int NINT = (D == 3) ? NC * NC * NC : NC * NC;
  static constexpr int NTAN = (D == 3) ? NC * NC : NC;              // cells on a face
  __host__ __device__ static constexpr int idx(int i, int j, int k) { return i + NG * (j + NG * k); }
  __host__ __device__ static constexpr int fidx(int i, int j, int k, int d) {
    return (i - 1) + NF * ((j - 1) + NF * (k - 1)) + d * NFACE;
  }
};

// ---- storage of a box's cells in a pool -------------------------------------
// Pools of FP64 values keep afivo's box layout (Geo::idx), so a box moves to
// and from the Fortran arrays as it is.  Pools of FP32 values (the mixed
// form's inner cycle, internal to the library) keep the interior cells of
// each red-black colour ((i + j + k) & 1) together, in lexicographic order,
// followed by the ghost cells in afivo's order: a half-sweep then reads and
// writes the cache lines of one colour only.  Both take Geo::NCELL values
// per box (STRIDE).  at(): position of any cell of the box; in_at(): of an
// interior cell; nb_at(): of the neighbour (i, j, k) + o e_d of an interior
// cell.
//
// afivo's layout, also that of every LDS copy of a box:
template <int D, int NC>
struct LayStd {
  using G = Geo<D, NC>;
  static constexpr bool SPLIT = false;
  static constexpr int STRIDE = G::NCELL;
  __host__ __device__ static constexpr int at(int i, int j, int k) { return G::idx(i, j, k); }
  __host__ __device__ static constexpr int in_at(int i, int j, int k) { return G::idx(i, j, k); }
  __host__ __device__ static int nb_at(int i, int j, int k, int d, int o) {
    constexpr int st[3] = {1, G::NG, G::NG * G::NG};
    return G::idx(i, j, k) + o * st[d];
  }
};
template <int D, int NC, typename T>
struct Lay : LayStd<D, NC> {};
template <int D, int NC>
struct Lay<D, NC, float> {
  using G = Geo<D, NC>;
  static constexpr bool SPLIT = true;
  static constexpr int HALF = G::NINT / 2;  // interior cells of one colour
  static constexpr int STRIDE = G::NCELL;
  __host__ __device__ static int in_at(int i, int j, int k) {
    const int q = (i - 1) + NC * ((j - 1) + NC * (D == 3 ? k - 1 : 0));
    return ((i + j + k) & 1) * HALF + (q >> 1);
  }
  // a ghost cell: its afivo position less the interior cells before it
  __host__ __device__ static int ghost_at(int i, int j, int k) {
    auto before = [](int v) { return v < 1 ? 0 : (v > NC ? NC : v - 1); };
    auto inside = [](int v) { return v >= 1 && v <= NC; };
    int n = before(j) * NC + (inside(j) ? before(i) : 0);
    if (D == 3) n = before(k) * NC * NC + (inside(k) ? n : 0);
    return G::NINT + G::idx(i, j, k) - n;
  }
  __host__ __device__ static int at(int i, int j, int k) {
    const bool in = i >= 1 && i <= NC && j >= 1 && j <= NC && (D == 2 || (k >= 1 && k <= NC));
    return in ? in_at(i, j, k) : ghost_at(i, j, k);
  }
  __host__ __device__ static int nb_at(int i, int j, int k, int d, int o) {
    constexpr int st[3] = {1, NC, NC * NC};
    int p[3] = {i, j, k};
    p[d] += o;
    if (p[d] < 1 || p[d] > NC) return ghost_at(p[0], p[1], p[2]);
    const int q = (i - 1) + NC * ((j - 1) + NC * (D == 3 ? k - 1 : 0));
    return (1 - ((i + j + k) & 1)) * HALF + ((q + o * st[d]) >> 1);
  }
};

// Interior cell number q of red-black colour c ((i + j + k) & 1 == c), in
// lexicographic order: row by row, every other cell (cells 2q and 2q + 1 of
// the box in lexicographic order are neighbours in one row, one of each
// colour).  Lay<D, NC, float>::in_at of this cell is c * NINT / 2 + q.
template <int D, int NC>
__host__ __device__ inline void colour_cell(int c, int q, int &i, int &j, int &k) {
  const int p = 2 * q;
  i = p % NC + 1;
  j = (p / NC) % NC + 1;
  k = (D == 3) ? p / (NC * NC) + 1 : 0;
  if (((i + j + k) & 1) != c) i += 1;
}

// Position of the neighbour (di, dj, dk) in DBox::nbmat, the reference's
// neighbor_mat(-1:1, -1:1[, -1:1]) in column-major order (3^D entries).
template <int D>
__host__ __device__ inline int nbmat_at(int di, int dj, int dk) {
  return (di + 1) + 3 * (dj + 1) + (D == 3 ? 9 * (dk + 1) : 0);
}

__host__ __device__ inline void tangential(int D, int d, int &t0, int &t1) {
  if (D == 2) {
    t0 = 1 - d;
    t1 = -1;
  } else {
    t0 = (d == 0) ? 1 : 0;
    t1 = (d == 2) ? 1 : 2;
  }
}

// Lookup: LT_get_loc / LT_get_col_at_loc (m_lookup_table.f90:330-406).
__device__ inline void lt_loc(const DTable &t, double x, int &ix, double &w) {
  double frac = (x - t.x_min) * t.inv_fac;
  if (frac <= 0) {
    ix = 1;
    w = 1;
  } else if (frac >= t.n - 1) {
    ix = t.n - 1;
    w = 0;
  } else {
    ix = (int)ceil(frac);
    w = ix - frac;
  }
}
__device__ inline double lt_at(const DTable &t, int c, int ix, double w) {
  const double *y = t.cols + (size_t)c * t.n;
  return w * y[ix - 1] + (1 - w) * y[ix];
}
// The same in the arithmetic type TF (identical to the above for double).
template <typename TF>
__device__ inline void lt_loc_t(const DTable &t, TF x, int &ix, TF &w) {
  TF frac = (x - (TF)t.x_min) * (TF)t.inv_fac;
  if (frac <= 0) {
    ix = 1;
    w = 1;
  } else if (frac >= t.n - 1) {
    ix = t.n - 1;
    w = 0;
  } else {
    ix = (int)ceil(frac);
    w = ix - frac;
  }
}
template <typename TF>
__device__ inline TF lt_at_t(const DTable &t, int c, int ix, TF w) {
  const double *y = t.cols + (size_t)c * t.n;
  return w * (TF)y[ix - 1] + (1 - w) * (TF)y[ix];
}

// Limiters (afivo/src/m_af_limiters.f90).
template <typename T>
__device__ inline T lim_gminmod(T a, T b, T theta) {
  if (a * b > 0) {
    T x = fabs(theta * a), y = fabs(theta * b), z = fabs(T(0.5) * (a + b));
    T m = fmin(fmin(x, y), z);
    return copysign(m, a);
  }
  return T(0);
}
template <typename T>
__device__ inline T lim_apply(T a, T b, int lim) {
  return (lim == LIM_MC) ? lim_gminmod<T>(a, b, T(2.0)) : lim_gminmod<T>(a, b, T(4 / 3.0));
}
template <typename T>
__device__ inline T lim_koren(T a, T b) {
  const T third = T(1 / 3.0);
  if (sizeof(T) < sizeof(double)) {
    // single precision: densities of 1e20 square past the float range, so
    // the same tests are made on magnitudes (a^2 <= c ab with ab > 0 is
    // |a| <= c |b|)
    if (!((a > 0 && b > 0) || (a < 0 && b < 0))) return T(0);
    const T fa = fabs(a), fb = fabs(b);
    if (fa <= T(0.25) * fb) return 2 * a;
    if (fa <= T(2.5) * fb) return third * (b + 2 * a);
    return 2 * b;
  }
  T aa = a * a, ab = a * b;
  if (ab <= 0) return T(0);
  if (aa <= T(0.25) * ab) return 2 * a;
  if (aa <= T(2.5) * ab) return third * (b + 2 * a);
  return 2 * b;
}

// Atomic max/min of non-negative doubles through their bit patterns (the
// IEEE order of non-negative values equals the unsigned-integer order).
__device__ inline void atomic_max_nonneg(unsigned long long *addr, double v) {
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}
__device__ inline void atomic_min_nonneg(unsigned long long *addr, double v) {
  atomicMin(addr, (unsigned long long)__double_as_longlong(v));
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) {
  if (threadIdx.x == 0) *p = v;
}
// The same slot of every run (all stripes) := v.
__global__ void k_set_run_slots(unsigned long long *p, unsigned long long v, int n_runs) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n_runs * RED_STRIPES) p[(size_t)q * RED_STRIDE] = v;
}
// out[run][slot] := max (min for the slots in min_mask) over the stripes.
__global__ void k_fold_stripes(const unsigned long long *__restrict__ red, unsigned long long *out, int n_runs,
                               unsigned min_mask) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_runs * RED_STRIDE) return;
  const int r = q / RED_STRIDE, s = q % RED_STRIDE;
  const bool mn = (min_mask >> s) & 1u;
  const unsigned long long *p = red + (size_t)r * RED_STRIPES * RED_STRIDE + s;
  unsigned long long v = p[0];
  for (int t = 1; t < RED_STRIPES; t++) {
    const unsigned long long w = p[(size_t)t * RED_STRIDE];
    v = mn ? (w < v ? w : v) : (w > v ? w : v);
  }
  out[q] = v;
}
// ---- device-driven field solve of a forest ---------------------------------
// Per-run record of one solve: max |rhs| over leaves, the stopping
// threshold of field_compute (src/m_field.f90:478-480), the last max
// |residual|, the cycles done, and whether the run has converged.
struct RunStep {
  double max_rhs, thr, res;
  int ncyc, done;
};
__device__ inline double u64_as_double(unsigned long long u) { return __longlong_as_double((long long)u); }
// From slot 0 (max |rhs|, folded): thresholds; every run of `base` active.
__global__ void k_run_thresholds(const unsigned long long *__restrict__ red_out, const double *__restrict__ volt,
                                 const double *__restrict__ min_dr, double rel, double len,
                                 const int *__restrict__ base, int *run_host, RunStep *st, int n_runs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_runs) return;
  const double conv_fac = 1e-10, min_residual = 1e-6;
  const double mr = u64_as_double(red_out[(size_t)r * RED_STRIDE]);
  const double a = mr * rel;
  const double t1 = (min_residual < a) ? a : min_residual;
  const double t2 = conv_fac * fabs(volt[r]) / (len * min_dr[r]);
  st[r].max_rhs = mr;
  st[r].thr = (t1 < t2) ? t2 : t1;
  st[r].res = 0;
  st[r].ncyc = 0;
  st[r].done = 0;
  run_host[r] = base[r];
}
// After a cycle (slot 0: max |residual|, folded): every active run counts
// the cycle and stops once its residual is below its threshold.
__global__ void k_run_converge(const unsigned long long *__restrict__ red_out, int *run_host, RunStep *st,
                               int n_runs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_runs || !run_host[r]) return;
  const double res = u64_as_double(red_out[(size_t)r * RED_STRIDE]);
  st[r].ncyc = st[r].ncyc + 1;
  st[r].res = res;
  if (res < st[r].thr) {
    st[r].done = 1;
    run_host[r] = 0;
  }
}
__global__ void k_run_copy(const int *__restrict__ from, int *to, int n_runs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n_runs) to[r] = from[r];
}

// active[r] := host_active[r] and lo <= (highest level of run r) <= hi.
__global__ void k_runs_in_levels(const int *__restrict__ host_active, const int *__restrict__ maxlvl, int *active,
                                 int lo, int hi, int n_runs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n_runs) active[r] = host_active[r] && maxlvl[r] >= lo && maxlvl[r] <= hi;
}
// The runs that, solved alone, would restrict the children on level lvl
// into their parents in a mixed-precision close (kind 0: the run has
// leaves below some level l in lc + 1 .. lvl, or refinement-boundary boxes
// on it), or fill the ghosts of level lvl (kind 1: refinement-boundary
// boxes on lvl + 1); bits holds per run the levels with leaves (n_runs
// words) and then the levels with refinement-boundary boxes.
__global__ void k_runs_close(const int *__restrict__ host_active, const int *__restrict__ bits, int *active, int kind,
                             int lvl, int lc, int n_runs) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_runs) return;
  const int leaf = bits[r], rb = bits[n_runs + r];
  bool on = false;
  if (kind == 0) {
    for (int l = lc + 1; l <= lvl; l++) on = on || ((leaf >> (l - 1)) & 1) || ((rb >> l) & 1);
  } else {
    on = (rb >> (lvl + 1)) & 1;
  }
  active[r] = host_active[r] && on;
}
// Box lists of a forest: the id of a box whose run is not active is stored
// negated, and every list kernel skips it; active runs get their ids back.
__global__ void k_mask_ids(int *ids, int n, const int *__restrict__ run_of, const int *__restrict__ active) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int v = ids[q];
  const int id = v < 0 ? -v : v;
  if (id == 0) return;
  ids[q] = active[run_of[id - 1]] ? id : -id;
}

// ---- work decomposition of the list kernels ---------------------------------
// A workgroup holds BB groups of BS threads; each group works on one work
// item of the list (a box, or one of the `bpb` parts of a box) at a time
// and moves on by the grid's item count, so the grid may be smaller than
// the work (persistent workgroups).  BB > 1 needs BS = 64: a group is then
// one wavefront and synchronises on its own (gsync), so groups never wait
// for each other and may leave the loop at different times.  BS is a power
// of two.
template <int BS>
__device__ inline int ltid() {
  return threadIdx.x & (BS - 1);
}
// The group index is uniform over a wavefront (a group is one wavefront,
// or the whole workgroup when BB = 1); readfirstlane tells the compiler so,
// which keeps the item's box id and box record in scalar registers.
template <int BS, int BB>
__device__ inline int lgrp() {
  if (BB == 1) return 0;
  return __builtin_amdgcn_readfirstlane(threadIdx.x / BS);
}
// The first work item of workgroup blockIdx.x when the grid's workgroups
// are dealt to the eight XCDs in turn (workgroup g to XCD g mod 8): XCD x
// gets a contiguous range of items instead of every eighth one, so boxes
// that are neighbours in the list share that XCD's L2 and a face line read
// by two of them is fetched from HBM once.  A bijection of [0, gridDim.x).
__device__ inline int xcd_item() {
  const int g = blockIdx.x, n = gridDim.x, x = g & 7, k = g >> 3, per = n >> 3, rem = n & 7;
  return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + k;
}
template <int BS, int BB>
__device__ inline void gsync() {
  static_assert(BB == 1 || BS == 64, "several groups per workgroup need one wavefront per group");
  if (BB == 1) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}
// first work item of this group and the stride between its items
template <int BS, int BB>
__device__ inline int item0() {
  return blockIdx.x * BB + lgrp<BS, BB>();
}
template <int BB>
__device__ inline int item_stride() {
  return gridDim.x * BB;
}

template <int BS>
__device__ inline double block_max(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < BS / 64; w++) r = fmax(r, sh[w]);
  __syncthreads();
  return r;
}
// max / min over one group (a wavefront: shuffles only; else the
// workgroup through LDS).  The result is in every thread of the group.
template <int BS, int BB>
__device__ inline double group_max(double v, double *sh) {
  if (BS == 64) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
  }
  return block_max<BS>(v, sh);
}
template <int BS>
__device__ inline double block_min(double v, double *sh);
template <int BS, int BB>
__device__ inline double group_min(double v, double *sh) {
  if (BS == 64) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
    return v;
  }
  return block_min<BS>(v, sh);
}

template <int BS>
__device__ inline double block_min(double v, double *sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = sh[0];
  for (int w = 1; w < BS / 64; w++) r = fmin(r, sh[w]);
  __syncthreads();
  return r;
}

}  // namespace afg
