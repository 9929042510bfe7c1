// afgpu: MI355X (gfx950) implementation of afivo-streamer's per-time-step
// hot path behind the C-ABI of include/afgpu.h.
//
// Host side of the library: box pools in HBM, per-level box lists, and the
// call sequences of the reference routines (mg_fas_vcycle, mg_fas_fmg,
// field_compute, forward_euler, photoi_set_src and the time loop).  All
// work is enqueued on one HIP stream per handle; the host reads back only
// the scalars the reference's control flow branches on (max |residual|,
// time-step limits).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <array>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "afgpu.h"
#include "afgpu_direct.h"
#include "afgpu_fluid.h"
#include "afgpu_chem.h"
#include "afgpu_diag.h"
#include "afgpu_refine.h"

using namespace afg;

namespace {

thread_local std::string g_err;

void set_err(const std::string &s) { g_err = s; }

#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_err(std::string(#expr) + ": " + hipGetErrorString(e_));                  \
      return AFGPU_ERR_HIP;                                                        \
    }                                                                              \
  } while (0)

#define RCHK(expr)              \
  do {                          \
    int r_ = (expr);            \
    if (r_ != AFGPU_OK) return r_; \
  } while (0)

struct List {
  int off = 0, n = 0;
};

// Classes of list kernels that share a work-decomposition shape.
enum KClass {
  KC_GSRB = 0,   // half-sweeps
  KC_RESID,      // residuals (FP64 mixed-precision residual, set_residual)
  KC_RESTRICT,   // restriction (with residual)
  KC_PROLONG,    // prolongation, coarse rhs
  KC_GC,         // ghost layers
  KC_BOXOP,      // whole-box elementwise operations and reductions
  KC_FLUX,       // drift-diffusion flux
  KC_UPDATE,     // density update, field, sources
  KC_N
};
// Shape of a class: BB boxes (work items) per workgroup, and a persistent
// grid of per_cu workgroups per CU (0: one workgroup per BB items).
struct Shape {
  int bb = 1;
  int per_cu = 0;
};

// Timing of one kernel class: accumulated from HIP event pairs recorded on
// the library's stream around each launch (also inside captured graphs,
// where the pair becomes two event-record nodes replayed with the graph).
struct KTimer {
  double total_ms = 0;
  double algo_bytes = 0;
  long long launches = 0;
};
struct TPair {
  hipEvent_t a, b;
  int kind;
  double bytes;
  int launches;  // back-to-back launches between the two events
};
// A captured launch sequence and the timing pairs it contains.
struct GraphRec {
  hipGraphExec_t exec = nullptr;
  std::vector<TPair> pairs;
};

}  // namespace

// A reaction list on the device (afgpu_chem_create).
struct afgpu_chem {
  int n_species = 0, n_reactions = 0;
  afgpu_reaction_t *d_r = nullptr;
  double *d_fld = nullptr, *d_energy = nullptr;
  ChemArgs args{};
  hipStream_t stream = nullptr;
  ~afgpu_chem() {
    if (stream) (void)hipStreamSynchronize(stream);
    (void)hipFree(d_r);
    (void)hipFree(d_fld);
    (void)hipFree(d_energy);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct afgpu_ctx;

// A variable of the refinement's data transfers: its pool, the index whose
// methods it has, and whether it prolongs with a limiter.
struct VarRef {
  double *p;
  int methods;
  bool limited;
};

// LDS of the general update (k_update_n): each thread's densities and
// derivatives of every species, and the reaction records.
constexpr size_t UPDATE_N_LDS_MAX = 64 * 1024;
static size_t update_n_lds(int bs, int n_sp, int n_reactions) {
  return 2 * (size_t)n_sp * bs * sizeof(double) + (size_t)n_reactions * sizeof(afgpu_reaction_t);
}

struct ImplBase {
  virtual ~ImplBase() {}
  virtual int gc_lvl(int lvl, int iv, bool corners) = 0;
  virtual int vcycle(int op, bool set_residual, int max_lvl) = 0;
  virtual int fmg(int op, bool set_residual, bool have_guess) = 0;
  virtual int maxabs(int iv, double *out) = 0;
  virtual int field_compute(int s_in, bool have_guess, int *n, double *res) = 0;
  virtual int field_from_potential() = 0;
  virtual int forward_euler(double dt, int s_deriv, int n_prev, const int *s_prev, const double *w_prev,
                            int s_out, int i_step, int n_steps, double *dt_lim) = 0;
  virtual int photoi() = 0;
  virtual int copy_states(bool backup) = 0;
  virtual int refine_cells(const afgpu_refine_t &rp, const int *ids, int n_ids, int *flags) = 0;
  // boxes removed or added by refinement: restriction from the listed
  // children, prolongation into the listed new boxes, their ghost cells
  virtual int restrict_ids(const int *child_ids, int n, double *pool) = 0;
  virtual int prolong_new(const int *ids, int n, const VarRef *vars, int n_vars) = 0;
  virtual int gc_ids(const int *ids, int n, const VarRef *vars, int n_vars) = 0;
  virtual int restrict_tree(int iv) = 0;
  // the same on any FP64 pool of cell values (a reaction list's species)
  virtual int restrict_pool(double *var) = 0;
  virtual int gc_pool(int lvl, double *var, const VarBc &m, bool corners) = 0;
  virtual int step_runs(afgpu_step_info_t *info) = 0;
  // per-leaf-box terms of the regression-log statistics (k_diag_box)
  virtual int diag_boxes(const VarPtrs &vars, double *part) = 0;
};

struct afgpu_ctx {
  int D = 2, nc = 8, coord = AFGPU_XYZ, precision = AFGPU_PREC_FP64;
  size_t ncell = 0, nface = 0;  // nface: all D blocks of one box
  hipStream_t stream = nullptr;
  hipStream_t cur = nullptr;  // stream of the launch helpers: `stream`, or a mode stream inside a fork
  // photoionization modes of the mixed-precision form run as parallel
  // branches (one stream each, default priority) inside one captured graph
  bool batch_modes = true;
  hipStream_t mstream[AFGPU_MAX_MODES] = {nullptr};
  hipEvent_t fork_ev = nullptr, join_ev[AFGPU_MAX_MODES] = {nullptr};
  struct ModeRes {
    float *cc32[4] = {nullptr};  // correction, its rhs, scratch, second buffer
    double *A0 = nullptr, *A1 = nullptr;  // uniform-level solve grids
    size_t a_cap = 0;
    int cap_boxes = 0;
  } mres[AFGPU_MAX_MODES];
  // tree
  int n_boxes = 0, cap_boxes = 0, highest_lvl = 0;
  // Forest of independent runs (a parameter study): every level-1 box roots
  // a run of its own.  A plain tree is run 0.  Each run has its own applied
  // voltage, reduction slots, convergence test and time step.
  int n_runs = 1;
  std::vector<double> run_volt;    // applied voltage per run
  double *d_run_volt = nullptr;
  std::vector<double> run_min_dr;  // finest cell size per run
  int *d_run_of = nullptr;         // run of every box (index id - 1)
  int *d_run_active = nullptr;     // runs the list kernels work on
  int *d_run_host = nullptr;       // runs the host has made active
  int *h_run_active = nullptr;     // pinned staging of the above
  int *d_run_maxlvl = nullptr;     // highest level of every run
  int *d_run_bits = nullptr;       // per run: levels with leaves (bit l), then levels with refinement-boundary boxes
  bool uneven = false;             // runs of different highest levels
  bool masked = false;             // some list entries are negated
  std::vector<int> base_active;    // runs taking part in the current step (empty: all)
  int red_cap_runs = 0;            // runs the reduction buffers hold
  double *d_dt_run = nullptr, *h_dt_run = nullptr;  // per-run dt of the stage advanced
  bool use_run_dt = false;
  std::vector<double> run_dt_lim;  // per-run dt limit of the last forward_euler
  std::vector<long long> run_leaf_cells;  // leaf cells per run
  std::vector<int> run_ncyc;       // per-run V-cycles of the last field solve
  std::vector<double> run_res;     // per-run final max |residual| of that solve
  struct RunTime {
    int it = 0;
    double time = 0, global_time = 0, global_dt = 0, dt = 0, frac_rejected = 0;
    bool started = false;
  };
  std::vector<RunTime> rt;         // time-loop state per run (ensemble stepping)
  RunStep *d_step = nullptr, *h_step = nullptr;  // device-driven field solve records
  int *d_run_ones = nullptr;       // every run active
  double *d_run_min_dr = nullptr;
  int step_k = 4;                  // cycles a captured ensemble step holds
  bool step_graph = false;         // ensemble steps as one captured graph
  std::vector<DBox> hboxes;
  DBox *d_boxes = nullptr;
  std::vector<int> h_lists;
  int *d_lists = nullptr;
  int lists_bb = 1;  // items per workgroup slot the list order is made for
  std::vector<std::vector<int>> src_ids, src_leaves, src_parents;  // per level, id order
  std::vector<int> src_refb;
  // per level l: the boxes across a refinement-boundary face of a level-(l+1)
  // box (the coarse neighbours whose stored ghost layer mg_sides_rb reads)
  std::vector<std::vector<int>> src_rbnb;
  std::vector<List> l_ids, l_leaves, l_parents, l_rbnb;
  std::vector<char> lvl_rb;  // level has boxes with a refinement-boundary face
  List l_all_leaves, l_all_ids, l_refb;
  // pools
  double *cc[AFGPU_N_CC] = {nullptr};
  double *fc[AFGPU_N_FC] = {nullptr};
  float *cc32[4] = {nullptr};  // mixed precision: correction, residual, scratch, second correction buffer
  // tables
  double *d_td = nullptr, *d_chem = nullptr;
  DTable td{}, chem{};
  bool have_tables = false;
  // settings
  afgpu_params_t p{};
  bool have_params = false;
  double voltage = 0;
  struct Op {
    double lambda2 = 0;
    bool ok = false;
    double *d_vinv = nullptr, *d_v = nullptr, *d_eig = nullptr;
  } ops[1 + AFGPU_MAX_MODES];
  // mixed precision: exact solve on the finest uniformly refined level
  int direct_max_n = 128;
  int mixed_smooth = 0;  // > 0: smoothing sweeps of the inner cycle (down = up)
  int block_smooth = 0;  // inner cycle: block smoother on levels with one workgroup per box
  bool direct_mfma = true;  // uniform-level transforms on the matrix cores
  bool photoi_inner_vcycle = false;  // photoionization inner cycle: V-cycle instead of FMG
  bool flux32 = true;  // mixed precision: FP32 flux stencil (FP64 accumulation)
  bool box_lds = false;  // mixed precision, 3D 16^3 boxes: whole boxes staged in LDS by the half-sweeps
  bool fuse_direct = false;  // uniform-level solve: last-axis transforms fused into one pass
  bool direct_f32 = false;   // uniform-level solve of the FP32 inner cycle in FP32 (else FP64; matrix cores)
  bool direct_fold = true;   // uniform-level solve: reflection-symmetric axes transformed by parity (half the products)
  // refinement flags: the boxes evaluated and their flags (device)
  // a reaction list in place of the standard model; with more species than
  // the three pools (or other charges), the species beyond the pools live
  // in xsp[(k - 3) * 3 + state] and the update is k_update_n
  const afgpu_chem *net = nullptr;
  int n_sp = 3;
  bool gen_sp = false;
  std::vector<int> sp_charge{-1, 1, -1};
  std::vector<double *> xsp;
  SpeciesArgs species_args() const {
    SpeciesArgs S{};
    S.n = n_sp;
    for (int k = 0; k < n_sp; k++) {
      S.charge[k] = sp_charge[k];
      for (int st = 0; st < 3; st++) S.sp[k][st] = k < 3 ? cc[AFGPU_CC_SPECIES(k) + st] : xsp[(k - 3) * 3 + st];
    }
    return S;
  }
  int *d_ref_ids = nullptr, *d_ref_flags = nullptr;
  int ref_cap = 0;
  // diagnostics: per-box terms and per-run results; box I/O staging
  double *d_diag = nullptr, *d_io = nullptr;
  size_t diag_cap = 0, io_cap = 0;
  // this GPU's address of the static scratch array (noalloc_scratch), looked
  // up once per device at afgpu_create
  double *scratch = nullptr;
  int photoi_cycles = 0;  // FMG cycles of the Helmholtz modes in the last photoionization solve
  std::vector<int> run_pcyc;  // the same per run of a forest
  struct Direct {
    bool set = false;  // lc and the map below are current for the tree
    int lc = 1, N = 0;
    int *d_map = nullptr;
    double *A0 = nullptr, *A1 = nullptr;
    size_t cap = 0;  // doubles in A0 / A1
    struct F {
      bool ok = false;
      double *vinv = nullptr, *v = nullptr, *vinvT = nullptr, *vT = nullptr, *eig = nullptr;
      float *vinv32 = nullptr, *v32 = nullptr, *eig32 = nullptr;  // (rounded copies)
      // parity-folded factors (k_lc_mfma FOLD): per axis N x N/2 forward,
      // N/2 x N backward, the eigenvalues in the folded modes' order
      bool fold[3] = {false, false, false};
      double *mf = nullptr, *bf = nullptr, *eigf = nullptr;
    } f[1 + AFGPU_MAX_MODES];
  } dir;
  // reductions
  unsigned long long *d_red = nullptr, *h_red = nullptr;  // [run][stripe][slot], host: [run][slot]
  unsigned long long *d_red_out = nullptr;                // stripes folded, [run][slot]
  // time loop
  int it = 0;
  double time = 0, global_time = 0, global_dt = 0, frac_rejected = 0, dt = 0;
  bool started = false;
  // timing
  bool timing = false;
  bool class_timing = false;  // time every list-kernel launch under 3 + its class
  KTimer timers[3 + KC_N];
  std::vector<hipEvent_t> ev_pool;  // events of eager launches, recycled
  size_t ev_used = 0;
  std::vector<TPair> pending;       // recorded, not yet read
  std::vector<TPair> *capture_pairs = nullptr;
  // work-decomposition shapes of the list-kernel classes (set from the
  // box size at creation, chosen by afgpu_autotune) and the CU count
  Shape shape[KC_N];
  int n_cu = 256;
  // captured launch sequences (HIP graphs), keyed by kind/operator/flags
  bool use_graphs = true;
  std::map<std::array<int, 4>, GraphRec> graphs;
  std::unique_ptr<ImplBase> impl;

  int n_states() const { return p.time_integrator == AFGPU_HEUNS_METHOD ? 2 : 1; }
  const int *list(const List &l) const { return d_lists + l.off; }
};

namespace {

VarBc var_methods(const afgpu_ctx &c, int iv, bool homogeneous = false) {
  VarBc m{};
  m.limiter = (c.D < 3) ? LIM_MC : LIM_GMINMOD43;
  for (int nb = 0; nb < 6; nb++) {
    m.type[nb] = BC_NEUMANN;
    m.val[nb] = 0.0;
  }
  const int last = c.D - 1;
  if (iv == AFGPU_CC_PHI) {
    m.rb = RB_MG;
    m.type[2 * last] = BC_DIRICHLET;
    m.type[2 * last + 1] = BC_DIRICHLET;
    m.val[2 * last + 1] = homogeneous ? 0.0 : c.voltage;
    if (!homogeneous && c.n_runs > 1) {  // per-run applied voltage
      m.run_face = 2 * last + 1;
      m.run_val = c.d_run_volt;
    }
  } else if (iv >= AFGPU_CC_HELMH && iv < AFGPU_CC_HELMH + AFGPU_MAX_MODES) {
    m.rb = RB_MG;
    m.type[2 * last] = BC_DIRICHLET;
    m.type[2 * last + 1] = BC_DIRICHLET;
  } else if (iv == AFGPU_CC_EFIELD) {
    m.rb = RB_INTERP;
  } else if (iv == AFGPU_CC_PHOTO) {
    m.rb = RB_INTERP;
    m.type[2 * last] = BC_DIRICHLET;
    m.type[2 * last + 1] = BC_DIRICHLET;
  } else if (iv == AFGPU_CC_RHS) {
    m.rb = RB_INTERP;  // af_bc_neumann_zero + af_gc_interp (src/streamer.f90:103-108)
  } else if (iv >= 0 && iv < 9) {
    m.rb = RB_INTERP_LIM;
    if (c.p.species_bc == AFGPU_SPECIES_BC_DIRICHLET_ZERO) {
      m.type[2 * last] = BC_DIRICHLET_COPY;
      m.type[2 * last + 1] = BC_DIRICHLET_COPY;
    }
  } else {
    m.rb = RB_NONE;
  }
  return m;
}

// Device lists, stored XCD-aware: the GPU deals consecutive workgroups to
// its 8 XCDs in turn, each XCD with its own L2, and a workgroup takes B
// consecutive items, so item w runs on XCD (w / B) % 8.  The id-ordered
// list v is permuted so that each XCD works through a contiguous run of
// boxes (siblings and near neighbours), whose shared faces stay in its L2.
void build_lists(afgpu_ctx &h, int B) {
  std::vector<int> &all = h.h_lists;
  all.clear();
  auto add = [&](const std::vector<int> &v) {
    constexpr int X = 8;
    List l;
    l.off = (int)all.size();
    l.n = (int)v.size();
    const int n = l.n;
    const int nslots = (n + B - 1) / B;
    int start[X + 1] = {0};
    for (int J = 0; J < nslots; J++) start[J % X + 1] += std::min(B, n - J * B);
    for (int x = 0; x < X; x++) start[x + 1] += start[x];
    for (int w = 0; w < n; w++) {
      const int J = w / B;
      all.push_back(v[start[J % X] + (J / X) * B + w % B]);
    }
    return l;
  };
  const int hl = (int)h.src_ids.size() - 1;
  h.l_ids.assign(hl + 1, List());
  h.l_leaves.assign(hl + 1, List());
  h.l_parents.assign(hl + 1, List());
  h.l_rbnb.assign(hl + 1, List());
  std::vector<int> all_leaves, all_ids;
  for (int l = 1; l <= hl; l++) {
    h.l_ids[l] = add(h.src_ids[l]);
    h.l_leaves[l] = add(h.src_leaves[l]);
    h.l_parents[l] = add(h.src_parents[l]);
    h.l_rbnb[l] = add(h.src_rbnb[l]);
    all_leaves.insert(all_leaves.end(), h.src_leaves[l].begin(), h.src_leaves[l].end());
    all_ids.insert(all_ids.end(), h.src_ids[l].begin(), h.src_ids[l].end());
  }
  h.l_all_leaves = add(all_leaves);
  h.l_all_ids = add(all_ids);
  h.l_refb = add(h.src_refb);
  h.lists_bb = B;
}

enum GraphKind {
  G_VCYCLE = 1,
  G_FMG = 2,
  G_FFP = 3,
  G_MP_RES = 4,
  G_MP_VCYCLE = 5,
  G_MP_GC = 6,
  G_MP_FMG = 7,
  G_MP_RES_MODES = 8,
  G_MP_FMG_MODES = 9,
  G_STEP_RUNS = 10
};
// reduction slots (per run): 0 residual max, 1-3 dt limits (3: a minimum),
// 8 + n photoionization mode n
constexpr int N_RED_SLOTS = RED_STRIDE;
constexpr unsigned RED_MIN_SLOTS = 1u << 3;
constexpr int RED_MODE = 8;

int op_phi(int op) { return op == AFGPU_OP_POISSON ? AFGPU_CC_PHI : AFGPU_CC_HELMH + (op - AFGPU_OP_HELMH); }

int timer_start(afgpu_ctx &c, int kind, double bytes, int launches = 1) {
  if (!c.timing) return AFGPU_OK;
  TPair tp{nullptr, nullptr, kind, bytes, launches};
  if (c.capture_pairs) {  // events owned by the graph being captured
    HIPCHK(hipEventCreate(&tp.a));
    HIPCHK(hipEventCreate(&tp.b));
    c.capture_pairs->push_back(tp);
  } else {
    while (c.ev_pool.size() < c.ev_used + 2) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      c.ev_pool.push_back(e);
    }
    tp.a = c.ev_pool[c.ev_used];
    tp.b = c.ev_pool[c.ev_used + 1];
    c.ev_used += 2;
    c.pending.push_back(tp);
  }
  HIPCHK(hipEventRecord(tp.a, c.cur));
  return AFGPU_OK;
}
// A pair of events that the kernel dispatch itself records
// (hipExtLaunchKernel): they bracket exactly the kernel's execution.
int timer_events(afgpu_ctx &c, int kind, double bytes, hipEvent_t *a, hipEvent_t *b) {
  while (c.ev_pool.size() < c.ev_used + 2) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    c.ev_pool.push_back(e);
  }
  TPair tp{c.ev_pool[c.ev_used], c.ev_pool[c.ev_used + 1], kind, bytes, 1};
  c.ev_used += 2;
  c.pending.push_back(tp);
  *a = tp.a;
  *b = tp.b;
  return AFGPU_OK;
}
int timer_stop(afgpu_ctx &c, int kind) {
  if (!c.timing) return AFGPU_OK;
  const TPair &tp = c.capture_pairs ? c.capture_pairs->back() : c.pending.back();
  (void)kind;
  HIPCHK(hipEventRecord(tp.b, c.cur));
  return AFGPU_OK;
}
// Accumulate the pairs recorded so far (the stream has been synchronised).
int timer_harvest(afgpu_ctx &c) {
  for (const TPair &tp : c.pending) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, tp.a, tp.b));
    KTimer &t = c.timers[tp.kind];
    t.total_ms += ms;
    t.algo_bytes += tp.bytes;
    t.launches += tp.launches;
  }
  c.pending.clear();
  c.ev_used = 0;
  return AFGPU_OK;
}
void drop_graphs(afgpu_ctx &c) {
  for (auto &kv : c.graphs) {
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    for (auto &tp : kv.second.pairs) {
      (void)hipEventDestroy(tp.a);
      (void)hipEventDestroy(tp.b);
    }
  }
  c.graphs.clear();
}
// Run `body` (a sequence of kernel launches on c.stream with no host
// synchronisation and no memset/memcpy) through a HIP graph captured on its
// first use and replayed afterwards.
template <typename F>
int run_graph(afgpu_ctx &c, std::array<int, 4> key, F body) {
  // HIP does not report elapsed times of events recorded by graph nodes, so
  // kernel timing runs the same sequences eagerly.
  if (!c.use_graphs || c.timing) return body();
  auto it = c.graphs.find(key);
  if (it == c.graphs.end()) {
    GraphRec rec;
    HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    c.capture_pairs = &rec.pairs;
    const int rc = body();
    c.capture_pairs = nullptr;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c.stream, &g);
    if (rc != AFGPU_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      for (auto &tp : rec.pairs) {
        (void)hipEventDestroy(tp.a);
        (void)hipEventDestroy(tp.b);
      }
      if (rc != AFGPU_OK) return rc;
      HIPCHK(e);
    }
    const hipError_t ei = hipGraphInstantiate(&rec.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ei);
    it = c.graphs.emplace(key, rec).first;
  }
  HIPCHK(hipGraphLaunch(it->second.exec, c.stream));
  if (c.timing) c.pending.insert(c.pending.end(), it->second.pairs.begin(), it->second.pairs.end());
  return AFGPU_OK;
}

template <int D, int NC, int BS>
struct Impl : ImplBase {
  using G = Geo<D, NC>;
  afgpu_ctx &c;
  explicit Impl(afgpu_ctx &cc_) : c(cc_) {}

  template <typename K, typename... Args>
  int launch(K kernel, int nblocks, Args... args) {
    if (nblocks <= 0) return AFGPU_OK;
    hipLaunchKernelGGL(kernel, dim3(nblocks), dim3(BS), 0, c.cur, args...);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }
  // launch() with the kernel's own execution timed under class `kind`
  // when kernel timing is on (eager runs only)
  template <typename K, typename... Args>
  int launch_timed(int kind, double bytes, K kernel, int nblocks, Args... args) {
    if (!c.timing || c.capture_pairs) return launch(kernel, nblocks, args...);
    if (nblocks <= 0) return AFGPU_OK;
    hipEvent_t a, b;
    RCHK(timer_events(c, kind, bytes, &a, &b));
    hipExtLaunchKernelGGL(kernel, dim3(nblocks), dim3(BS), 0, c.cur, a, b, 0, args...);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }

  // Run f with the boxes-per-workgroup count of class cls as a constant
  // (several boxes per workgroup only when one box is one wavefront).
  template <typename F>
  int with_bb(int cls, F f) {
    if constexpr (BS == 64) {
      switch (c.shape[cls].bb) {
        case 2: return f(std::integral_constant<int, 2>());
        case 4: return f(std::integral_constant<int, 4>());
        case 8: return f(std::integral_constant<int, 8>());
        default: return f(std::integral_constant<int, 1>());
      }
    } else {
      return f(std::integral_constant<int, 1>());
    }
  }
  int grid_for(int cls, int nw, int bb) const {
    int g = (nw + bb - 1) / bb;
    const int cap = c.shape[cls].per_cu * c.n_cu;
    return cap > 0 ? std::min(g, cap) : g;
  }
  // A list kernel (template argument BB) over nw work items, in the shape of
  // its class; the item count is the kernel's last argument.
  template <int BB, typename K, typename... Args>
  int items(int cls, K kernel, int nw, Args... args) {
    if (c.class_timing && c.timing && !c.capture_pairs) return items_timed<BB>(3 + cls, 0.0, cls, kernel, nw, args...);
    if (nw <= 0) return AFGPU_OK;
    hipLaunchKernelGGL(kernel, dim3(grid_for(cls, nw, BB)), dim3(BS * BB), 0, c.cur, args..., nw);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }
  // items() with the kernel's own execution timed under `kind` when kernel
  // timing is on (eager runs only; under the kernel class instead while the
  // classes are timed)
  template <int BB, typename K, typename... Args>
  int items_timed(int kind, double bytes, int cls, K kernel, int nw, Args... args) {
    return items_nt<BS * BB, BB>(kind, bytes, cls, kernel, nw, args...);
  }
  // items_timed() of a kernel with NT threads per workgroup
  template <int NT, int BB, typename K, typename... Args>
  int items_nt(int kind, double bytes, int cls, K kernel, int nw, Args... args) {
    if (!c.timing || c.capture_pairs) {
      if (nw <= 0) return AFGPU_OK;
      hipLaunchKernelGGL(kernel, dim3(grid_for(cls, nw, BB)), dim3(NT), 0, c.cur, args..., nw);
      HIPCHK(hipGetLastError());
      return AFGPU_OK;
    }
    if (nw <= 0) return AFGPU_OK;
    if (c.class_timing) kind = 3 + cls;
    hipEvent_t a, b;
    RCHK(timer_events(c, kind, bytes, &a, &b));
    hipExtLaunchKernelGGL(kernel, dim3(grid_for(cls, nw, BB)), dim3(NT), 0, c.cur, a, b, 0, args..., nw);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }

  // Workgroups per box: 1 on large levels; on levels with few boxes the
  // box's work items are spread over several workgroups.
  static int bpb(int n_boxes, int items) {
    if (n_boxes >= 256) return 1;
    int b = (512 + n_boxes - 1) / n_boxes;
    const int cap = std::max(1, items / 64);
    return std::max(1, std::min(b, cap));
  }

  template <typename T>
  int boxop(const List &l, const T *a, T *b, int op, T s = T(0)) {
    constexpr int STRIDE = Lay<D, NC, T>::STRIDE;
    const int k = bpb(l.n, STRIDE * (int)sizeof(T) / 16);  // (16-byte words)
    return with_bb(KC_BOXOP, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      return items<bb>(KC_BOXOP, k_boxop<STRIDE, BS, T, bb>, l.n * k, c.list(l), a, b, op, s, k);
    });
  }

  template <typename T>
  int gc_var(int lvl, T *var, const VarBc &m, bool corners) {
    return gc_list(c.l_ids[lvl], var, m, corners);
  }
  // the ghost layer of the boxes of list l (all of one level)
  template <typename T>
  int gc_list(const List &l, T *var, const VarBc &m, bool corners) {
    const int k = corners ? 1 : bpb(l.n, 2 * D * G::NTAN);
    if (k == 1)
      return with_bb(KC_GC, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_GC, k_gc_lds<D, NC, BS, T, bb>, l.n, c.list(l), (const DBox *)c.d_boxes, var, m,
                         (int)corners);
      });
    return items<1>(KC_GC, k_gc<D, NC, BS, T, 1>, l.n * k, c.list(l), (const DBox *)c.d_boxes, var, m, (int)corners, k);
  }
  int gc_lvl(int lvl, int iv, bool corners) override { return gc_var(lvl, c.cc[iv], var_methods(c, iv), corners); }
  int gc_tree(int iv) {
    for (int lvl = 1; lvl <= c.highest_lvl; lvl++) RCHK(gc_lvl(lvl, iv, true));
    return AFGPU_OK;
  }

  // --- multigrid (afivo/src/m_af_multigrid.f90) ---------------------------
  // The variables one multigrid solve works on.  FP64: the operator's own
  // phi with the shared rhs/tmp.  Mixed precision: the single-precision
  // correction e, its rhs (the FP64 residual, rounded) and scratch, with the
  // boundary conditions made homogeneous.
  template <typename T>
  struct Mg {
    int op;
    T *phi, *rhs, *tmp;
    VarBc m;
    T lambda2;
    T *alt;  // second buffer of the block smoother (single precision only)
  };
  Mg<double> mg64(int op) {
    const int iphi = op_phi(op);
    return {op, c.cc[iphi], c.cc[AFGPU_CC_RHS], c.cc[AFGPU_CC_TMP], var_methods(c, iphi), c.ops[op].lambda2,
            nullptr};
  }
  // single-precision pools of an operator: the photoionization modes each
  // have their own, so that they can be solved side by side
  float *f32(int op, int k) const {
    return op >= AFGPU_OP_HELMH ? c.mres[op - AFGPU_OP_HELMH].cc32[k] : c.cc32[k];
  }
  static int red_slot(int op) { return op >= AFGPU_OP_HELMH ? RED_MODE + (op - AFGPU_OP_HELMH) : 0; }
  Mg<float> mg32(int op) {
    return {op, f32(op, 0), f32(op, 1), f32(op, 2), var_methods(c, op_phi(op), true), (float)c.ops[op].lambda2,
            f32(op, 3)};
  }

  // Levels of the single-precision inner cycle that use the block smoother:
  // their down-sweep writes the second buffer, which the restriction and the
  // coarse correction then use, and their up-sweep writes phi back.
  template <typename T>
  bool use_block(int lvl) const {
    return std::is_same<T, float>::value && c.block_smooth && bpb(c.l_ids[lvl].n, G::NINT / 2) == 1;
  }
  template <typename T>
  T *level_buf(const Mg<T> &v, int lvl, bool after_down) const {
    return (after_down && use_block<T>(lvl)) ? v.alt : v.phi;
  }

  // stored_first: the first half-sweep reads the stored ghost layer (the
  // finest level of a V-cycle called on its own, whose layer the reference
  // smoother reads as the previous operation left it)
  template <typename T>
  int gsrb_boxes(const Mg<T> &v, int lvl, bool down, bool final_gc = true, bool stored_first = false) {
    int n_cycle = down ? c.p.mg_n_cycle_down : c.p.mg_n_cycle_up;
    if (std::is_same<T, float>::value && c.mixed_smooth > 0) n_cycle = c.mixed_smooth;
    const List &l = c.l_ids[lvl];
    // algorithmic bytes of a half-sweep: the other colour and the rhs of
    // the updated colour are read, the updated colour is written
    const double bytes = 1.5 * sizeof(T) * l.n * G::NINT;
    const int k = bpb(l.n, G::NINT / 2);
    const int cyl = c.coord == AFGPU_CYL;
    if (use_block<T>(lvl)) {
      T *in = down ? v.phi : v.alt, *out = down ? v.alt : v.phi;
      // algorithmic bytes of a block phase: in and rhs read, out written
      RCHK(launch_timed(0, 3.0 * sizeof(T) * l.n * G::NINT, k_smooth_block<D, NC, BS, T>, l.n, c.list(l),
                        (const DBox *)c.d_boxes, (const T *)in, out, (const T *)v.rhs, (const T *)v.phi, 2 * n_cycle,
                        v.lambda2, cyl, v.m));
      return final_gc ? gc_var(lvl, out, v.m, !down) : AFGPU_OK;
    }
    for (int n = 1; n <= 2 * n_cycle; n++) {
      const int stored = (stored_first && n == 1 && !Lay<D, NC, T>::SPLIT) ? 1 : 0;
      if constexpr (D == 2 && Lay<D, NC, T>::SPLIT && NC == 8) {
        if (k == 1 && !c.box_lds) {  // (a wavefront per eight boxes, four wavefronts per workgroup)
          RCHK((items_nt<256, 1>(0, bytes, KC_GSRB, k_gsrb_split2d<NC>, (l.n + 31) / 32, c.list(l),
                                 (const DBox *)c.d_boxes, v.phi, (const T *)v.rhs, n, v.lambda2, cyl, v.m, l.n)));
          continue;
        }
      }
      if constexpr (D == 3 && Lay<D, NC, T>::SPLIT && BS == NC * NC) {
        if (k == 1 && !c.box_lds) {
          RCHK((items_nt<GSRB_SPLIT_NT, 1>(0, bytes, KC_GSRB, k_gsrb_split<D, NC, GSRB_SPLIT_NT>, l.n, c.list(l),
                                           (const DBox *)c.d_boxes, v.phi, (const T *)v.rhs, n, v.lambda2, v.m)));
          continue;
        }
      }
      if (k == 1)
        RCHK(with_bb(KC_GSRB, [&](auto B) {
          constexpr int bb = decltype(B)::value;
          return items_timed<bb>(0, bytes, KC_GSRB, k_gsrb_lds<D, NC, BS, T, bb>, l.n, c.list(l),
                                 (const DBox *)c.d_boxes, v.phi, (const T *)v.rhs, n, v.lambda2, cyl, v.m, stored);
        }));
      else
        RCHK(with_bb(KC_GSRB, [&](auto B) {
          constexpr int bb = decltype(B)::value;
          return items_timed<bb>(0, bytes, KC_GSRB, k_gsrb<D, NC, BS, T, bb>, l.n * k, c.list(l),
                                 (const DBox *)c.d_boxes, v.phi, (const T *)v.rhs, n, v.lambda2, cyl, v.m, stored,
                                 k);
        }));
    }
    // ghost layer after the last half-sweep (corners on the final up-sweep)
    return final_gc ? gc_var(lvl, v.phi, v.m, !down) : AFGPU_OK;
  }

  // The single-precision inner cycle restricts with the child box staged in
  // LDS and its face ghosts formed in-kernel, so it needs no ghost fill
  // after the down-sweep.  (The FP64 form keeps the reference's fills: its
  // stored ghost layers are part of the state compared with the oracle.)
  template <typename T>
  bool restrict_in_lds(int lvl) const {
    constexpr int NQ = G::NINT / (1 << D);
    return std::is_same<T, float>::value && bpb(c.l_ids[lvl].n, NQ) == 1;
  }

  template <typename T>
  int update_coarse(const Mg<T> &v, int lvl, bool v_cycle) {
    const int cyl = c.coord == AFGPU_CYL;
    const List &l = c.l_ids[lvl];
    constexpr int NQ = G::NINT / (1 << D);
    const int k = bpb(l.n, NQ);
    if (restrict_in_lds<T>(lvl)) {
      RCHK(with_bb(KC_RESTRICT, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_RESTRICT, k_restrict_residual_lds<D, NC, BS, T, bb>, l.n, c.list(l),
                         (const DBox *)c.d_boxes, (const T *)level_buf(v, lvl, v_cycle), v.phi, (const T *)v.rhs, v.tmp,
                         v.lambda2, cyl, (int)!v_cycle, v.m);
      }));
    } else {
      // (a forest whose runs end on different levels fills every level: a
      // run's own finest level needs it, and for the other runs the fill
      // repeats the one the level above has just made, with the same values)
      if (!v_cycle && (lvl == c.highest_lvl || c.uneven)) RCHK(gc_var(lvl, v.phi, v.m, true));
      RCHK(with_bb(KC_RESTRICT, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_RESTRICT, k_restrict_residual<D, NC, BS, T, bb>, l.n * k, c.list(l),
                         (const DBox *)c.d_boxes, v.phi, (const T *)v.rhs, v.tmp, v.lambda2, cyl, (int)!v_cycle, k);
      }));
    }
    RCHK(gc_var(lvl - 1, v.phi, v.m, true));
    const List &pl = c.l_parents[lvl - 1];
    const int k2 = bpb(pl.n, G::NCELL);
    return with_bb(KC_PROLONG, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      return items<bb>(KC_PROLONG, k_coarse_rhs<D, NC, BS, T, bb>, pl.n * k2, c.list(pl), (const DBox *)c.d_boxes,
                       (const T *)v.phi, v.rhs, v.tmp, v.lambda2, cyl, (int)v_cycle, k2);
    });
  }

  // (target: the buffer holding the children's level, see use_block)
  template <typename T>
  int correct_children(const Mg<T> &v, int lvl_p, T *target = nullptr) {
    RCHK(boxop<T>(c.l_parents[lvl_p], v.phi, v.tmp, OP_SUB_FROM));
    const List &l = c.l_ids[lvl_p + 1];
    const int k = bpb(l.n, G::NINT);
    return with_bb(KC_PROLONG, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      if constexpr (Lay<D, NC, T>::SPLIT)
        return items<bb>(KC_PROLONG, k_prolong_add_split<D, NC, BS, bb>, l.n * k, c.list(l), (const DBox *)c.d_boxes,
                         (const T *)v.tmp, target ? target : v.phi, k);
      else
        return items<bb>(KC_PROLONG, k_prolong_add<D, NC, BS, T, bb>, l.n * k, c.list(l), (const DBox *)c.d_boxes,
                         (const T *)v.tmp, target ? target : v.phi, k);
    });
  }

  int ensure_factors(int op) {
    afgpu_ctx::Op &o = c.ops[op];
    if (o.ok) return AFGPU_OK;
    const DBox &b1 = c.hboxes[c.h_lists[c.l_ids[1].off] - 1];
    int bcd[6] = {0, 0, 0, 0, 0, 0};
    bcd[2 * (D - 1)] = 1;
    bcd[2 * (D - 1) + 1] = 1;
    const size_t nn = (size_t)D * NC * NC;
    std::vector<double> vinv(nn), v(nn), eig((size_t)D * NC);
    RCHK(afgpu_coarse_factors(D, NC, c.coord, b1.r_min, b1.dr, bcd, vinv.data(), v.data(), eig.data()));
    if (!o.d_vinv) {
      HIPCHK(hipMalloc(&o.d_vinv, nn * sizeof(double)));
      HIPCHK(hipMalloc(&o.d_v, nn * sizeof(double)));
      HIPCHK(hipMalloc(&o.d_eig, (size_t)D * NC * sizeof(double)));
    }
    HIPCHK(hipMemcpy(o.d_vinv, vinv.data(), nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(o.d_v, v.data(), nn * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(o.d_eig, eig.data(), (size_t)D * NC * sizeof(double), hipMemcpyHostToDevice));
    o.ok = true;
    return AFGPU_OK;
  }

  // The level-1 solve always runs in FP64 (single-precision storage is
  // widened on load and rounded on store).
  template <typename T>
  int solve_coarse(const Mg<T> &v) {
    const List &l1 = c.l_ids[1];  // one box per run
    const afgpu_ctx::Op &o = c.ops[v.op];
    hipLaunchKernelGGL((k_coarse_solve<D, NC, 1024, T>), dim3(l1.n), dim3(1024), 0, c.cur, c.list(l1),
                       (const DBox *)c.d_boxes, (const T *)v.rhs, v.phi, v.m, o.lambda2,
                       (int)(c.coord == AFGPU_CYL), (const double *)o.d_vinv, (const double *)o.d_v,
                       (const double *)o.d_eig);
    HIPCHK(hipGetLastError());
    return gc_var(1, v.phi, v.m, true);
  }

  int vcycle(int op, bool set_residual, int max_lvl) override {
    RCHK(ensure_factors(op));  // allocation and upload happen outside any capture
    return run_graph(c, {G_VCYCLE, op, (int)set_residual, max_lvl},
                     [&] { return vcycle_body(mg64(op), set_residual, max_lvl, 1, true); });
  }

  // FAS V-cycle from max_lvl down to the coarsest level lc (1: the level-1
  // box solve of the reference; > 1: the exact uniform-level solve of the
  // mixed-precision inner cycle) and back.  standalone: called on its own
  // (not from FMG, which fills the ghost layer before each V-cycle), so the
  // first half-sweep on max_lvl reads the stored ghost layer.
  template <typename T>
  int vcycle_body(const Mg<T> &v, bool set_residual, int max_lvl, int lc = 1, bool standalone = false) {
    for (int lvl = max_lvl; lvl >= lc + 1; lvl--) {
      RCHK(gsrb_boxes(v, lvl, true, !restrict_in_lds<T>(lvl), standalone && lvl == max_lvl));
      RCHK(update_coarse(v, lvl, true));
    }
    if (lc == 1)
      RCHK(solve_coarse(v));
    else
      RCHK(solve_level(v, lc));
    // (no ghost fill between the correction and the smoother: the fused
    // half-sweeps form every out-of-box neighbour themselves, and the
    // stored layer is rewritten after the last half-sweep)
    for (int lvl = lc + 1; lvl <= max_lvl; lvl++) {
      RCHK(correct_children(v, lvl - 1, level_buf(v, lvl, true)));
      // the single-precision correction's ghost layer on the finest level
      // is read by nothing after the cycle
      const bool final_gc = !(std::is_same<T, float>::value && lvl == c.highest_lvl);
      RCHK(gsrb_boxes(v, lvl, false, final_gc));
    }
    if (set_residual) RCHK(residual_pass(v, max_lvl, true));
    return AFGPU_OK;
  }
  // residual on every box of levels 1..max_lvl; the max |residual| over
  // leaves is reduced on the fly into slot 0 of each run (read back with
  // leaf_residual_max), cleared first with `reset`
  template <typename T>
  int residual_pass(const Mg<T> &v, int max_lvl, bool reset) {
    const int cyl = c.coord == AFGPU_CYL;
    if (reset) RCHK(set_red(0, 0ull));
    for (int lvl = 1; lvl <= max_lvl; lvl++) {
      const List &l = c.l_ids[lvl];
      const int k = bpb(l.n, G::NINT);
      RCHK(with_bb(KC_RESID, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_RESID, k_residual<D, NC, BS, T, bb>, l.n * k, c.list(l), (const DBox *)c.d_boxes,
                         (const T *)v.phi, (const T *)v.rhs, v.tmp, v.lambda2, cyl, c.d_red, k);
      }));
    }
    return AFGPU_OK;
  }

  // max |tmp| over leaf interiors as reduced by the last set_residual pass
  // (the max over the runs of a forest)
  int leaf_residual_max(double *out) {
    RCHK(read_red(1));
    *out = as_double(c.h_red[0]);
    for (int r = 1; r < c.n_runs; r++) *out = std::max(*out, run_red(r, 0));
    return AFGPU_OK;
  }
  // slot of run r as read by the last read_red
  double run_red(int r, int slot) const { return as_double(c.h_red[(size_t)r * N_RED_SLOTS + slot]); }

  // The runs taking part in the current step (all, unless a rejected step
  // is being redone for some runs only).
  std::vector<int> base_set() const {
    return c.base_active.empty() ? std::vector<int>(c.n_runs, 1) : c.base_active;
  }
  // Let the list kernels work on the runs with active[r] != 0 only (null:
  // the base set): the ids of the other runs are negated in every list and
  // in the uniform-level map.  A no-op for a plain tree.
  int set_active(const std::vector<int> *active) {
    if (c.n_runs == 1) return AFGPU_OK;
    const std::vector<int> base = base_set();
    if (!active) active = &base;
    bool all = true;
    for (int r = 0; r < c.n_runs; r++) {
      c.h_run_active[r] = (*active)[r] != 0;
      all = all && c.h_run_active[r];
    }
    if (all && !c.masked) return AFGPU_OK;
    HIPCHK(hipMemcpyAsync(c.d_run_host, c.h_run_active, c.n_runs * sizeof(int), hipMemcpyHostToDevice, c.stream));
    RCHK(mask_levels(0, std::numeric_limits<int>::max()));
    // (the staging buffer is rewritten by the next call)
    HIPCHK(hipStreamSynchronize(c.stream));
    c.masked = !all;
    return AFGPU_OK;
  }
  // Device side of the masks (capturable): the list kernels work on the
  // runs made active by the host whose highest level lies in [lo, hi].
  int mask_levels(int lo, int hi) {
    hipLaunchKernelGGL(k_runs_in_levels, dim3((c.n_runs + 63) / 64), dim3(64), 0, c.cur, (const int *)c.d_run_host,
                       (const int *)c.d_run_maxlvl, c.d_run_active, lo, hi, c.n_runs);
    HIPCHK(hipGetLastError());
    const int n = (int)c.h_lists.size();
    hipLaunchKernelGGL(k_mask_ids, dim3((n + 255) / 256), dim3(256), 0, c.cur, c.d_lists, n,
                       (const int *)c.d_run_of, (const int *)c.d_run_active);
    HIPCHK(hipGetLastError());
    if (c.dir.set && c.dir.lc > 1) {
      const int nm = c.l_ids[c.dir.lc].n;
      hipLaunchKernelGGL(k_mask_ids, dim3((nm + 255) / 256), dim3(256), 0, c.cur, c.dir.d_map, nm,
                         (const int *)c.d_run_of, (const int *)c.d_run_active);
      HIPCHK(hipGetLastError());
    }
    return AFGPU_OK;
  }

  // Device side of the masks of a mode's close in a forest (capturable):
  // the runs made active by the host that would restrict level lvl (kind
  // 0) or fill the ghosts of level lvl (kind 1) when solved alone.
  int mask_close(int kind, int lvl, int lc) {
    hipLaunchKernelGGL(k_runs_close, dim3((c.n_runs + 63) / 64), dim3(64), 0, c.cur, (const int *)c.d_run_host,
                       (const int *)c.d_run_bits, c.d_run_active, kind, lvl, lc, c.n_runs);
    HIPCHK(hipGetLastError());
    const int n = (int)c.h_lists.size();
    hipLaunchKernelGGL(k_mask_ids, dim3((n + 255) / 256), dim3(256), 0, c.cur, c.d_lists, n,
                       (const int *)c.d_run_of, (const int *)c.d_run_active);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }

  int fmg(int op, bool set_residual, bool have_guess) override {
    RCHK(ensure_factors(op));
    return run_graph(c, {G_FMG, op, (int)set_residual, (int)have_guess},
                     [&] { return fmg_body(mg64(op), set_residual, have_guess); });
  }

  template <typename T>
  int fmg_body(const Mg<T> &v, bool set_residual, bool have_guess, int lc = 1) {
    if (have_guess) {
      for (int lvl = c.highest_lvl; lvl >= lc + 1; lvl--) RCHK(update_coarse(v, lvl, false));
    } else {
      for (int lvl = c.highest_lvl; lvl >= lc + 1; lvl--) {  // init_phi_rhs
        const List &l = c.l_ids[lvl];
        RCHK(boxop<T>(l, v.phi, v.phi, OP_CLEAR));
        const int k = bpb(l.n, G::NINT / (1 << D));
        RCHK(with_bb(KC_RESTRICT, [&](auto B) {
          constexpr int bb = decltype(B)::value;
          return items<bb>(KC_RESTRICT, k_restrict<D, NC, BS, T, bb>, l.n * k, c.list(l), (const DBox *)c.d_boxes,
                           (const T *)v.rhs, v.rhs, 1, (int)(c.coord == AFGPU_CYL), k);
        }));
      }
    }
    // the single-precision correction starts from zero on its coarsest level
    // too (refined boxes there keep the last cycle's values otherwise)
    if (std::is_same<T, float>::value && !have_guess) RCHK(boxop<T>(c.l_ids[lc], v.phi, v.phi, OP_CLEAR));
    RCHK(boxop<T>(c.l_ids[lc], v.phi, v.tmp, OP_COPY));
    if (c.uneven) return fmg_up_uneven(v, set_residual, lc);
    RCHK(vcycle_body(v, set_residual && c.highest_lvl == lc, lc, lc));
    for (int lvl = lc + 1; lvl <= c.highest_lvl; lvl++) {
      RCHK(boxop<T>(c.l_ids[lvl], v.phi, v.tmp, OP_COPY));
      RCHK(correct_children(v, lvl - 1));
      RCHK(vcycle_body(v, set_residual && lvl == c.highest_lvl, lvl, lc));
    }
    return AFGPU_OK;
  }
  // The upward part of FMG for a forest whose runs end on different
  // levels: the V-cycle ending on level l involves only the runs that reach
  // l (a run's FMG ends on its own finest level), and each run's residual
  // is formed after its own last V-cycle.  The masks are set on the device,
  // so the sequence can be captured.
  template <typename T>
  int fmg_up_uneven(const Mg<T> &v, bool set_residual, int lc) {
    const int top = std::numeric_limits<int>::max();
    if (set_residual) RCHK(set_red(0, 0ull));
    RCHK(vcycle_body(v, false, lc, lc));
    if (set_residual) {
      RCHK(mask_levels(lc, lc));
      RCHK(residual_pass(v, lc, false));
    }
    for (int lvl = lc + 1; lvl <= c.highest_lvl; lvl++) {
      RCHK(mask_levels(lvl, top));
      RCHK(boxop<T>(c.l_ids[lvl], v.phi, v.tmp, OP_COPY));
      RCHK(correct_children(v, lvl - 1));
      RCHK(vcycle_body(v, false, lvl, lc));
      if (set_residual) {
        RCHK(mask_levels(lvl, lvl));
        RCHK(residual_pass(v, lvl, false));
      }
    }
    return mask_levels(0, top);
  }

  // --- exact solve on a uniformly refined level (mixed precision) ---------
  // The coarsest level of the inner cycle: the highest level lc such that
  // every level up to lc covers the domain and lc has at most direct_max_n
  // cells per axis (1 when no such level above 1 exists).
  int direct_level(int *lc_out) {
    afgpu_ctx::Direct &d = c.dir;
    if (!d.set) {
      // (a forest of 2D runs is solved as a batch of equal grids, one per
      // run; 3D forests use the level-1 solve)
      int lc = 1;
      for (int l = 2; l <= c.highest_lvl && (D == 2 || c.n_runs == 1); l++) {
        const long long want = (long long)c.n_runs << (D * (l - 1));
        const int n = NC << (l - 1);
        if (c.l_ids[l].n != want || n > c.direct_max_n || n < 16) break;
        lc = l;
      }
      d.lc = lc;
      d.N = NC << (lc - 1);
      if (lc > 1) {
        const int nb = 1 << (lc - 1);
        const int nbox = c.l_ids[lc].n;
        const int per = nbox / c.n_runs;  // boxes of one run's level
        std::vector<int> map(nbox, 0);
        for (int q = 0; q < nbox; q++) {
          const int id = c.h_lists[c.l_ids[lc].off + q];
          const DBox &b = c.hboxes[id - 1];
          int pos = 0, mul = 1;
          for (int dd = 0; dd < D; dd++) {
            pos += (b.ix[dd] - 1) * mul;
            mul *= nb;
          }
          if (pos < 0 || pos >= per || b.run < 0 || b.run >= c.n_runs) pos = -1;
          if (pos >= 0) pos += b.run * per;
          if (pos < 0 || pos >= nbox || map[pos] != 0) {
            set_err("uniform level: inconsistent box positions");
            return AFGPU_ERR_STATE;
          }
          map[pos] = id;
        }
        (void)hipFree(d.d_map);
        d.d_map = nullptr;
        HIPCHK(hipMalloc(&d.d_map, nbox * sizeof(int)));
        HIPCHK(hipMemcpy(d.d_map, map.data(), nbox * sizeof(int), hipMemcpyHostToDevice));
        size_t cells = c.n_runs;
        for (int dd = 0; dd < D; dd++) cells *= (size_t)d.N;
        if (cells > d.cap) {
          (void)hipFree(d.A0);
          (void)hipFree(d.A1);
          d.A0 = d.A1 = nullptr;
          HIPCHK(hipMalloc(&d.A0, cells * sizeof(double)));
          HIPCHK(hipMalloc(&d.A1, cells * sizeof(double)));
          d.cap = cells;
        }
      }
      for (auto &f : d.f) f.ok = false;
      d.set = true;
    }
    *lc_out = d.lc;
    return AFGPU_OK;
  }

  // 1-D eigen-factors of the level-lc operator (allocation and upload, so
  // this runs before any capture).
  int ensure_direct_factors(int op) {
    afgpu_ctx::Direct &d = c.dir;
    afgpu_ctx::Direct::F &f = d.f[op];
    if (f.ok || d.lc <= 1) return AFGPU_OK;
    const int N = d.N;
    const DBox &b1 = c.hboxes[c.h_lists[c.l_ids[1].off] - 1];
    const DBox &bl = c.hboxes[c.h_lists[c.l_ids[d.lc].off] - 1];
    int bcd[6] = {0, 0, 0, 0, 0, 0};
    bcd[2 * (D - 1)] = 1;
    bcd[2 * (D - 1) + 1] = 1;
    const size_t nn = (size_t)D * N * N;
    std::vector<double> vinv(nn), v(nn), eig((size_t)D * N), vinvT(nn), vT(nn);
    RCHK(afgpu_coarse_factors(D, N, c.coord, b1.r_min, bl.dr, bcd, vinv.data(), v.data(), eig.data()));
    for (int dd = 0; dd < D; dd++)
      for (int i = 0; i < N; i++)
        for (int m = 0; m < N; m++) {
          const size_t o = (size_t)dd * N * N;
          vinvT[o + (size_t)m * N + i] = vinv[o + (size_t)i * N + m];
          vT[o + (size_t)m * N + i] = v[o + (size_t)i * N + m];
        }
    // Axes whose 1-D operator is symmetric under reversal (Cartesian, the
    // same boundary condition at both ends) have even and odd eigenvectors:
    // the folded factors list the even modes first, then the odd ones.
    const int H = N / 2;
    std::vector<double> mf((size_t)D * N * H, 0.0), bf((size_t)D * H * N, 0.0), eigf(eig);
    for (int dd = 0; dd < D; dd++) {
      f.fold[dd] = false;
      if (N != 128 || (c.coord == AFGPU_CYL && dd == 0) || bcd[2 * dd] != bcd[2 * dd + 1]) continue;
      const double *Vd = v.data() + (size_t)dd * N * N;     // V[m][i]: component m of eigenvector i
      const double *Wd = vinv.data() + (size_t)dd * N * N;  // V^-1[i][m]
      std::vector<int> even, odd;
      bool ok = true;
      for (int i = 0; i < N && ok; i++) {
        double se = 0, so = 0, nv = 0, we = 0, wo = 0, nw = 0;
        for (int m = 0; m < N; m++) {
          const double a = Vd[(size_t)m * N + i], b = Vd[(size_t)(N - 1 - m) * N + i];
          const double x = Wd[(size_t)i * N + m], y = Wd[(size_t)i * N + N - 1 - m];
          se += (a - b) * (a - b);
          so += (a + b) * (a + b);
          nv += a * a;
          we += (x - y) * (x - y);
          wo += (x + y) * (x + y);
          nw += x * x;
        }
        // (the eigenvectors come out symmetric or antisymmetric to ~1e-10:
        // the other measure is then ~4 times the norm)
        if (se <= 1e-12 * nv && we <= 1e-12 * nw)
          even.push_back(i);
        else if (so <= 1e-12 * nv && wo <= 1e-12 * nw)
          odd.push_back(i);
        else
          ok = false;
      }
      if (!ok || (int)even.size() != H) continue;
      std::vector<int> mode(even);
      mode.insert(mode.end(), odd.begin(), odd.end());
      // exactly (anti)symmetric halves: the mean of each value and its mirror
      for (int r = 0; r < N; r++) {
        const double sg = (r < H) ? 1.0 : -1.0;
        const double *w = Wd + (size_t)mode[r] * N;
        for (int m = 0; m < H; m++) mf[(size_t)dd * N * H + (size_t)r * H + m] = 0.5 * (w[m] + sg * w[N - 1 - m]);
        eigf[(size_t)dd * N + r] = eig[(size_t)dd * N + mode[r]];
      }
      for (int m = 0; m < H; m++)
        for (int r = 0; r < N; r++) {
          const double sg = (r < H) ? 1.0 : -1.0;
          bf[(size_t)dd * H * N + (size_t)m * N + r] =
              0.5 * (Vd[(size_t)m * N + mode[r]] + sg * Vd[(size_t)(N - 1 - m) * N + mode[r]]);
        }
      f.fold[dd] = true;
    }
    double **dst[8] = {&f.vinv, &f.v, &f.vinvT, &f.vT, &f.eig, &f.mf, &f.bf, &f.eigf};
    const std::vector<double> *srcs[8] = {&vinv, &v, &vinvT, &vT, &eig, &mf, &bf, &eigf};
    for (int q = 0; q < 8; q++) {
      (void)hipFree(*dst[q]);
      *dst[q] = nullptr;
      HIPCHK(hipMalloc(dst[q], srcs[q]->size() * sizeof(double)));
      HIPCHK(hipMemcpy(*dst[q], srcs[q]->data(), srcs[q]->size() * sizeof(double), hipMemcpyHostToDevice));
    }
    float **dst32[3] = {&f.vinv32, &f.v32, &f.eig32};
    const std::vector<double> *srcs32[3] = {&vinv, &v, &eig};
    for (int q = 0; q < 3; q++) {
      std::vector<float> r(srcs32[q]->begin(), srcs32[q]->end());
      (void)hipFree(*dst32[q]);
      *dst32[q] = nullptr;
      HIPCHK(hipMalloc(dst32[q], r.size() * sizeof(float)));
      HIPCHK(hipMemcpy(*dst32[q], r.data(), r.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    f.ok = true;
    return AFGPU_OK;
  }

  // The transforms of the uniform-level solve on the grids A0 / A1 (the
  // level gathered into A0; *result: the grid holding the result).  On the matrix cores
  // with a box pool given (rhs_pool, phi_pool), the first x transform reads
  // the level's rhs from the pool and the last writes phi to it (gather and
  // scatter fused), and the last forward transform, the division and the
  // first backward one run as one pass when that pass has a workgroup per
  // CU at least.
  // wavefronts per 64 x 64 tile of the matrix-core transforms
  static constexpr int LC_WAVES = 8;
  template <int N, typename T>
  int direct_transforms(const afgpu_ctx::Direct::F &f, double lambda2, double *A0, double *A1,
                        const double **result, const T *rhs_pool = nullptr, T *phi_pool = nullptr) {
    const size_t NN = (size_t)N * N;
    // the grids of all runs are transformed together: the x rows of every
    // run, and the y lines with the run as the remaining coordinate
    const int batch = (D == 3) ? N : c.n_runs;
    const int nrows = (D == 3) ? N * N : N * c.n_runs;
    constexpr int TI = (N < 64) ? N : 64;
    constexpr int TR = (256 / TI) / (N / TI) * 16;
    const int row_blocks = (nrows + TR - 1) / TR;
    constexpr int TC = (N < 64) ? N : (N == 64 ? 64 : 32);
    const int line_blocks = (N / TC) * batch;
    // rows(): the x transform (MT: transposed matrix for the small-N
    // kernel, M itself for the tiled one); lines(): the y / z transforms
    constexpr bool TILED = (N == 64 || N == 128);
    constexpr int NM = TILED ? N : 64;
    const int gemm_row_blocks = (N / 64) * (nrows / 64);
    const int gemm_line_blocks = (N / 64) * (N / 64) * batch;
    const bool MC = TILED && c.direct_mfma;
    const int mfma_row_blocks = (N / 64) * (nrows / 64);
    const int mfma_line_blocks = (N / 64) * (N / 64) * batch;
    const bool pool = MC && rhs_pool && phi_pool;
    const int nb = N / NC;
    const int *map = c.dir.d_map;
    const int *nomap = nullptr;
    const T *noin = nullptr;
    T *noout = nullptr;
    // parity-folded transforms (FOLD 1 forward, 2 backward) of the axes
    // that allow them; the eigenvalue sums then follow the folded order
    const bool folding = MC && NM == 128 && c.direct_fold && (f.fold[0] || f.fold[1] || (D == 3 && f.fold[2]));
    const double *eigs = folding ? (const double *)f.eigf : (const double *)f.eig;
    auto rows_fold = [&](const double *in, double *out, int io, int fold) {
      if constexpr (NM == 128) {
        const double *Mx = (fold == 1) ? f.mf : f.bf;
        const int blocks = (fold == 2) ? nrows / 64 : mfma_row_blocks;
        if (io == 1)
          hipLaunchKernelGGL((k_lc_mfma<D, NM, true, 1, NC, T, LC_WAVES, 1>), dim3(blocks), dim3(64 * LC_WAVES), 0,
                             c.cur, in, out, Mx, 0, 0, eigs, lambda2, map, nb, rhs_pool, noout);
        else if (io == 2)
          hipLaunchKernelGGL((k_lc_mfma<D, NM, true, 2, NC, T, LC_WAVES, 2>), dim3(blocks), dim3(64 * LC_WAVES), 0,
                             c.cur, in, out, Mx, 0, 0, eigs, lambda2, map, nb, noin, phi_pool);
        else if (fold == 1)
          hipLaunchKernelGGL((k_lc_mfma<D, NM, true, 0, NC, T, LC_WAVES, 1>), dim3(blocks), dim3(64 * LC_WAVES), 0,
                             c.cur, in, out, Mx, 0, 0, eigs, lambda2, nomap, 0, noin, noout);
        else
          hipLaunchKernelGGL((k_lc_mfma<D, NM, true, 0, NC, T, LC_WAVES, 2>), dim3(blocks), dim3(64 * LC_WAVES), 0,
                             c.cur, in, out, Mx, 0, 0, eigs, lambda2, nomap, 0, noin, noout);
      }
      return hipGetLastError();
    };
    auto lines_fold = [&](const double *in, double *out, int axis, int div, int fold) {
      if constexpr (NM == 128) {
        const double *Mx = ((fold == 1) ? f.mf : f.bf) + (size_t)axis * N * N / 2;
        const int blocks = (fold == 2) ? mfma_line_blocks / (N / 64) : mfma_line_blocks;
        if (fold == 1)
          hipLaunchKernelGGL((k_lc_mfma<D, NM, false, 0, NC, T, LC_WAVES, 1>), dim3(blocks), dim3(64 * LC_WAVES),
                             0, c.cur, in, out, Mx, axis, div, eigs, lambda2, nomap, 0, noin, noout);
        else
          hipLaunchKernelGGL((k_lc_mfma<D, NM, false, 0, NC, T, LC_WAVES, 2>), dim3(blocks), dim3(64 * LC_WAVES),
                             0, c.cur, in, out, Mx, axis, div, eigs, lambda2, nomap, 0, noin, noout);
      }
      return hipGetLastError();
    };
    auto rows = [&](const double *in, double *out, const double *Mx, const double *MTx, int io) {
      if (folding && f.fold[0]) return rows_fold(in, out, io, Mx == f.vinv ? 1 : 2);
      if (MC && io == 1)
        hipLaunchKernelGGL((k_lc_mfma<D, NM, true, 1, NC, T, LC_WAVES>), dim3(mfma_row_blocks), dim3(64 * LC_WAVES), 0, c.cur, in, out,
                           Mx, 0, 0, eigs, lambda2, map, nb, rhs_pool, noout);
      else if (MC && io == 2)
        hipLaunchKernelGGL((k_lc_mfma<D, NM, true, 2, NC, T, LC_WAVES>), dim3(mfma_row_blocks), dim3(64 * LC_WAVES), 0, c.cur, in, out,
                           Mx, 0, 0, eigs, lambda2, map, nb, noin, phi_pool);
      else if (MC)
        hipLaunchKernelGGL((k_lc_mfma<D, NM, true, 0, NC, T, LC_WAVES>), dim3(mfma_row_blocks), dim3(64 * LC_WAVES), 0, c.cur, in, out,
                           Mx, 0, 0, eigs, lambda2, nomap, 0, noin, noout);
      else if (TILED)
        hipLaunchKernelGGL((k_lc_gemm<D, NM, true>), dim3(gemm_row_blocks), dim3(256), 0, c.cur, in, out, Mx, 0, 0,
                           eigs, lambda2);
      else
        hipLaunchKernelGGL((k_lc_rows<D, N>), dim3(row_blocks), dim3(256), 0, c.cur, in, out, MTx, nrows);
      return hipGetLastError();
    };
    auto lines = [&](const double *in, double *out, const double *M, int axis, int div) {
      if (folding && f.fold[axis]) return lines_fold(in, out, axis, div, M == f.vinv + (size_t)axis * NN ? 1 : 2);
      if (MC)
        hipLaunchKernelGGL((k_lc_mfma<D, NM, false, 0, NC, T, LC_WAVES>), dim3(mfma_line_blocks), dim3(64 * LC_WAVES), 0, c.cur, in,
                           out, M, axis, div, eigs, lambda2, nomap, 0, noin, noout);
      else if (TILED)
        hipLaunchKernelGGL((k_lc_gemm<D, NM, false>), dim3(gemm_line_blocks), dim3(256), 0, c.cur, in, out, M, axis,
                           div, eigs, lambda2);
      else
        hipLaunchKernelGGL((k_lc_lines<D, N>), dim3(line_blocks), dim3(256), 0, c.cur, in, out, M, axis, div,
                           eigs, lambda2);
      return hipGetLastError();
    };
    // the last forward, the division and the first backward transform
    // along the last axis: one pass, or two
    const int fused_blocks = (N / 64) * batch;
    const bool fuse = MC && c.fuse_direct && !folding && fused_blocks >= c.n_cu;
    const int last = D - 1;
    const double *vinv_l = f.vinv + (size_t)last * NN, *v_l = f.v + (size_t)last * NN;
    // ping-pong between the grids: `cur` holds the latest result
    double *cur = A0, *nxt = A1;
    auto step = [&](hipError_t e) {
      std::swap(cur, nxt);
      return e;
    };
    if (pool) {
      nxt = A1;
      HIPCHK(step(rows(nullptr, nxt, f.vinv, f.vinvT, 1)));
    } else {
      HIPCHK(step(rows(cur, nxt, f.vinv, f.vinvT, 0)));
    }
    if (D == 3) HIPCHK(step(lines(cur, nxt, f.vinv + NN, 1, 0)));
    if (fuse) {
      hipLaunchKernelGGL((k_lc_mfma_fused<D, NM>), dim3(fused_blocks), dim3(4 * NM), 0, c.cur, cur, nxt, vinv_l,
                         v_l, (const double *)f.eig, lambda2);
      HIPCHK(step(hipGetLastError()));
    } else {
      HIPCHK(step(lines(cur, nxt, vinv_l, last, 1)));
      HIPCHK(step(lines(cur, nxt, v_l, last, 0)));
    }
    if (D == 3) HIPCHK(step(lines(cur, nxt, f.v + NN, 1, 0)));
    if (pool) {
      HIPCHK(rows(cur, nullptr, f.v, f.vT, 2));
      *result = nullptr;
    } else {
      HIPCHK(step(rows(cur, nxt, f.v, f.vT, 0)));
      *result = cur;
    }
    return AFGPU_OK;
  }


  // The transforms of the uniform-level solve in FP32 on the matrix cores
  // (64 or 128 cells per axis): the first x transform reads the rhs from the
  // FP32 pool, the last writes phi to it; the grids A0 / A1 hold floats.
  template <int N>
  int direct_transforms32(const afgpu_ctx::Direct::F &f, float lambda2, float *A0, float *A1,
                          const float *rhs_pool, float *phi_pool) {
    static_assert(N == 64 || N == 128, "matrix-core transform for N = 64, 128");
    const size_t NN = (size_t)N * N;
    const int batch = (D == 3) ? N : c.n_runs;
    const int nrows = (D == 3) ? N * N : N * c.n_runs;
    const int row_blocks = (N / 64) * (nrows / 64);
    const int line_blocks = (N / 64) * (N / 64) * batch;
    const int nb = N / NC, last = D - 1;
    const int *map = c.dir.d_map, *nomap = nullptr;
    const float *noin = nullptr;
    float *noout = nullptr;
    float *cur = A0, *nxt = A1;
    auto lines = [&](const float *M, int axis, int div) {
      hipLaunchKernelGGL((k_lc_mfma32<D, N, false, 0, NC>), dim3(line_blocks), dim3(256), 0, c.cur, cur, nxt, M,
                         axis, div, (const float *)f.eig32, lambda2, nomap, 0, noin, noout);
      std::swap(cur, nxt);
      return hipGetLastError();
    };
    hipLaunchKernelGGL((k_lc_mfma32<D, N, true, 1, NC>), dim3(row_blocks), dim3(256), 0, c.cur, noin, cur,
                       (const float *)f.vinv32, 0, 0, (const float *)f.eig32, lambda2, map, nb, rhs_pool, noout);
    HIPCHK(hipGetLastError());
    if (D == 3) HIPCHK(lines(f.vinv32 + NN, 1, 0));
    HIPCHK(lines(f.vinv32 + (size_t)last * NN, last, 1));
    HIPCHK(lines(f.v32 + (size_t)last * NN, last, 0));
    if (D == 3) HIPCHK(lines(f.v32 + NN, 1, 0));
    hipLaunchKernelGGL((k_lc_mfma32<D, N, true, 2, NC>), dim3(row_blocks), dim3(256), 0, c.cur, (const float *)cur,
                       noout, (const float *)f.v32, 0, 0, (const float *)f.eig32, lambda2, map, nb, noin, phi_pool);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }

  // phi(lc) := L^-1 rhs(lc) on the whole uniform level (homogeneous
  // boundary conditions: the correction equation), then its ghost layer.
  template <typename T>
  int solve_level(const Mg<T> &v, int lc) {
    afgpu_ctx::Direct &d = c.dir;
    const afgpu_ctx::Direct::F &f = d.f[v.op];
    if (d.lc != lc || !f.ok) {
      set_err("uniform-level solve used before its set-up");
      return AFGPU_ERR_STATE;
    }
    const int nb = d.N / NC;
    const int nbox = c.l_ids[lc].n;
    const bool mode = v.op >= AFGPU_OP_HELMH;
    double *A0 = mode ? c.mres[v.op - AFGPU_OP_HELMH].A0 : d.A0;
    double *A1 = mode ? c.mres[v.op - AFGPU_OP_HELMH].A1 : d.A1;
    if (!A0 || !A1) {
      set_err("uniform-level solve: grids not allocated");
      return AFGPU_ERR_STATE;
    }
    // on the matrix cores (64 or 128 cells per axis) the transforms read
    // the rhs from the pool and write phi to it themselves
    const bool pool = c.direct_mfma && (d.N == 64 || d.N == 128);
    if (!pool) {
      hipLaunchKernelGGL((k_lc_gather<D, NC, BS, T>), dim3(nbox), dim3(BS), 0, c.cur, (const int *)d.d_map, nb,
                         (const T *)v.rhs, A0);
      HIPCHK(hipGetLastError());
    }
    const T *rp = pool ? (const T *)v.rhs : nullptr;
    T *pp = pool ? v.phi : nullptr;
    const double l2 = c.ops[v.op].lambda2;
    if constexpr (std::is_same<T, float>::value) {
      if (pool && c.direct_f32 && !c.fuse_direct) {
        float *F0 = (float *)A0, *F1 = (float *)A1;
        if (d.N == 64)
          RCHK(direct_transforms32<64>(f, (float)l2, F0, F1, rp, pp));
        else
          RCHK(direct_transforms32<128>(f, (float)l2, F0, F1, rp, pp));
        return gc_var(lc, v.phi, v.m, true);
      }
    }
    const double *res = nullptr;
    switch (d.N) {
      case 16: RCHK((direct_transforms<16, T>(f, l2, A0, A1, &res))); break;
      case 32: RCHK((direct_transforms<32, T>(f, l2, A0, A1, &res))); break;
      case 64: RCHK((direct_transforms<64, T>(f, l2, A0, A1, &res, rp, pp))); break;
      case 128: RCHK((direct_transforms<128, T>(f, l2, A0, A1, &res, rp, pp))); break;
      default: set_err("uniform-level solve: unsupported size"); return AFGPU_ERR_UNSUPPORTED;
    }
    if (!pool) {
      hipLaunchKernelGGL((k_lc_scatter<D, NC, BS, T>), dim3(nbox), dim3(BS), 0, c.cur, (const int *)d.d_map, nb,
                         res, v.phi);
      HIPCHK(hipGetLastError());
    }
    return gc_var(lc, v.phi, v.m, true);
  }

  // --- mixed precision: FP64 defect correction around FP32 cycles ---------
  // One closing pass of a correction cycle: (optionally) phi += e, e = 0 on
  // every box; then phi on the refined (covered) cells := restriction of
  // the finer phi, level by level, which is what every FAS cycle starts
  // with (update_coarse, m_af_multigrid.f90:691-740) and what the composite
  // residual of the coarse leaf cells next to refined boxes depends on, so
  // the iteration has the plain form's fixed point; then the FP64 residual
  // of phi on the leaves into the FP32 rhs of the next cycle, with its max
  // reduced into reduction slot 0.
  int mp_close(int op, bool correct, int lc) {
    const int iphi = op_phi(op);
    if (correct) {
      const List &l = c.l_all_leaves;
      RCHK(with_bb(KC_BOXOP, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_BOXOP, k_mp_correct_leaves<D, NC, BS, bb>, l.n, c.list(l), c.cc[iphi], f32(op, 0));
      }));
    }
    // (levels below lc are covered everywhere and not part of the
    // iteration).  The covered values are read by the residual of coarse
    // leaf cells beside refined boxes and by the refinement-boundary ghosts
    // of finer boxes; below the lowest level where either occurs they are
    // not read before mp_finish (a uniformly refined run needs none).
    // The refinement-boundary ghosts of a level-l box (mg_sides_rb) read
    // the level l-1 neighbour including its stored ghost layer, so that
    // layer is refreshed on those neighbours (l_rbnb; the other boxes'
    // stored layers are not read before mp_finish refreshes every level).
    // The finest leaves' own ghosts are formed inside the residual kernel.
    const VarBc m = var_methods(c, iphi);
    if (c.n_runs > 1 && op != AFGPU_OP_POISSON) {
      // A photoionization mode in a forest keeps no mp_finish, so its
      // covered cells stay as its closes leave them: each run restricts
      // and fills the levels it would alone (the forest's decision, taken
      // over all runs, would touch cells a run alone leaves as they are).
      const int top = std::numeric_limits<int>::max();
      for (int lvl = c.highest_lvl; lvl >= mp_restrict_lo(lc); lvl--) {
        RCHK(mask_close(0, lvl, lc));
        RCHK(restrict_levels(iphi, lvl, lvl));
      }
      for (int lvl = lc; lvl < c.highest_lvl; lvl++)
        if (c.lvl_rb[lvl + 1]) {
          RCHK(mask_close(1, lvl, lc));
          RCHK(gc_list(c.l_rbnb[lvl], c.cc[iphi], m, true));
        }
      RCHK(mask_levels(0, top));
    } else {
      RCHK(restrict_levels(iphi, c.highest_lvl, mp_restrict_lo(lc)));
      for (int lvl = lc; lvl < c.highest_lvl; lvl++)
        if (c.lvl_rb[lvl + 1]) RCHK(gc_list(c.l_rbnb[lvl], c.cc[iphi], m, true));
    }
    const int slot = red_slot(op);
    RCHK(set_red(slot, 0ull));
    const List &l = c.l_all_leaves;
    const int k = bpb(l.n, G::NINT);
    if (k == 1)
      return with_bb(KC_RESID, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_RESID, k_mp_residual_lds<D, NC, BS, bb>, l.n, c.list(l), (const DBox *)c.d_boxes,
                         (const double *)c.cc[iphi], (const double *)c.cc[AFGPU_CC_RHS], f32(op, 1), c.ops[op].lambda2,
                         (int)(c.coord == AFGPU_CYL), m, c.d_red + slot);
      });
    return with_bb(KC_RESID, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      return items<bb>(KC_RESID, k_mp_residual<D, NC, BS, bb>, l.n * k, c.list(l), (const DBox *)c.d_boxes,
                       (const double *)c.cc[iphi], (const double *)c.cc[AFGPU_CC_RHS], f32(op, 1), c.ops[op].lambda2,
                       (int)(c.coord == AFGPU_CYL), m, c.d_red + slot, k);
    });
  }

  // Lowest level whose phi mp_close restricts to the level below (levels
  // lo..highest; highest + 1: none).
  int mp_restrict_lo(int lc) const {
    for (int lvl = lc + 1; lvl <= c.highest_lvl; lvl++)
      if (c.l_leaves[lvl - 1].n > 0 || c.lvl_rb[lvl]) return lvl;
    return c.highest_lvl + 1;
  }

  // phi on refined cells := restriction of the finer phi (no geometry
  // weights, as update_coarse restricts phi), levels hi -> lo-1.
  int restrict_levels(int iphi, int hi, int lo) {
    for (int lvl = hi; lvl >= lo; lvl--) {
      const List &l = c.l_ids[lvl];
      const int k = bpb(l.n, G::NINT / (1 << D));
      RCHK(with_bb(KC_RESTRICT, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_RESTRICT, k_restrict<D, NC, BS, double, bb>, l.n * k, c.list(l), (const DBox *)c.d_boxes,
                         (const double *)c.cc[iphi], c.cc[iphi], 0, (int)(c.coord == AFGPU_CYL), k);
      }));
    }
    return AFGPU_OK;
  }

  // End of a mixed-precision solve: the levels that mp_close leaves
  // unrestricted get the restricted solution too, and the stored ghost
  // layers are refreshed everywhere.
  int mp_finish(int op, int lc) {
    return run_graph(c, {G_MP_GC, op, lc, 0}, [&] { return mp_finish_body(op, lc); });
  }
  int mp_finish_body(int op, int lc) {
    const int iphi = op_phi(op);
    RCHK(restrict_levels(iphi, mp_restrict_lo(lc) - 1, 2));
    return gc_tree(iphi);
  }

  // read slots 0..n-1 of every run
  int read_red(int n) {
    const size_t cnt = (size_t)(c.n_runs - 1) * N_RED_SLOTS + n;
    const int nq = c.n_runs * N_RED_SLOTS;
    hipLaunchKernelGGL(k_fold_stripes, dim3((nq + 255) / 256), dim3(256), 0, c.stream,
                       (const unsigned long long *)c.d_red, c.d_red_out, c.n_runs, RED_MIN_SLOTS);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(c.h_red, c.d_red_out, cnt * sizeof(unsigned long long), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    return timer_harvest(c);
  }
  // reduction slot i of every run := value (a kernel, so that it can live
  // in a graph)
  int set_red(int i, unsigned long long value) {
    hipLaunchKernelGGL(k_set_run_slots, dim3((c.n_runs * RED_STRIPES + 63) / 64), dim3(64), 0, c.cur, c.d_red + i,
                       value, c.n_runs);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }
  static double as_double(unsigned long long u) {
    double d;
    std::memcpy(&d, &u, sizeof(d));
    return d;
  }

  // max |u| over leaf interiors, per run (out[r])
  int maxabs_runs(int iv, std::vector<double> &out) {
    RCHK(set_red(0, 0ull));
    const int k = bpb(c.l_all_leaves.n, G::NINT);
    RCHK(with_bb(KC_BOXOP, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      return items<bb>(KC_BOXOP, k_maxabs<D, NC, BS, bb>, c.l_all_leaves.n * k, c.list(c.l_all_leaves),
                       (const DBox *)c.d_boxes, (const double *)c.cc[iv], c.d_red, k);
    }));
    RCHK(read_red(1));
    out.assign(c.n_runs, 0.0);
    for (int r = 0; r < c.n_runs; r++)
      out[r] = c.l_all_leaves.n > 0 ? run_red(r, 0) : -std::numeric_limits<double>::max() / 10;
    return AFGPU_OK;
  }
  int maxabs(int iv, double *out) override {
    std::vector<double> m;
    RCHK(maxabs_runs(iv, m));
    *out = *std::max_element(m.begin(), m.end());
    return AFGPU_OK;
  }

  // --- field (src/m_field.f90) ----------------------------------------------
  int field_from_potential() override {
    return run_graph(c, {G_FFP, 0, 0, 0}, [&] { return ffp_body(); });
  }
  int ffp_body() {
    RCHK(with_bb(KC_UPDATE, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      return items<bb>(KC_UPDATE, k_gradient<D, NC, BS, bb>, c.l_all_ids.n, c.list(c.l_all_ids),
                       (const DBox *)c.d_boxes, (const double *)c.cc[AFGPU_CC_PHI], c.fc[AFGPU_FC_FIELD],
                       c.cc[AFGPU_CC_EFIELD]);
    }));
    return gc_tree(AFGPU_CC_EFIELD);
  }

  // field_compute (src/m_field.f90:448-528).  Each run of a forest has its
  // own threshold and stops on its own test: the runs that have converged
  // are masked out of the following cycles, so every run does exactly the
  // cycles it would do alone.
  int field_compute(int s_in, bool have_guess, int *n_out, double *res_out) override {
    const int R = c.n_runs;
    RCHK(set_active(nullptr));
    if (c.gen_sp) {
      const List &l = c.l_all_leaves;
      if (l.n > 0) {
        hipLaunchKernelGGL((k_set_rhs_n<G::NCELL, BS>), dim3(std::min(l.n, 8 * c.n_cu)), dim3(BS), 0, c.cur,
                           c.list(l), c.species_args(), s_in, c.cc[AFGPU_CC_RHS], l.n);
        HIPCHK(hipGetLastError());
      }
    } else {
      RCHK(with_bb(KC_BOXOP, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_BOXOP, k_set_rhs<G::NCELL, BS, bb>, c.l_all_leaves.n, c.list(c.l_all_leaves),
                         (const double *)c.cc[AFGPU_CC_ELECTRON + s_in], (const double *)c.cc[AFGPU_CC_POS_ION + s_in],
                         (const double *)c.cc[AFGPU_CC_NEG_ION + s_in], c.cc[AFGPU_CC_RHS]);
      }));
    }
    std::vector<double> max_rhs;
    RCHK(maxabs_runs(AFGPU_CC_RHS, max_rhs));
    const double conv_fac = 1e-10, min_residual = 1e-6, max_residual = 1e8;
    std::vector<double> threshold(R);
    for (int r = 0; r < R; r++) {
      const double volt = (R > 1) ? c.run_volt[r] : c.voltage;
      threshold[r] = std::max(std::max(min_residual, max_rhs[r] * c.p.mg_max_rel_residual),
                              conv_fac * std::fabs(volt) / (c.p.domain_len_last * c.run_min_dr[r]));
    }
    c.run_ncyc.assign(R, 0);
    c.run_res.assign(R, 0.0);
    std::vector<int> active = base_set();
    int n_active = 0;
    for (int r = 0; r < R; r++) n_active += active[r] != 0;
    const int n_part = n_active;
    if (!have_guess) {
      std::vector<std::vector<double>> res(R, std::vector<double>(101, 0.0));
      int i;
      for (i = 1; i <= 100 && n_active > 0; i++) {
        RCHK(fmg(AFGPU_OP_POISSON, true, true));
        RCHK(read_red(1));
        for (int r = 0; r < R; r++) {
          if (!active[r]) continue;
          res[r][i] = run_red(r, 0);
          c.run_ncyc[r]++;
          c.run_res[r] = res[r][i];
          bool done = res[r][i] < threshold[r];
          if (!done && i > 2) {
            const double mn = std::min(std::min(res[r][i - 2], res[r][i - 1]), res[r][i]);
            const double mx = std::max(std::max(res[r][i - 2], res[r][i - 1]), res[r][i]);
            const double ratio = mn / mx;
            done = ratio < 2.0 && ratio > 0.5 && res[r][i] < max_residual;
          }
          if (done) {
            active[r] = 0;
            n_active--;
          }
        }
        RCHK(set_active(&active));
      }
      if (n_active > 0) {
        RCHK(set_active(nullptr));
        set_err("No convergence in initial field computation");
        return AFGPU_ERR_NO_CONVERGENCE;
      }
      active = base_set();
      n_active = n_part;
      RCHK(set_active(nullptr));
    }
    if (c.precision == AFGPU_PREC_MIXED) {
      RCHK(mixed_vcycles(threshold, active));
    } else {
      for (int i = 1; i <= c.p.mg_num_vcycles && n_active > 0; i++) {
        RCHK(vcycle(AFGPU_OP_POISSON, true, c.highest_lvl));
        RCHK(read_red(1));
        for (int r = 0; r < R; r++) {
          if (!active[r]) continue;
          c.run_res[r] = run_red(r, 0);
          c.run_ncyc[r]++;
          if (c.run_res[r] < threshold[r]) {
            active[r] = 0;
            n_active--;
          }
        }
        RCHK(set_active(&active));
      }
      RCHK(set_active(nullptr));
    }
    RCHK(field_from_potential());
    if (n_out) *n_out = *std::max_element(c.run_ncyc.begin(), c.run_ncyc.end());
    if (res_out) *res_out = *std::max_element(c.run_res.begin(), c.run_res.end());
    return AFGPU_OK;
  }

  // Mixed precision: FP64 defect correction around FP32 V-cycles.  Each
  // iteration solves L e = r approximately with one single-precision FAS
  // V-cycle from e = 0 (smoothing, restriction and prolongation in FP32;
  // the coarsest level solved exactly in FP64: the finest uniformly refined
  // level when there is one, else the level-1 box), adds e to phi in FP64
  // and forms the new residual in FP64; each run stops on the same FP64
  // leaf-residual test as the plain form (m_field.f90:515-521).  `active`
  // holds the runs still iterating (all of them on entry).
  int mixed_vcycles(const std::vector<double> &threshold, std::vector<int> &active) {
    const int op = AFGPU_OP_POISSON;
    const int R = c.n_runs;
    int lc;
    RCHK(ensure_factors(op));
    RCHK(direct_level(&lc));
    RCHK(ensure_direct_factors(op));
    RCHK(run_graph(c, {G_MP_RES, op, lc, 0}, [&] { return mp_close(op, false, lc); }));
    int n_active = 0;
    for (int r = 0; r < R; r++) n_active += active[r] != 0;
    for (int i = 1; i <= c.p.mg_num_vcycles && n_active > 0; i++) {
      RCHK(run_graph(c, {G_MP_VCYCLE, op, lc, c.highest_lvl}, [&] {
        RCHK(vcycle_body(mg32(op), false, c.highest_lvl, lc));
        return mp_close(op, true, lc);
      }));
      RCHK(read_red(1));
      for (int r = 0; r < R; r++) {
        if (!active[r]) continue;
        c.run_res[r] = run_red(r, 0);
        c.run_ncyc[r]++;
        if (c.run_res[r] < threshold[r]) {
          active[r] = 0;
          n_active--;
        }
      }
      RCHK(set_active(&active));
    }
    RCHK(set_active(nullptr));
    return mp_finish(op, lc);
  }

  // --- fluid (src/m_fluid.f90, afivo/src/m_af_flux_schemes.f90) -------------
  int forward_euler(double dt, int s_deriv, int n_prev, const int *s_prev, const double *w_prev, int s_out,
                    int i_step, int n_steps, double *dt_lim) override {
    const bool last_step = (i_step == n_steps);
    if (i_step > 1) RCHK(field_compute(s_deriv, true, nullptr, nullptr));
    RCHK(fe_launch(dt, s_deriv, n_prev, s_prev, w_prev, s_out, last_step));
    RCHK(read_red(4));
    return fe_limits(last_step, dt_lim);
  }
  // The kernels of one forward-Euler stage (no host synchronisation: can
  // be captured); the dt limits are left in reduction slots 1-3.
  int fe_launch(double dt, int s_deriv, int n_prev, const int *s_prev, const double *w_prev, int s_out,
                bool last_step) {
    const int cyl = c.coord == AFGPU_CYL;
    const int iv_e = AFGPU_CC_ELECTRON + s_deriv;
    // af_restrict_ref_boundary (afivo/src/m_af_restrict.f90:140-161)
    RCHK(with_bb(KC_RESTRICT, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      return items<bb>(KC_RESTRICT, k_restrict<D, NC, BS, double, bb>, c.l_refb.n, c.list(c.l_refb),
                       (const DBox *)c.d_boxes, (const double *)c.cc[iv_e], c.cc[iv_e], 1, cyl, 1);
    }));
    const double big = 1e100;
    unsigned long long bigbits;
    std::memcpy(&bigbits, &big, sizeof(big));
    RCHK(set_red(1, 0ull));
    RCHK(set_red(2, 0ull));
    RCHK(set_red(3, bigbits));
    const double N_inv = 1 / c.p.gas_number_density;
    const VarBc msp = var_methods(c, iv_e);
    for (int lvl = 1; lvl <= c.highest_lvl; lvl++) {
      const List &l = c.l_leaves[lvl];
      const double fbytes = (2.0 + 2 * D) * 8.0 * l.n * G::NINT;
      if (c.precision == AFGPU_PREC_MIXED && c.flux32)
        RCHK(with_bb(KC_FLUX, [&](auto B) {
          constexpr int bb = decltype(B)::value;
          return items_timed<bb>(1, fbytes, KC_FLUX, k_flux<D, NC, BS, float, bb>, l.n, c.list(l),
                                 (const DBox *)c.d_boxes, c.cc[iv_e], (const double *)c.cc[AFGPU_CC_EFIELD],
                                 (const double *)c.fc[AFGPU_FC_FIELD], c.fc[AFGPU_FC_FLUX], msp, c.td, N_inv,
                                 c.d_red + 1, c.d_red + 2);
        }));
      else
        RCHK(with_bb(KC_FLUX, [&](auto B) {
          constexpr int bb = decltype(B)::value;
          return items_timed<bb>(1, fbytes, KC_FLUX, k_flux<D, NC, BS, double, bb>, l.n, c.list(l),
                                 (const DBox *)c.d_boxes, c.cc[iv_e], (const double *)c.cc[AFGPU_CC_EFIELD],
                                 (const double *)c.fc[AFGPU_FC_FIELD], c.fc[AFGPU_FC_FLUX], msp, c.td, N_inv,
                                 c.d_red + 1, c.d_red + 2);
        }));
    }
    for (int lvl = 1; lvl <= c.highest_lvl - 1; lvl++) {
      const List &l = c.l_parents[lvl];
      RCHK(with_bb(KC_UPDATE, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_UPDATE, k_consistent_fluxes<D, NC, BS, bb>, l.n, c.list(l), (const DBox *)c.d_boxes,
                         c.fc[AFGPU_FC_FLUX], cyl);
      }));
    }
    UpdateArgs A{};
    for (int sp = 0; sp < 3; sp++)
      for (int s = 0; s < 3; s++) A.sp[sp][s] = c.cc[AFGPU_CC_SPECIES(sp) + s];
    A.efield = c.cc[AFGPU_CC_EFIELD];
    A.photo = c.cc[AFGPU_CC_PHOTO];
    A.flux = c.fc[AFGPU_FC_FLUX];
    A.chem = c.chem;
    A.dt = dt;
    A.dt_run = c.use_run_dt ? c.d_dt_run : nullptr;  // forest: dt_run[r] scaled by dt
    A.dt_scale = dt;
    A.N_inv = N_inv;
    A.s_deriv = s_deriv;
    A.s_out = s_out;
    A.n_prev = n_prev;
    for (int m = 0; m < n_prev; m++) {
      A.s_prev[m] = s_prev[m];
      A.w_prev[m] = w_prev[m];
    }
    A.last_step = last_step;
    A.chem_nmin_on = c.p.dt_chemistry_nmin > 0;
    A.chem_nmin = c.p.dt_chemistry_nmin;
    A.chem_limit_loss = c.p.dt_chemistry_limit_loss;
    A.photoi_enabled = c.p.photoi_enabled;
    A.photoi_species = c.p.photoi_species;
    A.cyl = cyl;
    A.use_net = c.net ? 1 : 0;
    if (c.net) A.net = c.net->args;
    if (c.gen_sp) {
      const List &l = c.l_all_leaves;
      if (l.n > 0) {
        const size_t lds = update_n_lds(BS, c.n_sp, c.net->n_reactions);
        if (lds > UPDATE_N_LDS_MAX) {
          set_err("reaction list too large for the update's LDS (species and reaction records)");
          return AFGPU_ERR_UNSUPPORTED;
        }
        hipLaunchKernelGGL((k_update_n<D, NC, BS>), dim3(std::min(l.n, 8 * c.n_cu)), dim3(BS), lds, c.cur, c.list(l),
                           (const DBox *)c.d_boxes, A, c.species_args(), c.d_red + 3, l.n);
        HIPCHK(hipGetLastError());
      }
      return AFGPU_OK;
    }
    RCHK(with_bb(KC_UPDATE, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      const double bytes = (7.0 + D + (c.p.photoi_enabled ? 1 : 0)) * 8.0 * c.l_all_leaves.n * G::NINT;
      if (A.use_net)
        return items_timed<bb>(2, bytes, KC_UPDATE, k_update<D, NC, BS, bb, true>, c.l_all_leaves.n,
                               c.list(c.l_all_leaves), (const DBox *)c.d_boxes, A, c.d_red + 3);
      return items_timed<bb>(2, bytes, KC_UPDATE, k_update<D, NC, BS, bb, false>, c.l_all_leaves.n,
                             c.list(c.l_all_leaves), (const DBox *)c.d_boxes, A, c.d_red + 3);
    }));
    return AFGPU_OK;
  }
  // dt limits of the stage from slots 1-3 as read back (src/m_fluid.f90:96-98)
  int fe_limits(bool last_step, double *dt_lim) {
    // per run (only the runs not masked out are meaningful)
    c.run_dt_lim.assign(c.n_runs, 1e100);
    for (int r = 0; r < c.n_runs; r++) {
      double lim[4];
      const double cfl_max = run_red(r, 1);
      const double sig_max = run_red(r, 2);
      lim[0] = c.l_all_leaves.n > 0 ? 1 / cfl_max : 1e100;
      lim[1] = c.l_all_leaves.n > 0 ? UC_eps0 / (UC_elem_charge * std::max(sig_max, 1e-100)) : 1e100;
      lim[2] = last_step ? run_red(r, 3) : 1e100;
      lim[3] = 1e100;
      lim[0] = lim[0] * c.p.dt_cfl_number;
      const double mv = std::min(std::min(std::min(lim[0], lim[1]), lim[2]), lim[3]);
      c.run_dt_lim[r] = std::min(c.p.dt_max, mv);
    }
    *dt_lim = c.run_dt_lim[0];
    return AFGPU_OK;
  }

  // --- photoionization (src/m_photoi.f90, src/m_photoi_helmh.f90) -----------
  // Per-mode single-precision pools and uniform-level grids (allocated
  // outside any capture; zero-filled, as the correction must start at 0).
  int ensure_mode_pools() {
    const afgpu_ctx::Direct &d = c.dir;
    for (int n = 0; n < c.p.photoi_n_modes; n++) {
      afgpu_ctx::ModeRes &r = c.mres[n];
      if (r.cap_boxes != c.cap_boxes) {
        for (int k = 0; k < 4; k++) {
          (void)hipFree(r.cc32[k]);
          r.cc32[k] = nullptr;
          HIPCHK(hipMalloc(&r.cc32[k], (size_t)c.cap_boxes * c.ncell * sizeof(float)));
          HIPCHK(hipMemsetAsync(r.cc32[k], 0, (size_t)c.cap_boxes * c.ncell * sizeof(float), c.stream));
        }
        r.cap_boxes = c.cap_boxes;
      }
      if (d.lc > 1 && r.a_cap < d.cap) {
        (void)hipFree(r.A0);
        (void)hipFree(r.A1);
        r.A0 = r.A1 = nullptr;
        HIPCHK(hipMalloc(&r.A0, d.cap * sizeof(double)));
        HIPCHK(hipMalloc(&r.A1, d.cap * sizeof(double)));
        r.a_cap = d.cap;
      }
    }
    return AFGPU_OK;
  }

  // Run body(op) for every photoionization mode in `mask`: as parallel
  // branches on the mode streams (forked from and joined back into the
  // handle's stream, so that a capture records them as branches of one
  // graph), or one after the other on the handle's stream.
  template <typename F>
  int fork_modes(int mask, F body) {
    int count = 0;
    for (int n = 0; n < c.p.photoi_n_modes; n++) count += (mask >> n) & 1;
    if (!c.batch_modes || count < 2) {
      for (int n = 0; n < c.p.photoi_n_modes; n++)
        if ((mask >> n) & 1) RCHK(body(AFGPU_OP_HELMH + n));
      return AFGPU_OK;
    }
    HIPCHK(hipEventRecord(c.fork_ev, c.stream));
    int rc = AFGPU_OK;
    for (int n = 0; n < c.p.photoi_n_modes && rc == AFGPU_OK; n++) {
      if (!((mask >> n) & 1)) continue;
      HIPCHK(hipStreamWaitEvent(c.mstream[n], c.fork_ev, 0));
      c.cur = c.mstream[n];
      rc = body(AFGPU_OP_HELMH + n);
      c.cur = c.stream;
      HIPCHK(hipEventRecord(c.join_ev[n], c.mstream[n]));
      HIPCHK(hipStreamWaitEvent(c.stream, c.join_ev[n], 0));
    }
    return rc;
  }

  // photoi_set_src (src/m_photoi.f90:153, src/m_photoi_helmh.f90:162-191)
  // for the runs of the base set.  In a forest each run stops every mode on
  // its own test (its own max |rhs|): a run that has converged is masked
  // out of the mode's further cycles, so every run does exactly the cycles
  // it does alone.
  int photoi() override {
    const int R = c.n_runs;
    RCHK(set_active(nullptr));
    RCHK(with_bb(KC_UPDATE, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      return items<bb>(KC_UPDATE, k_photoi_src<D, NC, BS, bb>, c.l_all_leaves.n, c.list(c.l_all_leaves),
                       (const double *)c.cc[AFGPU_CC_EFIELD], (const double *)c.cc[AFGPU_CC_ELECTRON],
                       c.cc[AFGPU_CC_RHS], c.td, c.p.gas_number_density, c.p.photoi_eta_quench);
    }));
    std::vector<double> max_rhs;
    RCHK(maxabs_runs(AFGPU_CC_RHS, max_rhs));
    for (double &m : max_rhs) m = std::max(m, std::sqrt(std::numeric_limits<double>::epsilon()));
    const int n_modes = c.p.photoi_n_modes;
    const std::vector<int> base = base_set();
    c.run_pcyc.assign(R, 0);
    c.photoi_cycles = 0;
    if (c.precision == AFGPU_PREC_MIXED && n_modes > 0) {
      // Each mode: FP64 defect correction around single-precision FMG
      // cycles on its correction (from e = 0, rhs = the FP64 residual of the
      // previous solution of the mode), stopping on the reference's test
      // (m_photoi_helmh.f90:184-189).  The modes are independent: on a
      // single run their cycles run side by side, each mode stopping on its
      // own; a forest's run masks are shared by all modes, so there the
      // modes take turns, each with its own masks.
      int lc;
      RCHK(direct_level(&lc));
      for (int n = 0; n < n_modes; n++) {
        RCHK(ensure_factors(AFGPU_OP_HELMH + n));
        RCHK(ensure_direct_factors(AFGPU_OP_HELMH + n));
      }
      RCHK(ensure_mode_pools());
      const int all = (1 << n_modes) - 1;
      for (int group = 0; group < (R == 1 ? 1 : n_modes); group++) {
        const int modes = (R == 1) ? all : (1 << group);
        int active = modes;
        std::vector<int> act = base;
        RCHK(run_graph(c, {G_MP_RES_MODES, active, lc, 0},
                       [&] { return fork_modes(active, [&](int op) { return mp_close(op, false, lc); }); }));
        for (int i = 1; i <= c.p.photoi_max_fmg_cycles && active; i++) {
          c.photoi_cycles += __builtin_popcount(active);
          for (int r = 0; r < R; r++)
            if (act[r]) c.run_pcyc[r] += __builtin_popcount(active);
          RCHK(run_graph(c, {G_MP_FMG_MODES, active, lc, 0}, [&] {
            return fork_modes(active, [&](int op) {
              if (c.photoi_inner_vcycle)
                RCHK(vcycle_body(mg32(op), false, c.highest_lvl, lc));
              else
                RCHK(fmg_body(mg32(op), false, false, lc));
              return mp_close(op, true, lc);
            });
          }));
          RCHK(read_red(N_RED_SLOTS));
          if (R == 1) {
            for (int n = 0; n < n_modes; n++)
              if ((active >> n) & 1) {
                const double residu = run_red(0, RED_MODE + n);
                if (residu / max_rhs[0] < c.p.photoi_max_rel_residual) active &= ~(1 << n);
              }
          } else {
            int left = 0;
            for (int r = 0; r < R; r++) {
              if (!act[r]) continue;
              if (run_red(r, RED_MODE + group) / max_rhs[r] < c.p.photoi_max_rel_residual)
                act[r] = 0;
              else
                left++;
            }
            if (left == 0) active = 0;
            RCHK(set_active(&act));
          }
        }
        RCHK(set_active(nullptr));
      }
      // No final ghost refresh of the modes: the next solve forms its
      // residual with in-kernel ghosts, and the photoionization source is
      // read on interior cells only (src/m_fluid.f90:437-439).
      return photo_sum();
    }
    for (int n = 0; n < n_modes; n++) {
      const int op = AFGPU_OP_HELMH + n;
      std::vector<int> act = base;
      int left = 0;
      for (int r = 0; r < R; r++) left += act[r] != 0;
      for (int i = 1; i <= c.p.photoi_max_fmg_cycles && left > 0; i++) {
        c.photoi_cycles++;
        RCHK(fmg(op, true, true));
        RCHK(read_red(1));
        for (int r = 0; r < R; r++) {
          if (!act[r]) continue;
          c.run_pcyc[r]++;
          if (run_red(r, 0) / max_rhs[r] < c.p.photoi_max_rel_residual) {
            act[r] = 0;
            left--;
          }
        }
        RCHK(set_active(&act));
      }
      RCHK(set_active(nullptr));
    }
    return photo_sum();
  }

  // photo := 0 on all boxes, then -= coeff_n * phi_n on the leaves for
  // every mode in turn, in one pass (no mode solve reads photo)
  int photo_sum() {
    PhotoModes pm{};
    pm.n = c.p.photoi_n_modes;
    static_assert(PHOTO_MAX_MODES == AFGPU_MAX_MODES, "mode count");
    for (int n = 0; n < pm.n; n++) {
      pm.phi[n] = c.cc[op_phi(AFGPU_OP_HELMH + n)];
      pm.coeff[n] = c.p.photoi_coeff[n];
    }
    return with_bb(KC_BOXOP, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      return items<bb>(KC_BOXOP, k_photo_sum<G::NCELL, BS, bb>, c.l_all_ids.n, c.list(c.l_all_ids),
                       (const DBox *)c.d_boxes, pm, c.cc[AFGPU_CC_PHOTO]);
    });
  }

  int refine_cells(const afgpu_refine_t &rp, const int *ids, int n_ids, int *flags) override {
    if (n_ids <= 0) return AFGPU_OK;
    hipLaunchKernelGGL((k_refine_cells<D, NC, BS>), dim3(std::min(n_ids, 8 * c.n_cu)), dim3(BS), 0, c.cur, ids,
                       (const DBox *)c.d_boxes, (const double *)c.cc[AFGPU_CC_EFIELD],
                       (const double *)c.cc[AFGPU_CC_ELECTRON], rp, c.td, c.p.gas_number_density, flags, n_ids);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }

  int restrict_ids(const int *child_ids, int n, double *pool) override {
    if (n <= 0) return AFGPU_OK;
    hipLaunchKernelGGL((k_restrict<D, NC, BS, double, 1>), dim3(std::min(n, 8 * c.n_cu)), dim3(BS), 0, c.cur,
                       child_ids, (const DBox *)c.d_boxes, (const double *)pool, pool, 1,
                       (int)(c.coord == AFGPU_CYL), 1, n);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }
  int prolong_new(const int *ids, int n, const VarRef *vars, int n_vars) override {
    if (n <= 0 || n_vars <= 0) return AFGPU_OK;
    if (n_vars > NEW_MAX_VARS) {
      set_err("prolongation of new boxes: too many variables");
      return AFGPU_ERR_ARG;
    }
    NewVars nv;
    nv.n = n_vars;
    for (int k = 0; k < n_vars; k++) {
      nv.v[k] = vars[k].p;
      nv.limiter[k] = vars[k].limited ? var_methods(c, vars[k].methods).limiter : -1;
    }
    const int nw = n * n_vars;
    hipLaunchKernelGGL((k_prolong_new<D, NC, BS>), dim3(std::min(nw, 8 * c.n_cu)), dim3(BS), 0, c.cur, ids,
                       (const DBox *)c.d_boxes, nv, nw);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }
  int diag_boxes(const VarPtrs &vars, double *part) override {
    const List &l = c.l_all_leaves;
    if (l.n <= 0) return AFGPU_OK;
    hipLaunchKernelGGL((k_diag_box<D, NC, BS>), dim3(std::min(l.n, 8 * c.n_cu)), dim3(BS), 0, c.cur, c.list(l),
                       (const DBox *)c.d_boxes, vars, (int)(c.coord == AFGPU_CYL), part, l.n);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }
  int gc_ids(const int *ids, int n, const VarRef *vars, int n_vars) override {
    if (n <= 0 || n_vars <= 0) return AFGPU_OK;
    if (n_vars > GC_MAX_VARS) {
      set_err("ghost cells of new boxes: too many variables");
      return AFGPU_ERR_ARG;
    }
    GcVars gv;
    gv.n = n_vars;
    for (int k = 0; k < n_vars; k++) {
      gv.v[k] = vars[k].p;
      gv.m[k] = var_methods(c, vars[k].methods);
    }
    const int nw = n * n_vars;
    hipLaunchKernelGGL((k_gc_lds_vars<D, NC, BS>), dim3(std::min(nw, 8 * c.n_cu)), dim3(BS), 0, c.cur, ids,
                       (const DBox *)c.d_boxes, gv, 1, nw);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }

  int copy_states(bool backup) override {
    const int ns = c.n_states();
    const SpeciesArgs S = c.species_args();
    for (int sp = 0; sp < S.n; sp++) {
      if (backup)
        RCHK(boxop(c.l_all_ids, S.sp[sp][0], S.sp[sp][ns], OP_COPY));
      else
        RCHK(boxop(c.l_all_ids, S.sp[sp][ns], S.sp[sp][0], OP_COPY));
    }
    if (backup)
      RCHK(boxop(c.l_all_ids, c.cc[AFGPU_CC_PHI], c.cc[AFGPU_CC_PHI_2], OP_COPY));
    else {
      RCHK(boxop(c.l_all_ids, c.cc[AFGPU_CC_PHI_2], c.cc[AFGPU_CC_PHI], OP_COPY));
      RCHK(field_from_potential());
    }
    return AFGPU_OK;
  }

  // --- ensembles: the field solve driven by the device ----------------------
  // Host-side set-up of the Poisson solve (factors, the uniform level):
  // allocations, so it runs before any capture.
  int field_setup() {
    RCHK(ensure_factors(AFGPU_OP_POISSON));
    if (c.precision == AFGPU_PREC_MIXED) {
      int lc;
      RCHK(direct_level(&lc));
      RCHK(ensure_direct_factors(AFGPU_OP_POISSON));
    }
    return AFGPU_OK;
  }
  // fold the stripes of every reduction slot (capturable)
  int fold_red() {
    const int nq = c.n_runs * N_RED_SLOTS;
    hipLaunchKernelGGL(k_fold_stripes, dim3((nq + 255) / 256), dim3(256), 0, c.cur,
                       (const unsigned long long *)c.d_red, c.d_red_out, c.n_runs, RED_MIN_SLOTS);
    HIPCHK(hipGetLastError());
    return AFGPU_OK;
  }
  // field_compute with a guess for every run of the forest as one
  // capturable sequence: the thresholds are formed on the device, and after
  // each of K cycles a kernel counts the cycle for every active run and
  // drops the converged ones, whose boxes the list kernels then skip (so
  // every run does the cycles it would do alone).  The records (RunStep)
  // tell the host which runs need more than K cycles.  Ends with all runs
  // active, the solve finished and the field computed.
  int field_solve_dev(int K) {
    const int op = AFGPU_OP_POISSON, R = c.n_runs, top = std::numeric_limits<int>::max();
    const dim3 rg((R + 63) / 64), rb(64);
    if (c.gen_sp) {
      const List &l = c.l_all_leaves;
      if (l.n > 0) {
        hipLaunchKernelGGL((k_set_rhs_n<G::NCELL, BS>), dim3(std::min(l.n, 8 * c.n_cu)), dim3(BS), 0, c.cur,
                           c.list(l), c.species_args(), 0, c.cc[AFGPU_CC_RHS], l.n);
        HIPCHK(hipGetLastError());
      }
    } else {
      RCHK(with_bb(KC_BOXOP, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_BOXOP, k_set_rhs<G::NCELL, BS, bb>, c.l_all_leaves.n, c.list(c.l_all_leaves),
                         (const double *)c.cc[AFGPU_CC_ELECTRON], (const double *)c.cc[AFGPU_CC_POS_ION],
                         (const double *)c.cc[AFGPU_CC_NEG_ION], c.cc[AFGPU_CC_RHS]);
      }));
    }
    RCHK(set_red(0, 0ull));
    const int k = bpb(c.l_all_leaves.n, G::NINT);
    RCHK(with_bb(KC_BOXOP, [&](auto B) {
      constexpr int bb = decltype(B)::value;
      return items<bb>(KC_BOXOP, k_maxabs<D, NC, BS, bb>, c.l_all_leaves.n * k, c.list(c.l_all_leaves),
                       (const DBox *)c.d_boxes, (const double *)c.cc[AFGPU_CC_RHS], c.d_red, k);
    }));
    RCHK(fold_red());
    hipLaunchKernelGGL(k_run_thresholds, rg, rb, 0, c.cur, (const unsigned long long *)c.d_red_out,
                       (const double *)c.d_run_volt, (const double *)c.d_run_min_dr, c.p.mg_max_rel_residual,
                       c.p.domain_len_last, (const int *)c.d_run_ones, c.d_run_host, c.d_step, R);
    HIPCHK(hipGetLastError());
    RCHK(mask_levels(0, top));
    const bool mixed = c.precision == AFGPU_PREC_MIXED;
    const int lc = mixed ? c.dir.lc : 1;
    if (mixed) RCHK(mp_close(op, false, lc));
    for (int i = 0; i < K; i++) {
      if (mixed) {
        RCHK(vcycle_body(mg32(op), false, c.highest_lvl, lc));
        RCHK(mp_close(op, true, lc));
      } else {
        RCHK(vcycle_body(mg64(op), true, c.highest_lvl, 1, true));
      }
      RCHK(fold_red());
      hipLaunchKernelGGL(k_run_converge, rg, rb, 0, c.cur, (const unsigned long long *)c.d_red_out, c.d_run_host,
                         c.d_step, R);
      HIPCHK(hipGetLastError());
      RCHK(mask_levels(0, top));
    }
    hipLaunchKernelGGL(k_run_copy, rg, rb, 0, c.cur, (const int *)c.d_run_ones, c.d_run_host, R);
    HIPCHK(hipGetLastError());
    RCHK(mask_levels(0, top));
    if (mixed) RCHK(mp_finish_body(op, lc));
    return ffp_body();
  }
  // The runs in `more` go on with the cycles of field_compute's loop from
  // where the device-driven solve left them (count and threshold in the
  // records), each to its own stop, then their solve is finished again.
  int continue_solve(const std::vector<int> &more) {
    const int op = AFGPU_OP_POISSON, R = c.n_runs;
    const bool mixed = c.precision == AFGPU_PREC_MIXED;
    const int lc = mixed ? c.dir.lc : 1;
    std::vector<int> active = more;
    int n_active = 0, done_cycles = 0;
    for (int r = 0; r < R; r++)
      if (active[r]) {
        n_active++;
        done_cycles = c.h_step[r].ncyc;  // (equal for all of them: active throughout)
      }
    RCHK(set_active(&active));
    for (int i = done_cycles + 1; i <= c.p.mg_num_vcycles && n_active > 0; i++) {
      if (mixed)
        RCHK(run_graph(c, {G_MP_VCYCLE, op, lc, c.highest_lvl}, [&] {
          RCHK(vcycle_body(mg32(op), false, c.highest_lvl, lc));
          return mp_close(op, true, lc);
        }));
      else
        RCHK(vcycle(op, true, c.highest_lvl));
      RCHK(read_red(1));
      for (int r = 0; r < R; r++) {
        if (!active[r]) continue;
        c.run_res[r] = run_red(r, 0);
        c.run_ncyc[r]++;
        if (c.run_res[r] < c.h_step[r].thr) {
          active[r] = 0;
          n_active--;
        }
      }
      RCHK(set_active(&active));
    }
    RCHK(set_active(&more));
    if (mixed) RCHK(mp_finish(op, lc));
    RCHK(field_from_potential());
    return set_active(nullptr);
  }

  // --- ensembles: one time step of every run of a forest -------------------
  // The time-loop body of src/streamer.f90:181-432 (as afgpu_step does it
  // for one run) for every run at once: each run has its own dt, its own
  // rejections and its own field-solve stopping test.  A rejected run is
  // restored and advanced again with the other runs masked out.
  int step_runs(afgpu_step_info_t *info) override {
    const int R = c.n_runs;
    for (auto &t : c.rt) {
      if (!t.started) {
        t.global_dt = c.p.dt_min;
        t.dt = t.global_dt;
        t.started = true;
      }
      t.it = t.it + 1;
    }
    // photoionization of the runs whose step count is due
    // (src/streamer.f90:236-239, as afgpu_step for one run)
    std::vector<int> photo(R, 0);
    if (c.p.photoi_enabled && c.p.photoi_per_steps > 0) {
      int n_photo = 0;
      for (int r = 0; r < R; r++) {
        photo[r] = c.rt[r].it % c.p.photoi_per_steps == 0;
        n_photo += photo[r];
      }
      if (n_photo > 0) {
        if (n_photo < R) c.base_active = photo;
        const int rc = photoi();
        c.base_active.clear();
        RCHK(rc);
        RCHK(set_active(nullptr));
      }
    }
    std::vector<double> dt(R), dt_lim(R, std::numeric_limits<double>::max());
    std::vector<int> todo(R, 1), n_att(R, 0), n_rej(R, 0);
    for (int r = 0; r < R; r++) dt[r] = c.rt[r].dt;
    const int max_attempts = 10;
    int n_todo = R;
    // Fast path: the first attempt of every run and its field solve as one
    // captured graph (backup, flux, update, K device-driven cycles, field),
    // then one read-back.  The solve of a run whose step turns out rejected
    // is undone by the restore; runs that need more than K cycles go on
    // afterwards.
    std::vector<int> solved(R, 0);
    bool slow_field = false;
    int n_first = 1;
    if (c.p.time_integrator == AFGPU_FORWARD_EULER && c.step_graph && c.use_graphs && !c.timing && R > 1) {
      RCHK(field_setup());
      for (int r = 0; r < R; r++) c.h_dt_run[r] = dt[r];
      HIPCHK(hipMemcpyAsync(c.d_dt_run, c.h_dt_run, R * sizeof(double), hipMemcpyHostToDevice, c.stream));
      const int K = std::max(1, std::min(c.step_k, c.p.mg_num_vcycles));
      c.use_run_dt = true;
      const int rc = run_graph(c, {G_STEP_RUNS, K, c.precision, 0}, [&] {
        RCHK(copy_states(true));
        const int sp[1] = {0};
        const double w[1] = {1.0};
        RCHK(fe_launch(1.0, 0, 1, sp, w, 0, true));
        return field_solve_dev(K);
      });
      c.use_run_dt = false;
      RCHK(rc);
      HIPCHK(hipMemcpyAsync(c.h_red, c.d_red_out, (size_t)R * N_RED_SLOTS * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, c.stream));
      HIPCHK(hipMemcpyAsync(c.h_step, c.d_step, (size_t)R * sizeof(RunStep), hipMemcpyDeviceToHost, c.stream));
      HIPCHK(hipStreamSynchronize(c.stream));
      double dl;
      RCHK(fe_limits(true, &dl));
      c.run_ncyc.assign(R, 0);
      c.run_res.assign(R, 0.0);
      std::vector<int> rejected(R, 0), more(R, 0);
      int n_rejected = 0, n_more = 0, kmax = 1;
      for (int r = 0; r < R; r++) {
        afgpu_ctx::RunTime &t = c.rt[r];
        t.time = t.time + dt[r];
        const double lim = c.run_dt_lim[r];
        dt_lim[r] = std::min(dt_lim[r], lim);
        n_att[r] = 1;
        if (dt[r] <= lim) {
          todo[r] = 0;
          n_todo--;
          solved[r] = 1;
          c.run_ncyc[r] = c.h_step[r].ncyc;
          c.run_res[r] = c.h_step[r].res;
          kmax = std::max(kmax, c.h_step[r].ncyc);
          if (!c.h_step[r].done) {
            more[r] = 1;
            n_more++;
          }
        } else {
          n_rej[r]++;
          dt[r] = c.p.dt_safety_factor * lim;
          t.time = t.global_time;
          rejected[r] = 1;
          n_rejected++;
        }
      }
      if (n_rejected > 0) {
        c.base_active = rejected;
        RCHK(set_active(nullptr));
        RCHK(copy_states(false));
        slow_field = true;
      }
      if (n_more > 0) {
        c.base_active.clear();
        RCHK(continue_solve(more));
        for (int r = 0; r < R; r++)
          if (more[r]) kmax = std::max(kmax, c.run_ncyc[r]);
        slow_field = true;
      }
      // the next steps' graph holds as many cycles as any step has needed
      // (non-decreasing, so that only a few graphs are ever captured)
      c.step_k = std::max(c.step_k, kmax);
      n_first = 2;
    }
    for (int n = n_first; n <= max_attempts && n_todo > 0; n++) {
      c.base_active = todo;
      RCHK(set_active(nullptr));
      RCHK(copy_states(true));
      for (int r = 0; r < R; r++) c.h_dt_run[r] = dt[r];
      // (the staging buffer is not rewritten before forward_euler's read-back
      // has synchronised the stream)
      HIPCHK(hipMemcpyAsync(c.d_dt_run, c.h_dt_run, R * sizeof(double), hipMemcpyHostToDevice, c.stream));
      c.use_run_dt = true;
      double dl;
      int rc;
      if (c.p.time_integrator == AFGPU_FORWARD_EULER) {
        const int sp[1] = {0};
        const double w[1] = {1.0};
        rc = forward_euler(1.0, 0, 1, sp, w, 0, 1, 1, &dl);
      } else {
        const int sp1[1] = {0};
        const double w1[1] = {1.0};
        rc = forward_euler(1.0, 0, 1, sp1, w1, 1, 1, 2, &dl);
        const int sp2[2] = {0, 1};
        const double w2[2] = {0.5, 0.5};
        if (rc == AFGPU_OK) rc = forward_euler(0.5, 1, 2, sp2, w2, 0, 2, 2, &dl);
      }
      c.use_run_dt = false;
      if (rc != AFGPU_OK) {
        c.base_active.clear();
        return rc;
      }
      std::vector<int> rejected(R, 0);
      int n_rejected = 0;
      for (int r = 0; r < R; r++) {
        if (!todo[r]) continue;
        afgpu_ctx::RunTime &t = c.rt[r];
        t.time = t.time + dt[r];
        const double lim = c.run_dt_lim[r];
        dt_lim[r] = std::min(dt_lim[r], lim);
        n_att[r] = n;
        if (dt[r] <= lim) {
          todo[r] = 0;
          n_todo--;
        } else {
          n_rej[r]++;
          dt[r] = c.p.dt_safety_factor * lim;
          t.time = t.global_time;
          rejected[r] = 1;
          n_rejected++;
        }
      }
      if (n_rejected > 0) {
        c.base_active = rejected;
        RCHK(set_active(nullptr));
        RCHK(copy_states(false));
      }
    }
    c.base_active.clear();
    RCHK(set_active(nullptr));
    if (n_todo > 0) {
      set_err("All time steps were rejected");
      return AFGPU_ERR_ALL_REJECTED;
    }
    // the field solve of every run whose accepted step has not had it
    std::vector<int> unsolved(R, 0);
    int n_unsolved = 0;
    for (int r = 0; r < R; r++)
      if (!solved[r]) {
        unsolved[r] = 1;
        n_unsolved++;
      }
    if (n_unsolved == R) {
      RCHK(field_compute(0, true, nullptr, nullptr));
    } else if (n_unsolved > 0) {
      const std::vector<int> ncyc = c.run_ncyc;
      const std::vector<double> res = c.run_res;
      c.base_active = unsolved;
      RCHK(field_compute(0, true, nullptr, nullptr));
      c.base_active.clear();
      RCHK(set_active(nullptr));
      for (int r = 0; r < R; r++)
        if (solved[r]) {
          c.run_ncyc[r] = ncyc[r];
          c.run_res[r] = res[r];
        }
    }
    std::vector<double> max_rhs(R, 0.0);
    if (info) {
      if (n_unsolved > 0 || slow_field || n_first == 1)
        RCHK(maxabs_runs(AFGPU_CC_RHS, max_rhs));
      else
        for (int r = 0; r < R; r++) max_rhs[r] = c.h_step[r].max_rhs;
    }
    bool too_small = false;
    for (int r = 0; r < R; r++) {
      afgpu_ctx::RunTime &t = c.rt[r];
      t.frac_rejected = 0.99 * t.frac_rejected;
      if (n_att[r] > 1) t.frac_rejected = t.frac_rejected + 0.01;
      const double dt_used = dt[r];
      double tmp = c.p.dt_max_growth_factor;
      if (t.frac_rejected > 0.1) tmp = 1.0;
      const double dtn = std::min(tmp * t.global_dt, c.p.dt_safety_factor * std::min(dt_lim[r], c.p.dt_max));
      t.global_dt = dtn;
      t.global_time = t.time;
      t.dt = dtn;
      too_small = too_small || t.global_dt < c.p.dt_min;
      if (info) {
        afgpu_step_info_t &f = info[r];
        f.it = t.it;
        f.time = t.time;
        f.dt_used = dt_used;
        f.dt_next = dtn;
        f.dt_lim = dt_lim[r];
        f.n_rejected = n_rej[r];
        f.n_vcycles = c.run_ncyc[r];
        f.residual = c.run_res[r];
        f.rel_residual = max_rhs[r] > 0 ? c.run_res[r] / max_rhs[r] : 0;
        f.photoi_done = photo[r];
        f.photoi_cycles = photo[r] ? c.run_pcyc[r] : 0;
        f.leaf_cells = c.run_leaf_cells[r];
      }
    }
    if (too_small) {
      set_err("dt too small");
      return AFGPU_ERR_DT_TOO_SMALL;
    }
    return AFGPU_OK;
  }

  int restrict_tree(int iv) override { return restrict_pool(c.cc[iv]); }
  int restrict_pool(double *var) override {
    for (int lvl = c.highest_lvl; lvl >= 2; lvl--) {
      const List &l = c.l_ids[lvl];
      RCHK(with_bb(KC_RESTRICT, [&](auto B) {
        constexpr int bb = decltype(B)::value;
        return items<bb>(KC_RESTRICT, k_restrict<D, NC, BS, double, bb>, l.n, c.list(l), (const DBox *)c.d_boxes,
                         (const double *)var, var, 1, (int)(c.coord == AFGPU_CYL), 1);
      }));
    }
    return AFGPU_OK;
  }
  int gc_pool(int lvl, double *var, const VarBc &m, bool corners) override { return gc_var(lvl, var, m, corners); }
};

// Workgroup size of the box kernels of a handle's (dimension, box size).
static int box_threads(int D, int nc) { return (D == 2 && nc == 8) ? 64 : 256; }

int make_impl(afgpu_ctx &c) {
  if (c.D == 2 && c.nc == 8)
    c.impl.reset(new Impl<2, 8, 64>(c));
  else if (c.D == 2 && c.nc == 16)
    c.impl.reset(new Impl<2, 16, 256>(c));
  else if (c.D == 3 && c.nc == 8)
    c.impl.reset(new Impl<3, 8, 256>(c));
  else if (c.D == 3 && c.nc == 16)
    c.impl.reset(new Impl<3, 16, 256>(c));
  else {
    set_err("box size must be 8 or 16");
    return AFGPU_ERR_UNSUPPORTED;
  }
  return AFGPU_OK;
}

int ready(afgpu_t *h, bool need_tables = false) {
  if (!h) {
    set_err("null handle");
    return AFGPU_ERR_ARG;
  }
  if (h->highest_lvl < 1 || !h->have_params) {
    set_err("tree or parameters not set");
    return AFGPU_ERR_STATE;
  }
  if (need_tables && !h->have_tables) {
    set_err("lookup tables not set");
    return AFGPU_ERR_STATE;
  }
  return AFGPU_OK;
}

int sync_stream(afgpu_t *h) {
  HIPCHK(hipStreamSynchronize(h->stream));
  return timer_harvest(*h);
}

}  // namespace

// Scratch of the entry points that allocate nothing (the per-box partial
// sums of afgpu_regression_values_noalloc): AFGPU_NOALLOC_SCRATCH_WORDS
// doubles (32 MiB) in the code object's device memory, one copy per GPU
// that the library runs on, held until the process ends.  Host code never
// passes this name to a kernel (on the host it is a stand-in, not a device
// address): it uses the address hipGetSymbolAddress returns for the current
// device, cached per device by noalloc_scratch_address.
static __device__ __align__(256) double noalloc_scratch[AFGPU_NOALLOC_SCRATCH_WORDS];

static int noalloc_scratch_address(double **out) {
  static std::mutex mu;
  static std::vector<double *> addr;  // by device ordinal
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if ((size_t)dev >= addr.size()) addr.resize(dev + 1, nullptr);
  if (!addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess || !p) {
      set_err(std::string("hipGetSymbolAddress(noalloc_scratch): ") + hipGetErrorString(e));
      return AFGPU_ERR_HIP;
    }
    addr[dev] = static_cast<double *>(p);
  }
  *out = addr[dev];
  return AFGPU_OK;
}

extern "C" {

const char *afgpu_last_error(void) { return g_err.c_str(); }

int afgpu_create(int ndim, int n_cell, int coord, int precision, afgpu_t **out) {
  if (!out || (ndim != 2 && ndim != 3) || (coord != AFGPU_XYZ && coord != AFGPU_CYL) ||
      (coord == AFGPU_CYL && ndim != 2) || (precision != AFGPU_PREC_FP64 && precision != AFGPU_PREC_MIXED)) {
    set_err("afgpu_create: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  std::unique_ptr<afgpu_t> h(new afgpu_t());
  h->D = ndim;
  h->nc = n_cell;
  h->coord = coord;
  h->precision = precision;
  h->ncell = 1;
  size_t nf = 1;
  for (int d = 0; d < ndim; d++) {
    h->ncell *= (n_cell + 2);
    nf *= (n_cell + 1);
  }
  h->nface = nf * ndim;
  RCHK(make_impl(*h));
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (ndev < 1) {
    set_err("no GPU device");
    return AFGPU_ERR_HIP;
  }
  {
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (ncu > 0) h->n_cu = ncu;
  }
  RCHK(noalloc_scratch_address(&h->scratch));
  // Default shapes: a box of 8 x 8 cells is one wavefront, and streaming
  // such boxes four to a workgroup from a persistent grid of 8 workgroups
  // per CU moves ~1.6x the bytes per second of one box per workgroup; larger
  // boxes fill a workgroup each.  afgpu_autotune measures the alternatives.
  for (auto &sh : h->shape) sh = (ndim == 2 && n_cell == 8) ? Shape{4, 8} : Shape{1, 0};
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  h->cur = h->stream;
  for (int n = 0; n < AFGPU_MAX_MODES; n++) {
    HIPCHK(hipStreamCreateWithFlags(&h->mstream[n], hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&h->join_ev[n], hipEventDisableTiming));
  }
  HIPCHK(hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming));
  *out = h.release();
  return AFGPU_OK;
}

int afgpu_destroy(afgpu_t *h) {
  if (!h) return AFGPU_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graphs(*h);
  for (auto e : h->ev_pool) (void)hipEventDestroy(e);
  for (int i = 0; i < AFGPU_N_CC; i++) (void)hipFree(h->cc[i]);
  for (int i = 0; i < AFGPU_N_FC; i++) (void)hipFree(h->fc[i]);
  for (int i = 0; i < 4; i++) (void)hipFree(h->cc32[i]);
  (void)hipFree(h->dir.d_map);
  (void)hipFree(h->d_ref_ids);
  (void)hipFree(h->d_ref_flags);
  (void)hipFree(h->d_diag);
  (void)hipFree(h->d_io);
  for (double *x : h->xsp) (void)hipFree(x);
  (void)hipFree(h->dir.A0);
  (void)hipFree(h->dir.A1);
  for (auto &f : h->dir.f) {
    (void)hipFree(f.vinv);
    (void)hipFree(f.v);
    (void)hipFree(f.vinvT);
    (void)hipFree(f.vT);
    (void)hipFree(f.eig);
    (void)hipFree(f.vinv32);
    (void)hipFree(f.v32);
    (void)hipFree(f.eig32);
    (void)hipFree(f.mf);
    (void)hipFree(f.bf);
    (void)hipFree(f.eigf);
  }
  (void)hipFree(h->d_boxes);
  (void)hipFree(h->d_lists);
  (void)hipFree(h->d_td);
  (void)hipFree(h->d_chem);
  (void)hipFree(h->d_red);
  (void)hipFree(h->d_red_out);
  if (h->h_red) (void)hipHostFree(h->h_red);
  (void)hipFree(h->d_run_of);
  (void)hipFree(h->d_run_volt);
  (void)hipFree(h->d_run_active);
  (void)hipFree(h->d_run_host);
  (void)hipFree(h->d_run_maxlvl);
  (void)hipFree(h->d_run_bits);
  (void)hipFree(h->d_dt_run);
  (void)hipFree(h->d_step);
  (void)hipFree(h->d_run_ones);
  (void)hipFree(h->d_run_min_dr);
  if (h->h_step) (void)hipHostFree(h->h_step);
  if (h->h_run_active) (void)hipHostFree(h->h_run_active);
  if (h->h_dt_run) (void)hipHostFree(h->h_dt_run);
  for (auto &o : h->ops) {
    (void)hipFree(o.d_vinv);
    (void)hipFree(o.d_v);
    (void)hipFree(o.d_eig);
  }
  for (int n = 0; n < AFGPU_MAX_MODES; n++) {
    if (h->mstream[n]) (void)hipStreamDestroy(h->mstream[n]);
    if (h->join_ev[n]) (void)hipEventDestroy(h->join_ev[n]);
    for (int k = 0; k < 4; k++) (void)hipFree(h->mres[n].cc32[k]);
    (void)hipFree(h->mres[n].A0);
    (void)hipFree(h->mres[n].A1);
  }
  if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return AFGPU_OK;
}

int afgpu_sync_tree(afgpu_t *h, int n_boxes, const int *lvl, const int *ix, const int *parent, const int *children,
                    const int *neighbors, const int *neighbor_mat, const double *r_min, const double *dr) {
  if (!h || n_boxes < 1 || !lvl || !ix || !parent || !children || !neighbors || !neighbor_mat || !r_min || !dr) {
    set_err("afgpu_sync_tree: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  const int D = h->D, nch = 1 << D, nnb = 2 * D, nm = (D == 3) ? 27 : 9;
  std::vector<DBox> hb(n_boxes);
  int hl = 0;
  for (int b = 0; b < n_boxes; b++) {
    DBox &x = hb[b];
    std::memset(&x, 0, sizeof(x));
    x.lvl = lvl[b];
    x.parent = parent[b];
    for (int d = 0; d < D; d++) {
      x.ix[d] = ix[b * D + d];
      x.r_min[d] = r_min[b * D + d];
      x.dr[d] = dr[b * D + d];
    }
    for (int q = 0; q < nch; q++) x.children[q] = children[b * nch + q];
    for (int q = 0; q < nnb; q++) x.nb[q] = neighbors[b * nnb + q];
    for (int q = 0; q < nm; q++) x.nbmat[q] = neighbor_mat[b * nm + q];
    // validate references
    auto bad = [&](int v) { return v > n_boxes; };
    bool err = bad(x.parent);
    for (int q = 0; q < nch; q++) err |= bad(x.children[q]) || x.children[q] < 0;
    for (int q = 0; q < nnb; q++) err |= bad(x.nb[q]);
    for (int q = 0; q < nm; q++) err |= bad(x.nbmat[q]);
    if (err || x.lvl < 0) {
      set_err("afgpu_sync_tree: box reference out of range");
      return AFGPU_ERR_ARG;
    }
    hl = std::max(hl, x.lvl);
  }
  // Level lists (lvl_t ids / leaves / parents) and derived lists.
  std::vector<std::vector<int>> ids(hl + 1), leaves(hl + 1), parents(hl + 1);
  std::vector<int> refb;
  for (int b = 0; b < n_boxes; b++) {
    const DBox &x = hb[b];
    if (x.lvl <= 0) continue;
    ids[x.lvl].push_back(b + 1);
    if (x.children[0] > 0)
      parents[x.lvl].push_back(b + 1);
    else
      leaves[x.lvl].push_back(b + 1);
  }
  std::vector<char> lvl_rb(hl + 2, 0);
  for (int l = 1; l <= hl; l++)
    for (int id : ids[l]) {
      const DBox &x = hb[id - 1];
      bool rb = false;
      for (int q = 0; q < nnb; q++) rb |= (x.nb[q] == 0);
      if (rb) lvl_rb[l] = 1;
      if (rb && x.parent > 0 && x.children[0] <= 0) refb.push_back(id);
    }
  if (hl < 1 || ids[1].empty()) {
    set_err("afgpu_sync_tree: no level-1 box");
    return AFGPU_ERR_ARG;
  }
  // Runs: every level-1 box roots a run of its own (coarse_grid_size =
  // box_size; several level-1 boxes form a forest of independent runs, all
  // with the same level-1 geometry); boxes inherit their parent's run.
  const int n_runs = (int)ids[1].size();
  for (int q = 0; q < n_runs; q++) {
    DBox &x = hb[ids[1][q] - 1];
    for (int f = 0; f < nnb; f++)
      if (x.nb[f] >= 0) {
        set_err("level 1 must consist of one box per run (coarse_grid_size = box_size)");
        return AFGPU_ERR_UNSUPPORTED;
      }
    const DBox &x0 = hb[ids[1][0] - 1];
    for (int d = 0; d < D; d++)
      if (x.r_min[d] != x0.r_min[d] || x.dr[d] != x0.dr[d]) {
        set_err("afgpu_sync_tree: the runs of a forest must share the level-1 box geometry");
        return AFGPU_ERR_UNSUPPORTED;
      }
    x.run = q;
  }
  for (int l = 2; l <= hl; l++)
    for (int id : ids[l]) {
      DBox &x = hb[id - 1];
      if (x.parent < 1) {
        set_err("box above level 1 without parent");
        return AFGPU_ERR_ARG;
      }
      x.run = hb[x.parent - 1].run;
    }
  std::vector<double> min_dr(n_runs, std::numeric_limits<double>::max());
  std::vector<int> run_of(n_boxes, 0), maxlvl(n_runs, 0);
  for (int b = 0; b < n_boxes; b++) {
    const DBox &x = hb[b];
    run_of[b] = x.run;
    if (x.lvl <= 0) continue;
    maxlvl[x.run] = std::max(maxlvl[x.run], x.lvl);
    for (int d = 0; d < D; d++) min_dr[x.run] = std::min(min_dr[x.run], x.dr[d]);
  }
  // Structural checks that every kernel relies on: a parent has all 2^D
  // children (and they point back), level-1 faces are physical, and every
  // refinement-boundary face has the parent's neighbour.
  for (int b = 0; b < n_boxes; b++) {
    const DBox &x = hb[b];
    if (x.lvl <= 0) continue;
    if (x.children[0] > 0)
      for (int q = 0; q < nch; q++) {
        const int cid = x.children[q];
        if (cid < 1 || hb[cid - 1].parent != b + 1 || hb[cid - 1].lvl != x.lvl + 1) {
          set_err("afgpu_sync_tree: inconsistent children");
          return AFGPU_ERR_ARG;
        }
      }
    for (int q = 0; q < nnb; q++) {
      const int nid = x.nb[q];
      if (nid > 0 && hb[nid - 1].lvl != x.lvl) {
        set_err("afgpu_sync_tree: neighbour on another level");
        return AFGPU_ERR_ARG;
      }
      if (x.lvl == 1 && nid == 0) {
        set_err("afgpu_sync_tree: level-1 box without neighbour or boundary");
        return AFGPU_ERR_ARG;
      }
    }
    for (int q = 0; q < nm; q++) {
      const int nid = x.nbmat[q];
      if (nid > 0 && hb[nid - 1].lvl != x.lvl) {
        set_err("afgpu_sync_tree: diagonal neighbour on another level");
        return AFGPU_ERR_ARG;
      }
    }
  }
  for (int l = 2; l <= hl; l++)
    for (int id : ids[l]) {
      const DBox &x = hb[id - 1];
      if (x.parent < 1) {
        set_err("box above level 1 without parent");
        return AFGPU_ERR_ARG;
      }
      for (int q = 0; q < nnb; q++)
        if (x.nb[q] == 0 && hb[x.parent - 1].nb[q] <= 0) {
          set_err("tree violates 2:1 balance at a refinement boundary");
          return AFGPU_ERR_ARG;
        }
    }
  for (int b = 0; b < n_boxes; b++) {
    DBox &x = hb[b];
    for (int q = 0; q < nnb; q++) x.pnb[q] = (x.nb[q] == 0 && x.parent > 0) ? hb[x.parent - 1].nb[q] : 0;
  }
  h->src_ids = ids;
  h->src_leaves = leaves;
  h->src_parents = parents;
  h->src_refb = refb;
  {
    std::vector<std::vector<int>> rbnb(hl + 1);
    for (int l = 2; l <= hl; l++)
      for (int id : ids[l]) {
        const DBox &x = hb[id - 1];
        for (int q = 0; q < nnb; q++)
          if (x.nb[q] == 0 && x.pnb[q] > 0) rbnb[l - 1].push_back(x.pnb[q]);
      }
    for (auto &v : rbnb) {
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
    }
    h->src_rbnb = rbnb;
  }
  build_lists(*h, h->shape[KC_GSRB].bb);
  const std::vector<int> &all = h->h_lists;
  HIPCHK(hipStreamSynchronize(h->stream));
  // Grow the pools, keeping the data of existing box ids.
  if (n_boxes > h->cap_boxes) {
    const int cap = std::max(n_boxes, h->cap_boxes + h->cap_boxes / 4);
    auto grow = [&](double *&ptr, size_t per_box) -> int {
      double *np = nullptr;
      HIPCHK(hipMalloc(&np, (size_t)cap * per_box * sizeof(double)));
      HIPCHK(hipMemsetAsync(np, 0, (size_t)cap * per_box * sizeof(double), h->stream));
      if (ptr) {
        HIPCHK(hipMemcpyAsync(np, ptr, (size_t)h->cap_boxes * per_box * sizeof(double), hipMemcpyDeviceToDevice,
                              h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipFree(ptr));
      }
      ptr = np;
      return AFGPU_OK;
    };
    for (int i = 0; i < AFGPU_N_CC; i++) RCHK(grow(h->cc[i], h->ncell));
    for (int i = 0; i < AFGPU_N_FC; i++) RCHK(grow(h->fc[i], h->nface));
    for (double *&x : h->xsp) RCHK(grow(x, h->ncell));
    if (h->precision == AFGPU_PREC_MIXED) {
      for (int i = 0; i < 4; i++) {
        (void)hipFree(h->cc32[i]);
        HIPCHK(hipMalloc(&h->cc32[i], (size_t)cap * h->ncell * sizeof(float)));
        HIPCHK(hipMemsetAsync(h->cc32[i], 0, (size_t)cap * h->ncell * sizeof(float), h->stream));
      }
    }
    h->cap_boxes = cap;
  }
  // The single-precision corrections start every cycle from zero on the
  // leaves; box roles may have changed, so clear them.
  if (h->precision == AFGPU_PREC_MIXED) {
    const size_t bytes = (size_t)h->cap_boxes * h->ncell * sizeof(float);
    for (int i = 0; i < 4; i++)
      if (h->cc32[i]) HIPCHK(hipMemsetAsync(h->cc32[i], 0, bytes, h->stream));
    for (auto &r : h->mres)
      for (int k = 0; k < 4; k++)
        if (r.cc32[k] && r.cap_boxes == h->cap_boxes) HIPCHK(hipMemsetAsync(r.cc32[k], 0, bytes, h->stream));
  }
  (void)hipFree(h->d_boxes);
  (void)hipFree(h->d_lists);
  h->d_boxes = nullptr;
  h->d_lists = nullptr;
  HIPCHK(hipMalloc(&h->d_boxes, (size_t)n_boxes * sizeof(DBox)));
  HIPCHK(hipMemcpy(h->d_boxes, hb.data(), (size_t)n_boxes * sizeof(DBox), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&h->d_lists, std::max<size_t>(1, all.size()) * sizeof(int)));
  if (!all.empty()) HIPCHK(hipMemcpy(h->d_lists, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice));
  // per-run buffers
  (void)hipFree(h->d_run_of);
  h->d_run_of = nullptr;
  HIPCHK(hipMalloc(&h->d_run_of, (size_t)n_boxes * sizeof(int)));
  HIPCHK(hipMemcpy(h->d_run_of, run_of.data(), (size_t)n_boxes * sizeof(int), hipMemcpyHostToDevice));
  if (n_runs > h->red_cap_runs) {
    (void)hipFree(h->d_red);
    (void)hipFree(h->d_red_out);
    (void)hipFree(h->d_run_volt);
    (void)hipFree(h->d_run_active);
    (void)hipFree(h->d_run_host);
    (void)hipFree(h->d_run_maxlvl);
    (void)hipFree(h->d_dt_run);
    (void)hipFree(h->d_step);
    (void)hipFree(h->d_run_ones);
    (void)hipFree(h->d_run_min_dr);
    if (h->h_step) (void)hipHostFree(h->h_step);
    h->d_step = h->h_step = nullptr;
    h->d_run_ones = nullptr;
    h->d_run_min_dr = nullptr;
    if (h->h_red) (void)hipHostFree(h->h_red);
    if (h->h_run_active) (void)hipHostFree(h->h_run_active);
    if (h->h_dt_run) (void)hipHostFree(h->h_dt_run);
    h->d_red = h->d_red_out = nullptr;
    h->d_run_volt = h->d_dt_run = nullptr;
    h->d_run_active = h->d_run_host = h->d_run_maxlvl = h->d_run_bits = nullptr;
    h->h_red = nullptr;
    h->h_run_active = nullptr;
    h->h_dt_run = nullptr;
    h->red_cap_runs = 0;
    const size_t nr = (size_t)n_runs;
    HIPCHK(hipMalloc(&h->d_red, nr * RED_STRIPES * N_RED_SLOTS * sizeof(unsigned long long)));
    HIPCHK(hipMalloc(&h->d_red_out, nr * N_RED_SLOTS * sizeof(unsigned long long)));
    HIPCHK(hipHostMalloc(&h->h_red, nr * N_RED_SLOTS * sizeof(unsigned long long), hipHostMallocDefault));
    HIPCHK(hipMalloc(&h->d_run_volt, nr * sizeof(double)));
    HIPCHK(hipMalloc(&h->d_run_active, nr * sizeof(int)));
    HIPCHK(hipMalloc(&h->d_run_host, nr * sizeof(int)));
    HIPCHK(hipMalloc(&h->d_run_maxlvl, nr * sizeof(int)));
    HIPCHK(hipMalloc(&h->d_run_bits, 2 * nr * sizeof(int)));
    HIPCHK(hipHostMalloc(&h->h_run_active, nr * sizeof(int), hipHostMallocDefault));
    HIPCHK(hipMalloc(&h->d_dt_run, nr * sizeof(double)));
    HIPCHK(hipHostMalloc(&h->h_dt_run, nr * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipMalloc(&h->d_step, nr * sizeof(RunStep)));
    HIPCHK(hipHostMalloc(&h->h_step, nr * sizeof(RunStep), hipHostMallocDefault));
    HIPCHK(hipMalloc(&h->d_run_ones, nr * sizeof(int)));
    HIPCHK(hipMalloc(&h->d_run_min_dr, nr * sizeof(double)));
    h->red_cap_runs = n_runs;
  }
  if ((int)h->run_volt.size() != n_runs) h->run_volt.assign(n_runs, h->voltage);
  HIPCHK(hipMemcpy(h->d_run_volt, h->run_volt.data(), (size_t)n_runs * sizeof(double), hipMemcpyHostToDevice));
  if ((int)h->rt.size() != n_runs) h->rt.assign(n_runs, afgpu_ctx::RunTime());
  HIPCHK(hipMemcpy(h->d_run_maxlvl, maxlvl.data(), (size_t)n_runs * sizeof(int), hipMemcpyHostToDevice));
  {
    // per run, the levels with leaves and the levels with refinement-
    // boundary boxes (the per-run form of the level lists and lvl_rb)
    std::vector<int> bits(2 * (size_t)n_runs, 0);
    for (int l = 1; l <= hl; l++) {
      for (int id : leaves[l]) bits[hb[id - 1].run] |= 1 << l;
      for (int id : ids[l]) {
        const DBox &x = hb[id - 1];
        for (int q = 0; q < nnb; q++)
          if (x.nb[q] == 0) bits[n_runs + x.run] |= 1 << l;
      }
    }
    HIPCHK(hipMemcpy(h->d_run_bits, bits.data(), bits.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  {
    std::vector<int> ones(n_runs, 1);
    HIPCHK(hipMemcpy(h->d_run_host, ones.data(), (size_t)n_runs * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_run_ones, ones.data(), (size_t)n_runs * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(h->d_run_min_dr, min_dr.data(), (size_t)n_runs * sizeof(double), hipMemcpyHostToDevice));
  }
  h->step_k = 4;
  h->uneven = *std::min_element(maxlvl.begin(), maxlvl.end()) != *std::max_element(maxlvl.begin(), maxlvl.end());
  h->n_runs = n_runs;
  h->lvl_rb = lvl_rb;
  h->run_min_dr = min_dr;
  h->run_leaf_cells.assign(n_runs, 0);
  for (int l = 1; l <= hl; l++)
    for (int id : leaves[l]) {
      long long cells = 1;
      for (int d = 0; d < D; d++) cells *= h->nc;
      h->run_leaf_cells[hb[id - 1].run] += cells;
    }
  h->masked = false;
  h->base_active.clear();
  h->hboxes = hb;
  h->n_boxes = n_boxes;
  h->highest_lvl = hl;
  for (auto &o : h->ops) o.ok = false;  // level-1 box may have changed
  h->dir.set = false;
  for (auto &f : h->dir.f) f.ok = false;
  drop_graphs(*h);
  return AFGPU_OK;
}

static int xfer_pool(afgpu_t *h, double *pool, size_t blk, int n_ids, const int *ids, double *data, bool put);
static int xfer(afgpu_t *h, bool cc, int iv, int n_ids, const int *ids, double *data, bool put) {
  if (!h || (cc && (iv < 0 || iv >= AFGPU_N_CC)) || (!cc && (iv < 0 || iv >= AFGPU_N_FC))) {
    set_err("transfer: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  return xfer_pool(h, cc ? h->cc[iv] : h->fc[iv], cc ? h->ncell : h->nface, n_ids, ids, data, put);
}
static int xfer_pool(afgpu_t *h, double *pool, size_t blk, int n_ids, const int *ids, double *data, bool put) {
  if (n_ids < 0 || (n_ids > 0 && (!ids || !data))) {
    set_err("transfer: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  if (n_ids == 0) return AFGPU_OK;
  for (int n = 0; n < n_ids; n++)
    if (ids[n] < 1 || ids[n] > h->n_boxes) {
      set_err("transfer: box id out of range");
      return AFGPU_ERR_ARG;
    }
  // Contiguous runs of ids move in one copy each.
  int n = 0;
  while (n < n_ids) {
    int m = n + 1;
    while (m < n_ids && ids[m] == ids[m - 1] + 1) m++;
    double *dev = pool + (size_t)(ids[n] - 1) * blk;
    double *host = data + (size_t)n * blk;
    const size_t bytes = (size_t)(m - n) * blk * sizeof(double);
    if (put)
      HIPCHK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, h->stream));
    else
      HIPCHK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, h->stream));
    n = m;
  }
  return sync_stream(h);
}

int afgpu_put_cc(afgpu_t *h, int iv, int n, const int *ids, const double *d) {
  return xfer(h, true, iv, n, ids, const_cast<double *>(d), true);
}
int afgpu_get_cc(afgpu_t *h, int iv, int n, const int *ids, double *d) { return xfer(h, true, iv, n, ids, d, false); }
int afgpu_put_fc(afgpu_t *h, int iv, int n, const int *ids, const double *d) {
  return xfer(h, false, iv, n, ids, const_cast<double *>(d), true);
}
int afgpu_get_fc(afgpu_t *h, int iv, int n, const int *ids, double *d) { return xfer(h, false, iv, n, ids, d, false); }

int afgpu_set_params(afgpu_t *h, const afgpu_params_t *p) {
  if (!h || !p || p->photoi_n_modes < 0 || p->photoi_n_modes > AFGPU_MAX_MODES ||
      (p->time_integrator != AFGPU_FORWARD_EULER && p->time_integrator != AFGPU_HEUNS_METHOD) ||
      p->photoi_species < 0 || p->photoi_species >= SP_MAX || p->mg_n_cycle_down < 1 || p->mg_n_cycle_up < 1) {
    set_err("afgpu_set_params: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  drop_graphs(*h);
  h->p = *p;
  h->have_params = true;
  h->ops[AFGPU_OP_POISSON].lambda2 = 0.0;
  for (int n = 0; n < p->photoi_n_modes; n++)
    h->ops[AFGPU_OP_HELMH + n].lambda2 = p->photoi_lambda[n] * p->photoi_lambda[n];
  return AFGPU_OK;
}

int afgpu_set_tables(afgpu_t *h, int n_rows, double td_x_min, double td_inv_fac, const double *td_cols,
                     double chem_x_min, double chem_inv_fac, const double *chem_cols) {
  if (!h || n_rows < 2 || !td_cols || !chem_cols) {
    set_err("afgpu_set_tables: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  (void)hipFree(h->d_td);
  (void)hipFree(h->d_chem);
  h->d_td = h->d_chem = nullptr;
  HIPCHK(hipMalloc(&h->d_td, 4 * (size_t)n_rows * sizeof(double)));
  HIPCHK(hipMalloc(&h->d_chem, 2 * (size_t)n_rows * sizeof(double)));
  HIPCHK(hipMemcpy(h->d_td, td_cols, 4 * (size_t)n_rows * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_chem, chem_cols, 2 * (size_t)n_rows * sizeof(double), hipMemcpyHostToDevice));
  drop_graphs(*h);
  h->td = DTable{h->d_td, n_rows, td_x_min, td_inv_fac};
  h->chem = DTable{h->d_chem, n_rows, chem_x_min, chem_inv_fac};
  h->have_tables = true;
  return AFGPU_OK;
}

int afgpu_set_voltage(afgpu_t *h, double v) {
  if (!h) return AFGPU_ERR_ARG;
  if (v != h->voltage) drop_graphs(*h);  // the boundary value is a captured argument
  h->voltage = v;
  if (!h->run_volt.empty()) {
    std::vector<double> all(h->run_volt.size(), v);
    return afgpu_set_run_voltages(h, (int)all.size(), all.data());
  }
  return AFGPU_OK;
}

int afgpu_num_runs(afgpu_t *h) { return h ? h->n_runs : 0; }

int afgpu_set_run_voltages(afgpu_t *h, int n_runs, const double *voltages) {
  if (!h || !voltages || h->highest_lvl < 1 || n_runs != h->n_runs) {
    set_err("afgpu_set_run_voltages: invalid arguments (tree synchronised? one value per run)");
    return AFGPU_ERR_ARG;
  }
  RCHK(sync_stream(h));  // kernels in flight may read the previous values
  h->run_volt.assign(voltages, voltages + n_runs);
  HIPCHK(hipMemcpy(h->d_run_volt, h->run_volt.data(), (size_t)n_runs * sizeof(double), hipMemcpyHostToDevice));
  return AFGPU_OK;
}

int afgpu_field_compute_runs(afgpu_t *h, int s_in, int have_guess, int n_runs, int *n_cycles, double *residual) {
  RCHK(ready(h));
  if (s_in < 0 || s_in > 2 || n_runs != h->n_runs) return AFGPU_ERR_ARG;
  RCHK(h->impl->field_compute(s_in, have_guess != 0, nullptr, nullptr));
  for (int r = 0; r < n_runs; r++) {
    if (n_cycles) n_cycles[r] = h->run_ncyc[r];
    if (residual) residual[r] = h->run_res[r];
  }
  return sync_stream(h);
}

int afgpu_get_run_time_state(afgpu_t *h, int run, int *it, double *time, double *global_dt) {
  if (!h || run < 0 || run >= (int)h->rt.size()) return AFGPU_ERR_ARG;
  const afgpu_ctx::RunTime &t = h->rt[run];
  if (it) *it = t.it;
  if (time) *time = t.time;
  if (global_dt) *global_dt = t.started ? t.global_dt : h->p.dt_min;
  return AFGPU_OK;
}

int afgpu_set_run_time_state(afgpu_t *h, int run, int it, double time, double global_dt) {
  if (!h || run < 0 || run >= (int)h->rt.size() || it < 0 || !(global_dt > 0)) return AFGPU_ERR_ARG;
  afgpu_ctx::RunTime &t = h->rt[run];
  t.it = it;
  t.time = t.global_time = time;
  t.global_dt = t.dt = global_dt;
  t.started = true;
  return AFGPU_OK;
}

static int check_photoi_species(const afgpu_t *h);

int afgpu_step_ensemble(afgpu_t *h, int n_runs, afgpu_step_info_t *info) {
  RCHK(ready(h, true));
  if (n_runs != h->n_runs) {
    set_err("afgpu_step_ensemble: n_runs differs from the forest's run count");
    return AFGPU_ERR_ARG;
  }
  RCHK(check_photoi_species(h));
  RCHK(h->impl->step_runs(info));
  return sync_stream(h);
}

int afgpu_mg_fas_vcycle(afgpu_t *h, int op, int set_residual, int max_lvl) {
  RCHK(ready(h));
  if (op < 0 || op > AFGPU_MAX_MODES) return AFGPU_ERR_ARG;
  if (max_lvl <= 0 || max_lvl > h->highest_lvl) max_lvl = h->highest_lvl;
  RCHK(h->impl->vcycle(op, set_residual != 0, max_lvl));
  return sync_stream(h);
}

int afgpu_mg_fas_fmg(afgpu_t *h, int op, int set_residual, int have_guess) {
  RCHK(ready(h));
  if (op < 0 || op > AFGPU_MAX_MODES) return AFGPU_ERR_ARG;
  RCHK(h->impl->fmg(op, set_residual != 0, have_guess != 0));
  return sync_stream(h);
}

int afgpu_tree_maxabs_cc(afgpu_t *h, int iv, double *out) {
  RCHK(ready(h));
  if (iv < 0 || iv >= AFGPU_N_CC || !out) return AFGPU_ERR_ARG;
  return h->impl->maxabs(iv, out);
}

// Species k of a reaction list (iv = AFGPU_VAR_SPECIES(k)): its state-0
// pool, with the densities' boundary and refinement-boundary methods.
static double *species_pool(const afgpu_t *h, int iv) {
  const int k = iv - AFGPU_VAR_SPECIES(0);
  const SpeciesArgs S = h->species_args();
  return (k >= 0 && k < S.n) ? S.sp[k][0] : nullptr;
}

int afgpu_gc_tree(afgpu_t *h, int iv) {
  RCHK(ready(h));
  if (double *sp = species_pool(h, iv)) {
    const VarBc m = var_methods(*h, AFGPU_CC_ELECTRON);
    for (int lvl = 1; lvl <= h->highest_lvl; lvl++) RCHK(h->impl->gc_pool(lvl, sp, m, true));
    return sync_stream(h);
  }
  if (iv < 0 || iv >= AFGPU_N_CC || var_methods(*h, iv).rb == RB_NONE) return AFGPU_ERR_ARG;
  for (int lvl = 1; lvl <= h->highest_lvl; lvl++) RCHK(h->impl->gc_lvl(lvl, iv, true));
  return sync_stream(h);
}

int afgpu_restrict_tree(afgpu_t *h, int iv) {
  RCHK(ready(h));
  if (double *sp = species_pool(h, iv)) {
    RCHK(h->impl->restrict_pool(sp));
    return sync_stream(h);
  }
  if (iv < 0 || iv >= AFGPU_N_CC) return AFGPU_ERR_ARG;
  RCHK(h->impl->restrict_tree(iv));
  return sync_stream(h);
}

int afgpu_field_compute(afgpu_t *h, int s_in, int have_guess, int *n_cycles, double *residual) {
  RCHK(ready(h));
  if (s_in < 0 || s_in > 2) return AFGPU_ERR_ARG;
  RCHK(h->impl->field_compute(s_in, have_guess != 0, n_cycles, residual));
  return sync_stream(h);
}

int afgpu_field_from_potential(afgpu_t *h) {
  RCHK(ready(h));
  RCHK(h->impl->field_from_potential());
  return sync_stream(h);
}

// The species receiving the photoionization source exists in the model
// (the three standard species, or those of the reaction list).
static int check_photoi_species(const afgpu_t *h) {
  if (h->p.photoi_enabled && h->p.photoi_species >= h->n_sp) {
    set_err("photoionization species " + std::to_string(h->p.photoi_species) + " is not among the " +
            std::to_string(h->n_sp) + " species of the model");
    return AFGPU_ERR_ARG;
  }
  return AFGPU_OK;
}

int afgpu_forward_euler(afgpu_t *h, double dt, int s_deriv, int n_prev, const int *s_prev, const double *w_prev,
                        int s_out, int i_step, int n_steps, double *dt_lim) {
  RCHK(ready(h, true));
  if (!dt_lim || n_prev < 1 || n_prev > 4 || !s_prev || !w_prev || s_deriv < 0 || s_deriv > 2 || s_out < 0 ||
      s_out > 2)
    return AFGPU_ERR_ARG;
  for (int m = 0; m < n_prev; m++)
    if (s_prev[m] < 0 || s_prev[m] > 2) return AFGPU_ERR_ARG;
  RCHK(check_photoi_species(h));
  return h->impl->forward_euler(dt, s_deriv, n_prev, s_prev, w_prev, s_out, i_step, n_steps, dt_lim);
}

static int advance(afgpu_t *h, double dt, double *dt_lim) {
  *dt_lim = 1e100;
  if (h->p.time_integrator == AFGPU_FORWARD_EULER) {
    const int sp[1] = {0};
    const double w[1] = {1.0};
    RCHK(h->impl->forward_euler(dt, 0, 1, sp, w, 0, 1, 1, dt_lim));
  } else {
    const int sp1[1] = {0};
    const double w1[1] = {1.0};
    RCHK(h->impl->forward_euler(dt, 0, 1, sp1, w1, 1, 1, 2, dt_lim));
    const int sp2[2] = {0, 1};
    const double w2[2] = {0.5, 0.5};
    RCHK(h->impl->forward_euler(0.5 * dt, 1, 2, sp2, w2, 0, 2, 2, dt_lim));
  }
  h->time = h->time + dt;
  return AFGPU_OK;
}

int afgpu_advance(afgpu_t *h, double dt, double *dt_lim) {
  RCHK(ready(h, true));
  if (!dt_lim) return AFGPU_ERR_ARG;
  RCHK(check_photoi_species(h));
  return advance(h, dt, dt_lim);
}

int afgpu_photoi_set_src(afgpu_t *h) {
  RCHK(ready(h, true));
  RCHK(check_photoi_species(h));
  RCHK(h->impl->photoi());
  return sync_stream(h);
}

static int ensure_ref_buffers(afgpu_t *h, int n);

// consistent_ref_flags (afivo/src/m_af_core.f90:929-1013): the cell flags
// of every leaf and of every box with a leaf child (the reference evaluates
// such a box once, from its first leaf child) on the device, the rest here
// in the reference's order.
int afgpu_refine_flags(afgpu_t *h, const afgpu_refine_t *rp, int n_boxes, int *ref_flags) {
  RCHK(ready(h, true));
  if (!rp || !ref_flags || n_boxes != (int)h->hboxes.size() || rp->buffer_width < 0 ||
      rp->buffer_width > h->nc || rp->n_seeds < 0 || rp->n_seeds > AFGPU_REF_MAX_ITEMS || rp->n_regions < 0 ||
      rp->n_regions > AFGPU_REF_MAX_ITEMS || rp->n_limits < 0 || rp->n_limits > AFGPU_REF_MAX_ITEMS) {
    set_err("afgpu_refine_flags: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  const std::vector<DBox> &hb = h->hboxes;
  auto has_children = [&](int id) { return hb[id - 1].children[0] > 0; };
  const int nch = 1 << h->D, nnb = 2 * h->D;
  std::vector<int> eval;
  for (size_t l = 1; l < h->src_leaves.size(); l++) eval.insert(eval.end(), h->src_leaves[l].begin(), h->src_leaves[l].end());
  for (size_t l = 1; l < h->src_parents.size(); l++)
    for (int id : h->src_parents[l]) {
      bool leaf_child = false;
      for (int q = 0; q < nch; q++) leaf_child = leaf_child || !has_children(hb[id - 1].children[q]);
      if (leaf_child) eval.push_back(id);
    }
  RCHK(ensure_ref_buffers(h, n_boxes));
  const int unset = std::numeric_limits<int>::min();
  h->cur = h->stream;
  HIPCHK(hipMemcpyAsync(h->d_ref_ids, eval.data(), eval.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_fill_int, dim3((n_boxes + 255) / 256), dim3(256), 0, h->stream, h->d_ref_flags, n_boxes, unset);
  HIPCHK(hipGetLastError());
  RCHK(h->impl->refine_cells(*rp, h->d_ref_ids, (int)eval.size(), h->d_ref_flags));
  std::vector<int> f(n_boxes);
  HIPCHK(hipMemcpyAsync(f.data(), h->d_ref_flags, (size_t)n_boxes * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  RCHK(sync_stream(h));
  // unset flags, the level cap (:987-1000)
  for (int &x : f)
    if (x == unset) x = REF_KEEP;
  const int hl = (int)h->src_ids.size() - 1;
  if (rp->max_lvl >= 1 && rp->max_lvl <= hl)
    for (int id : h->src_ids[rp->max_lvl])
      if (f[id - 1] == REF_DO) f[id - 1] = REF_KEEP;
  // ensure_two_one_balance (:1016-1057)
  for (int lvl = hl; lvl >= 1; lvl--)
    for (int id : h->src_leaves[lvl]) {
      int &x = f[id - 1];
      const DBox &b = hb[id - 1];
      if (x == REF_DO || x == AFGPU_REF_REFINE) {
        x = AFGPU_REF_REFINE;
        for (int nb = 0; nb < nnb; nb++)
          if (b.nb[nb] == 0) f[hb[b.parent - 1].nb[nb] - 1] = AFGPU_REF_REFINE;
      } else if (x == REF_RM) {
        for (int nb = 0; nb < nnb; nb++) {
          const int nb_id = b.nb[nb];
          if (nb_id > 0 && (has_children(nb_id) || f[nb_id - 1] > REF_KEEP)) {
            x = REF_KEEP;
            break;
          }
        }
      }
    }
  // handle_derefinement_flags (:1059-1088)
  for (int lvl = hl - 1; lvl >= 1; lvl--)
    for (int id : h->src_parents[lvl]) {
      const DBox &b = hb[id - 1];
      bool all_refined = true, all_rm = true;
      for (int q = 0; q < nch; q++) {
        all_refined = all_refined && has_children(b.children[q]);
        all_rm = all_rm && f[b.children[q] - 1] == REF_RM;
      }
      if (all_refined) continue;
      if (all_rm && f[id - 1] <= REF_KEEP) {
        f[id - 1] = AFGPU_REF_DEREFINE;
      } else {
        f[id - 1] = REF_KEEP;
        for (int q = 0; q < nch; q++) {
          int &x = f[b.children[q] - 1];
          if (x != AFGPU_REF_DEREFINE) x = std::max(x, REF_KEEP);
        }
      }
    }
  std::copy(f.begin(), f.end(), ref_flags);
  return AFGPU_OK;
}

// The variables with methods, in the order the streamer sets them
// (photoi_initialize, field_initialize, mg_init, src/streamer.f90:84-86 and
// 103-108), and whether they are prolonged with af_prolong_limit (the
// densities and rhs: the streamer's default prolongation,
// src/m_streamer.f90:402-417) or af_prolong_linear.
// (every species has the densities' methods, src/m_streamer.f90:402-417)
static std::vector<VarRef> auto_vars(const afgpu_ctx &c) {
  std::vector<VarRef> v;
  if (c.p.photoi_enabled) v.push_back({c.cc[AFGPU_CC_PHOTO], AFGPU_CC_PHOTO, false});
  v.push_back({c.cc[AFGPU_CC_EFIELD], AFGPU_CC_EFIELD, false});
  v.push_back({c.cc[AFGPU_CC_PHI], AFGPU_CC_PHI, false});
  const SpeciesArgs S = c.species_args();
  for (int k = 0; k < S.n; k++) v.push_back({S.sp[k][0], AFGPU_CC_SPECIES(k < 3 ? k : 0), true});
  v.push_back({c.cc[AFGPU_CC_RHS], AFGPU_CC_RHS, true});
  return v;
}
static int ensure_ref_buffers(afgpu_t *h, int n) {
  if (n <= h->ref_cap) return AFGPU_OK;
  (void)hipFree(h->d_ref_ids);
  (void)hipFree(h->d_ref_flags);
  h->d_ref_ids = h->d_ref_flags = nullptr;
  h->ref_cap = 0;
  HIPCHK(hipMalloc(&h->d_ref_ids, (size_t)n * sizeof(int)));
  HIPCHK(hipMalloc(&h->d_ref_flags, (size_t)n * sizeof(int)));
  h->ref_cap = n;
  return AFGPU_OK;
}

// auto_restrict (afivo/src/m_af_core.f90:825-840) for the listed boxes,
// before their children are removed
int afgpu_restrict_boxes(afgpu_t *h, int n_ids, const int *ids) {
  RCHK(ready(h));

  const int n_boxes = (int)h->hboxes.size(), nch = 1 << h->D;
  if (n_ids < 0 || (n_ids > 0 && !ids)) return AFGPU_ERR_ARG;
  std::vector<int> ch;
  for (int q = 0; q < n_ids; q++) {
    const int id = ids[q];
    if (id < 1 || id > n_boxes || h->hboxes[id - 1].lvl < 1 || h->hboxes[id - 1].children[0] <= 0) {
      set_err("afgpu_restrict_boxes: not a box with children");
      return AFGPU_ERR_ARG;
    }
    for (int k = 0; k < nch; k++) ch.push_back(h->hboxes[id - 1].children[k]);
  }
  if (ch.empty()) return AFGPU_OK;
  RCHK(ensure_ref_buffers(h, std::max((int)ch.size(), n_boxes)));
  h->cur = h->stream;
  HIPCHK(hipMemcpyAsync(h->d_ref_ids, ch.data(), ch.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
  for (const auto &v : auto_vars(*h)) RCHK(h->impl->restrict_ids(h->d_ref_ids, (int)ch.size(), v.p));
  return sync_stream(h);
}

// auto_prolong (afivo/src/m_af_core.f90:843-881) for the listed new boxes
// of the synchronised tree: level by level, the values from the parents,
// then the new boxes' ghost cells (af_gc_box with corners)
int afgpu_prolong_boxes(afgpu_t *h, int n_ids, const int *ids) {
  RCHK(ready(h));

  const int n_boxes = (int)h->hboxes.size();
  if (n_ids < 0 || (n_ids > 0 && !ids)) return AFGPU_ERR_ARG;
  std::vector<int> sorted;
  for (int q = 0; q < n_ids; q++) {
    const int id = ids[q];
    if (id < 1 || id > n_boxes || h->hboxes[id - 1].lvl < 2 || h->hboxes[id - 1].parent < 1) {
      set_err("afgpu_prolong_boxes: not a box with a parent");
      return AFGPU_ERR_ARG;
    }
    sorted.push_back(id);
  }
  if (sorted.empty()) return AFGPU_OK;
  std::stable_sort(sorted.begin(), sorted.end(),
                   [&](int a, int b) { return h->hboxes[a - 1].lvl < h->hboxes[b - 1].lvl; });
  RCHK(ensure_ref_buffers(h, std::max((int)sorted.size(), n_boxes)));
  h->cur = h->stream;
  HIPCHK(hipMemcpyAsync(h->d_ref_ids, sorted.data(), sorted.size() * sizeof(int), hipMemcpyHostToDevice,
                        h->stream));
  const auto vars = auto_vars(*h);
  for (size_t a = 0; a < sorted.size();) {
    size_t b = a;
    while (b < sorted.size() && h->hboxes[sorted[b] - 1].lvl == h->hboxes[sorted[a] - 1].lvl) b++;
    const int *d = h->d_ref_ids + a;
    const int n = (int)(b - a);
    // (in groups of at most NEW_MAX_VARS variables per launch)
    for (size_t k0 = 0; k0 < vars.size(); k0 += NEW_MAX_VARS) {
      const int nv = (int)std::min<size_t>(NEW_MAX_VARS, vars.size() - k0);
      RCHK(h->impl->prolong_new(d, n, vars.data() + k0, nv));
      RCHK(h->impl->gc_ids(d, n, vars.data() + k0, nv));
    }
    a = b;
  }
  return sync_stream(h);
}

// af_total_volume (afivo/src/m_af_types.f90:805-825) of one run: its
// level-1 box (coarse_grid_size = box_size)
static double run_volume(const afgpu_t *h, int run) {
  const DBox *b1 = nullptr;
  for (int id : h->src_ids[1])
    if (h->hboxes[id - 1].run == run) b1 = &h->hboxes[id - 1];
  const int D = h->D;
  double len[3];
  for (int d = 0; d < D; d++) len[d] = h->nc * b1->dr[d];
  if (D == 2 && h->coord == AFGPU_CYL) {
    const double pi = 3.141592653589793;  // acos(-1)
    const double r0 = b1->r_min[0], r1 = r0 + len[0];
    return pi * (r1 * r1 - r0 * r0) * len[D - 1];
  }
  double p = 1;
  for (int d = 0; d < D; d++) p *= len[d];
  return p;
}

static int check_vars(const afgpu_t *h, int n_vars, const int *ivs, VarPtrs &vp, const char *who) {
  if (n_vars < 1 || n_vars > AFGPU_DIAG_MAX_VARS || !ivs) {
    set_err(std::string(who) + ": 1 to AFGPU_DIAG_MAX_VARS variables");
    return AFGPU_ERR_ARG;
  }
  vp.n = n_vars;
  const SpeciesArgs S = h->species_args();
  for (int v = 0; v < n_vars; v++) {
    const int k = ivs[v] - AFGPU_VAR_SPECIES(0);
    if (k >= 0 && k < S.n) {
      vp.v[v] = S.sp[k][0];
      continue;
    }
    if (ivs[v] < 0 || ivs[v] >= AFGPU_N_CC) {
      set_err(std::string(who) + ": variable index out of range");
      return AFGPU_ERR_ARG;
    }
    vp.v[v] = h->cc[ivs[v]];
  }
  return AFGPU_OK;
}

// Doubles of scratch the regression statistics need: the per-leaf-box
// terms and the per-run results.
static size_t diag_words(const afgpu_t *h, int n_vars) {
  return (size_t)std::max(h->l_all_leaves.n, 1) * n_vars * 3 + (size_t)h->n_runs * n_vars * 3;
}

// The regression statistics with the scratch `work` (diag_words doubles).
static int regression_values_in(afgpu_t *h, const VarPtrs &vp, int n_vars, double *work, double *out) {
  const int R = h->n_runs, nw = h->l_all_leaves.n;
  double *part = work, *dout = work + (size_t)std::max(nw, 1) * n_vars * 3;
  h->cur = h->stream;
  RCHK(h->impl->diag_boxes(vp, part));
  hipLaunchKernelGGL(k_diag_fold<256>, dim3(R), dim3(256), 0, h->stream, h->list(h->l_all_leaves),
                     (const DBox *)h->d_boxes, (const double *)part, nw, n_vars, (int)(R == 1), dout);
  HIPCHK(hipGetLastError());
  std::vector<double> res((size_t)R * n_vars * 3);
  HIPCHK(hipMemcpyAsync(res.data(), dout, res.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  RCHK(sync_stream(h));
  for (int r = 0; r < R; r++) {
    const double vol = run_volume(h, r);
    for (int v = 0; v < n_vars; v++) {
      const double *x = &res[((size_t)r * n_vars + v) * 3];
      double *o = out + ((size_t)r * n_vars + v) * 3;
      o[0] = x[0] / vol;
      o[1] = x[1] / vol;
      o[2] = x[2];
    }
  }
  return AFGPU_OK;
}

// The statistics of output_regression_log (src/m_output.f90:797-810) per
// run: the per-box terms on the device, summed per run in a fixed order.
int afgpu_regression_values(afgpu_t *h, int n_vars, const int *ivs, double *out) {
  RCHK(ready(h));
  VarPtrs vp;
  RCHK(check_vars(h, n_vars, ivs, vp, "afgpu_regression_values"));
  if (!out) return AFGPU_ERR_ARG;
  const size_t need = diag_words(h, n_vars);
  if (need > h->diag_cap) {
    (void)hipFree(h->d_diag);
    h->d_diag = nullptr;
    h->diag_cap = 0;
    HIPCHK(hipMalloc(&h->d_diag, need * sizeof(double)));
    h->diag_cap = need;
  }
  return regression_values_in(h, vp, n_vars, h->d_diag, out);
}

// The same statistics with the scratch in the library's static device
// array (noalloc_scratch), so that the call allocates no device memory.
// Every call on one GPU shares that array: calls that overlap on one GPU
// race.  Nothing in it needs clearing (every per-box slot the fold reads is
// written first by the same call).
int afgpu_regression_values_noalloc(afgpu_t *h, int n_vars, const int *ivs, double *out) {
  RCHK(ready(h));
  VarPtrs vp;
  RCHK(check_vars(h, n_vars, ivs, vp, "afgpu_regression_values_noalloc"));
  if (!out) return AFGPU_ERR_ARG;
  if (!h->scratch) {
    set_err("afgpu_regression_values_noalloc: the static scratch array has no address on this device");
    return AFGPU_ERR_STATE;
  }
  const size_t need = diag_words(h, n_vars);
  if (need > AFGPU_NOALLOC_SCRATCH_WORDS) {
    set_err("afgpu_regression_values_noalloc: (leaf boxes + runs) * n_vars * 3 = " + std::to_string(need) +
            " doubles exceed the static scratch of AFGPU_NOALLOC_SCRATCH_WORDS = " +
            std::to_string((size_t)AFGPU_NOALLOC_SCRATCH_WORDS));
    return AFGPU_ERR_ARG;
  }
  return regression_values_in(h, vp, n_vars, h->scratch, out);
}

// Fortran's Ew.d with a three-digit exponent (E20.8e3): 0.dddddddd, the
// decimal exponent as E+eee, right-aligned in w characters.
static int fortran_e(char *dst, size_t cap, double x, int w, int d) {
  char tmp[64];
  if (!std::isfinite(x)) return std::snprintf(dst, cap, "%*s", w, std::isnan(x) ? "NaN" : (x > 0 ? "Infinity" : "-Infinity"));
  if (x == 0) {
    std::snprintf(tmp, sizeof tmp, "%s0.%0*dE+000", std::signbit(x) ? "-" : "", d, 0);
  } else {
    // d significant digits, correctly rounded: d.ddd...e+X, then 0.dddd E+(X+1)
    char sci[64];
    std::snprintf(sci, sizeof sci, "%.*e", d - 1, std::fabs(x));
    const char *e = std::strchr(sci, 'e');
    const int ex = std::atoi(e + 1) + 1;
    std::string digits;
    for (const char *p = sci; p < e; p++)
      if (*p != '.') digits += *p;
    std::snprintf(tmp, sizeof tmp, "%s0.%sE%c%03d", x < 0 ? "-" : "", digits.c_str(), ex < 0 ? '-' : '+',
                  ex < 0 ? -ex : ex);
  }
  return std::snprintf(dst, cap, "%*s", w, tmp);
}

int afgpu_format_regression_line(int out_cnt, int n_vals, const double *vals, char *buf, int buf_len) {
  if (n_vals < 0 || (n_vals > 0 && !vals) || !buf || buf_len < 1) return AFGPU_ERR_ARG;
  std::string line = std::to_string(out_cnt);
  for (int n = 0; n < n_vals; n++) {
    char f[64];
    fortran_e(f, sizeof f, vals[n], 20, 8);
    line += f;
  }
  if ((int)line.size() + 1 > buf_len) {
    set_err("afgpu_format_regression_line: buffer too short");
    return AFGPU_ERR_ARG;
  }
  std::memcpy(buf, line.c_str(), line.size() + 1);
  return (int)line.size();
}

// output_regression_log (src/m_output.f90:788-842) on a plain tree
int afgpu_write_regression_log(afgpu_t *h, const char *filename, int out_cnt, int n_vars, const int *ivs,
                               const char *const *names) {
  RCHK(ready(h));
  if (!filename || !names || out_cnt < 0) return AFGPU_ERR_ARG;
  if (h->n_runs != 1) {
    set_err("afgpu_write_regression_log: a forest of runs (use afgpu_regression_values)");
    return AFGPU_ERR_STATE;
  }
  for (int v = 0; v < n_vars; v++)
    if (!names[v]) return AFGPU_ERR_ARG;
  std::vector<double> vals(3 + 3 * (size_t)std::max(n_vars, 0));
  RCHK(afgpu_regression_values(h, n_vars, ivs, vals.data() + 3));
  // (per variable: sum, sum of squares, maximum) -> the log's column order
  std::vector<double> line(3 + 3 * (size_t)n_vars);
  line[0] = 0;  // (out_cnt is written as an integer)
  line[1] = h->time;
  line[2] = h->started ? h->global_dt : h->p.dt_min;
  for (int v = 0; v < n_vars; v++) {
    line[3 + v] = vals[3 + 3 * v];
    line[3 + n_vars + v] = vals[3 + 3 * v + 1];
    line[3 + 2 * n_vars + v] = vals[3 + 3 * v + 2];
  }
  std::FILE *f = std::fopen(filename, out_cnt == 0 ? "w" : "a");
  if (!f) {
    set_err(std::string("afgpu_write_regression_log: cannot open ") + filename);
    return AFGPU_ERR_ARG;
  }
  if (out_cnt == 0) {
    std::fputs("it time dt", f);
    for (int v = 0; v < n_vars; v++) std::fprintf(f, " sum(%s)", names[v]);
    for (int v = 0; v < n_vars; v++) std::fprintf(f, " sum(%s^2)", names[v]);
    for (int v = 0; v < n_vars; v++) std::fprintf(f, " max(%s)", names[v]);
    std::fputs("\n", f);
  }
  std::vector<char> buf(64 + 24 * line.size());
  const int len = afgpu_format_regression_line(out_cnt, (int)line.size() - 1, line.data() + 1, buf.data(),
                                               (int)buf.size());
  if (len < 0) {
    std::fclose(f);
    return len;
  }
  std::fputs(buf.data(), f);
  std::fputs("\n", f);
  return std::fclose(f) == 0 ? AFGPU_OK : AFGPU_ERR_ARG;
}

// Box data in af_write_tree's order (afivo/src/m_af_output.f90:90-114):
// cell variables (species included) and face variables (AFGPU_VAR_FACE)
static int box_io(afgpu_t *h, int n_vars, const int *ivs, int n_ids, const int *ids, double *data, bool put) {
  RCHK(ready(h));
  const char *who = put ? "afgpu_put_boxes" : "afgpu_get_boxes";
  if (n_vars < 1 || n_vars > AFGPU_DIAG_MAX_VARS || !ivs) {
    set_err(std::string(who) + ": 1 to AFGPU_DIAG_MAX_VARS variables");
    return AFGPU_ERR_ARG;
  }
  BoxIoVars bv;
  bv.n = n_vars;
  bv.total = 0;
  for (int v = 0; v < n_vars; v++) {
    const int f = ivs[v] - AFGPU_VAR_FACE(0);
    if (f >= 0 && f < AFGPU_N_FC) {
      bv.v[v] = h->fc[f];
      bv.len[v] = (int)h->nface;
    } else {
      VarPtrs one;
      RCHK(check_vars(h, 1, ivs + v, one, who));
      bv.v[v] = one.v[0];
      bv.len[v] = (int)h->ncell;
    }
    bv.off[v] = bv.total;
    bv.total += bv.len[v];
  }
  if (n_ids < 0 || (n_ids > 0 && (!ids || !data))) return AFGPU_ERR_ARG;
  if (n_ids == 0) return AFGPU_OK;
  for (int q = 0; q < n_ids; q++)
    if (ids[q] < 1 || ids[q] > h->n_boxes) {
      set_err("box I/O: box id out of range");
      return AFGPU_ERR_ARG;
    }
  const size_t words = (size_t)n_ids * bv.total;
  if (words > h->io_cap) {
    (void)hipFree(h->d_io);
    h->d_io = nullptr;
    h->io_cap = 0;
    HIPCHK(hipMalloc(&h->d_io, words * sizeof(double)));
    h->io_cap = words;
  }
  RCHK(ensure_ref_buffers(h, std::max(n_ids, h->n_boxes)));
  HIPCHK(hipMemcpyAsync(h->d_ref_ids, ids, (size_t)n_ids * sizeof(int), hipMemcpyHostToDevice, h->stream));
  const int nw = n_ids * n_vars;
  if (put) HIPCHK(hipMemcpyAsync(h->d_io, data, words * sizeof(double), hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_box_io, dim3(std::min(nw, 8 * h->n_cu)), dim3(256), 0, h->stream, (const int *)h->d_ref_ids,
                     bv, h->d_io, (int)put, nw);
  HIPCHK(hipGetLastError());
  if (!put) HIPCHK(hipMemcpyAsync(data, h->d_io, words * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  return sync_stream(h);
}

int afgpu_get_boxes(afgpu_t *h, int n_vars, const int *ivs, int n_ids, const int *ids, double *data) {
  return box_io(h, n_vars, ivs, n_ids, ids, data, false);
}

int afgpu_put_boxes(afgpu_t *h, int n_vars, const int *ivs, int n_ids, const int *ids, const double *data) {
  return box_io(h, n_vars, ivs, n_ids, ids, const_cast<double *>(data), true);
}

int afgpu_copy_current_state(afgpu_t *h) {
  RCHK(ready(h));
  RCHK(h->impl->copy_states(true));
  return sync_stream(h);
}

int afgpu_restore_previous_state(afgpu_t *h) {
  RCHK(ready(h));
  RCHK(h->impl->copy_states(false));
  return sync_stream(h);
}

int afgpu_step(afgpu_t *h, afgpu_step_info_t *info) {
  RCHK(ready(h, true));
  RCHK(check_photoi_species(h));
  if (h->n_runs > 1) {
    set_err("afgpu_step: the tree is a forest of runs (use afgpu_step_ensemble)");
    return AFGPU_ERR_STATE;
  }
  if (!h->started) {
    h->global_dt = h->p.dt_min;
    h->dt = h->global_dt;
    h->started = true;
  }
  h->it = h->it + 1;
  int photoi_done = 0;
  if (h->p.photoi_enabled && h->p.photoi_per_steps > 0 && h->it % h->p.photoi_per_steps == 0) {
    RCHK(h->impl->photoi());
    photoi_done = 1;
  }
  double dt_lim = std::numeric_limits<double>::max();
  double dt = h->dt;
  const int max_attempts = 10;
  int n, n_rej = 0;
  for (n = 1; n <= max_attempts; n++) {
    RCHK(h->impl->copy_states(true));
    double dt_lim_step;
    RCHK(advance(h, dt, &dt_lim_step));
    dt_lim = std::min(dt_lim, dt_lim_step);
    if (dt <= dt_lim_step) break;
    n_rej++;
    dt = h->p.dt_safety_factor * dt_lim_step;
    h->time = h->global_time;
    RCHK(h->impl->copy_states(false));
  }
  h->frac_rejected = 0.99 * h->frac_rejected;
  if (n > 1) h->frac_rejected = h->frac_rejected + 0.01;
  if (n == max_attempts + 1) {
    set_err("All time steps were rejected");
    return AFGPU_ERR_ALL_REJECTED;
  }
  const double dt_used = dt;
  int n_cyc = 0;
  double res = 0;
  RCHK(h->impl->field_compute(0, true, &n_cyc, &res));
  double tmp = h->p.dt_max_growth_factor;
  if (h->frac_rejected > 0.1) tmp = 1.0;
  dt = std::min(tmp * h->global_dt, h->p.dt_safety_factor * std::min(dt_lim, h->p.dt_max));
  h->global_dt = dt;
  h->global_time = h->time;
  h->dt = dt;
  if (info) {
    info->it = h->it;
    info->time = h->time;
    info->dt_used = dt_used;
    info->dt_next = dt;
    info->dt_lim = dt_lim;
    info->n_rejected = n_rej;
    info->n_vcycles = n_cyc;
    info->residual = res;
    double mr;
    RCHK(h->impl->maxabs(AFGPU_CC_RHS, &mr));
    info->rel_residual = mr > 0 ? res / mr : 0;
    info->photoi_done = photoi_done;
    info->leaf_cells = afgpu_num_leaf_cells(h);
    info->photoi_cycles = photoi_done ? h->photoi_cycles : 0;
  }
  RCHK(sync_stream(h));
  if (h->global_dt < h->p.dt_min) {
    set_err("dt too small");
    return AFGPU_ERR_DT_TOO_SMALL;
  }
  return AFGPU_OK;
}

int afgpu_get_time_state(afgpu_t *h, int *it, double *time, double *global_dt) {
  if (!h) return AFGPU_ERR_ARG;
  if (it) *it = h->it;
  if (time) *time = h->time;
  if (global_dt) *global_dt = h->started ? h->global_dt : h->p.dt_min;
  return AFGPU_OK;
}

int afgpu_set_time_state(afgpu_t *h, int it, double time, double global_dt) {
  if (!h || it < 0 || !(global_dt > 0)) return AFGPU_ERR_ARG;
  h->it = it;
  h->time = h->global_time = time;
  h->global_dt = h->dt = global_dt;
  h->started = true;
  return AFGPU_OK;
}

int afgpu_limit_step_dt(afgpu_t *h, double dt_cap) {
  if (!h || !(dt_cap >= 0)) return AFGPU_ERR_ARG;
  if (!h->have_params) {
    set_err("afgpu_limit_step_dt: parameters not set");
    return AFGPU_ERR_STATE;
  }
  if (!h->started) {  // (as afgpu_step starts: dt = global_dt = dt_min)
    h->global_dt = h->p.dt_min;
    h->dt = h->global_dt;
    h->started = true;
  }
  h->dt = std::min(h->dt, dt_cap);
  return AFGPU_OK;
}

long long afgpu_num_leaf_cells(afgpu_t *h) {
  if (!h) return 0;
  long long n = h->l_all_leaves.n;
  for (int d = 0; d < h->D; d++) n *= h->nc;
  return n;
}

int afgpu_kernel_timing(afgpu_t *h, int enable) {
  if (!h) return AFGPU_ERR_ARG;
  RCHK(sync_stream(h));
  if ((enable != 0) != h->timing) drop_graphs(*h);  // recapture with/without event nodes
  h->timing = enable != 0;
  return AFGPU_OK;
}

int afgpu_kernel_stats(afgpu_t *h, int kind, int reset, double *total_ms, long long *launches, double *algo_bytes) {
  if (!h || kind < 0 || kind > 2) return AFGPU_ERR_ARG;
  RCHK(sync_stream(h));
  KTimer &t = h->timers[kind];
  if (total_ms) *total_ms = t.total_ms;
  if (launches) *launches = t.launches;
  if (algo_bytes) *algo_bytes = t.algo_bytes;
  if (reset) t = KTimer();
  return AFGPU_OK;
}

// ---- work-decomposition shapes and their autotuner ---------------------------
static int upload_lists(afgpu_t *h, int bb) {
  RCHK(sync_stream(h));
  build_lists(*h, bb);
  if (!h->h_lists.empty())
    HIPCHK(hipMemcpy(h->d_lists, h->h_lists.data(), h->h_lists.size() * sizeof(int), hipMemcpyHostToDevice));
  h->masked = false;
  h->base_active.clear();
  drop_graphs(*h);
  return AFGPU_OK;
}

static bool shape_ok(const afgpu_t *h, int bb, int per_cu) {
  const bool wave_boxes = h->D == 2 && h->nc == 8;
  return (bb == 1 || ((bb == 2 || bb == 4 || bb == 8) && wave_boxes)) && per_cu >= 0 && per_cu <= 64;
}

int afgpu_get_shapes(afgpu_t *h, int n_classes, afgpu_shape_t *shapes) {
  if (!h || n_classes < 0 || n_classes > KC_N || (n_classes > 0 && !shapes)) return AFGPU_ERR_ARG;
  for (int k = 0; k < n_classes; k++) {
    shapes[k].boxes_per_workgroup = h->shape[k].bb;
    shapes[k].workgroups_per_cu = h->shape[k].per_cu;
    shapes[k].ms = 0;
  }
  return AFGPU_OK;
}

int afgpu_set_shapes(afgpu_t *h, int n_classes, const afgpu_shape_t *shapes) {
  if (!h || n_classes < 0 || n_classes > KC_N || (n_classes > 0 && !shapes)) return AFGPU_ERR_ARG;
  for (int k = 0; k < n_classes; k++)
    if (!shape_ok(h, shapes[k].boxes_per_workgroup, shapes[k].workgroups_per_cu)) {
      set_err("afgpu_set_shapes: unsupported shape");
      return AFGPU_ERR_ARG;
    }
  RCHK(sync_stream(h));
  for (int k = 0; k < n_classes; k++) h->shape[k] = Shape{shapes[k].boxes_per_workgroup, shapes[k].workgroups_per_cu};
  if (h->highest_lvl >= 1 && h->shape[KC_GSRB].bb != h->lists_bb) RCHK(upload_lists(h, h->shape[KC_GSRB].bb));
  drop_graphs(*h);
  return AFGPU_OK;
}

// Everything a time step changes, for trial steps.
struct Snapshot {
  std::vector<std::pair<void *, void *>> bufs;  // (pool, copy)
  std::vector<size_t> bytes;
  int it;
  double time, global_time, global_dt, dt, frac_rejected;
  bool started;
  std::vector<afgpu_ctx::RunTime> rt;
};
static int snap_add(Snapshot &sn, void *pool, size_t bytes) {
  if (!pool || bytes == 0) return AFGPU_OK;
  void *cp = nullptr;
  HIPCHK(hipMalloc(&cp, bytes));
  sn.bufs.push_back({pool, cp});
  sn.bytes.push_back(bytes);
  return AFGPU_OK;
}
static void snap_free(Snapshot &sn) {
  for (auto &b : sn.bufs) (void)hipFree(b.second);
  sn.bufs.clear();
}
static int snap_copy(afgpu_t *h, Snapshot &sn, bool save) {
  for (size_t q = 0; q < sn.bufs.size(); q++) {
    void *dst = save ? sn.bufs[q].second : sn.bufs[q].first;
    const void *src = save ? sn.bufs[q].first : sn.bufs[q].second;
    HIPCHK(hipMemcpyAsync(dst, src, sn.bytes[q], hipMemcpyDeviceToDevice, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  if (save) {
    sn.it = h->it;
    sn.time = h->time;
    sn.global_time = h->global_time;
    sn.global_dt = h->global_dt;
    sn.dt = h->dt;
    sn.frac_rejected = h->frac_rejected;
    sn.started = h->started;
    sn.rt = h->rt;
  } else {
    h->it = sn.it;
    h->time = sn.time;
    h->global_time = sn.global_time;
    h->global_dt = sn.global_dt;
    h->dt = sn.dt;
    h->frac_rejected = sn.frac_rejected;
    h->started = sn.started;
    h->rt = sn.rt;
  }
  return AFGPU_OK;
}

static int trial_step(afgpu_t *h) {
  if (h->n_runs > 1) {
    RCHK(h->impl->step_runs(nullptr));
    return sync_stream(h);
  }
  return afgpu_step(h, nullptr);
}

int afgpu_autotune(afgpu_t *h, int n_classes, afgpu_shape_t *shapes) {
  RCHK(ready(h, true));
  if (n_classes < 0 || n_classes > KC_N || (n_classes > 0 && !shapes)) return AFGPU_ERR_ARG;
  RCHK(sync_stream(h));
  Snapshot sn;
  int rc = AFGPU_OK;
  const size_t cells = (size_t)h->cap_boxes * h->ncell, faces = (size_t)h->cap_boxes * h->nface;
  for (int i = 0; i < AFGPU_N_CC && rc == AFGPU_OK; i++) rc = snap_add(sn, h->cc[i], cells * sizeof(double));
  for (int i = 0; i < AFGPU_N_FC && rc == AFGPU_OK; i++) rc = snap_add(sn, h->fc[i], faces * sizeof(double));
  // the species of a reaction list beyond the three standard pools
  for (size_t i = 0; i < h->xsp.size() && rc == AFGPU_OK; i++) rc = snap_add(sn, h->xsp[i], cells * sizeof(double));
  for (int i = 0; i < 4 && rc == AFGPU_OK; i++) rc = snap_add(sn, h->cc32[i], cells * sizeof(float));
  for (auto &r : h->mres)
    for (int i = 0; i < 4 && rc == AFGPU_OK; i++)
      if (r.cap_boxes == h->cap_boxes) rc = snap_add(sn, r.cc32[i], cells * sizeof(float));
  if (rc == AFGPU_OK) rc = snap_copy(h, sn, true);
  if (rc != AFGPU_OK) {
    snap_free(sn);
    return rc;
  }
  const bool wave_boxes = h->D == 2 && h->nc == 8;
  std::vector<Shape> cands;
  if (wave_boxes) {
    for (int bb : {1, 2, 4, 8})
      for (int pc : {0, 4, 8, 16, 32}) cands.push_back(Shape{bb, pc});
  } else {
    for (int pc : {0, 2, 4, 8}) cands.push_back(Shape{1, pc});
  }
  const Shape saved[KC_N] = {h->shape[0], h->shape[1], h->shape[2], h->shape[3],
                             h->shape[4], h->shape[5], h->shape[6], h->shape[7]};
  const bool timing0 = h->timing;
  double best[KC_N];
  Shape pick[KC_N];
  for (int k = 0; k < KC_N; k++) {
    best[k] = std::numeric_limits<double>::max();
    pick[k] = saved[k];
  }
  h->timing = true;
  h->class_timing = true;
  drop_graphs(*h);
  // one unmeasured step first (set-up work of the first solve), then one
  // step per candidate, each from the snapshot
  for (int q = -1; q < (int)cands.size() && rc == AFGPU_OK; q++) {
    const Shape sh = q < 0 ? saved[KC_GSRB] : cands[q];
    for (int k = 0; k < KC_N; k++) h->shape[k] = q < 0 ? saved[k] : sh;
    rc = upload_lists(h, h->shape[KC_GSRB].bb);
    for (int k = 0; k < KC_N; k++) h->timers[3 + k] = KTimer();
    if (rc == AFGPU_OK) rc = trial_step(h);
    if (rc == AFGPU_OK) rc = sync_stream(h);
    if (rc == AFGPU_OK && q >= 0)
      for (int k = 0; k < KC_N; k++) {
        const KTimer &t = h->timers[3 + k];
        if (t.launches > 0 && t.total_ms < best[k]) {
          best[k] = t.total_ms;
          pick[k] = sh;
        }
      }
    const int rc2 = snap_copy(h, sn, false);
    if (rc == AFGPU_OK) rc = rc2;
  }
  h->timing = timing0;
  h->class_timing = false;
  for (int k = 0; k < KC_N; k++) {
    h->timers[3 + k] = KTimer();
    h->shape[k] = (rc == AFGPU_OK) ? pick[k] : saved[k];
  }
  const int rc3 = upload_lists(h, h->shape[KC_GSRB].bb);
  snap_free(sn);
  RCHK(rc);
  RCHK(rc3);
  for (int k = 0; k < n_classes; k++) {
    shapes[k].boxes_per_workgroup = h->shape[k].bb;
    shapes[k].workgroups_per_cu = h->shape[k].per_cu;
    shapes[k].ms = best[k] < std::numeric_limits<double>::max() ? best[k] : 0.0;
  }
  return AFGPU_OK;
}

int afgpu_set_option(afgpu_t *h, int option, int value) {
  if (!h) return AFGPU_ERR_ARG;
  switch (option) {
    case AFGPU_OPT_GRAPHS:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->use_graphs = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_DIRECT_N:
      if (value > 128) {
        set_err("afgpu_set_option: direct solve limited to 128 cells per axis");
        return AFGPU_ERR_ARG;
      }
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->direct_max_n = value;
      h->dir.set = false;
      for (auto &f : h->dir.f) f.ok = false;
      return AFGPU_OK;
    case AFGPU_OPT_FUSE_DIRECT:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->fuse_direct = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_DIRECT_F32:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->direct_f32 = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_DIRECT_FOLD:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->direct_fold = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_STEP_GRAPH:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->step_graph = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_FLUX32:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->flux32 = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_BOX_LDS:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->box_lds = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_PHOTOI_VCYCLE:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->photoi_inner_vcycle = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_DIRECT_MFMA:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->direct_mfma = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_BATCH_MODES:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->batch_modes = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_BLOCK_SMOOTH:
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->block_smooth = value != 0;
      return AFGPU_OK;
    case AFGPU_OPT_MIXED_SMOOTH:
      if (value < 0 || value > 8) {
        set_err("afgpu_set_option: inner smoothing sweeps must be 0..8");
        return AFGPU_ERR_ARG;
      }
      RCHK(sync_stream(h));
      drop_graphs(*h);
      h->mixed_smooth = value;
      return AFGPU_OK;
    default:
      set_err("afgpu_set_option: unknown option");
      return AFGPU_ERR_ARG;
  }
}

}  // extern "C"

// ---- general chemistry networks (src/m_chemistry.f90:672-804) --------------

int afgpu_chem_create(int n_species, int n_reactions, const afgpu_reaction_t *reactions, int table_n,
                      double table_x_min, double table_x_max, int n_fld_cols, const double *fld_cols,
                      const double *energy_col, double gas_temperature, afgpu_chem_t **out) {
  if (!out || n_species < 1 || n_species > AFGPU_CHEM_MAX_SPECIES || n_reactions < 0 ||
      n_reactions > AFGPU_CHEM_MAX_REACTIONS || (n_reactions > 0 && !reactions) ||
      table_n < 2 || !(table_x_max > table_x_min) || n_fld_cols < 0 || (n_fld_cols > 0 && !fld_cols) || !(gas_temperature > 0)) {
    set_err("afgpu_chem_create: invalid arguments");
    return AFGPU_ERR_ARG;
  }
  bool need_te = false;
  for (int r = 0; r < n_reactions; r++) {
    const afgpu_reaction_t &R = reactions[r];
    const int t = R.rate_type;
    const bool known = (t >= AFGPU_RATE_TABULATED_FIELD && t <= AFGPU_RATE_K1) || (t >= AFGPU_RATE_K3 && t <= AFGPU_RATE_K16);
    if (!known) {
      set_err(t == AFGPU_RATE_TABULATED_ENERGY ? "afgpu_chem_create: energy tables need the energy equation"
                                               : "afgpu_chem_create: unknown rate type");
      return AFGPU_ERR_UNSUPPORTED;
    }
    if (R.n_in < 1 || R.n_in > AFGPU_CHEM_MAX_IN || R.n_out < 0 || R.n_out > AFGPU_CHEM_MAX_OUT ||
        (t == AFGPU_RATE_TABULATED_FIELD && (R.lookup_col < 0 || R.lookup_col >= n_fld_cols))) {
      set_err("afgpu_chem_create: invalid reaction");
      return AFGPU_ERR_ARG;
    }
    for (int q = 0; q < R.n_in; q++)
      if (R.ix_in[q] < 0 || R.ix_in[q] >= n_species) return set_err("afgpu_chem_create: species index"), AFGPU_ERR_ARG;
    for (int q = 0; q < R.n_out; q++)
      if (R.ix_out[q] < 0 || R.ix_out[q] >= n_species) return set_err("afgpu_chem_create: species index"), AFGPU_ERR_ARG;
    need_te = need_te || t == AFGPU_RATE_K1 || t == AFGPU_RATE_K3 || t == AFGPU_RATE_K16;
  }
  if (need_te && !energy_col) {
    set_err("afgpu_chem_create: the rate types with Te need the mean-energy column");
    return AFGPU_ERR_ARG;
  }
  std::unique_ptr<afgpu_chem> c(new afgpu_chem);
  c->n_species = n_species;
  c->n_reactions = n_reactions;
  if (n_reactions > 0) {
    HIPCHK(hipMalloc(&c->d_r, (size_t)n_reactions * sizeof(afgpu_reaction_t)));
    HIPCHK(hipMemcpy(c->d_r, reactions, (size_t)n_reactions * sizeof(afgpu_reaction_t), hipMemcpyHostToDevice));
  }
  if (n_fld_cols > 0) {
    HIPCHK(hipMalloc(&c->d_fld, (size_t)n_fld_cols * table_n * sizeof(double)));
    HIPCHK(hipMemcpy(c->d_fld, fld_cols, (size_t)n_fld_cols * table_n * sizeof(double), hipMemcpyHostToDevice));
  }
  if (energy_col) {
    HIPCHK(hipMalloc(&c->d_energy, (size_t)table_n * sizeof(double)));
    HIPCHK(hipMemcpy(c->d_energy, energy_col, (size_t)table_n * sizeof(double), hipMemcpyHostToDevice));
  }
  // LT_create (src/lookup_table_fortran/m_lookup_table.f90): linear spacing
  const double inv_fac = (table_n - 1) / (table_x_max - table_x_min);
  c->args.r = c->d_r;
  c->args.n_reactions = n_reactions;
  c->args.n_species = n_species;
  c->args.fld = DTable{c->d_fld, table_n, table_x_min, inv_fac};
  c->args.energy = DTable{c->d_energy, table_n, table_x_min, inv_fac};
  c->args.tg = gas_temperature;
  c->args.need_te = need_te ? 1 : 0;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  *out = c.release();
  return AFGPU_OK;
}

int afgpu_chem_destroy(afgpu_chem_t *c) {
  delete c;  // (the destructor releases the device arrays and the stream)
  return AFGPU_OK;
}

static int chem_launch(afgpu_chem_t *c, int n_cells, const double *fields, const double *dens, double *rates,
                       double *derivs, int mode) {
  if (!c || n_cells < 0) return AFGPU_ERR_ARG;
  if (n_cells == 0) return AFGPU_OK;
  if ((mode != 1 && !fields) || (mode != 0 && (!dens || !derivs)) || (mode != 2 && c->n_reactions > 0 && !rates)) {
    set_err("afgpu_chem: missing array");
    return AFGPU_ERR_ARG;
  }
  // LDS per workgroup within 64 KiB: the derivatives of 256 cells with
  // their densities staged for small networks, else of 64 cells (densities
  // staged while they fit), and the reaction records while they fit
  const size_t ns = (size_t)c->n_species, cap = 64 * 1024, d = sizeof(double);
  const size_t rb = (size_t)c->n_reactions * sizeof(afgpu_reaction_t);
  const dim3 g256((n_cells + 255) / 256), g64((n_cells + 63) / 64);
  if (rb + 2 * ns * 256 * d <= cap)
    hipLaunchKernelGGL((k_chem<256, true, true>), g256, dim3(256), rb + 2 * ns * 256 * d, c->stream, c->args,
                       n_cells, fields, dens, rates, derivs, mode);
  else if (rb + 2 * ns * 64 * d <= cap)
    hipLaunchKernelGGL((k_chem<64, true, true>), g64, dim3(64), rb + 2 * ns * 64 * d, c->stream, c->args, n_cells,
                       fields, dens, rates, derivs, mode);
  else if (rb + ns * 64 * d <= cap)
    hipLaunchKernelGGL((k_chem<64, false, true>), g64, dim3(64), rb + ns * 64 * d, c->stream, c->args, n_cells,
                       fields, dens, rates, derivs, mode);
  else
    hipLaunchKernelGGL((k_chem<64, false, false>), g64, dim3(64), ns * 64 * d, c->stream, c->args, n_cells, fields,
                       dens, rates, derivs, mode);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return AFGPU_OK;
}

int afgpu_chem_rates(afgpu_chem_t *c, int n_cells, const double *fields, double *rates) {
  return chem_launch(c, n_cells, fields, nullptr, rates, nullptr, 0);
}
int afgpu_chem_derivatives(afgpu_chem_t *c, int n_cells, const double *dens, double *rates, double *derivs) {
  return chem_launch(c, n_cells, nullptr, dens, rates, derivs, 1);
}
int afgpu_chem_source(afgpu_chem_t *c, int n_cells, const double *fields, const double *dens, double *derivs) {
  return chem_launch(c, n_cells, fields, dens, nullptr, derivs, 2);
}

int afgpu_device_alloc(size_t bytes, void **ptr) {
  if (!ptr) return AFGPU_ERR_ARG;
  *ptr = nullptr;
  HIPCHK(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
  return AFGPU_OK;
}
int afgpu_device_free(void *ptr) {
  if (ptr) HIPCHK(hipFree(ptr));
  return AFGPU_OK;
}
int afgpu_device_copy(void *dst, const void *src, size_t bytes, int to_device) {
  if ((!dst || !src) && bytes > 0) return AFGPU_ERR_ARG;
  if (bytes == 0) return AFGPU_OK;
  HIPCHK(hipMemcpy(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost));
  return AFGPU_OK;
}

// The time step's source terms from a reaction list (add_source_terms,
// src/m_fluid.f90:298-466).  Species 0 are the electrons; species 1 and 2
// live in the pools of the positive and negative ions, further species in
// pools of their own; charges[k] is species k's (NULL: -1, +1, -1 for a
// list over three species).  NULL returns to the standard model.  The
// chemistry handle must outlive its use by the step.
int afgpu_set_chemistry(afgpu_t *h, const afgpu_chem_t *chem, const int *charges) {
  if (!h) return AFGPU_ERR_ARG;
  const int n = chem ? chem->n_species : 3;
  if (chem && (n < 3 || n > SP_MAX || (n > 3 && !charges) || (charges && charges[0] != -1))) {
    set_err("afgpu_set_chemistry: 3 to 16 species, the electrons first, charges for more than three");
    return AFGPU_ERR_ARG;
  }
  std::vector<int> q{-1, 1, -1};
  if (chem && charges) q.assign(charges, charges + n);
  const bool gen = chem && (n > 3 || q != std::vector<int>{-1, 1, -1});
  if (gen && h->p.photoi_enabled && (h->p.photoi_species < 0 || h->p.photoi_species >= n)) {
    set_err("afgpu_set_chemistry: photoionization species out of range");
    return AFGPU_ERR_ARG;
  }
  if (gen && update_n_lds(box_threads(h->D, h->nc), n, chem->n_reactions) > UPDATE_N_LDS_MAX) {
    set_err("afgpu_set_chemistry: " + std::to_string(n) + " species and " + std::to_string(chem->n_reactions) +
            " reactions need more LDS than the update of boxes of this size has (64 KiB: 2 x species x " +
            std::to_string(box_threads(h->D, h->nc)) + " doubles plus the reaction records)");
    return AFGPU_ERR_UNSUPPORTED;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  drop_graphs(*h);  // (the list and the pools are arguments of captured updates)
  // pools of the species beyond the three, zero-filled
  const size_t want = gen ? (size_t)(n - 3) * 3 : 0;
  if (h->xsp.size() != want) {
    for (double *x : h->xsp) (void)hipFree(x);
    h->xsp.assign(want, nullptr);
    for (double *&x : h->xsp) {
      HIPCHK(hipMalloc(&x, std::max<size_t>(1, (size_t)h->cap_boxes * h->ncell) * sizeof(double)));
      HIPCHK(hipMemset(x, 0, std::max<size_t>(1, (size_t)h->cap_boxes * h->ncell) * sizeof(double)));
    }
  }
  h->net = chem;
  h->n_sp = gen ? n : 3;
  h->gen_sp = gen;
  h->sp_charge = gen ? q : std::vector<int>{-1, 1, -1};
  return AFGPU_OK;
}

// A species' cells (box by box, as afgpu_get_cc): species 0-2 are the pools
// of the electrons and the two ions, further species those of a reaction
// list (afgpu_set_chemistry).
static int species_xfer(afgpu_t *h, int k, int state, int n_ids, const int *ids, double *data, bool put) {
  if (!h || k < 0 || k >= h->n_sp || state < 0 || state > 2) {
    set_err("species transfer: invalid species or state");
    return AFGPU_ERR_ARG;
  }
  return xfer_pool(h, h->species_args().sp[k][state], h->ncell, n_ids, ids, data, put);
}
int afgpu_get_species(afgpu_t *h, int k, int state, int n_ids, const int *ids, double *data) {
  return species_xfer(h, k, state, n_ids, ids, data, false);
}
int afgpu_put_species(afgpu_t *h, int k, int state, int n_ids, const int *ids, const double *data) {
  return species_xfer(h, k, state, n_ids, ids, const_cast<double *>(data), true);
}
