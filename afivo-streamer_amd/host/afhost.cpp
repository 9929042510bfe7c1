// Host-side stand-in for the parts of afivo-streamer that stay on the CPU:
// the quadtree/octree of boxes (af_init, refinement with 2:1 balance), the
// initial conditions and the transport/chemistry lookup tables.  It produces
// exactly the data the C-ABI of the device path consumes (afgpu_sync_tree,
// afgpu_put_cc, afgpu_set_tables), so that tests and benchmarks can drive
// the path without the Fortran driver.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

namespace {

struct HBox {
  int lvl = 0;
  int ix[3] = {1, 1, 1};
  int parent = 0;
  int children[8] = {0};
  int nb[6] = {0};
  int nbmat[27] = {0};
  double r_min[3] = {0, 0, 0};
  double dr[3] = {0, 0, 0};
};

const int child_dix[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                             {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
const int child_adj_nb2[4][2] = {{1, 3}, {2, 4}, {1, 2}, {3, 4}};
const int child_adj_nb3[6][4] = {{1, 3, 5, 7}, {2, 4, 6, 8}, {1, 2, 5, 6},
                                 {3, 4, 7, 8}, {1, 2, 3, 4}, {5, 6, 7, 8}};

}  // namespace

struct afhost_tree {
  int D = 2, nc = 8, coord = 1;
  double r_base[3] = {0, 0, 0};
  double dr_base[3] = {0, 0, 0};
  std::vector<HBox> boxes;  // index id-1
  std::map<std::vector<int>, int> lookup;

  std::vector<int> key(int lvl, const int *ix) const {
    std::vector<int> k(1 + D);
    k[0] = lvl;
    for (int d = 0; d < D; d++) k[1 + d] = ix[d];
    return k;
  }
  int find(int lvl, const int *ix) const {
    auto it = lookup.find(key(lvl, ix));
    return it == lookup.end() ? 0 : it->second;
  }
  bool inside(int lvl, const int *ix) const {
    int n = 1 << (lvl - 1);
    for (int d = 0; d < D; d++)
      if (ix[d] < 1 || ix[d] > n) return false;
    return true;
  }
  // Neighbour information (m_af_types.f90:296-299): a same-level box, the
  // physical boundary (-1) or no box (0).
  void set_neighbors() {
    for (size_t b = 0; b < boxes.size(); b++) {
      HBox &x = boxes[b];
      if (x.lvl <= 0) continue;
      for (int nb = 0; nb < 2 * D; nb++) {
        int q[3] = {x.ix[0], x.ix[1], x.ix[2]};
        q[nb / 2] += (nb % 2) ? 1 : -1;
        x.nb[nb] = inside(x.lvl, q) ? find(x.lvl, q) : -1;
      }
      int kmax = (D == 3) ? 1 : 0;
      for (int dk = -kmax; dk <= kmax; dk++)
        for (int dj = -1; dj <= 1; dj++)
          for (int di = -1; di <= 1; di++) {
            int q[3] = {x.ix[0] + di, x.ix[1] + dj, x.ix[2] + dk};
            int v;
            if (di == 0 && dj == 0 && dk == 0)
              v = (int)b + 1;
            else
              v = inside(x.lvl, q) ? find(x.lvl, q) : -1;
            x.nbmat[(di + 1) + 3 * (dj + 1) + (D == 3 ? 9 * (dk + 1) : 0)] = v;
          }
    }
  }
  // Refinement of one leaf (m_af_core.f90:1195-1214): children get
  // dr = dr/2 and r_min = r_min + (dr/2 * dix) * nc.
  void refine_box(int id) {
    HBox p = boxes[id - 1];
    int nch = 1 << D;
    int first = (int)boxes.size() + 1;
    for (int i = 0; i < nch; i++) {
      HBox c;
      c.lvl = p.lvl + 1;
      for (int d = 0; d < D; d++) {
        c.ix[d] = 2 * (p.ix[d] - 1) + 1 + child_dix[i][d];
        c.dr[d] = 0.5 * p.dr[d];
        c.r_min[d] = p.r_min[d] + 0.5 * p.dr[d] * child_dix[i][d] * nc;
      }
      c.parent = id;
      boxes.push_back(c);
      lookup[key(c.lvl, c.ix)] = first + i;
      boxes[id - 1].children[i] = first + i;
    }
  }
  bool is_leaf(int id) const { return boxes[id - 1].lvl > 0 && boxes[id - 1].children[0] == 0; }
  int highest() const {
    int h = 0;
    for (const HBox &b : boxes) h = b.lvl > h ? b.lvl : h;
    return h;
  }
};

typedef struct afhost_tree afhost_tree_t;

extern "C" {

// af_init (m_af_core.f90:138-330) with coarse_grid_size = n_cell: level 1 is
// one box.
int afhost_tree_create(int ndim, int nc, int coord, const double *r_min, const double *domain_len,
                       afhost_tree_t **out) {
  if ((ndim != 2 && ndim != 3) || nc < 4 || (nc % 2) || !r_min || !domain_len || !out) return 1;
  afhost_tree_t *t = new afhost_tree_t();
  t->D = ndim;
  t->nc = nc;
  t->coord = coord;
  HBox b;
  b.lvl = 1;
  for (int d = 0; d < ndim; d++) {
    t->r_base[d] = r_min[d];
    t->dr_base[d] = (domain_len[d]) / nc;
    b.dr[d] = t->dr_base[d];
    b.r_min[d] = t->r_base[d] + (b.ix[d] - 1) * t->dr_base[d] * nc;
  }
  t->boxes.push_back(b);
  t->lookup[t->key(1, b.ix)] = 1;
  t->set_neighbors();
  *out = t;
  return 0;
}

int afhost_tree_destroy(afhost_tree_t *t) {
  delete t;
  return 0;
}

// Refine the given leaves, closing the set under afivo's face-neighbour 2:1
// balance rule (ensure_two_one_balance, m_af_core.f90:1016-1057).  Returns
// the number of boxes added, or -1 on error.
int afhost_refine(afhost_tree_t *t, int n, const int *ids) {
  if (!t || n < 0 || (n > 0 && !ids)) return -1;
  std::vector<char> flag(t->boxes.size() + 1, 0);
  for (int i = 0; i < n; i++) {
    if (ids[i] < 1 || ids[i] > (int)t->boxes.size() || !t->is_leaf(ids[i])) return -1;
    flag[ids[i]] = 1;
  }
  bool changed = true;
  while (changed) {
    changed = false;
    for (int lvl = t->highest(); lvl >= 1; lvl--)
      for (size_t b = 0; b < t->boxes.size(); b++) {
        const HBox &x = t->boxes[b];
        if (x.lvl != lvl || !flag[b + 1]) continue;
        for (int nb = 0; nb < 2 * t->D; nb++) {
          if (x.nb[nb] != 0) continue;
          int p_nb = t->boxes[x.parent - 1].nb[nb];
          if (p_nb > 0 && !flag[p_nb] && t->is_leaf(p_nb)) {
            flag[p_nb] = 1;
            changed = true;
          }
        }
      }
  }
  int added = 0;
  size_t nb0 = t->boxes.size();
  for (int lvl = 1; lvl <= t->highest(); lvl++)
    for (size_t b = 0; b < nb0; b++)
      if (flag[b + 1] && t->boxes[b].lvl == lvl && t->is_leaf((int)b + 1)) {
        t->refine_box((int)b + 1);
        added += 1 << t->D;
      }
  t->set_neighbors();
  return added;
}

// Removal of the children of the given boxes, whose children must all be
// leaves (remove_children, m_af_core.f90:1159-1192, for the boxes that
// af_adjust_refinement derefines): the children's slots become unused
// (level 0) and are not reused; the boxes become leaves.  Returns the
// number of boxes removed, or -1 on error.
int afhost_derefine(afhost_tree_t *t, int n, const int *ids) {
  if (!t || n < 0 || (n > 0 && !ids)) return -1;
  const int nch = 1 << t->D;
  for (int i = 0; i < n; i++) {
    if (ids[i] < 1 || ids[i] > (int)t->boxes.size() || t->boxes[ids[i] - 1].lvl < 1) return -1;
    const HBox &p = t->boxes[ids[i] - 1];
    if (p.children[0] == 0) return -1;
    for (int q = 0; q < nch; q++)
      if (!t->is_leaf(p.children[q])) return -1;
  }
  int removed = 0;
  for (int i = 0; i < n; i++) {
    HBox &p = t->boxes[ids[i] - 1];
    for (int q = 0; q < nch; q++) {
      const int c = p.children[q];
      t->lookup.erase(t->key(t->boxes[c - 1].lvl, t->boxes[c - 1].ix));
      t->boxes[c - 1] = HBox();
      t->boxes[c - 1].lvl = 0;
      p.children[q] = 0;
      removed++;
    }
  }
  t->set_neighbors();
  return removed;
}

// af_refine_up_to_lvl: refine every leaf until all leaves are at `lvl`.
int afhost_refine_uniform(afhost_tree_t *t, int lvl) {
  if (!t) return -1;
  int total = 0;
  while (true) {
    std::vector<int> ids;
    for (size_t b = 0; b < t->boxes.size(); b++)
      if (t->is_leaf((int)b + 1) && t->boxes[b].lvl < lvl) ids.push_back((int)b + 1);
    if (ids.empty()) break;
    int a = afhost_refine(t, (int)ids.size(), ids.data());
    if (a < 0) return -1;
    total += a;
  }
  return total;
}

int afhost_tree_info(afhost_tree_t *t, int *n_boxes, int *highest_lvl, long long *n_leaf_cells,
                     int *n_leaves) {
  if (!t) return 1;
  long long leaves = 0;
  for (size_t b = 0; b < t->boxes.size(); b++)
    if (t->is_leaf((int)b + 1)) leaves++;
  long long cells = leaves;
  for (int d = 0; d < t->D; d++) cells *= t->nc;
  if (n_boxes) *n_boxes = (int)t->boxes.size();
  if (highest_lvl) *highest_lvl = t->highest();
  if (n_leaf_cells) *n_leaf_cells = cells;
  if (n_leaves) *n_leaves = (int)leaves;
  return 0;
}

// Arrays in the layout of afgpu_sync_tree.
int afhost_tree_export(afhost_tree_t *t, int *lvl, int *ix, int *parent, int *children, int *neighbors,
                       int *nbmat, double *r_min, double *dr) {
  if (!t) return 1;
  const int D = t->D, nch = 1 << D, nnb = 2 * D, nm = (D == 3) ? 27 : 9;
  for (size_t b = 0; b < t->boxes.size(); b++) {
    const HBox &x = t->boxes[b];
    lvl[b] = x.lvl;
    parent[b] = x.parent;
    for (int d = 0; d < D; d++) {
      ix[b * D + d] = x.ix[d];
      r_min[b * D + d] = x.r_min[d];
      dr[b * D + d] = x.dr[d];
    }
    for (int q = 0; q < nch; q++) children[b * nch + q] = x.children[q];
    for (int q = 0; q < nnb; q++) neighbors[b * nnb + q] = x.nb[q];
    for (int q = 0; q < nm; q++) nbmat[b * nm + q] = x.nbmat[q];
  }
  return 0;
}

// Counter-based uniform variate in [0, 1) (splitmix64 finaliser), so that
// the noise does not depend on box order or thread count.
static double uniform01(uint64_t seed, uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  uint64_t z = seed ^ (a * 0x9E3779B97F4A7C15ULL) ^ (b * 0xC2B2AE3D27D4EB4FULL) ^
               (c * 0x165667B19E3779F9ULL) ^ (d * 0xD6E8FEB86659FD93ULL);
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Initial electron density on all (nc+2)^D cells of the given boxes, as
// init_cond_set_box (src/m_init_cond.f90:217-291): background with seeded
// multiplicative noise (1 + noise*(2U-1)) plus a Gaussian seed
// exp(-(d/w)^2) (src/m_geometry.f90:96-100) around seed_r0.  Positive ions
// equal electrons (neutral start), so only one array is produced.
int afhost_init_cond(afhost_tree_t *t, int n_ids, const int *ids, double background, double noise,
                     unsigned long long seed, double seed_density, double seed_width, const double *seed_r0,
                     double *out) {
  if (!t || (n_ids > 0 && (!ids || !out || !seed_r0))) return 1;
  const int D = t->D, nc = t->nc, ng = nc + 2;
  size_t ncell = 1;
  for (int d = 0; d < D; d++) ncell *= ng;
  for (int n = 0; n < n_ids; n++) {
    if (ids[n] < 1 || ids[n] > (int)t->boxes.size()) return 1;
    const HBox &x = t->boxes[ids[n] - 1];
    double *o = out + (size_t)n * ncell;
    int kmax = (D == 3) ? ng - 1 : 0;
    for (int k = 0; k <= kmax; k++)
      for (int j = 0; j < ng; j++)
        for (int i = 0; i < ng; i++) {
          int q[3] = {i, j, k};
          double r[3] = {0, 0, 0};
          long long g[3] = {0, 0, 0};
          for (int d = 0; d < D; d++) {
            r[d] = x.r_min[d] + (q[d] - 0.5) * x.dr[d];
            g[d] = (long long)(x.ix[d] - 1) * nc + q[d];
          }
          double dist2 = 0;
          for (int d = 0; d < D; d++) dist2 += (r[d] - seed_r0[d]) * (r[d] - seed_r0[d]);
          double dist = std::sqrt(dist2);
          double u = uniform01(seed, (uint64_t)x.lvl, (uint64_t)(g[0] + 7), (uint64_t)(g[1] + 7),
                               (uint64_t)(g[2] + 7));
          double bg = background * (1 + noise * (2 * u - 1));
          double val = std::exp(-(dist / seed_width) * (dist / seed_width)) * seed_density;
          o[i + (size_t)ng * (j + (size_t)ng * k)] = bg + val;
        }
  }
  return 0;
}

// LT_lin_interp_list (src/lookup_table_fortran/m_lookup_table.f90:160-180)
static double lin_interp(const std::vector<double> &x, const std::vector<double> &y, double xv) {
  const size_t n = x.size();
  if (xv <= x[0]) return y[0];
  if (xv >= x[n - 1]) return y[n - 1];
  size_t lo = 0, hi = n - 1;  // smallest ix with x[ix] >= xv
  while (lo < hi) {
    size_t mid = lo + ((hi - lo) >> 1);
    if (x[mid] >= xv)
      hi = mid;
    else
      lo = mid + 1;
  }
  size_t ix = lo;
  double temp = (xv - x[ix - 1]) / (x[ix] - x[ix - 1]);
  return (1 - temp) * y[ix - 1] + temp * y[ix];
}

// Old-style transport data -> td_tbl (mobility*N, diffusion*N, alpha/N,
// eta/N vs E/N in Td; src/m_transport_data.f90:85-126, LT_create
// m_lookup_table.f90:186-233) and chemtbl_fld with the ionization and
// attachment rate coefficients of the standard model
// (src/m_chemistry.f90:233-256, 345-372).  Columns are [col][row].
int afhost_transport_tables(int n_in, const double *E, const double *mu, const double *dif, const double *alpha,
                            const double *eta, double N, int table_size, double table_max_td,
                            double *td_x_min, double *td_inv_fac, double *td_cols, double *chem_x_min,
                            double *chem_inv_fac, double *chem_cols) {
  if (n_in < 2 || table_size < 2 || !E || !mu || !dif || !alpha || !eta) return 1;
  const double SI_to_Townsend = 1e21, Townsend_to_SI = 1e-21;
  std::vector<double> xx(n_in), y0(n_in), y1(n_in), y2(n_in), y3(n_in);
  for (int i = 0; i < n_in; i++) {
    xx[i] = E[i] * SI_to_Townsend / N;
    y0[i] = mu[i] * N;
    y1[i] = dif[i] * N;
    y2[i] = alpha[i] / N;
    y3[i] = eta[i] / N;
  }
  const double x_min = 0.0;
  const double x_max = (table_max_td < 0) ? xx[n_in - 1] : table_max_td;
  std::vector<double> x(table_size);
  const double tmp = 1.0 / (table_size - 1);
  for (int i = 0; i < table_size; i++) x[i] = (i * tmp);
  for (int i = 0; i < table_size; i++) x[i] = x_min + x[i] * (x_max - x_min);
  *td_x_min = x_min;
  *td_inv_fac = (table_size - 1) / (x_max - x_min);
  const std::vector<double> *ys[4] = {&y0, &y1, &y2, &y3};
  for (int c = 0; c < 4; c++)
    for (int i = 0; i < table_size; i++) td_cols[(size_t)c * table_size + i] = lin_interp(xx, *ys[c], x[i]);
  // chemtbl_fld on the same x range and size
  std::vector<double> ion(table_size), att(table_size);
  for (int i = 0; i < table_size; i++) {
    ion[i] = td_cols[2 * (size_t)table_size + i] * td_cols[i] * x[i] * Townsend_to_SI * N;
    att[i] = td_cols[3 * (size_t)table_size + i] * td_cols[i] * x[i] * Townsend_to_SI * N;
  }
  *chem_x_min = x[0];
  *chem_inv_fac = (table_size - 1) / (x[table_size - 1] - x[0]);
  for (int i = 0; i < table_size; i++) {
    chem_cols[i] = lin_interp(x, ion, x[i]);
    chem_cols[(size_t)table_size + i] = lin_interp(x, att, x[i]);
  }
  return 0;
}

}  // extern "C"
