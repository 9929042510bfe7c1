// Exact level-1 solve factors (host code, no GPU use).
//
// The reference hands the level-1 system to hypre PFMG with a 1e-6
// tolerance (afivo/src/m_af_multigrid.f90:266-291,
// afivo/src/m_coarse_solver.f90:393-439).  hypre is not available here, so
// the level-1 system is solved exactly instead.  Level 1 is a single box
// whose faces are all physical boundaries, and afivo's constant-coefficient
// 3/5/7-point stencil with the boundary conditions folded in
// (stencil_handle_boundaries, afivo/src/m_coarse_solver.f90:442-491) is then
// a Kronecker sum  A = T_1 (+) T_2 (+) T_3 - lambda.  Each 1-D operator T_d is
// diagonalised once, T_d = V_d diag(e_d) V_d^-1, and the solve is
//   phi = (V_1 x V_2 x V_3) diag(1/(e_1i + e_2j + e_3k - lambda))
//         (V_1^-1 x V_2^-1 x V_3^-1) b,
// applied axis by axis.  In cylindrical coordinates T_r is not symmetric
// (af_cyl_flux_factors, afivo/src/m_af_types.f90:1199-1211); it is brought
// to symmetric form by a diagonal similarity D T_r D^-1 first.
#include <cmath>
#include <cstring>
#include <vector>

#include "afgpu.h"

namespace {

// Cyclic Jacobi eigenvalue iteration for a symmetric n x n matrix (row
// major).  On return a holds the eigenvalues on its diagonal and q the
// orthonormal eigenvectors as columns.
void jacobi_eigen(int n, std::vector<double> &a, std::vector<double> &q) {
  q.assign(n * n, 0.0);
  for (int i = 0; i < n; i++) q[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        double v = a[i * n + j] * a[i * n + j];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-64 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; p++) {
      for (int r = p + 1; r < n; r++) {
        double apr = a[p * n + r];
        if (apr == 0.0) continue;
        double app = a[p * n + p], arr = a[r * n + r];
        double theta = (arr - app) / (2.0 * apr);
        double t = (theta >= 0 ? 1.0 : -1.0) /
                   (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {  // columns p, r
          double akp = a[k * n + p], akr = a[k * n + r];
          a[k * n + p] = c * akp - s * akr;
          a[k * n + r] = s * akp + c * akr;
        }
        for (int k = 0; k < n; k++) {  // rows p, r
          double apk = a[p * n + k], ark = a[r * n + k];
          a[p * n + k] = c * apk - s * ark;
          a[r * n + k] = s * apk + c * ark;
        }
        for (int k = 0; k < n; k++) {
          double qkp = q[k * n + p], qkr = q[k * n + r];
          q[k * n + p] = c * qkp - s * qkr;
          q[k * n + r] = s * qkp + c * qkr;
        }
      }
    }
  }
}

}  // namespace

extern "C" int afgpu_coarse_factors(int ndim, int nc, int coord,
                                    const double *r_min, const double *dr,
                                    const int *bc_dirichlet, double *vinv,
                                    double *v, double *eig) {
  if ((ndim != 2 && ndim != 3) || nc < 2 || nc > 256 || !r_min || !dr ||
      !bc_dirichlet || !vinv || !v || !eig)
    return AFGPU_ERR_ARG;
  if (coord == AFGPU_CYL && ndim != 2) return AFGPU_ERR_ARG;
  for (int d = 0; d < ndim; d++) {
    // Off-diagonal couplings of the 1-D operator along dimension d:
    // lo[i] multiplies phi(i-1), hi[i] multiplies phi(i+1).
    std::vector<double> lo(nc), hi(nc), T(nc * nc, 0.0);
    const double inv_dr2 = 1.0 / (dr[d] * dr[d]);
    for (int i = 0; i < nc; i++) {
      lo[i] = inv_dr2;
      hi[i] = inv_dr2;
      if (coord == AFGPU_CYL && d == 0) {
        const double r = r_min[0] + ((i + 1) - 0.5) * dr[0];
        const double inv_r = 1.0 / r;
        lo[i] = ((r - 0.5 * dr[0]) * inv_r) * inv_dr2;
        hi[i] = ((r + 0.5 * dr[0]) * inv_r) * inv_dr2;
      }
    }
    for (int i = 0; i < nc; i++) {
      double diag = -(lo[i] + hi[i]);
      if (i > 0) T[i * nc + i - 1] = lo[i];
      if (i < nc - 1) T[i * nc + i + 1] = hi[i];
      // Boundary folding: Dirichlet ghost = 2b - x1, Neumann ghost = x1 + ...
      if (i == 0) diag += bc_dirichlet[2 * d] ? -lo[i] : lo[i];
      if (i == nc - 1) diag += bc_dirichlet[2 * d + 1] ? -hi[i] : hi[i];
      T[i * nc + i] = diag;
    }
    // Similarity S = D T D^-1 is symmetric for d_{i+1} = d_i sqrt(T(i,i+1)/T(i+1,i)).
    std::vector<double> dsc(nc, 1.0), S(nc * nc, 0.0), Q;
    for (int i = 0; i + 1 < nc; i++) {
      const double up = T[i * nc + i + 1], dn = T[(i + 1) * nc + i];
      if (!(up > 0.0) || !(dn > 0.0)) return AFGPU_ERR_ARG;
      dsc[i + 1] = dsc[i] * std::sqrt(up / dn);
    }
    for (int i = 0; i < nc; i++) {
      S[i * nc + i] = T[i * nc + i];
      if (i + 1 < nc) {
        const double s = std::sqrt(T[i * nc + i + 1] * T[(i + 1) * nc + i]);
        S[i * nc + i + 1] = s;
        S[(i + 1) * nc + i] = s;
      }
    }
    jacobi_eigen(nc, S, Q);
    double *Vi = vinv + (size_t)d * nc * nc, *Vd = v + (size_t)d * nc * nc;
    for (int i = 0; i < nc; i++) {
      eig[d * nc + i] = S[i * nc + i];
      for (int m = 0; m < nc; m++) {
        // V = D^-1 Q  (V[i][m] = Q[i][m] / d_i),  V^-1 = Q^T D.
        Vd[i * nc + m] = Q[i * nc + m] / dsc[i];
        Vi[m * nc + i] = Q[i * nc + m] * dsc[i];
      }
    }
  }
  return AFGPU_OK;
}
