/* All-core CPU restatement of the mixed Darcy RT1-P0 assembly and of a preconditioned
 * MINRES solve, FP64, OpenMP. Test infrastructure: the "cpu-opt" baseline timed beside
 * the GPU and checked against the per-element Python restatement of the same algorithm.
 *
 * Element loop (examples/md_elliptic_scalar/md_elliptic_scalar.py:250-343 with
 * LaplaceDualWeakForm.evaluate_form, src/weak_forms/laplace_dual_weak_form.py:22-89):
 *   J_uu[i][j] = sum_q w_q |det| kinv (J phi_i / |det|) . (J phi_j / |det|)
 *   J_up[i]    = -sum_q w_q |det| (div_i / |det|) * 1,  J_pu = -J_up^T,
 *   r_p        = -sum_q w_q |det| f
 * with phi_i the reference RT1 functions times the element's DoF sign (contravariant Piola
 * with the positive determinant, src/geometry/mapping.py:39-61), then the COO scatter of
 * the driver summed into CSR slots in (row, column, element) order.
 * Local order: 4 flux DoFs then the pressure DoF; staged row-major 5 x 5 per element. */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <omp.h>

int ocpu_threads(void) { return omp_get_max_threads(); }

/* X (E, 4, 3) vertex coordinates; sign (E, 4); kinv, f (E); phi (nq, 4, 3) and div (4)
 * reference RT1 values / divergences; w (nq). Writes stage (E, 25) and rstage (E, 5). */
void ocpu_darcy_rt1p0(int64_t E, int nq, const double* X, const double* sign, const double* kinv,
                      const double* f, const double* phi, const double* div, const double* w,
                      double* stage, double* rstage) {
  double wsum = 0.0;
  for (int q = 0; q < nq; ++q) wsum += w[q];
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < E; ++e) {
    const double* x = X + e * 12;
    double J[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) J[r][c] = x[(c + 1) * 3 + r] - x[r];
    const double det = fabs(J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                            J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]));
    const double* sg = sign + e * 4;
    double M[4][4] = {{0}};
    for (int q = 0; q < nq; ++q) {
      double v[4][3];
      for (int i = 0; i < 4; ++i) {
        const double* p = phi + (q * 4 + i) * 3;
        for (int r = 0; r < 3; ++r) v[i][r] = sg[i] * (J[r][0] * p[0] + J[r][1] * p[1] + J[r][2] * p[2]) / det;
      }
      const double dw = det * w[q] * kinv[e];
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) M[i][j] += dw * (v[i][0] * v[j][0] + v[i][1] * v[j][1] + v[i][2] * v[j][2]);
    }
    double* s = stage + e * 25;
    for (int i = 0; i < 4; ++i) {
      for (int j = 0; j < 4; ++j) s[i * 5 + j] = M[i][j];
      const double b = wsum * sg[i] * div[i];  /* sum_q w |det| (div_i / |det|) */
      s[i * 5 + 4] = -b;
      s[20 + i] = b;
    }
    s[24] = 0.0;
    double* rs = rstage + e * 5;
    rs[0] = rs[1] = rs[2] = rs[3] = 0.0;
    rs[4] = -det * wsum * f[e];
  }
}

/* out[m] = sum of stage[perm[k]] for k in [segptr[m], segptr[m+1]), left to right. */
void ocpu_segsum(int64_t nseg, const int64_t* segptr, const int64_t* perm, const double* stage,
                 double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t m = 0; m < nseg; ++m) {
    const int64_t b = segptr[m], en = segptr[m + 1];
    double s = en > b ? stage[perm[b]] : 0.0;
    for (int64_t k = b + 1; k < en; ++k) s += stage[perm[k]];
    out[m] = s;
  }
}

static void spmv(int64_t n, const int64_t* rp, const int64_t* col, const double* val,
                 const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += val[k] * x[col[k]];
    y[i] = s;
  }
}

static double dot(int64_t n, const double* a, const double* b) {
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Preconditioned MINRES (Paige & Saunders 1975; the Lanczos recurrence with the SPD
 * diagonal preconditioner pinv and a QR update by Givens rotations) on the symmetric
 * system S A x = S b with S = diag(sgn), started from x = 0; stops when the recurrence's
 * residual estimate ||r||_{M^-1} falls below rtol * ||b||_{M^-1} or after maxit steps.
 * A in CSR (FP64). work: 7 n doubles. Returns the iteration count. */
int64_t ocpu_minres(int64_t n, const int64_t* rp, const int64_t* col, const double* val,
                    const double* sgn, const double* pinv, const double* b, double* x, double rtol,
                    int64_t maxit, double* work) {
  double *r1 = work, *r2 = work + n, *y = work + 2 * n, *v = work + 3 * n, *w = work + 4 * n,
         *w1 = work + 5 * n, *w2 = work + 6 * n;
  int64_t i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < n; ++i) {
    x[i] = 0.0;
    r1[i] = sgn[i] * b[i];
    r2[i] = r1[i];
    y[i] = pinv[i] * r1[i];
    w[i] = w1[i] = w2[i] = 0.0;
  }
  double beta1 = sqrt(dot(n, r1, y));
  if (beta1 == 0.0) return 0;
  double beta = beta1, oldb = 0.0, dbar = 0.0, epsln = 0.0, phibar = beta1;
  double cs = -1.0, sn = 0.0;
  int64_t it = 0;
  while (it < maxit) {
    ++it;
    const double s = 1.0 / beta;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) v[i] = s * y[i];
    spmv(n, rp, col, val, v, y);
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) y[i] *= sgn[i];
    if (it >= 2) {
      const double c = beta / oldb;
#pragma omp parallel for schedule(static)
      for (i = 0; i < n; ++i) y[i] -= c * r1[i];
    }
    const double alfa = dot(n, v, y);
    const double c2 = alfa / beta;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
      y[i] -= c2 * r2[i];
      r1[i] = r2[i];
      r2[i] = y[i];
      y[i] = pinv[i] * r2[i];
    }
    oldb = beta;
    beta = sqrt(dot(n, r2, y));
    const double oldeps = epsln;
    const double delta = cs * dbar + sn * alfa;
    const double gbar = sn * dbar - cs * alfa;
    epsln = sn * beta;
    dbar = -cs * beta;
    double gamma = sqrt(gbar * gbar + beta * beta);
    if (gamma < 1e-300) gamma = 1e-300;
    cs = gbar / gamma;
    sn = beta / gamma;
    const double phi = cs * phibar;
    phibar = sn * phibar;
    const double dn = 1.0 / gamma;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; ++i) {
      const double wn = (v[i] - oldeps * w1[i] - delta * w2[i]) * dn;
      w1[i] = w2[i];
      w2[i] = wn;
      x[i] += phi * wn;
    }
    if (phibar <= rtol * beta1) break;
  }
  return it;
}

/* r = b - A x, returns ||r||_2. */
double ocpu_residual(int64_t n, const int64_t* rp, const int64_t* col, const double* val,
                     const double* b, const double* x, double* r) {
  spmv(n, rp, col, val, x, r);
  double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i] - r[i];
    s += r[i] * r[i];
  }
  return sqrt(s);
}
