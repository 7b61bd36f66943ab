/*
 * inca_oracle.c -- CPU FP64 restatement of the INCA intracule hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, and the CPU baseline timed beside
 * the GPU).  Nothing in the product links, loads or calls this file.
 *
 * Reference: markelgtx/INCA_parallel, Fortran 90 + OpenMP.  Each function
 * cites the routine it restates.  The restatement keeps the reference's
 * arithmetic (the same Vandermonde + partial-pivot LU as reference LAPACK
 * dgesv, the same thresholds and filters; integer powers by repeated
 * multiplication, see ipow), so that on well-conditioned inputs it agrees
 * with the reference to rounding.
 *
 * How each part is pinned (tests/test_oracle.py):
 *  - whole path: the reference's CH4 goldens (integral, I(0), two points,
 *    the 101-row .rad scan and all 8000 cube voxels) at <= 1e-10 relative;
 *  - normprim: unit self-overlap of every normalised primitive;
 *  - gauherm: numpy.polynomial.hermite.hermgauss nodes and weights;
 *  - polycoef + grid loop: closed form for s-only quartets, the per-quartet
 *    sum rule  int f_q(u) d3u = S_ik S_jl  for s/p quartets, and an exact
 *    polynomial expansion for s/p/d quartets;
 *  - gauleg: numpy leggauss; lebedev (input): exactness on harmonics;
 *  - TFVC: partition of unity, inversion symmetry, integrals of normalised
 *    Gaussians;
 *  - centercalc: hand-computed centres for CH4 and two-heavy-atom cases;
 *  - smearing: the closed form of one point (a normal density on the
 *    reference's window, I_avg = P / (4 pi u^2)), linearity;
 *  - the |C| > 1e3 zeroing of ill-conditioned polycoef fits: a 60-digit
 *    audit (tests/golden/polycoef_audit.py) that classifies every fit as
 *    decided by the mathematics or by rounding; every decision the
 *    mathematics fixes is checked in both arithmetic modes (orc_set_arith)
 *    and in an FMA build (tests/test_polycoef_audit.py).
 */
#include "inca_oracle.h"

#include <float.h>
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI = 3.14159265358979323846;

/* The reference writes integer powers as x**dble(n) (libm pow).  Here they
 * are n-1 successive multiplications: libm pow is correctly rounded, so the
 * two differ by a few ulp at most, which matters only where the Vandermonde
 * fit is ill-conditioned -- and there the reference's own output depends on
 * its libm, BLAS and compiler.  Fractional powers are written with sqrt. */
static double ipow(double x, int e) {
  double r = 1.0;
  for (int k = 0; k < e; ++k) r = (k == 0) ? x : r * x;
  return r;
}

/* Arithmetic mode (orc_set_arith).  0, the default: integer powers by
 * repeated multiplication and Gauss-Hermite roots as correctly rounded
 * doubles -- the operation sequence the device pre-pass follows.  1: the
 * reference's literal forms through libm: x**dble(n) as pow(x, n)
 * (mod_intrastuff.f90:118-122, 153, 189-192), JA8 as one left-to-right
 * product with pi**1.5 and (2 aik)**-(L+1.5) (:117-122), the Gauss-Hermite
 * roots as 16-significant-digit decimals in the reference's node order
 * (:222-298; x_max of two nodes to 15 digits, :44-113) and the Hermite
 * polynomials of the weight formula written out in powers (:312-340).
 * Both are FP64 implementations of the same algorithm; on well-conditioned
 * fits they agree to rounding, and their spread on ill-conditioned fits is
 * what mode 1 exists to measure.  Set it before a run: it is global. */
static int g_arith = 0;
/* per-thread override of the mode (orc_fit_census), -1 when unset */
static _Thread_local int t_arith = -1;
static inline int arith_mode(void) { return t_arith >= 0 ? t_arith : g_arith; }
void orc_set_arith(int mode) { g_arith = mode; }
int orc_get_arith(void) { return g_arith; }

static double rpow(double x, int e) { return arith_mode() ? pow(x, (double)e) : ipow(x, e); }

/* Gauss-Hermite roots for n = 1..11 nodes (ascending).  Values are the
 * roots of H_n computed with numpy.polynomial.hermite.hermgauss; they are
 * the same constants mod_intrastuff.f90:222-298 hard-codes. */
static const double GH_ROOT[11][11] = {
    {0},
    {-0.70710678118654746, 0.70710678118654746},
    {-1.2247448713915889, 0, 1.2247448713915889},
    {-1.6506801238857847, -0.52464762327529035, 0.52464762327529035,
     1.6506801238857847},
    {-2.0201828704560856, -0.95857246461381851, 0, 0.95857246461381851,
     2.0201828704560856},
    {-2.3506049736744923, -1.335849074013697, -0.43607741192761651,
     0.43607741192761651, 1.335849074013697, 2.3506049736744923},
    {-2.6519613568352334, -1.6735516287674714, -0.81628788285896459, 0,
     0.81628788285896459, 1.6735516287674714, 2.6519613568352334},
    {-2.9306374202572441, -1.981656756695843, -1.1571937124467802,
     -0.38118699020732211, 0.38118699020732211, 1.1571937124467802,
     1.981656756695843, 2.9306374202572441},
    {-3.1909932017815277, -2.2665805845318432, -1.4685532892166679,
     -0.72355101875283756, 0, 0.72355101875283756, 1.4685532892166679,
     2.2665805845318432, 3.1909932017815277},
    {-3.4361591188377374, -2.5327316742327897, -1.7566836492998816,
     -1.0366108297895136, -0.34290132722370459, 0.34290132722370459,
     1.0366108297895136, 1.7566836492998816, 2.5327316742327897,
     3.4361591188377374},
    {-3.6684708465595826, -2.7832900997816514, -2.0259480158257555,
     -1.3265570844949328, -0.65680956688209979, 0, 0.65680956688209979,
     1.3265570844949328, 2.0259480158257555, 2.7832900997816514,
     3.6684708465595826}};

/* mod_numbers.f90:14-27 */
int orc_dfact(int a) {
  int d = 1;
  if (a % 2 == 0) {
    for (int i = 1; i <= a / 2; ++i) d *= 2 * i;
  } else {
    for (int i = 1; i <= (a + 1) / 2; ++i) d *= 2 * i - 1;
  }
  return d;
}

/* mod_intracule.f90:148-150 */
double orc_normprim(double alpha, int t, int m, int n) {
  double df = (double)(orc_dfact(2 * t - 1) * orc_dfact(2 * m - 1) *
                       orc_dfact(2 * n - 1));
  return pow(2.0 * alpha / PI, 0.75) *
         sqrt(pow(4.0 * alpha, (double)(t + m + n)) * (1.0 / df));
}

/* Physicists' Hermite polynomial H_{n-1}(x) by its three-term recurrence
 * (mod_intrastuff.f90:312-340 writes the same polynomials out). */
static double hermite_nm1(double x, int n) {
  double h0 = 1.0, h1 = 2.0 * x;
  if (n == 1) return h0;
  for (int k = 2; k < n; ++k) {
    double h2 = 2.0 * x * h1 - 2.0 * (k - 1) * h0;
    h0 = h1;
    h1 = h2;
  }
  return h1;
}

/* The root of H_n nearest x0, refined by Newton steps in long double
 * (64-bit significand) and rounded to `digits` significant decimal digits,
 * as the reference writes its root constants. */
static double root_decimal(double x0, int n, int digits) {
  long double x = x0;
  for (int it = 0; it < 4; ++it) {
    long double h0 = 1.0L, h1 = 2.0L * x; /* H_0, H_1 */
    for (int k = 2; k <= n; ++k) {
      long double h2 = 2.0L * x * h1 - 2.0L * (k - 1) * h0;
      h0 = h1;
      h1 = h2;
    }
    /* H_n' = 2 n H_(n-1) */
    if (n >= 1 && h0 != 0.0L) x -= h1 / (2.0L * n * h0);
  }
  char buf[64];
  snprintf(buf, sizeof buf, "%.*Lg", digits, x);
  return strtod(buf, NULL);
}

/* Gauss-Hermite nodes of mode 1: the roots to 16 significant digits
 * (0 exactly for the middle node), in the reference's order -- ascending,
 * except for five nodes: -r2, -r1, 0, r2, r1 with r2 > r1
 * (mod_intrastuff.f90:222-298). */
static void gh_roots_ref(int n, double *rh) {
  double r[11];
  for (int i = 0; i < n; ++i) {
    double x0 = GH_ROOT[n - 1][i];
    r[i] = (x0 == 0.0) ? 0.0 : root_decimal(x0, n, 16);
  }
  if (n == 5) {
    rh[0] = r[0];
    rh[1] = r[1];
    rh[2] = r[2];
    rh[3] = r[4];
    rh[4] = r[3];
  } else {
    for (int i = 0; i < n; ++i) rh[i] = r[i];
  }
}

/* H_(n-1)(x) written out in powers of x (physicists' Hermite polynomials,
 * coefficients from the three-term recurrence), each power through libm pow
 * as the reference's x**k (mod_intrastuff.f90:312-340). */
static double hermite_powers(double x, int n) {
  double c[12] = {1.0}, prev[12] = {0.0}, cur[12];
  int deg = 0; /* c holds H_deg */
  for (; deg < n - 1; ++deg) {
    for (int m = 0; m < 12; ++m) cur[m] = 0.0;
    for (int m = 0; m <= deg; ++m) cur[m + 1] += 2.0 * c[m];
    for (int m = 0; m < 12; ++m) cur[m] -= 2.0 * deg * prev[m];
    for (int m = 0; m < 12; ++m) {
      prev[m] = c[m];
      c[m] = cur[m];
    }
  }
  double h = 0.0;
  for (int m = deg; m >= 0; --m) {
    if (c[m] == 0.0) continue;
    h += (m == 0) ? c[m] : c[m] * (m == 1 ? x : pow(x, (double)m));
  }
  return h;
}

/* mod_intrastuff.f90:198-341.  n = floor(L/2)+1 nodes; weight
 * w_i = 2^(n-1) n! sqrt(pi) / (n^2 H_{n-1}(x_i)^2).  Returns n, or -1 for
 * L > 21 (the reference prints an error and continues with garbage). */
int orc_gauherm(int L, double *rh, double *wh) {
  int n;
  if (L > 0)
    n = (L % 2 == 0) ? L / 2 + 1 : (L + 1) / 2;
  else
    n = 1;
  if (n > 11) return -1;
  double factn = 1.0;
  for (int i = 2; i <= n; ++i) factn *= i;
  double cons = pow(2.0, (double)(n - 1)) * factn * sqrt(PI) / (double)(n * n);
  if (arith_mode()) {
    gh_roots_ref(n, rh);
    for (int i = 0; i < n; ++i) wh[i] = cons * (1.0 / pow(hermite_powers(rh[i], n), 2.0));
    return n;
  }
  for (int i = 0; i < n; ++i) {
    rh[i] = GH_ROOT[n - 1][i];
    double h = hermite_nm1(rh[i], n);
    wh[i] = cons * (1.0 / (h * h));
  }
  return n;
}

/* Largest Gauss-Hermite root for nn nodes (JA8's x_max, :44-113); mode 1
 * takes the reference's decimals (16 digits, 15 for two nodes). */
static double gh_xmax(int nn) {
  if (nn <= 1) return 0.0;
  if (nn > 11) nn = 11; /* undefined in the reference beyond 10 */
  if (arith_mode()) return root_decimal(GH_ROOT[nn - 1][nn - 1], nn, nn == 2 ? 15 : 16);
  return GH_ROOT[nn - 1][nn - 1];
}

/* mod_intrastuff.f90:32-123 */
double orc_ja8(int ti, int tk, int mi, int mk, int ni, int nk, double alfi,
               double alfk, double xi, double xk, double yi, double yk,
               double zi, double zk, double rik2, double aik, double eik) {
  double xm0 = gh_xmax(ti + tk + 1), xm1 = gh_xmax(mi + mk + 1),
         xm2 = gh_xmax(ni + nk + 1);
  double s = sqrt(2.0 / aik);
  double L = (double)(ti + tk + mi + mk + ni + nk);
  double dx = fabs(xk - xi), dy = fabs(yk - yi), dz = fabs(zk - zi);
  const double two_a = 2.0 * aik;
  if (arith_mode()) /* the reference's product, left to right (:117-122) */
    return pow(PI, 1.5) * pow(two_a, -(L + 1.5)) * exp(-2.0 * eik * rik2) *
           pow(xm0 + alfk * s * dx, 2.0 * ti) * pow(xm0 + alfi * s * dx, 2.0 * tk) *
           pow(xm1 + alfk * s * dy, 2.0 * mi) * pow(xm1 + alfi * s * dy, 2.0 * mk) *
           pow(xm2 + alfk * s * dz, 2.0 * ni) * pow(xm2 + alfi * s * dz, 2.0 * nk);
  double J = 5.568327996831708 * (1.0 / (ipow(two_a, (int)L + 1) * sqrt(two_a))) *
             exp(-2.0 * eik * rik2);
  J *= ipow(xm0 + alfk * s * dx, 2 * ti) * ipow(xm0 + alfi * s * dx, 2 * tk);
  J *= ipow(xm1 + alfk * s * dy, 2 * mi) * ipow(xm1 + alfi * s * dy, 2 * mk);
  J *= ipow(xm2 + alfk * s * dz, 2 * ni) * ipow(xm2 + alfi * s * dz, 2 * nk);
  return J;
}

/* Partial-pivot LU solve of a column-major np x np system, in the operation
 * order of reference LAPACK dgetrf2/dgetrs with reference BLAS (pivot =
 * first max |a|, multipliers scaled by the reciprocal pivot, updates
 * a_ij -= l_ik u_kj in increasing k, forward/back substitution column by
 * column).  Returns INFO (0, or k if U(k,k) == 0). */
static int lu_solve(int n, double *M, double *b) {
  int info = 0;
  int ipiv[32];
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = fabs(M[k + n * k]);
    for (int i = k + 1; i < n; ++i) {
      double v = fabs(M[i + n * k]);
      if (v > amax) {
        amax = v;
        p = i;
      }
    }
    ipiv[k] = p;
    if (M[p + n * k] != 0.0) {
      if (p != k)
        for (int j = 0; j < n; ++j) {
          double t = M[k + n * j];
          M[k + n * j] = M[p + n * j];
          M[p + n * j] = t;
        }
      double piv = M[k + n * k];
      if (fabs(piv) >= DBL_MIN) {
        double rp = 1.0 / piv;
        for (int i = k + 1; i < n; ++i) M[i + n * k] *= rp;
      } else {
        for (int i = k + 1; i < n; ++i) M[i + n * k] /= piv;
      }
    } else if (info == 0) {
      info = k + 1;
    }
    for (int j = k + 1; j < n; ++j) {
      double ukj = M[k + n * j];
      for (int i = k + 1; i < n; ++i) M[i + n * j] -= M[i + n * k] * ukj;
    }
  }
  if (info != 0) return info;
  for (int k = 0; k < n; ++k) {
    int p = ipiv[k];
    if (p != k) {
      double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
  }
  for (int k = 0; k < n; ++k)
    if (b[k] != 0.0)
      for (int i = k + 1; i < n; ++i) b[i] -= b[k] * M[i + n * k];
  for (int k = n - 1; k >= 0; --k)
    if (b[k] != 0.0) {
      b[k] /= M[k + n * k];
      for (int i = 0; i < k; ++i) b[i] -= b[k] * M[i + n * k];
    }
  return 0;
}

/* mod_intrastuff.f90:178-193 (Wijkl) */
static double wijkl(double ri, double rj, double rk, double rl, double rijkl,
                    int ti, int tj, int tk, int tl, double rhat, double r,
                    double invz, double alf) {
  double a = invz * rhat;
  double ap = (alf + 0.5) * r, am = (alf - 0.5) * r;
  return rpow(a + am + (rijkl - ri), ti) * rpow(a + ap + (rijkl - rj), tj) *
         rpow(a + am + (rijkl - rk), tk) * rpow(a + ap + (rijkl - rl), tl);
}

/* mod_intrastuff.f90:127-175 */
int orc_polycoef(double ri, double rj, double rk, double rl, double rik,
                 double rjl, double rijkl, int ti, int tj, int tk, int tl,
                 int np, double sqe, double invz, double alfijkl, double *C) {
  double rh[11], wh[11];
  int L = np - 1;
  int n = orc_gauherm(L, rh, wh);
  if (np == 1) {
    C[0] = wh[0];
    return 0;
  }
  if (n < 0) {
    for (int i = 0; i < np; ++i) C[i] = 0.0;
    return 1;
  }
  double M[22 * 22], xx[22];
  for (int i = 0; i < np; ++i) xx[i] = -0.5 + (double)(i + 1);
  for (int i = 0; i < np; ++i) {
    int l1 = L;
    for (int j = 0; j < np; ++j) {
      M[i + np * j] = rpow(sqe * (xx[i] + rik - rjl), l1);
      --l1;
    }
  }
  for (int i = 0; i < np; ++i) {
    double c = 0.0;
    for (int j = 0; j < n; ++j)
      c += wh[j] * wijkl(ri, rj, rk, rl, rijkl, ti, tj, tk, tl, rh[j], xx[i],
                         invz, alfijkl);
    C[i] = c;
  }
  int info = lu_solve(np, M, C);
  int bad = (info != 0);
  for (int i = 0; i < np && !bad; ++i)
    if (fabs(C[i]) > 1.0e3 || C[i] != C[i]) bad = 1;
  if (bad) {
    for (int i = 0; i < np; ++i) C[i] = 0.0;
    return 1;
  }
  return 0;
}

/* Exact coefficients of the polynomial polycoef fits: with
 * p(r) = sum_n w_n prod_a (invz t_n + s_a r + (Rijkl - R_a))^{t_a},
 * s = alfijkl - 1/2 for i,k and + 1/2 for j,l (mod_intrastuff.f90:178-193),
 * the fit is q(y) = p(r) on y = sqe (r + d), d = P_ik - P_jl.  Each factor
 * b + s r is (b - s d) + (s / sqe) y, so q is expanded directly in the
 * reference's own variable instead of interpolated; E[0] multiplies y^L.
 * No reference counterpart: the well-conditioned alternative to the
 * Vandermonde fit (same polynomial in exact arithmetic). */
static void exact_poly(const double R[4], double rijkl, const int t[4], int L,
                       double invz, double alf, double sqe, double d,
                       double *E) {
  double rh[11], wh[11];
  int n = orc_gauherm(L, rh, wh);
  double acc[22] = {0.0};
  for (int k = 0; k < n; ++k) {
    double poly[22] = {1.0};
    int deg = 0;
    const double a = invz * rh[k];
    for (int f = 0; f < 4; ++f) {
      const double s0 = (f == 0 || f == 2) ? alf - 0.5 : alf + 0.5;
      const double b = (a + (rijkl - R[f])) - s0 * d;
      const double sl = s0 / sqe;
      for (int e = 0; e < t[f]; ++e) {
        poly[deg + 1] = poly[deg] * sl;
        for (int m = deg; m >= 1; --m) poly[m] = poly[m] * b + poly[m - 1] * sl;
        poly[0] *= b;
        ++deg;
      }
    }
    for (int m = 0; m <= deg; ++m) acc[m] += wh[k] * poly[m];
  }
  for (int q = 0; q <= L; ++q) E[q] = acc[L - q];
}

/* mod_intracule.f90:307-312: the Gaussian-product quantities of a quartet
 * passing screen 1 (R[ax][0..3] = coordinate of primitives i, j, k, l). */
static void quartet_geometry(double alfi, double alfj, double alfk, double alfl,
                             double *const R[3], double *sqe, double *invz, double *alfijkl,
                             double P_ik[3], double P_jl[3], double P[3]) {
  double aik = alfi + alfk, ajl = alfj + alfl;
  double zeta = aik + ajl;
  double eijkl = (aik * ajl) * (1.0 / zeta);
  *sqe = sqrt(eijkl);
  for (int ax = 0; ax < 3; ++ax) {
    P_ik[ax] = (alfi * R[ax][0] + alfk * R[ax][2]) * (1.0 / aik);
    P_jl[ax] = (alfj * R[ax][1] + alfl * R[ax][3]) * (1.0 / ajl);
    P[ax] = (aik * P_ik[ax] + ajl * P_jl[ax]) * (1.0 / zeta);
  }
  *invz = 1.0 / sqrt(zeta);
  *alfijkl = 0.5 * (aik - ajl) * (1.0 / zeta);
}

static inline int TMN(const int32_t *tmn, int nprim, int i, int ax) {
  return tmn[i + nprim * ax];
}

/* mod_intracule.f90:282-344 */
void orc_prepare_quartet(int nprim, const double *alpha, const int32_t *tmn,
                         const double *x, const double *y, const double *z,
                         const double *normprim, double lim, int i, int j,
                         int k, int l, double D, orc_quartet *q) {
  memset(q, 0, sizeof(*q));
  --i;
  --j;
  --k;
  --l;
  double DMval = D;
  if (normprim) DMval = normprim[i] * normprim[j] * normprim[k] * normprim[l] * D;
  double alfi = alpha[i], alfj = alpha[j], alfk = alpha[k], alfl = alpha[l];
  double X[4] = {x[i], x[j], x[k], x[l]};
  double Y[4] = {y[i], y[j], y[k], y[l]};
  double Z[4] = {z[i], z[j], z[k], z[l]};
  int t[4][3];
  int idx[4] = {i, j, k, l};
  for (int a = 0; a < 4; ++a)
    for (int ax = 0; ax < 3; ++ax) t[a][ax] = TMN(tmn, nprim, idx[a], ax);
  double aik = alfi + alfk, ajl = alfj + alfl;
  double eik = alfi * alfk * (1.0 / aik), ejl = alfj * alfl * (1.0 / ajl);
  double rik2 = (X[0] - X[2]) * (X[0] - X[2]) + (Y[0] - Y[2]) * (Y[0] - Y[2]) +
                (Z[0] - Z[2]) * (Z[0] - Z[2]);
  double rjl2 = (X[1] - X[3]) * (X[1] - X[3]) + (Y[1] - Y[3]) * (Y[1] - Y[3]) +
                (Z[1] - Z[3]) * (Z[1] - Z[3]);
  double jik = orc_ja8(t[0][0], t[2][0], t[0][1], t[2][1], t[0][2], t[2][2],
                       alfi, alfk, X[0], X[2], Y[0], Y[2], Z[0], Z[2], rik2,
                       aik, eik);
  double jjl = orc_ja8(t[1][0], t[3][0], t[1][1], t[3][1], t[1][2], t[3][2],
                       alfj, alfl, X[1], X[3], Y[1], Y[3], Z[1], Z[3], rjl2,
                       ajl, ejl);
  q->screen1 = fabs(DMval) * sqrt(jik * jjl);
  if (!(q->screen1 >= lim)) return;
  q->pass1 = 1;
  double zeta = aik + ajl;
  double sqe, invz, alfijkl;
  double P_ik[3], P_jl[3], P[3];
  double *R[3] = {X, Y, Z};
  quartet_geometry(alfi, alfj, alfk, alfl, R, &sqe, &invz, &alfijkl, P_ik, P_jl, P);
  double A = DMval * (1.0 / (zeta * sqrt(zeta))) * exp(-eik * rik2 - ejl * rjl2);
  double U[3];
  for (int ax = 0; ax < 3; ++ax) {
    int L = t[0][ax] + t[1][ax] + t[2][ax] + t[3][ax];
    q->L[ax] = L;
    int np = L + 1;
    if (orc_polycoef(R[ax][0], R[ax][1], R[ax][2], R[ax][3], P_ik[ax],
                     P_jl[ax], P[ax], t[0][ax], t[1][ax], t[2][ax], t[3][ax],
                     np, sqe, invz, alfijkl, q->C[ax]))
      q->polyfail++;
    {
      const double Rq[4] = {R[ax][0], R[ax][1], R[ax][2], R[ax][3]};
      const int tq[4] = {t[0][ax], t[1][ax], t[2][ax], t[3][ax]};
      exact_poly(Rq, P[ax], tq, L, invz, alfijkl, sqe, P_ik[ax] - P_jl[ax],
                 q->E[ax]);
    }
    /* wmarray, mod_intracule.f90:136-143 */
    double u = 0.0;
    for (int p = 0; p < np; ++p) {
      double wm = (p == 0) ? 1.0
                           : pow((p + 1) * 0.5, (p + 1) * 0.5) *
                                 exp(-(double)(p + 1) * 0.5);
      u += q->C[ax][p] * wm;
    }
    U[ax] = u;
  }
  q->A = A;
  q->sqe = sqe;
  q->dx = P_ik[0] - P_jl[0];
  q->dy = P_ik[1] - P_jl[1];
  q->dz = P_ik[2] - P_jl[2];
  q->screen2 = fabs(A * U[0] * U[1] * U[2]);
  q->pass2 = (q->screen2 >= lim);
}

/* mod_intracule.f90:345-358 for one prepared quartet over all points.
 * exact != 0 evaluates the exact axis polynomials instead of the fits. */
static void grid_loop(const orc_quartet *q, int64_t G, const double *r,
                      double *acc, int exact) {
  const int nx = q->L[0] + 1, ny = q->L[1] + 1, nz = q->L[2] + 1;
  const double(*Cc)[22] = exact ? q->E : q->C;
  for (int64_t g = 0; g < G; ++g) {
    double rx = q->sqe * (r[3 * g] + q->dx);
    double ry = q->sqe * (r[3 * g + 1] + q->dy);
    double rz = q->sqe * (r[3 * g + 2] + q->dz);
    double vx = 0.0, vy = 0.0, vz = 0.0;
    for (int p = 0; p < nx; ++p) vx = vx * rx + Cc[0][p];
    for (int p = 0; p < ny; ++p) vy = vy * ry + Cc[1][p];
    for (int p = 0; p < nz; ++p) vz = vz * rz + Cc[2][p];
    acc[g] += q->A * exp(-(rx * rx + ry * ry + rz * rz)) * vx * vy * vz;
  }
}

int orc_intracule(int nprim, const double *alpha, const int32_t *tmn,
                  const double *x, const double *y, const double *z,
                  int normalize, double thresh, int64_t nrec,
                  const int32_t *rec_ijkl, const double *rec_d, int64_t G,
                  const double *r, double *ivec, int64_t *counts,
                  int nthreads, int dynamic) {
  return orc_intracule_ex(nprim, alpha, tmn, x, y, z, normalize, thresh, nrec,
                          rec_ijkl, rec_d, G, r, ivec, counts, nthreads,
                          dynamic, 0);
}

/* mod_intracule.f90:229-381 */
int orc_intracule_ex(int nprim, const double *alpha, const int32_t *tmn,
                     const double *x, const double *y, const double *z,
                     int normalize, double thresh, int64_t nrec,
                     const int32_t *rec_ijkl, const double *rec_d, int64_t G,
                     const double *r, double *ivec, int64_t *counts,
                     int nthreads, int dynamic, int exact) {
  /* the reference stops at the first invalid record (:239-240) */
  int64_t nvalid = 0;
  for (; nvalid < nrec; ++nvalid) {
    const int32_t *p = rec_ijkl + 4 * nvalid;
    if (p[0] < 1 || p[0] > nprim) break;
    if (p[1] < 1 || p[2] < 1 || p[3] < 1) break;
    if (p[1] > nprim || p[2] > nprim || p[3] > nprim) break;
  }
  double *normprim = NULL;
  if (normalize) {
    normprim = (double *)malloc(sizeof(double) * (size_t)nprim);
    for (int i = 0; i < nprim; ++i)
      normprim[i] = orc_normprim(alpha[i], TMN(tmn, nprim, i, 0),
                                 TMN(tmn, nprim, i, 1), TMN(tmn, nprim, i, 2));
  }
  /* mod_intracule.f90:163 */
  double lim = thresh * (1.0 / ((double)nprim * ((double)nprim + 1.0) * 0.5));
  for (int64_t g = 0; g < G; ++g) ivec[g] = 0.0;
  int64_t c1 = 0, c2 = 0, cf = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : c1, c2, cf)
  {
    double *acc = (double *)calloc((size_t)(G > 0 ? G : 1), sizeof(double));
    orc_quartet q;
    if (dynamic) {
#pragma omp for schedule(dynamic, 64)
      for (int64_t n = 0; n < nvalid; ++n) {
        const int32_t *p = rec_ijkl + 4 * n;
        orc_prepare_quartet(nprim, alpha, tmn, x, y, z, normprim, lim, p[0],
                            p[1], p[2], p[3], rec_d[n], &q);
        c1 += q.pass1;
        cf += q.polyfail;
        if (q.pass1 && q.pass2) {
          ++c2;
          grid_loop(&q, G, r, acc, exact);
        }
      }
    } else {
#pragma omp for schedule(static)
      for (int64_t n = 0; n < nvalid; ++n) {
        const int32_t *p = rec_ijkl + 4 * n;
        orc_prepare_quartet(nprim, alpha, tmn, x, y, z, normprim, lim, p[0],
                            p[1], p[2], p[3], rec_d[n], &q);
        c1 += q.pass1;
        cf += q.polyfail;
        if (q.pass1 && q.pass2) {
          ++c2;
          grid_loop(&q, G, r, acc, exact);
        }
      }
    }
    /* thread-ordered merge (the reference uses an unordered CRITICAL) */
#pragma omp for ordered schedule(static, 1)
    for (int t = 0; t < omp_get_num_threads(); ++t) {
#pragma omp ordered
      for (int64_t g = 0; g < G; ++g) ivec[g] += acc[g];
    }
    free(acc);
  }
  free(normprim);
  if (counts) {
    counts[0] = nvalid;
    counts[1] = c1;
    counts[2] = c2;
    counts[3] = cf;
  }
  return 0;
}

/* mod_quadratures.f90:37-86 (Newton iteration on P_n, EPS = 1e-14) */
void orc_gauleg(int n, double *t, double *w) {
  const double x1 = -1.0, x2 = 1.0, eps = 1.0e-14;
  double pi = acos(-1.0);
  int m = (int)(((double)n + 1.0) / 2.0);
  double xm = 0.5 * (x1 + x2), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; ++i) {
    double zz = cos(pi * ((double)i - 0.25) / ((double)n + 0.5));
    double z1, p1, p2, p3, pp;
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * zz * p2 - ((double)j - 1.0) * p3) / (double)j;
      }
      pp = (double)n * (zz * p1 - p2) / (zz * zz - 1.0);
      z1 = zz;
      zz = z1 - p1 / pp;
    } while (fabs(zz - z1) > eps);
    t[i - 1] = xm - xl * zz;
    t[n - i] = xm + xl * zz;
    w[i - 1] = 2.0 * xl / ((1.0 - zz * zz) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

/* f_k of mod_build_grid.f90:554-574 with k = 4 */
static double tfvc_fk(double v) {
  for (int i = 0; i < 4; ++i) v = 1.5 * v - 0.5 * (v * v * v);
  return v;
}

/* mod_build_grid.f90:474-575 (Salvador's TFVC) */
void orc_tfvc(int64_t npts, const double *pts, const int32_t *owner,
              int nquad, const double *cent, const double *Ps, double *wout) {
  double *Rij = (double *)malloc(sizeof(double) * nquad * nquad);
  double *Xi = (double *)malloc(sizeof(double) * nquad * nquad);
  for (int i = 0; i < nquad; ++i)
    for (int j = 0; j < nquad; ++j) {
      double dx = cent[3 * i] - cent[3 * j], dy = cent[3 * i + 1] - cent[3 * j + 1],
             dz = cent[3 * i + 2] - cent[3 * j + 2];
      Rij[i * nquad + j] = sqrt(dx * dx + dy * dy + dz * dz);
      Xi[i * nquad + j] = (Ps[j] > 1e-12) ? pow(Ps[i] / Ps[j], 1.0 / 3.0) : 1.0;
    }
#pragma omp parallel
  {
    double *P = (double *)malloc(sizeof(double) * nquad);
    double *rr = (double *)malloc(sizeof(double) * nquad);
#pragma omp for schedule(static)
    for (int64_t g = 0; g < npts; ++g) {
      const double *p = pts + 3 * g;
      for (int i = 0; i < nquad; ++i) {
        double dx = p[0] - cent[3 * i], dy = p[1] - cent[3 * i + 1],
               dz = p[2] - cent[3 * i + 2];
        rr[i] = sqrt(dx * dx + dy * dy + dz * dz);
      }
      double Ptot = 0.0;
      for (int i = 0; i < nquad; ++i) {
        double Pi = 1.0;
        for (int j = 0; j < nquad; ++j) {
          if (i == j) continue;
          double R = Rij[i * nquad + j];
          double mu = (R > 1e-12) ? (rr[i] - rr[j]) * (1.0 / R) : 0.0;
          double term = Xi[i * nquad + j] * (1.0 - mu);
          double v = (1.0 + mu - term) / (1.0 + mu + term);
          Pi *= 0.5 * (1.0 - tfvc_fk(v));
        }
        P[i] = Pi;
        Ptot += Pi;
      }
      double wo = P[owner[g]];
      if (Ptot > 1e-15) wo *= 1.0 / Ptot;
      wout[g] = wo;
    }
    free(P);
    free(rr);
  }
  free(Rij);
  free(Xi);
}

/* mod_build_grid.f90:13-277 */
int orc_gridpoints(int nquad, const double *cent, const int32_t *nradc,
                   const int32_t *nangc, const double *sfalpha,
                   const double *Ps, int nosym, const double *ang,
                   int64_t *maxgrid, int64_t *rgrid, int64_t *rrgrid,
                   double *r_out, double *w_out, double *wvee_out,
                   int32_t *map_out, int64_t cap) {
  const double pi = 4.0 * atan(1.0);
  const double tol = 1.0e-100, trsh2 = 1.0e-106;
  int64_t np = 0;
  for (int c = 0; c < nquad; ++c) np += (int64_t)nradc[c] * nangc[c];
  *maxgrid = np;
  if (cap < np) return -1;
  double *gr = (double *)malloc(sizeof(double) * 3 * (size_t)np);
  double *wt = (double *)malloc(sizeof(double) * (size_t)np);
  int32_t *own = (int32_t *)malloc(sizeof(int32_t) * (size_t)np);
  int64_t nk = 0;
  const double *angc = ang;
  for (int c = 0; c < nquad; ++c) {
    int nrad = nradc[c], nang = nangc[c];
    double *xl = (double *)malloc(sizeof(double) * nrad);
    double *wl = (double *)malloc(sizeof(double) * nrad);
    double *rad = (double *)malloc(sizeof(double) * nrad);
    orc_gauleg(nrad, xl, wl);
    for (int ir = 0; ir < nrad; ++ir) {
      rad[ir] = (1.0 + xl[ir]) / (1.0 - xl[ir]) * sfalpha[c];
      if (nquad > 1 && rad[ir] > 15.0) {
        wl[ir] = 0.0;
      } else {
        double om = 1.0 - xl[ir];
        wl[ir] = 2 * pi * (2.0 * sfalpha[c] / (om * om)) * wl[ir] * rad[ir] *
                 rad[ir];
      }
    }
    for (int ir = 0; ir < nrad; ++ir)
      for (int ia = 0; ia < nang; ++ia) {
        const double *a = angc + 4 * ia;
        double px = rad[ir] * a[0] + cent[3 * c];
        double py = rad[ir] * a[1] + cent[3 * c + 1];
        double pz = rad[ir] * a[2] + cent[3 * c + 2];
        double w;
        if (nosym) {
          w = a[3] * wl[ir];
        } else {
          if (!(pz >= -tol)) continue;
          w = (fabs(pz) < tol) ? a[3] * wl[ir] : 2.0 * a[3] * wl[ir];
        }
        gr[3 * nk] = px;
        gr[3 * nk + 1] = py;
        gr[3 * nk + 2] = pz;
        wt[nk] = w;
        own[nk] = c;
        ++nk;
      }
    angc += 4 * nang;
    free(xl);
    free(wl);
    free(rad);
  }
  *rgrid = nk;
  if (nquad > 1) {
    double *wb = (double *)malloc(sizeof(double) * (size_t)(nk > 0 ? nk : 1));
    orc_tfvc(nk, gr, own, nquad, cent, Ps, wb);
    for (int64_t g = 0; g < nk; ++g) wt[g] *= wb[g];
    free(wb);
  }
  int64_t m = 0;
  for (int64_t g = 0; g < nk; ++g) {
    if (!(wt[g] > trsh2)) continue;
    double d = sqrt(gr[3 * g] * gr[3 * g] + gr[3 * g + 1] * gr[3 * g + 1] +
                    gr[3 * g + 2] * gr[3 * g + 2]);
    r_out[3 * m] = gr[3 * g];
    r_out[3 * m + 1] = gr[3 * g + 1];
    r_out[3 * m + 2] = gr[3 * g + 2];
    w_out[m] = wt[g];
    wvee_out[m] = wt[g] * (1.0 / d);
    map_out[m] = own[g] + 1;
    ++m;
  }
  *rrgrid = m;
  free(gr);
  free(wt);
  free(own);
  return 0;
}

/* mod_build_grid.f90:281-408 */
int64_t orc_gridpoints2(int nblock, const double *tart, const double *step,
                        const int32_t *nang, int nosym, const double *ang,
                        int64_t *nradi, double *radi, int32_t *smn,
                        double *r_out, double *w_out, int64_t cap) {
  const double tol = sqrt(DBL_EPSILON);
  int64_t nr = 0, cnt = 0;
  for (int b = 0; b < nblock; ++b)
    nr += (int64_t)((tart[2 * b + 1] - tart[2 * b]) / step[b]) + 1;
  *nradi = nr;
  if (!radi) return 0;
  int64_t sr = 0;
  const double *angb = ang;
  for (int b = 0; b < nblock; ++b) {
    int npb = (int)((tart[2 * b + 1] - tart[2 * b]) / step[b]) + 1;
    for (int j = 0; j < npb; ++j) {
      double s = tart[2 * b] + step[b] * (double)j;
      radi[sr] = s;
      smn[sr] = 0;
      for (int ia = 0; ia < nang[b]; ++ia) {
        const double *a = angb + 4 * ia;
        if (nosym) continue; /* reference quirk: nosym keeps no point */
        double zz = s * a[2];
        if (!(zz >= -tol)) continue;
        if (cnt >= cap) return -1;
        smn[sr]++;
        r_out[3 * cnt] = s * a[0];
        r_out[3 * cnt + 1] = s * a[1];
        r_out[3 * cnt + 2] = zz;
        w_out[cnt] = (fabs(zz) < tol) ? a[3] : 2.0 * a[3];
        ++cnt;
      }
      ++sr;
    }
    angb += 4 * nang[b];
  }
  return cnt;
}

/* mod_build_grid.f90:412-440 */
int64_t orc_gridpoints3(const double *center, const double *step,
                        const int32_t *np, double *r_out) {
  double xm[3];
  for (int i = 0; i < 3; ++i) {
    if (fmod((double)np[i], 2.0) == 0.0)
      xm[i] = center[i] - (step[i] / 2.0) - (((double)np[i] - 2.0) / 2.0) * step[i];
    else
      xm[i] = center[i] - (((double)np[i] - 1.0) / 2.0) * step[i];
  }
  int64_t s = 0;
  if (r_out)
    for (int i = 0; i < np[0]; ++i)
      for (int j = 0; j < np[1]; ++j)
        for (int k = 0; k < np[2]; ++k) {
          r_out[3 * s] = xm[0] + step[0] * (double)i;
          r_out[3 * s + 1] = xm[1] + step[1] * (double)j;
          r_out[3 * s + 2] = xm[2] + step[2] * (double)k;
          ++s;
        }
  return (int64_t)np[0] * np[1] * np[2];
}

/* mod_autogrid.f90:142-185 (Bragg-Slater radii in Angstrom x 1.8897) */
double orc_bs_radius(int iz) {
  static const double r[36] = {0.35, 0.35, 1.45, 1.05, 0.85, 0.70, 0.65,
                               0.60, 0.50, 0.38, 1.80, 1.50, 1.25, 1.10,
                               1.00, 1.00, 1.00, 1.00, 2.20, 1.80, 1.60,
                               1.40, 1.35, 1.40, 1.40, 1.40, 1.35, 1.35,
                               1.35, 1.35, 1.30, 1.25, 1.15, 1.15, 1.15,
                               1.00};
  double v = (iz >= 1 && iz <= 36) ? r[iz - 1] : 1.00;
  return v * 1.8897;
}

/* mod_autogrid.f90:12-131 */
int orc_centercalc(int natoms, const int32_t *an, const double *cartes,
                   int nohydro, int betaone, double *cent, double *Ps,
                   double *sfalpha) {
  int nqmax = natoms * (natoms - 1) + 1;
  double *c1 = (double *)calloc((size_t)3 * nqmax, sizeof(double));
  double *ne = (double *)calloc((size_t)nqmax, sizeof(double));
  double *al = (double *)calloc((size_t)nqmax, sizeof(double));
  double tot = 0.0, wsum = 0.0;
  for (int i = 0; i < natoms; ++i) {
    double rj = orc_bs_radius(an[i]);
    double pairs = an[i] * (an[i] - 1.0) / 2.0;
    tot += pairs;
    wsum += pairs * rj;
  }
  ne[0] = tot;
  al[0] = (tot > 1e-15) ? wsum / tot : orc_bs_radius(1);
  int sm = 1;
#define CART(a, ax) cartes[(a) + natoms * (ax)]
  for (int j = 0; j < natoms - 1; ++j)
    for (int k = j + 1; k < natoms; ++k) {
      if (nohydro && !(an[j] > 1 && an[k] > 1)) continue;
      for (int ax = 0; ax < 3; ++ax) c1[3 * sm + ax] = CART(j, ax) - CART(k, ax);
      ne[sm] = an[j] * an[k] * 0.5;
      al[sm] = (orc_bs_radius(an[j]) + orc_bs_radius(an[k])) / 2.0;
      if (fabs(c1[3 * sm + 2]) < 0.5) c1[3 * sm + 2] = 0.0;
      int smp = sm++;
      for (int ax = 0; ax < 3; ++ax) c1[3 * sm + ax] = -c1[3 * smp + ax];
      ne[sm] = ne[smp];
      al[sm] = al[smp];
      ++sm;
    }
#undef CART
  nqmax = sm;
  /* merge, mod_autogrid.f90:83-99 (1-based i = 2..nqmax-1) */
  for (int i = 1; i < nqmax - 1; ++i) {
    double zs = fabs(c1[3 * i]) + fabs(c1[3 * i + 1]) + fabs(c1[3 * i + 2]);
    if (zs == 0.0) continue;
    for (int j = i + 1; j < nqmax; ++j) {
      int eq = fabs(c1[3 * i] - c1[3 * j]) < 0.1 &&
               fabs(c1[3 * i + 1] - c1[3 * j + 1]) < 0.1 &&
               fabs(c1[3 * i + 2] - c1[3 * j + 2]) < 0.1;
      if (eq) {
        al[i] = (al[i] * ne[i] + al[j] * ne[j]) / (ne[i] + ne[j]);
        ne[i] = ne[i] + ne[j];
        c1[3 * j] = c1[3 * j + 1] = c1[3 * j + 2] = 0.0;
        ne[j] = 0.0;
      }
    }
  }
  int nq = 0;
  for (int i = 0; i < nqmax; ++i) {
    double zs = fabs(c1[3 * i]) + fabs(c1[3 * i + 1]) + fabs(c1[3 * i + 2]);
    if (i == 0 || zs != 0.0) {
      for (int ax = 0; ax < 3; ++ax) cent[3 * nq + ax] = c1[3 * i + ax];
      Ps[nq] = ne[i];
      sfalpha[nq] = al[i];
      ++nq;
    }
  }
  if (betaone)
    for (int i = 0; i < nq; ++i) {
      sfalpha[i] = 1.0;
      Ps[i] = 1.0;
    }
  free(c1);
  free(ne);
  free(al);
  return nq;
}

/* mod_intracule.f90:576-649 */
void orc_smearing(int64_t G, const double *coords, const double *weights,
                  const double *density, double u_max, int n_bins,
                  double sigma, double *u, double *P, double *Iavg) {
  const double pi = 4.0 * atan(1.0);
  double bw = u_max / (double)(n_bins - 1);
  for (int j = 0; j < n_bins; ++j) {
    u[j] = (double)j * bw;
    P[j] = 0.0;
  }
  double pref = 1.0 / (sigma * sqrt(2.0 * pi));
  for (int64_t i = 0; i < G; ++i) {
    const double *c = coords + 3 * i;
    double ui = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    double val = weights[i] * density[i];
    int jmin = (int)((ui - 4.0 * sigma) / bw) + 1;
    int jmax = (int)((ui + 4.0 * sigma) / bw) + 1;
    if (jmin < 1) jmin = 1;
    if (jmax > n_bins) jmax = n_bins;
    for (int j = jmin; j <= jmax; ++j) {
      double dist = u[j - 1] - ui;
      double g = pref * exp(-(dist * dist) / (2.0 * sigma * sigma));
      P[j - 1] += val * g;
    }
  }
  for (int j = 0; j < n_bins; ++j)
    Iavg[j] = (u[j] > 1e-6) ? P[j] * (1.0 / (4.0 * pi * u[j] * u[j])) : 0.0;
}

/* The polynomial fits (mod_intrastuff.f90:127-175) of the records that pass
 * screen 1, as rows of their FP64 inputs: for every such record and every
 * axis of degree L >= 1, ORC_FIT_ROW doubles
 *   [record (0-based), axis, Ri, Rj, Rk, Rl, P_ik, P_jl, P, ti, tj, tk, tl,
 *    np = L + 1, sqe, invz, alfijkl]
 * (what polycoef receives at mod_intracule.f90:323-341).  Returns the number
 * of rows (rows may be NULL to count them), or -1 if cap is too small.
 * Test infrastructure for the roundoff audit of the fits. */
int64_t orc_fit_rows(int nprim, const double *alpha, const int32_t *tmn, const double *x,
                     const double *y, const double *z, int normalize, double thresh,
                     int64_t nrec, const int32_t *rec_ijkl, const double *rec_d, double *rows,
                     int64_t cap) {
  double *np_ = NULL;
  if (normalize) {
    np_ = (double *)malloc(sizeof(double) * nprim);
    for (int i = 0; i < nprim; ++i)
      np_[i] = orc_normprim(alpha[i], TMN(tmn, nprim, i, 0), TMN(tmn, nprim, i, 1),
                           TMN(tmn, nprim, i, 2));
  }
  const double lim = thresh / ((double)nprim * (nprim + 1) / 2.0);
  int64_t n = 0;
  for (int64_t r = 0; r < nrec; ++r) {
    const int32_t *q4 = rec_ijkl + 4 * r;
    int ok = q4[0] >= 1 && q4[0] <= nprim;
    for (int c = 1; c < 4; ++c) ok = ok && q4[c] >= 1 && q4[c] <= nprim;
    if (!ok) break;
    orc_quartet q;
    orc_prepare_quartet(nprim, alpha, tmn, x, y, z, np_, lim, q4[0], q4[1], q4[2], q4[3],
                        rec_d[r], &q);
    if (!q.pass1) continue;
    const int id[4] = {q4[0] - 1, q4[1] - 1, q4[2] - 1, q4[3] - 1};
    double X[4], Y[4], Z[4];
    for (int a = 0; a < 4; ++a) {
      X[a] = x[id[a]];
      Y[a] = y[id[a]];
      Z[a] = z[id[a]];
    }
    double *R[3] = {X, Y, Z};
    double sqe, invz, alf, P_ik[3], P_jl[3], P[3];
    quartet_geometry(alpha[id[0]], alpha[id[1]], alpha[id[2]], alpha[id[3]], R, &sqe, &invz,
                     &alf, P_ik, P_jl, P);
    for (int ax = 0; ax < 3; ++ax) {
      if (q.L[ax] < 1) continue;
      if (rows) {
        if (n >= cap) {
          free(np_);
          return -1;
        }
        double *w = rows + n * ORC_FIT_ROW;
        w[0] = (double)r;
        w[1] = ax;
        for (int a = 0; a < 4; ++a) w[2 + a] = R[ax][a];
        w[6] = P_ik[ax];
        w[7] = P_jl[ax];
        w[8] = P[ax];
        for (int a = 0; a < 4; ++a) w[9 + a] = TMN(tmn, nprim, id[a], ax);
        w[13] = q.L[ax] + 1;
        w[14] = sqe;
        w[15] = invz;
        w[16] = alf;
      }
      ++n;
    }
  }
  free(np_);
  return n;
}

/* orc_polycoef on every row of orc_fit_rows, in the current arithmetic
 * mode: C (n x 22, C[0] = highest power) and status (1: the reference would
 * zero the fit).  Returns the number of zeroed fits. */
int64_t orc_polycoef_rows(int64_t n, const double *rows, double *C, int32_t *status) {
  int64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int64_t i = 0; i < n; ++i) {
    const double *w = rows + i * ORC_FIT_ROW;
    double c[22] = {0.0};
    const int st = orc_polycoef(w[2], w[3], w[4], w[5], w[6], w[7], w[8], (int)w[9], (int)w[10],
                                (int)w[11], (int)w[12], (int)w[13], w[14], w[15], w[16], c);
    for (int k = 0; k < 22; ++k) C[i * 22 + k] = c[k];
    status[i] = st;
    bad += st;
  }
  return bad;
}

/* Census of the fits near the reference's |C| > 1e3 rule: the rows of
 * orc_fit_rows (records passing screen 1, axes of degree >= 1) whose fit is
 * zeroed, or has max |C| >= cmin, in arithmetic mode 0 or mode 1 of this
 * build.  Records are split over OpenMP threads in contiguous ranges and the
 * rows come out in record order.  Every record must hold valid indices.
 * Returns the number of rows written (at most cap), or -1 on a bad index
 * or allocation failure.  Test infrastructure for the roundoff audit. */
int64_t orc_fit_census(int nprim, const double *alpha, const int32_t *tmn, const double *x,
                       const double *y, const double *z, int normalize, double thresh,
                       int64_t nrec, const int32_t *rec_ijkl, const double *rec_d, double cmin,
                       double *rows, int64_t cap) {
  for (int64_t r = 0; r < 4 * nrec; ++r)
    if (rec_ijkl[r] < 1 || rec_ijkl[r] > nprim) return -1;
  double *np_ = NULL;
  if (normalize) {
    np_ = (double *)malloc(sizeof(double) * nprim);
    for (int i = 0; i < nprim; ++i)
      np_[i] = orc_normprim(alpha[i], TMN(tmn, nprim, i, 0), TMN(tmn, nprim, i, 1),
                           TMN(tmn, nprim, i, 2));
  }
  const double lim = thresh / ((double)nprim * (nprim + 1) / 2.0);
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  double **buf = (double **)calloc(nt, sizeof(double *));
  int64_t *cnt = (int64_t *)calloc(nt, sizeof(int64_t));
  int fail = 0;
#pragma omp parallel num_threads(nt)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    const int64_t r0 = nrec * t / nt, r1 = nrec * (t + 1) / nt;
    int64_t bcap = 1024;
    double *b = (double *)malloc(sizeof(double) * ORC_FIT_ROW * bcap);
    for (int64_t r = r0; r < r1 && b; ++r) {
      const int32_t *q4 = rec_ijkl + 4 * r;
      orc_quartet q;
      /* screen 1 in mode 0 (the rows' inputs do not depend on the mode) */
      t_arith = 0;
      orc_prepare_quartet(nprim, alpha, tmn, x, y, z, np_, lim, q4[0], q4[1], q4[2], q4[3],
                          rec_d[r], &q);
      if (!q.pass1) continue;
      const int id[4] = {q4[0] - 1, q4[1] - 1, q4[2] - 1, q4[3] - 1};
      double X[4], Y[4], Z[4];
      for (int a = 0; a < 4; ++a) {
        X[a] = x[id[a]];
        Y[a] = y[id[a]];
        Z[a] = z[id[a]];
      }
      double *R[3] = {X, Y, Z};
      double sqe, invz, alf, P_ik[3], P_jl[3], P[3];
      quartet_geometry(alpha[id[0]], alpha[id[1]], alpha[id[2]], alpha[id[3]], R, &sqe, &invz,
                       &alf, P_ik, P_jl, P);
      for (int ax = 0; ax < 3; ++ax) {
        if (q.L[ax] < 1) continue;
        int t4[4];
        for (int a = 0; a < 4; ++a) t4[a] = TMN(tmn, nprim, id[a], ax);
        int keep = 0;
        for (int mode = 0; mode < 2 && !keep; ++mode) {
          double c[22];
          t_arith = mode;
          const int st = orc_polycoef(R[ax][0], R[ax][1], R[ax][2], R[ax][3], P_ik[ax], P_jl[ax],
                                      P[ax], t4[0], t4[1], t4[2], t4[3], q.L[ax] + 1, sqe, invz,
                                      alf, c);
          double cm = 0.0;
          for (int k = 0; k <= q.L[ax]; ++k) cm = fmax(cm, fabs(c[k]));
          keep = st || cm >= cmin;
        }
        if (!keep) continue;
        if (cnt[t] == bcap) {
          bcap *= 2;
          double *nb = (double *)realloc(b, sizeof(double) * ORC_FIT_ROW * bcap);
          if (!nb) {
            free(b);
            b = NULL;
            break;
          }
          b = nb;
        }
        double *w = b + cnt[t] * ORC_FIT_ROW;
        w[0] = (double)r;
        w[1] = ax;
        for (int a = 0; a < 4; ++a) w[2 + a] = R[ax][a];
        w[6] = P_ik[ax];
        w[7] = P_jl[ax];
        w[8] = P[ax];
        for (int a = 0; a < 4; ++a) w[9 + a] = t4[a];
        w[13] = q.L[ax] + 1;
        w[14] = sqe;
        w[15] = invz;
        w[16] = alf;
        cnt[t]++;
      }
    }
    if (!b) {
#pragma omp atomic write
      fail = 1;
    }
    t_arith = -1;
    buf[t] = b;
  }
  int64_t n = 0;
  for (int t = 0; t < nt && !fail; ++t) {
    for (int64_t i = 0; i < cnt[t] && n < cap; ++i, ++n)
      if (rows) memcpy(rows + n * ORC_FIT_ROW, buf[t] + i * ORC_FIT_ROW, sizeof(double) * ORC_FIT_ROW);
  }
  for (int t = 0; t < nt; ++t) free(buf[t]);
  free(buf);
  free(cnt);
  free(np_);
  return fail ? -1 : n;
}

/* One-electron functions at explicit points (mod_functions.f90): Prim
 * (:17-31, an unnormalised Cartesian primitive at its centre), sd (:113-152,
 * its Laplacian), MoOr (:199-210, sum over |T| > 1e-10), Density (:312-329,
 * sum_i w_i MoOr_i^2, w_i = 2 closed shell), Lapl (:351-372, sum_i 2 w_i
 * MoOr_i sum_p T(i,p) sd_p).  kind 1 primitive `index`, 3 orbital `index`,
 * 5 density, 6 Laplacian (1-based index).  T(nmo, nprim) column-major. */
static double onee_prim(int p, const double *alpha, const int32_t *tmn, int nprim,
                        const double *cx, const double *cy, const double *cz, const double *r,
                        double *lap) {
  const double d[3] = {r[0] - cx[p], r[1] - cy[p], r[2] - cz[p]};
  const int t[3] = {TMN(tmn, nprim, p, 0), TMN(tmn, nprim, p, 1), TMN(tmn, nprim, p, 2)};
  const double a = alpha[p];
  const double e = exp(-a * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
  double f[3], f2[3];
  for (int ax = 0; ax < 3; ++ax) {
    const double x = d[ax];
    const int n = t[ax];
    f[ax] = ipow(x, n);
    /* d2/dx2 of x^n exp(-a x^2), without the exponential */
    double v = 4.0 * a * a * ipow(x, n + 2) - 2.0 * a * (2 * n + 1) * ipow(x, n);
    if (n >= 2) v += (double)(n * (n - 1)) * ipow(x, n - 2);
    f2[ax] = v;
  }
  if (lap)
    *lap = (f2[0] * f[1] * f[2] + f[0] * f2[1] * f[2] + f[0] * f[1] * f2[2]) * e;
  return f[0] * f[1] * f[2] * e;
}

void orc_onee(int kind, int index, int nprim, const double *alpha, const int32_t *tmn,
              const double *cx, const double *cy, const double *cz, int nmo, const double *T,
              const double *w, int64_t npts, const double *r, double *out) {
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < npts; ++g) {
    const double *rg = r + 3 * g;
    if (kind == 1) {
      out[g] = onee_prim(index - 1, alpha, tmn, nprim, cx, cy, cz, rg, NULL);
      continue;
    }
    double acc = 0.0;
    for (int i = (kind == 3 ? index - 1 : 0); i < (kind == 3 ? index : nmo); ++i) {
      double phi = 0.0, psi = 0.0;
      for (int p = 0; p < nprim; ++p) {
        double lp;
        const double v = onee_prim(p, alpha, tmn, nprim, cx, cy, cz, rg, &lp);
        const double c = T[i + (int64_t)nmo * p];
        if (fabs(c) > 1e-10) phi += c * v;
        psi += c * lp;
      }
      if (kind == 3) acc = phi;
      else if (kind == 5) acc += w[i] * phi * phi;
      else acc += 2.0 * w[i] * phi * psi;
    }
    out[g] = acc;
  }
}
