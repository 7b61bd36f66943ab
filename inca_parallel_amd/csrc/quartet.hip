// DM2 quartet preparation on the device.
//
// The reference handles each .dm2p record inside its OpenMP loop
// (mod_intracule.f90:276-344): normalise D, compute both Gaussian products
// and JA8 bounds, screen 1, the quartet prefactors, three polycoef fits
// (Gauss-Hermite sums + Vandermonde solve), screen 2.  Here:
//   k_pair_table   one thread per primitive pair (a,b): JA8, a+b, eR^2, P.
//                  Screen 1 then costs two table reads per record instead of
//                  two JA8 evaluations.
//   k_first_bad    the reader's stop rule (first record with a bad index).
//   k_screen1      one thread per record: screen 1 alone, a pass flag.
//   k_prep<NPM,S>  one thread per screen-1 survivor (compacted, file order):
//                  prefactors + polycoef + screen 2; writes the prepared
//                  record into the survivor's slot and a pass flag.
//                  S: records with every axis of degree <= 4, fits in
//                  registers; the others are deferred to a second launch
//                  over their compacted slots with the fit in scratch.
//   compaction     deterministic scan of the flags, then a gather of the
//                  surviving records in file order.
// Records are read straight from HBM as three 8-byte words per 24-byte
// record, so a wavefront streams 1536 contiguous bytes.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "common.hpp"

// The pre-pass follows the reference's operation sequence (separate multiply
// and add roundings, pow for integer powers) so that its coefficients agree
// with the FP64 restatement to the last bits where the Vandermonde systems
// are well conditioned and stay as close as roundoff allows where not.
#pragma clang fp contract(off)

namespace inca {
namespace {

constexpr int kT = 256;

// Gauss-Hermite nodes (roots of H_n, n = 1..11) and weights
// w = 2^(n-1) n! sqrt(pi) / (n^2 H_{n-1}(x)^2)  (mod_intrastuff.f90:198-341)
__constant__ double c_gh_x[11][11];
__constant__ double c_gh_w[11][11];
// w_m of the second screen (mod_intracule.f90:136-143)
__constant__ double c_wm[kNpMax];
// Gamma((k+1)/2) = int |y|^k exp(-y^2) dy over the real line
__constant__ double c_gmom[kNpMax];
// max over y of |y|^k exp(-y^2) = (k/2)^(k/2) exp(-k/2)  (1 for k = 0)
__constant__ double c_ymax[kNpMax];

const double kGhRoot[11][11] = {
    {0},
    {-0.70710678118654746, 0.70710678118654746},
    {-1.2247448713915889, 0, 1.2247448713915889},
    {-1.6506801238857847, -0.52464762327529035, 0.52464762327529035, 1.6506801238857847},
    {-2.0201828704560856, -0.95857246461381851, 0, 0.95857246461381851, 2.0201828704560856},
    {-2.3506049736744923, -1.335849074013697, -0.43607741192761651, 0.43607741192761651,
     1.335849074013697, 2.3506049736744923},
    {-2.6519613568352334, -1.6735516287674714, -0.81628788285896459, 0, 0.81628788285896459,
     1.6735516287674714, 2.6519613568352334},
    {-2.9306374202572441, -1.981656756695843, -1.1571937124467802, -0.38118699020732211,
     0.38118699020732211, 1.1571937124467802, 1.981656756695843, 2.9306374202572441},
    {-3.1909932017815277, -2.2665805845318432, -1.4685532892166679, -0.72355101875283756, 0,
     0.72355101875283756, 1.4685532892166679, 2.2665805845318432, 3.1909932017815277},
    {-3.4361591188377374, -2.5327316742327897, -1.7566836492998816, -1.0366108297895136,
     -0.34290132722370459, 0.34290132722370459, 1.0366108297895136, 1.7566836492998816,
     2.5327316742327897, 3.4361591188377374},
    {-3.6684708465595826, -2.7832900997816514, -2.0259480158257555, -1.3265570844949328,
     -0.65680956688209979, 0, 0.65680956688209979, 1.3265570844949328, 2.0259480158257555,
     2.7832900997816514, 3.6684708465595826}};

// Integer power x^e by e-1 successive multiplications.  The reference writes
// x**dble(e) (libm pow); the FP64 restatement uses this same sequence, so the
// Vandermonde systems the LU sees are bitwise those of the restatement.
__device__ __forceinline__ double ipow(double x, int e) {
  double r = 1.0;
  for (int k = 0; k < e; ++k) r = (k == 0) ? x : r * x;
  return r;
}

// ipow for exponents e <= E, as a loop of fixed length E with the same
// multiplications (no run-time trip count, so fully unrolled callers keep
// their values in registers)
template <int E>
__device__ __forceinline__ double ipow_upto(double x, int e) {
  double r = 1.0;
#pragma unroll
  for (int k = 0; k < E; ++k)
    if (k < e) r = (k == 0) ? x : r * x;
  return r;
}

// Largest Gauss-Hermite root for nn nodes: JA8's x_max (:44-113)
__device__ __forceinline__ double gh_xmax(int nn) {
  if (nn <= 1) return 0.0;
  if (nn > 11) nn = 11;
  return c_gh_x[nn - 1][nn - 1];
}

__global__ __launch_bounds__(kT) void k_pair_table(int nprim, const double *__restrict__ prim,
                                                   const int32_t *__restrict__ tmn,
                                                   PairEntry *__restrict__ pair) {
  int64_t id = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (id >= (int64_t)nprim * nprim) return;
  const int a = (int)(id / nprim), b = (int)(id % nprim);
  const double *al = prim, *X = prim + nprim, *Y = prim + 2 * nprim, *Z = prim + 3 * nprim;
  const double alfi = al[a], alfk = al[b];
  const double aik = alfi + alfk;
  const double eik = alfi * alfk * (1.0 / aik);
  const double dx = X[a] - X[b], dy = Y[a] - Y[b], dz = Z[a] - Z[b];
  const double rik2 = dx * dx + dy * dy + dz * dz;
  // JA8 (mod_intrastuff.f90:115-122)
  const int ti = tmn[a], mi = tmn[nprim + a], ni = tmn[2 * nprim + a];
  const int tk = tmn[b], mk = tmn[nprim + b], nk = tmn[2 * nprim + b];
  const double s = sqrt(2.0 / aik);
  const double x0 = gh_xmax(ti + tk + 1), x1 = gh_xmax(mi + mk + 1), x2 = gh_xmax(ni + nk + 1);
  const double L = (double)(ti + tk + mi + mk + ni + nk);
  const double adx = fabs(dx), ady = fabs(dy), adz = fabs(dz);
  // pi^(3/2) (2 aik)^-(L+3/2) as a literal and a division by an integer
  // power times a square root (no pow with a fractional exponent)
  const double two_a = 2.0 * aik;
  double J = 5.568327996831708 * (1.0 / (ipow(two_a, (int)L + 1) * sqrt(two_a))) *
             exp(-2.0 * eik * rik2);
  J *= ipow(x0 + alfk * s * adx, 2 * ti) * ipow(x0 + alfi * s * adx, 2 * tk);
  J *= ipow(x1 + alfk * s * ady, 2 * mi) * ipow(x1 + alfi * s * ady, 2 * mk);
  J *= ipow(x2 + alfk * s * adz, 2 * ni) * ipow(x2 + alfi * s * adz, 2 * nk);
  PairEntry e;
  e.J = J;
  e.aik = aik;
  e.eR2 = eik * rik2;
  e.P[0] = (alfi * X[a] + alfk * X[b]) * (1.0 / aik);
  e.P[1] = (alfi * Y[a] + alfk * Y[b]) * (1.0 / aik);
  e.P[2] = (alfi * Z[a] + alfk * Z[b]) * (1.0 / aik);
  e.pad[0] = e.pad[1] = 0.0;
  pair[id] = e;
}

__global__ __launch_bounds__(kT) void k_first_bad(const int32_t *__restrict__ rec, int64_t n,
                                                  int nprim,
                                                  unsigned long long *__restrict__ first) {
  int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= n) return;
  const int32_t *p = rec + 6 * r;
  int i = p[0], j = p[1], k = p[2], l = p[3];
  bool bad = i < 1 || i > nprim || j < 1 || k < 1 || l < 1 || j > nprim || k > nprim || l > nprim;
  if (bad) atomicMin(first, (unsigned long long)r);
}

// polycoef for one axis (mod_intrastuff.f90:127-194).  Returns 1 when the
// reference would zero the coefficients (INFO != 0 or any |C| > 1e3).
template <int NPM>
__device__ int polycoef(const double R[4], double Rik, double Rjl, double Rijkl, const int t[4],
                        int np, double sqe, double invz, double alf, double *C) {
  const int L = np - 1;
  const int n = L / 2 + 1;  // Gauss-Hermite nodes, exact for degree 2n-1 >= L
  if (np == 1) {
    C[0] = c_gh_w[0][0];
    return 0;
  }
  double M[NPM * NPM];
  for (int i = 0; i < np; ++i) {
    const double xx = -0.5 + (double)(i + 1);
    const double y = sqe * (xx + Rik - Rjl);
    // M(i, j) = y^(L - j)  (mod_intrastuff.f90:151-157)
    for (int j = 0; j <= L; ++j) M[i + np * j] = ipow(y, L - j);
    const double am = (alf - 0.5) * xx, ap = (alf + 0.5) * xx;
    double c = 0.0;
    for (int k = 0; k < n; ++k) {
      const double a = invz * c_gh_x[n - 1][k];
      const double w = ipow(a + am + (Rijkl - R[0]), t[0]) * ipow(a + ap + (Rijkl - R[1]), t[1]) *
                       ipow(a + am + (Rijkl - R[2]), t[2]) * ipow(a + ap + (Rijkl - R[3]), t[3]);
      c += c_gh_w[n - 1][k] * w;
    }
    C[i] = c;
  }
  // partial-pivot LU + solve in reference LAPACK order (dgetrf2 / dgetrs)
  int piv[NPM];
  int info = 0;
  for (int k = 0; k < np; ++k) {
    int p = k;
    double amax = fabs(M[k + np * k]);
    for (int i = k + 1; i < np; ++i) {
      double v = fabs(M[i + np * k]);
      if (v > amax) {
        amax = v;
        p = i;
      }
    }
    piv[k] = p;
    if (M[p + np * k] != 0.0) {
      if (p != k)
        for (int j = 0; j < np; ++j) {
          double tt = M[k + np * j];
          M[k + np * j] = M[p + np * j];
          M[p + np * j] = tt;
        }
      const double pv = M[k + np * k];
      if (fabs(pv) >= 2.2250738585072014e-308) {
        const double rp = 1.0 / pv;
        for (int i = k + 1; i < np; ++i) M[i + np * k] *= rp;
      } else {
        for (int i = k + 1; i < np; ++i) M[i + np * k] /= pv;
      }
    } else if (info == 0) {
      info = k + 1;
    }
    for (int j = k + 1; j < np; ++j) {
      const double u = M[k + np * j];
      for (int i = k + 1; i < np; ++i) M[i + np * j] -= M[i + np * k] * u;
    }
  }
  int bad = (info != 0);
  if (!bad) {
    for (int k = 0; k < np; ++k)
      if (piv[k] != k) {
        double tt = C[k];
        C[k] = C[piv[k]];
        C[piv[k]] = tt;
      }
    for (int k = 0; k < np; ++k)
      if (C[k] != 0.0)
        for (int i = k + 1; i < np; ++i) C[i] -= C[k] * M[i + np * k];
    for (int k = np - 1; k >= 0; --k)
      if (C[k] != 0.0) {
        C[k] /= M[k + np * k];
        for (int i = 0; i < k; ++i) C[i] -= C[k] * M[i + np * k];
      }
    for (int i = 0; i < np; ++i)
      if (!(fabs(C[i]) <= 1.0e3)) bad = 1;
  }
  if (bad)
    for (int i = 0; i < np; ++i) C[i] = 0.0;
  return bad;
}

// Exact coefficients (descending powers of y = sqe (r + d), d = P_ik - P_jl)
// of the polynomial polycoef fits: p(r) = sum_n w_n prod_a (invz t_n + s_a r +
// (Rijkl - R_a))^{t_a}, s = alf -+ 1/2 (mod_intrastuff.f90:178-193); each
// factor b + s r is expanded as (b - s d) + (s/sqe) y.  Same polynomial as
// the fit in exact arithmetic, without the Vandermonde solve.
template <int NPM>
__device__ void exact_poly(const double R[4], double Rijkl, const int t[4], int L, double invz,
                           double alf, double sqe, double d, double *E) {
  const int n = L / 2 + 1;
  double acc[NPM];
  for (int m = 0; m <= L; ++m) acc[m] = 0.0;
  for (int k = 0; k < n; ++k) {
    double poly[NPM];
    poly[0] = 1.0;
    int deg = 0;
    const double a = invz * c_gh_x[n - 1][k];
    for (int f = 0; f < 4; ++f) {
      const double s0 = (f == 0 || f == 2) ? alf - 0.5 : alf + 0.5;
      const double b = (a + (Rijkl - R[f])) - s0 * d;
      const double sl = s0 / sqe;
      for (int e = 0; e < t[f]; ++e) {
        poly[deg + 1] = poly[deg] * sl;
        for (int m = deg; m >= 1; --m) poly[m] = poly[m] * b + poly[m - 1] * sl;
        poly[0] *= b;
        ++deg;
      }
    }
    for (int m = 0; m <= deg; ++m) acc[m] += c_gh_w[n - 1][k] * poly[m];
  }
  for (int q = 0; q <= L; ++q) E[q] = acc[L - q];
}

// polycoef for an axis of degree L <= NP - 1 as an NP x NP system with
// compile-time indices, so that the matrix stays in registers.  The
// np = L + 1 equations sit in the leading block; the rest of the matrix is
// the identity and the rest of the right-hand side zero.  Partial pivoting
// never picks a padding row for a leading column (their entries there are
// zero and the search takes strictly larger values), the eliminations
// subtract exact zeros from the padding, and the solves skip the zero
// unknowns, so the leading coefficients come out of the same operations as
// in polycoef<NPM> on the np x np system (np = 1 included: the single
// Gauss-Hermite weight divided by one).  The pivot row is a per-lane value,
// so row exchanges are selects over the rows below.
template <int NP>
__device__ __forceinline__ int polycoef_pad(const double R[4], double Rik, double Rjl,
                                            double Rijkl, const int t[4], int L, double sqe,
                                            double invz, double alf, double (&C)[NP]) {
  constexpr int E = NP - 1;         // largest degree, and largest t[f]
  constexpr int NN = E / 2 + 1;     // most Gauss-Hermite nodes
  const int np = L + 1;
  const int n = L / 2 + 1;
  double M[NP][NP];  // M[i][j] = y_i^(L - j) in the leading block
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const double xx = -0.5 + (double)(i + 1);
    const double y = sqe * (xx + Rik - Rjl);
#pragma unroll
    for (int j = 0; j < NP; ++j)
      M[i][j] = (i < np && j < np) ? ipow_upto<E>(y, L - j) : (i == j ? 1.0 : 0.0);
    const double am = (alf - 0.5) * xx, ap = (alf + 0.5) * xx;
    double c = 0.0;
#pragma unroll
    for (int k = 0; k < NN; ++k) {
      if (k < n) {
        const double a = invz * c_gh_x[n - 1][k];
        const double w = ipow_upto<E>(a + am + (Rijkl - R[0]), t[0]) *
                         ipow_upto<E>(a + ap + (Rijkl - R[1]), t[1]) *
                         ipow_upto<E>(a + am + (Rijkl - R[2]), t[2]) *
                         ipow_upto<E>(a + ap + (Rijkl - R[3]), t[3]);
        c += c_gh_w[n - 1][k] * w;
      }
    }
    C[i] = (i < np) ? c : 0.0;
  }
  int piv[NP];
  int info = 0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    int p = k;
    double amax = fabs(M[k][k]);
#pragma unroll
    for (int i = k + 1; i < NP; ++i) {
      const double v = fabs(M[i][k]);
      if (v > amax) {
        amax = v;
        p = i;
      }
    }
    piv[k] = p;
    // Exchange rows k and p.  A zero pivot column leaves p = k (the search
    // takes strictly larger values), so the exchange is unconditional.
#pragma unroll
    for (int j = 0; j < NP; ++j) {
      double r = M[k][j];
#pragma unroll
      for (int i = k + 1; i < NP; ++i) r = (p == i) ? M[i][j] : r;
      const double a = M[k][j];
#pragma unroll
      for (int i = k + 1; i < NP; ++i) M[i][j] = (p == i) ? a : M[i][j];
      M[k][j] = r;
    }
    const double pv = M[k][k];
    if (pv != 0.0) {
      if (fabs(pv) >= 2.2250738585072014e-308) {
        const double rp = 1.0 / pv;
#pragma unroll
        for (int i = k + 1; i < NP; ++i) M[i][k] *= rp;
      } else {
#pragma unroll
        for (int i = k + 1; i < NP; ++i) M[i][k] /= pv;
      }
    } else if (info == 0) {
      info = k + 1;
    }
#pragma unroll
    for (int j = k + 1; j < NP; ++j) {
      const double u = M[k][j];
#pragma unroll
      for (int i = k + 1; i < NP; ++i) M[i][j] -= M[i][k] * u;
    }
  }
  int bad = (info != 0);
  if (!bad) {
#pragma unroll
    for (int k = 0; k < NP; ++k)
#pragma unroll
      for (int i = k + 1; i < NP; ++i) {
        const bool sw = (piv[k] == i);
        const double a = C[k], b = C[i];
        C[k] = sw ? b : a;
        C[i] = sw ? a : b;
      }
#pragma unroll
    for (int k = 0; k < NP; ++k)
      if (C[k] != 0.0)
#pragma unroll
        for (int i = k + 1; i < NP; ++i) C[i] -= C[k] * M[i][k];
#pragma unroll
    for (int k = NP - 1; k >= 0; --k)
      if (C[k] != 0.0) {
        C[k] /= M[k][k];
#pragma unroll
        for (int i = 0; i < k; ++i) C[i] -= C[k] * M[i][k];
      }
#pragma unroll
    for (int i = 0; i < NP; ++i)
      if (!(fabs(C[i]) <= 1.0e3)) bad = 1;
  }
  if (bad)
#pragma unroll
    for (int i = 0; i < NP; ++i) C[i] = 0.0;
  return bad;
}

// polycoef_pad at size K (K - 1 >= L), its coefficients copied into the
// first K entries of C (the callers read C[0..L] only).
template <int K, int NP>
__device__ __forceinline__ int fit_into(const double R[4], double Rik, double Rjl, double Pijkl,
                                        const int t[4], int L, double sqe, double invz,
                                        double alf, double (&C)[NP]) {
  double c[K];
  const int bad = polycoef_pad<K>(R, Rik, Rjl, Pijkl, t, L, sqe, invz, alf, c);
#pragma unroll
  for (int q = 0; q < K; ++q) C[q] = c[q];
  return bad;
}

// The fit of an axis of degree L <= NP - 1 at size L + 1 up to 5, padded to
// NP beyond.
template <int NP>
__device__ __forceinline__ int fit_sized(const double R[4], double Rik, double Rjl, double Pijkl,
                                         const int t[4], int L, double sqe, double invz,
                                         double alf, double (&C)[NP]) {
  if constexpr (NP >= 2)
    if (L == 0) return fit_into<1, NP>(R, Rik, Rjl, Pijkl, t, L, sqe, invz, alf, C);
  if constexpr (NP >= 3)
    if (L == 1) return fit_into<2, NP>(R, Rik, Rjl, Pijkl, t, L, sqe, invz, alf, C);
  if constexpr (NP >= 4)
    if (L == 2) return fit_into<3, NP>(R, Rik, Rjl, Pijkl, t, L, sqe, invz, alf, C);
  if constexpr (NP >= 5)
    if (L == 3) return fit_into<4, NP>(R, Rik, Rjl, Pijkl, t, L, sqe, invz, alf, C);
  if constexpr (NP >= 6)
    if (L == 4) return fit_into<5, NP>(R, Rik, Rjl, Pijkl, t, L, sqe, invz, alf, C);
  return polycoef_pad<NP>(R, Rik, Rjl, Pijkl, t, L, sqe, invz, alf, C);
}

// One axis of degree L <= NP - 1 through polycoef_pad: second-screen factor,
// L1 bound and integral factors, and the coefficients written to out[0..L]
// (fitted polynomials only; the exact-polynomial mode takes polycoef<NPM>).
template <int NP>
__device__ __forceinline__ int axis_pad(const double R[4], double Rik, double Rjl, double Pijkl,
                                        const int t[4], int L, double sqe, double invz,
                                        double alf, double scale, double &U, double &B,
                                        double &M, double *__restrict__ out) {
  // the fit at the axis's own size, np = L + 1: the padded solve performs
  // the np x np solve's operations, so every size gives the same bits; the
  // pre-pass hands out records sorted by degrees, so the lanes of a wave
  // mostly take one branch
  double C[NP];
  const int bad = fit_sized<NP>(R, Rik, Rjl, Pijkl, t, L, sqe, invz, alf, C);
  double u = 0.0;
#pragma unroll
  for (int q = 0; q < NP; ++q)
    if (q <= L) u += C[q] * c_wm[q];
  U *= u;
  double bax = 0.0, max_ = 0.0;
#pragma unroll
  for (int q = 0; q < NP; ++q)
    if (q <= L) {
      const double g = c_gmom[L - q];
      bax += fabs(C[q]) * g;
      if (((L - q) & 1) == 0) max_ += C[q] * g;
    }
  B *= bax * (1.0 / sqe);
  M *= max_ * (1.0 / sqe);
#pragma unroll
  for (int q = 0; q < NP; ++q)
    if (q <= L) out[q] = scale * C[q];
  return bad;
}

// Radius in rp beyond which the term A exp(-|rp|^2) Vx Vy Vz stays below
// 2^-cut_log2 of its peak bound |A| prod_axis sum_k |c_k| ymax(k): with
// |rp_axis| <= |rp| = sqrt(s), |term| <= |A| e^-s prod_axis H_axis(sqrt(s)),
// H(t) = sum_k |c_k| t^k, so the smallest s with
//   F(s) = sum_axis log(H_axis(sqrt s) / P_axis) - s + s0 <= 0,
// s0 = cut_log2 ln 2, is wanted.  d/ds log H(sqrt s) <= L / (2 s), so F
// falls at least at rate 1 - Lsum/(2s) > 0 for s >= s0 > Lsum/2, which makes
// the fixed-point iterate plus one Newton-type correction a safe bound.
// c holds the three axis polynomials back to back (highest power first,
// the x axis scaled by A, which the ratios do not see).  s0 <= 0: no cutoff.
__device__ double cutoff_radius(const double *c, const int Ls[3], double s0) {
  if (!(s0 > 0.0)) return HUGE_VAL;
  double Pax[3];
  int off = 0;
  for (int ax = 0; ax < 3; ++ax) {
    const int L = Ls[ax];
    double p = 0.0;
    for (int q = 0; q <= L; ++q) p += fabs(c[off + q]) * c_ymax[L - q];
    if (!(p > 0.0)) return 0.0;  // identically zero term
    Pax[ax] = p;
    off += L + 1;
  }
  const int Lsum = Ls[0] + Ls[1] + Ls[2];
  // three fixed-point steps from s0; the correction after them makes the
  // bound hold whatever the iterate (F is decreasing beyond s0 > Lsum/2)
  double s = s0, F = 0.0;
  for (int it = 0; it < 3; ++it) {
    const double t = sqrt(s);
    double ratio = 1.0;
    off = 0;
    for (int ax = 0; ax < 3; ++ax) {
      const int L = Ls[ax];
      double h = 0.0;
      for (int q = 0; q <= L; ++q) h = h * t + fabs(c[off + q]);
      ratio *= h / Pax[ax];
      off += L + 1;
    }
    F = log(ratio) - s + s0;
    if (it < 2) s = s + F;
  }
  if (F > 0.0) s += F / (1.0 - Lsum / (2.0 * s));
  const double r = sqrt(s) * (1.0 + 1e-12);
  return (r >= 0.0) ? r : HUGE_VAL;
}

// Screen 1 alone (the expression of k_prep, mod_intracule.f90:297-305):
// pass[n] = 1 for the records that pass, so that k_prep can run over the
// compacted survivors (about half of the records) with no lane idle on a
// failed one.
__global__ __launch_bounds__(kT) void k_screen1(const int32_t *__restrict__ rec, int64_t n,
                                                int nprim, const double *__restrict__ prim,
                                                const PairEntry *__restrict__ pair, double lim,
                                                int normalize, uint8_t *__restrict__ pass) {
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (g >= n) return;
  const int32_t *p = rec + 6 * g;
  const int i = p[0] - 1, j = p[1] - 1, k = p[2] - 1, l = p[3] - 1;
  double D;
  __builtin_memcpy(&D, p + 4, sizeof(double));
  const double *nrm = prim + 4 * nprim;
  double DMval = D;
  if (normalize) DMval = nrm[i] * nrm[j] * nrm[k] * nrm[l] * D;
  const double Jik = pair[(int64_t)i * nprim + k].J, Jjl = pair[(int64_t)j * nprim + l].J;
  pass[g] = (uint8_t)(fabs(DMval) * sqrt(Jik * Jjl) >= lim);
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

// The group words of record p (1-based primitive indices i, j, k, l in
// words 0-3): sqe and sqe (P_ik - P_jl), by the operations k_prep stores
// them with, so that equal inputs give the bits of k_prep's row words.
__device__ __forceinline__ void group_words(const int32_t *p, int nprim,
                                            const PairEntry *__restrict__ pair, double w[4]) {
  const int i = p[0] - 1, j = p[1] - 1, k = p[2] - 1, l = p[3] - 1;
  const PairEntry pik = pair[(int64_t)i * nprim + k];
  const PairEntry pjl = pair[(int64_t)j * nprim + l];
  const double aik = pik.aik, ajl = pjl.aik;
  const double zeta = aik + ajl;
  const double eijkl = (aik * ajl) * (1.0 / zeta);
  const double sqe = sqrt(eijkl);
  w[0] = sqe;
  w[1] = sqe * (pik.P[0] - pjl.P[0]);
  w[2] = sqe * (pik.P[1] - pjl.P[1]);
  w[3] = sqe * (pik.P[2] - pjl.P[2]);
}

// Group keys of the screen-1 survivors s < n1 (record idx[s]): a 32-bit
// hash of the bits of sqe and sqe (P_ik - P_jl), computed by k_prep's own
// operations, so that the records of a shell-quartet group (equal words)
// share a key; val[s] = idx[s] (increasing in s) for the stable sort.
__global__ __launch_bounds__(kT) void k_slot_keys(const int32_t *__restrict__ rec, int64_t n1,
                                                  const int64_t *__restrict__ idx, int nprim,
                                                  const PairEntry *__restrict__ pair,
                                                  uint32_t *__restrict__ key,
                                                  uint32_t *__restrict__ val) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s >= n1) return;
  const int64_t r = idx[s];
  double w[4];
  group_words(rec + 6 * r, nprim, pair, w);
  uint64_t h = 0x9e3779b97f4a7c15ull;
#pragma unroll
  for (int q = 0; q < 4; ++q) h = mix64(h ^ (uint64_t)__double_as_longlong(w[q]));
  key[s] = (uint32_t)(h >> 32);
  val[s] = (uint32_t)r;
}

// Slot j after the sort holds record order[j]: its 24 bytes are copied to
// out[j], so that the fits read their records in slot order, and head[j] = 1
// where slot j starts a run of equal group words (slot j - 1 has other
// words): the group boundaries of the rows k_prep writes for these slots,
// found here instead of from the rows.  The previous slot's words come from
// the neighbouring thread through LDS, or, for a block's first thread, from
// its own read of that record.
__global__ __launch_bounds__(kT) void k_slot_records(const uint32_t *__restrict__ order,
                                                     int64_t n1, const int32_t *__restrict__ rec,
                                                     int nprim,
                                                     const PairEntry *__restrict__ pair,
                                                     int32_t *__restrict__ out,
                                                     uint8_t *__restrict__ head) {
  __shared__ double wl[4][kT];
  const int t = threadIdx.x;
  const int64_t j = (int64_t)blockIdx.x * kT + t;
  double w[4] = {0.0, 0.0, 0.0, 0.0};
  if (j < n1) {
    const int32_t *src = rec + 6 * (int64_t)order[j];
    const int64_t *s8 = reinterpret_cast<const int64_t *>(src);
    int64_t *dst = reinterpret_cast<int64_t *>(out + 6 * j);
    dst[0] = s8[0];
    dst[1] = s8[1];
    dst[2] = s8[2];
    group_words(src, nprim, pair, w);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) wl[q][t] = w[q];
  __syncthreads();
  if (j >= n1) return;
  bool h = j == 0;
  if (!h) {
    double b[4];
    if (t > 0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) b[q] = wl[q][t - 1];
    } else {
      group_words(rec + 6 * (int64_t)order[j - 1], nprim, pair, b);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) h = h || __double_as_longlong(w[q]) != __double_as_longlong(b[q]);
  }
  head[j] = (uint8_t)h;
}

// Fit order: key[s] = Lx | Ly << 4 | Lz << 8 | max(Lx, Ly, Lz) << 12 of
// slot s (record rec + 6 rmap[s], or s when rmap is null), val[s] = s, for
// a stable sort that gives each largest degree its own launch and lets the
// lanes of a wave fit axes of one degree.
__global__ __launch_bounds__(kT) void k_slot_class(const int32_t *__restrict__ rec, int64_t n1,
                                                   const int64_t *__restrict__ rmap,
                                                   const int32_t *__restrict__ tmn, int nprim,
                                                   uint32_t *__restrict__ key,
                                                   uint32_t *__restrict__ val) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s >= n1) return;
  const int32_t *p = rec + 6 * (rmap ? rmap[s] : s);
  const int i = p[0] - 1, j = p[1] - 1, k = p[2] - 1, l = p[3] - 1;
  uint32_t c = 0, lmax = 0;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int32_t *tt = tmn + ax * nprim;
    const uint32_t L = (uint32_t)min(tt[i] + tt[j] + tt[k] + tt[l], 15);
    c |= L << (4 * ax);
    lmax = max(lmax, L);
  }
  key[s] = c | (lmax << 12);
  val[s] = (uint32_t)s;
}

// bnd[v] = first position of the sorted keys whose largest degree is v
// (slots of degree v: bnd[v] .. the next present start; bnd filled with n1
// beforehand).
__global__ __launch_bounds__(kT) void k_degree_bounds(const uint32_t *__restrict__ key,
                                                      int64_t n1, int64_t *__restrict__ bnd) {
  const int64_t j = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (j >= n1) return;
  const uint32_t v = key[j] >> 12;
  if (j == 0 || (key[j - 1] >> 12) != v) bnd[v] = j;
}

__global__ void k_fill_i64(int64_t *__restrict__ a, int n, int64_t v) {
  if ((int)threadIdx.x < n) a[threadIdx.x] = v;
}

__global__ __launch_bounds__(kT) void k_u32_to_i64(const uint32_t *__restrict__ a, int64_t n,
                                                   int64_t *__restrict__ b) {
  const int64_t s = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (s < n) b[s] = a[s];
}

// One thread per record; the output record layout is described at kQHead.
//
// Slots: the lanes take the slots idx[0..nrun) (slots 0..nrun when idx is
// null); slot s holds record rmap[s] (record s when rmap is null), and its
// row, flag and deferral flag are written at s.  SMALL: the records
// whose three axes all have degree below NPM, each axis by the padded
// compile-time-size fit (axis_pad), so nothing is indexed at run time and
// the matrices stay in registers (fitted polynomials only); every other
// record is marked in `defer` and left alone.  Otherwise each axis by the
// run-time-size fit, whose matrix lives in scratch.
template <int NPM, bool SMALL>
__global__ __launch_bounds__(kT) void k_prep(
    const int32_t *__restrict__ rec, int64_t nrun, const int64_t *__restrict__ idx,
    const int64_t *__restrict__ rmap, int nprim,
    const double *__restrict__ prim, const int32_t *__restrict__ tmn,
    const PairEntry *__restrict__ pair, double lim, int normalize, int qstride, int exact,
    double cut_s0, double *__restrict__ qout, uint8_t *__restrict__ flag,
    uint8_t *__restrict__ defer, unsigned long long *__restrict__ cnt,
    double *__restrict__ msum) {
  static_assert(!SMALL || NPM <= 9, "the compile-time fits cover degrees 0-8");
  const int64_t g = (int64_t)blockIdx.x * kT + threadIdx.x;
  int pass1 = 0, pass2 = 0, nfail = 0;
  double mom = 0.0;  // signed whole-space integral of a surviving term
  bool run = g < nrun;
  int64_t n = g;
  if (run && idx) n = idx[g];
  int64_t r = n;
  if (run && rmap) r = rmap[n];
  const int32_t *p = rec + 6 * r;
  int i = 0, j = 0, k = 0, l = 0;
  if (run) {
    i = p[0] - 1, j = p[1] - 1, k = p[2] - 1, l = p[3] - 1;
    if (SMALL) {
      int lmax = 0;
      for (int ax = 0; ax < 3; ++ax) {
        const int *tt = tmn + ax * nprim;
        lmax = max(lmax, tt[i] + tt[j] + tt[k] + tt[l]);
      }
      const bool big = lmax >= NPM;
      if (defer) defer[n] = (uint8_t)big;
      run = !big;
    }
  }
  if (run) {
    double D;
    __builtin_memcpy(&D, p + 4, sizeof(double));
    const double *nrm = prim + 4 * nprim;
    double DMval = D;
    if (normalize) DMval = nrm[i] * nrm[j] * nrm[k] * nrm[l] * D;
    const PairEntry pik = pair[(int64_t)i * nprim + k];
    const PairEntry pjl = pair[(int64_t)j * nprim + l];
    const double screen1 = fabs(DMval) * sqrt(pik.J * pjl.J);
    if (screen1 >= lim) {
      pass1 = 1;
      const double aik = pik.aik, ajl = pjl.aik;
      const double zeta = aik + ajl;
      const double eijkl = (aik * ajl) * (1.0 / zeta);
      const double sqe = sqrt(eijkl);
      const double invz = 1.0 / sqrt(zeta);
      const double alf = 0.5 * (aik - ajl) * (1.0 / zeta);
      const double A = DMval * (1.0 / (zeta * sqrt(zeta))) * exp(-pik.eR2 - pjl.eR2);
      double *out = qout + n * (int64_t)qstride;
      int off = kQHead;
      int Ls[3];
      double U = A;
      // B: bound on the L1 norm of the term, M: its integral over all space
      // (per axis, int q(y) exp(-y^2) dr with dr = dy / sqe)
      double B = fabs(A), M = A;
      double C[NPM];
#pragma unroll 1
      for (int ax = 0; ax < 3; ++ax) {
        const double *Rx = prim + (1 + ax) * nprim;
        const double R[4] = {Rx[i], Rx[j], Rx[k], Rx[l]};
        const int *tt = tmn + ax * nprim;
        const int t[4] = {tt[i], tt[j], tt[k], tt[l]};
        const int L = t[0] + t[1] + t[2] + t[3];
        Ls[ax] = L;
        const double Pijkl = (aik * pik.P[ax] + ajl * pjl.P[ax]) * (1.0 / zeta);
        if (SMALL) {
          const double scale = (ax == 0) ? A : 1.0;
          const double Rik = pik.P[ax], Rjl = pjl.P[ax];
          nfail += axis_pad<NPM>(R, Rik, Rjl, Pijkl, t, L, sqe, invz, alf, scale, U, B, M,
                                 out + off);
          off += L + 1;
          continue;
        }
        nfail += polycoef<NPM>(R, pik.P[ax], pjl.P[ax], Pijkl, t, L + 1, sqe, invz, alf, C);
        double u = 0.0;
        for (int q = 0; q <= L; ++q) u += C[q] * c_wm[q];
        U *= u;
        if (exact) exact_poly<NPM>(R, Pijkl, t, L, invz, alf, sqe, pik.P[ax] - pjl.P[ax], C);
        double bax = 0.0, max_ = 0.0;
        for (int q = 0; q <= L; ++q) {
          const double g = c_gmom[L - q];
          bax += fabs(C[q]) * g;
          if (((L - q) & 1) == 0) max_ += C[q] * g;
        }
        B *= bax * (1.0 / sqe);
        M *= max_ * (1.0 / sqe);
        const double scale = (ax == 0) ? A : 1.0;
        for (int q = 0; q <= L; ++q) out[off + q] = scale * C[q];
        off += L + 1;
      }
      if (fabs(U) >= lim) {
        pass2 = 1;
        out[0] = sqe;
        out[1] = sqe * (pik.P[0] - pjl.P[0]);
        out[2] = sqe * (pik.P[1] - pjl.P[1]);
        out[3] = sqe * (pik.P[2] - pjl.P[2]);
        // class word: L per axis in bits 0-23, the axes of nonzero degree
        // (the evaluation's case) in bits 24-26
        const int axes = (Ls[0] > 0) | ((Ls[1] > 0) << 1) | ((Ls[2] > 0) << 2);
        out[4] = __longlong_as_double(
            (long long)(Ls[0] | (Ls[1] << 8) | (Ls[2] << 16) | (axes << 24)));
        out[5] = B;
        out[6] = cutoff_radius(out + kQHead, Ls, cut_s0);
        out[7] = 0.0;
        mom = M;
      } else {
        // dropped by screen 2: the group words and the mark only
        out[0] = sqe;
        out[1] = sqe * (pik.P[0] - pjl.P[0]);
        out[2] = sqe * (pik.P[1] - pjl.P[1]);
        out[3] = sqe * (pik.P[2] - pjl.P[2]);
        out[4] = __longlong_as_double((long long)kQDropped);
        out[5] = 0.0;
        out[6] = 0.0;
        out[7] = 0.0;
      }
    }
    flag[n] = (uint8_t)pass2;
  }
  // wave-aggregated counters
  unsigned long long b1 = __ballot(pass1);
  int wsum = nfail;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) wsum += __shfl_xor(wsum, d, 64);
  if (__lane_id() == 0) {
    if (b1) atomicAdd(&cnt[1], (unsigned long long)__popcll(b1));
    if (wsum) atomicAdd(&cnt[3], (unsigned long long)wsum);
  }
  if (msum) {  // fixed-order block sum of the term integrals
    __shared__ double wred[kT / 64];
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) mom += __shfl_down(mom, d, 64);
    if (__lane_id() == 0) wred[threadIdx.x / 64] = mom;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int w = 0; w < kT / 64; ++w) s += wred[w];
      msum[blockIdx.x] = s;
    }
  }
}

// Sum of n block partials by one block of kSumT threads in a fixed order:
// thread t keeps four running sums, of the partials i = t + kSumT (4 k + u)
// for u = 0..3 (independent chains, so that several loads are in flight),
// adds them as (s0 + s1) + (s2 + s3), and a tree over the threads follows.
constexpr int kSumT = 1024;
__global__ __launch_bounds__(kSumT) void k_sum_fixed(const double *__restrict__ x, int64_t n,
                                                     double *__restrict__ out) {
  __shared__ double red[kSumT];
  const int t = threadIdx.x;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t i = t; i < n; i += 4 * (int64_t)kSumT) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t k = i + (int64_t)u * kSumT;
      if (k < n) s[u] += x[k];
    }
  }
  red[t] = (s[0] + s[1]) + (s[2] + s[3]);
  __syncthreads();
  for (int h = kSumT / 2; h > 0; h >>= 1) {
    if (t < h) red[t] += red[t + h];
    __syncthreads();
  }
  if (t == 0) out[0] = red[0];
}

__global__ __launch_bounds__(kT) void k_gather_records(const double *__restrict__ src,
                                                       const int64_t *__restrict__ idx,
                                                       int64_t n, int stride,
                                                       double *__restrict__ dst) {
  // one thread per (record, word); only the header and the record's
  // coefficient words are copied (the slots past its degrees are never
  // written, so they carry nothing)
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= n * stride) return;
  const int64_t r = t / stride;
  const int w = (int)(t - r * stride);
  const double *row = src + idx[r] * stride;
  if (w >= kQHead) {
    const unsigned cls = (unsigned)__double_as_longlong(row[4]);
    if (w >= kQHead + 3 + (int)(cls & 0xff) + (int)((cls >> 8) & 0xff) + (int)((cls >> 16) & 0xff))
      return;
  }
  dst[t] = row[w];
}

// flag[r] = 1 for the slots that passed screen 2 (no kQDropped mark).
__global__ __launch_bounds__(kT) void k_live_flags(const double *__restrict__ q, int stride,
                                                   int64_t n, uint8_t *__restrict__ flag) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= n) return;
  flag[r] = (uint8_t)!((unsigned)__double_as_longlong(q[r * stride + 4]) & kQDropped);
}

// Histogram of the binary exponent of B (header word 5) over the records
// whose flag is set (all of them when flag is null), LDS per block, then one
// global atomic per non-empty bin.
constexpr int kMagBins = 2100;  // ilogb in [-1075, 1024]
constexpr int kMagOff = 1075;

__global__ __launch_bounds__(kT) void k_mag_hist(const double *__restrict__ qtmp, int stride,
                                                 const uint8_t *__restrict__ flag, int64_t n,
                                                 unsigned long long *__restrict__ hist) {
  __shared__ unsigned int h[kMagBins];
  for (int b = threadIdx.x; b < kMagBins; b += kT) h[b] = 0;
  __syncthreads();
  for (int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x; r < n; r += (int64_t)gridDim.x * kT) {
    if (flag && !flag[r]) continue;
    const double u = qtmp[r * stride + 5];
    int b = (u > 0.0) ? ilogb(u) + kMagOff : 0;
    b = b < 0 ? 0 : (b >= kMagBins ? kMagBins - 1 : b);
    atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < kMagBins; b += kT)
    if (h[b]) atomicAdd(&hist[b], (unsigned long long)h[b]);
}

// heavy (B >= tau) or light (B < tau) among the records that passed (all of
// them when pass is null)
__global__ __launch_bounds__(kT) void k_split_flags(const double *__restrict__ qtmp, int stride,
                                                    const uint8_t *__restrict__ pass, int64_t n,
                                                    double tau, int heavy,
                                                    uint8_t *__restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (r >= n) return;
  uint8_t f = 0;
  if (!pass || pass[r]) {
    const bool big = qtmp[r * stride + 5] >= tau;
    f = heavy ? big : !big;
  }
  out[r] = f;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

// Contexts on several host threads may prepare at once: the first upload
// per device is made under a lock.
bool g_const_ready[64] = {false};
std::mutex g_const_mu;

int upload_constants(inca_ctx *ctx) {
  std::lock_guard<std::mutex> lock(g_const_mu);
  if (ctx->device >= 0 && ctx->device < 64 && g_const_ready[ctx->device]) return 0;
  double w[11][11] = {};
  const double sqrtpi = std::sqrt(std::acos(-1.0));
  for (int n = 1; n <= 11; ++n) {
    double factn = 1.0;
    for (int i = 2; i <= n; ++i) factn *= i;
    const double cons = std::pow(2.0, (double)(n - 1)) * factn * sqrtpi / (double)(n * n);
    for (int i = 0; i < n; ++i) {
      const double x = kGhRoot[n - 1][i];
      double h0 = 1.0, h1 = 2.0 * x;  // H_{n-1}(x) by recurrence
      double h = h0;
      if (n >= 2) {
        for (int k = 2; k < n; ++k) {
          double h2 = 2.0 * x * h1 - 2.0 * (k - 1) * h0;
          h0 = h1;
          h1 = h2;
        }
        h = h1;
      }
      w[n - 1][i] = cons * (1.0 / (h * h));
    }
  }
  double wm[kNpMax], gm[kNpMax];
  for (int p = 1; p <= kNpMax; ++p)
    wm[p - 1] = (p == 1) ? 1.0 : std::pow(p * 0.5, p * 0.5) * std::exp(-(double)p * 0.5);
  for (int k = 0; k < kNpMax; ++k) gm[k] = std::tgamma(0.5 * (k + 1));
  double ym[kNpMax];
  for (int k = 0; k < kNpMax; ++k)
    ym[k] = (k == 0) ? 1.0 : std::pow(0.5 * k, 0.5 * k) * std::exp(-0.5 * k);
  INCA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gh_x), kGhRoot, sizeof(kGhRoot)));
  INCA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gh_w), w, sizeof(w)));
  INCA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_wm), wm, sizeof(wm)));
  INCA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_gmom), gm, sizeof(gm)));
  INCA_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_ymax), ym, sizeof(ym)));
  if (ctx->device >= 0 && ctx->device < 64) g_const_ready[ctx->device] = true;
  return 0;
}

}  // namespace

int build_pair_table(inca_ctx *ctx) {
  int rc;
  if ((rc = upload_constants(ctx))) return rc;
  const int64_t np2 = (int64_t)ctx->nprim * ctx->nprim;
  if (ctx->d_pair) (void)hipFree(ctx->d_pair);
  ctx->d_pair = nullptr;
  INCA_HIP(hipMalloc((void **)&ctx->d_pair, (size_t)np2 * sizeof(PairEntry)));
  hipLaunchKernelGGL(k_pair_table, dim3(nblk(np2)), dim3(kT), 0, ctx->stream, ctx->nprim,
                     ctx->d_prim, ctx->d_tmn, ctx->d_pair);
  INCA_HIP(hipGetLastError());
  return 0;
}

// Grow a device buffer to hold `need` elements, keeping its first `used`.
// The new size is 1.5 times the old one, or `need` when that does not fit
// beside the old buffer; when `need` does not fit either, the kept elements
// wait in pinned host memory while the old buffer is freed, so that a set
// of prepared quartets as large as the free device memory can be built.
template <class T>
int grow_keep(inca_ctx *ctx, T *&ptr, size_t &cap, size_t need, size_t used) {
  if (need <= cap && ptr) return 0;
  size_t n = std::max<size_t>(need, cap + cap / 2);
  T *p = nullptr;
  hipError_t e = hipMalloc((void **)&p, (n ? n : 1) * sizeof(T));
  if (e != hipSuccess && n > need) {
    (void)hipGetLastError();
    n = need;
    e = hipMalloc((void **)&p, (n ? n : 1) * sizeof(T));
  }
  if (e != hipSuccess && used && ptr) {
    (void)hipGetLastError();
    T *h = nullptr;
    e = hipHostMalloc((void **)&h, used * sizeof(T), hipHostMallocDefault);
    if (e != hipSuccess)
      return set_err(ctx, INCA_ENOMEM, "pinned staging of %zu bytes failed: %s", used * sizeof(T),
                     hipGetErrorString(e));
    e = hipMemcpyAsync(h, ptr, used * sizeof(T), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      (void)hipHostFree(h);
      return hip_fail(ctx, e, "buffer growth (to host)", __LINE__);
    }
    (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    e = hipMalloc((void **)&p, (n ? n : 1) * sizeof(T));
    if (e != hipSuccess) {
      (void)hipHostFree(h);
      return set_err(ctx, INCA_ENOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                     hipGetErrorString(e));
    }
    e = hipMemcpyAsync(p, h, used * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipHostFree(h);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return hip_fail(ctx, e, "buffer growth (from host)", __LINE__);
    }
    ptr = p;
    cap = n;
    return 0;
  }
  if (e != hipSuccess)
    return set_err(ctx, INCA_ENOMEM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T),
                   hipGetErrorString(e));
  if (used && ptr) {
    e = hipMemcpyAsync(p, ptr, used * sizeof(T), hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return hip_fail(ctx, e, "buffer growth", __LINE__);
    }
  }
  if (ptr) (void)hipFree(ptr);
  ptr = p;
  cap = n;
  return 0;
}

// The pre-pass runs over chunks of ctx->prep_chunk records (default 2^26).
// A chunk's prepared slots (one record of `stride` doubles per screen-1
// survivor) live in d_qtmp, and the screen-2 survivors of every chunk are
// appended to d_q in slot order, so a stream of any length is prepared in
// bounded scratch (at most 2^26 slots: 19 GB for s/p/d bases; the
// radial-plot benchmark's 65.9M records are one chunk with 31.9M slots).
// Slot order is file order, or with grouping on, the order of a stable
// sort by group key: the records of a shell-quartet group then sit
// together (in file order), ready for group_rows.

int prepare_quartets(inca_ctx *ctx, double thresh, int precision, int64_t *counts) {
  hipStream_t st = ctx->stream;
  ctx->rho_h = 0.0;  // word 7 of the new records holds no rho yet
  ctx->rows_sparse = 0;
  const int nprim = ctx->nprim;
  const int64_t nrec = ctx->nrec;
  int rc;
  // stride: header words + up to 3 (L+1) coefficients, even for 16-B rows
  const int la = ctx->lmax_axis;
  int stride = kQHead + 3 * (la + 1);
  stride += stride & 1;
  ctx->qstride = stride;
  unsigned long long hcnt[4] = {(unsigned long long)nrec, 0, 0, 0};
  INCA_HIP(hipMemcpyAsync(ctx->d_cnt, hcnt, sizeof(hcnt), hipMemcpyHostToDevice, st));
  if (nrec > 0) {
    hipLaunchKernelGGL(k_first_bad, dim3(nblk(nrec)), dim3(kT), 0, st,
                       (const int32_t *)ctx->d_rec, nrec, nprim, ctx->d_cnt);
    INCA_HIP(hipGetLastError());
  }
  INCA_HIP(hipMemcpyAsync(hcnt, ctx->d_cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
  INCA_HIP(hipStreamSynchronize(st));
  const int64_t nvalid = (int64_t)hcnt[0];
  const int64_t chunk = std::max<int64_t>(1, std::min(nvalid, ctx->prep_chunk));
  if ((rc = ensure(ctx, ctx->d_flag, ctx->flag_cap, (size_t)chunk))) return rc;
  if ((rc = ensure(ctx, ctx->d_idx, ctx->idx_cap, (size_t)chunk))) return rc;
  // mod_intracule.f90:163
  const double lim = thresh * (1.0 / ((double)nprim * ((double)nprim + 1.0) * 0.5));
  // block partials of sum M (the small-degree pass, then the deferred
  // records' pass), then the chunk's total after them
  const int64_t nb1 = nblk(chunk);
  // temporaries from the context's scratch block: the term-integral
  // partials, the deferral flags and the mixed split's histogram
  Carve cv;
  const bool split = !ctx->poly_exact;
  const bool deferred = split && la > 4;
  // grouping: the slots are ordered by group key, so the rows of a group
  // come out adjacent (file order within the group)
  const bool sort_groups = groupable(ctx) && chunk <= (int64_t)UINT32_MAX;
  // the register-resident fits take their slots in degree order, one launch
  // per largest degree (bases up to d functions)
  const bool sort_class = split && la <= 8 && chunk <= (int64_t)UINT32_MAX;
  const int64_t ns = (sort_groups || sort_class) ? chunk : 1;
  const size_t sort_bytes = (sort_groups || sort_class) ? sort_slots_bytes(chunk) : 0;
  // (one launch per largest degree adds at most 9 partial blocks)
  const int64_t ntot = 2 * nb1 + 16;  // the chunk's total, after the partials
  const size_t o_mom = cv.add<double>(ntot + 1), o_def = cv.add<uint8_t>(chunk),
               o_hist = cv.add<unsigned long long>(kMagBins), o_pass = cv.add<uint8_t>(chunk),
               o_dslot = cv.add<int64_t>(deferred ? chunk : 1), o_kin = cv.add<uint32_t>(ns),
               o_kout = cv.add<uint32_t>(ns), o_vin = cv.add<uint32_t>(ns),
               o_vout = cv.add<uint32_t>(ns), o_srec = cv.add<int32_t>(6 * (sort_groups ? ns : 1)),
               o_head = cv.add<uint8_t>(sort_groups ? ns : 1),
               o_fit = cv.add<int64_t>(sort_class ? ns : 1), o_bnd = cv.add<int64_t>(16),
               o_sort = cv.add<char>((int64_t)sort_bytes);
  if ((rc = ensure(ctx, ctx->d_tmp, ctx->tmp_cap, cv.size))) return rc;
  double *d_mom = Carve::at<double>(ctx->d_tmp, o_mom);
  // Fitted polynomials: records whose axes are all of degree <= 4 (most
  // of them) go through the register-resident fits; the rest are compacted
  // in slot order and handled by the scratch kernel sized for the basis.
  // Exact polynomials: the scratch kernel takes every record.
  uint8_t *d_defer = deferred ? Carve::at<uint8_t>(ctx->d_tmp, o_def) : nullptr;
  uint8_t *d_pass = Carve::at<uint8_t>(ctx->d_tmp, o_pass);
  int64_t *d_dslot = Carve::at<int64_t>(ctx->d_tmp, o_dslot);
  uint32_t *d_kin = Carve::at<uint32_t>(ctx->d_tmp, o_kin);
  uint32_t *d_kout = Carve::at<uint32_t>(ctx->d_tmp, o_kout);
  uint32_t *d_vin = Carve::at<uint32_t>(ctx->d_tmp, o_vin);
  uint32_t *d_vout = Carve::at<uint32_t>(ctx->d_tmp, o_vout);
  int32_t *d_srec = Carve::at<int32_t>(ctx->d_tmp, o_srec);
  uint8_t *d_head = Carve::at<uint8_t>(ctx->d_tmp, o_head);  // group starts of the sorted slots
  int64_t *d_fit = Carve::at<int64_t>(ctx->d_tmp, o_fit);
  int64_t *d_bnd = Carve::at<int64_t>(ctx->d_tmp, o_bnd);
  char *d_sort = Carve::at<char>(ctx->d_tmp, o_sort);
  double mom_total = 0.0;  // chunk totals added in file order
  int64_t q2 = 0;
  // groups without the records (inca_set_keep_records(ctx, 0)): grouped
  // chunk by chunk, so the prepared rows never exceed one chunk
  const bool groups_only = !ctx->keep_records && groupable(ctx);
  if (groups_only) reset_groups(ctx);
  // one chunk, sorted by group and kept: its slots become the records as
  // they are (screen-2 failures marked), compacted only when a record-by-
  // record use asks (need_records); the groups are built from them directly
  const bool sparse = sort_groups && !groups_only && nvalid <= chunk;
  for (int64_t r0 = 0; r0 < nvalid; r0 += chunk) {
    const int64_t n = std::min(chunk, nvalid - r0);
    const int32_t *recs = (const int32_t *)ctx->d_rec + 6 * r0;
    hipError_t e = hipSuccess;
    // screen 1 first: the fits run over its compacted survivors only.  Slot
    // s (0..n1) holds survivor d_idx[s] (file order); rows, flags and
    // deferral flags are indexed by slot, so d_qtmp holds n1 rows.
    hipLaunchKernelGGL(k_screen1, dim3(nblk(n)), dim3(kT), 0, st, recs, n, nprim, ctx->d_prim,
                       ctx->d_pair, lim, ctx->normalize, d_pass);
    INCA_HIP(hipGetLastError());
    int64_t n1 = 0;
    if ((rc = compact_flags(ctx, d_pass, n, ctx->d_idx, &n1))) return rc;
    if ((rc = ensure(ctx, ctx->d_qtmp, ctx->qtmp_cap, (size_t)(n1 > 0 ? n1 : 1) * stride)))
      return rc;
    // the fits read slot s's record at srec + 6 rmap[s] (rmap null: s)
    const int32_t *srec = recs;
    const int64_t *slot_rec = ctx->d_idx;
    if (sort_groups && n1 > 0) {
      hipLaunchKernelGGL(k_slot_keys, dim3(nblk(n1)), dim3(kT), 0, st, recs, n1,
                         (const int64_t *)ctx->d_idx, nprim, ctx->d_pair, d_kin, d_vin);
      INCA_HIP(hipGetLastError());
      if ((rc = sort_slots(ctx, d_sort, sort_bytes, d_kin, d_kout, d_vin, d_vout, n1))) return rc;
      hipLaunchKernelGGL(k_slot_records, dim3(nblk(n1)), dim3(kT), 0, st, d_vout, n1, recs,
                         nprim, ctx->d_pair, d_srec, d_head);
      INCA_HIP(hipGetLastError());
      srec = d_srec;
      slot_rec = nullptr;
    }
    int64_t nbt = 0;                // block partials of the term integral,
    const double *sum_from = d_mom;  // from sum_from in a fixed order
    if (sort_class) {
      // fit order: the slots stably sorted by (largest degree, degrees); one
      // launch per largest degree v, its fits in registers at size v + 1
      // (9 for v = 5..8), its partials after the previous launch's
      if (n1 > 0) {
        hipLaunchKernelGGL(k_slot_class, dim3(nblk(n1)), dim3(kT), 0, st, srec, n1, slot_rec,
                           ctx->d_tmn, nprim, d_kin, d_vin);
        INCA_HIP(hipGetLastError());
        if ((rc = sort_slots(ctx, d_sort, sort_bytes, d_kin, d_kout, d_vin, d_vout, n1, 16u)))
          return rc;
        hipLaunchKernelGGL(k_u32_to_i64, dim3(nblk(n1)), dim3(kT), 0, st, d_vout, n1, d_fit);
        hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(64), 0, st, d_bnd, 16, n1);
        hipLaunchKernelGGL(k_degree_bounds, dim3(nblk(n1)), dim3(kT), 0, st, d_kout, n1, d_bnd);
        INCA_HIP(hipGetLastError());
        int64_t bnd[16];
        INCA_HIP(hipMemcpyAsync(bnd, d_bnd, sizeof(bnd), hipMemcpyDeviceToHost, st));
        INCA_HIP(hipStreamSynchronize(st));
        for (int v = 0; v <= 8 && e == hipSuccess; ++v) {
          const int64_t b0 = bnd[v];
          if (b0 >= n1) continue;
          int64_t b1 = n1;
          for (int u = v + 1; u < 16; ++u)
            if (bnd[u] < b1) b1 = bnd[u];
          const int64_t cnt = b1 - b0;
          if (cnt <= 0) continue;
#define INCA_PREP_V(NP)                                                                       \
  hipLaunchKernelGGL((k_prep<NP, true>), dim3(nblk(cnt)), dim3(kT), 0, st, srec, cnt,          \
                     (const int64_t *)(d_fit + b0), slot_rec, nprim, ctx->d_prim, ctx->d_tmn,  \
                     ctx->d_pair, lim, ctx->normalize, stride, 0,                              \
                     ctx->cut_log2 * 0.69314718055994531, ctx->d_qtmp, ctx->d_flag,            \
                     (uint8_t *)nullptr, ctx->d_cnt, d_mom + nbt)
          switch (v) {
            case 0: INCA_PREP_V(1); break;
            case 1: INCA_PREP_V(2); break;
            case 2: INCA_PREP_V(3); break;
            case 3: INCA_PREP_V(4); break;
            case 4: INCA_PREP_V(5); break;
            default: INCA_PREP_V(9); break;
          }
#undef INCA_PREP_V
          e = hipGetLastError();
          nbt += (int64_t)nblk(cnt);
        }
      }
    } else {
      // the small-degree pass's nblk(n1) partials end at d_mom + nb1, where
      // the deferred pass's begin, so that one fixed-order sum covers both
      double *small_mom = d_mom + (nb1 - (int64_t)nblk(n1));
      if (split && n1 > 0) {
        hipLaunchKernelGGL((k_prep<5, true>), dim3(nblk(n1)), dim3(kT), 0, st, srec, n1,
                           (const int64_t *)nullptr, slot_rec, nprim, ctx->d_prim, ctx->d_tmn,
                           ctx->d_pair, lim, ctx->normalize, stride, 0,
                           ctx->cut_log2 * 0.69314718055994531, ctx->d_qtmp, ctx->d_flag,
                           d_defer, ctx->d_cnt, small_mom);
        e = hipGetLastError();
      }
      int64_t nrun = split ? 0 : n1;  // records for the scratch kernel
      if (e == hipSuccess && d_defer && n1 > 0) {
        rc = compact_flags(ctx, d_defer, n1, d_dslot, &nrun);
        if (rc) return rc;
      }
      const int64_t *run_idx = split ? d_dslot : nullptr;
      double *run_mom = split ? d_mom + nb1 : d_mom;
      if (e == hipSuccess && nrun > 0) {
#define INCA_PREP(NPM)                                                                        \
  hipLaunchKernelGGL((k_prep<NPM, false>), dim3(nblk(nrun)), dim3(kT), 0, st, srec, nrun,    \
                     run_idx, slot_rec, nprim, ctx->d_prim, ctx->d_tmn, ctx->d_pair, lim,     \
                     ctx->normalize, stride, ctx->poly_exact,                                 \
                     ctx->cut_log2 * 0.69314718055994531, ctx->d_qtmp, ctx->d_flag,           \
                     (uint8_t *)nullptr, ctx->d_cnt, run_mom)
        if (la <= 4) INCA_PREP(5);
        else if (la <= 8) INCA_PREP(9);
        else if (la <= 12) INCA_PREP(13);
        else if (la <= 16) INCA_PREP(17);
        else INCA_PREP(22);
#undef INCA_PREP
        e = hipGetLastError();
      }
      nbt = split ? (n1 > 0 ? (int64_t)nblk(n1) : 0) + (nrun > 0 ? (int64_t)nblk(nrun) : 0)
                  : (nrun > 0 ? (int64_t)nblk(nrun) : 0);
      sum_from = split ? small_mom : d_mom;
    }
    double chunk_mom = 0.0;
    if (e == hipSuccess && nbt > 0) {
      hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(kSumT), 0, st, sum_from, nbt, d_mom + ntot);
      e = hipGetLastError();
      if (e == hipSuccess)
        e = hipMemcpyAsync(&chunk_mom, d_mom + ntot, sizeof(double), hipMemcpyDeviceToHost,
                           st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e, "quartet preparation", __LINE__);
    mom_total += chunk_mom;
    // survivors of the chunk, appended in file order; groups only: the
    // chunk's survivors are grouped and their rows dropped (groups are
    // independent terms, so a shell quartet split between chunks is two
    // groups)
    int64_t qc = 0;
    if ((rc = compact_flags(ctx, ctx->d_flag, n1, ctx->d_idx, &qc))) return rc;
    if (groups_only) {
      // the chunk's slots grouped as they are (screen-2 failures skipped)
      if (qc > 0 && (rc = group_rows(ctx, ctx->d_qtmp, n1, sort_groups ? d_head : nullptr)))
        return rc;
    } else if (sparse) {
      std::swap(ctx->d_q, ctx->d_qtmp);
      std::swap(ctx->q_cap, ctx->qtmp_cap);
      ctx->rows_sparse = n1;
    } else {
      if ((rc = grow_keep(ctx, ctx->d_q, ctx->q_cap, (size_t)(q2 + qc) * stride,
                          (size_t)q2 * stride)))
        return rc;
      if (qc > 0) {
        hipLaunchKernelGGL(k_gather_records, dim3(nblk(qc * stride)), dim3(kT), 0, st,
                           ctx->d_qtmp, ctx->d_idx, qc, stride, ctx->d_q + q2 * stride);
        INCA_HIP(hipGetLastError());
      }
    }
    q2 += qc;
  }
  // in the grids' convention: their weights carry 2 pi where the volume
  // element has 4 pi (mod_build_grid.f90:132)
  ctx->term_integral = 0.5 * mom_total;
  if (groups_only) {
    // only the groups stay: the rows of the last chunk go, and every
    // record-by-record use reports the release (need_records)
    if (ctx->d_q) (void)hipFree(ctx->d_q);
    ctx->d_q = nullptr;
    ctx->q_cap = 0;
    ctx->q2 = q2;
    ctx->q_fp64 = q2;
    ctx->mixed_tau = 0.0;
    ctx->prep_precision = precision;
    ctx->rec_split_pending = 0;
    INCA_HIP(hipMemcpyAsync(hcnt, ctx->d_cnt, sizeof(hcnt), hipMemcpyDeviceToHost, st));
    INCA_HIP(hipStreamSynchronize(st));
    counts[0] = nvalid;
    counts[1] = (int64_t)hcnt[1];
    counts[2] = q2;
    counts[3] = (int64_t)hcnt[3];
    return finish_groups(ctx, precision);
  }
  if (!ctx->rows_sparse && (rc = grow_keep(ctx, ctx->d_q, ctx->q_cap,
                                           (size_t)(q2 > 0 ? q2 : 1) * stride,
                                           (size_t)q2 * stride)))
    return rc;
  ctx->q_fp64 = q2;
  ctx->mixed_tau = 0.0;
  ctx->prep_precision = precision;
  // the records' FP64/FP32 split (split_records) waits until a record-by-
  // record mixed evaluation needs it: grouped evaluations split the groups
  ctx->rec_split_pending = precision == INCA_MIXED && q2 > 0 && ctx->mixed_tail > 0.0;
  INCA_HIP(hipMemcpyAsync(hcnt, ctx->d_cnt, sizeof(hcnt), hipMemcpyDeviceToHost, st));
  INCA_HIP(hipStreamSynchronize(st));
  ctx->q2 = q2;
  counts[0] = nvalid;
  counts[1] = (int64_t)hcnt[1];
  counts[2] = q2;
  counts[3] = (int64_t)hcnt[3];
  // sparse rows are the single chunk's sorted slots: their group starts
  // are known
  if ((rc = build_groups(ctx, precision, ctx->rows_sparse ? d_head : nullptr))) return rc;
  // without groups the evaluation runs record by record: split now
  return ctx->ng > 0 ? 0 : split_records(ctx);
}


// The records' precision split of a mixed preparation.  FP32 roundoff in a
// term is bounded by a few ulps of its L1 norm, bounded by B; the lightest
// terms go to the FP32 range while their summed B stays within mixed_tail *
// |sum M| (sum M being the integral of I over all space), the rest stays
// FP64 and comes first.  Done once, when a record-by-record mixed
// evaluation (or a query of the split) first needs it.
int need_records(inca_ctx *ctx) {
  if (ctx->q2 > 0 && !ctx->d_q)
    return set_err(ctx, INCA_EINVAL,
                   "the prepared records were released (inca_release_records): only the "
                   "evaluation by shell-quartet groups remains; prepare again");
  if (ctx->rows_sparse == 0) return 0;
  // compact the kept slots: the screen-2 survivors in slot order
  hipStream_t st = ctx->stream;
  const int64_t n = ctx->rows_sparse, q2 = ctx->q2;
  const int stride = ctx->qstride;
  int rc;
  if ((rc = ensure(ctx, ctx->d_flag, ctx->flag_cap, (size_t)n)) ||
      (rc = ensure(ctx, ctx->d_idx, ctx->idx_cap, (size_t)n)) ||
      (rc = ensure(ctx, ctx->d_qtmp, ctx->qtmp_cap, (size_t)(q2 > 0 ? q2 : 1) * stride)))
    return rc;
  hipLaunchKernelGGL(k_live_flags, dim3(nblk(n)), dim3(kT), 0, st, ctx->d_q, stride, n,
                     ctx->d_flag);
  INCA_HIP(hipGetLastError());
  int64_t qc = 0;
  if ((rc = compact_flags(ctx, ctx->d_flag, n, ctx->d_idx, &qc))) return rc;
  if (qc != q2) return set_err(ctx, INCA_EINVAL, "record compaction: %lld rows for %lld records",
                               (long long)qc, (long long)q2);
  if (qc > 0) {
    hipLaunchKernelGGL(k_gather_records, dim3(nblk(qc * stride)), dim3(kT), 0, st, ctx->d_q,
                       ctx->d_idx, qc, stride, ctx->d_qtmp);
    INCA_HIP(hipGetLastError());
  }
  INCA_HIP(hipStreamSynchronize(st));
  std::swap(ctx->d_q, ctx->d_qtmp);
  std::swap(ctx->q_cap, ctx->qtmp_cap);
  ctx->rows_sparse = 0;
  return 0;
}

int split_records(inca_ctx *ctx) {
  if (!ctx->rec_split_pending) return 0;
  if (int rc = need_records(ctx)) return rc;
  ctx->rec_split_pending = 0;
  hipStream_t st = ctx->stream;
  const int64_t q2 = ctx->q2;
  const int stride = ctx->qstride;
  int rc = 0;
  Carve cv;
  const size_t o_hist = cv.add<unsigned long long>(kMagBins);
  if ((rc = ensure(ctx, ctx->d_tmp, ctx->tmp_cap, cv.size))) return rc;
  {
    // Precision by magnitude.  FP32 roundoff in a term is bounded by a few
    // ulps of its L1 norm, bounded by B; the lightest terms go to the FP32
    // kernel while their summed B stays within mixed_tail * |sum M| (sum M
    // being the integral of I over all space), the rest stays FP64.
    unsigned long long *d_h = Carve::at<unsigned long long>(ctx->d_tmp, o_hist);
    std::vector<unsigned long long> h(kMagBins, 0);
    INCA_HIP(hipMemcpyAsync(d_h, h.data(), kMagBins * sizeof(unsigned long long),
                            hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_mag_hist, dim3(1024), dim3(kT), 0, st, ctx->d_q, stride,
                       (const uint8_t *)nullptr, q2, d_h);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(h.data(), d_h, kMagBins * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(ctx, e, "magnitude histogram", __LINE__);
    // mass of a bin ~ count * 2^(exponent + 1/2)
    std::vector<double> mass(kMagBins);
    double total = 0.0;
    for (int b = 0; b < kMagBins; ++b) {
      // upper bound of a bin's mass: count * 2^(exponent + 1), with the scale
      // capped so that 2^1024 never appears (empty bins stay 0)
      mass[b] = h[b] ? (double)h[b] * std::ldexp(2.0, std::min(b - kMagOff, 1000)) : 0.0;
      total += mass[b];
    }
    const double budget = ctx->mixed_tail >= 1.0
                              ? HUGE_VAL
                              : ctx->mixed_tail * std::min(std::fabs(ctx->term_integral), total);
    double cum = 0.0;
    int cut = 0;
    while (cut < kMagBins && cum + mass[cut] <= budget) cum += mass[cut++];
    const double tau = std::ldexp(1.0, cut - kMagOff);
    // stable partition of the survivors into d_qtmp (heavy, then light), which
    // then becomes the quartet list.  Without device room for that second
    // copy every survivor stays FP64 (mixed_tau 0, no FP32 range).
    if ((rc = ensure(ctx, ctx->d_qtmp, ctx->qtmp_cap, (size_t)q2 * stride)) ||
        (rc = ensure(ctx, ctx->d_flag, ctx->flag_cap, (size_t)q2)) ||
        (rc = ensure(ctx, ctx->d_idx, ctx->idx_cap, (size_t)q2))) {
      if (rc != INCA_ENOMEM) return rc;
      (void)hipGetLastError();
      rc = 0;
    } else {
      ctx->mixed_tau = tau;
      int64_t nh = 0, nl = 0;
      hipLaunchKernelGGL(k_split_flags, dim3(nblk(q2)), dim3(kT), 0, st, ctx->d_q, stride,
                         (const uint8_t *)nullptr, q2, tau, 1, ctx->d_flag);
      rc = compact_flags(ctx, ctx->d_flag, q2, ctx->d_idx, &nh);
      if (!rc && nh > 0)
        hipLaunchKernelGGL(k_gather_records, dim3(nblk(nh * stride)), dim3(kT), 0, st, ctx->d_q,
                           ctx->d_idx, nh, stride, ctx->d_qtmp);
      if (!rc) {
        hipLaunchKernelGGL(k_split_flags, dim3(nblk(q2)), dim3(kT), 0, st, ctx->d_q, stride,
                           (const uint8_t *)nullptr, q2, tau, 0, ctx->d_flag);
        rc = compact_flags(ctx, ctx->d_flag, q2, ctx->d_idx, &nl);
      }
      if (!rc && nl > 0)
        hipLaunchKernelGGL(k_gather_records, dim3(nblk(nl * stride)), dim3(kT), 0, st, ctx->d_q,
                           ctx->d_idx, nl, stride, ctx->d_qtmp + nh * stride);
      e = hipStreamSynchronize(st);
      if (rc) return rc;
      if (e != hipSuccess) return hip_fail(ctx, e, "mixed split", __LINE__);
      std::swap(ctx->d_q, ctx->d_qtmp);
      std::swap(ctx->q_cap, ctx->qtmp_cap);
      ctx->q_fp64 = nh;
    }
  }
  return 0;
}

}  // namespace inca
