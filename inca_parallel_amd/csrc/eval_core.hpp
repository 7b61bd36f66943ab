// Device arithmetic of the hot grid loop (mod_intracule.f90:345-358), shared
// by the single-context kernels and the batched persistent kernel so both
// evaluate every term with the same operation sequence.
#pragma once
#include <cmath>

#include "common.hpp"

namespace inca {

constexpr int kT = 256;  // lanes per block in every evaluation kernel

// 2^(j/256), j = 0..255, each the double nearest the exact value (computed
// with 80-digit decimal arithmetic).  Kernels that evaluate exp_neg stage it
// in LDS (stage_exp_table, 2 KB) and read it with a per-lane index.
constexpr int kExpTable = 256;
static __constant__ double c_exp2_256[kExpTable] = {
    1, 1.0027112750502025, 1.0054299011128027, 1.0081558981184175,
    1.0108892860517005, 1.0136300849514894, 1.0163783149109531, 1.0191339960777379,
    1.0218971486541166, 1.0246677928971357, 1.0274459491187637, 1.030231637686041,
    1.0330248790212284, 1.0358256936019572, 1.0386341019613787, 1.0414501246883161,
    1.0442737824274138, 1.0471050958792898, 1.0499440858006872, 1.0527907730046264,
    1.0556451783605572, 1.0585073227945128, 1.0613772272892621, 1.0642549128844645,
    1.0671404006768237, 1.0700337118202419, 1.0729348675259756, 1.075843889062791,
    1.0787607977571199, 1.0816856149932152, 1.0846183622133092, 1.0875590609177697,
    1.0905077326652577, 1.0934643990728858, 1.0964290818163769, 1.0994018026302219,
    1.1023825833078409, 1.1053714457017412, 1.1083684117236787, 1.1113735033448175,
    1.1143867425958924, 1.1174081515673693, 1.1204377524096067, 1.1234755673330199,
    1.1265216186082418, 1.1295759285662881, 1.1326385195987192, 1.1357094141578055,
    1.1387886347566916, 1.1418762039695616, 1.1449721444318042, 1.1480764788401789,
    1.1511892299529827, 1.1543104205902159, 1.1574400736337511, 1.1605782120274988,
    1.1637248587775775, 1.1668800369524817, 1.1700437696832502, 1.1732160801636373,
    1.1763969916502812, 1.1795865274628758, 1.182784710984341, 1.1859915656609938,
    1.189207115002721, 1.1924313825831512, 1.1956643920398273, 1.1989061670743806,
    1.2021567314527031, 1.2054161090051239, 1.2086843236265816, 1.2119613992768012,
    1.215247359980469, 1.2185422298274085, 1.2218460329727576, 1.2251587936371455,
    1.22848053610687, 1.2318112847340759, 1.2351510639369334, 1.2384998981998165,
    1.241857812073484, 1.245224830175258, 1.2486009771892048, 1.2519862778663162,
    1.2553807570246911, 1.2587844395497165, 1.2621973503942507, 1.2656195145788063,
    1.2690509571917332, 1.2724917033894028, 1.275941778396392, 1.2794012075056693,
    1.2828700160787783, 1.2863482295460256, 1.2898358734066657, 1.2933329732290895,
    1.2968395546510096, 1.3003556433796506, 1.3038812651919358, 1.3074164459346773,
    1.3109612115247644, 1.3145155879493546, 1.318079601266064, 1.3216532776031575,
    1.3252366431597413, 1.3288297242059544, 1.3324325470831615, 1.3360451382041458,
    1.3396675240533029, 1.3432997311868353, 1.3469417862329458, 1.3505937158920345,
    1.3542555469368927, 1.3579273062129011, 1.3616090206382248, 1.3653007172040119,
    1.3690024229745905, 1.3727141650876684, 1.3764359707545302, 1.380167867260238,
    1.383909881963832, 1.3876620422985291, 1.3914243757719262, 1.3951969099662003,
    1.3989796725383112, 1.4027726912202048, 1.4065759938190154, 1.4103896082172707,
    1.4142135623730951, 1.4180478843204152, 1.4218926021691656, 1.4257477441054942,
    1.42961333839197, 1.4334894133677889, 1.4373759974489824, 1.4412731191286257,
    1.4451808069770467, 1.449099089642035, 1.4530279958490526, 1.4569675544014438,
    1.460917794180647, 1.4648787441464057, 1.4688504333369818, 1.4728328908693675,
    1.4768261459394993, 1.4808302278224719, 1.4848451658727524, 1.488870989524397,
    1.4929077282912648, 1.4969554117672355, 1.5010140696264256, 1.5050837316234065,
    1.5091644275934228, 1.5132561874526098, 1.5173590411982147, 1.5214730189088146,
    1.5255981507445384, 1.529734466947287, 1.5338819978409559, 1.5380407738316568,
    1.5422108254079407, 1.5463921831410214, 1.550584877685, 1.5547889397770887,
    1.5590044002378369, 1.5632312899713576, 1.567469639965553, 1.5717194812923414,
    1.5759808451078865, 1.5802537626528246, 1.5845382652524937, 1.588834384317164,
    1.593142151342267, 1.5974615979086271, 1.6017927556826934, 1.606135656416771,
    1.6104903319492543, 1.6148568142048607, 1.6192351351948637, 1.6236253270173289,
    1.6280274218573478, 1.632441451987275, 1.6368674497669644, 1.6413054476440063,
    1.6457554781539649, 1.6502175739206177, 1.6546917676561943, 1.6591780921616162,
    1.6636765803267364, 1.6681872651305825, 1.6727101796415966, 1.6772453570178785,
    1.681792830507429, 1.6863526334483934, 1.6909247992693053, 1.6955093614893326,
    1.7001063537185235, 1.7047158096580513, 1.7093377631004629, 1.713972247929926,
    1.7186192981224779, 1.723278947746274, 1.7279512309618377, 1.7326361820223111,
    1.7373338352737062, 1.7420442251551564, 1.746767386199169, 1.7515033530318782,
    1.7562521603732995, 1.7610138430375839, 1.7657884359332727, 1.7705759740635547,
    1.7753764925265212, 1.7801900265154245, 1.785016611318935, 1.789856282321401,
    1.7947090750031072, 1.7995750249405351, 1.8044541678066239, 1.809346539371032,
    1.8142521755003989, 1.8191711121586085, 1.8241033854070534, 1.8290490314048973,
    1.8340080864093424, 1.8389805867758937, 1.843966568958626, 1.8489660695104508,
    1.8539791250833855, 1.8590057724288205, 1.864046048397789, 1.8690999899412386,
    1.8741676341103, 1.8792490180565602, 1.8843441790323345, 1.8894531543909392,
    1.8945759815869656, 1.8997126981765553, 1.9048633418176741, 1.9100279502703899,
    1.9152065613971474, 1.9203992131630474, 1.925605943636125, 1.9308267909876271,
    1.9360617934922943, 1.9413109895286405, 1.9465744175792332, 1.9518521162309783,
    1.9571441241754002, 1.9624504802089273, 1.9677712232331759, 1.9731063922552343,
    1.9784560263879509, 1.9838201648502194, 1.9891988469672663, 1.9945921121709402};

// Copy the table into the block's LDS; every thread of the block must call it.
__device__ __forceinline__ void stage_exp_table(double *lds) {
  for (int i = threadIdx.x; i < kExpTable; i += blockDim.x) lds[i] = c_exp2_256[i];
  __syncthreads();
}

// exp(-s) for s >= 0 in FP64, 11 FP64 operations:
//   s is clamped at 1000 (exp(-1000) is 0 in double);
//   t = -256 s / ln 2 + 1.5 2^52 rounds to an integer k in one FMA (k < 0,
//     |k| < 2^19) and leaves k in the low word of t; kd = t - 1.5 2^52;
//   r = -s - k ln2/256 by a Cody-Waite split of ln2/256 (hi: 22 significant
//     bits, so k hi is exact), |r| <= ln2/512;
//   e^r by its degree-4 Taylor polynomial (truncation error
//     (ln2/512)^5/120 = 3.8e-17);
//   e^-s = e^r 2^(j/256) 2^n with k = 256 n + j, j in [0, 255], the table
//     entry from LDS and 2^n by ldexp.
__device__ __forceinline__ double exp_neg(double s, const double *__restrict__ tbl) {
  const double kMagic = 6755399441055744.0;  // 1.5 2^52
  s = fmin(s, 1000.0);
  const double t = fma(s, -369.3299304675746, kMagic);  // -256 / ln 2
  const int k = __double2loint(t);
  const double kd = t - kMagic;
  double r = fma(kd, -0.0027076052501797676, -s);    // ln2/256, high part
  r = fma(kd, -9.238825187562685e-10, r);          // ln2/256, low part
  double p = fma(r, 1.0 / 24.0, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p * tbl[k & (kExpTable - 1)], k >> 8);
}

struct QHeader {
  double sqe, ax, ay, az, rcut;
  int Lx, Ly, Lz, axes;  // axes: bit a set when axis a has L > 0
};

__device__ __forceinline__ QHeader load_header(const double *__restrict__ rec) {
  QHeader h;
  h.sqe = rec[0];
  h.ax = rec[1];
  h.ay = rec[2];
  h.az = rec[3];
  h.rcut = rec[6];
  const unsigned cls = (unsigned)__double_as_longlong(rec[4]);
  h.Lx = cls & 0xff;
  h.Ly = (cls >> 8) & 0xff;
  h.Lz = (cls >> 16) & 0xff;
  h.axes = (cls >> 24) & 7;
  return h;
}

// Horner in rp of an axis polynomial of degree L >= 1 (c[0] = highest power)
// at the P points of a lane.  The first step is fma(p, c0, c1): one copy of
// c1 into a vector register per quartet instead of P copies of c0.
template <int P>
__device__ __forceinline__ void horner(const double *__restrict__ c, int L, const double (&p)[P],
                                       double (&v)[P]) {
  const double c0 = c[0], c1 = c[1];
#pragma unroll
  for (int k = 0; k < P; ++k) v[k] = fma(p[k], c0, c1);
  for (int q = 2; q <= L; ++q) {
    const double cq = c[q];
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = fma(v[k], p[k], cq);
  }
}

// One quartet whose axes with L > 0 are exactly those flagged MX, MY, MZ.
// Axes of degree 0 contribute their constant; the constants multiply into
// one wave-uniform factor S, and the per-point product takes one operation
// per remaining factor, the last one fused into the accumulation.
template <int P, bool MX, bool MY, bool MZ>
__device__ __forceinline__ void accumulate_case(const double *__restrict__ rec, const QHeader &h,
                                                const double (&x)[P], const double (&y)[P],
                                                const double (&z)[P], double (&acc)[P],
                                                const double *__restrict__ tbl) {
  const double *__restrict__ cx = rec + kQHead;
  const double *__restrict__ cy = cx + h.Lx + 1;
  const double *__restrict__ cz = cy + h.Ly + 1;
  double px[P], py[P], pz[P], vx[P], vy[P], vz[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    px[k] = fma(h.sqe, x[k], h.ax);
    py[k] = fma(h.sqe, y[k], h.ay);
    pz[k] = fma(h.sqe, z[k], h.az);
  }
  if (MX) horner<P>(cx, h.Lx, px, vx);
  if (MY) horner<P>(cy, h.Ly, py, vy);
  if (MZ) horner<P>(cz, h.Lz, pz, vz);
  double S = 1.0;
  if (!MX) S = cx[0];
  if (!MY) S = MX ? cy[0] : S * cy[0];
  if (!MZ) S = (MX && MY) ? cz[0] : S * cz[0];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const double s = fma(px[k], px[k], fma(py[k], py[k], pz[k] * pz[k]));
    const double e = exp_neg(s, tbl);
    if (MX && MY && MZ) {
      acc[k] = fma(e * vx[k] * vy[k], vz[k], acc[k]);
    } else {
      double f = e * S;
      double last;
      if (MZ) {
        if (MX) f *= vx[k];
        if (MY) f *= vy[k];
        last = vz[k];
      } else if (MY) {
        if (MX) f *= vx[k];
        last = vy[k];
      } else if (MX) {
        last = vx[k];
      } else {
        last = 1.0;
      }
      acc[k] = (MX || MY || MZ) ? fma(f, last, acc[k]) : acc[k] + f;
    }
  }
}

// One prepared quartet at the P points of a lane:
// acc[k] += A exp(-|rp|^2) Vx Vy Vz with rp = sqe r + sqe (P_ik - P_jl);
// tbl is the LDS copy of c_exp2_64.
template <int P>
__device__ __forceinline__ void accumulate_fp64(const double *__restrict__ rec, const QHeader &h,
                                                const double (&x)[P], const double (&y)[P],
                                                const double (&z)[P], double (&acc)[P],
                                                const double *__restrict__ tbl) {
  switch (h.axes) {
    case 0: accumulate_case<P, false, false, false>(rec, h, x, y, z, acc, tbl); break;
    case 1: accumulate_case<P, true, false, false>(rec, h, x, y, z, acc, tbl); break;
    case 2: accumulate_case<P, false, true, false>(rec, h, x, y, z, acc, tbl); break;
    case 3: accumulate_case<P, true, true, false>(rec, h, x, y, z, acc, tbl); break;
    case 4: accumulate_case<P, false, false, true>(rec, h, x, y, z, acc, tbl); break;
    case 5: accumulate_case<P, true, false, true>(rec, h, x, y, z, acc, tbl); break;
    case 6: accumulate_case<P, false, true, true>(rec, h, x, y, z, acc, tbl); break;
    default: accumulate_case<P, true, true, true>(rec, h, x, y, z, acc, tbl); break;
  }
}

// Ray tiles of a radial scan: the P points of a lane are r_k = o + k h w
// (k < P) for the lane's first point o, its unit direction w and the scan's
// step h, so along the lane
//   rp_k = rp_0 + k d,  d = sqe h w,  s_k = |rp_k|^2,
//   s_(k+1) - s_k = D_k,  D_0 = 2 d.rp_0 + |d|^2,  D_(k+1) - D_k = 2 |d|^2,
// and exp(-s) follows the recurrence
//   e_(k+1) = e_k q_k,  q_(k+1) = q_k rho,  q_0 = exp(-D_0),
// rho = exp(-2 sqe^2 h^2) (record word 7, set for the scan's step): two
// exponentials per lane and quartet and two products per further point, in
// place of one exponential per point.  The points are the ideal ones,
// o + k h w; the grid's own differ from them by the rounding of their
// coordinates (~1e-15 bohr), which moves s by at most 2 |rp| sqe 1e-15.
//
// Range.  With s_0 <= 690 every factor stays a normal double while the
// values matter: e_0 >= e^-690; q_0 <= e^690 (D_0 >= -s_0); s is convex in
// k, so a product that underflows on the way (s_k > 708) is never followed
// by a larger value.  A lane with s_0 > 690 has |rp_k| >= sqrt(690) -
// (P - 1) sqe h, and the recurrence is used only when that is beyond the
// quartet's cutoff radius (sqe (P - 1) h + rcut <= sqrt(690)): such a lane's
// points are all below the cutoff's 2^-cut bound, and its recurrence, with
// D_0 clamped at -700, yields values between 0 and the true ones.  Other
// quartets (tight ones, or all of them when the cutoff is off) take
// exp(-s_k) point by point at the same ideal points.
constexpr double kRayMaxRp = 26.26;  // sqrt(690)

// Horner in rp of one axis polynomial (degree L >= 1) at one point.
__device__ __forceinline__ double horner1(const double *__restrict__ c, int L, double p) {
  double v = fma(p, c[0], c[1]);
  for (int q = 2; q <= L; ++q) v = fma(v, p, c[q]);
  return v;
}

// One axis polynomial (degree L >= 1, c[0] the highest power) at the ray
// points p_k = p0 + k d, k = k0 .. k0 + N - 1.  Degree 1 is written in k
// itself, V_k = V_0 + k c0 d (one FMA per point instead of forming p_k and
// then Horner); higher degrees use Horner on p_k.  Point 0 is p0 itself.
template <int N>
__device__ __forceinline__ void axis_ray(const double *__restrict__ c, int L, double p0, double d,
                                         int k0, double (&v)[N]) {
  if (L == 1) {
    const double c0 = c[0];
    const double b1 = c0 * d;
    const double v0 = fma(p0, c0, c[1]);
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = (k0 + k == 0) ? v0 : fma((double)(k0 + k), b1, v0);
    return;
  }
  double p[N];
#pragma unroll
  for (int k = 0; k < N; ++k) p[k] = (k0 + k == 0) ? p0 : fma((double)(k0 + k), d, p0);
  horner<N>(c, L, p, v);
}

// Product of the axis polynomials with L > 0 (MX, MY, MZ) at the N lane
// points k0 .. k0 + N - 1, one axis at a time.  ZLINE: only z varies along
// the lane.
template <int N, bool ZLINE, bool MX, bool MY, bool MZ>
__device__ __forceinline__ void ray_poly(const double *__restrict__ cx,
                                         const double *__restrict__ cy,
                                         const double *__restrict__ cz, const QHeader &h,
                                         double p0x, double p0y, double p0z, double dx, double dy,
                                         double dz, int k0, double (&w)[N]) {
  double v[N];
  if (ZLINE) {
    if (MZ) axis_ray<N>(cz, h.Lz, p0z, dz, k0, w);
    return;
  }
  if (MX) axis_ray<N>(cx, h.Lx, p0x, dx, k0, w);
  if (MY) {
    axis_ray<N>(cy, h.Ly, p0y, dy, k0, v);
#pragma unroll
    for (int k = 0; k < N; ++k) {
      w[k] = MX ? w[k] * v[k] : v[k];
      // keep the product here: sunk past the z axis, it holds x, y and z
      // values live together and spills at 8 points per lane
      if (MX && MZ) asm volatile("" : "+v"(w[k]));
    }
  }
  if (MZ) {
    axis_ray<N>(cz, h.Lz, p0z, dz, k0, v);
#pragma unroll
    for (int k = 0; k < N; ++k) w[k] = (MX || MY) ? w[k] * v[k] : v[k];
  }
}

// One quartet on a lane of P points o + k h w (k = 0 .. P-1) whose axes
// with L > 0 are exactly those flagged MX, MY, MZ.  exp(-s_k) follows the
// two-product recurrence e_{k+1} = e_k q_k, q_{k+1} = q_k rho from two
// exponentials at the lane's start (rho = exp(-2 sqe^2 h^2), record word 7).
// ZLINE: the lane runs along z (a cube line), so the x and y polynomials
// are constant and fold into the uniform factor S.
template <int P, bool ZLINE, bool MX, bool MY, bool MZ>
__device__ __forceinline__ void ray_case(const double *__restrict__ rec, const QHeader &h,
                                         const double (&o)[3], const double (&hw)[3], double h2,
                                         double (&acc)[P], const double *__restrict__ tbl) {
  static_assert(P >= 2, "a ray needs two points");
  const double *__restrict__ cx = rec + kQHead;
  const double *__restrict__ cy = cx + h.Lx + 1;
  const double *__restrict__ cz = cy + h.Ly + 1;
  const double p0x = fma(h.sqe, o[0], h.ax), p0y = fma(h.sqe, o[1], h.ay),
               p0z = fma(h.sqe, o[2], h.az);
  // constants of the degree-0 axes
  double S = 1.0;
  if (!MX) S = cx[0];
  if (!MY) S = MX ? cy[0] : S * cy[0];
  if (!MZ) S = (MX && MY) ? cz[0] : S * cz[0];
  double dot, dx = 0.0, dy = 0.0, dz;
  constexpr bool perpoint = ZLINE ? MZ : (MX || MY || MZ);  // w varies along the lane
  if (ZLINE) {
    // x and y polynomials once per lane, folded into S
    if (MX) S *= horner1(cx, h.Lx, p0x);
    if (MY) S *= horner1(cy, h.Ly, p0y);
    dz = h.sqe * hw[2];
    dot = dz * p0z;
  } else {
    dx = h.sqe * hw[0];
    dy = h.sqe * hw[1];
    dz = h.sqe * hw[2];
    dot = fma(dx, p0x, fma(dy, p0y, dz * p0z));
  }
  double w[P];
  ray_poly<P, ZLINE, MX, MY, MZ>(cx, cy, cz, h, p0x, p0y, p0z, dx, dy, dz, 0, w);
  const double s0 = fma(p0x, p0x, fma(p0y, p0y, p0z * p0z));
  const double dd = h.sqe * h.sqe * h2;  // |d|^2 = (sqe h)^2
  const double rho = rec[7];
  double e = exp_neg(s0, tbl);
  double qk = exp_neg(fmax(fma(2.0, dot, dd), -700.0), tbl);
  // S multiplies e unless every factor is already in w
  if (ZLINE || !(MX && MY && MZ)) e *= S;
#pragma unroll
  for (int k = 0; k < P; ++k) {
    acc[k] = perpoint ? fma(e, w[k], acc[k]) : acc[k] + e;
    if (k + 1 < P) e *= qk;
    if (k + 2 < P) qk *= rho;
  }
}

// The same quartet point by point (exp(-s_k) at each ideal point), for the
// quartets the recurrence does not take; one point's values live at a time.
template <int P>
__device__ __forceinline__ void ray_direct(const double *__restrict__ rec, const QHeader &h,
                                        const double (&o)[3], const double (&hw)[3],
                                        double (&acc)[P], const double *__restrict__ tbl) {
  const double *__restrict__ c[3];
  c[0] = rec + kQHead;
  c[1] = c[0] + h.Lx + 1;
  c[2] = c[1] + h.Ly + 1;
  const int L[3] = {h.Lx, h.Ly, h.Lz};
  const double a[3] = {h.ax, h.ay, h.az};
  // a rolled loop over the points (the index is wave-uniform, so the
  // accumulator is picked by a uniform branch)
#pragma unroll 1
  for (int k = 0; k < P; ++k) {
    double pa[3], s = 0.0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      pa[ax] = fma(h.sqe, fma((double)k, hw[ax], o[ax]), a[ax]);
      s = fma(pa[ax], pa[ax], s);
    }
    double f = exp_neg(s, tbl);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) f *= L[ax] > 0 ? horner1(c[ax], L[ax], pa[ax]) : c[ax][0];
#pragma unroll
    for (int j = 0; j < P; ++j)
      if (j == k) acc[j] += f;
  }
}

template <int P, bool ZLINE>
__device__ __forceinline__ void ray_any(const double *__restrict__ rec, const QHeader &h,
                                        const double (&o)[3], const double (&hw)[3], double span,
                                        double h2, double (&acc)[P],
                                        const double *__restrict__ tbl) {
  if (!(fma(h.sqe, span, h.rcut) <= kRayMaxRp)) {
    ray_direct<P>(rec, h, o, hw, acc, tbl);
    return;
  }
  switch (h.axes) {
    case 0: ray_case<P, ZLINE, false, false, false>(rec, h, o, hw, h2, acc, tbl); break;
    case 1: ray_case<P, ZLINE, true, false, false>(rec, h, o, hw, h2, acc, tbl); break;
    case 2: ray_case<P, ZLINE, false, true, false>(rec, h, o, hw, h2, acc, tbl); break;
    case 3: ray_case<P, ZLINE, true, true, false>(rec, h, o, hw, h2, acc, tbl); break;
    case 4: ray_case<P, ZLINE, false, false, true>(rec, h, o, hw, h2, acc, tbl); break;
    case 5: ray_case<P, ZLINE, true, false, true>(rec, h, o, hw, h2, acc, tbl); break;
    case 6: ray_case<P, ZLINE, false, true, true>(rec, h, o, hw, h2, acc, tbl); break;
    default: ray_case<P, ZLINE, true, true, true>(rec, h, o, hw, h2, acc, tbl); break;
  }
}

// Mixed precision: rp in FP64 (sqe r + sqe d cancels), everything after it
// in FP32 -- the hardware exp2 and FP32 Horner -- and each term added to an
// FP64 accumulator.  The body is specialised on the axes of nonzero degree
// like the FP64 kernel's.
template <int P, bool MX, bool MY, bool MZ>
__device__ __forceinline__ void accumulate_mixed(const double *__restrict__ rec, const QHeader &h,
                                                 const double (&x)[P], const double (&y)[P],
                                                 const double (&z)[P], double (&acc)[P]) {
  // -log2(e) split so that s*log2(e) is formed to FP32 rounding without the
  // 1.3e-8 relative bias of a single-float constant
  const float kNL2eHi = -1.44269502162933349609375f;
  const float kNL2eLo = -1.925963033500011e-08f;
  const double *__restrict__ cx = rec + kQHead;
  const double *__restrict__ cy = cx + h.Lx + 1;
  const double *__restrict__ cz = cy + h.Ly + 1;
  float px[P], py[P], pz[P], vx[P], vy[P], vz[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    px[k] = (float)fma(h.sqe, x[k], h.ax);
    py[k] = (float)fma(h.sqe, y[k], h.ay);
    pz[k] = (float)fma(h.sqe, z[k], h.az);
  }
  auto horner32 = [&](const double *__restrict__ c, int L, const float(&p)[P], float(&v)[P]) {
    const float c0 = (float)c[0], c1 = (float)c[1];
#pragma unroll
    for (int k = 0; k < P; ++k) v[k] = fmaf(p[k], c0, c1);
    for (int j = 2; j <= L; ++j) {
      const float cj = (float)c[j];
#pragma unroll
      for (int k = 0; k < P; ++k) v[k] = fmaf(v[k], p[k], cj);
    }
  };
  if (MX) horner32(cx, h.Lx, px, vx);
  if (MY) horner32(cy, h.Ly, py, vy);
  if (MZ) horner32(cz, h.Lz, pz, vz);
  float S = 1.0f;
  if (!MX) S = (float)cx[0];
  if (!MY) S = MX ? (float)cy[0] : S * (float)cy[0];
  if (!MZ) S = (MX && MY) ? (float)cz[0] : S * (float)cz[0];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const float s = fmaf(px[k], px[k], fmaf(py[k], py[k], pz[k] * pz[k]));
    float f = __builtin_amdgcn_exp2f(fmaf(s, kNL2eHi, s * kNL2eLo));
    if (!(MX && MY && MZ)) f *= S;
    if (MX) f *= vx[k];
    if (MY) f *= vy[k];
    if (MZ) f *= vz[k];
    acc[k] += (double)f;
  }
}

template <int P>
__device__ __forceinline__ void accumulate_mixed_any(const double *__restrict__ rec,
                                                     const QHeader &h, const double (&x)[P],
                                                     const double (&y)[P], const double (&z)[P],
                                                     double (&acc)[P]) {
  switch (h.axes) {
    case 0: accumulate_mixed<P, false, false, false>(rec, h, x, y, z, acc); break;
    case 1: accumulate_mixed<P, true, false, false>(rec, h, x, y, z, acc); break;
    case 2: accumulate_mixed<P, false, true, false>(rec, h, x, y, z, acc); break;
    case 3: accumulate_mixed<P, true, true, false>(rec, h, x, y, z, acc); break;
    case 4: accumulate_mixed<P, false, false, true>(rec, h, x, y, z, acc); break;
    case 5: accumulate_mixed<P, true, false, true>(rec, h, x, y, z, acc); break;
    case 6: accumulate_mixed<P, false, true, true>(rec, h, x, y, z, acc); break;
    default: accumulate_mixed<P, true, true, true>(rec, h, x, y, z, acc); break;
  }
}

// ---- shell-quartet groups (group.hip) ----
// A group's term at rp is exp(-|rp|^2) P(rp) with P(x, y, z) = sum K_abc
// x^a y^b z^c over a <= Lx, b <= Ly, c <= Lz, a + b + c <= Ls, evaluated by
// nested Horner: for a = Lx..0 the polynomial in (y, z) of the x^a terms,
// itself Horner in y over polynomials in z, the coefficients in that order
// in the group's pool (highest powers first).  The coefficients are
// wave-uniform (scalar loads); every loop has a uniform trip count.
struct GHeader {
  double sqe, ax, ay, az;
  int Lx, Ly, Lz, Ls;
  int64_t off;
};

__device__ __forceinline__ GHeader load_gheader(const double *__restrict__ g) {
  GHeader h;
  h.sqe = g[0];
  h.ax = g[1];
  h.ay = g[2];
  h.az = g[3];
  const uint64_t cls = (uint64_t)__double_as_longlong(g[4]);
  h.Lx = (int)(cls & 0xff);
  h.Ly = (int)((cls >> 8) & 0xff);
  h.Lz = (int)((cls >> 16) & 0xff);
  h.Ls = (int)((cls >> 32) & 0xff);
  h.off = (int64_t)__double_as_longlong(g[8]);
  return h;
}

__device__ __forceinline__ double mad(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float mad(float a, float b, float c) { return fmaf(a, b, c); }

template <int P, typename T, typename C>
__device__ __forceinline__ void group_poly(const C *__restrict__ c, const GHeader &h,
                                           const T (&px)[P], const T (&py)[P], const T (&pz)[P],
                                           T (&vx)[P]) {
  for (int a = h.Lx; a >= 0; --a) {
    const int bm = min(h.Ly, h.Ls - a);
    T vy[P];
    for (int b = bm; b >= 0; --b) {
      const int cm = min(h.Lz, h.Ls - a - b);
      T vz[P];
      if (cm == 0) {
        const T c0 = (T)c[0];
#pragma unroll
        for (int k = 0; k < P; ++k) vz[k] = c0;
        c += 1;
      } else {
        const T c0 = (T)c[0], c1 = (T)c[1];
#pragma unroll
        for (int k = 0; k < P; ++k) vz[k] = mad(pz[k], c0, c1);
        for (int j = 2; j <= cm; ++j) {
          const T cj = (T)c[j];
#pragma unroll
          for (int k = 0; k < P; ++k) vz[k] = mad(vz[k], pz[k], cj);
        }
        c += cm + 1;
      }
      if (b == bm) {
#pragma unroll
        for (int k = 0; k < P; ++k) vy[k] = vz[k];
      } else {
#pragma unroll
        for (int k = 0; k < P; ++k) vy[k] = mad(vy[k], py[k], vz[k]);
      }
    }
    if (a == h.Lx) {
#pragma unroll
      for (int k = 0; k < P; ++k) vx[k] = vy[k];
    } else {
#pragma unroll
      for (int k = 0; k < P; ++k) vx[k] = mad(vx[k], px[k], vy[k]);
    }
  }
}

// The same polynomial with every monomial of total degree <= LS present
// (Lx = Ly = Lz = Ls = LS, as the group builder lays out groups of total
// degree up to 8): the loops unroll, the coefficient offsets are constants
// and the loads merge into wide scalar loads.
template <int P, int LS, typename T, typename C>
__device__ __forceinline__ void group_poly_fixed(const C *__restrict__ c, const T (&px)[P],
                                                 const T (&py)[P], const T (&pz)[P],
                                                 T (&vx)[P]) {
  int i = 0;
#pragma unroll
  for (int a = LS; a >= 0; --a) {
    T vy[P];
#pragma unroll
    for (int b = LS - a; b >= 0; --b) {
      const int cm = LS - a - b;
      T vz[P];
      if (cm == 0) {
        const T c0 = (T)c[i];
#pragma unroll
        for (int k = 0; k < P; ++k) vz[k] = c0;
        i += 1;
      } else {
        const T c0 = (T)c[i], c1 = (T)c[i + 1];
#pragma unroll
        for (int k = 0; k < P; ++k) vz[k] = mad(pz[k], c0, c1);
#pragma unroll
        for (int j = 2; j <= cm; ++j) {
          const T cj = (T)c[i + j];
#pragma unroll
          for (int k = 0; k < P; ++k) vz[k] = mad(vz[k], pz[k], cj);
        }
        i += cm + 1;
      }
      if (b == LS - a) {
#pragma unroll
        for (int k = 0; k < P; ++k) vy[k] = vz[k];
      } else {
#pragma unroll
        for (int k = 0; k < P; ++k) vy[k] = mad(vy[k], py[k], vz[k]);
      }
    }
    if (a == LS) {
#pragma unroll
      for (int k = 0; k < P; ++k) vx[k] = vy[k];
    } else {
#pragma unroll
      for (int k = 0; k < P; ++k) vx[k] = mad(vx[k], px[k], vy[k]);
    }
  }
}

// Dispatch on the group's layout: dense total degree 0..6 unrolled (7 and 8
// unrolled would hold 120-165 coefficients and set the register count of
// the whole kernel), any other layout by the general loops.
template <int P, typename T, typename C>
__device__ __forceinline__ void group_poly_any(const C *__restrict__ c, const GHeader &h,
                                               const T (&px)[P], const T (&py)[P],
                                               const T (&pz)[P], T (&v)[P]) {
  const bool dense = h.Lx == h.Ls && h.Ly == h.Ls && h.Lz == h.Ls;
  switch (dense ? h.Ls : -1) {
    case 0:
#pragma unroll
      for (int k = 0; k < P; ++k) v[k] = (T)c[0];
      break;
    case 1: group_poly_fixed<P, 1, T, C>(c, px, py, pz, v); break;
    case 2: group_poly_fixed<P, 2, T, C>(c, px, py, pz, v); break;
    case 3: group_poly_fixed<P, 3, T, C>(c, px, py, pz, v); break;
    case 4: group_poly_fixed<P, 4, T, C>(c, px, py, pz, v); break;
    case 5: group_poly_fixed<P, 5, T, C>(c, px, py, pz, v); break;
    case 6: group_poly_fixed<P, 6, T, C>(c, px, py, pz, v); break;
    default: group_poly<P, T, C>(c, h, px, py, pz, v); break;
  }
}

// One group at the P points of a lane, FP64: acc[k] += exp(-|rp|^2) P(rp).
template <int P>
__device__ __forceinline__ void group_fp64(const double *__restrict__ gh,
                                           const double *__restrict__ pool, const double (&x)[P],
                                           const double (&y)[P], const double (&z)[P],
                                           double (&acc)[P], const double *__restrict__ tbl) {
  const GHeader h = load_gheader(gh);
  double px[P], py[P], pz[P], v[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    px[k] = fma(h.sqe, x[k], h.ax);
    py[k] = fma(h.sqe, y[k], h.ay);
    pz[k] = fma(h.sqe, z[k], h.az);
  }
  group_poly_any<P, double, double>(pool + h.off, h, px, py, pz, v);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const double s = fma(px[k], px[k], fma(py[k], py[k], pz[k] * pz[k]));
    acc[k] = fma(exp_neg(s, tbl), v[k], acc[k]);
  }
}

// The polynomial at the four ray points t = 4 H .. 4 H + 3.
template <int H>
__device__ __forceinline__ void ray_half(const double *__restrict__ pool, const GHeader &h,
                                         double p0x, double p0y, double p0z, double dx,
                                         double dy, double dz, double (&w)[4]) {
  double px[4], py[4], pz[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int kk = 4 * H + k;
    px[k] = kk == 0 ? p0x : fma((double)kk, dx, p0x);
    py[k] = kk == 0 ? p0y : fma((double)kk, dy, p0y);
    pz[k] = kk == 0 ? p0z : fma((double)kk, dz, p0z);
  }
  group_poly_any<4, double, double>(pool + h.off, h, px, py, pz, w);
}

// The polynomial at the single ray point t = T.
template <int T>
__device__ __forceinline__ void ray_point(const double *__restrict__ pool, const GHeader &h,
                                          double p0x, double p0y, double p0z, double dx,
                                          double dy, double dz, double &v) {
  const double px[1] = {fma((double)T, dx, p0x)}, py[1] = {fma((double)T, dy, p0y)},
               pz[1] = {fma((double)T, dz, p0z)};
  double w[1];
  group_poly_any<1, double, double>(pool + h.off, h, px, py, pz, w);
  v = w[0];
}

// Values at t = 4..7 of a polynomial of degree L in t from its values
// v[0..L] at t = 0..L: the values themselves up to t = L, then steps of the
// backward-difference table at t = L (nabla^L constant).
template <int L>
__device__ __forceinline__ void ray_extrapolate(const double (&v)[6], double (&w)[4]) {
  double a[L + 1], n[L + 1];
#pragma unroll
  for (int i = 0; i <= L; ++i) a[i] = v[i];
  n[0] = a[L];
#pragma unroll
  for (int k = 1; k <= L; ++k) {
#pragma unroll
    for (int i = L; i >= k; --i) a[i] = a[i] - a[i - 1];
    n[k] = a[L];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (4 + k <= L) {
      w[k] = v[4 + k];
    } else {
#pragma unroll
      for (int j = L - 1; j >= 0; --j) n[j] += n[j + 1];
      w[k] = n[0];
    }
  }
}

// acc[4 H + k] += exp(-|rp_t|^2) w[k] for t = 4 H + k: by the recurrence
// (rec: e and qk carried from point to point) or point by point.
template <int H>
__device__ __forceinline__ void ray_accumulate(bool rec, double &e, double &qk, double rho,
                                               const double (&w)[4], double p0x, double p0y,
                                               double p0z, double dx, double dy, double dz,
                                               double (&acc)[8],
                                               const double *__restrict__ tbl) {
  if (rec) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[4 * H + k] = fma(e, w[k], acc[4 * H + k]);
      e *= qk;
      qk *= rho;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int kk = 4 * H + k;
      const double px = kk == 0 ? p0x : fma((double)kk, dx, p0x);
      const double py = kk == 0 ? p0y : fma((double)kk, dy, p0y);
      const double pz = kk == 0 ? p0z : fma((double)kk, dz, p0z);
      const double s = fma(px, px, fma(py, py, pz * pz));
      acc[kk] = fma(exp_neg(s, tbl), w[k], acc[kk]);
    }
  }
}

// One group on a ray lane of P points o + k h w (k = 0 .. P-1): rp_k =
// rp_0 + k d with d = sqe h w, the polynomial at each point, and exp(-s_k)
// by the two-product recurrence of the record ray kernel (e_{k+1} = e_k q_k,
// q_{k+1} = q_k rho, rho = exp(-2 sqe^2 h^2) in header word 7), under the
// same range condition (sqe (P-1) h + rcut <= sqrt(690)); groups beyond it
// take exp(-s_k) point by point at the same ideal points.
template <int P, bool DIFF = true>
__device__ __forceinline__ void group_ray(const double *__restrict__ gh,
                                          const double *__restrict__ pool, const double (&o)[3],
                                          const double (&hw)[3], double span, double h2,
                                          double (&acc)[P], const double *__restrict__ tbl) {
  const GHeader h = load_gheader(gh);
  const double rcut = gh[6];
  const double p0x = fma(h.sqe, o[0], h.ax), p0y = fma(h.sqe, o[1], h.ay),
               p0z = fma(h.sqe, o[2], h.az);
  const double dx = h.sqe * hw[0], dy = h.sqe * hw[1], dz = h.sqe * hw[2];
  if constexpr (P == 8) {
    // eight points as two sequential halves of four: the two anchor
    // exponentials spread over eight points, the polynomial keeps the
    // registers of four
    const bool rec = fma(h.sqe, span, rcut) <= kRayMaxRp;
    const double rho = gh[7];
    double e = 0.0, qk = 0.0;
    if (rec) {
      const double s0 = fma(p0x, p0x, fma(p0y, p0y, p0z * p0z));
      const double dot = fma(dx, p0x, fma(dy, p0y, dz * p0z));
      const double dd = h.sqe * h.sqe * h2;  // |d|^2 = (sqe h)^2
      e = exp_neg(s0, tbl);
      qk = exp_neg(fmax(fma(2.0, dot, dd), -700.0), tbl);
    }
    // Along the ray the group's polynomial is one of degree Ls in the step
    // index t, so for Ls <= 3 its values at t = 4..7 follow from those at
    // t = 0..3 by the backward-difference table at t = 3 (three additions
    // per point in place of M fused multiply-adds).  The two halves are
    // separate regions of code, so their polynomials do not interleave.
    // DIFF false (cube lines): Horner at every point -- a line that passes
    // a group's centre after points where its polynomial is large
    // amplifies the table's roundoff (3e-13 of the maximum on a test cube).
    double w[4];
    ray_half<0>(pool, h, p0x, p0y, p0z, dx, dy, dz, w);
    ray_accumulate<0>(rec, e, qk, rho, w, p0x, p0y, p0z, dx, dy, dz, acc, tbl);
    if (DIFF && h.Ls <= 3) {
      const double d1 = w[3] - w[2], d1b = w[2] - w[1], d1c = w[1] - w[0];
      const double d2 = d1 - d1b, d2b = d1b - d1c;
      // differences of degree above Ls are zero, not roundoff
      double pv = w[3];
      double n1 = h.Ls >= 1 ? d1 : 0.0;
      double n2 = h.Ls >= 2 ? d2 : 0.0;
      const double n3 = h.Ls >= 3 ? d2 - d2b : 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        n2 += n3;
        n1 += n2;
        pv += n1;
        w[k] = pv;
      }
    } else if (DIFF && h.Ls <= 5) {
      // degrees 4 and 5: the polynomial at t = 4 (and 5) by Horner, the
      // remaining points from the difference table of order Ls at t = Ls
      double v[6] = {w[0], w[1], w[2], w[3], 0.0, 0.0};
      ray_point<4>(pool, h, p0x, p0y, p0z, dx, dy, dz, v[4]);
      if (h.Ls == 4) {
        ray_extrapolate<4>(v, w);
      } else {
        ray_point<5>(pool, h, p0x, p0y, p0z, dx, dy, dz, v[5]);
        ray_extrapolate<5>(v, w);
      }
    } else {
      ray_half<1>(pool, h, p0x, p0y, p0z, dx, dy, dz, w);
    }
    ray_accumulate<1>(rec, e, qk, rho, w, p0x, p0y, p0z, dx, dy, dz, acc, tbl);
    return;
  }
  double px[P], py[P], pz[P], v[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    px[k] = k == 0 ? p0x : fma((double)k, dx, p0x);
    py[k] = k == 0 ? p0y : fma((double)k, dy, p0y);
    pz[k] = k == 0 ? p0z : fma((double)k, dz, p0z);
  }
  group_poly_any<P, double, double>(pool + h.off, h, px, py, pz, v);
  if (!(fma(h.sqe, span, rcut) <= kRayMaxRp)) {
#pragma unroll
    for (int k = 0; k < P; ++k) {
      const double s = fma(px[k], px[k], fma(py[k], py[k], pz[k] * pz[k]));
      acc[k] = fma(exp_neg(s, tbl), v[k], acc[k]);
    }
    return;
  }
  const double s0 = fma(p0x, p0x, fma(p0y, p0y, p0z * p0z));
  const double dot = fma(dx, p0x, fma(dy, p0y, dz * p0z));
  const double dd = h.sqe * h.sqe * h2;  // |d|^2 = (sqe h)^2
  const double rho = gh[7];
  double e = exp_neg(s0, tbl);
  double qk = exp_neg(fmax(fma(2.0, dot, dd), -700.0), tbl);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    acc[k] = fma(e, v[k], acc[k]);
    if (k + 1 < P) e *= qk;
    if (k + 2 < P) qk *= rho;
  }
}

// The same in mixed precision: rp in FP64, then FP32 (hardware exp2, the
// FP32 copy of the coefficients), each term added to the FP64 accumulator.
template <int P>
__device__ __forceinline__ void group_mixed(const double *__restrict__ gh,
                                            const float *__restrict__ pool32, const double (&x)[P],
                                            const double (&y)[P], const double (&z)[P],
                                            double (&acc)[P]) {
  const float kNL2eHi = -1.44269502162933349609375f;
  const float kNL2eLo = -1.925963033500011e-08f;
  const GHeader h = load_gheader(gh);
  float px[P], py[P], pz[P], v[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    px[k] = (float)fma(h.sqe, x[k], h.ax);
    py[k] = (float)fma(h.sqe, y[k], h.ay);
    pz[k] = (float)fma(h.sqe, z[k], h.az);
  }
  group_poly_any<P, float, float>(pool32 + h.off, h, px, py, pz, v);
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const float s = fmaf(px[k], px[k], fmaf(py[k], py[k], pz[k] * pz[k]));
    const float e = __builtin_amdgcn_exp2f(fmaf(s, kNL2eHi, s * kNL2eLo));
    acc[k] += (double)(e * v[k]);
  }
}

// XCD-aware block order.  The device dispatches workgroups to its 8 XCDs
// round-robin by linear index, and each XCD has its own L2.  With `xcd` set,
// block L runs on XCD L mod 8 and the blocks of XCD x take the chunks
// c = x (mod 8) in turn, every point tile of one chunk before the next, so a
// chunk of quartets is streamed from HBM into one L2 and reused there by all
// tiles.  The grid then has tiles x round_up(nch, 8) blocks, and a block
// whose chunk is past the end has no work.  Without it (fewer than 8
// chunks), blocks are tile-major.
constexpr int kXcds = 8;
struct BlockTask {
  int64_t tile_block, chunk;
};
__device__ __forceinline__ BlockTask block_task(int64_t tiles, int xcd) {
  const int64_t L = blockIdx.x;
  if (!xcd) return {L % tiles, L / tiles};
  const int64_t k = L / kXcds;
  return {k % tiles, (k / tiles) * kXcds + (L % kXcds)};
}

// ---- per-point cutoff: tile bounds and skip masks (shared by the
// single-context and batched evaluations) ----
namespace {

// Tile bounds.  Wave w of block b owns the 64 P consecutive grid points
// from (4 b + w) 64 P on, or, with an index list, the points
// idx[(t P + k) 64 + lane] (-1: none).  For each such tile: an anchor c and
// the range [rho_min, rho_max] of |r - c| over its points (words 0..4) and
// whether it holds any point (word 5).  Anchors: the origin for radial
// scans (the points lie on spheres about it), the owning centre for
// multicentre grids when the whole tile belongs to one centre, else the
// tile's centroid.
constexpr int kTB = 8;  // doubles per tile record

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__global__ __launch_bounds__(kT) void k_tile_bounds(const double *__restrict__ gx,
                                                    const double *__restrict__ gy,
                                                    const double *__restrict__ gz, int64_t G,
                                                    int P, int mode,
                                                    const int32_t *__restrict__ map,
                                                    const double *__restrict__ gcent,
                                                    int64_t ntile, double *__restrict__ tb,
                                                    const int32_t *__restrict__ idx) {
  const int lane = threadIdx.x & 63;
  const int64_t t = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (t >= ntile) return;  // uniform per wave
  const int64_t g0 = t * 64 * P;
  auto point = [&](int k) -> int64_t {
    if (idx) return idx[(t * P + k) * 64 + lane];
    const int64_t g = g0 + (int64_t)k * 64 + lane;
    return g < G ? g : -1;
  };
  const int first = (mode == 1 && !idx && g0 < G) ? map[g0] : 0;
  double sx = 0.0, sy = 0.0, sz = 0.0, n = 0.0;
  int same = 1;
  for (int k = 0; k < P; ++k) {
    const int64_t g = point(k);
    if (g >= 0) {
      sx += gx[g];
      sy += gy[g];
      sz += gz[g];
      n += 1.0;
      if (mode == 1 && map[g] != first) same = 0;
    }
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  sz = wave_sum(sz);
  n = wave_sum(n);
  same = __all(same);
  double cx = 0.0, cy = 0.0, cz = 0.0;
  if (mode == 1 && same && first > 0) {
    cx = gcent[3 * (first - 1)];
    cy = gcent[3 * (first - 1) + 1];
    cz = gcent[3 * (first - 1) + 2];
  } else if (mode != 0 && n > 0.0) {
    cx = sx / n;
    cy = sy / n;
    cz = sz / n;
  }
  double lo = HUGE_VAL, hi = 0.0;
  for (int k = 0; k < P; ++k) {
    const int64_t g = point(k);
    if (g >= 0) {
      const double dx = gx[g] - cx, dy = gy[g] - cy, dz = gz[g] - cz;
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      lo = fmin(lo, d);
      hi = fmax(hi, d);
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, d, 64));
    hi = fmax(hi, __shfl_xor(hi, d, 64));
  }
  if (lane == 0) {
    double *o = tb + t * kTB;
    o[0] = cx;
    o[1] = cy;
    o[2] = cz;
    // widened by far more than the roundoff of the distances
    o[3] = n > 0.0 ? lo * (1.0 - 1e-12) : 0.0;
    o[4] = n > 0.0 ? hi * (1.0 + 1e-12) + 1e-300 : 0.0;
    o[5] = n;
    o[6] = 0.0;
    o[7] = 0.0;
  }
}

__global__ __launch_bounds__(kT) void k_fill_tiles(int64_t ntile, double *__restrict__ tb) {
  const int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x;
  if (t >= ntile) return;
  for (int w = 0; w < kTB; ++w) tb[t * kTB + w] = (w == 5) ? 1.0 : 0.0;
}

// True when every point of the tile lies beyond the quartet's cutoff radius:
// |rp| = |sqe (r - c) + u|, u = sqe c + sqe d, is at least
// max(|u| - sqe rho_max, sqe rho_min - |u|) on the tile.
__device__ __forceinline__ bool tile_far(const QHeader &h, double cx, double cy, double cz,
                                         double rlo, double rhi) {
  const double ux = fma(h.sqe, cx, h.ax), uy = fma(h.sqe, cy, h.ay), uz = fma(h.sqe, cz, h.az);
  const double u2 = fma(ux, ux, fma(uy, uy, uz * uz));
  const double o = fma(h.sqe, rhi, h.rcut);
  const double i = fma(h.sqe, rlo, -h.rcut);
  return (u2 >= o * o) || (i > 0.0 && u2 <= i * i);
}

// Skip mask: bit i of mask[w][t] is set when quartet 64 w + i is beyond its
// cutoff radius at every point of wave tile t.  One wave per 64-quartet word:
// lane i holds quartet 64 w + i, a ballot per tile gives the word, and the
// words of 64 consecutive tiles are stored by the 64 lanes at once.
//
// wcount[w] (optional) receives the evaluations the mask keeps for the word:
// the sum over tiles of kept quartets x the tile's points.
//
// wflop[w] (optional, with the operations per point in record word `fword`):
// the sum over the word's kept (record, tile) pairs of points x operations.
__global__ __launch_bounds__(kT) void k_skip_mask(const double *__restrict__ q, int qstride,
                                                  int64_t nq, const double *__restrict__ tb,
                                                  int64_t ntile, int64_t nwords,
                                                  unsigned long long *__restrict__ mask,
                                                  unsigned long long *__restrict__ wcount,
                                                  int fword = -1,
                                                  double *__restrict__ wflop = nullptr) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * (kT / 64) + (threadIdx.x >> 6);
  if (w >= nwords) return;  // uniform per wave
  const int64_t qi = w * 64 + lane;
  const bool valid = qi < nq;
  const unsigned long long vbits = __ballot(valid);
  QHeader h{};
  if (valid) h = load_header(q + qi * qstride);
  unsigned long long kept = 0;
  double mine = 0.0;  // points of the tiles this lane's record keeps
  for (int64_t t0 = 0; t0 < ntile; t0 += 64) {
    unsigned long long keep = 0;
    const int nt = (int)((ntile - t0) < 64 ? (ntile - t0) : 64);
    for (int j = 0; j < nt; ++j) {
      const double *__restrict__ tbw = tb + (t0 + j) * kTB;
      const double npts = tbw[5];
      const bool far = valid && npts > 0.0 && tile_far(h, tbw[0], tbw[1], tbw[2], tbw[3], tbw[4]);
      const unsigned long long m = __ballot(far);
      if (lane == j) keep = m;
      kept += (unsigned long long)__popcll(vbits & ~m) * (unsigned long long)npts;
      if (valid && !far) mine += npts;
    }
    if (lane < nt) mask[w * ntile + t0 + lane] = keep;
  }
  if (wcount && lane == 0) wcount[w] = kept;
  if (wflop) {
    const double f = wave_sum(valid ? mine * q[qi * qstride + fword] : 0.0);
    if (lane == 0) wflop[w] = f;
  }
}

// Fixed-order sum of n doubles by one block, added to out[0].
__global__ __launch_bounds__(kT) void k_sum_f64(const double *__restrict__ x, int64_t n,
                                                double *__restrict__ out) {
  __shared__ double red[kT];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kT) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] += red[0];
}

// Fixed-order sum of n counters by one block.
__global__ __launch_bounds__(kT) void k_sum_u64(const unsigned long long *__restrict__ x,
                                                int64_t n, unsigned long long *__restrict__ out) {
  __shared__ unsigned long long red[kT];
  unsigned long long s = 0;
  for (int64_t i = threadIdx.x; i < n; i += kT) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = kT / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] += red[0];
}

// Calls f(qi) for the quartets of [a, b) that the skip mask keeps for
// `tile` (all of them when mask is null).  The mask words are wave-uniform,
// so the iteration runs on scalar bit scans.
template <class F>
__device__ __forceinline__ void for_live(const unsigned long long *__restrict__ mask, int64_t ntile,
                                         int64_t tile, int64_t a, int64_t b, F &&f) {
  if (!mask) {
    for (int64_t qi = a; qi < b; ++qi) f(qi);
    return;
  }
  for (int64_t wd = a >> 6; (wd << 6) < b; ++wd) {
    const int64_t base = wd << 6;
    unsigned long long live = ~mask[wd * ntile + tile];
    if (base < a) live &= ~0ull << (a - base);
    if (b - base < 64) live &= (1ull << (b - base)) - 1ull;
    while (live) {
      const int i = __builtin_ctzll(live);
      live &= live - 1ull;
      f(base + i);
    }
  }
}

}  // namespace

}  // namespace inca
