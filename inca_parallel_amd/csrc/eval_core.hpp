// Device arithmetic of the hot grid loop (mod_intracule.f90:345-358), shared
// by the single-context kernels and the batched persistent kernel so both
// evaluate every term with the same operation sequence.
#pragma once
#include <cmath>

#include "common.hpp"

namespace inca {

constexpr int kT = 256;  // lanes per block in every evaluation kernel

// 2^(j/64), j = 0..63, each the double nearest the exact value (computed
// with 80-digit decimal arithmetic).  Kernels that evaluate exp_neg stage it
// in LDS (stage_exp_table) and read it with a per-lane index.
static __constant__ double c_exp2_64[64] = {
    1, 1.0108892860517005, 1.0218971486541166, 1.0330248790212284,
    1.0442737824274138, 1.0556451783605572, 1.0671404006768237, 1.0787607977571199,
    1.0905077326652577, 1.1023825833078409, 1.1143867425958924, 1.1265216186082418,
    1.1387886347566916, 1.1511892299529827, 1.1637248587775775, 1.1763969916502812,
    1.189207115002721, 1.2021567314527031, 1.215247359980469, 1.22848053610687,
    1.241857812073484, 1.2553807570246911, 1.2690509571917332, 1.2828700160787783,
    1.2968395546510096, 1.3109612115247644, 1.3252366431597413, 1.3396675240533029,
    1.3542555469368927, 1.3690024229745905, 1.383909881963832, 1.3989796725383112,
    1.4142135623730951, 1.42961333839197, 1.4451808069770467, 1.460917794180647,
    1.4768261459394993, 1.4929077282912648, 1.5091644275934228, 1.5255981507445384,
    1.5422108254079407, 1.5590044002378369, 1.5759808451078865, 1.593142151342267,
    1.6104903319492543, 1.6280274218573478, 1.6457554781539649, 1.6636765803267364,
    1.681792830507429, 1.7001063537185235, 1.7186192981224779, 1.7373338352737062,
    1.7562521603732995, 1.7753764925265212, 1.7947090750031072, 1.8142521755003989,
    1.8340080864093424, 1.8539791250833855, 1.8741676341103, 1.8945759815869656,
    1.9152065613971474, 1.9360617934922943, 1.9571441241754002, 1.9784560263879509};

// Copy the table into the block's LDS; every thread of the block must call it.
__device__ __forceinline__ void stage_exp_table(double *lds) {
  if (threadIdx.x < 64) lds[threadIdx.x] = c_exp2_64[threadIdx.x];
  __syncthreads();
}

// exp(-s) for s >= 0 in FP64.  k = rint(-64 s / ln 2) (s clamped at 1000,
// where exp(-s) is 0 in double), r = -s - k ln2/64 by a Cody-Waite split of
// ln2/64 (hi: 22 significant bits, so k hi is exact for |k| < 2^31), |r| <=
// ln2/128; e^r by its degree-5 Taylor polynomial (truncation error
// (ln2/128)^6/720 = 3.5e-17); then e^-s = e^r 2^(j/64) 2^n with k = 64 n + j,
// j in [0, 63], from the LDS table.  13 FP64 operations.
__device__ __forceinline__ double exp_neg(double s, const double *__restrict__ tbl) {
  s = fmin(s, 1000.0);
  const double kd = rint(s * -92.33248261689366);  // -64 / ln 2
  const int k = (int)kd;
  double r = fma(kd, -0.01083042100071907, -s);    // ln2/64, high part
  r = fma(kd, -3.695530075025074e-09, r);          // ln2/64, low part
  double p = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p * tbl[k & 63], k >> 6);
}

struct QHeader {
  double sqe, ax, ay, az, rcut;
  int Lx, Ly, Lz;
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
  const int mask = (h.Lx > 0) | ((h.Ly > 0) << 1) | ((h.Lz > 0) << 2);
  switch (mask) {
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

}  // namespace inca
