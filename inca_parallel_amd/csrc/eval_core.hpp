// Device arithmetic of the hot grid loop (mod_intracule.f90:345-358), shared
// by the single-context kernels and the batched persistent kernel so both
// evaluate every term with the same operation sequence.
#pragma once
#include <cmath>

#include "common.hpp"

namespace inca {

constexpr int kT = 256;  // lanes per block in every evaluation kernel

__device__ __forceinline__ double exp_neg(double s) {
  // exp(-s) for s >= 0 (s <= 1000 clamp keeps k inside int range; e^-1000 = 0)
  s = fmin(s, 1000.0);
  const double kLog2e = 1.4426950408889634;
  const double kLn2Hi = 6.93147180369123816490e-01;  // Cody-Waite split of ln 2
  const double kLn2Lo = 1.90821492927058770002e-10;
  const double k = rint(s * kLog2e);
  double r = fma(k, kLn2Hi, -s);
  r = fma(k, kLn2Lo, r);
  // Taylor coefficients 1/n!, n = 12 .. 0
  double p = 2.08767569878680989792e-09;
  p = fma(p, r, 2.50521083854417187751e-08);
  p = fma(p, r, 2.75573192239858906526e-07);
  p = fma(p, r, 2.75573192239858906526e-06);
  p = fma(p, r, 2.48015873015873015873e-05);
  p = fma(p, r, 1.98412698412698412698e-04);
  p = fma(p, r, 1.38888888888888888889e-03);
  p = fma(p, r, 8.33333333333333333333e-03);
  p = fma(p, r, 4.16666666666666666667e-02);
  p = fma(p, r, 1.66666666666666666667e-01);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, -(int)k);
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

// One prepared quartet at the P points of a lane:
// acc[k] += A exp(-|rp|^2) Vx Vy Vz with rp = sqe r + sqe (P_ik - P_jl).
template <int P>
__device__ __forceinline__ void accumulate_fp64(const double *__restrict__ rec, const QHeader &h,
                                                const double (&x)[P], const double (&y)[P],
                                                const double (&z)[P], double (&acc)[P]) {
  const double *__restrict__ cx = rec + kQHead;
  const double *__restrict__ cy = cx + h.Lx + 1;
  const double *__restrict__ cz = cy + h.Ly + 1;
  double px[P], py[P], pz[P], vx[P], vy[P], vz[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    px[k] = fma(h.sqe, x[k], h.ax);
    py[k] = fma(h.sqe, y[k], h.ay);
    pz[k] = fma(h.sqe, z[k], h.az);
    vx[k] = cx[0];
    vy[k] = cy[0];
    vz[k] = cz[0];
  }
  for (int p = 1; p <= h.Lx; ++p) {
    const double c = cx[p];
#pragma unroll
    for (int k = 0; k < P; ++k) vx[k] = fma(vx[k], px[k], c);
  }
  for (int p = 1; p <= h.Ly; ++p) {
    const double c = cy[p];
#pragma unroll
    for (int k = 0; k < P; ++k) vy[k] = fma(vy[k], py[k], c);
  }
  for (int p = 1; p <= h.Lz; ++p) {
    const double c = cz[p];
#pragma unroll
    for (int k = 0; k < P; ++k) vz[k] = fma(vz[k], pz[k], c);
  }
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const double s = fma(px[k], px[k], fma(py[k], py[k], pz[k] * pz[k]));
    acc[k] = fma(exp_neg(s) * vx[k] * vy[k], vz[k], acc[k]);
  }
}

}  // namespace inca
