// Integration centres of the multicentre intracule grid (host code).
//
// Restates mod_autogrid.f90:12-185 (centercalc + get_bs_radius) for callers
// of the C-ABI that have no Fortran host: the origin carries the
// same-atom pairs, each ordered heavy-atom pair (j,k) contributes the two
// intracule centres +-(R_j - R_k) (z snapped to 0 below 0.5 bohr), centres
// closer than 0.1 bohr on every axis merge with population-weighted radii,
// and Bragg-Slater radii (angstrom x 1.8897) scale the radial maps.  A
// caller may pass its own per-atom radii instead of the Bragg-Slater table:
// the centre formulas stay those of the reference.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace inca {

// Bragg-Slater radii in angstrom, Z = 1..36 (mod_autogrid.f90:142-185); 1.00
// for heavier elements.
double bragg_slater_bohr(int z) {
  static const double kAngstrom[36] = {
      0.35, 0.35, 1.45, 1.05, 0.85, 0.70, 0.65, 0.60, 0.50, 0.38, 1.80, 1.50,
      1.25, 1.10, 1.00, 1.00, 1.00, 1.00, 2.20, 1.80, 1.60, 1.40, 1.35, 1.40,
      1.40, 1.40, 1.35, 1.35, 1.35, 1.35, 1.30, 1.25, 1.15, 1.15, 1.15, 1.00};
  const double a = (z >= 1 && z <= 36) ? kAngstrom[z - 1] : 1.00;
  return a * 1.8897;
}

namespace {

struct Centre {
  double c[3];
  double pop;    // electron-pair population (the TFVC size parameter Ps)
  double alpha;  // radial scale
};

bool at_origin(const Centre &q) { return std::fabs(q.c[0]) + std::fabs(q.c[1]) + std::fabs(q.c[2]) == 0.0; }

}  // namespace

int autogrid(int natoms, const int32_t *an, const double *cartes, int nohydro, int betaone,
             const double *radius, std::vector<double> &cent, std::vector<double> &Ps,
             std::vector<double> &sfalpha) {
  auto rad = [&](int a) { return radius ? radius[a] : bragg_slater_bohr(an[a]); };
  auto xyz = [&](int a, int ax) { return cartes[a + (size_t)natoms * ax]; };
  std::vector<Centre> cs;
  // origin: same-atom pairs, radius weighted by their count
  double pairs = 0.0, wr = 0.0;
  for (int a = 0; a < natoms; ++a) {
    const double p = an[a] * (an[a] - 1.0) / 2.0;
    pairs += p;
    wr += p * rad(a);
  }
  cs.push_back(Centre{{0.0, 0.0, 0.0}, pairs, pairs > 1e-15 ? wr / pairs : bragg_slater_bohr(1)});
  for (int j = 0; j + 1 < natoms; ++j)
    for (int k = j + 1; k < natoms; ++k) {
      if (nohydro && !(an[j] > 1 && an[k] > 1)) continue;
      Centre q;
      for (int ax = 0; ax < 3; ++ax) q.c[ax] = xyz(j, ax) - xyz(k, ax);
      if (std::fabs(q.c[2]) < 0.5) q.c[2] = 0.0;
      q.pop = an[j] * an[k] * 0.5;
      q.alpha = 0.5 * (rad(j) + rad(k));
      cs.push_back(q);
      for (double &v : q.c) v = -v;
      cs.push_back(q);
    }
  // merge (the origin and the last centre take no part as the first member,
  // as in the reference's 2..nqmax-1 loop)
  const int n = (int)cs.size();
  for (int i = 1; i < n - 1; ++i) {
    if (at_origin(cs[i])) continue;
    for (int j = i + 1; j < n; ++j) {
      bool near = true;
      for (int ax = 0; ax < 3; ++ax) near = near && std::fabs(cs[i].c[ax] - cs[j].c[ax]) < 0.1;
      if (!near) continue;
      cs[i].alpha = (cs[i].alpha * cs[i].pop + cs[j].alpha * cs[j].pop) / (cs[i].pop + cs[j].pop);
      cs[i].pop += cs[j].pop;
      cs[j] = Centre{{0.0, 0.0, 0.0}, 0.0, cs[j].alpha};
    }
  }
  cent.clear();
  Ps.clear();
  sfalpha.clear();
  for (int i = 0; i < n; ++i) {
    if (i > 0 && at_origin(cs[i])) continue;
    cent.insert(cent.end(), cs[i].c, cs[i].c + 3);
    Ps.push_back(betaone ? 1.0 : cs[i].pop);
    sfalpha.push_back(betaone ? 1.0 : cs[i].alpha);
  }
  return (int)Ps.size();
}

// Per-atom descriptors of the learned autogrid (16 values, each scaled to
// order one): element, exponent range and angular momentum of the atom's
// primitives, its neighbourhood, and molecule-wide exponent range.  prim_atom
// is 1-based (geninfo::Ra); tmn(nprim,3) column-major.
void atom_features(int natoms, const int32_t *an, const double *cartes, int nprim,
                   const double *alpha, const int32_t *tmn, const int32_t *prim_atom, float *f) {
  auto xyz = [&](int a, int ax) { return cartes[a + (size_t)natoms * ax]; };
  double gmin = HUGE_VAL, gmax = 0.0;
  for (int i = 0; i < nprim; ++i) {
    gmin = std::fmin(gmin, alpha[i]);
    gmax = std::fmax(gmax, alpha[i]);
  }
  if (nprim == 0) gmin = gmax = 1.0;
  for (int a = 0; a < natoms; ++a) {
    double amin = HUGE_VAL, amax = 0.0, cnt = 0.0;
    int lmax = 0;
    for (int i = 0; i < nprim; ++i) {
      if (prim_atom[i] != a + 1) continue;
      amin = std::fmin(amin, alpha[i]);
      amax = std::fmax(amax, alpha[i]);
      cnt += 1.0;
      lmax = std::max(lmax, tmn[i] + tmn[i + nprim] + tmn[i + 2 * nprim]);
    }
    if (cnt == 0.0) amin = amax = 1.0;
    double dmin = 0.0, dsum = 0.0, near = 0.0, coul = 0.0;
    bool first = true;
    for (int b = 0; b < natoms; ++b) {
      if (b == a) continue;
      double d2 = 0.0;
      for (int ax = 0; ax < 3; ++ax) {
        const double d = xyz(a, ax) - xyz(b, ax);
        d2 += d * d;
      }
      const double d = std::sqrt(d2);
      dmin = first ? d : std::fmin(dmin, d);
      first = false;
      dsum += d;
      near += (d < 4.0) ? 1.0 : 0.0;
      coul += (d > 1e-8) ? an[b] / d : 0.0;
    }
    const int z = an[a];
    const int period = z <= 2 ? 1 : (z <= 10 ? 2 : (z <= 18 ? 3 : 4));
    float *o = f + (size_t)a * 16;
    o[0] = (float)(z / 10.0);
    o[1] = (float)(z == 1 ? 1.0 : 0.0);
    o[2] = (float)(period / 3.0);
    o[3] = (float)(std::log10(amin) / 3.0);
    o[4] = (float)(std::log10(amax) / 3.0);
    o[5] = (float)(cnt / 20.0);
    o[6] = (float)(lmax / 2.0);
    o[7] = (float)(dmin / 5.0);
    o[8] = (float)(natoms > 1 ? dsum / (natoms - 1) / 5.0 : 0.0);
    o[9] = (float)(near / 10.0);
    o[10] = (float)(natoms / 20.0);
    o[11] = (float)(bragg_slater_bohr(z) / 3.0);
    o[12] = (float)(coul / 10.0);
    o[13] = (float)(std::log10(gmin) / 3.0);
    o[14] = (float)(std::log10(gmax) / 3.0);
    o[15] = 1.0f;
  }
}

}  // namespace inca
