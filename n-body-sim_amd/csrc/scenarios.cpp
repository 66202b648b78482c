// Seeded scenario generators (host only).  Portable and deterministic: own
// xoshiro256** stream seeded through splitmix64, own Box-Muller, no
// implementation-defined std::*_distribution.
#include <algorithm>
#include <cmath>
#include <vector>
#include <cstdint>

#include "../../include/nbody.h"

namespace {

struct Rng {  // xoshiro256** (Blackman & Vigna 2018), splitmix64 seeding
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (auto& w : s) {
      seed += 0x9e3779b97f4a7c15ULL;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      w = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
  double uniform(double lo, double hi) { return lo + (hi - lo) * u01(); }
  double gauss(double mean, double sd) {  // Box-Muller (1958), one draw pair
    const double u1 = 1.0 - u01();         // (0, 1]
    const double u2 = u01();
    return mean + sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  void unit_vector(double& x, double& y, double& z) {  // isotropic
    const double c = 2.0 * u01() - 1.0;
    const double ph = 2.0 * M_PI * u01();
    const double s = std::sqrt(std::fmax(0.0, 1.0 - c * c));
    x = s * std::cos(ph);
    y = s * std::sin(ph);
    z = c;
  }
};

// Plummer sphere (Aarseth, Henon & Wielen 1974, A&A 37, 183): radii from the
// inverted cumulative mass profile, speeds by von Neumann rejection from
// g(q) = q^2 (1 - q^2)^(7/2) as a fraction of the local escape speed.
// Radii beyond 20 scale radii are redrawn.  Writes bodies [0, n).
void plummer_into(Rng& rng, int64_t n, double a, double mtot, double g,
                  double* m, double* x, double* y, double* z, double* vx,
                  double* vy, double* vz) {
  const double mi = mtot / (double)n;
  const double vscale = std::sqrt(g * mtot / a);
  for (int64_t i = 0; i < n; ++i) {
    double r;
    do {
      double X;
      do X = rng.u01(); while (X <= 0.0);
      r = 1.0 / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
    } while (!(r <= 20.0));
    double ux, uy, uz;
    rng.unit_vector(ux, uy, uz);
    double q, yv;
    do {
      q = rng.u01();
      yv = 0.1 * rng.u01();
    } while (yv >= q * q * std::pow(1.0 - q * q, 3.5));
    const double ve = std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
    double wx, wy, wz;
    rng.unit_vector(wx, wy, wz);
    m[i] = mi;
    x[i] = a * r * ux;
    y[i] = a * r * uy;
    z[i] = a * r * uz;
    vx[i] = vscale * q * ve * wx;
    vy[i] = vscale * q * ve * wy;
    vz[i] = vscale * q * ve * wz;
  }
}

void to_com_frame(int64_t n, const double* m, double* x, double* y, double* z,
                  double* vx, double* vy, double* vz) {
  double M = 0, c[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = 0; i < n; ++i) {
    M += m[i];
    c[0] += m[i] * x[i];
    c[1] += m[i] * y[i];
    c[2] += m[i] * z[i];
    c[3] += m[i] * vx[i];
    c[4] += m[i] * vy[i];
    c[5] += m[i] * vz[i];
  }
  if (M == 0.0) return;
  for (auto& v : c) v /= M;
  for (int64_t i = 0; i < n; ++i) {
    x[i] -= c[0];
    y[i] -= c[1];
    z[i] -= c[2];
    vx[i] -= c[3];
    vy[i] -= c[4];
    vz[i] -= c[5];
  }
}

bool ok_ptrs(const double* a, const double* b, const double* c,
             const double* d, const double* e, const double* f,
             const double* g) {
  return a && b && c && d && e && f && g;
}

// v > 0 and finite (false for NaN).
bool positive(double v) { return v > 0.0 && std::isfinite(v); }

}  // namespace

extern "C" int nbody_scenario_plummer(int64_t n, uint64_t seed, double a,
                                      double total_mass, double g, double* m,
                                      double* x, double* y, double* z,
                                      double* vx, double* vy, double* vz) {
  if (n <= 0 || !positive(a) || !positive(total_mass) || !positive(g) ||
      !ok_ptrs(m, x, y, z, vx, vy, vz))
    return NBODY_E_INVALID;
  Rng rng(seed);
  plummer_into(rng, n, a, total_mass, g, m, x, y, z, vx, vy, vz);
  to_com_frame(n, m, x, y, z, vx, vy, vz);
  return NBODY_OK;
}

// Star system after gui/generators.py:11-98: a central mass at rest, then
// bodies on circular orbits of radius r about it in the x-y plane, inclined
// about x by `inc` and rotated about z by a random node longitude.  Here the
// inclination is drawn uniform in [0, max_incl) per body and masses are
// log-uniform in [m_lo, m_hi].
extern "C" int nbody_scenario_star_system(int64_t n, uint64_t seed,
                                          double central_mass, double m_lo,
                                          double m_hi, double r_lo,
                                          double r_hi, double max_incl_deg,
                                          double g, double* m, double* x,
                                          double* y, double* z, double* vx,
                                          double* vy, double* vz) {
  if (n <= 0 || !positive(central_mass) || !positive(m_lo) ||
      !(m_hi >= m_lo) || !std::isfinite(m_hi) || !positive(r_lo) ||
      !(r_hi >= r_lo) || !std::isfinite(r_hi) || !(max_incl_deg >= 0) ||
      !std::isfinite(max_incl_deg) || !positive(g) ||
      !ok_ptrs(m, x, y, z, vx, vy, vz))
    return NBODY_E_INVALID;
  Rng rng(seed);
  m[0] = central_mass;
  x[0] = y[0] = z[0] = vx[0] = vy[0] = vz[0] = 0.0;
  const double llo = std::log(m_lo), lhi = std::log(m_hi);
  const double imax = max_incl_deg * M_PI / 180.0;
  for (int64_t i = 1; i < n; ++i) {
    const double r = rng.uniform(r_lo, r_hi);
    const double th = rng.uniform(0.0, 2.0 * M_PI);
    const double v = std::sqrt(g * central_mass / r);
    const double px = r * std::cos(th), py = r * std::sin(th);
    const double qx = -v * std::sin(th), qy = v * std::cos(th);
    const double inc = rng.uniform(0.0, imax);
    const double om = rng.uniform(0.0, 2.0 * M_PI);
    const double ci = std::cos(inc), si = std::sin(inc);
    const double co = std::cos(om), so = std::sin(om);
    // R = Rz(om) * Rx(inc) applied to (px, py, 0) and (qx, qy, 0)
    const double py_i = ci * py, pz_i = si * py;
    const double qy_i = ci * qy, qz_i = si * qy;
    x[i] = co * px - so * py_i;
    y[i] = so * px + co * py_i;
    z[i] = pz_i;
    vx[i] = co * qx - so * qy_i;
    vy[i] = so * qx + co * qy_i;
    vz[i] = qz_i;
    m[i] = std::exp(rng.uniform(llo, lhi));
  }
  return NBODY_OK;
}

extern "C" int nbody_scenario_colliding(int64_t n, uint64_t seed,
                                        double separation, double rel_speed,
                                        double g, double* m, double* x,
                                        double* y, double* z, double* vx,
                                        double* vy, double* vz) {
  if (n < 2 || !positive(g) || !std::isfinite(separation) ||
      !std::isfinite(rel_speed) || !ok_ptrs(m, x, y, z, vx, vy, vz))
    return NBODY_E_INVALID;
  Rng rng(seed);
  const int64_t h = n / 2;
  const double mtot = 1.0;  // masses 1/n each; each sphere holds its count/n
  int64_t starts[2] = {0, h}, counts[2] = {h, n - h};
  const double cx[2] = {-0.5 * separation, 0.5 * separation};
  const double cv[2] = {0.5 * rel_speed, -0.5 * rel_speed};
  for (int s = 0; s < 2; ++s) {
    const int64_t o = starts[s], c = counts[s];
    const double msph = mtot * (double)c / (double)n;
    plummer_into(rng, c, 1.0, msph, g, m + o, x + o, y + o, z + o, vx + o,
                 vy + o, vz + o);
    to_com_frame(c, m + o, x + o, y + o, z + o, vx + o, vy + o, vz + o);
    for (int64_t i = o; i < o + c; ++i) {
      x[i] += cx[s];
      vx[i] += cv[s];
    }
  }
  return NBODY_OK;
}

// Disc galaxy after the GUI's generate_disc_galaxy (gui/generators.py:
// 101-153): radii radius * U^density_power, uniform azimuth, a thin
// Gaussian layer (sd 0.01 radius); bodies written in order of radius, each
// on the circular speed of the mass inside it, softened,
// sqrt(g * m_enc * r / (r^2 + softening^2)), plus 10 % velocity dispersion
// in the plane and 5 % vertically.  Draws follow the GUI's order: all
// positions, then the velocities in radius order.
extern "C" int nbody_scenario_disc(int64_t n, uint64_t seed, double radius,
                                   double body_mass, double density_power,
                                   double softening, double g, double* m,
                                   double* x, double* y, double* z,
                                   double* vx, double* vy, double* vz) {
  if (n <= 0 || !positive(radius) || !positive(body_mass) ||
      !positive(density_power) || !(softening >= 0) ||
      !std::isfinite(softening) || !positive(g) ||
      !ok_ptrs(m, x, y, z, vx, vy, vz))
    return NBODY_E_INVALID;
  Rng rng(seed);
  std::vector<double> r(n), th(n), h(n);
  for (int64_t i = 0; i < n; ++i) {
    r[i] = radius * std::pow(rng.u01(), density_power);
    th[i] = rng.uniform(0.0, 2.0 * M_PI);
    h[i] = rng.gauss(0.0, radius * 0.01);
  }
  std::vector<int64_t> order(n);
  for (int64_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int64_t a, int64_t b) { return r[a] < r[b]; });
  for (int64_t k = 0; k < n; ++k) {
    const int64_t i = order[k];
    const double menc = (double)(k + 1) * body_mass;
    const double d2 = r[i] * r[i] + softening * softening;
    // a body drawn at the centre with no softening rests there (0 / 0)
    const double v = d2 > 0.0 ? std::sqrt(g * menc * r[i] / d2) : 0.0;
    const double disp = v * 0.1;
    m[k] = body_mass;
    x[k] = r[i] * std::cos(th[i]);
    y[k] = r[i] * std::sin(th[i]);
    z[k] = h[i];
    vx[k] = -v * std::sin(th[i]) + rng.gauss(0.0, disp);
    vy[k] = v * std::cos(th[i]) + rng.gauss(0.0, disp);
    vz[k] = rng.gauss(0.0, disp * 0.5);
  }
  return NBODY_OK;
}
