// ensemble.hpp -- droplet ensembles created on the device from a seed.
//
// A counter-based generator gives every droplet of an ensemble its own random
// numbers from (seed, droplet index) alone, so an ensemble of any size is made
// in one pass over HBM, the same on every run and independent of how the work
// is split: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011), a 10-round multiply/xor bijection of
// a 128-bit counter under a 64-bit key.  Droplet i uses counter (i, 0, 0, j)
// for its j-th block of four 32-bit words; the key is the seed.
//
// The droplet is then created the way create_particle does it
// (particles.f90:2226-2300): position uniform over the domain with the
// height between r and z(nnz) - r, radius and salt mass from the ensemble's
// distributions, vp = uf(xp), Tp = Tf(xp) (the FP64 gather), rc =
// crit_radius(m_s, kappa_s, Tp).
#pragma once
#include "device_common.hpp"

namespace ntlp {

struct U4 {
  uint32_t v[4];
};

__host__ __device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  constexpr uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.v[0];
    const uint64_t p1 = (uint64_t)M1 * c.v[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    U4 o;
    o.v[0] = hi1 ^ c.v[1] ^ k0;
    o.v[1] = lo1;
    o.v[2] = hi0 ^ c.v[3] ^ k1;
    o.v[3] = lo0;
    c = o;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// A double in (0, 1) from two words: 53 random bits plus one half, scaled by
// 2^-53 (exact).
__host__ __device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Distributions of an ensemble (ntlp_ensemble in include/ntlp_gpu.h).
enum { kDistLognormal = 0, kDistLogUniform = 1 };
enum { kSaltFixed = 0, kSaltDryLognormal = 1, kSaltChoice3 = 2 };

struct DevEnsemble {
  uint64_t seed;
  int radius_kind, radius2_kind;
  double radius_p[2], radius2_p[2], frac2;
  int salt_kind, pad_;
  double salt_p[3];
  double kappa_s;
  int64_t mult;
  int64_t first;  // counter index of the ensemble's first droplet
};

// The uniforms of droplet i: u[j] in (0, 1) from words 2(j%2), 2(j%2)+1 of
// Philox block j/2 (counter (i, 0, 0, j/2)).  Their use: u0, u1, u2 the
// position, u3 and u4 the radius (u4 only for a lognormal), u5 the choice of
// the radius component, u6 and u7 the salt (u7 only for a lognormal dry
// radius).
constexpr int kDraws = 8;

__host__ __device__ __forceinline__ void droplet_draws(uint64_t seed, uint64_t i, double u[kDraws]) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int b = 0; b < kDraws / 2; ++b) {
    U4 c = {{(uint32_t)i, (uint32_t)(i >> 32), 0u, (uint32_t)b}};
    const U4 w = philox4x32_10(c, k0, k1);
    u[2 * b] = u01(w.v[0], w.v[1]);
    u[2 * b + 1] = u01(w.v[2], w.v[3]);
  }
}

// standard normal from two uniforms (Box-Muller, cosine branch)
__device__ __forceinline__ double std_normal(double u1, double u2) {
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__device__ __forceinline__ double draw_radius(int kind, const double p[2], double ua, double ub) {
  if (kind == kDistLogUniform) return exp(log(p[0]) + ua * (log(p[1]) - log(p[0])));
  return p[0] * exp(p[1] * std_normal(ua, ub));
}

// The ensemble's droplet i: position, radius and salt mass (the rest of its
// state comes from the gather at its position).
struct NewDroplet {
  double x, y, z, r, m_s;
};

__device__ __forceinline__ NewDroplet make_droplet(const DevEnsemble& e, const DevGrid& g, const double* s_z,
                                                   int64_t i, double rhos, double pi2) {
  double u[kDraws];
  droplet_draws(e.seed, (uint64_t)(e.first + i), u);
  NewDroplet n;
  const bool second = e.frac2 > 0.0 && u[5] < e.frac2;
  n.r = second ? draw_radius(e.radius2_kind, e.radius2_p, u[3], u[4])
               : draw_radius(e.radius_kind, e.radius_p, u[3], u[4]);
  n.x = u[0] * g.xl;
  n.y = u[1] * g.yl;
  n.z = n.r + u[2] * (s_z[g.nnz] - 2.0 * n.r);
  if (e.salt_kind == kSaltFixed) {
    n.m_s = e.salt_p[0];
  } else if (e.salt_kind == kSaltChoice3) {
    n.m_s = u[6] < 1.0 / 3.0 ? e.salt_p[0] : (u[6] < 2.0 / 3.0 ? e.salt_p[1] : e.salt_p[2]);
  } else {
    // dry radius lognormal; m_s = rhos (4/3) pi r_d^3 (particles.f90:2498-2499)
    const double rd = e.salt_p[0] * exp(e.salt_p[1] * std_normal(u[6], u[7]));
    n.m_s = rhos * (2.0 / 3.0 * pi2) * (rd * rd * rd);
  }
  return n;
}

}  // namespace ntlp
