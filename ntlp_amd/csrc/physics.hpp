// physics.hpp -- per-droplet physics of the droplet step on gfx950.
//
// FP64 ("reference") paths keep the operation order of the Fortran routines
// they restate; loop invariants are hoisted (Sc**(1/3), Pr**(1/3), taup0,
// mod_magnus(Tf), |u_f - v|), which changes no result.  Dead reference
// computations whose values no caller reads are not formed: the velocity
// residual of ie_vrt_nd and the backward evaluations of jacob_approx_2d.
//
// The analytic-Jacobian Newton (NTLP_PREC_MIXED, NTLP_PREC_FP64_AJ) solves
// the same backward-Euler equations (the residual ie_vrt_nd of
// particle_update_BE, particles.f90:4263-4327) with the exact 2x2 Jacobian
// from the chain rule,
// the same damping, guards, tolerance and Levenberg-Marquardt fallback.
#pragma once
#include "device_common.hpp"

namespace ntlp {

enum { PREC_FP64 = 0, PREC_MIXED = 1, PREC_FP64_AJ = 2 };

struct Drop {
  double x, y, z, vx, vy, vz, Tp, r, m_s, kappa, rc, res, actres, numact;
};

// Values the reference stores in the droplet record besides the hot state.
struct Diag {
  double uf[3], Tf, qinf, qstar;
  double vold[3], Tpold, rold;
  double xrhs[3], vrhs[3], Tprhs_s, Tprhs_L, radrhs;
  // u level of the gather position; a BE step packs the first w level above
  // it (the flux statistics' kc0) into the same register: see pack_levels
  int kpt;
};

// kpt (-2..nnz+1) and kc0 (-1..nnz+1) of one droplet in one int (nnz + 2 <=
// 1024 levels): one register held across the radius solve instead of two.
__device__ __forceinline__ int pack_levels(int kpt, int kc0) { return (kpt + 2) | ((kc0 + 1) << 11); }
__device__ __forceinline__ int packed_kpt(int v) { return (v & 2047) - 2; }
__device__ __forceinline__ int packed_kc0(int v) { return (v >> 11) - 1; }

// Per-thread Newton bookkeeping (particles.f90:3035-3039, 3128, 3147-3174).
struct Tally {
  int num100, numimpos, actnum, denum;
  int gn_iters, lm_iters, solves, hist;  // hist: GN iterations of this solve, 0 = LM used, -1 none
};

__device__ __forceinline__ double mod_magnus(double T) {  // les.F:10764-10774
  return 610.94 * exp((17.6257 * (T - 273.15)) / (243.04 + (T - 273.15)));
}
__device__ __forceinline__ double func_p_base(const DevConsts& C, double z) {  // les.F:10787-10795
  return C.surf_p * pow(1.0 - z * C.grav / C.Cpa / C.tsfcc1, C.Cpa / C.Rd);
}
__device__ __forceinline__ double exner(const DevConsts& C, double p) {  // les.F:10775-10784
  return pow(C.surf_p / p, -C.Rd / C.Cpa);
}

// Air density and the potential-to-actual temperature conversion of Tf
// (particles.f90:3064-3072).
__device__ __forceinline__ double air_rho(const DevConsts& C, const DevFlags& fl, double z, double& Tf) {
  if (fl.iexner == 1) {
    double pb = func_p_base(C, z);
    Tf = Tf * exner(C, pb);
    return pb / C.Rd / Tf;
  }
  return C.surf_rho;
}

// particle_bcs_nonperiodic, icase=0 (2683-2812), then particle_bcs_periodic
// (2814-2843).
__device__ __forceinline__ void apply_bcs(const DevGrid& g, const double* s_z, Drop& d) {
  const double top = s_z[g.nnz] - d.r;
  const double bot = 0.0 + d.r;
  if (d.z > top) {
    d.z = top - (d.z - top);
    d.vz = -d.vz;
  } else if (d.z < bot) {
    d.z = bot + (bot - d.z);
    d.vz = -d.vz;
  }
  if (d.x > g.xl) d.x = d.x - g.xl;
  else if (d.x < 0) d.x = g.xl + d.x;
  if (d.y > g.yl) d.y = d.y - g.yl;
  else if (d.y < 0) d.y = g.yl + d.y;
}

// ----------------------------------------------------------- Newton solve

// FP64 reciprocal: hardware estimate + two Newton-Raphson refinements.
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// FP32 quotient through the hardware reciprocal (a few ulp; the mixed step's
// FP32 drag, Koehler guess and q* carry FP32 rounding anyway).
__device__ __forceinline__ float fdiv32(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// mod_magnus with the quotient through rcp64 (closed-form and mixed steps).
__device__ __forceinline__ double mod_magnus_fast(double T) {
  const double t = T - 273.15;
  return 610.94 * exp(17.6257 * t * rcp64(243.04 + t));
}

// Inputs of ie_vrt_nd that stay fixed during one solve.
struct NewtonIn {
  double r, Tp, Tf, qinf, m_s, kappa, rhoa, esa, taup0, h, diffnorm;
  // reference-order invariants: 18*rhoa*nuf, (kappa*m_s*rhow)/rhos, taup0/r,
  // r/taup0, taup0/Tp, 1/Tf
  double c18, ksol, t0_r, r_t0, t0_Tp, iTf;
};

// ie_vrt_nd (particles.f90:4263-4327) in the reference's operation order.
__device__ __forceinline__ void resid_ref(const DevConsts& C, const NewtonIn& P, double tempr,
                                          double tempt, double F[2]) {
  const double rnext = tempr * P.r;
  const double Tnext = tempt * P.Tp;
  const double dnext = rnext * 2.;
  const double VolP = (2. / 3.) * C.pi2 * (rnext * rnext * rnext);
  const double rhop = (P.m_s + VolP * C.rhow) / VolP;
  const double Rep = dnext * P.diffnorm / C.nuf;
  const double taup = (rhop * (dnext * dnext)) / P.c18;
  const double qstr = (C.Mw / (C.Ru * Tnext * P.rhoa)) * P.esa *
                      exp((C.LvMw_Ru * (P.iTf - (1. / Tnext))) + (C.Mw2Gam / (C.Ru_rhow * rnext * Tnext)) -
                          (P.ksol / (VolP * rhop - P.m_s)));
  const double sRe = sqrt(Rep);
  const double Shp = 2. + 0.6 * sRe * C.Sc13;
  double rprime = (1. / 9.) * (Shp / C.Sc) * (rhop / C.rhow) * (rnext / taup) * (P.qinf - qstr);
  rprime = rprime * P.t0_r;
  const double Nup = 2. + 0.6 * sRe * C.Pra13;
  double Tprime = -((1. / 3.) * (Nup / C.Pra) * C.CpaCpp * (rhop / C.rhow) * (1. / taup) * (Tnext - P.Tf)) +
                  C.Lv3 * (1. / (rnext * C.Cpp)) * rprime * P.r_t0;
  Tprime = Tprime * P.t0_Tp;
  F[0] = rnext / P.r - 1.0 - P.h * rprime;
  F[1] = Tnext / P.Tp - 1.0 - P.h * Tprime;
}

// Reference residual + jacob_approx_2d (4224-4252): forward differences with
// step 1e-12 around the residual, which the reference evaluates twice.
struct RefEval {
  const DevConsts& C;
  const NewtonIn& P;
  static constexpr bool kSquare = false;  // reference (J^T J)^-1 J^T step
  __device__ __forceinline__ void f(double tr, double tt, double F[2]) const { resid_ref(C, P, tr, tt, F); }
  __device__ __forceinline__ void fj(double tr, double tt, double F[2], double J[2][2]) const {
    const double diff = 1e-12;
    double fx[2];
    resid_ref(C, P, tr, tt, F);
    resid_ref(C, P, tr + diff, tt, fx);
    J[0][0] = (fx[0] - F[0]) / diff;
    J[1][0] = (fx[1] - F[1]) / diff;
    resid_ref(C, P, tr, tt + diff, fx);
    J[0][1] = (fx[0] - F[0]) / diff;
    J[1][1] = (fx[1] - F[1]) / diff;
  }
};

// The same equations in closed form with their exact Jacobian.  With
// rn = tr*r, Tn = tt*Tp the reference's groupings reduce to
//   h*rprime = X = hA*Sh*(qinf - q*)/rn,       hA = dt ρa ν / (2 Sc ρw r)
//   h*Tprime = Y = -hB*Nu*(Tn - Tf)/rn^2 + Cc*X/rn,
//              hB = 3/2 dt (Cpa/Cpp) ρa ν / (Pr ρw Tp),  Cc = 3 Lv r / (Cpp Tp)
//   q* = K/Tn * exp(E),  K = Mw esa / (Ru ρa),
//   E  = Lv Mw/Ru (Tn - Tf)/(Tf Tn) + 2 Mw Γ/(Ru ρw rn Tn) - κ m_s/(ρs V(rn))
// (using ρp V - m_s = ρw V; taup0 cancels), F = (tr - 1 - X, tt - 1 - Y).
struct AJIn {
  double r, Tp, Tf, qinf, K, csol, LMRiTf, sq, hA, hB, Cc;
};

struct AJEval {
  const DevConsts& C;
  const AJIn& P;
  static constexpr bool kSquare = true;  // J^-1 F, equal to the reference step
  __device__ __forceinline__ void eval(double tr, double tt, double F[2], double J[2][2], bool jac) const {
    const double rn = tr * P.r, Tn = tt * P.Tp;
    const double irn = rcp64(rn), iTn = rcp64(Tn);
    const double irn2 = irn * irn;
    const double s = P.sq * sqrt(rn);  // sqrt(Re)
    const double Sh = 2.0 + C.sc6 * s, Nu = 2.0 + C.pr6 * s;
    const double curv = C.Mw2Gam_Ru_rhow * irn * iTn;
    const double solu = P.csol * irn2 * irn;
    const double dT = Tn - P.Tf;
    const double E = P.LMRiTf * dT * iTn + curv - solu;
    const double qs = P.K * iTn * exp(E);
    const double dq = P.qinf - qs;
    const double X = P.hA * Sh * dq * irn;
    const double Y = -P.hB * Nu * dT * irn2 + P.Cc * X * irn;
    F[0] = tr - 1.0 - X;
    F[1] = tt - 1.0 - Y;
    if (!jac) return;
    const double dE_r = (3.0 * solu - curv) * irn;
    const double dE_T = (C.LvMw_Ru * iTn - curv) * iTn;
    const double dqs_r = qs * dE_r, dqs_T = qs * (dE_T - iTn);
    const double dSh = 0.5 * (Sh - 2.0) * irn, dNu = 0.5 * (Nu - 2.0) * irn;
    const double dX_r = P.hA * irn * (dSh * dq - Sh * dqs_r - Sh * dq * irn);
    const double dX_T = -P.hA * Sh * dqs_T * irn;
    const double dY_r = -P.hB * irn2 * (dNu * dT - 2.0 * Nu * dT * irn) + P.Cc * irn * (dX_r - X * irn);
    const double dY_T = -P.hB * Nu * irn2 + P.Cc * dX_T * irn;
    J[0][0] = 1.0 - P.r * dX_r;
    J[0][1] = -P.Tp * dX_T;
    J[1][0] = -P.r * dY_r;
    J[1][1] = 1.0 - P.Tp * dY_T;
  }
  __device__ __forceinline__ void f(double tr, double tt, double F[2]) const {
    double J[2][2];
    eval(tr, tt, F, J, false);
  }
  __device__ __forceinline__ void fj(double tr, double tt, double F[2], double J[2][2]) const {
    eval(tr, tt, F, J, true);
  }
};

// gauss_newton_2d (4071-4142): damped Gauss-Newton, backtracking while a
// component is negative or NaN, failure below 0.01 or on NaN, convergence on
// the undamped correction norm < 1e-8, at most 50 iterations.
template <class Ev>
__device__ __forceinline__ int gauss_newton(const Ev& ev, const double vec1[2], double vec2[2], int& iters) {
  int iterations = 0, flag = 0;
  double v1[2] = {vec1[0], vec1[1]};
  vec2[0] = v1[0];
  vec2[1] = v1[1];
  while (iterations < 50) {
    iterations++;
    double fv1[2], J[2][2], fJ[2][2];
    ev.fj(v1[0], v1[1], fv1, J);
    if constexpr (Ev::kSquare) {
      // (J^T J)^-1 J^T = J^-1 for the square system; det(J^T J) = det(J)^2
      const double detJ = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      if (fabs(detJ * detJ) < 1.0e-10) { flag = 1; break; }
      const double s = rcp64(detJ);
      fJ[0][0] = s * J[1][1];
      fJ[0][1] = -s * J[0][1];
      fJ[1][0] = -s * J[1][0];
      fJ[1][1] = s * J[0][0];
    } else {
      double fancy[2][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) fancy[i][j] = J[0][i] * J[0][j] + J[1][i] * J[1][j];
      const double det = fancy[0][0] * fancy[1][1] - fancy[0][1] * fancy[1][0];
      if (fabs(det) < 1.0e-10) { flag = 1; break; }
      const double s = 1. / det;
      const double inv00 = s * fancy[1][1], inv10 = s * (-fancy[1][0]);
      const double inv01 = s * (-fancy[0][1]), inv11 = s * fancy[0][0];
      fJ[0][0] = inv00 * J[0][0] + inv01 * J[0][1];
      fJ[0][1] = inv00 * J[1][0] + inv01 * J[1][1];
      fJ[1][0] = inv10 * J[0][0] + inv11 * J[0][1];
      fJ[1][1] = inv10 * J[1][0] + inv11 * J[1][1];
    }
    const double c0 = fJ[0][0] * fv1[0] + fJ[0][1] * fv1[1];
    const double c1 = fJ[1][0] * fv1[0] + fJ[1][1] * fv1[1];
    vec2[0] = v1[0] - c0;
    vec2[1] = v1[1] - c1;
    int counts = 0;
    double relax = 1.0;
    while ((vec2[0] < 0) || (vec2[1] < 0) || isnan(vec2[0])) {
      counts++;
      relax = relax * 0.5;
      vec2[0] = v1[0] - c0 * relax;
      vec2[1] = v1[1] - c1 * relax;
      if (counts > 10) break;
    }
    if (vec2[0] < 0.01 || vec2[1] < 0.01 || isnan(vec2[0]) || isnan(vec2[1])) { flag = 1; break; }
    if (Ev::kSquare ? (c0 * c0 + c1 * c1 < 1.0e-16) : (sqrt(c0 * c0 + c1 * c1) < 1.0e-8)) break;
    v1[0] = vec2[0];
    v1[1] = vec2[1];
  }
  if (iterations == 50) flag = 1;
  if (isnan(vec2[0]) || vec2[0] < 0 || isnan(vec2[1]) || vec2[1] < 0) flag = 1;
  iters = iterations;
  return flag;
}

// LV_solver (4143-4223): Levenberg-Marquardt, lambda 1e-3 halved on success
// and doubled otherwise, merit |(F1^2/2, F2^2/2)|.  vec2 keeps what the failed
// Gauss-Newton left, as the shared actual argument rt_zeroes does.
template <class Ev>
__device__ __forceinline__ int lv_solver(const Ev& ev, const double vec1[2], double vec2[2], int& iters) {
  int iterations = 0, flag = 0;
  double v1[2] = {vec1[0], vec1[1]};
  double lambda = 0.001;
  while (iterations < 1000) {
    iterations++;
    double J[2][2], fv1[2];
    ev.fj(v1[0], v1[1], fv1, J);
    double gm[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) gm[i][j] = (J[0][i] * J[0][j] + J[1][i] * J[1][j]) + lambda * (i == j ? 1.0 : 0.0);
    const double gradC0 = J[0][0] * fv1[0] + J[1][0] * fv1[1];
    const double gradC1 = J[0][1] * fv1[0] + J[1][1] * fv1[1];
    const double C0 = 0.5 * fv1[0] * fv1[0], C1 = 0.5 * fv1[1] * fv1[1];
    const double det = gm[0][0] * gm[1][1] - gm[0][1] * gm[1][0];
    if (fabs(det) < 1.0e-10) { flag = 1; break; }
    const double s = 1. / det;
    const double i00 = s * gm[1][1], i10 = s * (-gm[1][0]), i01 = s * (-gm[0][1]), i11 = s * gm[0][0];
    const double c0 = i00 * gradC0 + i01 * gradC1;
    const double c1 = i10 * gradC0 + i11 * gradC1;
    vec2[0] = v1[0] - c0;
    vec2[1] = v1[1] - c1;
    if (sqrt(c0 * c0 + c1 * c1) < 1.0e-8) break;
    if (vec2[0] < 0.0 || vec2[1] < 0.0 || isnan(vec2[0]) || isnan(vec2[1])) { flag = 1; break; }
    double fv2[2];
    ev.f(vec2[0], vec2[1], fv2);
    const double n0 = 0.5 * fv2[0] * fv2[0], n1 = 0.5 * fv2[1] * fv2[1];
    if (sqrt(n0 * n0 + n1 * n1) < sqrt(C0 * C0 + C1 * C1)) {
      v1[0] = vec2[0];
      v1[1] = vec2[1];
      lambda = lambda / 2.0;
    } else {
      lambda = lambda * 2.0;
    }
  }
  if (iterations == 1000) flag = 1;
  if (vec2[0] < 0 || vec2[1] < 0) flag = 1;
  iters = iterations;
  return flag;
}

// rad_solver2 (4328-4371): Koehler equilibrium radius from Cardano on
// r^3 + a r^2 + c = 0, in FP64.
__device__ __forceinline__ double rad_solver2(const DevConsts& C, double Tf, double rhoa, double qinf,
                                              double kappa, double m_s, double r, double esa, int& mflag) {
  mflag = 0;
  const double L = log((C.Ru * Tf * rhoa * qinf) / (C.Mw * esa));
  const double a = -((2 * C.Mw * C.Gam) / (C.Ru * C.rhow * Tf) / L);
  const double c = (kappa * m_s) / ((2.0 / 3.0) * C.pi2 * C.rhos) / L;
  const double Q = (a * a) / 9.0;
  const double R = (2.0 * (a * a * a) + 27.0 * c) / 54.0;
  const double Q3 = Q * Q * Q;
  const double M = R * R - Q3;
  double guess;
  if (M < 0) {
    const double theta = acos(R / sqrt(Q3));
    guess = -(2 * sqrt(Q) * cos((theta - C.pi2) / 3.0)) - a / 3.0;
    if (guess < 0) guess = -(2 * sqrt(Q) * cos((theta + C.pi2) / 3.0)) - a / 3.0;
  } else {
    const double S = -(R / fabs(R)) * cbrt(fabs(R) + sqrt(M));
    const double T = Q / S;
    guess = S + T - a / 3.0;
    if (guess < 0) {
      guess = r;
      mflag = 1;
    }
  }
  return guess;
}

// The same guess for the mixed-precision solve.  The cubic's coefficients,
// its discriminant M = R^2 - Q^3 and therefore the choice of branch are the
// FP64 ones.  Where the roots lie well apart (|M| > 1e-4 Q^3 with three real
// roots, M > 1e-4 R^2 with one), acos/cos or cbrt run in FP32 and one FP64
// Newton step on r^3 + a r^2 + c polishes the root back to FP64 accuracy;
// near the double root, where Cardano's argument loses its digits, the
// FP64 expression is used as it stands.
__device__ __forceinline__ double cubic_polish(double g, double a, double c) {
  const double f = g * g * (g + a) + c, fp = g * (3.0 * g + 2.0 * a);
  return fp != 0.0 ? g - f * rcp64(fp) : g;
}

// ln(A/B) for A, B > 0.  Near A = B (|u| <= 1/4 with u = (A-B)/(A+B), a
// saturation ratio in [0.6, 1.67]) through ln(A/B) = 2 atanh(u), whose series
// u (1 + u^2/3 + u^4/5 + ...) is summed to u^26 (the next term is below
// 1e-17 of the sum); A - B is exact there (Sterbenz), so the result carries a
// few ulps of L itself rather than the ulp of A/B that log(A/B) carries.
// Elsewhere log(A/B).
__device__ __forceinline__ double log_ratio(double A, double B) {
  const double u = (A - B) * rcp64(A + B);
  if (fabs(u) > 0.25) return log(A * rcp64(B));
  const double w = u * u;
  double p = 1.0 / 27;
  p = fma(p, w, 1.0 / 25);
  p = fma(p, w, 1.0 / 23);
  p = fma(p, w, 1.0 / 21);
  p = fma(p, w, 1.0 / 19);
  p = fma(p, w, 1.0 / 17);
  p = fma(p, w, 1.0 / 15);
  p = fma(p, w, 1.0 / 13);
  p = fma(p, w, 1.0 / 11);
  p = fma(p, w, 1.0 / 9);
  p = fma(p, w, 1.0 / 7);
  p = fma(p, w, 1.0 / 5);
  p = fma(p, w, 1.0 / 3);
  p = fma(p, w, 1.0);
  return 2.0 * u * p;
}

__device__ __forceinline__ double rad_solver2_mixed(const DevConsts& C, double Tf, double rhoa, double qinf,
                                                    double kappa, double m_s, double r, double esa, int& mflag) {
  mflag = 0;
  const double L = log_ratio(C.Ru * Tf * rhoa * qinf, C.Mw * esa);
  const double iL = rcp64(L);
  const double a = -(C.Mw2Gam_Ru_rhow * rcp64(Tf)) * iL;
  const double c = kappa * m_s * C.isol * iL;
  const double Q = (a * a) * (1.0 / 9.0);
  const double R = (2.0 * (a * a * a) + 27.0 * c) * (1.0 / 54.0);
  const double Q3 = Q * Q * Q;
  const double M = R * R - Q3;
  const double a3 = a * (1.0 / 3.0);
  double guess;
  if (M < 0) {
    const double sQ = sqrt(Q), sQ2 = 2.0 * sQ;
    if (-M > 1.0e-4 * Q3) {
      const float th = acosf((float)(R * rcp64(Q * sQ)));  // R / sqrt(Q^3)
      const float p2 = (float)C.pi2;
      guess = -sQ2 * (double)cosf((th - p2) * (1.0f / 3.0f)) - a3;
      if (guess < 0) guess = -sQ2 * (double)cosf((th + p2) * (1.0f / 3.0f)) - a3;
      guess = cubic_polish(guess, a, c);
    } else {
      const double theta = acos(R / sqrt(Q3));
      guess = -(sQ2 * cos((theta - C.pi2) / 3.0)) - a3;
      if (guess < 0) guess = -(sQ2 * cos((theta + C.pi2) / 3.0)) - a3;
    }
  } else {
    double S;
    if (M > 1.0e-4 * (R * R)) S = (double)cbrtf((float)(fabs(R) + sqrt(M)));
    else S = cbrt(fabs(R) + sqrt(M));
    S = (R > 0) ? -S : S;
    guess = S + Q * rcp64(S) - a3;
    if (M > 1.0e-4 * (R * R)) guess = cubic_polish(guess, a, c);
    if (guess < 0) {
      guess = r;
      mflag = 1;
    }
  }
  return guess;
}

// Radius/temperature block of particle_update_BE (3103-3178) on a droplet
// whose velocity is already the new one.  Updates d.r, d.Tp, activation
// bookkeeping and the tally.  PREC_FP64: reference residual and FD
// Jacobian.  Otherwise the closed-form residual with its exact Jacobian.
// Every precision starts from the reference's Koehler guess to FP64
// accuracy (the mixed solve through rad_solver2_mixed, whose branch choice
// and result are the FP64 ones): the BE equations can have more than one
// solution near activation, and which one Newton (or Levenberg-Marquardt)
// reaches depends on where it starts; a guess formed wholly in FP32
// (Cardano's discriminant cancels near the Koehler maximum) sent a few
// hundred droplets per step of the stiff activation case into failed solves.
template <int PREC>
__device__ __forceinline__ void be_microphysics(const DevConsts& C, const DevFlags& fl, Drop& d,
                                                double Tf, double qinf, double rhoa, double diffnorm,
                                                double Rep, double esa, Tally& tl) {
  const double dt = C.dt;
  if (fl.ievap == 1 && qinf > 0.0) {
    int mflag = 0, flag, it1 = 0, it2 = 0;
    double rt_start[2] = {1.0, 1.0}, rt_zeroes[2];
    // reciprocals of the closed-form solve (the reference-order FP64 solve divides)
    const double ir = (PREC == PREC_FP64) ? 0.0 : rcp64(d.r);
    const double iTp = (PREC == PREC_FP64) ? 0.0 : rcp64(d.Tp);
    {  // the reference's start: the Koehler guess
      const double guess = (PREC == PREC_MIXED)
                               ? rad_solver2_mixed(C, Tf, rhoa, qinf, d.kappa, d.m_s, d.r, esa, mflag)
                               : rad_solver2(C, Tf, rhoa, qinf, d.kappa, d.m_s, d.r, esa, mflag);
      if (mflag == 0) {
        rt_start[0] = (PREC == PREC_FP64) ? guess / d.r : guess * ir;
        rt_start[1] = (PREC == PREC_FP64) ? Tf / d.Tp : Tf * iTp;
      }
    }
    if (PREC == PREC_FP64) {
      const double ri = C.radius_init;
      NewtonIn P;
      P.r = d.r;
      P.Tp = d.Tp;
      P.Tf = Tf;
      P.qinf = qinf;
      P.m_s = d.m_s;
      P.kappa = d.kappa;
      P.rhoa = rhoa;
      P.esa = esa;
      P.taup0 = (((d.m_s) / ((2. / 3.) * C.pi2 * (ri * ri * ri)) + C.rhow) * ((ri * 2) * (ri * 2))) /
                (18 * rhoa * C.nuf);
      P.h = dt / P.taup0;
      P.diffnorm = diffnorm;
      P.c18 = 18.0 * rhoa * C.nuf;
      P.ksol = d.kappa * d.m_s * C.rhow / C.rhos;
      P.t0_r = P.taup0 / d.r;
      P.r_t0 = d.r / P.taup0;
      P.t0_Tp = P.taup0 / d.Tp;
      P.iTf = 1. / Tf;
      const RefEval ev{C, P};
      flag = gauss_newton(ev, rt_start, rt_zeroes, it1);
      tl.gn_iters += it1;
      tl.solves += 1;
      tl.hist = it1;
      if (flag == 1) {
        tl.num100 += 1;
        tl.hist = 0;
        flag = lv_solver(ev, rt_start, rt_zeroes, it2);
        tl.lm_iters += it2;
      }
    } else {
      AJIn P;
      P.r = d.r;
      P.Tp = d.Tp;
      P.Tf = Tf;
      P.qinf = qinf;
      const double irhoa = (fl.iexner == 1) ? rcp64(rhoa) : C.isurf_rho;  // rhoa = surf_rho unless iexner
      P.K = C.Mw_Ru * esa * irhoa;
      P.csol = d.kappa * d.m_s * C.isol;
      P.LMRiTf = C.LvMw_Ru * rcp64(Tf);
      P.sq = sqrt(diffnorm * (2.0 * C.inuf));
      P.hA = C.hA0 * rhoa * ir;
      P.hB = C.hB0 * rhoa * iTp;
      P.Cc = C.Lv3_Cpp * d.r * iTp;
      const AJEval ev{C, P};
      flag = gauss_newton(ev, rt_start, rt_zeroes, it1);
      tl.gn_iters += it1;
      tl.solves += 1;
      tl.hist = it1;
      if (flag == 1) {
        tl.num100 += 1;
        tl.hist = 0;
        flag = lv_solver(ev, rt_start, rt_zeroes, it2);
        tl.lm_iters += it2;
      }
    }
    if ((flag == 1) || isnan(rt_zeroes[0]) || (rt_zeroes[0] * d.r < 0) || isnan(rt_zeroes[1]) ||
        (rt_zeroes[1] < 0) || (rt_zeroes[0] * d.r > 1.0e-2)) {
      tl.numimpos += 1;
      rt_zeroes[0] = 1.0;
      rt_zeroes[1] = Tf / d.Tp;
    }
    if (d.r > d.rc && d.r * rt_zeroes[0] < d.rc) {
      tl.denum += 1;
    } else if (d.r < d.rc && d.r * rt_zeroes[0] > d.rc) {
      tl.actnum += 1;
      d.numact = d.numact + 1.0;
      d.actres = 0.0;
    }
    d.r = rt_zeroes[0] * d.r;
    d.Tp = rt_zeroes[1] * d.Tp;
  } else {
    const double r = d.r;
    const double Volp = C.pi2 * 2.0 / 3.0 * (r * r * r);
    const double rhop = (d.m_s + Volp * C.rhow) / Volp;
    const double taup_i = 18.0 * rhoa * C.nuf / rhop / ((2.0 * r) * (2.0 * r));
    const double Nup = 2.0 + 0.6 * sqrt(Rep) * C.Pra13;
    const double tmp_coeff = Nup / 3.0 / C.Pra * C.CpaCpp * rhop / C.rhow * taup_i;
    d.Tp = (d.Tp + tmp_coeff * dt * Tf) / (1 + dt * tmp_coeff);
  }
}

// qstar of the updated droplet (particles.f90:3192-3201), FP64.
__device__ __forceinline__ double qstar_f64(const DevConsts& C, const Drop& d, double Tf, double rhoa,
                                            double einf) {
  const double r = d.r;
  const double Volp = C.pi2 * 2.0 / 3.0 * (r * r * r);
  const double rhop = (d.m_s + Volp * C.rhow) / Volp;
  return (C.Mw / (C.Ru * d.Tp * rhoa)) * einf *
         exp(((C.Lv * C.Mw / C.Ru) * ((1. / Tf) - (1. / d.Tp))) + ((2. * C.Mw * C.Gam) / (C.Ru * C.rhow * r * d.Tp)) -
             ((d.kappa * d.m_s * C.rhow / C.rhos) / (Volp * rhop - d.m_s)));
}

// The same diagnostic in FP32 (mixed precision): 1/Tf - 1/Tp formed as
// (Tp - Tf)/(Tf Tp) from the FP64 difference.
__device__ __forceinline__ double qstar_f32(const DevConsts& C, const Drop& d, double Tf, double rhoa,
                                            double einf) {
  const float r = (float)d.r, Tp = (float)d.Tp;
  const float V = (float)C.pi2_23 * r * r * r;
  const float E = (float)C.LvMw_Ru * fdiv32((float)(d.Tp - Tf), (float)Tf * Tp) +
                  fdiv32((float)C.Mw2Gam_Ru_rhow, r * Tp) - fdiv32((float)(d.kappa * d.m_s * C.irhos), V);
  return (double)(fdiv32((float)C.Mw_Ru, Tp * (float)rhoa) * (float)einf * __expf(E));
}

// ---------------------------------------------------------- BE droplet

// What the radius/temperature solve of one droplet needs from the advance.
struct BeHand {
  double Tf, qinf, rhoa;
  double diffnorm;  // |u_f - v| with the new velocity (ie_vrt_nd's Reynolds number)
  double Rep;       // Reynolds number of the pre-step velocity (no-evaporation branch)
};

// First half of one droplet of particle_update_BE (3047-3100): interpolation,
// it<=1 velocity set-up, air density, then the implicit velocity and the
// position advanced with the old velocity.  PREC_MIXED: drag and the velocity
// update in FP32, positions in FP64.
template <int PREC, int ILIN>
__device__ __forceinline__ BeHand be_advance(const DevGrid& g, const DevConsts& C, const DevFlags& fl,
                                             const DevFields& F, const double* s_z, const double* s_zz,
                                             Drop& d, int it, Diag& o, int kguess,
                                             const FieldBox* bx = nullptr) {
  const double dt = C.dt;
  o.vold[0] = d.vx;
  o.vold[1] = d.vy;
  o.vold[2] = d.vz;
  o.Tpold = d.Tp;
  o.rold = d.r;
  const Cell c = kguess >= 0 ? cell_near(g, s_z, s_zz, d.x, d.y, d.z, kguess) : cell_of(g, s_z, s_zz, d.x, d.y, d.z);
  // minloc(z, z > xp3i) - 1 of the pre-step height (flux statistics)
  o.kpt = pack_levels(c.kpt, (c.kwpt == -2) ? -1 : c.kwpt + 1);
  Fluid f;
  if (ILIN != 1) f = gather6_f64(g, F, s_z, s_zz, d.x, d.y, d.z, c);
  else if (bx) f = gather_lin_staged<PREC == PREC_MIXED>(g, F, *bx, s_z, s_zz, d.x, d.y, d.z, c);
  else if (PREC == PREC_MIXED) f = gather_lin_mixed<true>(g, F, s_z, s_zz, d.x, d.y, d.z, c);
  else f = gather_lin_f64<true>(g, F, s_z, s_zz, d.x, d.y, d.z, c);
  if (it <= 1) {
    d.vx = f.u[0];
    d.vy = f.u[1];
    d.vz = f.u[2];
  }
  BeHand h;
  double Tf = f.T;
  h.rhoa = air_rho(C, fl, d.z, Tf);
  h.Tf = Tf;
  h.qinf = f.q;
  const double r = d.r;
  if (PREC == PREC_MIXED) {
    const float rf = (float)r;
    const float a0 = (float)(d.vx - f.u[0]), a1 = (float)(d.vy - f.u[1]), a2 = (float)(d.vz - f.u[2]);
    const float Volp = (float)C.pi2_23 * rf * rf * rf;
    const float rhopf = fdiv32((float)d.m_s + Volp * (float)C.rhow, Volp);
    const float tpi = fdiv32((float)(h.rhoa * C.nuf18), rhopf * ((2.0f * rf) * (2.0f * rf)));
    h.Rep = (double)(2.0f * rf * sqrtf(a0 * a0 + a1 * a1 + a2 * a2) * (float)C.inuf);
    d.x = d.x + dt * d.vx;
    d.y = d.y + dt * d.vy;
    d.z = d.z + dt * d.vz;
    const float dtf = (float)dt, iden = __builtin_amdgcn_rcpf(1.0f + dtf * tpi);
    const float nx = ((float)d.vx + tpi * dtf * (float)f.u[0] + dtf * (float)C.g[0]) * iden;
    const float ny = ((float)d.vy + tpi * dtf * (float)f.u[1] + dtf * (float)C.g[1]) * iden;
    const float nz = ((float)d.vz + tpi * dtf * (float)f.u[2] + dtf * (float)C.g[2]) * iden;
    d.vx = nx;
    d.vy = ny;
    d.vz = nz;
    const float b0 = (float)f.u[0] - nx, b1 = (float)f.u[1] - ny, b2 = (float)f.u[2] - nz;
    h.diffnorm = (double)sqrtf(b0 * b0 + b1 * b1 + b2 * b2);
  } else {
    const double a0 = d.vx - f.u[0], a1 = d.vy - f.u[1], a2 = d.vz - f.u[2];
    const double diffnorm = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    const double Volp = C.pi2 * 2.0 / 3.0 * (r * r * r);
    const double rhop = (d.m_s + Volp * C.rhow) / Volp;
    const double taup_i = 18.0 * h.rhoa * C.nuf / rhop / ((2.0 * r) * (2.0 * r));
    h.Rep = 2.0 * r * diffnorm / C.nuf;
    d.x = d.x + dt * d.vx;
    d.y = d.y + dt * d.vy;
    d.z = d.z + dt * d.vz;
    const double den = (1 + dt * 1.0 * taup_i);
    d.vx = (d.vx + taup_i * dt * 1.0 * f.u[0] + dt * C.g[0]) / den;
    d.vy = (d.vy + taup_i * dt * 1.0 * f.u[1] + dt * C.g[1]) / den;
    d.vz = (d.vz + taup_i * dt * 1.0 * f.u[2] + dt * C.g[2]) / den;
    const double b0 = f.u[0] - d.vx, b1 = f.u[1] - d.vy, b2 = f.u[2] - d.vz;
    h.diffnorm = sqrt(b0 * b0 + b1 * b1 + b2 * b2);
  }
  o.uf[0] = f.u[0];
  o.uf[1] = f.u[1];
  o.uf[2] = f.u[2];
  o.Tf = Tf;
  o.qinf = f.q;
  return h;
}

// The flux statistics of particle_update_BE (3208-3238) for one droplet,
// after its advance and radius/temperature update and before the walls.
// With kc0 / kc1 the first w level above the pre-step / new height
// (minloc(z, z > .) - 1; kc1 = nnz+1 above zl = z(nnz) and 0 below the
// ground), a droplet that moved up adds mult, its mass rhop*Volp*mult and its
// enthalpy rhop*Volp*Cpp*Tp*mult to each level iz = kc0..kc1-1, one that moved
// down subtracts them from iz = kc1..kc0-1 (kc is monotone in the height, so
// the direction is the sign of kc1 - kc0).  flux = pflux(0:nnz) |
// pmassflux(0:nnz) | penegflux(0:nnz); levels outside 0..nnz are skipped.
// Few droplets cross a level in one step, and they add with FP64 atomics.
__device__ __forceinline__ void be_flux(const DevGrid& g, const DevConsts& C, const double* s_z, const Drop& d,
                                        int kc0, int64_t mult, double* flux) {
  const int len = g.nnz + 2, nlev = g.nnz + 1;
  int kc1;
  if (d.z > s_z[g.nnz]) kc1 = g.nnz + 1;
  else if (d.z < 0.0) kc1 = 0;
  else if (kc0 >= 1 && kc0 < len && s_z[kc0 - 1] <= d.z && d.z < s_z[kc0]) kc1 = kc0;
  else {
    const int p = first_gt(s_z, len, g.nsteps, d.z);
    kc1 = (p == len) ? -1 : p;
  }
  if (kc1 == kc0) return;
  const double r = d.r;
  const double Volp = C.pi2 * 2.0 / 3.0 * (r * r * r);
  const double rhop = (d.m_s + Volp * C.rhow) / Volp;
  const double m = (double)mult;
  const double sgn = kc1 > kc0 ? 1.0 : -1.0;
  const int lo = kc1 > kc0 ? kc0 : kc1, hi = kc1 > kc0 ? kc1 : kc0;
  for (int iz = max(lo, 0); iz < min(hi, nlev); ++iz) {
    atomicAdd(&flux[iz], sgn * m);
    atomicAdd(&flux[nlev + iz], sgn * (rhop * Volp * m));
    atomicAdd(&flux[2 * nlev + iz], sgn * (rhop * Volp * C.Cpp * d.Tp * m));
  }
}

// Second half (3103-3238 and the wall reflection of particle_bcs_nonperiodic):
// radius/temperature solve, q* (a diagnostic of the record: formed when
// with_qstar, i.e. when the caller stores it), residence times, flux
// statistics (flux not null; levels: the packed levels be_advance left), then
// walls at the new radius.
template <int PREC, bool LAZY = false>
__device__ __forceinline__ void be_finish(const DevGrid& g, const DevConsts& C, const DevFlags& fl,
                                          const double* s_z, Drop& d, const BeHand& h, double& qstar,
                                          Tally& tl, int levels, double* flux = nullptr, int64_t mult = 1,
                                          bool with_qstar = true) {
  const double dt = C.dt;
  const double esa = (PREC == PREC_FP64) ? mod_magnus(h.Tf) : mod_magnus_fast(h.Tf);
  be_microphysics<PREC>(C, fl, d, h.Tf, h.qinf, h.rhoa, h.diffnorm, h.Rep, esa, tl);
  qstar = 0.0;
  if (with_qstar)
    qstar = (PREC == PREC_MIXED) ? qstar_f32(C, d, h.Tf, h.rhoa, esa) : qstar_f64(C, d, h.Tf, h.rhoa, esa);
  if (!LAZY) {  // LAZY: left to the end of the call (k_res_settle)
    d.res = d.res + dt;
    d.actres = d.actres + dt;
  }
  if (flux) be_flux(g, C, s_z, d, packed_kc0(levels), mult, flux);
  apply_bcs(g, s_z, d);
}

// One droplet of particle_update_BE (3043-3238) plus its boundary conditions.
template <int PREC, int ILIN, bool LAZY = false>
__device__ __forceinline__ void be_droplet(const DevGrid& g, const DevConsts& C, const DevFlags& fl,
                                           const DevFields& F, const double* s_z, const double* s_zz,
                                           Drop& d, int it, Diag& o, Tally& tl, int kguess = -1,
                                           double* flux = nullptr, int64_t mult = 1,
                                           const FieldBox* bx = nullptr, bool with_qstar = true) {
  const BeHand h = be_advance<PREC, ILIN>(g, C, fl, F, s_z, s_zz, d, it, o, kguess, bx);
  be_finish<PREC, LAZY>(g, C, fl, s_z, d, h, o.qstar, tl, o.kpt, flux, mult, with_qstar);
  o.kpt = packed_kpt(o.kpt);
}

// ---------------------------------------------------------- RK3 droplet

// One droplet of one stage of particle_update_rk3 (2879-2973) plus boundary
// conditions.  The stage reads the stored right-hand sides (o.xrhs ...) of
// the previous stage.  MIXED: drag, Reynolds/Nusselt/Sherwood numbers, q* and
// the explicit right-hand sides in FP32; positions, velocities, temperature
// and radius accumulate in FP64.
template <int PREC, int ILIN>
__device__ __forceinline__ void rk3_droplet(const DevGrid& g, const DevConsts& C, const DevFlags& fl,
                                            const DevFields& F, const double* s_z, const double* s_zz,
                                            Drop& d, int istage, int it, Diag& o, int kguess = -1,
                                            const FieldBox* bx = nullptr) {
  const double dt = C.dt;
  const double zs = C.zetas[istage - 1], gs = C.gama[istage - 1];
  if (istage == 1) {
    o.vold[0] = d.vx;
    o.vold[1] = d.vy;
    o.vold[2] = d.vz;
    o.Tpold = d.Tp;
    o.rold = d.r;
  }
  const Cell c = kguess >= 0 ? cell_near(g, s_z, s_zz, d.x, d.y, d.z, kguess) : cell_of(g, s_z, s_zz, d.x, d.y, d.z);
  o.kpt = c.kpt;
  Fluid f;
  if (ILIN != 1) f = gather6_f64(g, F, s_z, s_zz, d.x, d.y, d.z, c);
  else if (bx) f = gather_lin_staged<PREC == PREC_MIXED>(g, F, *bx, s_z, s_zz, d.x, d.y, d.z, c);
  else if (PREC == PREC_MIXED) f = gather_lin_mixed<true>(g, F, s_z, s_zz, d.x, d.y, d.z, c);
  else f = gather_lin_f64<true>(g, F, s_z, s_zz, d.x, d.y, d.z, c);
  double Tf = f.T;
  const double rhoa = air_rho(C, fl, d.z, Tf);
  if (it <= 1) {
    d.vx = f.u[0];
    d.vy = f.u[1];
    d.vz = f.u[2];
    d.Tp = Tf;
  }
  const double r = d.r;
  double xtmp[3], vtmp[3];
  xtmp[0] = d.x + dt * zs * o.xrhs[0];
  xtmp[1] = d.y + dt * zs * o.xrhs[1];
  xtmp[2] = d.z + dt * zs * o.xrhs[2];
  vtmp[0] = d.vx + dt * zs * o.vrhs[0];
  vtmp[1] = d.vy + dt * zs * o.vrhs[1];
  vtmp[2] = d.vz + dt * zs * o.vrhs[2];
  double Tptmp = d.Tp + dt * zs * o.Tprhs_s;
  Tptmp = Tptmp + dt * zs * o.Tprhs_L;
  const double radiustmp = r + dt * zs * o.radrhs;
  o.xrhs[0] = d.vx;
  o.xrhs[1] = d.vy;
  o.xrhs[2] = d.vz;
  if (PREC == PREC_MIXED) {
    const float rf = (float)r;
    const float a0 = (float)(d.vx - f.u[0]), a1 = (float)(d.vy - f.u[1]), a2 = (float)(d.vz - f.u[2]);
    const float Rep = 2.0f * rf * sqrtf(a0 * a0 + a1 * a1 + a2 * a2) * (float)C.inuf;
    const float Volp = (float)C.pi2_23 * rf * rf * rf;
    const float rhop = ((float)d.m_s + Volp * (float)C.rhow) / Volp;
    const float taup_i = fdiv32((float)(rhoa * C.nuf18), rhop * ((2.0f * rf) * (2.0f * rf)));
    const float corrfac = 1.0f + 0.15f * __powf(Rep, 0.687f);
    const float sRe = sqrtf(Rep);
    const float Nup = 2.0f + 0.6f * sRe * (float)C.Pra13;
    const float Shp = 2.0f + 0.6f * sRe * (float)C.Sc13;
    const float einf = (float)mod_magnus(Tf);
    const float Tp = (float)d.Tp;
    const float E = (float)C.LvMw_Ru * fdiv32((float)(d.Tp - Tf), (float)Tf * Tp) +
                    fdiv32((float)C.Mw2Gam_Ru_rhow, rf * Tp) - fdiv32((float)(d.kappa * d.m_s * C.irhos), Volp);
    const float qstar = fdiv32((float)C.Mw_Ru * einf * __expf(E), Tp * (float)rhoa);
    o.qstar = qstar;
    const float ct = corrfac * taup_i;
    o.vrhs[0] = (double)(ct * (float)(f.u[0] - d.vx) + (float)C.g[0]);
    o.vrhs[1] = (double)(ct * (float)(f.u[1] - d.vy) + (float)C.g[1]);
    o.vrhs[2] = (double)(ct * (float)(f.u[2] - d.vz) + (float)C.g[2]);
    const float radrhs = (fl.ievap == 1)
                             ? Shp / 9.0f / (float)C.Sc * rhop / (float)C.rhow * rf * taup_i * (float)(f.q - (double)qstar)
                             : 0.0f;
    o.radrhs = radrhs;
    o.Tprhs_s = (double)(-(Nup / 3.0f / (float)C.Pra * (float)C.CpaCpp * rhop / (float)C.rhow * taup_i * (float)(d.Tp - Tf)));
    o.Tprhs_L = (double)(fdiv32((float)C.Lv3_Cpp, rf) * radrhs);
  } else {
    const double a0 = d.vx - f.u[0], a1 = d.vy - f.u[1], a2 = d.vz - f.u[2];
    const double diffnorm = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    const double Rep = 2.0 * r * diffnorm / C.nuf;
    const double Volp = C.pi2 * 2.0 / 3.0 * (r * r * r);
    const double rhop = (d.m_s + Volp * C.rhow) / Volp;
    const double taup_i = 18.0 * rhoa * C.nuf / rhop / ((2.0 * r) * (2.0 * r));
    const double corrfac = (1.0 + 0.15 * pow(Rep, 0.687));
    const double sRe = sqrt(Rep);
    const double Nup = 2.0 + 0.6 * sRe * C.Pra13;
    const double Shp = 2.0 + 0.6 * sRe * C.Sc13;
    const double einf = mod_magnus(Tf);
    const double Eff_C = 2.0 * C.Mw * C.Gam / (C.Ru * C.rhow * r * d.Tp);
    const double Eff_S = d.kappa * d.m_s * C.rhow / C.rhos / (Volp * rhop - d.m_s);
    const double estar = einf * exp(C.Mw * C.Lv / C.Ru * (1.0 / Tf - 1.0 / d.Tp) + Eff_C - Eff_S);
    const double qstar = C.Mw / C.Ru * estar / d.Tp / rhoa;
    o.qstar = qstar;
    o.vrhs[0] = corrfac * taup_i * (f.u[0] - d.vx) + C.g[0];
    o.vrhs[1] = corrfac * taup_i * (f.u[1] - d.vy) + C.g[1];
    o.vrhs[2] = corrfac * taup_i * (f.u[2] - d.vz) + C.g[2];
    o.radrhs = (fl.ievap == 1) ? Shp / 9.0 / C.Sc * rhop / C.rhow * r * taup_i * (f.q - qstar) : 0.0;
    o.Tprhs_s = -(Nup / 3.0 / C.Pra * C.CpaCpp * rhop / C.rhow * taup_i * (d.Tp - Tf));
    o.Tprhs_L = 3.0 * C.Lv / C.Cpp / r * o.radrhs;
  }
  d.x = xtmp[0] + dt * gs * o.xrhs[0];
  d.y = xtmp[1] + dt * gs * o.xrhs[1];
  d.z = xtmp[2] + dt * gs * o.xrhs[2];
  d.vx = vtmp[0] + dt * gs * o.vrhs[0];
  d.vy = vtmp[1] + dt * gs * o.vrhs[1];
  d.vz = vtmp[2] + dt * gs * o.vrhs[2];
  d.Tp = Tptmp + dt * gs * o.Tprhs_s;
  d.Tp = d.Tp + dt * gs * o.Tprhs_L;
  d.r = radiustmp + dt * gs * o.radrhs;
  if (istage == 3) d.res = d.res + dt;
  o.uf[0] = f.u[0];
  o.uf[1] = f.u[1];
  o.uf[2] = f.u[2];
  o.Tf = Tf;
  o.qinf = f.q;
  apply_bcs(g, s_z, d);
}

// -------------------------------------------------------- coupling amounts

// The amounts below for a mixed-precision step, which stores them in FP32:
// the changes over the step (v - v_old, r - r_old, Tp - Tp_old) and Tf - Tp
// are formed in FP64, everything after them in FP32 with hardware
// reciprocals.  The enthalpy amount is vap * Cpv/Cpa * (Tf - Tp) * mult, the
// reference's -A*Tp + A*Tf without the cancellation of its two terms.
__device__ __forceinline__ void coupling_amounts_f32(const DevConsts& C, const DevFlags& fl, const Drop& d,
                                                     double Tf, const double vold[3], double Tpold,
                                                     double rold, int64_t mult, double a[6]) {
  const float idt = (float)C.idt;
  const float r = (float)d.r;
  const float radrhs = (float)(d.r - rold) * idt;
  const float V = (float)C.pi2_23 * (r * r * r);
  const float rhop = (float)d.m_s * __builtin_amdgcn_rcpf(V) + (float)C.rhow;
  const double rhoa = (fl.iexner == 1) ? func_p_base(C, d.z) / C.Rd / Tf : C.surf_rho;
  const float irhoa = (fl.iexner == 1) ? (float)rcp64(rhoa) : (float)C.isurf_rho;
  const float rm = (float)mult;
  const float vap = (float)(C.rhow * C.pi2 * 2) * irhoa * (r * r) * radrhs;
  for (int k = 0; k < 6; ++k) a[k] = 0.0;
  if (fl.icouple == 1) {
    const float pm = rhop * V * irhoa;
    a[0] = (double)(-(pm * ((float)(d.vx - vold[0]) * idt - (float)C.g[0]) + vap * (float)d.vx) * rm);
    a[1] = (double)(-(pm * ((float)(d.vy - vold[1]) * idt - (float)C.g[1]) + vap * (float)d.vy) * rm);
    a[2] = (double)(-(pm * ((float)(d.vz - vold[2]) * idt - (float)C.g[2]) + vap * (float)d.vz) * rm);
  }
  if (fl.iTcouple == 1) {
    const float ir = __builtin_amdgcn_rcpf(r);
    const float Tprhs_s = (float)(d.Tp - Tpold) * idt - (float)C.Lv3_Cpp * ir * radrhs;
    const float taup_i = 4.5f * (float)(rhoa * C.nuf) * __builtin_amdgcn_rcpf(rhop) * ir * ir;
    const float itp = __builtin_amdgcn_rcpf(rhop * (float)C.CpaCpp * taup_i);
    a[3] = (double)(-(Tprhs_s * 6.0f * (float)C.rhow * itp * (float)(C.pi2 / 2.0) * r * (float)C.nuf) * rm);
  }
  if (fl.iHcouple == 1) {
    a[4] = (double)(-vap * rm);
    a[5] = (double)(vap * (float)C.Cpv_Cpa * (float)(Tf - d.Tp) * rm);
  }
}

// Per-droplet source amounts of particle_coupling_update (1187-1264), in the
// reference's operations: the node contribution is amount * wtt / dV.
// a[0..2] momentum (icouple), a[3] heat (iTcouple), a[4] vapour and a[5]
// enthalpy (iHcouple).
__device__ __forceinline__ void coupling_amounts(const DevConsts& C, const DevFlags& fl, const Drop& d,
                                                 double Tf, const double vold[3], double Tpold,
                                                 double rold, int64_t mult, double a[6]) {
  const double dt = C.dt;
  const double r = d.r;
  const double V = C.pi2_23 * (r * r * r);  // C.pi2 * 2.0 / 3.0 * (r * r * r)
  const double vrhs0 = (d.vx - vold[0]) / dt;
  const double vrhs1 = (d.vy - vold[1]) / dt;
  const double vrhs2 = (d.vz - vold[2]) / dt;
  const double radrhs = (r - rold) / dt;
  const double Tprhs_L = 3.0 * C.Lv / C.Cpp / r * radrhs;
  const double Tprhs_s = (d.Tp - Tpold) / dt - Tprhs_L;
  const double rhoa = (fl.iexner == 1) ? func_p_base(C, d.z) / C.Rd / Tf : C.surf_rho;
  const double rhop = (d.m_s + V * C.rhow) / V;
  const double taup_i = 18.0 * rhoa * C.nuf / rhop / ((2.0 * r) * (2.0 * r));
  const double irhoa = 1.0 / rhoa;
  const double partmass = rhop * 2.0 / 3.0 * C.pi2 * (r * r * r);
  const double rm = (double)mult;
  const double vap = C.rhow * irhoa * C.pi2 * 2 * (r * r) * radrhs;
  for (int k = 0; k < 6; ++k) a[k] = 0.0;
  if (fl.icouple == 1) {
    a[0] = -(partmass * irhoa * (vrhs0 - C.g[0]) + vap * d.vx) * rm;
    a[1] = -(partmass * irhoa * (vrhs1 - C.g[1]) + vap * d.vy) * rm;
    a[2] = -(partmass * irhoa * (vrhs2 - C.g[2]) + vap * d.vz) * rm;
  }
  if (fl.iTcouple == 1) {
    const double itp = 1.0 / (rhop * C.CpaCpp * taup_i);
    a[3] = -(Tprhs_s * 6.0 * C.rhow * itp * (C.pi2 / 2.0) * r * C.nuf) * rm;
  }
  if (fl.iHcouple == 1) {
    a[4] = -vap * rm;
    a[5] = (-vap * C.Cpv / C.Cpa * d.Tp + vap * C.Cpv / C.Cpa * Tf) * rm;
  }
}

}  // namespace ntlp
