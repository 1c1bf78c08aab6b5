// device_common.hpp -- device-side data structures, cell indexing and the
// trilinear / 6th-order gathers of the droplet step (gfx950).
//
// Index conventions are the reference's (particles.f90): ipt/jpt are 1-based
// horizontal node indices, kpt/kwpt are u-level / w-level vertical indices in
// 0..nnz+1, ext arrays are (0:nnz+1, -1:nny+3, -1:nnx+3) with z fastest.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>

namespace ntlp {

constexpr int kBlock = 256;       // threads per workgroup (4 wave64)
constexpr int kMaxLevels = 1024;  // nnz + 2 <= kMaxLevels (LDS copies of z, zz)

struct DevGrid {
  int nnx, nny, nnz;
  int interleaved;  // zz(k) < z(k) < zz(k+1) strictly: one search yields kpt and kwpt
  int nsteps;       // steps of the fixed-length binary search over nnz+2 levels
  int pad_;
  float rdx32, rdy32;  // 1/dx, 1/dy rounded to FP32 (mixed-precision gather weights)
  double dx, dy, xl, yl;
  const double *z, *zz, *dzw, *dzu;  // device copies of the grid metadata
  const float *rdzw32, *rdzu32;      // 1/dzw, 1/dzu rounded to FP32
  const double *rdzu64, *rdVu64;     // 1/dzu and 1/(dx*dy*dzu) in FP64
  double rdx64, rdy64;               // 1/dx, 1/dy in FP64
};

struct DevConsts {
  double dt, nuf, Cpa, Cpp, CpaCpp, Pra, Sc, Rd, rhow, rhos, Mw, Ru, Gam, Lv, Cpv;
  double surf_rho, surf_p, tsfcc1, grav, radius_init, pi2;
  double g[3], zetas[3], gama[3];
  double Sc13, Pra13;  // Sc**(1./3.), Pra**(1./3.): loop invariants of ie_vrt_nd
  // constant subexpressions, formed on the host in the reference's order
  double LvMw_Ru;         // (Lv*Mw)/Ru
  double Mw2Gam;          // (2*Mw)*Gam
  double Ru_rhow;         // Ru*rhow
  double Lv3;             // 3*Lv
  double Mw_Ru;           // Mw/Ru
  double Mw2Gam_Ru_rhow;  // 2 Mw Gam / (Ru rhow)
  double isol;            // 1 / ((2/3) pi2 rhos)
  // constants of the mixed-precision step and the closed-form solve, formed
  // once on the host (reciprocals and products that the FP64 reference-order
  // path does not use)
  double pi2_23;          // pi2*2.0/3.0 (evaluated left to right)
  double inuf;            // 1/nuf
  double idt;             // 1/dt
  double Lv3_Cpp;         // 3 Lv / Cpp
  double Cpv_Cpa;         // Cpv / Cpa
  double hA0;             // 0.5 dt nuf / (Sc rhow)
  double hB0;             // 1.5 dt (Cpa/Cpp) nuf / (Pra rhow)
  double Ru_Mw;           // Ru / Mw
  double irhos;           // 1 / rhos
  double nuf18;           // 18 nuf
  double isurf_rho;       // 1 / surf_rho
  // 0.6 Sc^(1/3), 0.6 Pra^(1/3) of Sherwood/Nusselt: read from the constants
  // (scalar registers) rather than kept per droplet in vector registers
  double sc6, pr6;
};

struct DevFlags {
  int ilin, icouple, iTcouple, iHcouple, ievap, iexner, icase, pad_;
};

struct DevFields {
  const double *u, *v, *w, *T, *q;  // ext layout, FP64
  const float *u32, *v32, *w32;     // FP32 copies of u, v, w (mixed precision gathers)
  int64_t sj, si;                   // element strides of j and i (k stride is 1)
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ int64_t ext_off(const DevFields& F, int k, int j, int i) {
  return (int64_t)k + F.sj * (int64_t)(j + 1) + F.si * (int64_t)(i + 1);
}

// Number of leading elements of a(0:len-1) that are not greater than x, i.e.
// the 0-based index of the first element > x, or len when there is none.  For
// a monotone array this is exactly what Fortran's
// minloc(a, 1, mask=(a .gt. x)) selects (as a 1-based position minus one);
// a NaN x compares false everywhere and yields len, like the empty mask.
// Fixed trip count (branch-free, lane-uniform).
__device__ __forceinline__ int first_gt(const double* a, int len, int nsteps, double x) {
  int pos = 0;
  for (int s = nsteps - 1; s >= 0; --s) {
    int cand = pos + (1 << s);
    if (cand <= len && !(a[cand - 1] > x)) pos = cand;
  }
  return pos;
}

struct Cell {
  int ipt, jpt, kpt, kwpt;
};

// Cell indices of particles.f90:780-788: ipt = floor(x/dx)+1,
// jpt = floor(y/dy)+1, kpt = minloc(zz, zz>z) - 2, kwpt = minloc(z, z>z) - 2
// (minloc = 0 when empty, giving -2).  s_z / s_zz are LDS copies.
__device__ __forceinline__ Cell cell_of(const DevGrid& g, const double* s_z, const double* s_zz,
                                        double x, double y, double z) {
  Cell c;
  c.ipt = (int)floor(x / g.dx) + 1;
  c.jpt = (int)floor(y / g.dy) + 1;
  const int len = g.nnz + 2;
  int pzz = first_gt(s_zz, len, g.nsteps, z);
  int pz;
  if (g.interleaved) {
    // z(m-2) < zz(m-1) <= x < zz(m) < z(m): the first z above x is m-1 or m.
    pz = (pzz >= 1 && s_z[pzz - 1] > z) ? pzz - 1 : pzz;
  } else {
    pz = first_gt(s_z, len, g.nsteps, z);
  }
  c.kpt = (pzz == len) ? -2 : pzz - 1;
  c.kwpt = (pz == len) ? -2 : pz - 1;
  return c;
}

// The same indices when a nearby u level is known (the droplet's previous
// cell): two comparisons confirm kpt = kg, i.e. zz(kg) <= z < zz(kg+1), and
// only otherwise the binary search runs, so the result is always the
// reference's minloc.
__device__ __forceinline__ Cell cell_near(const DevGrid& g, const double* s_z, const double* s_zz,
                                          double x, double y, double z, int kg) {
  Cell c;
  c.ipt = (int)floor(x / g.dx) + 1;
  c.jpt = (int)floor(y / g.dy) + 1;
  const int len = g.nnz + 2;
  kg = clampi(kg, 0, g.nnz);
  int pzz;
  if (s_zz[kg] <= z && s_zz[kg + 1] > z) pzz = kg + 1;
  else if (kg + 2 < len && s_zz[kg + 1] <= z && s_zz[kg + 2] > z) pzz = kg + 2;
  else if (kg >= 1 && s_zz[kg - 1] <= z && s_zz[kg] > z) pzz = kg;
  else pzz = first_gt(s_zz, len, g.nsteps, z);
  int pz;
  if (g.interleaved) {
    pz = (pzz >= 1 && s_z[pzz - 1] > z) ? pzz - 1 : pzz;
  } else {
    pz = first_gt(s_z, len, g.nsteps, z);
  }
  c.kpt = (pzz == len) ? -2 : pzz - 1;
  c.kwpt = (pz == len) ? -2 : pz - 1;
  return c;
}

// w-cell index of calc_tauc_min (particles.f90:3334): minloc(z, z>z) - 1.
__device__ __forceinline__ int wcell_of(const DevGrid& g, const double* s_z, double z) {
  const int len = g.nnz + 2;
  int pz = first_gt(s_z, len, g.nsteps, z);
  return (pz == len) ? -1 : pz;
}

// u-cell sort key: (ipt-1, jpt-1, kpt) clamped into the box
// [0,nnx] x [0,nny] x [0,nnz] that every in-domain droplet occupies, k fastest
// so that consecutive keys walk a z column of the ext arrays.
__device__ __forceinline__ int ukey_of(const DevGrid& g, const Cell& c) {
  int i0 = clampi(c.ipt - 1, 0, g.nnx);
  int j0 = clampi(c.jpt - 1, 0, g.nny);
  int k0 = clampi(c.kpt, 0, g.nnz);
  return k0 + (g.nnz + 1) * (j0 + (g.nny + 1) * i0);
}

// Load the vertical grid into LDS (call by all threads, then __syncthreads).
__device__ __forceinline__ void stage_grid(const DevGrid& g, double* s_z, double* s_zz) {
  for (int k = threadIdx.x; k < g.nnz + 2; k += blockDim.x) {
    s_z[k] = g.z[k];
    s_zz[k] = g.zz[k];
  }
}

// ---------------------------------------------------------------- gathers

struct Fluid {
  double u[3];
  double T, q;
};

// Node values of the trilinear gathers: the ext fields in HBM, or a
// workgroup's box of them staged in LDS.  The corner column (i, j) of a
// droplet's cell starts at element col00 + sj*j + si*i and level k sits at
// +k.  Element indices are 32-bit and the byte offset of an element is
// formed in 32 bits (the ext arrays are limited to 4 GB), so that a load is
// a scalar base plus one 32-bit vector offset instead of a 64-bit vector
// address per corner.
struct NodeView {
  const double *u, *v, *w, *T, *q;
  const float *u32, *v32, *w32;
  unsigned sj, si;
};

template <typename T>
__device__ __forceinline__ T node_at(const T* base, unsigned idx) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + idx * (unsigned)sizeof(T));
}

__device__ __forceinline__ NodeView hbm_view(const DevFields& F) {
  NodeView V;
  V.u = F.u;
  V.v = F.v;
  V.w = F.w;
  V.T = F.T;
  V.q = F.q;
  V.u32 = F.u32;
  V.v32 = F.v32;
  V.w32 = F.w32;
  V.sj = (unsigned)F.sj;
  V.si = (unsigned)F.si;
  return V;
}

// element index of the droplet's (0, 0) corner column in the ext arrays
__device__ __forceinline__ unsigned hbm_col00(const DevGrid& g, const DevFields& F, const Cell& c) {
  const int ic = clampi(c.ipt, -1, g.nnx + 2), jc = clampi(c.jpt, -1, g.nny + 2);
  return (unsigned)F.sj * (unsigned)(jc + 1) + (unsigned)F.si * (unsigned)(ic + 1);
}

// uf_interp_lin (particles.f90:767-847) in FP64, same product order
// ((f*wtx)*wty)*wtz and (i,j,k) accumulation order.  The zeroth-order wall
// override of 831-842 leaves the node (ipt+1, jpt+1) values in uf(1:2), Tf and
// qinf.  Indices are clamped only where the reference would leave its arrays.
template <bool TQ>
__device__ __forceinline__ Fluid gather_lin_f64_v(const DevGrid& g, const NodeView& V, unsigned col00,
                                                  const double* s_z, const double* s_zz, double x, double y,
                                                  double z, const Cell& c) {
  const int nnz = g.nnz, kmax = nnz + 1;
  const int kpt = c.kpt, kwpt = c.kwpt;
  const double dzu1 = g.dzu[clampi(kpt + 1, 0, nnz + 2)];
  const double dzw1 = g.dzw[clampi(kwpt + 1, 0, nnz + 2)];
  const int ku[2] = {clampi(kpt, 0, kmax), clampi(kpt + 1, 0, kmax)};
  const int kw[2] = {clampi(kwpt, 0, kmax), clampi(kwpt + 1, 0, kmax)};
  double wtz[2], wtw[2], wtx[2], wty[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    wtz[k] = 1.0 - fabs(z - s_zz[ku[k]]) / dzu1;
    wtw[k] = 1.0 - fabs(z - s_z[kw[k]]) / dzw1;
    wtx[k] = 1.0 - fabs(x - g.dx * (double)(k + c.ipt - 1)) / g.dx;
    wty[k] = 1.0 - fabs(y - g.dy * (double)(k + c.jpt - 1)) / g.dy;
  }
  Fluid f = {{0.0, 0.0, 0.0}, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const unsigned col = col00 + V.sj * j + V.si * i;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const unsigned o = col + ku[k], ow = col + kw[k];
        f.u[0] = f.u[0] + node_at(V.u, o) * wtx[i] * wty[j] * wtz[k];
        f.u[1] = f.u[1] + node_at(V.v, o) * wtx[i] * wty[j] * wtz[k];
        f.u[2] = f.u[2] + node_at(V.w, ow) * wtx[i] * wty[j] * wtw[k];
        if (TQ) {
          f.T = f.T + node_at(V.T, o) * wtx[i] * wty[j] * wtz[k];
          f.q = f.q + node_at(V.q, o) * wtx[i] * wty[j] * wtz[k];
        }
      }
    }
  if (kpt == nnz || kpt == 0) {
    const unsigned o = col00 + V.sj + V.si + (unsigned)(kpt == 0 ? 1 : kpt);
    f.u[0] = node_at(V.u, o);
    f.u[1] = node_at(V.v, o);
    if (TQ) {
      f.T = node_at(V.T, o);
      f.q = node_at(V.q, o);
    }
  }
  return f;
}

template <bool TQ>
__device__ __forceinline__ Fluid gather_lin_f64(const DevGrid& g, const DevFields& F,
                                                const double* s_z, const double* s_zz, double x,
                                                double y, double z, const Cell& c) {
  return gather_lin_f64_v<TQ>(g, hbm_view(F), hbm_col00(g, F, c), s_z, s_zz, x, y, z, c);
}

// Levels k and k+1 of one ext column in one load (the arrays are level
// fastest, so the two nodes are adjacent): an 8-byte load of two FP32 nodes
// or a 16-byte load of two FP64 nodes, at the element's own alignment.
__device__ __forceinline__ void node_pair(const float* base, unsigned idx, float& lo, float& hi) {
  const char* p = static_cast<const char*>(
      __builtin_assume_aligned(reinterpret_cast<const char*>(base) + idx * 4u, 4));
  float v[2];
  __builtin_memcpy(v, p, 8);
  lo = v[0];
  hi = v[1];
}
__device__ __forceinline__ void node_pair(const double* base, unsigned idx, double& lo, double& hi) {
  const char* p = static_cast<const char*>(
      __builtin_assume_aligned(reinterpret_cast<const char*>(base) + idx * 8u, 8));
  double v[2];
  __builtin_memcpy(v, p, 16);
  lo = v[0];
  hi = v[1];
}

// Mixed-precision trilinear gather: FP32 fractional weights (the FP64 offset
// of the droplet from its cell origin times the FP32 reciprocal cell size), u, v, w from FP32 field
// copies accumulated in FP32; T and q from the FP64 fields in the form
// f000 + sum w_n (f_n - f000), which equals the plain weighted sum because the
// trilinear weights sum to one, and keeps the FP32 weight rounding off the
// 300 K offset of T.  Same wall override as the FP64 gather.
// PAIR (HBM): each corner column's two levels come in one load (node_pair)
// from levels kp, kp+1 with kp = min(ku0, kmax-1), which stay inside the
// column; a droplet whose levels clamp to one (outside the domain
// vertically) has both vertical weights moved onto that level.  For every
// other droplet the values and the arithmetic are those of the two-load form
// (LDS), which the staged gather therefore uses only for such droplets.
template <bool TQ, bool PAIR = false>
__device__ __forceinline__ Fluid gather_lin_mixed_v(const DevGrid& g, const NodeView& V, unsigned col00,
                                                    const double* s_z, const double* s_zz,
                                                    double x, double y, double z, const Cell& c) {
  const int nnz = g.nnz, kmax = nnz + 1;
  const int kpt = c.kpt, kwpt = c.kwpt;
  const float rdzu1 = g.rdzu32[clampi(kpt + 1, 0, nnz + 2)];
  const float rdzw1 = g.rdzw32[clampi(kwpt + 1, 0, nnz + 2)];
  const int ku0 = clampi(kpt, 0, kmax), ku1 = clampi(kpt + 1, 0, kmax);
  const int kw0 = clampi(kwpt, 0, kmax), kw1 = clampi(kwpt + 1, 0, kmax);
  const float fx = (float)(x - g.dx * (double)(c.ipt - 1)) * g.rdx32;
  const float fy = (float)(y - g.dy * (double)(c.jpt - 1)) * g.rdy32;
  const float fz = (float)(z - s_zz[ku0]) * rdzu1;
  const float fw = (float)(z - s_z[kw0]) * rdzw1;
  const float wtx[2] = {1.0f - fx, fx}, wty[2] = {1.0f - fy, fy};
  float wtz[2] = {1.0f - fz, fz}, wtw[2] = {1.0f - fw, fw};
  float u0 = 0.f, u1 = 0.f, u2 = 0.f;
  double T = 0.0, q = 0.0, T0 = 0.0, q0 = 0.0;
  if (PAIR) {
    const int kp = min(ku0, kmax - 1), kq = min(kw0, kmax - 1);
    if (ku1 == ku0) {  // clamped: one level, the lower (kp) or the upper (kp+1)
      const float s = wtz[0] + wtz[1];
      wtz[0] = (ku0 == kp) ? s : 0.0f;
      wtz[1] = (ku0 == kp) ? 0.0f : s;
    }
    if (kw1 == kw0) {
      const float s = wtw[0] + wtw[1];
      wtw[0] = (kw0 == kq) ? s : 0.0f;
      wtw[1] = (kw0 == kq) ? 0.0f : s;
    }
    double Tn[2][2][2], qn[2][2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const unsigned col = col00 + V.sj * j + V.si * i;
        const float wxy = wtx[i] * wty[j];
        const float a0 = wxy * wtz[0], a1 = wxy * wtz[1];
        float na, nb;
        node_pair(V.u32, col + kp, na, nb);
        u0 += na * a0 + nb * a1;
        node_pair(V.v32, col + kp, na, nb);
        u1 += na * a0 + nb * a1;
        node_pair(V.w32, col + kq, na, nb);
        u2 += na * (wxy * wtw[0]) + nb * (wxy * wtw[1]);
        if (TQ) {
          node_pair(V.T, col + kp, Tn[i][j][0], Tn[i][j][1]);
          node_pair(V.q, col + kp, qn[i][j][0], qn[i][j][1]);
        }
      }
    if (TQ) {
      T0 = (ku0 == kp) ? Tn[0][0][0] : Tn[0][0][1];
      q0 = (ku0 == kp) ? qn[0][0][0] : qn[0][0][1];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const float wxy = wtx[i] * wty[j];
          const float a0 = wxy * wtz[0], a1 = wxy * wtz[1];
          T += (Tn[i][j][0] - T0) * (double)a0 + (Tn[i][j][1] - T0) * (double)a1;
          q += (qn[i][j][0] - q0) * (double)a0 + (qn[i][j][1] - q0) * (double)a1;
        }
    }
  } else {
    if (TQ) {
      T0 = node_at(V.T, col00 + ku0);
      q0 = node_at(V.q, col00 + ku0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const unsigned col = col00 + V.sj * j + V.si * i;
        const float wxy = wtx[i] * wty[j];
        const float a0 = wxy * wtz[0], a1 = wxy * wtz[1];
        u0 += node_at(V.u32, col + ku0) * a0 + node_at(V.u32, col + ku1) * a1;
        u1 += node_at(V.v32, col + ku0) * a0 + node_at(V.v32, col + ku1) * a1;
        u2 += node_at(V.w32, col + kw0) * (wxy * wtw[0]) + node_at(V.w32, col + kw1) * (wxy * wtw[1]);
        if (TQ) {
          T += (node_at(V.T, col + ku0) - T0) * (double)a0 + (node_at(V.T, col + ku1) - T0) * (double)a1;
          q += (node_at(V.q, col + ku0) - q0) * (double)a0 + (node_at(V.q, col + ku1) - q0) * (double)a1;
        }
      }
  }
  Fluid f;
  f.u[0] = u0;
  f.u[1] = u1;
  f.u[2] = u2;
  f.T = T0 + T;
  f.q = q0 + q;
  if (kpt == nnz || kpt == 0) {
    const unsigned o = col00 + V.sj + V.si + (unsigned)(kpt == 0 ? 1 : kpt);
    f.u[0] = node_at(V.u32, o);
    f.u[1] = node_at(V.v32, o);
    if (TQ) {
      f.T = node_at(V.T, o);
      f.q = node_at(V.q, o);
    }
  }
  return f;
}

template <bool TQ>
__device__ __forceinline__ Fluid gather_lin_mixed(const DevGrid& g, const DevFields& F,
                                                  const double* s_z, const double* s_zz,
                                                  double x, double y, double z, const Cell& c) {
  return gather_lin_mixed_v<TQ, true>(g, hbm_view(F), hbm_col00(g, F, c), s_z, s_zz, x, y, z, c);
}

// ------------------------------------------------- LDS-staged gather

// The field box of a workgroup: the ext nodes of the corner columns of every
// cell its droplets occupy, staged in LDS before the gather.  Box column
// (ii, jj) holds ext column ipt = i0 + 1 + ii, jpt = j0 + 1 + jj (i0, j0: the
// smallest cell indices ipt-1, jpt-1 of the workgroup's droplets) at levels
// k0 .. k0 + nk - 1, level fastest.  ni == 0: no box (every droplet gathers
// from HBM).
struct FieldBox {
  int i0, j0, k0, ni, nj, nk;
  NodeView V;  // the box's node arrays in LDS (sj = nk, si = nj * nk)
};

// bytes per staged node: FP32 u, v, w and FP64 T, q (mixed), or five FP64
__host__ __device__ constexpr int box_node_bytes(bool mixed) { return mixed ? 28 : 40; }

// True when every node the trilinear gather of cell c reads is in the box.
__device__ __forceinline__ bool box_has(const FieldBox& b, const DevGrid& g, const Cell& c) {
  const int ii = c.ipt - 1 - b.i0, jj = c.jpt - 1 - b.j0;
  const int kmax = g.nnz + 1, k1 = b.k0 + b.nk - 1;
  const int ku0 = clampi(c.kpt, 0, kmax), ku1 = clampi(c.kpt + 1, 0, kmax);
  const int kw0 = clampi(c.kwpt, 0, kmax), kw1 = clampi(c.kwpt + 1, 0, kmax);
  // a droplet whose levels clamp to one is gathered from HBM (the paired
  // loads there weight it differently from the two-load form)
  return b.ni > 0 && ii >= 0 && ii <= b.ni - 2 && jj >= 0 && jj <= b.nj - 2 && ku0 >= b.k0 && ku1 <= k1 &&
         kw0 >= b.k0 && kw1 <= k1 && ku1 != ku0 && kw1 != kw0;
}

// Trilinear gather of cell c from the box when it holds the cell's nodes,
// from HBM otherwise; both read the same node values with the same
// arithmetic, so the result does not depend on where the nodes came from.
template <bool MIXED>
__device__ __forceinline__ Fluid gather_lin_staged(const DevGrid& g, const DevFields& F, const FieldBox& b,
                                                   const double* s_z, const double* s_zz, double x, double y,
                                                   double z, const Cell& c) {
  // one path per wave: the box when it holds the cells of all the wave's
  // droplets (in the workgroup that the box was made for, every in-domain
  // droplet's), HBM otherwise
  if (__all(box_has(b, g, c))) {
    const unsigned col00 = (unsigned)(((c.ipt - 1 - b.i0) * b.nj + (c.jpt - 1 - b.j0)) * b.nk - b.k0);
    return MIXED ? gather_lin_mixed_v<true>(g, b.V, col00, s_z, s_zz, x, y, z, c)
                 : gather_lin_f64_v<true>(g, b.V, col00, s_z, s_zz, x, y, z, c);
  }
  return MIXED ? gather_lin_mixed<true>(g, F, s_z, s_zz, x, y, z, c) : gather_lin_f64<true>(g, F, s_z, s_zz, x, y, z, c);
}

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
  return v;
}

// Stages the field box of the workgroup's droplets (u-cell keys, k fastest;
// key < 0 or !live: no droplet) into s_box (box_bytes of LDS).  All threads
// of the workgroup call it (it holds two barriers); s_mm: 6 * (blockDim/64)
// ints of LDS.  Returns ni == 0 when the box would not fit.
template <bool MIXED>
__device__ FieldBox stage_box(const DevGrid& g, const DevFields& F, int key, bool live, double* s_box,
                              int box_bytes, int* s_mm) {
  const int nz1 = g.nnz + 1, ny1 = g.nny + 1;
  int v[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MAX, INT_MAX, INT_MAX};
  if (live && key >= 0) {
    const int k = key % nz1, rest = key / nz1;
    const int j = rest % ny1, i = rest / ny1;
    v[0] = i, v[1] = -i, v[2] = j, v[3] = -j, v[4] = k, v[5] = -k;
  }
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;  // waves of the workgroup
#pragma unroll
  for (int q = 0; q < 6; ++q) v[q] = wave_min_i32(v[q]);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 6; ++q) s_mm[w * 6 + q] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    int m = s_mm[q];
    for (int k = 1; k < nw; ++k) m = min(m, s_mm[k * 6 + q]);
    v[q] = __builtin_amdgcn_readfirstlane(m);
  }
  FieldBox b;
  b.ni = 0;
  if (v[0] == INT_MAX) return b;  // no droplet: the same for every thread
  b.i0 = v[0];
  b.j0 = v[2];
  b.ni = -v[1] - v[0] + 2;
  b.nj = -v[3] - v[2] + 2;
  b.k0 = max(v[4] - 1, 0);
  b.nk = min(-v[5] + 1, g.nnz + 1) - b.k0 + 1;
  // in 64 bits: a workgroup whose droplets sit at both ends of the domain (a
  // droplet that crossed the periodic boundary since the slots were last
  // sorted) spans the whole grid, whose byte count overflows 32 bits
  const int64_t box_size = (int64_t)b.ni * b.nj * b.nk * box_node_bytes(MIXED);
  if (box_size > box_bytes) {
    b.ni = 0;
    return b;
  }
  const int nodes = b.ni * b.nj * b.nk;
  double* sT = s_box;
  double* sq = s_box + nodes;
  b.V.T = sT;
  b.V.q = sq;
  b.V.sj = (unsigned)b.nk;
  b.V.si = (unsigned)(b.nj * b.nk);
  float *su32 = nullptr, *sv32 = nullptr, *sw32 = nullptr;
  double *su = nullptr, *sv = nullptr, *sw = nullptr;
  if (MIXED) {
    su32 = reinterpret_cast<float*>(s_box + 2 * nodes);
    sv32 = su32 + nodes;
    sw32 = sv32 + nodes;
  } else {
    su = s_box + 2 * nodes;
    sv = su + nodes;
    sw = sv + nodes;
  }
  b.V.u32 = su32;
  b.V.v32 = sv32;
  b.V.w32 = sw32;
  b.V.u = su;
  b.V.v = sv;
  b.V.w = sw;
  // consecutive threads take consecutive levels of a box column (coalesced)
  const unsigned gsj = (unsigned)F.sj, gsi = (unsigned)F.si;
  const unsigned g0 = gsj * (unsigned)(b.j0 + 2) + gsi * (unsigned)(b.i0 + 2) + (unsigned)b.k0;
  for (int e = threadIdx.x; e < nodes; e += blockDim.x) {
    const int col = e / b.nk, l = e - col * b.nk;
    const int ii = col / b.nj, jj = col - ii * b.nj;
    const unsigned o = g0 + gsj * (unsigned)jj + gsi * (unsigned)ii + (unsigned)l;
    sT[e] = node_at(F.T, o);
    sq[e] = node_at(F.q, o);
    if (MIXED) {
      su32[e] = node_at(F.u32, o);
      sv32[e] = node_at(F.v32, o);
      sw32[e] = node_at(F.w32, o);
    } else {
      su[e] = node_at(F.u, o);
      sv[e] = node_at(F.v, o);
      sw[e] = node_at(F.w, o);
    }
  }
  __syncthreads();
  return b;
}

// uf_interp (particles.f90:554-765): 6th-order Lagrange on 6x6x6 nodes with
// the reference's one-sided truncation near the walls.  FP64 in all
// precisions.  The out-of-domain branches with first=last=0 leave the weights
// zero (the reference evaluates one stray iteration outside its arrays there).
__device__ __noinline__ Fluid gather6_f64(const DevGrid& g, const DevFields& F, const double* s_z,
                                          const double* s_zz, double x, double y, double z,
                                          const Cell& c) {
  const int nnz = g.nnz;
  int kuv[6], kwv[6];
  double wt[4][6];
  kuv[2] = c.kpt;
  kuv[3] = kuv[2] + 1; kuv[4] = kuv[3] + 1; kuv[5] = kuv[4] + 1;
  kuv[1] = kuv[2] - 1; kuv[0] = kuv[1] - 1;
  kwv[2] = c.kwpt;
  kwv[3] = kwv[2] + 1; kwv[4] = kwv[3] + 1; kwv[5] = kwv[4] + 1;
  kwv[1] = kwv[2] - 1; kwv[0] = kwv[1] - 1;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 6; ++b) wt[a][b] = 0.0;
  {
    const double dxv[2] = {g.dx, g.dy};
    const double xp[2] = {x, y};
    const int base[2] = {c.ipt, c.jpt};
    for (int d = 0; d < 2; ++d)
      for (int j = 0; j < 6; ++j) {
        double xj = dxv[d] * (double)(base[d] + j - 3);
        double pj = 1.0;
        for (int k = 0; k < 6; ++k) {
          double xk = dxv[d] * (double)(base[d] + k - 3);
          if (j != k) pj = pj * (xp[d] - xk) / (xj - xk);
        }
        wt[d][j] = pj;
      }
  }
  int first, last;
  if (kuv[2] == 1) { first = 3; last = 4; kuv[0] = 1; kuv[1] = 1; }
  else if (kuv[2] == 0) { first = 4; last = 5; kuv[0] = 1; kuv[1] = 1; kuv[2] = 1; }
  else if (kuv[2] < 0) { first = 0; last = 0; }
  else if (kuv[2] == 2) { first = 2; last = 5; }
  else if (kuv[2] == nnz) { first = 2; last = 3; kuv[3] = nnz; kuv[4] = nnz; kuv[5] = nnz; }
  else if (kuv[2] > nnz) { first = 0; last = 0; }
  else if (kuv[2] == nnz - 1) { first = 3; last = 4; kuv[4] = nnz; kuv[5] = nnz; }
  else if (kuv[2] == nnz - 2) { first = 2; last = 5; }
  else { first = 1; last = 6; }
  if (first > 0)
    for (int j = first; j <= last; ++j) {
      double xj = s_zz[clampi(kuv[j - 1], 0, nnz + 1)];
      double pj = 1.0;
      for (int k = first; k <= last; ++k) {
        double xk = s_zz[clampi(kuv[k - 1], 0, nnz + 1)];
        if (j != k) pj = pj * (z - xk) / (xj - xk);
      }
      wt[2][j - 1] = pj;
    }
  if (kwv[2] == 0) { first = 3; last = 4; kwv[0] = 1; kwv[1] = 1; }
  else if (kwv[2] < 0) { first = 0; last = 0; kwv[0] = 1; kwv[1] = 1; kwv[2] = 1; }
  else if (kwv[2] == 1) { first = 2; last = 5; kwv[0] = 1; }
  else if (kwv[2] == nnz - 1) { first = 3; last = 4; kwv[4] = nnz; kwv[5] = nnz; }
  else if (kwv[2] >= nnz) { first = 0; last = 0; kwv[2] = nnz; kwv[3] = nnz; kwv[4] = nnz; kwv[5] = nnz; }
  else if (kwv[2] == nnz - 2) { first = 2; last = 5; kwv[5] = nnz; }
  else { first = 1; last = 6; }
  if (first > 0)
    for (int j = first; j <= last; ++j) {
      double xj = s_z[clampi(kwv[j - 1], 0, nnz + 1)];
      double pj = 1.0;
      for (int k = first; k <= last; ++k) {
        double xk = s_z[clampi(kwv[k - 1], 0, nnz + 1)];
        if (j != k) pj = pj * (z - xk) / (xj - xk);
      }
      wt[3][j - 1] = pj;
    }
  Fluid f = {{0.0, 0.0, 0.0}, 0.0, 0.0};
  for (int k = 0; k < 6; ++k) {
    const int izuv = clampi(kuv[k], 0, nnz + 1), izw = clampi(kwv[k], 0, nnz + 1);
    for (int j = 0; j < 6; ++j) {
      const int iy = clampi(c.jpt + j - 2, -1, g.nny + 3);
      for (int i = 0; i < 6; ++i) {
        const int ix = clampi(c.ipt + i - 2, -1, g.nnx + 3);
        const int64_t o = ext_off(F, izuv, iy, ix), ow = ext_off(F, izw, iy, ix);
        f.u[0] = f.u[0] + F.u[o] * wt[0][i] * wt[1][j] * wt[2][k];
        f.u[1] = f.u[1] + F.v[o] * wt[0][i] * wt[1][j] * wt[2][k];
        f.u[2] = f.u[2] + F.w[ow] * wt[0][i] * wt[1][j] * wt[3][k];
        f.T = f.T + F.T[o] * wt[0][i] * wt[1][j] * wt[2][k];
        f.q = f.q + F.q[o] * wt[0][i] * wt[1][j] * wt[2][k];
      }
    }
  }
  return f;
}

}  // namespace ntlp
