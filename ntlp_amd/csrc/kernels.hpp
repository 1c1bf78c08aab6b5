// kernels.hpp -- HIP kernels of the droplet step (gfx950, wave64).
//
// Device layout of the droplets (capacity `cap`, slot t):
//   hot[hix(c, t)]   14 FP64 columns the step reads and writes every step
//                    (kHot* below), kept in cell-sorted slot order and
//                    ping-ponged between two buffers by the fused step.  The
//                    columns are tiled by wave: the 14 columns of 64
//                    consecutive slots form one 7 KB block, so that a wave's
//                    column reads and writes stay within a page or two
//                    instead of touching 14 pages (28 across the ping-pong)
//                    whose translations would thrash the L1 TLB;
//   mult[t], home[t] multiplicity and the droplet's put-order index;
//   cold[c*cap + h]  the remaining FP64 record columns, in put order h, only
//                    touched through home[t] (diagnostics, RK3 stage
//                    registers, pre-step values for the coupling).
#pragma once
#include <algorithm>

#include "ensemble.hpp"
#include "physics.hpp"

namespace ntlp {

// hot column indices
enum {
  kHX = 0, kHY, kHZ, kHVX, kHVY, kHVZ, kHTP, kHR, kHMS, kHKAPPA, kHRC, kHRES, kHACTRES, kHNUMACT, kNHot
};
// cold column indices
enum {
  kCUFX = 0, kCUFY, kCUFZ, kCXRHSX, kCXRHSY, kCXRHSZ, kCVRHSX, kCVRHSY, kCVRHSZ, kCTPRHS_S, kCTPRHS_L,
  kCTF, kCRADRHS, kCQINF, kCQSTAR, kCVXOLD, kCVYOLD, kCVZOLD, kCTPOLD, kCRADIUSOLD, kNCold
};

// Tally slots in the per-step counter rows.
enum { kTNum100 = 0, kTNumImpos, kTActnum, kTDenum, kTGnIters, kTLmIters, kTSolves, kTSteps, kTHist0, kTallyWidth = kTHist0 + 64 };

// Wave tiles of the droplet columns (see above): element (c, t) of a tiled
// array with `cols` columns.
constexpr int kTileRows = 64;
__host__ __device__ __forceinline__ int64_t tix(int cols, int c, int64_t t) {
  return ((t >> 6) * cols + c) * kTileRows + (t & (kTileRows - 1));
}
__host__ __device__ __forceinline__ int64_t hix(int c, int64_t t) { return tix(kNHot, c, t); }
__host__ __device__ __forceinline__ int64_t aix(int f, int64_t t) { return tix(6, f, t); }
// elements of a tiled array of `cols` columns holding `cap` slots
inline int64_t tiled_elems(int64_t cap, int cols) { return (cap + kTileRows - 1) / kTileRows * kTileRows * cols; }

struct DropletBufs {
  double* hot;
  int64_t* mult;
  int32_t* home;
};

// Serving pool: the slot of a request (home >> kServeIdxBits) and the index
// of the droplet in it (home & kServeIdxMask) travel in `home`.
constexpr int kServeIdxBits = 23;
constexpr int32_t kServeIdxMask = (1 << kServeIdxBits) - 1;
constexpr int kServeSlots = 256;

// Per-iteration control block of a serving pool (device resident, written by
// the host before each step so that one captured graph serves every pool size).
struct ServeCtl {
  int64_t hw;      // slots in use in the current buffer (staying and leaving)
  int64_t n_live;  // droplets stepped in this iteration
  int64_t out_off[kServeSlots];  // first output row of each request slot
  int32_t last[kServeSlots];     // 1: this iteration is the request's last step
};

struct StepArgs {
  DevGrid g;
  DevConsts C;
  // the same constants in device memory: the droplet kernels read them
  // through this pointer where they are used, so that the ~100 dwords of
  // DevConsts are not held in scalar registers across the Newton solve (they
  // spilled to vector-register lanes and were reloaded every iteration)
  const DevConsts* Cp;
  DevFlags fl;
  DevFields F;
  int64_t n, cap;
  DropletBufs src, dst;  // fused step reads src through perm, writes dst in slot order
  const int32_t* perm;   // sorted index -> src slot (nullptr: identity)
  double* cold;          // put-order cold columns
  double* amounts;       // [6][cap] coupling amounts in dst slot order (FP32 in mixed precision)
  const int32_t* keys_in; // u-cell key of the current position (src slot order), or null
  int32_t* keys;          // [cap] u-cell key of the new position (dst slot order)
  int32_t* counts;       // u-cell histogram (must be zero on entry)
  unsigned long long* tally;  // kTallySlots x kTallyWidth counter slots
  const ServeCtl* serve;      // serving pool control (SERVE kernels only)
  double* out;                // serving output rows: [kNHot][out_cap]
  int64_t* out_mult;
  int64_t out_cap;
  int it;
  int istage;
  int write_cold;        // write the record diagnostics (final step of a fused call)
  int inplace;           // update slot t in place (src == dst, no perm, keys == keys_in)
  double* diag;          // [kNCold][cap] diagnostics of the final step in dst slot order
  int lazy_j;            // BE library steps: index of this step within its call
  double* flux;          // BE flux statistics [3][nnz+1] (zero on entry), or null
  int box_bytes;         // LDS for the staged field box of k_step (0: gather from HBM)
};

// Dynamic LDS of k_step: z and zz, then the field box.  The box is sized so
// that four 256-thread workgroups (four waves per SIMD, what the step's
// register use allows) share a CU's 160 KB.
constexpr int kStepLdsPerBlock = 38912;
inline int step_box_bytes(int nnz, bool staged) {
  const int rest = kStepLdsPerBlock - 16 * (nnz + 2) - 4 * (kTallyWidth + 1) - 4 * 6 * (kBlock / 64);
  return staged ? std::max(0, std::min(32768, rest)) : 0;
}
inline size_t step_lds_bytes(int nnz, int box_bytes) { return (size_t)16 * (nnz + 2) + (size_t)box_bytes; }

// Workgroup -> droplet chunk mapping that keeps consecutive chunks on one XCD:
// gfx950 deals workgroups to its 8 XCDs round-robin, so workgroup b lands on
// XCD b%8; giving XCD x the contiguous range x*(nb/8) ... keeps neighbouring
// (cell-sorted) droplets, and so the same field columns, in one L2.
__device__ __forceinline__ int64_t xcd_block() {
  const unsigned nb = gridDim.x;  // launched as a multiple of 8
  const unsigned b = blockIdx.x;
  return (int64_t)(b & 7u) * (nb >> 3) + (b >> 3);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// --------------------------------------------------------------- tallies

__device__ __forceinline__ long long wave_sum(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Newton tallies are summed per workgroup (ballots and 32-bit wave sums, then
// LDS) and added to one of kTallySlots counter slots, each on its own cache
// lines, chosen by workgroup; k_tally_close folds the slots into the step's
// row.  (One shared row would put every workgroup's atomics on one L2 line.)
constexpr int kTallySlots = 64;

__device__ __forceinline__ int wave_sum_i32(int v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// s_red: LDS scratch of kTallyWidth + 1 ints (the sums, then an arrival
// counter), zeroed before the kernel's first __syncthreads.  Each wave adds
// its counts into LDS; the last wave of the workgroup to arrive adds the
// workgroup's sums to its slot.  There is no workgroup barrier, so a wave
// whose droplets converged quickly exits without waiting for the others.
__device__ __forceinline__ void flush_tally(const Tally& tl, unsigned long long* slots, int* s_red) {
  const int c_n100 = __popcll(__ballot(tl.num100 != 0));
  const int c_imp = __popcll(__ballot(tl.numimpos != 0));
  const int c_act = __popcll(__ballot(tl.actnum != 0));
  const int c_de = __popcll(__ballot(tl.denum != 0));
  const int c_sol = __popcll(__ballot(tl.solves != 0));
  const int c_gn = wave_sum_i32(tl.gn_iters);
  const int c_lm = __ballot(tl.lm_iters != 0) ? wave_sum_i32(tl.lm_iters) : 0;
  if (lane_id() == 0) {
    const int v[7] = {c_n100, c_imp, c_act, c_de, c_gn, c_lm, c_sol};
#pragma unroll
    for (int k = 0; k < 7; ++k)
      if (v[k]) atomicAdd(&s_red[k], v[k]);
  }
  // iteration histogram: one LDS add per distinct value in the wave (a wave
  // holds two to four), not one per lane onto the same few words
  {
    const int hv = tl.hist < 63 ? tl.hist : 63;
    unsigned long long rem = __ballot(tl.hist >= 0);
    while (rem) {
      const int lead = __ffsll((long long)rem) - 1;
      const int h = __shfl(hv, lead, 64);
      const unsigned long long same = __ballot(tl.hist >= 0 && hv == h);
      if (lane_id() == lead) atomicAdd(&s_red[kTHist0 + h], __popcll(same));
      rem &= ~same;
    }
  }
  __threadfence_block();  // this wave's adds precede its arrival
  int last = 0;
  if (lane_id() == 0) last = atomicAdd(&s_red[kTallyWidth], 1) == (int)(blockDim.x >> 6) - 1;
  last = __shfl(last, 0, 64);
  if (last) {
    __threadfence_block();
    for (int c = lane_id(); c < kTallyWidth; c += 64) {
      const int v = atomicAdd(&s_red[c], 0);
      if (v) atomicAdd(&slots[(blockIdx.x % kTallySlots) * kTallyWidth + c], (unsigned long long)v);
    }
  }
}

__device__ __forceinline__ void zero_tally_lds(int* s_red) {
  for (int i = threadIdx.x; i <= kTallyWidth; i += blockDim.x) s_red[i] = 0;
}

// Folds the slots into the step's row and into the running totals (and
// counts the step), clearing the slots; clears `flux_next` (nflux values),
// the flux statistics buffer of the next BE step, when it is not null.
__global__ void k_tally_close(unsigned long long* slots, unsigned long long* row, unsigned long long* totals,
                              int steps, double* flux_next, int nflux) {
  for (int c = threadIdx.x; c < kTallyWidth; c += blockDim.x) {
    unsigned long long s = 0;
    for (int k = 0; k < kTallySlots; ++k) {
      s += slots[k * kTallyWidth + c];
      slots[k * kTallyWidth + c] = 0ull;
    }
    s += (c == kTSteps ? (unsigned long long)steps : 0ull);
    row[c] = s;
    totals[c] += s;
  }
  if (flux_next)
    for (int i = threadIdx.x; i < nflux; i += blockDim.x) flux_next[i] = 0.0;
}

// ------------------------------------------------------------ load/store

// with_res false: the residence times stay in HBM (d.res = d.actres = 0)
__device__ __forceinline__ Drop load_drop(const double* hot, int64_t s, bool with_res = true) {
  Drop d;
  if (!with_res) {
    d.x = hot[hix(kHX, s)];
    d.y = hot[hix(kHY, s)];
    d.z = hot[hix(kHZ, s)];
    d.vx = hot[hix(kHVX, s)];
    d.vy = hot[hix(kHVY, s)];
    d.vz = hot[hix(kHVZ, s)];
    d.Tp = hot[hix(kHTP, s)];
    d.r = hot[hix(kHR, s)];
    d.m_s = hot[hix(kHMS, s)];
    d.kappa = hot[hix(kHKAPPA, s)];
    d.rc = hot[hix(kHRC, s)];
    d.res = d.actres = 0.0;
    d.numact = hot[hix(kHNUMACT, s)];
    return d;
  }
  d.x = hot[hix(kHX, s)];
  d.y = hot[hix(kHY, s)];
  d.z = hot[hix(kHZ, s)];
  d.vx = hot[hix(kHVX, s)];
  d.vy = hot[hix(kHVY, s)];
  d.vz = hot[hix(kHVZ, s)];
  d.Tp = hot[hix(kHTP, s)];
  d.r = hot[hix(kHR, s)];
  d.m_s = hot[hix(kHMS, s)];
  d.kappa = hot[hix(kHKAPPA, s)];
  d.rc = hot[hix(kHRC, s)];
  d.res = hot[hix(kHRES, s)];
  d.actres = hot[hix(kHACTRES, s)];
  d.numact = hot[hix(kHNUMACT, s)];
  return d;
}

__device__ __forceinline__ void store_drop(double* hot, int64_t s, const Drop& d) {
  hot[hix(kHX, s)] = d.x;
  hot[hix(kHY, s)] = d.y;
  hot[hix(kHZ, s)] = d.z;
  hot[hix(kHVX, s)] = d.vx;
  hot[hix(kHVY, s)] = d.vy;
  hot[hix(kHVZ, s)] = d.vz;
  hot[hix(kHTP, s)] = d.Tp;
  hot[hix(kHR, s)] = d.r;
  hot[hix(kHMS, s)] = d.m_s;
  hot[hix(kHKAPPA, s)] = d.kappa;
  hot[hix(kHRC, s)] = d.rc;
  hot[hix(kHRES, s)] = d.res;
  hot[hix(kHACTRES, s)] = d.actres;
  hot[hix(kHNUMACT, s)] = d.numact;
}

// The hot columns as plain column blocks out[c*ld + row] (serving results).
__device__ __forceinline__ void store_drop_rows(double* out, int64_t ld, int64_t row, const Drop& d) {
  const double v[kNHot] = {d.x, d.y, d.z, d.vx, d.vy, d.vz, d.Tp, d.r, d.m_s, d.kappa, d.rc, d.res, d.actres, d.numact};
#pragma unroll
  for (int c = 0; c < kNHot; ++c) out[c * ld + row] = v[c];
}

// State that moves with a droplet in the fused step but is constant in it.
// numact changes only when the droplet activates: with write_numact false it
// is left as it is (the in-place step stores it only when it changed).
__device__ __forceinline__ void store_dynamic(double* hot, int64_t s, const Drop& d, bool write_numact = true,
                                              bool write_res = true) {
  hot[hix(kHX, s)] = d.x;
  hot[hix(kHY, s)] = d.y;
  hot[hix(kHZ, s)] = d.z;
  hot[hix(kHVX, s)] = d.vx;
  hot[hix(kHVY, s)] = d.vy;
  hot[hix(kHVZ, s)] = d.vz;
  hot[hix(kHTP, s)] = d.Tp;
  hot[hix(kHR, s)] = d.r;
  if (write_res) {
    hot[hix(kHRES, s)] = d.res;
    hot[hix(kHACTRES, s)] = d.actres;
  }
  if (write_numact) hot[hix(kHNUMACT, s)] = d.numact;
}

// Residence times over the steps of one multi-step BE call.  res and actres
// grow by dt every step (particles.f90:3204-3205) and actres restarts from 0
// when the droplet activates (3172).  The call's step kernels leave both in
// HBM; an activation in step j of the call leaves j NaN-boxed in actres, and
// k_res_settle then makes the call's additions, one dt at a time in step
// order, so the bits are those of per-step updates.  This keeps 32 bytes per
// droplet and step out of the step kernel's traffic.
constexpr int kActMarkHi = 0x7FFDEAD0;
__device__ __forceinline__ double act_mark(int j) { return __hiloint2double(kActMarkHi, j); }

__global__ void __launch_bounds__(kBlock) k_res_settle(double* hot, int64_t n, int nsteps, double dt) {
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBlock) {
    double res = hot[hix(kHRES, t)], act = hot[hix(kHACTRES, t)];
    int j0 = 0;
    if (__double2hiint(act) == kActMarkHi) {  // activated in step j0 of the call
      j0 = __double2loint(act);
      act = 0.0;
    }
    for (int k = 0; k < nsteps; ++k) {
      res = res + dt;
      if (k >= j0) act = act + dt;
    }
    hot[hix(kHRES, t)] = res;
    hot[hix(kHACTRES, t)] = act;
  }
}

__device__ __forceinline__ void load_rhs(const double* cold, int64_t cap, int64_t h, Diag& o) {
  o.xrhs[0] = cold[kCXRHSX * cap + h];
  o.xrhs[1] = cold[kCXRHSY * cap + h];
  o.xrhs[2] = cold[kCXRHSZ * cap + h];
  o.vrhs[0] = cold[kCVRHSX * cap + h];
  o.vrhs[1] = cold[kCVRHSY * cap + h];
  o.vrhs[2] = cold[kCVRHSZ * cap + h];
  o.Tprhs_s = cold[kCTPRHS_S * cap + h];
  o.Tprhs_L = cold[kCTPRHS_L * cap + h];
  o.radrhs = cold[kCRADRHS * cap + h];
}

__device__ __forceinline__ void store_gathered(double* cold, int64_t cap, int64_t h, const Diag& o) {
  cold[kCUFX * cap + h] = o.uf[0];
  cold[kCUFY * cap + h] = o.uf[1];
  cold[kCUFZ * cap + h] = o.uf[2];
  cold[kCTF * cap + h] = o.Tf;
  cold[kCQINF * cap + h] = o.qinf;
  cold[kCQSTAR * cap + h] = o.qstar;
}

__device__ __forceinline__ void store_old(double* cold, int64_t cap, int64_t h, const Diag& o) {
  cold[kCVXOLD * cap + h] = o.vold[0];
  cold[kCVYOLD * cap + h] = o.vold[1];
  cold[kCVZOLD * cap + h] = o.vold[2];
  cold[kCTPOLD * cap + h] = o.Tpold;
  cold[kCRADIUSOLD * cap + h] = o.rold;
}

__device__ __forceinline__ void store_rhs(double* cold, int64_t cap, int64_t h, const Diag& o) {
  cold[kCXRHSX * cap + h] = o.xrhs[0];
  cold[kCXRHSY * cap + h] = o.xrhs[1];
  cold[kCXRHSZ * cap + h] = o.xrhs[2];
  cold[kCVRHSX * cap + h] = o.vrhs[0];
  cold[kCVRHSY * cap + h] = o.vrhs[1];
  cold[kCVRHSZ * cap + h] = o.vrhs[2];
  cold[kCTPRHS_S * cap + h] = o.Tprhs_s;
  cold[kCTPRHS_L * cap + h] = o.Tprhs_L;
  cold[kCRADRHS * cap + h] = o.radrhs;
}

// Wave-aggregated histogram increment for runs of equal keys in lane order
// (the droplets are cell sorted, so runs are long): one atomic per run.
// Histogram of the u-cell keys of a wave: one atomic per run of equal keys.
// A negative key (a droplet leaving the pool) is not counted.
__device__ __forceinline__ void count_key(int* counts, int key, bool valid) {
  const int lane = lane_id();
  const bool live = valid && key >= 0;
  const int prev = __shfl_up(key, 1, 64);
  const bool head = live && (lane == 0 || prev != key);
  // a run ends before the next head or the next lane that is not counted
  const unsigned long long brk = __ballot(head || !live);
  if (head) {
    const unsigned long long after = (lane == 63) ? 0ull : (brk >> (lane + 1)) << (lane + 1);
    const int end = after ? (__ffsll((long long)after) - 1) : 64;
    atomicAdd(&counts[key], end - lane);
  }
}

// ---------------------------------------------------------- fused step

// One whole droplet step of a fused call: gather + advance (MODE) + boundary
// conditions + coupling amounts + new u-cell key and histogram.  Reads the
// droplet through perm from the source buffers and writes it to slot t of
// the destination buffers, so that the destination is cell sorted.
// inplace: a step between two such reorderings updates slot t where it is
// (the slots stay nearly cell sorted for many steps: a droplet moves a small
// fraction of a cell per step); only the state that changes is written back.
//
// SERVE: the step of a serving pool.  The droplet count comes from the control
// block; a droplet whose request takes its last step here is written to the
// request's output rows and leaves the pool (key -1: the next sort drops it);
// no coupling amounts are formed.
// LAZY (BE steps of the library): the residence times stay in HBM until
// k_res_settle at the end of the call.
template <int MODE, int PREC, int ILIN, bool SERVE = false, bool LAZY = false>
__global__ void __launch_bounds__(kBlock) k_step(StepArgs a) {
  const DevConsts& C = *a.Cp;
  extern __shared__ double s_dyn[];  // z(0:nnz+1) | zz(0:nnz+1) | field box (a.box_bytes)
  __shared__ int s_red[kTallyWidth + 1];
  __shared__ int s_mm[6 * (kBlock / 64)];
  const int nlev = a.g.nnz + 2;
  double* s_z = s_dyn;
  double* s_zz = s_dyn + nlev;
  const int64_t t = xcd_block() * kBlock + threadIdx.x;
  const bool valid = t < (SERVE ? a.serve->n_live : a.n);
  Tally tl = {0, 0, 0, 0, 0, 0, 0, -1};
  int key = -1;
  const bool inplace = !SERVE && a.inplace;
  static_assert(!LAZY || (MODE == 2 && !SERVE), "lazy residence times: library BE steps only");
  // the droplet's loads are issued before the levels and the field box are
  // staged (a barrier orders every memory access around it), so that they are
  // in flight while those load
  int64_t s = t;
  int kin = -1;
  Drop d;
  int64_t mult = 0;
  int32_t home = 0;
  if (valid) {
    s = (a.perm && !inplace) ? (int64_t)a.perm[t] : t;
    if (a.keys_in) kin = a.keys_in[s];
    d = load_drop(a.src.hot, s, !LAZY);
    mult = a.src.mult[s];
    home = inplace ? 0 : a.src.home[s];
  }
  stage_grid(a.g, s_z, s_zz);
  zero_tally_lds(s_red);
  __syncthreads();
  // LDS-staged gather: the trilinear field nodes of the workgroup's cells
  FieldBox bx;
  bx.ni = 0;
  if (ILIN == 1 && a.box_bytes > 0 && a.keys_in)
    bx = stage_box<PREC == PREC_MIXED>(a.g, a.F, kin, valid, s_dyn + 2 * nlev, a.box_bytes, s_mm);
  if (valid) {
    const double numact0 = d.numact;
    Diag o;
    const int kg = a.keys_in ? kin % (a.g.nnz + 1) : -1;
    if (MODE == 1) {  // three particle_update_rk3 stages; stage registers stay in VGPRs
      o.xrhs[0] = o.xrhs[1] = o.xrhs[2] = 0.0;  // zetas(1) = 0 discards the stale values
      o.vrhs[0] = o.vrhs[1] = o.vrhs[2] = 0.0;
      o.Tprhs_s = o.Tprhs_L = o.radrhs = 0.0;
      o.kpt = kg;
      for (int st = 1; st <= 3; ++st)
        rk3_droplet<PREC, ILIN>(a.g, C, a.fl, a.F, s_z, s_zz, d, st, a.it, o, o.kpt, &bx);
    } else {
      // q* is a diagnostic of the record: formed on the step that stores it
      be_droplet<PREC, ILIN, LAZY>(a.g, C, a.fl, a.F, s_z, s_zz, d, a.it, o, tl, kg, SERVE ? nullptr : a.flux,
                                   mult, &bx, !SERVE && a.write_cold);
    }
    if (SERVE && a.serve->last[home >> kServeIdxBits]) {
      const int64_t row = a.serve->out_off[home >> kServeIdxBits] + (home & kServeIdxMask);
      store_drop_rows(a.out, a.out_cap, row, d);
      a.out_mult[row] = mult;
      a.keys[t] = -1;
    } else {
      if (!SERVE) {
        double am[6];
        if (PREC == PREC_MIXED) coupling_amounts_f32(C, a.fl, d, o.Tf, o.vold, o.Tpold, o.rold, mult, am);
        else coupling_amounts(C, a.fl, d, o.Tf, o.vold, o.Tpold, o.rold, mult, am);
        if (PREC == PREC_MIXED) {  // stored in FP32, summed in FP64 by the scatter
          float* af = reinterpret_cast<float*>(a.amounts);
#pragma unroll
          for (int k = 0; k < 6; ++k) af[aix(k, t)] = (float)am[k];
        } else {
#pragma unroll
          for (int k = 0; k < 6; ++k) a.amounts[aix(k, t)] = am[k];
        }
      }
      store_dynamic(a.dst.hot, t, d, !inplace || d.numact != numact0, !LAZY);
      if (LAZY) {
        if (!inplace) {  // carried to the new slot as they are
          a.dst.hot[hix(kHRES, t)] = a.src.hot[hix(kHRES, s)];
          a.dst.hot[hix(kHACTRES, t)] = a.src.hot[hix(kHACTRES, s)];
        }
        if (tl.actnum != 0) a.dst.hot[hix(kHACTRES, t)] = act_mark(a.lazy_j);
      }
      if (!inplace) {
        a.dst.hot[hix(kHMS, t)] = d.m_s;
        a.dst.hot[hix(kHKAPPA, t)] = d.kappa;
        a.dst.hot[hix(kHRC, t)] = d.rc;
        a.dst.mult[t] = mult;
        a.dst.home[t] = home;
      }
      const Cell c = cell_near(a.g, s_z, s_zz, d.x, d.y, d.z, o.kpt);
      key = ukey_of(a.g, c);
      a.keys[t] = key;
    }
    if (!SERVE && a.write_cold) {  // coalesced in slot order; put order is restored on demand
      store_gathered(a.diag, a.cap, t, o);
      store_old(a.diag, a.cap, t, o);
      if (MODE == 1) store_rhs(a.diag, a.cap, t, o);
    }
  }
  count_key(a.counts, key, valid);
  if (MODE == 2) flush_tally(tl, a.tally, s_red);
}

// ------------------------------------------------- reference entry points

// particle_update_rk3(istage) / particle_update_BE in place on the current
// buffers (no reordering), reading and writing the record columns.
template <int MODE, int PREC, int ILIN>
__global__ void __launch_bounds__(kBlock) k_update(StepArgs a) {
  const DevConsts& C = *a.Cp;
  __shared__ double s_z[kMaxLevels], s_zz[kMaxLevels];
  __shared__ int s_red[kTallyWidth + 1];
  const int64_t t = xcd_block() * kBlock + threadIdx.x;
  Tally tl = {0, 0, 0, 0, 0, 0, 0, -1};
  // the droplet's loads go out before the barrier that follows the staging
  Drop d;
  int32_t home = 0;
  if (t < a.n) {
    d = load_drop(a.src.hot, t);
    home = a.src.home[t];
  }
  stage_grid(a.g, s_z, s_zz);
  zero_tally_lds(s_red);
  __syncthreads();
  if (t < a.n) {
    Diag o;
    if (MODE == 1) {
      load_rhs(a.cold, a.cap, home, o);
      rk3_droplet<PREC, ILIN>(a.g, C, a.fl, a.F, s_z, s_zz, d, a.istage, a.it, o);
      store_rhs(a.cold, a.cap, home, o);
      if (a.istage == 1) store_old(a.cold, a.cap, home, o);
    } else {
      be_droplet<PREC, ILIN>(a.g, C, a.fl, a.F, s_z, s_zz, d, a.it, o, tl, -1, a.flux, a.src.mult[t]);
      store_old(a.cold, a.cap, home, o);
    }
    store_gathered(a.cold, a.cap, home, o);
    store_dynamic(a.src.hot, t, d);
  }
  if (MODE == 2) flush_tally(tl, a.tally, s_red);
}

// uf_interp_lin / uf_interp for every droplet, optionally vp=uf and Tp=Tf.
template <int ILIN>
__global__ void __launch_bounds__(kBlock) k_interp(StepArgs a, int set_vp, int set_tp) {
  __shared__ double s_z[kMaxLevels], s_zz[kMaxLevels];
  const int64_t t = xcd_block() * kBlock + threadIdx.x;
  double x = 0.0, y = 0.0, z = 0.0;  // loaded before the barrier that follows the staging
  if (t < a.n) {
    x = a.src.hot[hix(kHX, t)];
    y = a.src.hot[hix(kHY, t)];
    z = a.src.hot[hix(kHZ, t)];
  }
  stage_grid(a.g, s_z, s_zz);
  __syncthreads();
  if (t >= a.n) return;
  const Cell c = cell_of(a.g, s_z, s_zz, x, y, z);
  Fluid f = (ILIN == 1) ? gather_lin_f64<true>(a.g, a.F, s_z, s_zz, x, y, z, c)
                        : gather6_f64(a.g, a.F, s_z, s_zz, x, y, z, c);
  const int32_t h = a.src.home[t];
  a.cold[kCUFX * a.cap + h] = f.u[0];
  a.cold[kCUFY * a.cap + h] = f.u[1];
  a.cold[kCUFZ * a.cap + h] = f.u[2];
  a.cold[kCTF * a.cap + h] = f.T;
  a.cold[kCQINF * a.cap + h] = f.q;
  if (set_vp) {
    a.src.hot[hix(kHVX, t)] = f.u[0];
    a.src.hot[hix(kHVY, t)] = f.u[1];
    a.src.hot[hix(kHVZ, t)] = f.u[2];
  }
  if (set_tp) a.src.hot[hix(kHTP, t)] = f.T;
}

// Coupling amounts of the current state from the stored pre-step values
// (particle_coupling_update's per-droplet part).
__global__ void __launch_bounds__(kBlock) k_amounts(StepArgs a) {
  const int64_t t = xcd_block() * kBlock + threadIdx.x;
  if (t >= a.n) return;
  const Drop d = load_drop(a.src.hot, t);
  const int32_t h = a.src.home[t];
  const double vold[3] = {a.cold[kCVXOLD * a.cap + h], a.cold[kCVYOLD * a.cap + h], a.cold[kCVZOLD * a.cap + h]};
  double am[6];
  coupling_amounts(a.C, a.fl, d, a.cold[kCTF * a.cap + h], vold, a.cold[kCTPOLD * a.cap + h],
                   a.cold[kCRADIUSOLD * a.cap + h], a.src.mult[t], am);
#pragma unroll
  for (int k = 0; k < 6; ++k) a.amounts[aix(k, t)] = am[k];
}

// -------------------------------------------------------- coupling scatter

// The scatters read the per-droplet amounts as AT: double, or float where a
// mixed-precision step stored them (their sums stay FP64).
struct ScatterArgs {
  DevGrid g;
  const int32_t* keys;    // u-cell keys of the positions (slot order)
  const double* hot;      // positions (slot order)
  const void* amounts;    // [6][cap] of AT
  int64_t n, cap;
  double* src;            // 6 fields (0:nnz+1, 1:nny+1, 1:nnx+1)
  int64_t nsrc;           // elements per field
  unsigned fmask;         // bit f set: field f has nonzero amounts
};

// Deposit geometry of one droplet (particles.f90:1179-1212): node offsets of
// its 8 corners in the source arrays and the weights wtt/dV.  The nodes hang
// off the u cell (ipt, jpt, kpt) of the droplet's sort key (indices clamped
// into [1,nnx+1] x [1,nny+1] x [0,nnz], which only positions outside the
// domain need); node nnx+2 (x == xl exactly) is the periodic image of 2.
struct Deposit {
  int key;
  int64_t node[8];
  double w[8];
};

__device__ __forceinline__ void key_nodes(const DevGrid& g, int key, int64_t node[8]) {
  const int nz1 = g.nnz + 1, ny1 = g.nny + 1;
  const int k0 = key % nz1;
  const int j0 = (key / nz1) % ny1;
  const int i0 = key / (nz1 * ny1);
  const int64_t sj = g.nnz + 2, si = (int64_t)(g.nnz + 2) * (g.nny + 1);
  int ix[2], iy[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    int xi = i0 + 1 + k, yj = j0 + 1 + k;
    if (xi > g.nnx + 1) xi -= g.nnx;
    if (yj > g.nny + 1) yj -= g.nny;
    ix[k] = xi;
    iy[k] = yj;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) node[(i * 2 + j) * 2 + k] = (int64_t)(k0 + k) + sj * (iy[j] - 1) + si * (ix[i] - 1);
}

// weight wtt/dV of a droplet of u cell (ipt, jpt, kpt) at corner (ci, cj, ck),
// with the reference's expression (particles.f90:1197-1205)
__device__ __forceinline__ double corner_weight(const DevGrid& g, const double* s_zz, int ipt, int jpt, int kpt,
                                                int ci, int cj, int ck, double x, double y, double z) {
  const double dzu1 = g.dzu[kpt + 1];
  const double dV = g.dx * g.dy * dzu1;
  const double wtx = 1.0 - fabs(x - g.dx * (double)(ci + ipt - 1)) / g.dx;
  const double wty = 1.0 - fabs(y - g.dy * (double)(cj + jpt - 1)) / g.dy;
  const double wtz = 1.0 - fabs(z - s_zz[kpt + ck]) / dzu1;
  return wtx * wty * wtz / dV;
}

// The deposit of a droplet in u cell `key` (the clamped cell; every in-domain
// droplet's own cell).
__device__ __forceinline__ Deposit deposit_key(const DevGrid& g, const double* s_zz, int key, double x,
                                               double y, double z) {
  Deposit dp;
  dp.key = key;
  const int nz1 = g.nnz + 1, ny1 = g.nny + 1;
  const int kpt = key % nz1;
  const int jpt = (key / nz1) % ny1 + 1;
  const int ipt = key / (nz1 * ny1) + 1;
  key_nodes(g, dp.key, dp.node);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) dp.w[(i * 2 + j) * 2 + k] = corner_weight(g, s_zz, ipt, jpt, kpt, i, j, k, x, y, z);
  return dp;
}

__device__ __forceinline__ Deposit deposit_of(const DevGrid& g, const double* s_z, const double* s_zz,
                                              double x, double y, double z) {
  return deposit_key(g, s_zz, ukey_of(g, cell_of(g, s_z, s_zz, x, y, z)), x, y, z);
}

// Two-way coupling scatter: every droplet deposits amount*wtt/dV on its 8
// nodes.  The droplets are cell sorted, so within a wave runs of equal u-cell
// keys share all 8 nodes:
//  1. a segmented inclusive scan (log2 of the longest run shuffle steps) sums
//     each run into its last lane;
//  2. the run sums are compacted to lanes 0..nseg-1 in key order, so that one
//     atomic instruction per (field, corner) covers consecutive cells of a z
//     column, i.e. consecutive addresses of the k-fastest source arrays;
//  3. a run whose cell sits directly above the previous run's cell shares its
//     bottom nodes with that run's top nodes, which are folded in first.
// Each wave then issues at most 48 FP64 global atomic instructions.
template <typename AT>
__global__ void __launch_bounds__(kBlock) k_scatter(ScatterArgs a) {
  const AT* amounts = static_cast<const AT*>(a.amounts);
  __shared__ double s_z[kMaxLevels], s_zz[kMaxLevels];
  __shared__ int s_tail[kBlock];
  stage_grid(a.g, s_z, s_zz);
  __syncthreads();
  const int64_t t = xcd_block() * kBlock + threadIdx.x;
  const bool valid = t < a.n;
  const int lane = lane_id();
  int* my_tail = s_tail + (threadIdx.x & ~63);
  Deposit dp;
  if (valid) {
    const double x = a.hot[hix(kHX, t)], y = a.hot[hix(kHY, t)], z = a.hot[hix(kHZ, t)];
    dp = a.keys ? deposit_key(a.g, s_zz, a.keys[t], x, y, z) : deposit_of(a.g, s_z, s_zz, x, y, z);
  } else {
    dp.key = -1 - lane;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      dp.node[c] = 0;
      dp.w[c] = 0.0;
    }
  }
  const int prev = __shfl_up(dp.key, 1, 64);
  const bool head = (lane == 0) || (prev != dp.key);
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const int seg_start = 63 - __clzll((long long)(heads & upto));
  const bool tail = (lane == 63) || ((heads >> (lane + 1)) & 1ull);
  const unsigned long long tails = __ballot(tail && valid);
  const int nseg = __popcll(tails);
  if (tail && valid) my_tail[__popcll(tails & ((1ull << lane) - 1ull))] = lane;
  int maxlen = lane - seg_start + 1;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, o, 64));
  int nsteps = 0;
  while ((1 << nsteps) < maxlen) ++nsteps;
  __syncthreads();
  // compacted run s lives in lane s
  const bool act = lane < nseg;
  const int src_lane = act ? my_tail[lane] : 0;
  const int skey = __shfl(dp.key, src_lane, 64);
  const int pkey = __shfl_up(skey, 1, 64);
  const int nkey = __shfl_down(skey, 1, 64);
  const int nz1 = a.g.nnz + 1;
  // the cell directly below / above in the same z column
  const bool below = act && lane > 0 && pkey == skey - 1 && (skey % nz1) != 0;
  const bool above = act && lane + 1 < nseg && nkey == skey + 1 && (nkey % nz1) != 0;
  int64_t node[8];
  key_nodes(a.g, act ? skey : 0, node);
  double am6[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // the six loads issued together
  if (valid) {
#pragma unroll
    for (int f = 0; f < 6; ++f) am6[f] = (double)amounts[aix(f, t)];
  }
  for (int f = 0; f < 6; ++f) {
    if (!((a.fmask >> f) & 1u)) continue;
    const double am = am6[f];
    double v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) v[c] = am * dp.w[c];
    for (int s = 0; s < nsteps; ++s) {
      const int off = 1 << s;
      const bool take = (lane - off) >= seg_start;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const double o = __shfl_up(v[c], off, 64);
        if (take) v[c] += o;
      }
    }
    double* fld = a.src + (int64_t)f * a.nsrc;
#pragma unroll
    for (int col = 0; col < 4; ++col) {
      double bot = __shfl(v[2 * col], src_lane, 64);
      const double top = __shfl(v[2 * col + 1], src_lane, 64);
      const double top_below = __shfl_up(top, 1, 64);
      if (below) bot += top_below;
      if (act) atomicAdd(&fld[node[2 * col]], bot);
      if (act && !above) atomicAdd(&fld[node[2 * col + 1]], top);
    }
  }
}

// Deterministic coupling without atomics: each source node (one thread; the
// threads of a wave are consecutive k of one node column) sums, in a fixed
// order, the deposits of the droplets of the u cells that touch it, walking
// the fresh cell sort (offsets + perm).  Every node is written exactly once,
// so the fields need no clearing, and the result is the same on every run.
struct PullArgs {
  DevGrid g;
  const double* hot;
  const void* amounts;     // [6][cap] of AT
  const int32_t* perm;     // sorted index -> slot
  const int32_t* offsets;  // u-cell offsets
  int64_t cap;
  double* src;
  int64_t nsrc;
  unsigned fmask;
};

template <typename AT>
__global__ void __launch_bounds__(kBlock) k_scatter_pull(PullArgs a) {
  const AT* amounts = static_cast<const AT*>(a.amounts);
  __shared__ double s_zz[kMaxLevels];
  for (int k = threadIdx.x; k < a.g.nnz + 2; k += blockDim.x) s_zz[k] = a.g.zz[k];
  __syncthreads();
  const int64_t node = xcd_block() * kBlock + threadIdx.x;
  if (node >= a.nsrc) return;
  const int nz2 = a.g.nnz + 2;
  const int kz = (int)(node % nz2);
  const int64_t rest = node / nz2;
  const int iy = (int)(rest % (a.g.nny + 1)) + 1;
  const int ix = (int)(rest / (a.g.nny + 1)) + 1;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  // cells along x touching node ix: cell i0 = ix-1-ci at corner ci; node 2
  // also takes corner 1 of cell nnx (its periodic image nnx+2)
  int cx[3], ci[3], ncx = 0, cy[3], cj[3], ncy = 0;
  for (int i = 0; i < 2; ++i) {
    const int i0 = ix - 1 - i;
    if (i0 >= 0 && i0 <= a.g.nnx) { cx[ncx] = i0; ci[ncx] = i; ++ncx; }
  }
  if (ix == 2) { cx[ncx] = a.g.nnx; ci[ncx] = 1; ++ncx; }
  for (int j = 0; j < 2; ++j) {
    const int j0 = iy - 1 - j;
    if (j0 >= 0 && j0 <= a.g.nny) { cy[ncy] = j0; cj[ncy] = j; ++ncy; }
  }
  if (iy == 2) { cy[ncy] = a.g.nny; cj[ncy] = 1; ++ncy; }
  for (int xa = 0; xa < ncx; ++xa)
    for (int ya = 0; ya < ncy; ++ya)
      for (int ck = 0; ck < 2; ++ck) {
        const int k0 = kz - ck;
        if (k0 < 0 || k0 > a.g.nnz) continue;
        const int key = k0 + (a.g.nnz + 1) * (cy[ya] + (a.g.nny + 1) * cx[xa]);
        const int b = a.offsets[key], e = a.offsets[key + 1];
        for (int p = b; p < e; ++p) {
          const int64_t s = a.perm[p];
          const double w = corner_weight(a.g, s_zz, cx[xa] + 1, cy[ya] + 1, k0, ci[xa], cj[ya], ck,
                                         a.hot[hix(kHX, s)], a.hot[hix(kHY, s)], a.hot[hix(kHZ, s)]);
          double am[6];  // loads issued together, masked after they land
#pragma unroll
          for (int f = 0; f < 6; ++f) am[f] = (double)amounts[aix(f, s)];
#pragma unroll
          for (int f = 0; f < 6; ++f)
            if ((a.fmask >> f) & 1u) acc[f] += am[f] * w;
        }
      }
#pragma unroll
  for (int f = 0; f < 6; ++f) a.src[(int64_t)f * a.nsrc + node] = acc[f];
}

// Tiled coupling scatter without global atomics.  Each workgroup owns a box
// of TI x TJ x TK source nodes (node columns ix, iy; levels kz) and keeps its
// 6 fields in LDS.  The droplets that can reach the box are those of the u
// cells ix-2..ix-1 along x (plus cell nnx for the periodic image of node 2),
// likewise along y, and kz-1..kz along z; in the fresh cell sort (offsets +
// perm) the cells kz-1..kz of one (x, y) cell column are one contiguous run,
// so the box reads (TI+2)(TJ+2) runs at most.  The runs are laid end to end
// and dealt to the threads kU droplets at a time: a thread first loads the
// kU slots from perm, then the kU droplets' positions, keys and amounts, and
// only then adds their in-box corners into LDS.  The kernel is bound by the
// latency of those loads and by the FP64 LDS atomics, so it is kept small (67
// VGPRs, 25 KB of LDS) to keep six workgroups resident per CU.  The box is
// finally stored with plain coalesced writes.  Every node belongs to exactly
// one box, so the fields need no clearing.  A droplet near a box face is read
// by the neighbouring boxes too: the read amplification is
// (TI+1)(TJ+1)(TK+1)/(TI TJ TK).
struct TileArgs {
  DevGrid g;
  const double* hot;       // positions (slot order)
  const void* amounts;     // [6][cap] of AT
  const int32_t* keys;     // u-cell keys (slot order)
  const int32_t* perm;     // sorted index -> slot
  const int32_t* offsets;  // u-cell offsets of the sort
  int64_t cap;
  double* src;
  int64_t nsrc;
  unsigned fmask;
  int ntx, nty, ntz;  // boxes along x, y, z
};

// Exclusive prefix over the block of one int per thread (kBlock threads);
// returns the thread's prefix and leaves the block total in *total.
__device__ __forceinline__ int block_exclusive_scan(int v, int* s_wave, int* total) {
  const int lane = lane_id(), w = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) s_wave[w] = inc;
  __syncthreads();
  int base = 0, sum = 0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k) {
    base += (k < w) ? s_wave[k] : 0;
    sum += s_wave[k];
  }
  *total = sum;
  return base + inc - v;
}

// LDS slot of box node (li, lj, lk): row li*TJ + (lj ^ li), level
// lk ^ ((li + 3 lj) mod TK), a permutation of each x plane's rows and of each
// row's levels.  Without it the 8 corners that neighbouring lanes add land on
// few LDS banks (the slot's bank depends on lj and lk only); with it the
// scatter measured 0.93 ms against 0.96 (swizzling the levels alone: 0.94).
template <int TI, int TJ, int TK>
__device__ __forceinline__ int box_slot(int li, int lj, int lk) {
  static_assert((TJ & (TJ - 1)) == 0 && (TK & (TK - 1)) == 0 && TI <= TJ, "power-of-two box rows and levels");
  return (li * TJ + (lj ^ (li & (TJ - 1)))) * TK + (lk ^ ((li + lj * 3) & (TK - 1)));
}

template <int TI, int TJ, int TK, typename AT>
__global__ void __launch_bounds__(kBlock) k_scatter_box(TileArgs a) {
  const AT* amounts = static_cast<const AT*>(a.amounts);
  // node (li, lj, lk) of the box at box_slot(li, lj, lk) (25 KB of LDS for
  // 8x8x8: six blocks fit a CU; an odd pitch, against bank conflicts, made
  // it 28 KB and five blocks, which measured slower than the permutation)
  constexpr int kNodes = TI * TJ * TK;
  constexpr int kMaxSeg = (TI + 2) * (TJ + 2);
  constexpr int kU = 1;  // droplets per thread per pass (1 to 8 measured; 1 fastest)
  static_assert(kMaxSeg <= kBlock, "one cell column per thread in the run table");
  __shared__ double s_acc[6 * kNodes];
  __shared__ int s_beg[kMaxSeg], s_pre[kMaxSeg + 1];
  __shared__ int s_wave[kBlock / 64];
  // the box's cell levels klo..khi: zz(k) (also k = khi + 1), 1/dzu(k+1) and
  // 1/dV(k+1), so that a droplet's weights wait on no further global load
  __shared__ double s_lev[3][TK + 2];
  const int64_t b = xcd_block();
  if (b >= (int64_t)a.ntx * a.nty * a.ntz) return;  // whole block
  // boxes with k fastest: consecutive boxes of one XCD share their droplets'
  // runs and node columns in L2 (the orders with x or y fastest, and boxes
  // of 1-2 x 8 x 32-130 nodes, measured 1.6-3.7x slower)
  const int kt = (int)(b % a.ntz);
  const int jt = (int)((b / a.ntz) % a.nty);
  const int it = (int)(b / ((int64_t)a.ntz * a.nty));
  const int nnx = a.g.nnx, nny = a.g.nny, nnz = a.g.nnz;
  const int X0 = 1 + it * TI, X1 = min(X0 + TI - 1, nnx + 1);
  const int Y0 = 1 + jt * TJ, Y1 = min(Y0 + TJ - 1, nny + 1);
  const int K0 = kt * TK, K1 = min(K0 + TK - 1, nnz + 1);
  for (int l = threadIdx.x; l < 6 * kNodes; l += blockDim.x) s_acc[l] = 0.0;
  // x cells whose nodes reach [X0, X1]: node ix takes cells ix-1 and ix-2;
  // node 2 also takes cell nnx through the periodic image nnx+2
  const int xlo = max(0, X0 - 2), xhi = min(nnx, X1 - 1);
  const int ylo = max(0, Y0 - 2), yhi = min(nny, Y1 - 1);
  const int ncx = xhi - xlo + 1 + ((X0 <= 2 && X1 >= 2 && xhi < nnx) ? 1 : 0);
  const int ncy = yhi - ylo + 1 + ((Y0 <= 2 && Y1 >= 2 && yhi < nny) ? 1 : 0);
  const int nseg = ncx * ncy;
  const int klo = max(0, K0 - 1), khi = min(nnz, K1);
  const int nz1 = nnz + 1, ny1 = nny + 1;
  if (threadIdx.x < khi - klo + 2) {
    const int k = klo + (int)threadIdx.x;
    s_lev[0][threadIdx.x] = a.g.zz[k];
    if (k <= khi) {
      s_lev[1][threadIdx.x] = a.g.rdzu64[k + 1];  // 1.0 / dzu(k+1), formed on the host
      s_lev[2][threadIdx.x] = a.g.rdVu64[k + 1];  // 1.0 / (dx * dy * dzu(k+1)), likewise
    }
  }
  int len = 0, beg = 0;
  if (threadIdx.x < nseg) {
    const int ci = threadIdx.x / ncy, cj = threadIdx.x % ncy;
    const int cx = (ci <= xhi - xlo) ? xlo + ci : nnx;
    const int cy = (cj <= yhi - ylo) ? ylo + cj : nny;
    const int base = nz1 * (cy + ny1 * cx);
    beg = a.offsets[base + klo];
    len = a.offsets[base + khi + 1] - beg;
  }
  int total;
  const int pre = block_exclusive_scan(len, s_wave, &total);
  if (threadIdx.x < nseg) {
    s_beg[threadIdx.x] = beg;
    s_pre[threadIdx.x] = pre;
  }
  if (threadIdx.x == 0) s_pre[nseg] = total;
  __syncthreads();
  const double dx = a.g.dx, dy = a.g.dy;
  const double rdx = a.g.rdx64, rdy = a.g.rdy64;  // 1.0 / dx, 1.0 / dy
  const int rot = lane_id() & 7;
  int step0 = 1;  // the largest power of two with 2 * step0 >= nseg
  while (2 * step0 < nseg) step0 *= 2;
  for (int p0 = threadIdx.x; p0 < total; p0 += kU * kBlock) {
    // the run of p: the last run whose prefix is <= p, so that
    // s_pre[lo] <= p < s_pre[lo + 1] (runs of length 0 repeat a prefix value
    // and are never chosen); a fixed number of steps, the droplets' searches
    // interleaved
    int lo[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) lo[u] = 0;
    for (int st = step0; st > 0; st >>= 1) {
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int p = p0 + u * kBlock;
        if (lo[u] + st < nseg && s_pre[lo[u] + st] <= p) lo[u] += st;
      }
    }
    int32_t slot[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int p = p0 + u * kBlock;
      slot[u] = p < total ? a.perm[s_beg[lo[u]] + (p - s_pre[lo[u]])] : -1;
    }
    int key[kU];
    double x[kU], y[kU], z[kU], am[kU][6];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (slot[u] < 0) continue;
      const int64_t t = slot[u];
      key[u] = a.keys[t];
      x[u] = a.hot[hix(kHX, t)];
      y[u] = a.hot[hix(kHY, t)];
      z[u] = a.hot[hix(kHZ, t)];
      // the six amount loads are issued together with x, y and z (every
      // step writes all six); a load under the field-mask test would be
      // compiled as a branch that waits for it before the next one is issued
#pragma unroll
      for (int f = 0; f < 6; ++f) am[u][f] = (double)amounts[aix(f, t)];
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int f = 0; f < 6; ++f)
        if (!((a.fmask >> f) & 1u)) am[u][f] = 0.0;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      if (slot[u] < 0) continue;
      const int kpt = key[u] % nz1;
      const int j0 = (key[u] / nz1) % ny1;
      const int i0 = key[u] / (nz1 * ny1);
      // corner weights of particles.f90:1197-1205 (divisions by dx, dy,
      // dzu(kpt+1) and dV as products with their reciprocals)
      const int kl = kpt - klo;  // the droplets come from levels klo..khi
      const double rdz = s_lev[1][kl];
      const double rdV = s_lev[2][kl];
      double wtx[2], wty[2], wtz[2];
      int lx[2], ly[2], lz[2];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        wtx[c] = 1.0 - fabs(x[u] - dx * (double)(c + i0)) * rdx;
        wty[c] = 1.0 - fabs(y[u] - dy * (double)(c + j0)) * rdy;
        wtz[c] = 1.0 - fabs(z[u] - s_lev[0][kl + c]) * rdz;
        int xi = i0 + 1 + c, yj = j0 + 1 + c;
        if (xi > nnx + 1) xi -= nnx;
        if (yj > nny + 1) yj -= nny;
        lx[c] = (xi >= X0 && xi <= X1) ? xi - X0 : -1;
        ly[c] = (yj >= Y0 && yj <= Y1) ? yj - Y0 : -1;
        const int kz = kpt + c;
        lz[c] = (kz >= K0 && kz <= K1) ? kz - K0 : -1;
      }
      // Neighbouring lanes hold droplets of the same or adjacent cells, whose
      // corners are the same LDS words; each lane starts at a different
      // corner so that one atomic instruction rarely hits a word twice.
#pragma unroll
      for (int s8 = 0; s8 < 8; ++s8) {
        const int cc = (s8 + rot) & 7;
        const int ci = cc >> 2, cj = (cc >> 1) & 1, ck = cc & 1;
        const int li = ci ? lx[1] : lx[0], lj = cj ? ly[1] : ly[0], lk = ck ? lz[1] : lz[0];
        if (li < 0 || lj < 0 || lk < 0) continue;
        const double w = (ci ? wtx[1] : wtx[0]) * (cj ? wty[1] : wty[0]) * (ck ? wtz[1] : wtz[0]) * rdV;
        const int l = box_slot<TI, TJ, TK>(li, lj, lk);
#pragma unroll
        for (int f = 0; f < 6; ++f)
          if ((a.fmask >> f) & 1u) atomicAdd(&s_acc[f * kNodes + l], am[u][f] * w);
      }
    }
  }
  __syncthreads();
  // store the box: z fastest, so consecutive threads write consecutive levels
  const int bk = K1 - K0 + 1, nj = Y1 - Y0 + 1, ni = X1 - X0 + 1;
  const int64_t sj = nnz + 2, si = (int64_t)(nnz + 2) * (nny + 1);
  for (int l = threadIdx.x; l < ni * nj * bk; l += blockDim.x) {
    const int lk = l % bk, lj = (l / bk) % nj, li = l / (bk * nj);
    const int64_t node = (int64_t)(K0 + lk) + sj * (Y0 + lj - 1) + si * (X0 + li - 1);
    const int ll = box_slot<TI, TJ, TK>(li, lj, lk);
#pragma unroll
    for (int f = 0; f < 6; ++f) a.src[(int64_t)f * a.nsrc + node] = s_acc[f * kNodes + ll];
  }
}

// ------------------------------------------- coupling exchange and apply

// particle_coupling_exchange (particles.f90:1066-1149) on one periodic
// domain, fused with its ztox_trans: the +1 ghost plane x = nnx+1 folds onto
// x = 1 (all rows, the ghost row included), then the ghost row y = nny+1 onto
// y = 1, in the reference's order of additions; the result is written in the
// flow solver's x-pencil layout xsrc[f][k][iy][ix] (k = 0..nnz+1).  One
// workgroup transposes a 64 (k) x 32 (ix) tile of one row iy through LDS so
// that both the k-fastest reads and the x-fastest writes are coalesced.
constexpr int kFoldK = 64, kFoldX = 32;

__global__ void __launch_bounds__(kBlock) k_fold_sources(const double* src, int64_t nsrc, int nnx, int nny, int nnz,
                                                         double* xsrc) {
  __shared__ double tile[kFoldX][kFoldK + 1];
  const int nz2 = nnz + 2;
  const int ntk = (nz2 + kFoldK - 1) / kFoldK, ntx = (nnx + kFoldX - 1) / kFoldX;
  int64_t b = blockIdx.x;
  const int xt = (int)(b % ntx);
  b /= ntx;
  const int kt = (int)(b % ntk);
  b /= ntk;
  const int iy = (int)(b % nny) + 1;
  const int f = (int)(b / nny);
  const double* s = src + (int64_t)f * nsrc;
  const int64_t sj = nz2, si = (int64_t)nz2 * (nny + 1);
  const int K0 = kt * kFoldK, X0 = 1 + xt * kFoldX;
  for (int e = threadIdx.x; e < kFoldK * kFoldX; e += blockDim.x) {
    const int kk = e % kFoldK, xx = e / kFoldK;
    const int k = K0 + kk, ix = X0 + xx;
    if (k < nz2 && ix <= nnx) {
      double v = s[k + sj * (iy - 1) + si * (ix - 1)];
      if (ix == 1) v = v + s[k + sj * (iy - 1) + si * nnx];
      if (iy == 1) {
        double w = s[k + sj * nny + si * (ix - 1)];
        if (ix == 1) w = w + s[k + sj * nny + si * nnx];
        v = v + w;
      }
      tile[xx][kk] = v;
    }
  }
  __syncthreads();
  double* o = xsrc + (int64_t)f * nz2 * nny * nnx;
  for (int e = threadIdx.x; e < kFoldK * kFoldX; e += blockDim.x) {
    const int xx = e % kFoldX, kk = e / kFoldX;
    const int k = K0 + kk, ix = X0 + xx;
    if (k < nz2 && ix <= nnx) o[(ix - 1) + (int64_t)nnx * ((iy - 1) + (int64_t)nny * k)] = tile[xx][kk];
  }
}

// apply_particle_coupling (particles.f90:1278-1313): t(ix,iy,1,iz) and
// t(ix,iy,2,iz) += dt * sources for iz = 1..nnz, on T and q arrays in the
// x-pencil layout (1:nnx+2, 1:nny, 0:nnz+1) or the core layout (0:nnz+1,
// 1:nny, 1:nnx).
__global__ void k_apply_coupling(double* T, double* q, const double* xsrc, DevGrid g, DevConsts C, DevFlags fl,
                                 int xpencil) {
  const int nx = g.nnx, ny = g.nny, nz = g.nnz;
  const int64_t total = (int64_t)nx * ny * nz;
  const int64_t nf = (int64_t)(nz + 2) * ny * nx;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    int ix, iy, iz;
    if (xpencil) {
      ix = (int)(e % nx) + 1;
      iy = (int)((e / nx) % ny) + 1;
      iz = (int)(e / ((int64_t)nx * ny)) + 1;
    } else {
      iz = (int)(e % nz) + 1;
      iy = (int)((e / nz) % ny) + 1;
      ix = (int)(e / ((int64_t)nz * ny)) + 1;
    }
    const int64_t xs = (ix - 1) + (int64_t)nx * ((iy - 1) + (int64_t)ny * iz);
    const int64_t t = xpencil ? (ix - 1) + (int64_t)(nx + 2) * ((iy - 1) + (int64_t)ny * iz)
                              : iz + (int64_t)(nz + 2) * ((iy - 1) + (int64_t)ny * (ix - 1));
    double Tv = T[t];
    if (fl.iTcouple == 1) {
      if (fl.iexner) Tv = Tv + C.dt * (xsrc[3 * nf + xs] / exner(C, func_p_base(C, g.zz[iz])));
      else Tv = Tv + C.dt * xsrc[3 * nf + xs];
    }
    if (fl.iHcouple == 1) {
      Tv = Tv + C.dt * xsrc[5 * nf + xs];
      q[t] = q[t] + C.dt * xsrc[4 * nf + xs];
    }
    T[t] = Tv;
  }
}

// ---------------------------------------------------------------- binning

struct BinArgs {
  DevGrid g;
  const double* hot;
  int64_t n, cap;
  int32_t* keys;
  int32_t* counts;
  int32_t* cellidx;  // optional [n][4] in slot order
};

__global__ void __launch_bounds__(kBlock) k_keys(BinArgs a) {
  __shared__ double s_z[kMaxLevels], s_zz[kMaxLevels];
  stage_grid(a.g, s_z, s_zz);
  __syncthreads();
  const int64_t t = xcd_block() * kBlock + threadIdx.x;
  const bool valid = t < a.n;
  int key = -1;
  if (valid) {
    const Cell c = cell_of(a.g, s_z, s_zz, a.hot[hix(kHX, t)], a.hot[hix(kHY, t)], a.hot[hix(kHZ, t)]);
    key = ukey_of(a.g, c);
    a.keys[t] = key;
    if (a.cellidx) {
      a.cellidx[4 * t + 0] = c.ipt;
      a.cellidx[4 * t + 1] = c.jpt;
      a.cellidx[4 * t + 2] = c.kpt;
      a.cellidx[4 * t + 3] = c.kwpt;
    }
  }
  count_key(a.counts, key, valid);
}

// calc_tauc_min's num_array over w cells (particles.f90:3320-3339).
__global__ void __launch_bounds__(kBlock) k_wcells(BinArgs a, int32_t* num) {
  __shared__ double s_z[kMaxLevels], s_zz[kMaxLevels];
  stage_grid(a.g, s_z, s_zz);
  __syncthreads();
  const int64_t t = xcd_block() * kBlock + threadIdx.x;
  if (t >= a.n) return;
  const double x = a.hot[hix(kHX, t)], y = a.hot[hix(kHY, t)], z = a.hot[hix(kHZ, t)];
  const Cell c = cell_of(a.g, s_z, s_zz, x, y, z);
  const int kc = wcell_of(a.g, s_z, z);
  if (c.ipt < 1 || c.ipt > a.g.nnx || c.jpt < 1 || c.jpt > a.g.nny || kc < 1 || kc > a.g.nnz) return;
  atomicAdd(&num[(c.ipt - 1) + (int64_t)a.g.nnx * ((c.jpt - 1) + (int64_t)a.g.nny * (kc - 1))], 1);
}

// ------------------------------------------------------------ statistics

// calc_tauc_min (particles.f90:3310-3368), first half: radius sum, count and
// multiplicity sum per w cell (ipt, jpt, kc); droplets outside (1:nnx,
// 1:nny, 1:nnz) are not counted.
__global__ void __launch_bounds__(kBlock) k_tauc_accum(BinArgs a, const int64_t* mult, double* rad, int32_t* num,
                                                       double* nc) {
  __shared__ double s_z[kMaxLevels], s_zz[kMaxLevels];
  stage_grid(a.g, s_z, s_zz);
  __syncthreads();
  const int64_t t = xcd_block() * kBlock + threadIdx.x;
  if (t >= a.n) return;
  const double x = a.hot[hix(kHX, t)], y = a.hot[hix(kHY, t)], z = a.hot[hix(kHZ, t)];
  const Cell c = cell_of(a.g, s_z, s_zz, x, y, z);
  const int kc = wcell_of(a.g, s_z, z);
  if (c.ipt < 1 || c.ipt > a.g.nnx || c.jpt < 1 || c.jpt > a.g.nny || kc < 1 || kc > a.g.nnz) return;
  const int64_t e = (c.ipt - 1) + (int64_t)a.g.nnx * ((c.jpt - 1) + (int64_t)a.g.nny * (kc - 1));
  atomicAdd(&rad[e], a.hot[hix(kHR, t)]);
  atomicAdd(&num[e], 1);
  atomicAdd(&nc[e], (double)mult[t]);
}

// Second half: the smallest 1/(denom + 1e-10) over occupied cells, as the
// bit pattern of a positive double (ordered like the value) in *out.
__global__ void __launch_bounds__(kBlock) k_tauc_min(DevGrid g, DevConsts C, const double* rad, const int32_t* num,
                                                     const double* nc, unsigned long long* out) {
  __shared__ double s_min[kBlock / 64];
  const int64_t ncell = (int64_t)g.nnx * g.nny * g.nnz;
  double m = 1.0e10;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ncell; e += (int64_t)gridDim.x * blockDim.x) {
    if (num[e] != 0) {
      const int iz = (int)(e / ((int64_t)g.nnx * g.nny)) + 1;
      const double radavg = rad[e] / (double)num[e];
      const double Nc = nc[e] / (g.dx * g.dy * g.dzw[iz]);
      const double denom = C.pi2 * (C.nuf / C.Sc) * 2.0 * radavg * Nc;
      m = fmin(m, 1.0 / (denom + 1.0e-10));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmin(m, __shfl_xor(m, o, 64));
  if (lane_id() == 0) s_min[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmin(m, s_min[w]);
    atomicMin(out, (unsigned long long)__double_as_longlong(m));
  }
}

// particle_xy_stats (particles.f90:3370-3770): the droplet loops.  Each of
// gridDim.x workgroups walks its own contiguous run of slots and keeps
//  - the 0th-order statistics per thread (counts, 6 sums, 6 maxima, 6
//    minima, starting from the reference's initial values), reduced in LDS;
//  - quantities q0..q0+nq-1 of the 24 per-level sums of statsvec1 in LDS;
// and writes them to its row of `part`; LWP(ipt, jpt) is added with global
// atomics.  k_xy_reduce then sums the rows in a fixed order.
constexpr int kStats0 = 21;
struct StatsArgs {
  DevGrid g;
  DevConsts C;
  const double* hot;
  const int64_t* mult;
  const int32_t* home;
  const double* cold;
  int64_t n, cap;
  int q0, nq;        // per-level quantities of this pass
  double* part0;     // [gridDim.x][kStats0] (pass 0 only)
  double* part1;     // [gridDim.x][nq][nnz+2]
  double* lwp;       // [nny][nnx] (pass 0 only)
};

__global__ void __launch_bounds__(kBlock) k_xy_stats(StatsArgs a) {
  extern __shared__ double s_lvl[];  // [nq][nnz+2]
  __shared__ double s_z[kMaxLevels], s_zz[kMaxLevels];
  __shared__ double s_red[kBlock / 64][kStats0];
  stage_grid(a.g, s_z, s_zz);
  const int nz2 = a.g.nnz + 2;
  for (int i = threadIdx.x; i < a.nq * nz2; i += blockDim.x) s_lvl[i] = 0.0;
  __syncthreads();
  const bool first = a.q0 == 0;
  double v0[kStats0] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0.0, 0.0, 0.0, -1000.0, -1000.0, -1000.0,
                        1000.0, 1000.0, 1000.0, 1000.0, 1000.0, 1000.0};
  const int64_t chunk = (a.n + gridDim.x - 1) / gridDim.x;
  const int64_t t0 = (int64_t)blockIdx.x * chunk, t1 = min(a.n, t0 + chunk);
  const double pi = 4.0 * atan(1.0);
  for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const Drop d = load_drop(a.hot, t);
    const int64_t h = a.home[t];
    const double uf0 = a.cold[kCUFX * a.cap + h], uf1 = a.cold[kCUFY * a.cap + h], uf2 = a.cold[kCUFZ * a.cap + h];
    const double Tf = a.cold[kCTF * a.cap + h], qinf = a.cold[kCQINF * a.cap + h];
    const double m = (double)a.mult[t];
    const double mw = a.C.rhow * 4.0 / 3.0 * pi * pow(d.r, 3);
    if (first) {
      v0[0] += 1.0;
      if (d.r > d.rc) v0[1] += 1.0;
      else v0[2] += 1.0;
      const double Volp = a.C.pi2 * 2.0 / 3.0 * pow(d.r, 3);
      const double rhop = (d.m_s + Volp * a.C.rhow) / Volp;
      const double d0 = d.vx - uf0, d1 = d.vy - uf1, d2 = d.vz - uf2;
      v0[3] += 2.0 * d.r * sqrt(d0 * d0 + d1 * d1 + d2 * d2) / a.C.nuf;
      v0[4] += Volp * a.C.rhow;
      v0[5] += Volp * rhop;
      v0[6] += Volp;
      v0[7] += d.r;
      v0[8] += d.r * d.r;
      const double mm[6] = {d.r, d.Tp, qinf, d.vx, d.vy, d.vz};
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        v0[9 + k] = fmax(v0[9 + k], mm[k]);
        v0[15 + k] = fmin(v0[15 + k], mm[k]);
      }
    }
    const Cell c = cell_of(a.g, s_z, s_zz, d.x, d.y, d.z);
    const int kc = wcell_of(a.g, s_z, d.z);
    if (kc >= 0) {
      const double q[24] = {1.0, d.vx, d.vy, d.vz, d.vx * d.vx, d.vy * d.vy, d.vz * d.vz, d.vx * d.vz,
                            uf0, uf1, uf2, uf0 * uf0, uf1 * uf1, uf2 * uf2, d.Tp, d.Tp * d.Tp, Tf,
                            d.r, d.r * d.r, qinf, d.Tp * d.vz, m, m * mw, a.cold[kCQSTAR * a.cap + h]};
#pragma unroll
      for (int k = 0; k < 24; ++k)
        if (k >= a.q0 && k < a.q0 + a.nq) atomicAdd(&s_lvl[(k - a.q0) * nz2 + kc], q[k]);
    }
    if (first && c.ipt >= 1 && c.ipt <= a.g.nnx && c.jpt >= 1 && c.jpt <= a.g.nny)
      atomicAdd(&a.lwp[(c.ipt - 1) + (int64_t)a.g.nnx * (c.jpt - 1)], m * mw);
  }
  if (first) {
#pragma unroll
    for (int k = 0; k < kStats0; ++k) {
      double v = v0[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_xor(v, o, 64);
        v = (k < 9) ? v + w : (k < 15 ? fmax(v, w) : fmin(v, w));
      }
      if (lane_id() == 0) s_red[threadIdx.x >> 6][k] = v;
    }
  }
  __syncthreads();
  if (first && threadIdx.x < kStats0) {
    const int k = threadIdx.x;
    double v = s_red[0][k];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      v = (k < 9) ? v + s_red[w][k] : (k < 15 ? fmax(v, s_red[w][k]) : fmin(v, s_red[w][k]));
    a.part0[(int64_t)blockIdx.x * kStats0 + k] = v;
  }
  for (int i = threadIdx.x; i < a.nq * nz2; i += blockDim.x) a.part1[(int64_t)blockIdx.x * a.nq * nz2 + i] = s_lvl[i];
}

// Sums (maxima, minima) of the workgroup rows in row order.
__global__ void k_xy_reduce(const double* part, int rows, int width, double* out, int kind_split0, int kind_split1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  double v = part[i];
  for (int r = 1; r < rows; ++r) {
    const double w = part[(int64_t)r * width + i];
    v = (i < kind_split0) ? v + w : (i < kind_split1 ? fmax(v, w) : fmin(v, w));
  }
  out[i] = v;
}

// Exclusive scan of the u-cell histogram: per-block sums, one-block scan of
// the block sums, then a block-local scan with the block's prefix.
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int block_exclusive_scan(int v, int* s_wave, int& total) {
  const int lane = lane_id(), wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[wid] = x;
  __syncthreads();
  if (wid == 0) {
    int w = (lane < nw) ? s_wave[lane] : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      int y = __shfl_up(w, o, 64);
      if (lane >= o) w += y;
    }
    if (lane < nw) s_wave[lane] = w;
  }
  __syncthreads();
  total = s_wave[nw - 1];
  const int before = (wid > 0) ? s_wave[wid - 1] : 0;
  __syncthreads();
  return before + x - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const int32_t* in, int64_t n, int32_t* bsum) {
  __shared__ int s_wave[16];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int v = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + (int64_t)k * kScanThreads + threadIdx.x;
    if (i < n) v += in[i];
  }
  int total;
  block_exclusive_scan(v, s_wave, total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_bsum(int32_t* bsum, int64_t nb) {
  __shared__ int s_wave[16];
  int carry = 0;
  for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
    const int64_t i = b0 + threadIdx.x;
    const int v = (i < nb) ? bsum[i] : 0;
    int total;
    const int ex = block_exclusive_scan(v, s_wave, total);
    if (i < nb) bsum[i] = carry + ex;
    carry += total;
  }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_down(const int32_t* in, int64_t n, const int32_t* bsum,
                                                            int32_t* out) {
  __shared__ int s_wave[16];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  // thread-contiguous items for a cheap serial prefix
  const int64_t i0 = base + (int64_t)threadIdx.x * kScanItems;
  static_assert(kScanItems == 8, "two 16-byte vectors per thread");
  // 16-byte loads and stores away from the end (in and out come from
  // hipMalloc, so in + i0 and out + i0 are 32-byte aligned)
  const bool whole = i0 + kScanItems <= n;
  int v[kScanItems], sum = 0;
  if (whole) {
    const int4 a = reinterpret_cast<const int4*>(in + i0)[0], b = reinterpret_cast<const int4*>(in + i0)[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] = (i0 + k < n) ? in[i0 + k] : 0;
  }
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) sum += v[k];
  int total;
  int run = block_exclusive_scan(sum, s_wave, total) + bsum[blockIdx.x];
  int o[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    o[k] = run;
    run += v[k];
  }
  if (whole) {
    reinterpret_cast<int4*>(out + i0)[0] = make_int4(o[0], o[1], o[2], o[3]);
    reinterpret_cast<int4*>(out + i0)[1] = make_int4(o[4], o[5], o[6], o[7]);
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
      if (i0 + k < n) out[i0 + k] = o[k];
  }
  // grand total at out[n]
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = bsum[blockIdx.x] + total;
}

// Counting-sort scatter: slot t goes to offsets[key] + its rank among the
// cell's droplets, taken from counts[key] (the cell's count, run down to zero
// by atomicSub).  A run of equal keys in lane order takes one atomic; slots
// with a negative key are skipped; n_dev, when set, holds the slot count
// (graph replays with a changing pool).  Each wave
// handles kBinChunks consecutive 64-droplet chunks: it loads all their keys,
// then issues all their atomics, then writes perm, so that the atomics' round
// trips overlap instead of following one another (launched with
// bin_blocks(n) workgroups).
constexpr int kBinChunks = 4;  // 0.073 ms at 20M slots; 1 chunk: 0.097, 8: 0.077, 16: 0.089
__global__ void __launch_bounds__(kBlock) k_bin_scatter(const int32_t* keys, int64_t n, const int32_t* offsets,
                                                        int32_t* counts, int32_t* perm, const int64_t* n_dev) {
  if (n_dev) n = *n_dev;
  const int lane = lane_id();
  const int64_t w0 = (xcd_block() * (kBlock / 64) + (threadIdx.x >> 6)) * (64 * kBinChunks);
  int key[kBinChunks];
#pragma unroll
  for (int r = 0; r < kBinChunks; ++r) {
    const int64_t t = w0 + r * 64 + lane;
    key[r] = t < n ? keys[t] : -1;
  }
  int seg[kBinChunks], base[kBinChunks];
#pragma unroll
  for (int r = 0; r < kBinChunks; ++r) {
    const bool live = key[r] >= 0;
    const int prev = __shfl_up(key[r], 1, 64);
    const bool head = live && (lane == 0 || prev != key[r]);
    const unsigned long long heads = __ballot(head);
    const unsigned long long brk = __ballot(head || !live);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    seg[r] = max(0, 63 - __clzll((long long)(heads & upto)));
    base[r] = 0;
    if (head) {
      const unsigned long long after = (lane == 63) ? 0ull : (brk >> (lane + 1)) << (lane + 1);
      const int end = after ? (__ffsll((long long)after) - 1) : 64;
      const int len = end - lane;
      base[r] = offsets[key[r]] - len + atomicSub(&counts[key[r]], len);
    }
  }
#pragma unroll
  for (int r = 0; r < kBinChunks; ++r) {
    const int b = __shfl(base[r], seg[r], 64);
    if (key[r] >= 0) perm[b + (lane - seg[r])] = (int32_t)(w0 + r * 64 + lane);
  }
}

// Makes the counting sort stable and deterministic: orders every cell's
// segment of the permutation by tie[perm[p]] (or perm[p] when tie is null).
__global__ void __launch_bounds__(kBlock) k_cell_sort(const int32_t* offsets, int64_t ncell, int32_t* perm,
                                                      const int32_t* tie) {
  const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (c >= ncell) return;
  const int b = offsets[c], e = offsets[c + 1];
  if (e - b < 2) return;
  for (int i = b + 1; i < e; ++i) {
    const int32_t v = perm[i];
    const int32_t kv = tie ? tie[v] : v;
    int j = i - 1;
    while (j >= b) {
      const int32_t u = perm[j];
      if ((tie ? tie[u] : u) <= kv) break;
      perm[j + 1] = u;
      --j;
    }
    perm[j + 1] = v;
  }
}

// ------------------------------------------------------------ utilities

// plain device copy, one 16-byte element per thread (bandwidth reference)
__global__ void __launch_bounds__(kBlock) k_copy16(const double2* __restrict__ a, double2* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_fill_f64(double* p, int64_t n, double v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0ull;
}
__global__ void k_iota(int32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = (int32_t)i;
}

// Halo fill of fill_ext on one periodic domain: ext(k, j, i) for
// j in -1..nny+3, i in -1..nnx+3 from core(k, 1..nny, 1..nnx).
__global__ void k_fill_ext(const double* core, double* ext, int nnx, int nny, int nnz) {
  const int64_t nz2 = nnz + 2;
  const int64_t total = nz2 * (nny + 5) * (int64_t)(nnx + 5);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(e % nz2);
    const int64_t r = e / nz2;
    const int j = (int)(r % (nny + 5)) - 1;
    const int i = (int)(r / (nny + 5)) - 1;
    const int jc = ((j - 1) % nny + nny) % nny + 1;
    const int ic = ((i - 1) % nnx + nnx) % nnx + 1;
    ext[e] = core[k + nz2 * ((jc - 1) + (int64_t)nny * (ic - 1))];
  }
}

// xtoz_trans + fill_ext from the flow solver's x-fastest array
// u(1:nnx+2, 1:nny, 0:nnz+1) (les.F:8394-8397, 10060-10131) to ext layout.
__global__ void k_xpencil_to_ext(const double* xp, double* ext, int nnx, int nny, int nnz) {
  const int64_t nz2 = nnz + 2;
  const int64_t total = nz2 * (nny + 5) * (int64_t)(nnx + 5);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int k = (int)(e % nz2);
    const int64_t r = e / nz2;
    const int j = (int)(r % (nny + 5)) - 1;
    const int i = (int)(r / (nny + 5)) - 1;
    const int jc = ((j - 1) % nny + nny) % nny + 1;
    const int ic = ((i - 1) % nnx + nnx) % nnx + 1;
    ext[e] = xp[(ic - 1) + (int64_t)(nnx + 2) * ((jc - 1) + (int64_t)nny * k)];
  }
}

__global__ void k_to_f32(const double* in, float* out, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) out[i] = (float)in[i];
}

// Critical radius at droplet creation (create_particle, particles.f90:2264-2266):
// rc = the table radius maximising the Koehler curve without its exponent
// (crit_radius, 5307-5327) over the 1024 log-spaced radii of setup_radval
// (5294-5305); the first maximum wins, like maxloc.
__constant__ double c_radval[1024];

// crit_radius (particles.f90:5307-5327): the table radius of largest
// equilibrium supersaturation; 0 when ievap /= 1 (create_particle, 2264-2266).
__device__ __forceinline__ double crit_radius_of(const DevConsts& C, const DevFlags& fl, double m_s, double kappa,
                                                 double Tf) {
  if (fl.ievap != 1) return 0.0;
  const double A = 2 * C.Mw * C.Gam / C.Ru / C.rhow / Tf;
  const double B = kappa * m_s;
  const double D = C.rhos * C.pi2 * 2.0 / 3.0;
  // A/r - B/(D r^3) rises to its single maximum at r* = sqrt(3B/(D A)) and
  // falls after it, by about 1% of its value per table step, so the first
  // maximum over the table lies within a few entries of r*'s position in the
  // table (radval(i) = 10^(-8 + 5(i-1)/1024)); scanning that window with the
  // reference's expression and first-maximum rule gives the full scan's
  // answer.  B = 0: the curve only falls, the answer is the first entry.
  int i0 = 0, i1 = 1023;
  if (B == 0.0) {
    i1 = 9;
  } else if (B > 0.0 && A > 0.0) {
    const double fi = (log10(sqrt(3.0 * B / (D * A))) + 8.0) * (1024.0 / 5.0);
    if (isfinite(fi)) {  // else keep the full scan
      const int ic = (int)floor(fmin(fmax(fi, -16.0), 1040.0));
      i0 = min(max(ic - 4, 0), 1014);
      i1 = max(min(ic + 5, 1023), 9);
    }
  }
  double best = 0.0;
  int ib = i0;
  for (int i = i0; i <= i1; ++i) {
    const double rv = c_radval[i];
    const double ss = A / rv - B / (D * (rv * rv * rv));
    if (i == i0 || ss > best) {
      best = ss;
      ib = i;
    }
  }
  return c_radval[ib];
}

__global__ void __launch_bounds__(kBlock) k_crit_radius(StepArgs a) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  double* hot = a.src.hot;
  hot[hix(kHRC, t)] = crit_radius_of(a.C, a.fl, hot[hix(kHMS, t)], hot[hix(kHKAPPA, t)], hot[hix(kHTP, t)]);
}

// Admission of one request into a serving pool: create_particle's state for
// each droplet from (x, y, z, radius, m_s, kappa_s, mult) -- vp = uf(xp),
// Tp = Tf(xp) (the it <= 1 set-up of particles.f90:2907-2910, with the FP64
// gather of uf_interp_lin/uf_interp), rc = crit_radius, residence times and
// activation count zero -- appended at slots base.. of the pool, with its
// u-cell key counted for the next sort.
struct AdmitArgs {
  DevGrid g;
  DevConsts C;
  DevFlags fl;
  DevFields F;
  const double* in;        // [6][m]: x, y, z, radius, m_s, kappa_s (null: from ens)
  const int64_t* in_mult;  // [m]
  DevEnsemble ens;         // the request's ensemble when in is null
  int64_t m, cap, base;
  int32_t tag;             // request slot << kServeIdxBits
  double* hot;
  int64_t* mult;
  int32_t* home;
  int32_t* keys;
  int32_t* counts;
};

template <int ILIN>
__global__ void __launch_bounds__(kBlock) k_serve_admit(AdmitArgs a) {
  __shared__ double s_z[kMaxLevels], s_zz[kMaxLevels];
  stage_grid(a.g, s_z, s_zz);
  __syncthreads();
  const int64_t t = xcd_block() * kBlock + threadIdx.x;
  const bool valid = t < a.m;
  int key = -1;
  if (valid) {
    double x, y, z, r, m_s, kappa;
    int64_t mult;
    if (a.in) {
      x = a.in[t];
      y = a.in[a.m + t];
      z = a.in[2 * a.m + t];
      r = a.in[3 * a.m + t];
      m_s = a.in[4 * a.m + t];
      kappa = a.in[5 * a.m + t];
      mult = a.in_mult[t];
    } else {
      const NewDroplet nd = make_droplet(a.ens, a.g, s_z, t, a.C.rhos, a.C.pi2);
      x = nd.x;
      y = nd.y;
      z = nd.z;
      r = nd.r;
      m_s = nd.m_s;
      kappa = a.ens.kappa_s;
      mult = a.ens.mult;
    }
    const Cell c = cell_of(a.g, s_z, s_zz, x, y, z);
    const Fluid f = (ILIN == 1) ? gather_lin_f64<true>(a.g, a.F, s_z, s_zz, x, y, z, c)
                                : gather6_f64(a.g, a.F, s_z, s_zz, x, y, z, c);
    Drop d;
    d.x = x;
    d.y = y;
    d.z = z;
    d.vx = f.u[0];
    d.vy = f.u[1];
    d.vz = f.u[2];
    d.Tp = f.T;
    d.r = r;
    d.m_s = m_s;
    d.kappa = kappa;
    d.rc = crit_radius_of(a.C, a.fl, m_s, kappa, f.T);
    d.res = 0.0;
    d.actres = 0.0;
    d.numact = 0.0;
    const int64_t s = a.base + t;
    store_drop(a.hot, s, d);
    a.mult[s] = mult;
    a.home[s] = a.tag | (int32_t)t;
    key = ukey_of(a.g, c);
    a.keys[s] = key;
  }
  count_key(a.counts, key, valid);
}

// ntlp_gpu_particles_generate: droplets 0..n-1 of ensemble e created on the
// device (create_particle, particles.f90:2226-2300: vp = uf, Tp = Tf from the
// gather at the new position, rc = crit_radius), in put order.
struct GenArgs {
  DevGrid g;
  DevConsts C;
  DevFlags fl;
  DevFields F;
  const DevEnsemble* ens;  // [nens]
  const int64_t* start;    // [nens + 1]: first droplet of each ensemble, then n
  int nens;
  int64_t n, cap;
  double* hot;
  int64_t* mult;
  int32_t* home;
  double* cold;
  int32_t* pidx;
  int32_t* procidx;
};

template <int ILIN>
__global__ void __launch_bounds__(kBlock) k_generate(GenArgs a) {
  __shared__ double s_z[kMaxLevels], s_zz[kMaxLevels];
  stage_grid(a.g, s_z, s_zz);
  __syncthreads();
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= a.n) return;
  int lo = 0, hi = a.nens;  // the ensemble of droplet t: start[lo] <= t < start[lo + 1]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a.start[mid] <= t) lo = mid;
    else hi = mid;
  }
  const DevEnsemble& e = a.ens[lo];
  const int64_t i = t - a.start[lo];
  const NewDroplet nd = make_droplet(e, a.g, s_z, i, a.C.rhos, a.C.pi2);
  const Cell c = cell_of(a.g, s_z, s_zz, nd.x, nd.y, nd.z);
  const Fluid f = (ILIN == 1) ? gather_lin_f64<true>(a.g, a.F, s_z, s_zz, nd.x, nd.y, nd.z, c)
                              : gather6_f64(a.g, a.F, s_z, s_zz, nd.x, nd.y, nd.z, c);
  Drop d;
  d.x = nd.x;
  d.y = nd.y;
  d.z = nd.z;
  d.vx = f.u[0];
  d.vy = f.u[1];
  d.vz = f.u[2];
  d.Tp = f.T;
  d.r = nd.r;
  d.m_s = nd.m_s;
  d.kappa = e.kappa_s;
  d.rc = crit_radius_of(a.C, a.fl, nd.m_s, e.kappa_s, f.T);
  d.res = d.actres = d.numact = 0.0;
  store_drop(a.hot, t, d);
  a.mult[t] = e.mult;
  a.home[t] = (int32_t)t;
  a.pidx[t] = (int32_t)(e.first + i + 1);
  a.procidx[t] = 0;
  for (int k = 0; k < kNCold; ++k) a.cold[(int64_t)k * a.cap + t] = 0.0;
  a.cold[kCUFX * a.cap + t] = f.u[0];
  a.cold[kCUFY * a.cap + t] = f.u[1];
  a.cold[kCUFZ * a.cap + t] = f.u[2];
  a.cold[kCTF * a.cap + t] = f.T;
  a.cold[kCQINF * a.cap + t] = f.q;
}

// a plain column into hot column c (slot order)
__global__ void k_pack_hot(const double* col, int c, int64_t n, double* hot) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) hot[hix(c, t)] = col[t];
}
// hot column c -> put order (home == nullptr: slot order is put order)
__global__ void k_unpack_hot(const double* hot, int c, const int32_t* home, int64_t n, double* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[home ? home[t] : t] = hot[hix(c, t)];
}
// Diagnostic columns (mask bit c = cold column c) from slot order to put order.
__global__ void k_flush_diag(const double* diag, const int32_t* home, int64_t n, int64_t cap, unsigned mask,
                             double* cold) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t h = home[t];
  for (int c = 0; c < kNCold; ++c)
    if ((mask >> c) & 1u) cold[c * cap + h] = diag[c * cap + t];
}

// slot-order column -> put order via home
__global__ void k_scatter_col_i64(const int64_t* col, const int32_t* home, int64_t n, int64_t* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[home[t]] = col[t];
}

// slot-order values -> put order via home
__global__ void k_scatter_home_i32x4(const int32_t* in, const int32_t* home, int64_t n, int32_t* out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t h = home[t];
  for (int k = 0; k < 4; ++k) out[4 * h + k] = in[4 * t + k];
}
__global__ void k_map_home(int32_t* perm, const int32_t* home, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  perm[t] = home[perm[t]];
}

}  // namespace ntlp
