// ntlp_gpu.hip -- C ABI of the MI355X droplet step (see include/ntlp_gpu.h).
//
// One process-wide context on the current HIP device, mirroring the module
// globals of particles.f90.  All work runs on the context's own stream
// (default priority).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ntlp_gpu.h"
#include "kernels.hpp"

using namespace ntlp;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(call)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (call);                                                                      \
    if (e_ != hipSuccess) return fail(NTLP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHK(expr)        \
  do {                   \
    int rc_ = (expr);    \
    if (rc_ != 0) return rc_; \
  } while (0)

struct Timed {
  std::string name;
  hipEvent_t a, b;
};

constexpr int64_t kTallyRows = 4096;

struct Ctx {
  bool init = false;
  hipStream_t stream = nullptr;
  DevGrid g{};
  DevConsts C{};
  DevFlags fl{};
  double* d_grid = nullptr;  // z | zz | dzw | dzu | FP32 1/dzw, 1/dzu
  DevConsts* d_consts = nullptr;  // device copy of C
  // fields
  double* ext[5] = {};
  float* f32[3] = {};
  double* field_stage = nullptr;
  bool have_fields = false;
  int64_t ext_elems = 0;
  // droplets
  int64_t n = 0, cap = 0;
  double* hot[2] = {};
  int64_t* mult[2] = {};
  int32_t* home[2] = {};
  int cur = 0;
  bool home_identity = true;
  double* cold = nullptr;
  int32_t* pidx = nullptr;
  int32_t* procidx = nullptr;
  double* amounts = nullptr;
  bool amounts_valid = false;  // amounts belong to the last step of the current slots
  bool amounts_f32 = false;    // ... and a mixed-precision step stored them in FP32
  double* diag = nullptr;      // last step's diagnostics in slot order (kNCold columns)
  unsigned diag_mask = 0;      // cold columns waiting in diag
  int32_t* keys[2] = {};
  int32_t* perm = nullptr;
  bool perm_valid = false;  // perm sorts the current slots by u cell
  bool perm_canonical = false;  // ... and within a cell by droplet (home)
  bool slots_sorted = false;    // the slots were last written in cell order
  int since_reorder = 0;        // fused steps since then
  bool keys_valid = false;  // keys[cur] are the cells of the current positions
  double* col_tmp = nullptr;
  // bins
  int32_t* counts = nullptr;
  int32_t* offsets = nullptr;
  int32_t* bsum = nullptr;
  int64_t ncell = 0, nbsum = 0;
  // coupling sources
  double* src = nullptr;
  int64_t nsrc = 0;
  bool src_valid = false;    // src holds the last coupling update
  double* xsrc = nullptr;    // folded x-pencil sources of the last exchange
  bool xsrc_valid = false;
  // counters
  unsigned long long* tally = nullptr;        // kTallyRows per-step rows
  unsigned long long* tally_slots = nullptr;  // kTallySlots x kTallyWidth
  unsigned long long* tally_totals = nullptr; // kTallyWidth running totals since the last reset
  int64_t step_row = 0;
  // flux statistics of BE steps: two buffers of 3*(nnz+1) doubles, used by
  // BE steps in turn (step j adds into flux[j%2]; its tally close clears the
  // other one for step j+1)
  double* flux = nullptr;
  int64_t be_steps = 0;
  // a failed multi-step call left the droplet state inconsistent
  bool state_bad = false;
  // a serving loop is open on these fields
  bool serving = false;
  // the fused step stages each workgroup's trilinear field nodes in LDS
  // (NTLP_OPT_STAGED_GATHER); off, every gather reads through the caches
  bool stage_gather = false;
  // timing
  bool timing = false;
  std::vector<Timed> pending;
  std::map<std::string, std::pair<int64_t, double>> totals;
};

Ctx ctx;

int hot_of(int c) {
  switch (c) {
    case NTLP_X: return kHX;
    case NTLP_Y: return kHY;
    case NTLP_Z: return kHZ;
    case NTLP_VX: return kHVX;
    case NTLP_VY: return kHVY;
    case NTLP_VZ: return kHVZ;
    case NTLP_TP: return kHTP;
    case NTLP_RADIUS: return kHR;
    case NTLP_M_S: return kHMS;
    case NTLP_KAPPA_S: return kHKAPPA;
    case NTLP_RC: return kHRC;
    case NTLP_RES: return kHRES;
    case NTLP_ACTRES: return kHACTRES;
    case NTLP_NUMACT: return kHNUMACT;
    default: return -1;
  }
}

int cold_of(int c) {
  switch (c) {
    case NTLP_UFX: return kCUFX;
    case NTLP_UFY: return kCUFY;
    case NTLP_UFZ: return kCUFZ;
    case NTLP_XRHSX: return kCXRHSX;
    case NTLP_XRHSY: return kCXRHSY;
    case NTLP_XRHSZ: return kCXRHSZ;
    case NTLP_VRHSX: return kCVRHSX;
    case NTLP_VRHSY: return kCVRHSY;
    case NTLP_VRHSZ: return kCVRHSZ;
    case NTLP_TPRHS_S: return kCTPRHS_S;
    case NTLP_TPRHS_L: return kCTPRHS_L;
    case NTLP_TF: return kCTF;
    case NTLP_RADRHS: return kCRADRHS;
    case NTLP_QINF: return kCQINF;
    case NTLP_QSTAR: return kCQSTAR;
    case NTLP_VXOLD: return kCVXOLD;
    case NTLP_VYOLD: return kCVYOLD;
    case NTLP_VZOLD: return kCVZOLD;
    case NTLP_TPOLD: return kCTPOLD;
    case NTLP_RADIUSOLD: return kCRADIUSOLD;
    default: return -1;
  }
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

template <typename T>
int dalloc(T*& p, int64_t count, const char* what) {
  dfree(p);
  if (count <= 0) count = 1;
  hipError_t e = hipMalloc((void**)&p, sizeof(T) * (size_t)count);
  if (e != hipSuccess) {
    p = nullptr;
    return fail(NTLP_E_NOMEM, std::string("hipMalloc failed for ") + what + " (" +
                                  std::to_string(sizeof(T) * (size_t)count) + " bytes): " + hipGetErrorString(e));
  }
  return 0;
}

unsigned grid8(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  b = (b + 7) / 8 * 8;
  return (unsigned)std::max<int64_t>(b, 8);
}
// workgroups of k_bin_scatter over n slots (kBinChunks wave chunks each)
unsigned bin_blocks(int64_t n) {
  int64_t b = (n + kBlock * kBinChunks - 1) / (kBlock * kBinChunks);
  b = (b + 7) / 8 * 8;
  return (unsigned)std::max<int64_t>(b, 8);
}
unsigned grid_flat(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  return (unsigned)std::min<int64_t>(std::max<int64_t>(b, 1), 1 << 20);
}

// Records HIP events around a launch when timing is on.
template <typename F>
int launch(const char* name, F&& f) {
  hipEvent_t a = nullptr, b = nullptr;
  if (ctx.timing) {
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, ctx.stream));
  }
  f();
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(NTLP_E_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(e));
  if (ctx.timing) {
    HIPCHK(hipEventRecord(b, ctx.stream));
    ctx.pending.push_back({name, a, b});
  }
  return 0;
}

int resolve_timers() {
  if (ctx.pending.empty()) return 0;
  HIPCHK(hipStreamSynchronize(ctx.stream));
  for (auto& t : ctx.pending) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
    auto& e = ctx.totals[t.name];
    e.first += 1;
    e.second += ms;
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  ctx.pending.clear();
  return 0;
}

int need_init() {
  if (!ctx.init) return fail(NTLP_E_NOT_INIT, "ntlp_gpu_init has not been called");
  return 0;
}
int need_fields() {
  CHK(need_init());
  if (!ctx.have_fields) return fail(NTLP_E_STATE, "no fields: call ntlp_gpu_set_fields first");
  return 0;
}
// Fields and a consistent droplet population, which may be empty: a rank
// whose particle list is empty runs every update as a loop over no droplets
// (the entry points return at once, with zero sources and counters).
int need_drops() {
  CHK(need_fields());
  if (ctx.state_bad)
    return fail(NTLP_E_STATE, "a failed step call left the droplets inconsistent: put them again");
  return 0;
}

int fill_f64(double* p, int64_t n, double v) {
  return launch("fill", [&] { k_fill_f64<<<grid_flat(n), kBlock, 0, ctx.stream>>>(p, n, v); });
}
int fill_i32(int32_t* p, int64_t n, int32_t v) {
  return launch("fill", [&] { k_fill_i32<<<grid_flat(n), kBlock, 0, ctx.stream>>>(p, n, v); });
}

void release_all() {
  for (auto& p : ctx.ext) dfree(p);
  for (auto& p : ctx.f32) dfree(p);
  dfree(ctx.field_stage);
  for (int b = 0; b < 2; ++b) {
    dfree(ctx.hot[b]);
    dfree(ctx.mult[b]);
    dfree(ctx.home[b]);
  }
  dfree(ctx.cold);
  dfree(ctx.pidx);
  dfree(ctx.procidx);
  dfree(ctx.amounts);
  dfree(ctx.diag);
  dfree(ctx.keys[0]);
  dfree(ctx.keys[1]);
  dfree(ctx.perm);
  dfree(ctx.col_tmp);
  dfree(ctx.counts);
  dfree(ctx.offsets);
  dfree(ctx.bsum);
  dfree(ctx.src);
  dfree(ctx.xsrc);
  dfree(ctx.tally);
  dfree(ctx.tally_slots);
  dfree(ctx.tally_totals);
  dfree(ctx.flux);
  dfree(ctx.d_grid);
  dfree(ctx.d_consts);
  for (auto& t : ctx.pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  ctx.pending.clear();
  ctx.totals.clear();
  if (ctx.stream) (void)hipStreamDestroy(ctx.stream);
  ctx = Ctx();
}

DevFields fields_view() {
  DevFields F;
  F.u = ctx.ext[0];
  F.v = ctx.ext[1];
  F.w = ctx.ext[2];
  F.T = ctx.ext[3];
  F.q = ctx.ext[4];
  F.u32 = ctx.f32[0];
  F.v32 = ctx.f32[1];
  F.w32 = ctx.f32[2];
  F.sj = ctx.g.nnz + 2;
  F.si = (int64_t)(ctx.g.nnz + 2) * (ctx.g.nny + 5);
  return F;
}

StepArgs base_args() {
  StepArgs a;
  memset(&a, 0, sizeof(a));
  a.g = ctx.g;
  a.C = ctx.C;
  a.Cp = ctx.d_consts;
  a.fl = ctx.fl;
  a.F = fields_view();
  a.n = ctx.n;
  a.cap = ctx.cap;
  a.src = {ctx.hot[ctx.cur], ctx.mult[ctx.cur], ctx.home[ctx.cur]};
  a.dst = {ctx.hot[1 - ctx.cur], ctx.mult[1 - ctx.cur], ctx.home[1 - ctx.cur]};
  a.perm = nullptr;
  a.cold = ctx.cold;
  a.amounts = ctx.amounts;
  a.diag = ctx.diag;
  a.keys_in = nullptr;
  a.keys = ctx.keys[ctx.cur];
  a.counts = ctx.counts;
  a.tally = ctx.tally_slots;
  a.box_bytes = step_box_bytes(ctx.g.nnz, ctx.stage_gather && ctx.fl.ilin == 1);
  return a;
}

// Constant subexpressions of the step, formed once from the namelist values
// (in the reference's order where the FP64 path uses them).
void derive_consts(DevConsts& C) {
  C.Sc13 = std::pow(C.Sc, 1. / 3.);
  C.Pra13 = std::pow(C.Pra, 1. / 3.);
  C.LvMw_Ru = C.Lv * C.Mw / C.Ru;
  C.Mw2Gam = 2. * C.Mw * C.Gam;
  C.Ru_rhow = C.Ru * C.rhow;
  C.Lv3 = 3. * C.Lv;
  C.Mw_Ru = C.Mw / C.Ru;
  C.Mw2Gam_Ru_rhow = 2. * C.Mw * C.Gam / (C.Ru * C.rhow);
  C.isol = 1.0 / ((2.0 / 3.0) * C.pi2 * C.rhos);
  C.pi2_23 = C.pi2 * 2.0 / 3.0;
  C.inuf = 1.0 / C.nuf;
  C.idt = 1.0 / C.dt;
  C.Lv3_Cpp = 3. * C.Lv / C.Cpp;
  C.Cpv_Cpa = C.Cpv / C.Cpa;
  C.hA0 = 0.5 * C.dt * C.nuf / (C.Sc * C.rhow);
  C.hB0 = 1.5 * C.dt * C.CpaCpp * C.nuf / (C.Pra * C.rhow);
  C.Ru_Mw = C.Ru / C.Mw;
  C.irhos = 1.0 / C.rhos;
  C.nuf18 = 18.0 * C.nuf;
  C.isurf_rho = 1.0 / C.surf_rho;
  C.sc6 = 0.6 * C.Sc13;
  C.pr6 = 0.6 * C.Pra13;
}

// the device copy of ctx.C, in stream order with the kernels that read it
int upload_consts() {
  HIPCHK(hipMemcpyAsync(ctx.d_consts, &ctx.C, sizeof(DevConsts), hipMemcpyHostToDevice, ctx.stream));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return 0;
}

// exclusive scan counts -> offsets (ncell + 1 entries)
int scan_counts() {
  const int64_t nb = (ctx.ncell + kScanTile - 1) / kScanTile;
  CHK(launch("bin_scan", [&] { k_scan_reduce<<<(unsigned)nb, kScanThreads, 0, ctx.stream>>>(ctx.counts, ctx.ncell, ctx.bsum); }));
  CHK(launch("bin_scan", [&] { k_scan_bsum<<<1, kScanThreads, 0, ctx.stream>>>(ctx.bsum, nb); }));
  CHK(launch("bin_scan", [&] {
    k_scan_down<<<(unsigned)nb, kScanThreads, 0, ctx.stream>>>(ctx.counts, ctx.ncell, ctx.bsum, ctx.offsets);
  }));
  return 0;
}

// keys + counts of the current slot order must be in place; leaves perm
// sorted by u cell (ties by `tie`, or by slot) and counts zeroed.
// canonical: order the droplets within each cell by `tie` (or by slot), so
// that the permutation does not depend on the order of the scatter's
// atomics.  The fused step does not need that (a droplet's arithmetic does
// not depend on its slot); the bin queries and the deterministic coupling do.
int finish_bin(const int32_t* tie, bool canonical = true) {
  CHK(scan_counts());
  CHK(launch("bin_scatter", [&] {
    k_bin_scatter<<<bin_blocks(ctx.n), kBlock, 0, ctx.stream>>>(ctx.keys[ctx.cur], ctx.n, ctx.offsets, ctx.counts, ctx.perm,
                                                              nullptr);
  }));
  if (canonical)
    CHK(launch("bin_cellsort", [&] {
      k_cell_sort<<<grid_flat(ctx.ncell), kBlock, 0, ctx.stream>>>(ctx.offsets, ctx.ncell, ctx.perm, tie);
    }));
  ctx.perm_valid = true;
  ctx.perm_canonical = canonical && tie == ctx.home[ctx.cur];
  return 0;
}

// orders the current permutation within cells by droplet identity (home)
int canonical_perm() {
  if (ctx.perm_canonical) return 0;
  const int32_t* home = ctx.home[ctx.cur];
  CHK(launch("bin_cellsort", [&] {
    k_cell_sort<<<grid_flat(ctx.ncell), kBlock, 0, ctx.stream>>>(ctx.offsets, ctx.ncell, ctx.perm, home);
  }));
  ctx.perm_canonical = true;
  return 0;
}

// bin the current slot order from the positions
int bin_current(int32_t* cellidx_dev, const int32_t* tie) {
  BinArgs b;
  b.g = ctx.g;
  b.hot = ctx.hot[ctx.cur];
  b.n = ctx.n;
  b.cap = ctx.cap;
  b.keys = ctx.keys[ctx.cur];
  b.counts = ctx.counts;
  b.cellidx = cellidx_dev;
  CHK(launch("bin_keys", [&] { k_keys<<<grid8(ctx.n), kBlock, 0, ctx.stream>>>(b); }));
  ctx.keys_valid = true;
  return finish_bin(tie);
}

unsigned src_mask() {
  unsigned m = 0;
  if (ctx.fl.icouple == 1) m |= 7u;
  if (ctx.fl.iTcouple == 1) m |= 8u;
  if (ctx.fl.iHcouple == 1) m |= 48u;
  return m;
}

// wave-segmented sums of equal-key runs, then FP64 global atomics (clears the
// fields first); works in slot order, sorted or not
int scatter_atomic() {
  ctx.src_valid = true;
  ctx.xsrc_valid = false;
  CHK(fill_f64(ctx.src, 6 * ctx.nsrc, 0.0));
  ScatterArgs s;
  s.g = ctx.g;
  s.keys = ctx.keys_valid ? ctx.keys[ctx.cur] : nullptr;
  s.hot = ctx.hot[ctx.cur];
  s.amounts = ctx.amounts;
  s.n = ctx.n;
  s.cap = ctx.cap;
  s.src = ctx.src;
  s.nsrc = ctx.nsrc;
  s.fmask = src_mask();
  return launch("scatter_seg", [&] {
    if (ctx.amounts_f32) k_scatter<float><<<grid8(ctx.n), kBlock, 0, ctx.stream>>>(s);
    else k_scatter<double><<<grid8(ctx.n), kBlock, 0, ctx.stream>>>(s);
  });
}

// atomic-free deterministic scatter over the current cell sort (perm and
// offsets must describe the current slots)
int pull_sorted() {
  ctx.src_valid = true;
  ctx.xsrc_valid = false;
  PullArgs p;
  p.g = ctx.g;
  p.hot = ctx.hot[ctx.cur];
  p.amounts = ctx.amounts;
  p.perm = ctx.perm;
  p.offsets = ctx.offsets;
  p.cap = ctx.cap;
  p.src = ctx.src;
  p.nsrc = ctx.nsrc;
  p.fmask = src_mask();
  return launch("scatter_pull", [&] {
    if (ctx.amounts_f32) k_scatter_pull<float><<<grid8(ctx.nsrc), kBlock, 0, ctx.stream>>>(p);
    else k_scatter_pull<double><<<grid8(ctx.nsrc), kBlock, 0, ctx.stream>>>(p);
  });
}

int scatter_pull() {
  if (!ctx.perm_valid || !ctx.keys_valid) CHK(bin_current(nullptr, nullptr));
  CHK(canonical_perm());  // a fixed summation order
  return pull_sorted();
}

// node boxes of the tiled scatter
constexpr int kTileI = 8, kTileJ = 8, kTileK = 8;

// tiled scatter over the current cell sort (keys, perm and offsets must
// describe the current slots); writes every node of the 6 fields
int tile_sorted() {
  ctx.src_valid = true;
  ctx.xsrc_valid = false;
  TileArgs t;
  t.g = ctx.g;
  t.hot = ctx.hot[ctx.cur];
  t.amounts = ctx.amounts;
  t.keys = ctx.keys[ctx.cur];
  t.perm = ctx.perm;
  t.offsets = ctx.offsets;
  t.cap = ctx.cap;
  t.src = ctx.src;
  t.nsrc = ctx.nsrc;
  t.fmask = src_mask();
  t.ntx = (ctx.g.nnx + 1 + kTileI - 1) / kTileI;
  t.nty = (ctx.g.nny + 1 + kTileJ - 1) / kTileJ;
  t.ntz = (ctx.g.nnz + 2 + kTileK - 1) / kTileK;
  const int64_t nt = (int64_t)t.ntx * t.nty * t.ntz;
  const unsigned grid = (unsigned)((nt + 7) / 8 * 8);
  return launch("scatter", [&] {
    if (ctx.amounts_f32) k_scatter_box<kTileI, kTileJ, kTileK, float><<<grid, kBlock, 0, ctx.stream>>>(t);
    else k_scatter_box<kTileI, kTileJ, kTileK, double><<<grid, kBlock, 0, ctx.stream>>>(t);
  });
}

int scatter_tile() {
  if (!ctx.perm_valid || !ctx.keys_valid) CHK(bin_current(nullptr, nullptr));
  return tile_sorted();
}

// cold columns of the diagnostics the last fused step left in slot order
constexpr unsigned kDiagBE = (1u << kCUFX) | (1u << kCUFY) | (1u << kCUFZ) | (1u << kCTF) | (1u << kCQINF) |
                             (1u << kCQSTAR) | (1u << kCVXOLD) | (1u << kCVYOLD) | (1u << kCVZOLD) |
                             (1u << kCTPOLD) | (1u << kCRADIUSOLD);
constexpr unsigned kDiagRK3 = kDiagBE | (1u << kCXRHSX) | (1u << kCXRHSY) | (1u << kCXRHSZ) | (1u << kCVRHSX) |
                              (1u << kCVRHSY) | (1u << kCVRHSZ) | (1u << kCTPRHS_S) | (1u << kCTPRHS_L) |
                              (1u << kCRADRHS);

// puts the pending diagnostics into the put-order cold columns
int flush_diag() {
  if (!ctx.diag_mask) return 0;
  const unsigned mask = ctx.diag_mask;
  const double* diag = ctx.diag;
  const int32_t* home = ctx.home[ctx.cur];
  double* cold = ctx.cold;
  CHK(launch("flush_diag", [&] {
    k_flush_diag<<<grid_flat(ctx.n), kBlock, 0, ctx.stream>>>(diag, home, ctx.n, ctx.cap, mask, cold);
  }));
  ctx.diag_mask = 0;
  return 0;
}

// ---------------------------------------------------- droplet statistics

// particle_xy_stats runs in passes of nq_max per-level quantities; every
// workgroup (row) writes kStats0 whole-domain values and nq_max*(nnz+2)
// per-level values per pass, and k_xy_reduce sums the rows in row order.
constexpr int kStatsRowsMax = 256;
constexpr int kStatsWidthMax = 6144;  // nq_max * (nnz+2) of one pass (nnz+2 <= kMaxLevels)
constexpr int64_t kStatsScratch = (int64_t)kStatsRowsMax * (kStats0 + kStatsWidthMax);

// Workspace of ntlp_gpu_particle_xy_stats_noalloc: its partial rows, 12.6 MB
// per GPU, held until the process ends.  Host code never passes this name to
// a kernel (on the host it is only a stand-in for the device symbol): it uses
// the device address that ntlp_gpu_init looks up on each device.
static __device__ __align__(256) double scratch[kStatsScratch];

constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` on each device

// Looks up the device address of `scratch` on the current device (once per
// device; called by ntlp_gpu_init with the device current).
int setup_scratch() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(NTLP_E_RANGE, "device ordinal beyond the workspace table");
  if (g_scratch_dev[dev]) return 0;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(NTLP_E_HIP, "hipGetSymbolAddress returned a null workspace address");
  g_scratch_dev[dev] = static_cast<double*>(p);
  return 0;
}

int stats_scratch(double** ws) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices || !g_scratch_dev[dev])
    return fail(NTLP_E_STATE, "no workspace address on this device: call ntlp_gpu_init with it current");
  *ws = g_scratch_dev[dev];
  return 0;
}

struct StatsShape {
  int nz2, rows, nq_max;
};

StatsShape stats_shape() {
  StatsShape sh;
  sh.nz2 = ctx.g.nnz + 2;
  sh.rows = (int)std::min<int64_t>(kStatsRowsMax, std::max<int64_t>(1, (ctx.n + kBlock - 1) / kBlock));
  sh.nq_max = std::max(1, std::min(24, kStatsWidthMax / sh.nz2));
  return sh;
}

// Enqueues the statistics passes: partial rows into part0 [rows][kStats0] and
// part1 [rows][nq_max][nnz+2], results into out0 [kStats0], out1 [24][nnz+2]
// and lw [nny][nnx] (all device memory).  timed: through launch() (HIP events
// when timing is on); otherwise the kernels are launched directly, so that the
// call creates nothing.
int stats_launch(const StatsShape& sh, double* part0, double* part1, double* out0, double* out1, double* lw,
                 bool timed) {
  auto go = [&](const char* name, auto&& f) -> int {
    if (timed) return launch(name, f);
    f();
    HIPCHK(hipGetLastError());
    return 0;
  };
  const int64_t nlw = (int64_t)ctx.g.nnx * ctx.g.nny;
  CHK(go("fill", [&] { k_fill_f64<<<grid_flat(nlw), kBlock, 0, ctx.stream>>>(lw, nlw, 0.0); }));
  for (int q0 = 0; q0 < 24; q0 += sh.nq_max) {
    StatsArgs a;
    a.g = ctx.g;
    a.C = ctx.C;
    a.hot = ctx.hot[ctx.cur];
    a.mult = ctx.mult[ctx.cur];
    a.home = ctx.home[ctx.cur];
    a.cold = ctx.cold;
    a.n = ctx.n;
    a.cap = ctx.cap;
    a.q0 = q0;
    a.nq = std::min(sh.nq_max, 24 - q0);
    a.part0 = part0;
    a.part1 = part1;
    a.lwp = lw;
    const size_t lds = sizeof(double) * a.nq * sh.nz2;
    const int width = a.nq * sh.nz2;
    double* dst = out1 + (int64_t)q0 * sh.nz2;
    CHK(go("xy_stats", [&] { k_xy_stats<<<sh.rows, kBlock, lds, ctx.stream>>>(a); }));
    CHK(go("xy_stats", [&] {
      k_xy_reduce<<<(width + kBlock - 1) / kBlock, kBlock, 0, ctx.stream>>>(part1, sh.rows, width, dst, width, width);
    }));
    if (q0 == 0)
      CHK(go("xy_stats", [&] { k_xy_reduce<<<1, kBlock, 0, ctx.stream>>>(part0, sh.rows, kStats0, out0, 9, 15); }));
  }
  return 0;
}

template <int MODE, int PREC, int ILIN>
void launch_update_t(const StepArgs& a, unsigned grid) {
  k_update<MODE, PREC, ILIN><<<grid, kBlock, 0, ctx.stream>>>(a);
}
template <int MODE, int PREC, int ILIN>
void launch_step_t(const StepArgs& a, unsigned grid, bool lazy) {
  const size_t lds = step_lds_bytes(a.g.nnz, a.box_bytes);
  if constexpr (MODE == 2) {
    if (lazy) {
      k_step<MODE, PREC, ILIN, false, true><<<grid, kBlock, lds, ctx.stream>>>(a);
      return;
    }
  }
  k_step<MODE, PREC, ILIN><<<grid, kBlock, lds, ctx.stream>>>(a);
}

template <int MODE>
int dispatch_update(const StepArgs& a, int prec) {
  const unsigned gr = grid8(a.n);
  const bool lin = ctx.fl.ilin == 1;
  const char* name = MODE == 1 ? "update_rk3" : "update_be";
  return launch(name, [&] {
    if (prec == PREC_FP64) lin ? launch_update_t<MODE, 0, 1>(a, gr) : launch_update_t<MODE, 0, 0>(a, gr);
    else if (prec == PREC_MIXED) lin ? launch_update_t<MODE, 1, 1>(a, gr) : launch_update_t<MODE, 1, 0>(a, gr);
    else lin ? launch_update_t<MODE, 2, 1>(a, gr) : launch_update_t<MODE, 2, 0>(a, gr);
  });
}

// Launch names by variant: reading through the cell sort ("_sorted") or in
// place, and the last step of a call that also writes the diagnostics
// ("_last"), since their memory traffic differs.
template <int MODE>
int dispatch_step(const StepArgs& a, int prec, bool lazy) {
  const unsigned gr = grid8(a.n);
  const bool lin = ctx.fl.ilin == 1;
  static const char* const names[2][4] = {
      {"step_rk3", "step_rk3_last", "step_rk3_sorted", "step_rk3_sorted_last"},
      {"step_be", "step_be_last", "step_be_sorted", "step_be_sorted_last"}};
  const char* name = names[MODE == 1 ? 0 : 1][(a.inplace ? 0 : 2) + (a.write_cold ? 1 : 0)];
  return launch(name, [&] {
    if (prec == PREC_FP64) lin ? launch_step_t<MODE, 0, 1>(a, gr, lazy) : launch_step_t<MODE, 0, 0>(a, gr, lazy);
    else if (prec == PREC_MIXED) lin ? launch_step_t<MODE, 1, 1>(a, gr, lazy) : launch_step_t<MODE, 1, 0>(a, gr, lazy);
    else lin ? launch_step_t<MODE, 2, 1>(a, gr, lazy) : launch_step_t<MODE, 2, 0>(a, gr, lazy);
  });
}

int flux_len() { return 3 * (ctx.g.nnz + 1); }
// flux statistics buffer of BE step j
double* flux_buf(int64_t j) { return ctx.flux + (j % 2) * flux_len(); }

// closes the current step: slots -> a fresh per-step row and the totals; a
// BE step also clears the flux buffer of the next BE step
int close_step(bool be) {
  ctx.step_row++;
  unsigned long long* row = ctx.tally + (ctx.step_row % kTallyRows) * kTallyWidth;
  unsigned long long* slots = ctx.tally_slots;
  unsigned long long* totals = ctx.tally_totals;
  double* next = be ? flux_buf(ctx.be_steps + 1) : nullptr;
  const int nf = flux_len();
  CHK(launch("tally", [&] { k_tally_close<<<1, 128, 0, ctx.stream>>>(slots, row, totals, 1, next, nf); }));
  if (be) ctx.be_steps++;
  return 0;
}

// nsteps fused steps: each reads the droplets through the cell sort, writes
// them in slot order with their new keys, coupling amounts and histogram,
// then sorts for the next step and (scatter) deposits the sources.  The last
// step leaves its diagnostics in slot order (put order is restored when a
// caller needs the cold columns).
// Steps between two reorderings of the slots into cell order.
constexpr int kReorderEvery = 16;
// Calls of fewer BE steps update the residence times in the step kernel: the
// settle pass reads and writes the 32 bytes per droplet that a lazy step
// saves, so it pays from about two steps on.
constexpr int kLazyMinSteps = 3;

// residence-time additions of the first `steps` steps of a lazy BE call
int settle(int steps) {
  double* hot = ctx.hot[ctx.cur];
  const int64_t n = ctx.n;
  const double dt = ctx.C.dt;
  return launch("res_settle", [&] { k_res_settle<<<grid_flat(n), kBlock, 0, ctx.stream>>>(hot, n, steps, dt); });
}

// nsteps steps of an empty population: each closes its counter row (zero,
// one step) and, for BE, its zero flux statistics; the coupling sources are
// zero.
int empty_steps(int mode, int nsteps, bool scatter) {
  for (int k = 0; k < nsteps; ++k) CHK(close_step(mode == NTLP_MODE_BE));
  if (scatter) {
    CHK(fill_f64(ctx.src, 6 * ctx.nsrc, 0.0));
    ctx.src_valid = true;
    ctx.xsrc_valid = false;
  }
  ctx.amounts_valid = true;
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return 0;
}

int run_fused_steps(int mode, int prec, int nsteps, int it0, bool scatter, bool lazy, int& done) {
  const unsigned mask = mode == NTLP_MODE_RK3 ? kDiagRK3 : kDiagBE;
  if (ctx.diag_mask & ~mask) CHK(flush_diag());
  if (!ctx.perm_valid || !ctx.keys_valid) CHK(bin_current(nullptr, nullptr));
  ctx.amounts_valid = false;
  ctx.amounts_f32 = prec == PREC_MIXED;  // the step kernels' amounts
  for (int k = 0; k < nsteps; ++k) {
    const bool reorder = !ctx.slots_sorted || ctx.since_reorder + 1 >= kReorderEvery;
    StepArgs a = base_args();
    a.it = it0 + k;
    a.write_cold = (k == nsteps - 1) ? 1 : 0;
    a.lazy_j = k;
    if (mode == NTLP_MODE_BE) a.flux = flux_buf(ctx.be_steps);
    if (reorder) {  // read through the sort, write slot t of the other buffer
      a.perm = ctx.perm;
      a.keys_in = ctx.keys[ctx.cur];
      a.keys = ctx.keys[1 - ctx.cur];
    } else {
      a.inplace = 1;
      a.dst = a.src;
      a.keys_in = ctx.keys[ctx.cur];
      a.keys = ctx.keys[ctx.cur];
    }
    if (mode == NTLP_MODE_RK3) CHK(dispatch_step<1>(a, prec, false));
    else CHK(dispatch_step<2>(a, prec, lazy));
    if (reorder) {
      ctx.cur = 1 - ctx.cur;
      ctx.home_identity = false;
      ctx.slots_sorted = true;
      ctx.since_reorder = 0;
    } else {
      ctx.since_reorder++;
    }
    ctx.keys_valid = true;
    done = k + 1;
    CHK(close_step(mode == NTLP_MODE_BE));
    CHK(finish_bin(nullptr, false));
    if (scatter && src_mask()) CHK(tile_sorted());  // no coupling flag: the sources stay zero
  }
  if (lazy) CHK(settle(nsteps));  // the call's residence-time additions
  ctx.diag_mask = mask;
  ctx.amounts_valid = true;
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return 0;
}

// nsteps fused steps.  If a launch fails part way through a lazy BE call, the
// residence times of the steps already made are settled (or, if that fails
// too, the droplets are marked inconsistent until the next put).
int run_fused(int mode, int prec, int nsteps, int it0, bool scatter) {
  const bool lazy = mode == NTLP_MODE_BE && nsteps >= kLazyMinSteps;
  int done = 0;
  const int rc = run_fused_steps(mode, prec, nsteps, it0, scatter, lazy, done);
  if (rc != 0 && lazy && done > 0) {
    const std::string msg = g_err;
    if (done == nsteps || settle(done) != 0 || hipStreamSynchronize(ctx.stream) != hipSuccess) ctx.state_bad = true;
    g_err = msg;
  }
  if (rc != 0) {
    ctx.perm_valid = ctx.keys_valid = ctx.amounts_valid = false;
    ctx.diag_mask = 0;
  }
  return rc;
}

}  // namespace

// ===================================================================== ABI

extern "C" {

const char* ntlp_gpu_last_error(void) { return g_err.c_str(); }

int ntlp_gpu_init(const ntlp_grid* grid, const ntlp_consts* c, const ntlp_flags* f) {
  if (ctx.serving) return fail(NTLP_E_STATE, "close the serving loop first");
  if (!grid || !c || !f) return fail(NTLP_E_INVALID, "null argument to ntlp_gpu_init");
  if (grid->nnx < 2 || grid->nny < 2 || grid->nnz < 3) return fail(NTLP_E_INVALID, "grid too small (nnx,nny >= 2, nnz >= 3)");
  if (grid->nnz + 2 > kMaxLevels) return fail(NTLP_E_RANGE, "nnz + 2 exceeds the LDS grid limit of 1024 levels");
  if (!(grid->dx > 0) || !(grid->dy > 0)) return fail(NTLP_E_INVALID, "dx and dy must be positive");
  if (!grid->z || !grid->zz || !grid->dzw || !grid->dzu) return fail(NTLP_E_INVALID, "null grid metadata");
  if (f->icase != 0) return fail(NTLP_E_UNSUPPORTED, "only icase=0 (reflecting walls) is supported");
  const int64_t cells_u = (int64_t)(grid->nnx + 1) * (grid->nny + 1) * (grid->nnz + 1);
  if (cells_u >= (int64_t)INT32_MAX) return fail(NTLP_E_RANGE, "grid has too many cells for 32-bit cell keys");
  // the trilinear gathers address the ext arrays with 32-bit byte offsets
  if ((int64_t)(grid->nnz + 2) * (grid->nny + 5) * (grid->nnx + 5) * 8 > (int64_t)UINT32_MAX)
    return fail(NTLP_E_RANGE, "ext field arrays beyond 4 GB each (nnx*nny*nnz above about 5.3e8)");
  // monotone z and zz are required by the minloc semantics
  for (int k = 1; k < grid->nnz + 2; ++k)
    if (!(grid->z[k] > grid->z[k - 1]) || !(grid->zz[k] > grid->zz[k - 1]))
      return fail(NTLP_E_INVALID, "z and zz must be strictly increasing");
  // every argument is checked before the device is touched: a rejected call
  // leaves an existing context as it was
  release_all();
  CHK(setup_scratch());
  HIPCHK(hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking));
  const int nz = grid->nnz;
  const int nlev = nz + 2, nd = nz + 3;
  CHK(dalloc(ctx.d_grid, 2 * nlev + 5 * nd, "grid"));
  HIPCHK(hipMemcpy(ctx.d_grid, grid->z, sizeof(double) * nlev, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx.d_grid + nlev, grid->zz, sizeof(double) * nlev, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx.d_grid + 2 * nlev, grid->dzw, sizeof(double) * nd, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx.d_grid + 2 * nlev + nd, grid->dzu, sizeof(double) * nd, hipMemcpyHostToDevice));
  {
    std::vector<float> rcp(2 * nd);
    for (int k = 0; k < nd; ++k) {
      rcp[k] = (float)(1.0 / grid->dzw[k]);
      rcp[nd + k] = (float)(1.0 / grid->dzu[k]);
    }
    HIPCHK(hipMemcpy(ctx.d_grid + 2 * nlev + 2 * nd, rcp.data(), sizeof(float) * 2 * nd, hipMemcpyHostToDevice));
    std::vector<double> r64(2 * nd);
    for (int k = 0; k < nd; ++k) {
      r64[k] = 1.0 / grid->dzu[k];
      r64[nd + k] = 1.0 / (grid->dx * grid->dy * grid->dzu[k]);
    }
    HIPCHK(hipMemcpy(ctx.d_grid + 2 * nlev + 3 * nd, r64.data(), sizeof(double) * 2 * nd, hipMemcpyHostToDevice));
  }
  DevGrid& g = ctx.g;
  g.nnx = grid->nnx;
  g.nny = grid->nny;
  g.nnz = nz;
  g.dx = grid->dx;
  g.dy = grid->dy;
  g.xl = grid->xl;
  g.yl = grid->yl;
  g.z = ctx.d_grid;
  g.zz = ctx.d_grid + nlev;
  g.dzw = ctx.d_grid + 2 * nlev;
  g.dzu = ctx.d_grid + 2 * nlev + nd;
  g.rdzw32 = reinterpret_cast<const float*>(ctx.d_grid + 2 * nlev + 2 * nd);
  g.rdzu32 = g.rdzw32 + nd;
  g.rdzu64 = ctx.d_grid + 2 * nlev + 3 * nd;
  g.rdVu64 = ctx.d_grid + 2 * nlev + 4 * nd;
  g.rdx64 = 1.0 / grid->dx;
  g.rdy64 = 1.0 / grid->dy;
  g.rdx32 = (float)(1.0 / grid->dx);
  g.rdy32 = (float)(1.0 / grid->dy);
  int steps = 0;
  while ((1 << steps) <= nlev) ++steps;
  g.nsteps = steps;
  bool inter = true;
  for (int k = 0; k < nlev; ++k) {
    if (!(grid->zz[k] < grid->z[k])) inter = false;
    if (k >= 1 && !(grid->z[k - 1] < grid->zz[k])) inter = false;
  }
  g.interleaved = inter ? 1 : 0;
  DevConsts& C = ctx.C;
  C.dt = c->dt;
  C.nuf = c->nuf;
  C.Cpa = c->Cpa;
  C.Cpp = c->Cpp;
  C.CpaCpp = c->CpaCpp;
  C.Pra = c->Pra;
  C.Sc = c->Sc;
  C.Rd = c->Rd;
  C.rhow = c->rhow;
  C.rhos = c->rhos;
  C.Mw = c->Mw;
  C.Ru = c->Ru;
  C.Gam = c->Gam;
  C.Lv = c->Lv;
  C.Cpv = c->Cpv;
  C.surf_rho = c->surf_rho;
  C.surf_p = c->surf_p;
  C.tsfcc1 = c->tsfcc1;
  C.grav = c->grav;
  C.radius_init = c->radius_init;
  C.pi2 = c->pi2;
  for (int d = 0; d < 3; ++d) {
    C.g[d] = c->part_grav[d];
    C.zetas[d] = c->zetas[d];
    C.gama[d] = c->gama[d];
  }
  derive_consts(C);
  CHK(dalloc(ctx.d_consts, 1, "constants"));
  DevFlags& fl = ctx.fl;
  fl.ilin = f->ilin;
  fl.icouple = f->icouple;
  fl.iTcouple = f->iTcouple;
  fl.iHcouple = f->iHcouple;
  fl.ievap = f->ievap;
  fl.iexner = f->iexner;
  fl.icase = f->icase;
  ctx.ncell = cells_u;
  ctx.nbsum = (cells_u + kScanTile - 1) / kScanTile;
  CHK(dalloc(ctx.counts, ctx.ncell, "counts"));
  CHK(dalloc(ctx.offsets, ctx.ncell + 1, "offsets"));
  CHK(dalloc(ctx.bsum, ctx.nbsum, "scan block sums"));
  ctx.nsrc = (int64_t)(nz + 2) * (grid->nny + 1) * (grid->nnx + 1);
  CHK(dalloc(ctx.src, 6 * ctx.nsrc, "coupling sources"));
  CHK(dalloc(ctx.tally, kTallyRows * kTallyWidth, "counters"));
  CHK(dalloc(ctx.tally_slots, kTallySlots * kTallyWidth, "counter slots"));
  CHK(dalloc(ctx.tally_totals, kTallyWidth, "counter totals"));
  CHK(dalloc(ctx.flux, 2 * 3 * (nz + 1), "flux statistics"));
  ctx.ext_elems = (int64_t)(nz + 2) * (grid->nny + 5) * (grid->nnx + 5);
  ctx.init = true;
  CHK(upload_consts());
  CHK(fill_i32(ctx.counts, ctx.ncell, 0));
  CHK(fill_f64(ctx.src, 6 * ctx.nsrc, 0.0));
  CHK(launch("fill", [&] { k_fill_u64<<<64, kBlock, 0, ctx.stream>>>(ctx.tally, kTallyRows * kTallyWidth); }));
  CHK(launch("fill", [&] { k_fill_u64<<<8, kBlock, 0, ctx.stream>>>(ctx.tally_slots, kTallySlots * kTallyWidth); }));
  CHK(launch("fill", [&] { k_fill_u64<<<1, kBlock, 0, ctx.stream>>>(ctx.tally_totals, kTallyWidth); }));
  CHK(fill_f64(ctx.flux, 2 * 3 * (nz + 1), 0.0));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

int ntlp_gpu_finalize(void) {
  if (ctx.serving) return fail(NTLP_E_STATE, "close the serving loop first");
  if (ctx.stream) (void)hipStreamSynchronize(ctx.stream);
  release_all();
  return NTLP_OK;
}

int ntlp_gpu_set_dt(double dt) {
  CHK(need_init());
  if (ctx.serving) return fail(NTLP_E_STATE, "the serving loop steps with the current dt: close it first");
  if (!(dt > 0)) return fail(NTLP_E_INVALID, "dt must be positive");
  ctx.C.dt = dt;
  derive_consts(ctx.C);
  return upload_consts();
}

int ntlp_gpu_set_fields(const double* u, const double* v, const double* w, const double* T, const double* q,
                        int32_t layout) {
  CHK(need_init());
  if (ctx.serving) return fail(NTLP_E_STATE, "the serving loop reads the fields: close it first");
  const double* in[5] = {u, v, w, T, q};
  for (auto p : in)
    if (!p) return fail(NTLP_E_INVALID, "null field pointer");
  const int nx = ctx.g.nnx, ny = ctx.g.nny, nz = ctx.g.nnz;
  int64_t in_elems;
  if (layout == NTLP_LAYOUT_EXT) in_elems = ctx.ext_elems;
  else if (layout == NTLP_LAYOUT_CORE) in_elems = (int64_t)(nz + 2) * ny * nx;
  else if (layout == NTLP_LAYOUT_XPENCIL) in_elems = (int64_t)(nx + 2) * ny * (nz + 2);
  else return fail(NTLP_E_INVALID, "unknown field layout");
  for (int f = 0; f < 5; ++f)
    if (!ctx.ext[f]) CHK(dalloc(ctx.ext[f], ctx.ext_elems, "ext field"));
  for (int f = 0; f < 3; ++f)
    if (!ctx.f32[f]) CHK(dalloc(ctx.f32[f], ctx.ext_elems, "fp32 field copy"));
  if (layout != NTLP_LAYOUT_EXT && !ctx.field_stage) CHK(dalloc(ctx.field_stage, in_elems, "field staging"));
  for (int f = 0; f < 5; ++f) {
    if (layout == NTLP_LAYOUT_EXT) {
      HIPCHK(hipMemcpyAsync(ctx.ext[f], in[f], sizeof(double) * in_elems, hipMemcpyHostToDevice, ctx.stream));
    } else {
      HIPCHK(hipMemcpyAsync(ctx.field_stage, in[f], sizeof(double) * in_elems, hipMemcpyHostToDevice, ctx.stream));
      double* dst = ctx.ext[f];
      const double* stg = ctx.field_stage;
      if (layout == NTLP_LAYOUT_CORE)
        CHK(launch("fill_ext", [&] { k_fill_ext<<<grid_flat(ctx.ext_elems), kBlock, 0, ctx.stream>>>(stg, dst, nx, ny, nz); }));
      else
        CHK(launch("fill_ext", [&] {
          k_xpencil_to_ext<<<grid_flat(ctx.ext_elems), kBlock, 0, ctx.stream>>>(stg, dst, nx, ny, nz);
        }));
    }
    if (f < 3) {
      double* s = ctx.ext[f];
      float* o = ctx.f32[f];
      CHK(launch("to_f32", [&] { k_to_f32<<<grid_flat(ctx.ext_elems), kBlock, 0, ctx.stream>>>(s, o, ctx.ext_elems); }));
    }
    HIPCHK(hipStreamSynchronize(ctx.stream));  // staging buffer reuse
  }
  dfree(ctx.field_stage);
  ctx.have_fields = true;
  return NTLP_OK;
}

namespace {
// device buffers for n droplets (kept when they are large enough); resets the
// slot bookkeeping to put order
int ensure_capacity(int64_t n) {
  HIPCHK(hipStreamSynchronize(ctx.stream));
  if (n > ctx.cap) {
    for (int b = 0; b < 2; ++b) {
      CHK(dalloc(ctx.hot[b], tiled_elems(n, kNHot), "hot state"));
      CHK(dalloc(ctx.mult[b], n, "mult"));
      CHK(dalloc(ctx.home[b], n, "home"));
    }
    CHK(dalloc(ctx.cold, kNCold * n, "cold state"));
    CHK(dalloc(ctx.pidx, n, "pidx"));
    CHK(dalloc(ctx.procidx, n, "procidx"));
    CHK(dalloc(ctx.amounts, tiled_elems(n, 6), "coupling amounts"));
    CHK(dalloc(ctx.diag, kNCold * n, "step diagnostics"));
    CHK(dalloc(ctx.keys[0], n, "keys"));
    CHK(dalloc(ctx.keys[1], n, "keys"));
    CHK(dalloc(ctx.perm, n, "perm"));
    CHK(dalloc(ctx.col_tmp, n, "column staging"));
    ctx.cap = n;
  }
  ctx.n = n;
  ctx.cur = 0;
  ctx.state_bad = false;
  ctx.home_identity = true;
  ctx.diag_mask = 0;
  ctx.amounts_valid = false;
  ctx.slots_sorted = false;
  ctx.perm_valid = false;
  ctx.keys_valid = false;
  return 0;
}

// the critical-radius table of setup_radval (particles.f90:5294-5305)
int upload_radval(hipStream_t stream) {
  static double table[1024];
  const double radstart = -8.0, radend = -3.0, dr = (radstart - radend) / 1024.0;
  for (int i = 1; i <= 1024; ++i) table[i - 1] = std::pow(10.0, radstart - (double)(i - 1) * dr);
  HIPCHK(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_radval), table, sizeof(table), 0, hipMemcpyHostToDevice, stream));
  return 0;
}

int check_ensemble(const ntlp_ensemble* e) {
  if (!e) return fail(NTLP_E_INVALID, "null ensemble");
  auto radius_ok = [](int kind, const double* p) {
    if (kind == NTLP_DIST_LOGNORMAL) return p[0] > 0 && p[1] >= 0;
    if (kind == NTLP_DIST_LOGUNIFORM) return p[0] > 0 && p[1] >= p[0];
    return false;
  };
  if (!radius_ok(e->radius_kind, e->radius_p)) return fail(NTLP_E_INVALID, "bad radius distribution");
  if (e->frac2 < 0 || e->frac2 > 1) return fail(NTLP_E_INVALID, "frac2 must lie in [0, 1]");
  if (e->frac2 > 0 && !radius_ok(e->radius2_kind, e->radius2_p))
    return fail(NTLP_E_INVALID, "bad second radius distribution");
  if (e->salt_kind != NTLP_SALT_FIXED && e->salt_kind != NTLP_SALT_DRY_LOGNORMAL && e->salt_kind != NTLP_SALT_CHOICE3)
    return fail(NTLP_E_INVALID, "unknown salt distribution");
  if (e->mult < 1 || e->first < 0) return fail(NTLP_E_INVALID, "mult must be >= 1 and first >= 0");
  return 0;
}

DevEnsemble dev_ensemble(const ntlp_ensemble* e) {
  DevEnsemble d;
  memset(&d, 0, sizeof(d));
  d.seed = e->seed;
  d.radius_kind = e->radius_kind;
  d.radius2_kind = e->radius2_kind;
  for (int k = 0; k < 2; ++k) {
    d.radius_p[k] = e->radius_p[k];
    d.radius2_p[k] = e->radius2_p[k];
  }
  d.frac2 = e->frac2;
  d.salt_kind = e->salt_kind;
  for (int k = 0; k < 3; ++k) d.salt_p[k] = e->salt_p[k];
  d.kappa_s = e->kappa_s;
  d.mult = e->mult;
  d.first = e->first;
  return d;
}
}  // namespace

int ntlp_gpu_particles_put(const double* soa, const int64_t* mult, const int32_t* pidx, const int32_t* procidx,
                           int64_t n) {
  CHK(need_init());
  if (n < 0 || (n > 0 && (!soa || !mult))) return fail(NTLP_E_INVALID, "bad droplet arrays");
  if (n >= (int64_t)INT32_MAX) return fail(NTLP_E_RANGE, "more than 2^31-1 droplets");
  CHK(ensure_capacity(n));
  if (n == 0) return NTLP_OK;
  for (int c = 0; c < NTLP_NREAL; ++c) {
    if (hot_of(c) >= 0) {
      double* tmp = ctx.col_tmp;
      double* hot = ctx.hot[0];
      const int hc = hot_of(c);
      HIPCHK(hipMemcpyAsync(tmp, soa + (int64_t)c * n, sizeof(double) * n, hipMemcpyHostToDevice, ctx.stream));
      CHK(launch("put", [&] { k_pack_hot<<<grid_flat(n), kBlock, 0, ctx.stream>>>(tmp, hc, n, hot); }));
      HIPCHK(hipStreamSynchronize(ctx.stream));  // col_tmp reuse
    } else {
      HIPCHK(hipMemcpyAsync(ctx.cold + (int64_t)cold_of(c) * ctx.cap, soa + (int64_t)c * n, sizeof(double) * n,
                            hipMemcpyHostToDevice, ctx.stream));
    }
  }
  HIPCHK(hipMemcpyAsync(ctx.mult[0], mult, sizeof(int64_t) * n, hipMemcpyHostToDevice, ctx.stream));
  if (pidx) HIPCHK(hipMemcpyAsync(ctx.pidx, pidx, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx.stream));
  else CHK(fill_i32(ctx.pidx, n, 0));
  if (procidx) HIPCHK(hipMemcpyAsync(ctx.procidx, procidx, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx.stream));
  else CHK(fill_i32(ctx.procidx, n, 0));
  CHK(launch("iota", [&] { k_iota<<<grid_flat(n), kBlock, 0, ctx.stream>>>(ctx.home[0], n); }));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

int ntlp_gpu_particles_get(double* soa, int64_t* mult, int32_t* pidx, int32_t* procidx, int64_t n) {
  CHK(need_init());
  if (n != ctx.n) return fail(NTLP_E_INVALID, "n does not match the droplet count on the device");
  if (ctx.state_bad)
    return fail(NTLP_E_STATE, "a failed step call left the droplets inconsistent: put them again");
  if (soa) CHK(flush_diag());
  if (n == 0) return NTLP_OK;
  const int64_t cap = ctx.cap;
  const int32_t* home = ctx.home[ctx.cur];
  if (soa) {
    for (int c = 0; c < NTLP_NREAL; ++c) {
      if (hot_of(c) >= 0) {
        const double* hot = ctx.hot[ctx.cur];
        const int hc = hot_of(c);
        double* tmp = ctx.col_tmp;
        const int32_t* hm = ctx.home_identity ? nullptr : home;
        CHK(launch("get", [&] { k_unpack_hot<<<grid_flat(n), kBlock, 0, ctx.stream>>>(hot, hc, hm, n, tmp); }));
        HIPCHK(hipMemcpyAsync(soa + (int64_t)c * n, tmp, sizeof(double) * n, hipMemcpyDeviceToHost, ctx.stream));
        HIPCHK(hipStreamSynchronize(ctx.stream));  // col_tmp reuse
      } else {
        HIPCHK(hipMemcpyAsync(soa + (int64_t)c * n, ctx.cold + (int64_t)cold_of(c) * cap, sizeof(double) * n,
                              hipMemcpyDeviceToHost, ctx.stream));
      }
    }
  }
  if (mult) {
    const int64_t* col = ctx.mult[ctx.cur];
    if (!ctx.home_identity) {
      int64_t* tmp = (int64_t*)ctx.col_tmp;
      CHK(launch("get", [&] { k_scatter_col_i64<<<grid_flat(n), kBlock, 0, ctx.stream>>>(col, home, n, tmp); }));
      col = tmp;
    }
    HIPCHK(hipMemcpyAsync(mult, col, sizeof(int64_t) * n, hipMemcpyDeviceToHost, ctx.stream));
    HIPCHK(hipStreamSynchronize(ctx.stream));
  }
  if (pidx) HIPCHK(hipMemcpyAsync(pidx, ctx.pidx, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx.stream));
  if (procidx) HIPCHK(hipMemcpyAsync(procidx, ctx.procidx, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx.stream));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

int64_t ntlp_gpu_particles_count(void) { return ctx.init ? ctx.n : 0; }

// 352-byte record of `type particle` in the ifort layout (particles.f90:76-85,
// 1929-1950): 4 int32, 39 real(8), mult and two pointer slots.
namespace {
constexpr int kRecReals = 39;
// record real index -> NTLP column (-1: dist, u_sub(3), sigm_s, not kept)
const int kRecCol[kRecReals] = {
    NTLP_VX, NTLP_VY, NTLP_VZ, NTLP_X, NTLP_Y, NTLP_Z, NTLP_UFX, NTLP_UFY, NTLP_UFZ,
    NTLP_XRHSX, NTLP_XRHSY, NTLP_XRHSZ, NTLP_VRHSX, NTLP_VRHSY, NTLP_VRHSZ,
    NTLP_TP, NTLP_TPRHS_S, NTLP_TPRHS_L, NTLP_TF, NTLP_RADIUS, NTLP_RADRHS, NTLP_QINF, NTLP_QSTAR,
    -1 /*dist*/, NTLP_RES, NTLP_M_S, NTLP_KAPPA_S, NTLP_RC, NTLP_ACTRES, NTLP_NUMACT,
    -1, -1, -1 /*u_sub*/, -1 /*sigm_s*/, NTLP_VXOLD, NTLP_VYOLD, NTLP_VZOLD, NTLP_TPOLD, NTLP_RADIUSOLD};
}  // namespace

int ntlp_gpu_particles_put_records(const void* records, int64_t n) {
  if (n < 0 || (n > 0 && !records)) return fail(NTLP_E_INVALID, "bad record buffer");
  std::vector<double> soa((size_t)n * NTLP_NREAL, 0.0);
  std::vector<int64_t> mult((size_t)n);
  std::vector<int32_t> pidx((size_t)n), procidx((size_t)n);
  const unsigned char* base = (const unsigned char*)records;
  for (int64_t i = 0; i < n; ++i) {
    const unsigned char* r = base + i * NTLP_RECORD_BYTES;
    int32_t ints[4];
    double reals[kRecReals];
    memcpy(ints, r, 16);
    memcpy(reals, r + 16, 8 * kRecReals);
    memcpy(&mult[i], r + 16 + 8 * kRecReals, 8);
    pidx[i] = ints[0];
    procidx[i] = ints[1];
    for (int k = 0; k < kRecReals; ++k)
      if (kRecCol[k] >= 0) soa[(size_t)kRecCol[k] * n + i] = reals[k];
  }
  return ntlp_gpu_particles_put(soa.data(), mult.data(), pidx.data(), procidx.data(), n);
}

int ntlp_gpu_particles_get_records(void* records, int64_t n) {
  CHK(need_init());
  if (n != ctx.n || (n > 0 && !records)) return fail(NTLP_E_INVALID, "bad record buffer");
  std::vector<double> soa((size_t)n * NTLP_NREAL);
  std::vector<int64_t> mult((size_t)n);
  std::vector<int32_t> pidx((size_t)n), procidx((size_t)n);
  CHK(ntlp_gpu_particles_get(soa.data(), mult.data(), pidx.data(), procidx.data(), n));
  unsigned char* base = (unsigned char*)records;
  for (int64_t i = 0; i < n; ++i) {
    unsigned char* r = base + i * NTLP_RECORD_BYTES;
    const int32_t ints[4] = {pidx[i], procidx[i], -1, -1};
    double reals[kRecReals];
    for (int k = 0; k < kRecReals; ++k) reals[k] = kRecCol[k] >= 0 ? soa[(size_t)kRecCol[k] * n + i] : 0.0;
    memcpy(r, ints, 16);
    memcpy(r + 16, reals, 8 * kRecReals);
    memcpy(r + 16 + 8 * kRecReals, &mult[i], 8);
    memset(r + 16 + 8 * kRecReals + 8, 0, 16);
  }
  return NTLP_OK;
}

// save_particles (particles.f90:1988-2046) on one rank: the int32 droplet
// count at offset 0, then one 352-byte record per droplet.
int ntlp_gpu_save_particles(const char* path) {
  CHK(need_init());
  if (!path) return fail(NTLP_E_INVALID, "null path");
  if (ctx.n > (int64_t)INT32_MAX) return fail(NTLP_E_RANGE, "the file's count is an int32");
  std::vector<unsigned char> rec((size_t)ctx.n * NTLP_RECORD_BYTES);
  CHK(ntlp_gpu_particles_get_records(rec.data(), ctx.n));
  FILE* f = fopen(path, "wb");
  if (!f) return fail(NTLP_E_INVALID, std::string("cannot open ") + path + " for writing");
  const int32_t count = (int32_t)ctx.n;
  bool ok = fwrite(&count, 4, 1, f) == 1;
  if (ok && !rec.empty()) ok = fwrite(rec.data(), 1, rec.size(), f) == rec.size();
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(NTLP_E_INVALID, std::string("write to ") + path + " failed");
  return NTLP_OK;
}

// read_part_res (particles.f90:2048-2146) on one rank: reads the count and
// the records in chunks of 5 million and keeps those inside the rank's open
// box xmin < x < xmax, ymin < y < ymax (here 0 < x < xl, 0 < y < yl), as the
// reference does; the records advance by 352 bytes, the writer's stride.
int ntlp_gpu_read_part_res(const char* path, int64_t* kept) {
  CHK(need_init());
  if (!path) return fail(NTLP_E_INVALID, "null path");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(NTLP_E_INVALID, std::string("cannot open ") + path);
  int32_t count = 0;
  if (fread(&count, 4, 1, f) != 1 || count < 0) {
    fclose(f);
    return fail(NTLP_E_INVALID, std::string(path) + ": no droplet count");
  }
  std::vector<unsigned char> keep;
  const int64_t chunk = 5000000;
  std::vector<unsigned char> buf;
  int64_t left = count, nkeep = 0;
  constexpr int kX = 16 + 3 * 8, kY = 16 + 4 * 8;  // xp(1), xp(2) follow vp(1:3)
  while (left > 0) {
    const int64_t m = std::min(left, chunk);
    buf.resize((size_t)m * NTLP_RECORD_BYTES);
    if (fread(buf.data(), 1, buf.size(), f) != buf.size()) {
      fclose(f);
      return fail(NTLP_E_INVALID, std::string(path) + ": fewer records than its count");
    }
    for (int64_t i = 0; i < m; ++i) {
      const unsigned char* r = buf.data() + i * NTLP_RECORD_BYTES;
      double x, y;
      memcpy(&x, r + kX, 8);
      memcpy(&y, r + kY, 8);
      if (y > 0.0 && y < ctx.g.yl && x > 0.0 && x < ctx.g.xl) {
        keep.insert(keep.end(), r, r + NTLP_RECORD_BYTES);
        ++nkeep;
      }
    }
    left -= m;
  }
  fclose(f);
  CHK(ntlp_gpu_particles_put_records(keep.data(), nkeep));
  if (kept) *kept = nkeep;
  return NTLP_OK;
}

int ntlp_gpu_particles_generate(const ntlp_ensemble* ens, const int64_t* counts, int32_t nens) {
  CHK(need_fields());
  if (!ens || !counts || nens < 1 || nens > NTLP_MAX_ENSEMBLES)
    return fail(NTLP_E_INVALID, "1 to " + std::to_string(NTLP_MAX_ENSEMBLES) + " ensembles with their counts");
  std::vector<DevEnsemble> de((size_t)nens);
  std::vector<int64_t> start((size_t)nens + 1, 0);
  for (int k = 0; k < nens; ++k) {
    CHK(check_ensemble(&ens[k]));
    if (counts[k] < 0) return fail(NTLP_E_INVALID, "negative droplet count");
    de[k] = dev_ensemble(&ens[k]);
    start[k + 1] = start[k] + counts[k];
  }
  const int64_t n = start[nens];
  if (n < 1 || n >= (int64_t)INT32_MAX) return fail(NTLP_E_RANGE, "the ensembles must hold 1 .. 2^31-2 droplets");
  CHK(ensure_capacity(n));
  CHK(upload_radval(ctx.stream));
  DevEnsemble* d_ens = nullptr;
  int64_t* d_start = nullptr;
  CHK(dalloc(d_ens, nens, "ensembles"));
  int rc = dalloc(d_start, nens + 1, "ensemble starts");
  if (!rc && (hipMemcpyAsync(d_ens, de.data(), sizeof(DevEnsemble) * nens, hipMemcpyHostToDevice, ctx.stream) !=
                  hipSuccess ||
              hipMemcpyAsync(d_start, start.data(), sizeof(int64_t) * (nens + 1), hipMemcpyHostToDevice, ctx.stream) !=
                  hipSuccess))
    rc = fail(NTLP_E_HIP, "ensemble upload failed");
  GenArgs a;
  a.g = ctx.g;
  a.C = ctx.C;
  a.fl = ctx.fl;
  a.F = fields_view();
  a.ens = d_ens;
  a.start = d_start;
  a.nens = nens;
  a.n = n;
  a.cap = ctx.cap;
  a.hot = ctx.hot[0];
  a.mult = ctx.mult[0];
  a.home = ctx.home[0];
  a.cold = ctx.cold;
  a.pidx = ctx.pidx;
  a.procidx = ctx.procidx;
  const unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
  const bool lin = ctx.fl.ilin == 1;
  if (!rc)
    rc = launch("generate", [&] {
      if (lin) k_generate<1><<<grid, kBlock, 0, ctx.stream>>>(a);
      else k_generate<0><<<grid, kBlock, 0, ctx.stream>>>(a);
    });
  if (hipStreamSynchronize(ctx.stream) != hipSuccess && !rc) rc = fail(NTLP_E_HIP, "ensemble generation failed");
  dfree(d_ens);
  dfree(d_start);
  return rc;
}

int ntlp_gpu_particles_get_column(int32_t column, double* out, int64_t n) {
  CHK(need_init());
  if (n != ctx.n || !out) return fail(NTLP_E_INVALID, "n does not match the droplet count, or null output");
  if (column < 0 || column >= NTLP_NREAL) return fail(NTLP_E_INVALID, "unknown column");
  if (ctx.state_bad)
    return fail(NTLP_E_STATE, "a failed step call left the droplets inconsistent: put them again");
  if (n == 0) return NTLP_OK;
  if (hot_of(column) >= 0) {
    const double* hot = ctx.hot[ctx.cur];
    const int hc = hot_of(column);
    double* tmp = ctx.col_tmp;
    const int32_t* hm = ctx.home_identity ? nullptr : ctx.home[ctx.cur];
    CHK(launch("get", [&] { k_unpack_hot<<<grid_flat(n), kBlock, 0, ctx.stream>>>(hot, hc, hm, n, tmp); }));
    HIPCHK(hipMemcpyAsync(out, tmp, sizeof(double) * n, hipMemcpyDeviceToHost, ctx.stream));
  } else {
    CHK(flush_diag());
    HIPCHK(hipMemcpyAsync(out, ctx.cold + (int64_t)cold_of(column) * ctx.cap, sizeof(double) * n,
                          hipMemcpyDeviceToHost, ctx.stream));
  }
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

int ntlp_gpu_crit_radius(void) {
  CHK(need_init());
  if (ctx.n == 0) return NTLP_OK;
  CHK(upload_radval(ctx.stream));
  StepArgs a = base_args();
  CHK(launch("crit_radius", [&] { k_crit_radius<<<grid_flat(ctx.n), kBlock, 0, ctx.stream>>>(a); }));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

int ntlp_gpu_interp(int32_t set_vp, int32_t set_tp) {
  CHK(need_drops());
  if (ctx.n == 0) return NTLP_OK;
  CHK(flush_diag());
  if (set_vp || set_tp) ctx.amounts_valid = false;
  StepArgs a = base_args();
  const bool lin = ctx.fl.ilin == 1;
  CHK(launch("interp", [&] {
    if (lin) k_interp<1><<<grid8(ctx.n), kBlock, 0, ctx.stream>>>(a, set_vp, set_tp);
    else k_interp<0><<<grid8(ctx.n), kBlock, 0, ctx.stream>>>(a, set_vp, set_tp);
  }));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

static int check_prec(int prec) {
  if (prec != PREC_FP64 && prec != PREC_MIXED && prec != PREC_FP64_AJ) return fail(NTLP_E_INVALID, "unknown precision");
  return 0;
}

int ntlp_gpu_particle_update_rk3_prec(int32_t istage, int32_t it, int32_t prec) {
  CHK(need_drops());
  CHK(check_prec(prec));
  if (istage < 1 || istage > 3) return fail(NTLP_E_INVALID, "istage must be 1, 2 or 3");
  if (ctx.n == 0) return NTLP_OK;
  CHK(flush_diag());
  ctx.keys_valid = false;
  ctx.amounts_valid = false;
  StepArgs a = base_args();
  a.istage = istage;
  a.it = it;
  CHK(dispatch_update<1>(a, prec));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

int ntlp_gpu_particle_update_rk3(int32_t istage, int32_t it) {
  return ntlp_gpu_particle_update_rk3_prec(istage, it, PREC_FP64);
}

// One BE step with the fused machinery (cell sorted, coupling amounts kept for
// particle_coupling_update); the droplets keep their put order to callers.
int ntlp_gpu_particle_update_be_prec(int32_t it, int32_t prec) {
  CHK(need_drops());
  CHK(check_prec(prec));
  if (ctx.n == 0) return empty_steps(NTLP_MODE_BE, 1, false);
  return run_fused(NTLP_MODE_BE, prec, 1, it, false);
}

int ntlp_gpu_particle_update_be(int32_t it) { return ntlp_gpu_particle_update_be_prec(it, PREC_FP64); }

int ntlp_gpu_particle_coupling_update(double* src_out, int32_t deterministic) {
  CHK(need_drops());
  if (deterministic < 0 || deterministic > 2)
    return fail(NTLP_E_INVALID, "coupling method " + std::to_string(deterministic) +
                                    ": 0 (box), 1 (fixed order) or 2 (segmented atomics)");
  if (ctx.n == 0) {  // no droplets: zero sources
    CHK(fill_f64(ctx.src, 6 * ctx.nsrc, 0.0));
    ctx.src_valid = true;
    ctx.xsrc_valid = false;
    if (src_out)
      HIPCHK(hipMemcpyAsync(src_out, ctx.src, sizeof(double) * 6 * ctx.nsrc, hipMemcpyDeviceToHost, ctx.stream));
    HIPCHK(hipStreamSynchronize(ctx.stream));
    return NTLP_OK;
  }
  if (!ctx.amounts_valid) {  // amounts from the state and its old values in cold
    CHK(flush_diag());
    StepArgs a = base_args();
    CHK(launch("amounts", [&] { k_amounts<<<grid8(ctx.n), kBlock, 0, ctx.stream>>>(a); }));
    ctx.amounts_valid = true;
    ctx.amounts_f32 = false;
  }
  if (deterministic == 1) CHK(scatter_pull());
  else if (deterministic == 2) CHK(scatter_atomic());
  else CHK(scatter_tile());
  if (src_out)
    HIPCHK(hipMemcpyAsync(src_out, ctx.src, sizeof(double) * 6 * ctx.nsrc, hipMemcpyDeviceToHost, ctx.stream));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

int ntlp_gpu_particle_coupling_exchange(double* partsrc, double* partTsrc, double* partHsrc, double* partTEsrc) {
  CHK(need_init());
  if (!ctx.src_valid) return fail(NTLP_E_STATE, "no coupling sources: run a coupling update or a fused step first");
  const int nx = ctx.g.nnx, ny = ctx.g.nny, nz = ctx.g.nnz;
  const int64_t nf = (int64_t)(nz + 2) * ny * nx;
  if (!ctx.xsrc) CHK(dalloc(ctx.xsrc, 6 * nf, "exchanged sources"));
  const int64_t blocks = 6LL * ny * ((nz + 2 + kFoldK - 1) / kFoldK) * ((nx + kFoldX - 1) / kFoldX);
  double* xs = ctx.xsrc;
  const double* src = ctx.src;
  const int64_t nsrc = ctx.nsrc;
  CHK(launch("exchange", [&] { k_fold_sources<<<(unsigned)blocks, kBlock, 0, ctx.stream>>>(src, nsrc, nx, ny, nz, xs); }));
  double* outs[4] = {partsrc, partTsrc, partHsrc, partTEsrc};
  const int first[4] = {0, 3, 4, 5}, count[4] = {3, 1, 1, 1};
  for (int i = 0; i < 4; ++i)
    if (outs[i])
      HIPCHK(hipMemcpyAsync(outs[i], ctx.xsrc + first[i] * nf, sizeof(double) * count[i] * nf, hipMemcpyDeviceToHost,
                            ctx.stream));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  ctx.xsrc_valid = true;
  return NTLP_OK;
}

int ntlp_gpu_apply_particle_coupling(double* T, double* q, int32_t layout) {
  CHK(need_fields());
  if (ctx.serving) return fail(NTLP_E_STATE, "the serving loop reads the fields: close it first");
  if (!ctx.xsrc_valid) return fail(NTLP_E_STATE, "no exchanged sources: call ntlp_gpu_particle_coupling_exchange first");
  if (!T || !q) return fail(NTLP_E_INVALID, "null field pointer");
  if (layout != NTLP_LAYOUT_XPENCIL && layout != NTLP_LAYOUT_CORE)
    return fail(NTLP_E_INVALID, "apply_particle_coupling takes x-pencil or core arrays");
  const int nx = ctx.g.nnx, ny = ctx.g.nny, nz = ctx.g.nnz;
  const int64_t elems = layout == NTLP_LAYOUT_XPENCIL ? (int64_t)(nx + 2) * ny * (nz + 2) : (int64_t)(nz + 2) * ny * nx;
  double* stage[2] = {nullptr, nullptr};
  int rc = dalloc(stage[0], elems, "field staging");
  if (!rc) rc = dalloc(stage[1], elems, "field staging");
  double* host[2] = {T, q};
  for (int i = 0; i < 2 && !rc; ++i)
    if (hipMemcpyAsync(stage[i], host[i], sizeof(double) * elems, hipMemcpyHostToDevice, ctx.stream) != hipSuccess)
      rc = fail(NTLP_E_HIP, "upload of T/q failed");
  if (!rc) {
    const int xp = layout == NTLP_LAYOUT_XPENCIL;
    rc = launch("apply_coupling", [&] {
      k_apply_coupling<<<grid_flat((int64_t)nx * ny * nz), kBlock, 0, ctx.stream>>>(stage[0], stage[1], ctx.xsrc,
                                                                                    ctx.g, ctx.C, ctx.fl, xp);
    });
  }
  for (int i = 0; i < 2 && !rc; ++i) {
    if (hipMemcpyAsync(host[i], stage[i], sizeof(double) * elems, hipMemcpyDeviceToHost, ctx.stream) != hipSuccess)
      rc = fail(NTLP_E_HIP, "download of T/q failed");
    // the resident T (ext 3) and q (ext 4) the next step gathers
    double* dst = ctx.ext[3 + i];
    const double* stg = stage[i];
    if (!rc)
      rc = launch("fill_ext", [&] {
        if (layout == NTLP_LAYOUT_CORE) k_fill_ext<<<grid_flat(ctx.ext_elems), kBlock, 0, ctx.stream>>>(stg, dst, nx, ny, nz);
        else k_xpencil_to_ext<<<grid_flat(ctx.ext_elems), kBlock, 0, ctx.stream>>>(stg, dst, nx, ny, nz);
      });
  }
  if (hipStreamSynchronize(ctx.stream) != hipSuccess && !rc) rc = fail(NTLP_E_HIP, "apply_particle_coupling failed");
  dfree(stage[0]);
  dfree(stage[1]);
  return rc;
}

int ntlp_gpu_get_sources(double* src_out) {
  CHK(need_init());
  if (!src_out) return fail(NTLP_E_INVALID, "null output");
  HIPCHK(hipMemcpyAsync(src_out, ctx.src, sizeof(double) * 6 * ctx.nsrc, hipMemcpyDeviceToHost, ctx.stream));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

int ntlp_gpu_bin(int32_t* cellidx, int32_t* counts, int32_t* offsets, int32_t* perm) {
  CHK(need_init());
  if (ctx.n == 0) {  // every cell empty
    if (counts) std::fill(counts, counts + ctx.ncell, 0);
    if (offsets) std::fill(offsets, offsets + ctx.ncell + 1, 0);
    return NTLP_OK;
  }
  const int64_t n = ctx.n;
  int32_t* ci_dev = nullptr;
  int32_t* ci_put = nullptr;
  if (cellidx) {
    CHK(dalloc(ci_dev, 4 * n, "cell indices"));
    CHK(dalloc(ci_put, 4 * n, "cell indices"));
  }
  int rc = bin_current(ci_dev, ctx.home[ctx.cur]);
  if (rc == 0 && cellidx) {
    const int32_t* home = ctx.home[ctx.cur];
    rc = launch("bin_out", [&] { k_scatter_home_i32x4<<<grid_flat(n), kBlock, 0, ctx.stream>>>(ci_dev, home, n, ci_put); });
    if (rc == 0 && hipMemcpyAsync(cellidx, ci_put, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost, ctx.stream) != hipSuccess)
      rc = fail(NTLP_E_HIP, "copy of cell indices failed");
  }
  if (rc == 0 && offsets &&
      hipMemcpyAsync(offsets, ctx.offsets, sizeof(int32_t) * (ctx.ncell + 1), hipMemcpyDeviceToHost, ctx.stream) != hipSuccess)
    rc = fail(NTLP_E_HIP, "copy of offsets failed");
  if (rc == 0 && counts) {
    std::vector<int32_t> off((size_t)ctx.ncell + 1);
    if (hipMemcpyAsync(off.data(), ctx.offsets, sizeof(int32_t) * (ctx.ncell + 1), hipMemcpyDeviceToHost, ctx.stream) != hipSuccess ||
        hipStreamSynchronize(ctx.stream) != hipSuccess)
      rc = fail(NTLP_E_HIP, "copy of offsets failed");
    else
      for (int64_t c = 0; c < ctx.ncell; ++c) counts[c] = off[c + 1] - off[c];
  }
  if (rc == 0 && perm) {
    // slot indices -> put-order indices, on a copy (the slot permutation
    // stays valid for the next fused step)
    int32_t* pp = nullptr;
    rc = dalloc(pp, n, "perm copy");
    if (rc == 0 && hipMemcpyAsync(pp, ctx.perm, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, ctx.stream) != hipSuccess)
      rc = fail(NTLP_E_HIP, "perm copy failed");
    const int32_t* home = ctx.home[ctx.cur];
    if (rc == 0) rc = launch("bin_out", [&] { k_map_home<<<grid_flat(n), kBlock, 0, ctx.stream>>>(pp, home, n); });
    if (rc == 0 && hipMemcpyAsync(perm, pp, sizeof(int32_t) * n, hipMemcpyDeviceToHost, ctx.stream) != hipSuccess)
      rc = fail(NTLP_E_HIP, "copy of perm failed");
    if (hipStreamSynchronize(ctx.stream) != hipSuccess && rc == 0) rc = fail(NTLP_E_HIP, "sync failed");
    dfree(pp);
  }
  if (hipStreamSynchronize(ctx.stream) != hipSuccess && rc == 0) rc = fail(NTLP_E_HIP, "sync failed");
  dfree(ci_dev);
  dfree(ci_put);
  return rc;
}

int ntlp_gpu_wcell_counts(int32_t* num_array) {
  CHK(need_init());
  if (!num_array) return fail(NTLP_E_INVALID, "null output");
  const int64_t nw = (int64_t)ctx.g.nnx * ctx.g.nny * ctx.g.nnz;
  int32_t* d = nullptr;
  CHK(dalloc(d, nw, "w-cell counts"));
  int rc = fill_i32(d, nw, 0);
  if (rc == 0 && ctx.n > 0) {
    BinArgs b;
    b.g = ctx.g;
    b.hot = ctx.hot[ctx.cur];
    b.n = ctx.n;
    b.cap = ctx.cap;
    b.keys = nullptr;
    b.counts = nullptr;
    b.cellidx = nullptr;
    rc = launch("wcells", [&] { k_wcells<<<grid8(ctx.n), kBlock, 0, ctx.stream>>>(b, d); });
  }
  if (rc == 0 && hipMemcpyAsync(num_array, d, sizeof(int32_t) * nw, hipMemcpyDeviceToHost, ctx.stream) != hipSuccess)
    rc = fail(NTLP_E_HIP, "copy failed");
  if (hipStreamSynchronize(ctx.stream) != hipSuccess && rc == 0) rc = fail(NTLP_E_HIP, "sync failed");
  dfree(d);
  return rc;
}

int ntlp_gpu_calc_tauc_min(double* tauc_min) {
  CHK(need_drops());
  if (!tauc_min) return fail(NTLP_E_INVALID, "null output");
  if (ctx.n == 0) {  // the start value 1e10 (particles.f90:3327), lowered by no droplet, times the
    *tauc_min = 1.0e10 * 5.0;  // fudge factor 5 (3363)
    return NTLP_OK;
  }
  const int64_t ncell = (int64_t)ctx.g.nnx * ctx.g.nny * ctx.g.nnz;
  double *rad = nullptr, *nc = nullptr;
  int32_t* num = nullptr;
  unsigned long long* res = nullptr;
  int rc = dalloc(rad, ncell, "tauc sums");
  if (!rc) rc = dalloc(nc, ncell, "tauc sums");
  if (!rc) rc = dalloc(num, ncell, "tauc counts");
  if (!rc) rc = dalloc(res, 1, "tauc result");
  const double start = 1.0e10;
  if (!rc && (hipMemsetAsync(rad, 0, sizeof(double) * ncell, ctx.stream) != hipSuccess ||
              hipMemsetAsync(nc, 0, sizeof(double) * ncell, ctx.stream) != hipSuccess ||
              hipMemsetAsync(num, 0, sizeof(int32_t) * ncell, ctx.stream) != hipSuccess ||
              hipMemcpyAsync(res, &start, sizeof(double), hipMemcpyHostToDevice, ctx.stream) != hipSuccess))
    rc = fail(NTLP_E_HIP, "tauc_min set-up failed");
  if (!rc) {
    BinArgs b;
    b.g = ctx.g;
    b.hot = ctx.hot[ctx.cur];
    b.n = ctx.n;
    b.cap = ctx.cap;
    const int64_t* mult = ctx.mult[ctx.cur];
    rc = launch("tauc", [&] { k_tauc_accum<<<grid8(ctx.n), kBlock, 0, ctx.stream>>>(b, mult, rad, num, nc); });
  }
  if (!rc)
    rc = launch("tauc", [&] { k_tauc_min<<<grid_flat(ncell), kBlock, 0, ctx.stream>>>(ctx.g, ctx.C, rad, num, nc, res); });
  double m = 0.0;
  if (!rc && (hipMemcpyAsync(&m, res, sizeof(double), hipMemcpyDeviceToHost, ctx.stream) != hipSuccess ||
              hipStreamSynchronize(ctx.stream) != hipSuccess))
    rc = fail(NTLP_E_HIP, "tauc_min failed");
  dfree(rad);
  dfree(nc);
  dfree(num);
  dfree(res);
  if (rc) return rc;
  *tauc_min = m * 5.0;
  return NTLP_OK;
}

int ntlp_gpu_particle_xy_stats(double* s0, double* s1, double* lwp) {
  CHK(need_drops());
  CHK(flush_diag());
  const StatsShape sh = stats_shape();
  double *part0 = nullptr, *part1 = nullptr, *out0 = nullptr, *out1 = nullptr, *lw = nullptr;
  int rc = dalloc(part0, (int64_t)sh.rows * kStats0, "statistics rows");
  if (!rc) rc = dalloc(part1, (int64_t)sh.rows * sh.nq_max * sh.nz2, "statistics rows");
  if (!rc) rc = dalloc(out0, kStats0, "statistics");
  if (!rc) rc = dalloc(out1, 24 * sh.nz2, "statistics");
  if (!rc) rc = dalloc(lw, (int64_t)ctx.g.nnx * ctx.g.nny, "LWP");
  if (!rc) rc = stats_launch(sh, part0, part1, out0, out1, lw, true);
  if (!rc && ((s0 && hipMemcpyAsync(s0, out0, sizeof(double) * kStats0, hipMemcpyDeviceToHost, ctx.stream) != hipSuccess) ||
              (s1 && hipMemcpyAsync(s1, out1, sizeof(double) * 24 * sh.nz2, hipMemcpyDeviceToHost, ctx.stream) != hipSuccess) ||
              (lwp && hipMemcpyAsync(lwp, lw, sizeof(double) * ctx.g.nnx * ctx.g.nny, hipMemcpyDeviceToHost,
                                     ctx.stream) != hipSuccess) ||
              hipStreamSynchronize(ctx.stream) != hipSuccess))
    rc = fail(NTLP_E_HIP, "statistics copy failed");
  dfree(part0);
  dfree(part1);
  dfree(out0);
  dfree(out1);
  dfree(lw);
  return rc;
}

int ntlp_gpu_particle_xy_stats_noalloc(double* s0, double* s1, double* lwp) {
  CHK(need_drops());
  if (!s0 || !s1 || !lwp) return fail(NTLP_E_INVALID, "null output (all three device outputs are required)");
  double* ws = nullptr;
  CHK(stats_scratch(&ws));
  const StatsShape sh = stats_shape();
  const int64_t need = (int64_t)sh.rows * (kStats0 + (int64_t)sh.nq_max * sh.nz2);
  if (need > kStatsScratch)
    return fail(NTLP_E_RANGE, "statistics rows need " + std::to_string(need * 8) + " bytes of workspace; the static "
                "workspace holds " + std::to_string(kStatsScratch * 8));
  if (ctx.diag_mask) {  // pending diagnostics (uf, Tf, qinf, qstar) into the put-order columns
    k_flush_diag<<<grid_flat(ctx.n), kBlock, 0, ctx.stream>>>(ctx.diag, ctx.home[ctx.cur], ctx.n, ctx.cap,
                                                              ctx.diag_mask, ctx.cold);
    HIPCHK(hipGetLastError());
    ctx.diag_mask = 0;
  }
  return stats_launch(sh, ws, ws + (int64_t)sh.rows * kStats0, s0, s1, lwp, false);
}

int64_t ntlp_gpu_num_ucells(void) { return ctx.init ? ctx.ncell : 0; }

int ntlp_gpu_step(int32_t mode, int32_t prec, int32_t nsteps, int32_t it0) {
  CHK(need_drops());
  CHK(check_prec(prec));
  if (mode != NTLP_MODE_RK3 && mode != NTLP_MODE_BE) return fail(NTLP_E_INVALID, "unknown mode");
  if (nsteps < 1) return fail(NTLP_E_INVALID, "nsteps must be >= 1");
  if (ctx.n == 0) return empty_steps(mode, nsteps, src_mask() != 0);
  return run_fused(mode, prec, nsteps, it0, true);
}

int ntlp_gpu_counters(ntlp_counters* out, int32_t reset) {
  CHK(need_init());
  unsigned long long row[kTallyWidth];
  HIPCHK(hipMemcpyAsync(row, ctx.tally_totals, sizeof(row), hipMemcpyDeviceToHost, ctx.stream));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  if (out) {
    memset(out, 0, sizeof(*out));
    out->num100 = row[kTNum100];
    out->numimpos = row[kTNumImpos];
    out->actnum = row[kTActnum];
    out->denum = row[kTDenum];
    out->gn_iters = row[kTGnIters];
    out->lm_iters = row[kTLmIters];
    out->solves = row[kTSolves];
    out->steps = row[kTSteps];
    for (int h = 0; h < NTLP_GN_HIST; ++h) out->gn_hist[h] = row[kTHist0 + h];
  }
  if (reset) {
    CHK(launch("fill", [&] { k_fill_u64<<<64, kBlock, 0, ctx.stream>>>(ctx.tally, kTallyRows * kTallyWidth); }));
    CHK(launch("fill", [&] { k_fill_u64<<<8, kBlock, 0, ctx.stream>>>(ctx.tally_slots, kTallySlots * kTallyWidth); }));
    CHK(launch("fill", [&] { k_fill_u64<<<1, kBlock, 0, ctx.stream>>>(ctx.tally_totals, kTallyWidth); }));
    HIPCHK(hipStreamSynchronize(ctx.stream));
    ctx.step_row = 0;
  }
  return NTLP_OK;
}

int ntlp_gpu_be_flux(double* pflux, double* pmassflux, double* penegflux) {
  CHK(need_init());
  if (ctx.be_steps == 0) return fail(NTLP_E_STATE, "no backward-Euler step has run");
  const int nlev = ctx.g.nnz + 1;
  const double* f = flux_buf(ctx.be_steps - 1);
  double* outs[3] = {pflux, pmassflux, penegflux};
  for (int k = 0; k < 3; ++k)
    if (outs[k])
      HIPCHK(hipMemcpyAsync(outs[k], f + (int64_t)k * nlev, sizeof(double) * nlev, hipMemcpyDeviceToHost, ctx.stream));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

int ntlp_gpu_step_counters(int64_t* rows_out, int32_t max_rows) {
  CHK(need_init());
  if (!rows_out || max_rows <= 0) return fail(NTLP_E_INVALID, "bad output");
  const int64_t nr = std::min<int64_t>(std::min<int64_t>(ctx.step_row, kTallyRows - 1), max_rows);
  std::vector<unsigned long long> rows((size_t)kTallyRows * kTallyWidth);
  HIPCHK(hipMemcpyAsync(rows.data(), ctx.tally, sizeof(unsigned long long) * rows.size(), hipMemcpyDeviceToHost, ctx.stream));
  HIPCHK(hipStreamSynchronize(ctx.stream));
  // rows 1..step_row hold the steps in order (row 0 is the pre-step row)
  for (int64_t r = 0; r < nr; ++r) {
    const int64_t src = (ctx.step_row - nr + 1 + r) % kTallyRows;
    for (int k = 0; k < kTallyWidth; ++k) rows_out[r * kTallyWidth + k] = (int64_t)rows[src * kTallyWidth + k];
  }
  return (int)nr;
}

int ntlp_gpu_timing(int32_t enable) {
  CHK(need_init());
  CHK(resolve_timers());
  ctx.totals.clear();
  ctx.timing = enable != 0;
  return NTLP_OK;
}

int ntlp_gpu_kernel_times(char* names, int64_t* launches, double* total_ms, int32_t cap) {
  CHK(need_init());
  CHK(resolve_timers());
  int k = 0;
  for (auto& e : ctx.totals) {
    if (k >= cap) break;
    if (names) {
      memset(names + 32 * k, 0, 32);
      strncpy(names + 32 * k, e.first.c_str(), 31);
    }
    if (launches) launches[k] = e.second.first;
    if (total_ms) total_ms[k] = e.second.second;
    ++k;
  }
  return k;
}

int ntlp_gpu_copy_bandwidth(int64_t bytes, int32_t reps, double* gbps) {
  CHK(need_init());
  if (bytes < (1 << 20) || bytes > ((int64_t)1 << 36) || reps < 1 || !gbps)
    return fail(NTLP_E_INVALID, "copy_bandwidth needs 1 MiB to 64 GiB and reps >= 1");
  const int64_t n = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  int rc = dalloc(a, n, "copy source");
  if (!rc) rc = dalloc(b, n, "copy destination");
  std::vector<float> ms;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(NTLP_E_HIP, "hipEventCreate failed");
  const unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
  for (int r = -1; r < reps && !rc; ++r) {  // r = -1: warm-up
    if (hipEventRecord(e0, ctx.stream) != hipSuccess) rc = fail(NTLP_E_HIP, "hipEventRecord failed");
    k_copy16<<<grid, kBlock, 0, ctx.stream>>>(a, b, n);
    if (!rc && hipGetLastError() != hipSuccess) rc = fail(NTLP_E_HIP, "copy kernel launch failed");
    if (!rc && hipEventRecord(e1, ctx.stream) != hipSuccess) rc = fail(NTLP_E_HIP, "hipEventRecord failed");
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = fail(NTLP_E_HIP, "copy kernel failed");
    float t = 0.f;
    if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = fail(NTLP_E_HIP, "hipEventElapsedTime failed");
    if (!rc && r >= 0) ms.push_back(t);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dfree(a);
  dfree(b);
  if (rc) return rc;
  std::sort(ms.begin(), ms.end());
  *gbps = 2.0 * 16.0 * (double)n / (ms[ms.size() / 2] * 1e-3) / 1e9;
  return NTLP_OK;
}

int ntlp_gpu_set_option(int32_t option, int64_t value) {
  CHK(need_init());
  if (ctx.serving) return fail(NTLP_E_STATE, "close the serving loop first");
  switch (option) {
    case NTLP_OPT_STAGED_GATHER:
      if (value != 0 && value != 1) return fail(NTLP_E_INVALID, "NTLP_OPT_STAGED_GATHER takes 0 or 1");
      ctx.stage_gather = value == 1;
      return NTLP_OK;
    default:
      return fail(NTLP_E_INVALID, "unknown option " + std::to_string(option));
  }
}

int ntlp_gpu_sync(void) {
  CHK(need_init());
  HIPCHK(hipStreamSynchronize(ctx.stream));
  return NTLP_OK;
}

}  // extern "C"

#include "serve.hpp"
