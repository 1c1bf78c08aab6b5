// serve.hpp -- continuous-batching serving loop of the droplet step.
//
// Requests are droplet ensembles (x, y, z, radius, m_s, kappa_s, mult) that
// each advance a fixed number of backward-Euler steps against the fields
// resident in the main context.  One engine thread owns a droplet pool on the
// device and its own stream:
//
//   every iteration
//     accept  queued requests, FIFO, while a request slot is free, the pool
//             has room (capacity = HBM budget / bytes per droplet) and the
//             input ring has space; start their upload on the upload stream;
//     admit   requests whose upload has landed: append their droplets to the
//             pool (k_serve_admit);
//     step    write the control block (pool size, which requests take their
//             last step, their rows in the output ring), then replay one
//             captured hipGraph per (pool size class, buffer parity):
//             counting sort of the pool by u cell and the fused step of every
//             droplet; droplets of finishing requests go to the output ring
//             and drop out of the next sort, so the pool compacts itself;
//     retire  download finished requests on the download stream and signal
//             their tickets when the copies land.
//
// Uploads, steps and downloads run on three streams (all of default
// priority), so a large request's copies overlap the steps of the others.
// Requests that arrive while others are in the pool join at the next
// iteration, and their steps run in the same kernels as everyone else's.  A
// droplet's arithmetic does not depend on the other droplets in the pool
// (the fields are fixed and the coupling scatter is not formed), so a
// request's result is the same alone and in any mix.
//
// This file is included by ntlp_gpu.hip after the context helpers.
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;

// Device bytes per pool droplet: hot state and mult/home/keys twice
// (ping-pong), perm, admission staging (6 FP64 + mult) and output rows.
constexpr int64_t kServeBytesPerDroplet =
    2 * (kNHot * 8 + 8 + 4 + 4) + 4 + (6 * 8 + 8) + (kNHot * 8 + 8);
static_assert(kServeBytesPerDroplet == NTLP_SERVE_BYTES_PER_DROPLET, "pool bytes per droplet");
constexpr int kServeCtlRing = 4;

enum { kReqQueued = 0, kReqLoading, kReqActive, kReqLeaving, kReqDone, kReqFailed };

// FIFO allocator of droplet rows in a device ring: blocks are released in the
// order they were taken (copies on one stream complete in order).  A block
// that does not fit before the end of the ring starts at row 0; the rows it
// skips are held with it.
struct Ring {
  int64_t size = 0, head = 0, tail = 0, used = 0;
  std::deque<std::pair<int64_t, int64_t>> blocks;  // (first held row, rows held)
  int64_t alloc(int64_t n) {
    if (n < 1 || n > size) return -1;
    if (used == 0) head = tail = 0;
    if (head >= tail && used < size) {
      if (size - head >= n) {
        blocks.push_back({head, n});
        used += n;
        head += n;
        return head - n;
      }
      if (tail >= n) {
        blocks.push_back({head, size - head + n});
        used += size - head + n;
        head = n;
        return 0;
      }
      return -1;
    }
    if (head < tail && tail - head >= n) {
      blocks.push_back({head, n});
      used += n;
      head += n;
      return head - n;
    }
    return -1;
  }
  void release() {
    const auto b = blocks.front();
    blocks.pop_front();
    used -= b.second;
    tail = (b.first + b.second) % size;
    if (used == 0) head = tail = 0;
  }
};

struct SReq {
  int64_t ticket = 0;
  int64_t n = 0;
  const double* in = nullptr;
  const int64_t* in_mult = nullptr;
  bool generated = false;  // created on the device from ens (no upload)
  DevEnsemble ens{};
  double* out = nullptr;
  int64_t* out_mult = nullptr;
  int64_t ld_in = 0, ld_out = 0;
  Clock::time_point t_submit, t_admit, t_done;
  int slot = -1;
  int steps_done = 0;
  int state = kReqQueued;
  int64_t in_row = -1, out_row = -1;  // rows in the input / output rings
  hipEvent_t ev_in = nullptr;         // upload landed
};

struct Serve {
  bool open = false;
  int device = 0;
  hipStream_t stream = nullptr;  // steps
  hipStream_t up = nullptr;      // uploads
  hipStream_t down = nullptr;    // downloads
  int steps = 10, prec = PREC_MIXED, mode = NTLP_MODE_BE;
  bool use_graph = true;
  int64_t cap = 0;
  // device pool
  double* hot[2] = {};
  int64_t* mult[2] = {};
  int32_t* home[2] = {};
  int32_t* keys[2] = {};
  int32_t* perm = nullptr;
  int32_t* counts = nullptr;
  int32_t* offsets = nullptr;
  int32_t* bsum = nullptr;
  double* in = nullptr;
  int64_t* in_mult = nullptr;
  double* out = nullptr;
  int64_t* out_mult = nullptr;
  ServeCtl* ctl_dev = nullptr;
  unsigned long long* tally = nullptr;
  // host-side pool state (engine thread only)
  ServeCtl* ctl_host[kServeCtlRing] = {};
  hipEvent_t ctl_ev[kServeCtlRing] = {};
  int ctl_i = 0;
  int cur = 0;
  int64_t hw = 0;
  std::map<std::pair<int64_t, int>, hipGraphExec_t> graphs;
  std::vector<SReq*> active;
  std::deque<SReq*> loading;  // accepted, upload in flight
  bool slot_used[kServeSlots] = {};
  int64_t reserved = 0;       // pool rows promised to loading requests
  Ring in_ring, out_ring;
  struct Pending {
    hipEvent_t ev;
    std::vector<SReq*> reqs;
  };
  std::deque<Pending> pending;   // downloads in flight
  std::deque<hipEvent_t> admitted;  // admissions that still read the input ring
  // shared with the client threads
  std::thread th;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  std::deque<SReq*> queue;
  std::unordered_map<int64_t, SReq*> reqs;
  int64_t next_ticket = 1;
  int err = 0;
  std::string err_msg;
  // statistics
  int64_t submitted = 0, completed = 0, iterations = 0, droplet_steps = 0, max_pool = 0, graph_launches = 0;
  double ms_engine = 0, ms_copy_calls = 0, ms_wait_download = 0, ms_wait_upload = 0, ms_wait_control = 0;
  ntlp_serve_totals snap = {};  // the engine's counters as of its last iteration (under mu)
  Clock::time_point t_engine0;
};

// Adds the time since construction to a host-time total (engine thread).
struct HostTimer {
  double& acc;
  Clock::time_point t0 = Clock::now();
  explicit HostTimer(double& a) : acc(a) {}
  ~HostTimer() { acc += std::chrono::duration<double, std::milli>(Clock::now() - t0).count(); }
};

Serve srv;

int serve_fail(int code, const std::string& msg) {
  std::lock_guard<std::mutex> lk(srv.mu);
  if (!srv.err) {
    srv.err = code;
    srv.err_msg = msg;
  }
  return code;
}

#define SHIP(call)                                                                                 \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess) return serve_fail(NTLP_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

void serve_release() {
  for (auto& kv : srv.graphs) (void)hipGraphExecDestroy(kv.second);
  srv.graphs.clear();
  for (int b = 0; b < 2; ++b) {
    dfree(srv.hot[b]);
    dfree(srv.mult[b]);
    dfree(srv.home[b]);
    dfree(srv.keys[b]);
  }
  dfree(srv.perm);
  dfree(srv.counts);
  dfree(srv.offsets);
  dfree(srv.bsum);
  dfree(srv.in);
  dfree(srv.in_mult);
  dfree(srv.out);
  dfree(srv.out_mult);
  dfree(srv.ctl_dev);
  dfree(srv.tally);
  for (int i = 0; i < kServeCtlRing; ++i) {
    if (srv.ctl_host[i]) (void)hipHostFree(srv.ctl_host[i]);
    if (srv.ctl_ev[i]) (void)hipEventDestroy(srv.ctl_ev[i]);
  }
  for (auto& p : srv.pending) (void)hipEventDestroy(p.ev);
  for (auto& e : srv.admitted) (void)hipEventDestroy(e);
  for (auto& kv : srv.reqs) {
    if (kv.second->ev_in) (void)hipEventDestroy(kv.second->ev_in);
    delete kv.second;
  }
  if (srv.stream) (void)hipStreamDestroy(srv.stream);
  if (srv.up) (void)hipStreamDestroy(srv.up);
  if (srv.down) (void)hipStreamDestroy(srv.down);
  srv.~Serve();
  new (&srv) Serve();
}

StepArgs serve_step_args(int cur) {
  StepArgs a;
  memset(&a, 0, sizeof(a));
  a.g = ctx.g;
  a.C = ctx.C;
  a.Cp = ctx.d_consts;
  a.fl = ctx.fl;
  a.F = fields_view();
  a.n = srv.cap;
  a.cap = srv.cap;
  a.src = {srv.hot[cur], srv.mult[cur], srv.home[cur]};
  a.dst = {srv.hot[1 - cur], srv.mult[1 - cur], srv.home[1 - cur]};
  a.perm = srv.perm;
  a.keys_in = srv.keys[cur];
  a.keys = srv.keys[1 - cur];
  a.counts = srv.counts;
  a.tally = srv.tally;
  a.serve = srv.ctl_dev;
  a.out = srv.out;
  a.out_mult = srv.out_mult;
  a.out_cap = srv.cap;
  a.it = 2;  // droplets are created with vp = uf and Tp = Tf at admission
  a.box_bytes = step_box_bytes(ctx.g.nnz, ctx.stage_gather && ctx.fl.ilin == 1);
  return a;
}

template <int PREC>
void serve_launch_step(const StepArgs& a, unsigned grid) {
  const size_t lds = step_lds_bytes(a.g.nnz, a.box_bytes);
  if (srv.mode == NTLP_MODE_RK3) {
    if (ctx.fl.ilin == 1) k_step<1, PREC, 1, true><<<grid, kBlock, lds, srv.stream>>>(a);
    else k_step<1, PREC, 0, true><<<grid, kBlock, lds, srv.stream>>>(a);
  } else {
    if (ctx.fl.ilin == 1) k_step<2, PREC, 1, true><<<grid, kBlock, lds, srv.stream>>>(a);
    else k_step<2, PREC, 0, true><<<grid, kBlock, lds, srv.stream>>>(a);
  }
}

// The per-iteration kernels for a pool of at most `size` slots: counting sort
// of the current buffer (the counts were gathered by the previous step and
// the admissions; the order within a cell is left to the atomics, since no
// droplet's arithmetic depends on it), then the step.  Sizes come from the
// control block.
void serve_enqueue_step(int64_t size, int cur) {
  const int64_t nb = (ctx.ncell + kScanTile - 1) / kScanTile;
  k_scan_reduce<<<(unsigned)nb, kScanThreads, 0, srv.stream>>>(srv.counts, ctx.ncell, srv.bsum);
  k_scan_bsum<<<1, kScanThreads, 0, srv.stream>>>(srv.bsum, nb);
  k_scan_down<<<(unsigned)nb, kScanThreads, 0, srv.stream>>>(srv.counts, ctx.ncell, srv.bsum, srv.offsets);
  k_bin_scatter<<<bin_blocks(size), kBlock, 0, srv.stream>>>(srv.keys[cur], size, srv.offsets, srv.counts, srv.perm,
                                                        &srv.ctl_dev->hw);
  const StepArgs a = serve_step_args(cur);
  if (srv.prec == PREC_FP64) serve_launch_step<PREC_FP64>(a, grid8(size));
  else if (srv.prec == PREC_MIXED) serve_launch_step<PREC_MIXED>(a, grid8(size));
  else serve_launch_step<PREC_FP64_AJ>(a, grid8(size));
}

int serve_run_step(int64_t slots) {
  int64_t size = 4096;
  while (size < slots) size *= 2;
  size = std::min(size, srv.cap);
  if (!srv.use_graph) {
    serve_enqueue_step(size, srv.cur);
    SHIP(hipGetLastError());
    return 0;
  }
  const auto key = std::make_pair(size, srv.cur);
  auto it = srv.graphs.find(key);
  if (it == srv.graphs.end()) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    SHIP(hipStreamBeginCapture(srv.stream, hipStreamCaptureModeThreadLocal));
    serve_enqueue_step(size, srv.cur);
    const hipError_t launch_err = hipGetLastError();
    const hipError_t end_err = hipStreamEndCapture(srv.stream, &graph);
    SHIP(launch_err);
    SHIP(end_err);
    const hipError_t inst_err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    SHIP(inst_err);
    it = srv.graphs.emplace(key, exec).first;
  }
  SHIP(hipGraphLaunch(it->second, srv.stream));
  srv.graph_launches++;
  return 0;
}

// Completed downloads: their requests are done and their output rows free.
// block: wait for the oldest download if none has landed.
int serve_retire(bool block) {
  while (!srv.pending.empty()) {
    auto& p = srv.pending.front();
    const hipError_t q = hipEventQuery(p.ev);
    if (q == hipErrorNotReady) {
      if (!block && srv.pending.size() <= 8) break;
      HostTimer w(srv.ms_wait_download);
      SHIP(hipEventSynchronize(p.ev));
    } else if (q != hipSuccess) {
      return serve_fail(NTLP_E_HIP, std::string("serving download failed: ") + hipGetErrorString(q));
    }
    block = false;
    const auto now = Clock::now();
    {
      std::lock_guard<std::mutex> lk(srv.mu);
      for (SReq* r : p.reqs) {
        r->state = kReqDone;
        r->t_done = now;
        srv.completed++;
      }
    }
    for (size_t i = 0; i < p.reqs.size(); ++i) srv.out_ring.release();
    (void)hipEventDestroy(p.ev);
    srv.pending.pop_front();
    srv.cv_done.notify_all();
  }
  // input rows whose admission kernels have run
  while (!srv.admitted.empty()) {
    const hipError_t q = hipEventQuery(srv.admitted.front());
    if (q == hipErrorNotReady) break;
    if (q != hipSuccess) return serve_fail(NTLP_E_HIP, std::string("serving admission failed: ") + hipGetErrorString(q));
    (void)hipEventDestroy(srv.admitted.front());
    srv.admitted.pop_front();
    srv.in_ring.release();
  }
  return 0;
}

// Starts the uploads of newly accepted requests.
int serve_upload(std::vector<SReq*>& accepted) {
  HostTimer w(srv.ms_copy_calls);
  for (SReq* r : accepted) {
    if (r->generated) {  // nothing to upload: admitted at once
      srv.loading.push_back(r);
      continue;
    }
    SHIP(hipMemcpy2DAsync(srv.in + 6 * r->in_row, sizeof(double) * r->n, r->in, sizeof(double) * r->ld_in,
                          sizeof(double) * r->n, 6, hipMemcpyHostToDevice, srv.up));
    SHIP(hipMemcpyAsync(srv.in_mult + r->in_row, r->in_mult, sizeof(int64_t) * r->n, hipMemcpyHostToDevice, srv.up));
    SHIP(hipEventCreateWithFlags(&r->ev_in, hipEventDisableTiming));
    SHIP(hipEventRecord(r->ev_in, srv.up));
    srv.loading.push_back(r);
  }
  return 0;
}

// Appends the requests whose upload has landed (waits for the first one when
// the pool would otherwise have nothing to step).
int serve_admit() {
  while (!srv.loading.empty()) {
    SReq* r = srv.loading.front();
    const hipError_t q = r->generated ? hipSuccess : hipEventQuery(r->ev_in);
    if (q == hipErrorNotReady) {
      if (!srv.active.empty()) break;
      HostTimer w(srv.ms_wait_upload);
      SHIP(hipEventSynchronize(r->ev_in));
    } else if (q != hipSuccess) {
      return serve_fail(NTLP_E_HIP, std::string("serving upload failed: ") + hipGetErrorString(q));
    }
    if (!r->generated) SHIP(hipStreamWaitEvent(srv.stream, r->ev_in, 0));
    AdmitArgs a;
    a.g = ctx.g;
    a.C = ctx.C;
    a.fl = ctx.fl;
    a.F = fields_view();
    a.in = r->generated ? nullptr : srv.in + 6 * r->in_row;
    a.in_mult = r->generated ? nullptr : srv.in_mult + r->in_row;
    a.ens = r->ens;
    a.m = r->n;
    a.cap = srv.cap;
    a.base = srv.hw;
    a.tag = (int32_t)(r->slot << kServeIdxBits);
    a.hot = srv.hot[srv.cur];
    a.mult = srv.mult[srv.cur];
    a.home = srv.home[srv.cur];
    a.keys = srv.keys[srv.cur];
    a.counts = srv.counts;
    if (ctx.fl.ilin == 1) k_serve_admit<1><<<grid8(r->n), kBlock, 0, srv.stream>>>(a);
    else k_serve_admit<0><<<grid8(r->n), kBlock, 0, srv.stream>>>(a);
    SHIP(hipGetLastError());
    if (!r->generated) {  // its input-ring rows are free once the admission has run
      hipEvent_t ev;
      SHIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      SHIP(hipEventRecord(ev, srv.stream));
      srv.admitted.push_back(ev);
    }
    if (r->ev_in) (void)hipEventDestroy(r->ev_in);
    r->ev_in = nullptr;
    srv.loading.pop_front();
    srv.hw += r->n;
    srv.reserved -= r->n;
    r->t_admit = Clock::now();
    r->state = kReqActive;
    srv.active.push_back(r);
  }
  return 0;
}

// Copies the engine thread's counters where ntlp_serve_stats reads them.
void serve_publish() {
  std::lock_guard<std::mutex> lk(srv.mu);
  ntlp_serve_totals& t = srv.snap;
  t.iterations = srv.iterations;
  t.droplet_steps = srv.droplet_steps;
  t.max_pool = srv.max_pool;
  t.graph_launches = srv.graph_launches;
  t.graphs = (int64_t)srv.graphs.size();
  t.capacity = srv.cap;
  t.ms_engine = std::chrono::duration<double, std::milli>(Clock::now() - srv.t_engine0).count();
  t.ms_copy_calls = srv.ms_copy_calls;
  t.ms_wait_download = srv.ms_wait_download;
  t.ms_wait_upload = srv.ms_wait_upload;
  t.ms_wait_control = srv.ms_wait_control;
}

// One step of the pool and the downloads of the requests it finishes.
int serve_step() {
  const int ci = srv.ctl_i;
  srv.ctl_i = (srv.ctl_i + 1) % kServeCtlRing;
  {
    HostTimer w(srv.ms_wait_control);
    SHIP(hipEventSynchronize(srv.ctl_ev[ci]));
  }
  ServeCtl* c = srv.ctl_host[ci];
  memset(c, 0, sizeof(ServeCtl));
  c->hw = srv.hw;
  int64_t n_step = 0;
  std::vector<SReq*> leaving;
  for (SReq* r : srv.active) {
    n_step += r->n;
    if (r->steps_done + 1 == srv.steps) {
      int64_t row;
      while ((row = srv.out_ring.alloc(r->n)) < 0) {
        if (srv.pending.empty()) return serve_fail(NTLP_E_RANGE, "output ring exhausted");
        CHK(serve_retire(true));
      }
      r->out_row = row;
      c->last[r->slot] = 1;
      c->out_off[r->slot] = row;
      leaving.push_back(r);
    }
  }
  c->n_live = n_step;
  SHIP(hipMemcpyAsync(srv.ctl_dev, c, sizeof(ServeCtl), hipMemcpyHostToDevice, srv.stream));
  SHIP(hipEventRecord(srv.ctl_ev[ci], srv.stream));
  CHK(serve_run_step(srv.hw));
  if (!leaving.empty()) {
    HostTimer w(srv.ms_copy_calls);
    hipEvent_t stepped, landed;
    SHIP(hipEventCreateWithFlags(&stepped, hipEventDisableTiming));
    SHIP(hipEventRecord(stepped, srv.stream));
    SHIP(hipStreamWaitEvent(srv.down, stepped, 0));
    (void)hipEventDestroy(stepped);
    for (SReq* r : leaving) {
      SHIP(hipMemcpy2DAsync(r->out, sizeof(double) * r->ld_out, srv.out + r->out_row, sizeof(double) * srv.cap,
                            sizeof(double) * r->n, kNHot, hipMemcpyDeviceToHost, srv.down));
      SHIP(hipMemcpyAsync(r->out_mult, srv.out_mult + r->out_row, sizeof(int64_t) * r->n, hipMemcpyDeviceToHost,
                          srv.down));
      r->state = kReqLeaving;
    }
    SHIP(hipEventCreateWithFlags(&landed, hipEventDisableTiming));
    SHIP(hipEventRecord(landed, srv.down));
    srv.pending.push_back({landed, leaving});
  }
  for (SReq* r : srv.active) r->steps_done++;
  std::vector<SReq*> stay;
  {
    std::lock_guard<std::mutex> lk(srv.mu);
    for (SReq* r : srv.active) {
      if (r->state == kReqLeaving) srv.slot_used[r->slot] = false;
      else stay.push_back(r);
    }
  }
  srv.active.swap(stay);
  srv.max_pool = std::max(srv.max_pool, srv.hw);
  // the step wrote every stepped droplet to slots 0..n_step-1 (leaving ones
  // with key -1, dropped by the next sort); an empty pool starts over
  srv.hw = srv.active.empty() ? 0 : n_step;
  srv.cur = 1 - srv.cur;
  srv.iterations++;
  srv.droplet_steps += n_step;
  serve_publish();
  return 0;
}

void serve_loop() {
  if (hipSetDevice(srv.device) != hipSuccess) {
    serve_fail(NTLP_E_HIP, "hipSetDevice in the serving thread failed");
    return;
  }
  HostTimer whole(srv.ms_engine);
  srv.t_engine0 = whole.t0;
  for (;;) {
    std::vector<SReq*> accepted;
    {
      std::unique_lock<std::mutex> lk(srv.mu);
      srv.cv_work.wait(lk, [] {
        return srv.stop || srv.err || !srv.queue.empty() || !srv.active.empty() || !srv.loading.empty() ||
               !srv.pending.empty() || !srv.admitted.empty();
      });
      if (srv.err) break;
      if (srv.stop && srv.queue.empty() && srv.active.empty() && srv.loading.empty() && srv.pending.empty() &&
          srv.admitted.empty())
        break;
      while (!srv.queue.empty()) {  // FIFO: the head waits for a slot, pool room and input rows
        SReq* r = srv.queue.front();
        int slot = -1;
        for (int k = 0; k < kServeSlots; ++k)
          if (!srv.slot_used[k]) {
            slot = k;
            break;
          }
        if (slot < 0 || r->n > srv.cap - srv.hw - srv.reserved) break;
        const int64_t row = r->generated ? 0 : srv.in_ring.alloc(r->n);
        if (row < 0) break;
        srv.queue.pop_front();
        srv.slot_used[slot] = true;
        srv.reserved += r->n;
        r->slot = slot;
        r->in_row = row;
        r->state = kReqLoading;
        accepted.push_back(r);
      }
    }
    int rc = serve_upload(accepted);
    if (!rc) rc = serve_admit();
    if (!rc) {
      if (!srv.active.empty()) rc = serve_step();
      if (!rc) rc = serve_retire(srv.active.empty() && srv.loading.empty() && !srv.pending.empty());
      if (!rc && srv.active.empty() && srv.loading.empty() && srv.pending.empty() && !srv.admitted.empty())
        rc = (hipEventSynchronize(srv.admitted.front()) == hipSuccess) ? serve_retire(false)
                                                                     : serve_fail(NTLP_E_HIP, "admission failed");
    }
    if (rc) break;
  }
  // on an error every request still in flight fails
  std::lock_guard<std::mutex> lk(srv.mu);
  if (srv.err) {
    for (auto& kv : srv.reqs)
      if (kv.second->state != kReqDone) kv.second->state = kReqFailed;
  }
  srv.cv_done.notify_all();
}

}  // namespace

extern "C" {

int ntlp_serve_open(const ntlp_serve_config* cfg, int64_t* capacity) {
  CHK(need_fields());
  if (srv.open) return fail(NTLP_E_STATE, "the serving loop is already open");
  if (!cfg) return fail(NTLP_E_INVALID, "null config");
  CHK(check_prec(cfg->precision));
  if (cfg->steps < 1) return fail(NTLP_E_INVALID, "steps must be >= 1");
  if (cfg->mode != 0 && cfg->mode != NTLP_MODE_BE && cfg->mode != NTLP_MODE_RK3)
    return fail(NTLP_E_INVALID, "serving mode must be NTLP_MODE_BE (or 0) or NTLP_MODE_RK3");
  const int64_t cap = cfg->hbm_budget_bytes / kServeBytesPerDroplet;
  if (cap < 4096) return fail(NTLP_E_RANGE, "HBM budget below 4096 droplets");
  if (cap > INT32_MAX) return fail(NTLP_E_RANGE, "HBM budget above 2^31 droplets");
  srv.device = 0;
  HIPCHK(hipGetDevice(&srv.device));
  srv.steps = cfg->steps;
  srv.prec = cfg->precision;
  srv.mode = cfg->mode == NTLP_MODE_RK3 ? NTLP_MODE_RK3 : NTLP_MODE_BE;
  srv.use_graph = cfg->use_graph != 0;
  srv.cap = cap;
  int rc = 0;
  auto alloc_all = [&]() -> int {
    HIPCHK(hipStreamCreateWithFlags(&srv.stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&srv.up, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&srv.down, hipStreamNonBlocking));
    srv.in_ring.size = cap;
    srv.out_ring.size = cap;
    for (int b = 0; b < 2; ++b) {
      CHK(dalloc(srv.hot[b], tiled_elems(cap, kNHot), "serving pool"));
      CHK(dalloc(srv.mult[b], cap, "serving pool"));
      CHK(dalloc(srv.home[b], cap, "serving pool"));
      CHK(dalloc(srv.keys[b], cap, "serving pool"));
    }
    CHK(dalloc(srv.perm, cap, "serving pool"));
    CHK(dalloc(srv.counts, ctx.ncell, "serving bins"));
    CHK(dalloc(srv.offsets, ctx.ncell + 1, "serving bins"));
    CHK(dalloc(srv.bsum, ctx.nbsum, "serving bins"));
    CHK(dalloc(srv.in, 6 * cap, "serving admission staging"));
    CHK(dalloc(srv.in_mult, cap, "serving admission staging"));
    CHK(dalloc(srv.out, kNHot * cap, "serving output rows"));
    CHK(dalloc(srv.out_mult, cap, "serving output rows"));
    CHK(dalloc(srv.ctl_dev, 1, "serving control"));
    CHK(dalloc(srv.tally, kTallySlots * kTallyWidth, "serving counters"));
    for (int i = 0; i < kServeCtlRing; ++i) {
      HIPCHK(hipHostMalloc((void**)&srv.ctl_host[i], sizeof(ServeCtl), hipHostMallocDefault));
      HIPCHK(hipEventCreateWithFlags(&srv.ctl_ev[i], hipEventDisableTiming));
    }
    HIPCHK(hipMemsetAsync(srv.counts, 0, sizeof(int32_t) * ctx.ncell, srv.stream));
    HIPCHK(hipMemsetAsync(srv.tally, 0, sizeof(unsigned long long) * kTallySlots * kTallyWidth, srv.stream));
    HIPCHK(hipMemsetAsync(srv.keys[0], 0xff, sizeof(int32_t) * cap, srv.stream));
    HIPCHK(hipMemsetAsync(srv.keys[1], 0xff, sizeof(int32_t) * cap, srv.stream));
    CHK(upload_radval(srv.stream));
    HIPCHK(hipStreamSynchronize(srv.stream));
    return 0;
  };
  rc = alloc_all();
  if (rc) {
    const std::string msg = g_err;
    serve_release();
    return fail(rc, msg);
  }
  srv.open = true;
  ctx.serving = true;
  srv.th = std::thread(serve_loop);
  if (capacity) *capacity = cap;
  return NTLP_OK;
}

int ntlp_serve_submit(const double* in, int64_t ld_in, const int64_t* in_mult, int64_t n, double* out,
                      int64_t ld_out, int64_t* out_mult, int64_t* ticket) {
  if (!srv.open) return fail(NTLP_E_STATE, "the serving loop is not open");
  if (!in || !in_mult || !out || !out_mult || !ticket) return fail(NTLP_E_INVALID, "null argument");
  if (n < 1) return fail(NTLP_E_INVALID, "a request needs at least one droplet");
  if (ld_in == 0) ld_in = n;
  if (ld_out == 0) ld_out = n;
  if (ld_in < n || ld_out < n) return fail(NTLP_E_INVALID, "column stride smaller than the request");
  if (n > srv.cap || n > kServeIdxMask + 1)
    return fail(NTLP_E_RANGE, "request larger than the pool (" + std::to_string(srv.cap) + " droplets) or 2^23");
  SReq* r = new SReq();
  r->n = n;
  r->in = in;
  r->in_mult = in_mult;
  r->out = out;
  r->out_mult = out_mult;
  r->ld_in = ld_in;
  r->ld_out = ld_out;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv.mu);
    if (srv.err) {
      delete r;
      return fail(srv.err, srv.err_msg);
    }
    r->ticket = srv.next_ticket++;
    srv.reqs[r->ticket] = r;
    srv.queue.push_back(r);
    srv.submitted++;
  }
  srv.cv_work.notify_one();
  *ticket = r->ticket;
  return NTLP_OK;
}

int ntlp_serve_submit_ensemble(const ntlp_ensemble* ens, int64_t n, double* out, int64_t ld_out, int64_t* out_mult,
                               int64_t* ticket) {
  if (!srv.open) return fail(NTLP_E_STATE, "the serving loop is not open");
  CHK(check_ensemble(ens));
  if (!out || !out_mult || !ticket) return fail(NTLP_E_INVALID, "null argument");
  if (n < 1) return fail(NTLP_E_INVALID, "a request needs at least one droplet");
  if (ld_out == 0) ld_out = n;
  if (ld_out < n) return fail(NTLP_E_INVALID, "column stride smaller than the request");
  if (n > srv.cap || n > kServeIdxMask + 1)
    return fail(NTLP_E_RANGE, "request larger than the pool (" + std::to_string(srv.cap) + " droplets) or 2^23");
  SReq* r = new SReq();
  r->n = n;
  r->generated = true;
  r->ens = dev_ensemble(ens);
  r->out = out;
  r->out_mult = out_mult;
  r->ld_in = n;
  r->ld_out = ld_out;
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(srv.mu);
    if (srv.err) {
      delete r;
      return fail(srv.err, srv.err_msg);
    }
    r->ticket = srv.next_ticket++;
    srv.reqs[r->ticket] = r;
    srv.queue.push_back(r);
    srv.submitted++;
  }
  srv.cv_work.notify_one();
  *ticket = r->ticket;
  return NTLP_OK;
}

int ntlp_serve_poll(int64_t ticket) {
  if (!srv.open) return fail(NTLP_E_STATE, "the serving loop is not open");
  std::lock_guard<std::mutex> lk(srv.mu);
  auto it = srv.reqs.find(ticket);
  if (it == srv.reqs.end()) return fail(NTLP_E_INVALID, "unknown ticket");
  if (it->second->state == kReqFailed) return fail(srv.err ? srv.err : NTLP_E_HIP, srv.err_msg);
  return it->second->state == kReqDone ? 1 : 0;
}

int ntlp_serve_wait(int64_t ticket, double* times_ms) {
  if (!srv.open) return fail(NTLP_E_STATE, "the serving loop is not open");
  std::unique_lock<std::mutex> lk(srv.mu);
  auto it = srv.reqs.find(ticket);
  if (it == srv.reqs.end()) return fail(NTLP_E_INVALID, "unknown ticket");
  SReq* r = it->second;
  srv.cv_done.wait(lk, [r] { return r->state == kReqDone || r->state == kReqFailed; });
  if (r->state == kReqFailed) return fail(srv.err ? srv.err : NTLP_E_HIP, srv.err_msg);
  if (times_ms) {
    times_ms[0] = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    times_ms[1] = std::chrono::duration<double, std::milli>(r->t_admit - r->t_submit).count();
  }
  srv.reqs.erase(it);
  delete r;
  return NTLP_OK;
}

int ntlp_serve_stats(ntlp_serve_totals* out) {
  if (!srv.open) return fail(NTLP_E_STATE, "the serving loop is not open");
  if (!out) return fail(NTLP_E_INVALID, "null output");
  std::lock_guard<std::mutex> lk(srv.mu);
  *out = srv.snap;
  out->submitted = srv.submitted;
  out->completed = srv.completed;
  out->queued = (int64_t)srv.queue.size();
  out->capacity = srv.cap;
  memset(&out->newton, 0, sizeof(out->newton));
  return NTLP_OK;
}

int ntlp_serve_close(ntlp_serve_totals* final_stats) {
  if (!srv.open) return fail(NTLP_E_STATE, "the serving loop is not open");
  {
    std::lock_guard<std::mutex> lk(srv.mu);
    srv.stop = true;
  }
  srv.cv_work.notify_all();
  if (srv.th.joinable()) srv.th.join();
  const int err = srv.err;
  const std::string msg = srv.err_msg;
  if (final_stats) {
    memset(final_stats, 0, sizeof(*final_stats));
    final_stats->submitted = srv.submitted;
    final_stats->completed = srv.completed;
    final_stats->iterations = srv.iterations;
    final_stats->droplet_steps = srv.droplet_steps;
    final_stats->max_pool = srv.max_pool;
    final_stats->graph_launches = srv.graph_launches;
    final_stats->graphs = (int64_t)srv.graphs.size();
    final_stats->capacity = srv.cap;
    final_stats->ms_engine = srv.ms_engine;
    final_stats->ms_copy_calls = srv.ms_copy_calls;
    final_stats->ms_wait_download = srv.ms_wait_download;
    final_stats->ms_wait_upload = srv.ms_wait_upload;
    final_stats->ms_wait_control = srv.ms_wait_control;
    std::vector<unsigned long long> t((size_t)kTallySlots * kTallyWidth);
    if (hipMemcpy(t.data(), srv.tally, sizeof(unsigned long long) * t.size(), hipMemcpyDeviceToHost) == hipSuccess) {
      int64_t v[kTallyWidth] = {};
      for (int s = 0; s < kTallySlots; ++s)
        for (int c = 0; c < kTallyWidth; ++c) v[c] += (int64_t)t[(size_t)s * kTallyWidth + c];
      ntlp_counters& nc = final_stats->newton;
      nc.num100 = v[kTNum100];
      nc.numimpos = v[kTNumImpos];
      nc.actnum = v[kTActnum];
      nc.denum = v[kTDenum];
      nc.gn_iters = v[kTGnIters];
      nc.lm_iters = v[kTLmIters];
      nc.solves = v[kTSolves];
      nc.steps = srv.iterations;
      for (int k = 0; k < NTLP_GN_HIST; ++k) nc.gn_hist[k] = v[kTHist0 + k];
    }
  }
  if (srv.stream) (void)hipStreamSynchronize(srv.stream);
  if (srv.up) (void)hipStreamSynchronize(srv.up);
  if (srv.down) (void)hipStreamSynchronize(srv.down);
  serve_release();
  ctx.serving = false;
  if (err) return fail(err, msg);
  return NTLP_OK;
}

int ntlp_host_alloc(int64_t bytes, void** ptr) {
  if (!ptr || bytes < 0) return fail(NTLP_E_INVALID, "bad argument");
  *ptr = nullptr;
  HIPCHK(hipHostMalloc(ptr, (size_t)std::max<int64_t>(bytes, 1), hipHostMallocDefault));
  return NTLP_OK;
}

int ntlp_host_free(void* ptr) {
  if (ptr) HIPCHK(hipHostFree(ptr));
  return NTLP_OK;
}

}  // extern "C"
