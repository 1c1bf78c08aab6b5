"""Droplet-step benchmark on one MI355X.

Workload (default): 20M droplets (60% haze at activation) on the
256x256x128 tanh-stretched grid, backward-Euler droplet step (the
ipart_method=2 step every NTLP case runs once per LES RK3 step), two-way
coupling on.  Each timed step = gather + advance with the Newton radius and
temperature solve + boundary conditions + coupling scatter + counting sort.

Reports droplet-updates/s for the mixed-precision step (value) beside the
FP64 step, the roofline of the dominant kernel from HIP events on the
library's stream, and the oracle timed on a bounded sample on the host.

  python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--workload 3]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(name: str, n: int, grid, precision: int) -> float | None:
    """Algorithmic HBM bytes of one launch of a fused-step kernel over n
    droplets (each byte the kernel must move once).  `am` is the size of one
    coupling amount: 4 bytes where a mixed-precision step stores them, else 8.

    step_be (in place, residence times left in HBM until the end of the call):
      reads 12 of the 14 FP64 hot columns, mult (8) and the u-cell key (4);
      writes the 8 columns that change every step (numact and actres only when
      a droplet activates), the 6 coupling amounts and the new key; reads once
      the gathered fields over the ext array (u, v, w in FP32 for mixed
      precision, else FP64; T, q FP64).
    step_rk3 (in place): reads all 14 hot columns, mult and the key; writes 10
      columns (the 8 above and the residence times), the amounts and the key.
    *_sorted (the step that reorders into cell order): reads also the sort
      permutation (4) and home (4), writes all 14 hot columns, mult and home
      into the other buffer.
    res_settle (once per call): reads and writes the two residence-time
      columns.
    *_last: the last step of a call also writes the 11 BE diagnostics (the
      gathered fluid state and the pre-step values), 88 bytes per droplet (RK3
      adds the 9 stage registers).
    scatter: reads the sort permutation, the key, x, y, z and the 6 amounts of
      each droplet; writes each of the 6 FP64 source fields once.
    """
    ext = (grid.nnz + 2) * (grid.nny + 5) * (grid.nnx + 5)
    nodes = (grid.nnz + 2) * (grid.nny + 1) * (grid.nnx + 1)
    per_cell = (3 * 4 + 2 * 8) if precision == 1 else 5 * 8
    am = 4 if precision == 1 else 8
    if name.startswith("step_be") or name.startswith("step_rk3"):
        rk3 = name.startswith("step_rk3")
        if "_sorted" in name:
            per = (4 + 14 * 8 + 8 + 4 + 4) + (14 * 8 + 8 + 4 + 6 * am + 4)
        elif rk3:
            per = (14 * 8 + 8 + 4) + (10 * 8 + 6 * am + 4)
        else:
            per = (12 * 8 + 8 + 4) + (8 * 8 + 6 * am + 4)
        if name.endswith("_last"):
            per += (20 if rk3 else 11) * 8
        return float(n) * per + float(ext) * per_cell
    if name == "res_settle":
        return float(n) * 4 * 8
    if name == "scatter":
        return float(n) * (4 + 4 + 3 * 8 + 6 * am) + 6 * 8 * float(nodes)
    return None


FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 vector
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 vector


def pmc_compute(kernel: str, workload: str, avg_ms: float, precision_key: str = "step_be"):
    """The compute side of a step kernel's roofline: FP64 and FP32 FLOPs per
    launch from the committed rocprofv3 instruction-mix summary (the
    SQ_INSTS_VALU_FLOPS_* counters, wave instructions x 64 lanes) over the
    kernel's average launch time measured in this run, when that summary covers
    this workload; with the VALU instruction mix that bounds it."""
    path = os.path.join(ROOT, "profiles", "step_scatter_20M_256x256x128_valu_mix.json")
    if not os.path.exists(path) or not avg_ms:
        return None
    with open(path) as fh:
        pm = json.load(fh)
    k = pm.get("kernels", {}).get(precision_key)
    if pm.get("workload") != workload or k is None:
        return None
    t = avg_ms * 1e-3
    f64, f32 = k["fp64_flops"] / t / 1e12, k["fp32_flops"] / t / 1e12
    return {"unit": "TFLOP/s", "fp64_achieved": f64, "fp64_peak": FP64_PEAK_TFLOPS, "fp64_frac": f64 / FP64_PEAK_TFLOPS,
            "fp32_achieved": f32, "fp32_peak": FP32_PEAK_TFLOPS, "fp32_frac": f32 / FP32_PEAK_TFLOPS,
            "valu_insts_per_launch": k["valu_insts"],
            "valu_mix": {"fp64": k["fp64_insts"] / k["valu_insts"], "fp32": k["fp32_insts"] / k["valu_insts"],
                         "int": k["int_insts"] / k["valu_insts"], "cvt": k["cvt_insts"] / k["valu_insts"],
                         "other": k["other_insts"] / k["valu_insts"]}}


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return "unknown"


def pmc_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE, scaled by the factors a plain copy of
    known size gives), when that summary covers this workload."""
    path = os.path.join(ROOT, "profiles", "step_scatter_20M_256x256x128_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        pm = json.load(fh)
    if pm.get("workload") != workload or kernel not in pm.get("kernels", {}):
        return None
    k = pm["kernels"][kernel]
    return k["fetch_kib"] * 1024 * pm["fetch_scale"] + k["write_kib"] * 1024 * pm["write_scale"]


def prepare(wl_id: int, n: int | None, seed: int):
    """The workload's grid, constants and fields (host); its droplets are drawn
    on the device (synth.ensemble)."""
    from ntlp_amd import synth

    t = time.time()
    wl = synth.workload(wl_id, n=n, seed=seed, droplets=False)
    log(f"[bench] workload {wl_id}: {wl.name}; fields built in {time.time() - t:.1f}s")
    return wl


def workload_size(wl_id: int, n: int | None) -> int:
    return n if n is not None else {1: 100_000, 2: 2_000_000, 3: 20_000_000, 4: 2_000_000, 5: 100_000_000}[wl_id]


def init_device(wl, wl_id: int, n: int, seed: int, flags=None):
    """Fields resident and n droplets of the workload's ensemble created on the
    device (vp = uf, Tp = Tf, rc as create_particle sets them)."""
    from ntlp_amd import DropletStep, synth

    s = DropletStep(wl.grid, wl.consts, flags or wl.flags)
    s.set_fields(*wl.fields)
    s.generate(synth.ensemble(wl_id, seed=seed), n)
    return s


def run_steps(s, init_state, mode, prec, steps, warmup, it0=2):
    """Times `steps` fused steps (one call) after `warmup` untimed ones, from
    init_state (a Particles, or a (ensemble, n) pair drawn on the device)."""
    restore(s, init_state)
    s.counters(reset=True)
    if warmup:
        s.step(mode=mode, precision=prec, nsteps=warmup, it0=it0)
    s.sync()
    t0 = time.perf_counter()
    s.step(mode=mode, precision=prec, nsteps=steps, it0=it0 + warmup)
    s.sync()
    dt = time.perf_counter() - t0
    return dt


def restore(s, init_state):
    if isinstance(init_state, tuple):
        s.generate(*init_state)
    else:
        s.put(init_state)


def kernel_timing(s, init_state, mode, prec, steps, warmup):
    """Per-kernel HIP-event times of the same steps (ms per launch)."""
    restore(s, init_state)
    if warmup:
        s.step(mode=mode, precision=prec, nsteps=warmup, it0=2)
    s.timing(True)
    s.step(mode=mode, precision=prec, nsteps=steps, it0=2 + warmup)
    kt = s.kernel_times()
    s.timing(False)
    return kt


def roofline_of(kt, n, grid, prec):
    per_kernel = {}
    for k, (nl, tot) in kt.items():
        b = kernel_bytes(k, n, grid, prec)
        if b is not None:
            per_kernel[k] = {"avg_ms": tot / nl, "launches": nl, "algorithmic_bytes": b,
                             "GBps": b / (tot / nl * 1e-3) / 1e9, "frac_of_peak": b / (tot / nl * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return per_kernel


def cpu_baseline(wl_id: int, seed: int, n_full: int, target_s: float = 8.0):
    """The oracle (FP64 C restatement of the reference loop: BE step + coupling
    deposit) on a bounded sample of the same workload -- same grid and fields,
    the first droplets of the same ensemble -- single-threaded and over all
    host cores (OpenMP).  Each step also zeroes the full source fields
    (particles.f90:1167-1170), a per-step cost that a small sample would
    otherwise spread over few droplets: it is timed apart and the full
    workload's CPU step is estimated as t_droplets * n_full / n_sample +
    t_zero."""
    from ntlp_amd import COL, synth
    from oracle import Oracle

    def leg(n_probe: int, par: bool):
        wl = synth.workload(wl_id, n=n_probe, seed=seed)
        ora = Oracle(wl.grid, wl.consts, wl.flags)
        ora.set_fields(*wl.fields)
        p = wl.droplets
        ora.interp(p.soa)
        for d in "xyz":
            p["v" + d] = p["uf" + d]
        p["Tp"] = p["Tf"]
        p["rc"] = ora.crit_radius(p.soa, COL)
        upd = ora.update_be_par if par else ora.update_be
        cpl = ora.coupling_par if par else ora.coupling
        t0 = time.perf_counter()
        cpl(p.soa[:, :0].copy(), p.mult[:0])  # zeroing alone
        t_zero = time.perf_counter() - t0
        t0 = time.perf_counter()
        upd(p.soa, 2)
        cpl(p.soa, p.mult)
        t_probe = time.perf_counter() - t0
        reps = int(min(max(1, target_s / max(t_probe, 1e-3)), 200))
        t0 = time.perf_counter()
        for r in range(reps):
            upd(p.soa, 3 + r)
            cpl(p.soa, p.mult)
        t = time.perf_counter() - t0
        t_drop = max(t / reps - t_zero, 1e-12)
        rate = n_full / (t_drop * n_full / n_probe + t_zero)
        return rate, reps, t, t_zero, ora.max_threads() if par else 1

    r1, reps1, t1, z1, _ = leg(20000, False)
    rp, repsp, tp, zp, cores = leg(200000, True)
    return {"value": rp, "unit": "droplet-updates/s", "cores": cores, "kind": "port",
            "cpu_model": cpu_model(), "single_thread_value": r1,
            "sample": f"oracle (C, FP64, -O2): BE step + coupling deposit of the same ensemble and fields; "
                      f"{cores} OpenMP threads on 200000 droplets x {repsp} steps ({tp:.1f} s) and 1 thread on "
                      f"20000 droplets x {reps1} steps ({t1:.1f} s); zeroing the sources costs {zp * 1e3:.0f} ms "
                      f"({z1 * 1e3:.0f} ms on 1 thread) per step and is added once per step of {n_full} droplets"}


def serve_bench(seed: int, duration_s: float, rate: float, steps: int = 10, budget_bytes: int = 32 << 30,
                nmax: int = 4_000_000, mode: int | None = None, dt: float | None = None):
    """The serving stream against resident 128^3 fields: Poisson arrivals at
    `rate` req/s for `duration_s`, request sizes log-uniform on [1e3, 4e6]
    droplets, each request its own ensemble -- its own seed, radius median
    drawn from 2-20 um (sigma_ln 0.4), aerosol median from 0.02-0.2 um
    (sigma_ln 0.5) -- drawn on the device at admission
    (ntlp_serve_submit_ensemble) and advancing `steps` mixed-precision steps
    of `mode` (BE by default; RK3 with a `dt` at which the explicit stages are
    stable for the ensembles' smallest droplets).  Results (14 columns and
    mult) land in a pinned FIFO ring of 3 x 4M rows.  Phases:

      stream     Poisson arrivals; p50/p99 latency from submission to result
                 in host memory;
      saturated  a second set of requests all arriving at t = 0 (a closed
                 loop bounded by the result ring): sustained throughput;
      batched    the same requests through the plain entry points, in batches
                 that fill the result ring: the batch's ensembles generated
                 together (ntlp_gpu_particles_generate), `steps` fused steps,
                 the 14 result columns and mult copied into the same pinned
                 ring -- the comparator of the sustained throughput.
    """
    import ctypes
    import queue
    import threading

    from ntlp_amd import (COL, MODE_BE, PREC_MIXED, SERVE_OUT, DropletStep, Ensemble, Flags, PinnedArray, ServeLoop,
                          SALT_DRY_LOGNORMAL, synth)

    mode = MODE_BE if mode is None else mode
    wl = synth.workload(4, n=1, seed=seed, droplets=False)
    if dt is not None:
        wl.consts.dt = dt
    flags = Flags(icouple=0, iTcouple=0, iHcouple=0)
    s = DropletStep(wl.grid, wl.consts, flags)
    s.set_fields(*wl.fields)
    ring_rows = 3 * nmax
    outbuf = PinnedArray((14, ring_rows))
    outmult = PinnedArray((ring_rows,), np.int64)

    class OutRing:  # FIFO rows of the pinned result buffer
        def __init__(self):
            self.head, self.blocks, self.cv = 0, [], threading.Condition()

        def take(self, n):
            with self.cv:
                while True:
                    start = self.head if self.head + n <= ring_rows else 0
                    if all(start + n <= a or start >= b for a, b in self.blocks):
                        self.blocks.append((start, start + n))
                        self.head = start + n
                        return start
                    self.cv.wait()

        def give(self):
            with self.cv:
                self.blocks.pop(0)
                self.cv.notify_all()

    rng = np.random.default_rng(seed + 77)

    def draw(k, n0):
        sz = np.exp(rng.uniform(np.log(1e3), np.log(4e6), k)).astype(np.int64)
        ens = [Ensemble(seed=1_000_003 * seed + n0 + i, radius_p=(rng.uniform(2e-6, 20e-6), 0.4),
                        salt_kind=SALT_DRY_LOGNORMAL, salt_p=(rng.uniform(0.02e-6, 0.2e-6), 0.5)) for i in range(k)]
        return sz, ens

    loop = ServeLoop(s, budget_bytes, steps=steps, precision=PREC_MIXED, mode=mode)

    def submit(ring, e, n):
        row = ring.take(n)
        return loop.submit_ensemble(e, n, outbuf.array[:, row:row + n], outmult.array[row:row + n])

    def run_stream(arrivals, sizes, ens):
        ring = OutRing()
        done_q = queue.Queue()
        lat, qd, comp = [], [], []

        def waiter():
            while True:
                item = done_q.get()
                if item is None:
                    return
                a, b = loop.wait(item)
                lat.append(a)
                qd.append(b)
                comp.append(time.perf_counter())
                ring.give()

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for ta, n, e in zip(arrivals, sizes, ens):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            done_q.put(submit(ring, e, int(n)))
        done_q.put(None)
        th.join()
        return np.array(lat), np.array(qd), t0, max(comp)

    def delta(a, b):
        return {k: b[k] - a[k] for k in ("iterations", "droplet_steps", "graph_launches")}

    nreq = max(1, int(rate * duration_s))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    sizes, ens = draw(nreq, 0)
    st0 = loop.stats()
    lat, qd, t0, t1 = run_stream(arrivals, sizes, ens)
    st1 = loop.stats()
    d = delta(st0, st1)
    work = float(sizes.sum()) * steps
    stream = {"requests": int(nreq), "rate_req_per_s": rate, "duration_s": float(arrivals[-1]),
              "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
              "p50_queue_ms": float(np.percentile(qd, 50)), "p99_queue_ms": float(np.percentile(qd, 99)),
              "throughput_du_s": work / (t1 - t0), "offered_du_s": rate * float(sizes.mean()) * steps,
              "mean_pool": d["droplet_steps"] / max(d["iterations"], 1), "iterations": d["iterations"]}
    log(f"[serve] stream: p50 {stream['p50_ms']:.1f} ms, p99 {stream['p99_ms']:.1f} ms, "
        f"{stream['throughput_du_s']:.3g} du/s")

    nsat = max(16, nreq // 4)
    sizes_s, ens_s = draw(nsat, 10_000_000)
    lat_s, _, t0, t1 = run_stream(np.zeros(nsat), sizes_s, ens_s)
    st2 = loop.stats()
    d = delta(st1, st2)
    sat_tp = float(sizes_s.sum()) * steps / (t1 - t0)
    mean_pool = d["droplet_steps"] / max(d["iterations"], 1)
    tot = loop.close()
    log(f"[serve] saturated {sat_tp:.3g} du/s (mean pool {mean_pool:.3g}); engine thread (ms) {tot['engine_ms']}")

    # the same requests through the plain entry points, in batches filling the ring
    lib = s.lib
    cols = [COL[c] for c in SERVE_OUT]
    tb = time.perf_counter()
    i, nb = 0, 0
    while i < nsat:
        j, rows = i, 0
        while j < nsat and rows + sizes_s[j] <= ring_rows:
            rows += int(sizes_s[j])
            j += 1
        s.generate(ens_s[i:j], sizes_s[i:j])
        s.step(mode, PREC_MIXED, nsteps=steps, it0=2)
        for k, c in enumerate(cols):
            s._check(lib.ntlp_gpu_particles_get_column(c, outbuf.array[k].ctypes.data_as(ctypes.c_void_p), rows))
        s._check(lib.ntlp_gpu_particles_get(None, outmult.array.ctypes.data_as(ctypes.c_void_p), None, None, rows))
        i, nb = j, nb + 1
    batched = float(sizes_s.sum()) * steps / (time.perf_counter() - tb)
    log(f"[serve] plain-path batches {batched:.3g} du/s ({nb} batches)")
    saturated = {"requests": int(nsat), "throughput_du_s": sat_tp, "mean_pool": mean_pool,
                 "iterations": d["iterations"],
                 "p50_ms": float(np.percentile(lat_s, 50)), "p99_ms": float(np.percentile(lat_s, 99)),
                 "batched_du_s": batched, "batches": nb,
                 "sustained_over_batched": sat_tp / batched,
                 "host_link_bound_du_s": steps * 56e9 / (14 * 8 + 8)}
    s.close()
    outbuf.free()
    outmult.free()
    mname = "BE" if mode == MODE_BE else f"RK3 (dt = {wl.consts.dt:g} s)"
    return {"workload": "128^3 resident fields, Poisson arrivals, sizes log-uniform 1e3-4e6, each request its own "
                        f"seed, radius median U(2,20) um and aerosol median U(0.02,0.2) um drawn on the device, "
                        f"{steps} {mname} steps each, mixed precision, HBM budget {budget_bytes >> 30} GiB",
            "stream": stream, "saturated": saturated, "graphs": tot["graphs"], "capacity": tot["capacity"],
            "newton": tot["newton"]}


def rk3_bench(seed: int, steps: int = 10, warmup: int = 2):
    """RK3 (ipart_method = 1, three gather + advance stages per step) on the
    128^3 workload's 2M droplets, FP64 and mixed, at dt = 2e-5 s: at the
    workload's own 5e-4 s the explicit drag of droplets below ~4 um has
    dt/tau_p > 2.5 and the scheme diverges (in the reference as here), so the
    throughput is measured where RK3 is stable for the whole population."""
    from ntlp_amd import MODE_RK3, PREC_FP64, PREC_MIXED, synth

    wl = prepare(2, None, seed)
    n = workload_size(2, None)
    s = init_device(wl, 2, n, seed)
    s.set_dt(2e-5)
    init = (synth.ensemble(2, seed=seed), n)
    out = {"workload": f"{wl.name} at dt = 2e-5 s (RK3 stable for every droplet)", "droplets": n}
    for prec, nm in ((PREC_MIXED, "mixed"), (PREC_FP64, "fp64")):
        t = run_steps(s, init, MODE_RK3, prec, steps, warmup)
        kt = kernel_timing(s, init, MODE_RK3, prec, steps, warmup)
        out[nm] = {"value": n * steps / t, "ms_per_step": 1e3 * t / steps,
                   "kernels": roofline_of(kt, n, wl.grid, prec)}
        log(f"[rk3] {nm}: {n * steps / t:.3g} du/s, " + ", ".join(f"{k}={v[1] / v[0]:.3f}ms" for k, v in kt.items()))
    out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    s.close()
    return out


def config5_bench(seed: int, steps: int = 20, warmup: int = 2):
    """The large single-GPU case: 512x512x256, 100M droplets (lognormal 10 um,
    trimodal salt) drawn on the device, `steps` mixed BE steps with coupling,
    HBM GB/s per kernel."""
    from ntlp_amd import MODE_BE, PREC_MIXED, synth

    wl = prepare(5, None, seed)
    n = workload_size(5, None)
    s = init_device(wl, 5, n, seed)
    init = (synth.ensemble(5, seed=seed), n)
    t = run_steps(s, init, MODE_BE, PREC_MIXED, steps, warmup)
    kt = kernel_timing(s, init, MODE_BE, PREC_MIXED, steps, warmup)
    s.close()
    out = {"workload": wl.name, "droplets": n, "steps": steps, "value": n * steps / t,
           "ms_per_step": 1e3 * t / steps, "kernels": roofline_of(kt, n, wl.grid, PREC_MIXED)}
    log(f"[config5] {out['value']:.3g} du/s, {out['ms_per_step']:.2f} ms/step; " +
        ", ".join(f"{k}={v['GBps']:.0f}GB/s" for k, v in out["kernels"].items()))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", type=int, default=3)
    ap.add_argument("--droplets", type=int, default=None)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-rk3", action="store_true")
    ap.add_argument("--no-config5", action="store_true")
    ap.add_argument("--no-onestep", action="store_true")
    ap.add_argument("--serve-only", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--no-copy", action="store_true",
                    help="skip the device-copy bandwidth measurement")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        from ntlp_amd import Consts, DropletStep, Flags
        from ntlp_amd.grid import Grid

        s = DropletStep(Grid.uniform(8, 8, 8, 0.01, 0.01, 0.01), Consts(), Flags())
        gbps = s.copy_bandwidth(1.0, 3)  # 1 GiB read + 1 GiB written per launch (k_copy16), 4 launches
        s.close()
        print(json.dumps({"copy_probe_bytes_read": 1 << 30, "copy_probe_bytes_written": 1 << 30, "launches": 4,
                          "GBps": gbps}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (the droplet step shards by "
              "droplets with an all-reduce of the coupling fields; that path is not built).", file=sys.stderr)
        sys.exit(2)

    from ntlp_amd import MODE_BE, MODE_RK3, PREC_FP64, PREC_FP64_AJ, PREC_MIXED, synth

    if args.serve_only:
        print(json.dumps({"serve": serve_bench(args.seed, args.serve_seconds, args.serve_rate),
                          "serve_rk3": serve_bench(args.seed, args.serve_seconds, args.serve_rate, mode=MODE_RK3,
                                                   dt=1e-6)}), flush=True)
        return

    wl = prepare(args.workload, args.droplets, args.seed)
    n = workload_size(args.workload, args.droplets)
    s = init_device(wl, args.workload, n, args.seed)
    init = (synth.ensemble(args.workload, seed=args.seed), n)
    g = wl.grid
    log(f"[bench] {n} droplets on {g.nnx}x{g.nny}x{g.nnz} drawn on the device")

    # mixed precision: the headline
    t_mixed = run_steps(s, init, MODE_BE, PREC_MIXED, args.steps, args.warmup)
    cnt_mixed = s.counters()
    rows = s.step_counters()
    end_mixed = {c: s.column(c) for c in ("radius", "Tp", "z", "vz")}
    log(f"[bench] mixed: {args.steps} steps in {t_mixed:.3f}s")
    ktimes = kernel_timing(s, init, MODE_BE, PREC_MIXED, args.steps, args.warmup)
    log("[bench] kernel times (mixed): " + ", ".join(f"{k}={v[1] / v[0]:.3f}ms x{v[0]}" for k, v in ktimes.items()))

    fp64 = None
    accuracy = None
    if not args.no_fp64:
        t_fp64 = run_steps(s, init, MODE_BE, PREC_FP64, args.steps, args.warmup)
        cnt_fp64 = s.counters()
        end_fp64 = {c: s.column(c) for c in ("radius", "Tp", "z", "vz")}
        log(f"[bench] fp64: {args.steps} steps in {t_fp64:.3f}s")
        ktimes64 = kernel_timing(s, init, MODE_BE, PREC_FP64, args.steps, args.warmup)
        log("[bench] kernel times (fp64): " + ", ".join(f"{k}={v[1] / v[0]:.3f}ms x{v[0]}" for k, v in ktimes64.items()))
        # the same steps with the exact-Jacobian FP64 solve (the mixed path's
        # solver in FP64): it tells the droplets where the reference's
        # FD-Jacobian solve failed (and reset the droplet) from precision effects
        run_steps(s, init, MODE_BE, PREC_FP64_AJ, args.steps, args.warmup)
        end_aj = s.column("radius")
        rr = np.abs(end_mixed["radius"] / end_fp64["radius"] - 1.0)
        r_aj = np.abs(end_mixed["radius"] / end_aj - 1.0)
        fd_fail = np.abs(end_fp64["radius"] / end_aj - 1.0) > 1e-5
        fp64 = {
            "value": n * args.steps / t_fp64, "ms_per_step": 1e3 * t_fp64 / args.steps,
            "newton_iters_per_solve": cnt_fp64["gn_iters"] / max(cnt_fp64["solves"], 1),
            "lm_fallbacks": cnt_fp64["num100"], "failed_solves": cnt_fp64["numimpos"],
            "kernel_ms": {k: v[1] / v[0] for k, v in ktimes64.items()},
        }
        if "step_be" in ktimes64:
            fp64["compute"] = pmc_compute("step_be", wl.name, ktimes64["step_be"][1] / ktimes64["step_be"][0],
                                          "step_be_fp64")
        accuracy = {"radius_max_rel_err_mixed_vs_fp64": float(rr.max()),
                    "radius_p99_rel_err_mixed_vs_fp64": float(np.quantile(rr, 0.99)),
                    "radius_max_rel_err_mixed_vs_fp64_fd_solve_converged": float(rr[~fd_fail].max()),
                    "droplets_where_fp64_fd_solve_failed": int(fd_fail.sum()),
                    "radius_max_rel_err_mixed_vs_fp64_exact_jacobian": float(r_aj.max()),
                    "Tp_max_abs_err_K": float(np.abs(end_mixed["Tp"] - end_fp64["Tp"]).max()),
                    "z_max_abs_err_m": float(np.abs(end_mixed["z"] - end_fp64["z"]).max()),
                    "vz_max_abs_err_m_s": float(np.abs(end_mixed["vz"] - end_fp64["vz"]).max())}

    # the drop-in path: one particle_update_BE (+ particle_coupling_update) call per step
    onestep = None
    if not args.no_onestep:
        onestep = {}
        for prec, nm in ((PREC_MIXED, "mixed"), (PREC_FP64, "fp64")):
            restore(s, init)
            s.particle_update_BE(it=2, precision=prec)
            s.sync()
            k1 = 10
            t0 = time.perf_counter()
            for k in range(k1):
                s.particle_update_BE(it=3 + k, precision=prec)
                s.particle_coupling_update(deterministic=0, out=False)
            s.sync()
            t1 = (time.perf_counter() - t0) / k1
            onestep[nm] = {"ms_per_call": 1e3 * t1, "value": n / t1}
            log(f"[bench] one-step calls ({nm}): {1e3 * t1:.3f} ms per particle_update_BE + coupling")

    # roofline of the dominant kernel (mixed precision): algorithmic bytes per
    # launch over its average HIP-event duration on the library's stream
    dom_name, (dom_launches, dom_ms) = max(ktimes.items(), key=lambda kv: kv[1][1])
    avg_ms = dom_ms / dom_launches
    per_kernel = roofline_of(ktimes, n, g, PREC_MIXED)
    traffic_alg = kernel_bytes(dom_name, n, g, PREC_MIXED)
    roofline = {"kernel": dom_name, "bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                "avg_launch_ms": avg_ms, "launches": dom_launches, "achieved": None, "frac": None,
                "algorithmic_bytes": traffic_alg}
    if traffic_alg:
        ach = traffic_alg / (avg_ms * 1e-3) / 1e9
        roofline.update({"achieved": ach, "frac": ach / HBM_PEAK_GBS})
    roofline["traffic"] = pmc_traffic(dom_name, wl.name)
    # the step kernel is also FP64-vector work: its FLOP rate against the
    # vector peaks, from the committed instruction-mix counters
    roofline["compute"] = pmc_compute(dom_name, wl.name, avg_ms) if dom_name == "step_be" else None
    if not args.no_copy:
        copy_gbps = s.copy_bandwidth(2.0, 10)
        log(f"[bench] device copy {copy_gbps:.0f} GB/s")
        roofline["copy_GBps"] = copy_gbps  # achievable: a plain 2 GiB device copy (16-byte accesses) on this GPU
        if roofline["achieved"]:
            roofline["frac_of_copy"] = roofline["achieved"] / copy_gbps
    roofline["kernels"] = per_kernel
    s.close()

    rk3 = None if args.no_rk3 else rk3_bench(args.seed)
    config5 = None if args.no_config5 else config5_bench(args.seed)

    cpu = None
    if not args.no_cpu:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(args.workload, args.seed, n)
        log(f"[bench] cpu baseline {cpu['value']:.3g} droplet-updates/s")

    serve = serve_rk3 = None
    if not args.no_serve:
        log("[bench] serving stream ...")
        serve = serve_bench(args.seed, args.serve_seconds, args.serve_rate)
        # the same stream with RK3 steps: dt = 1e-6 s keeps the explicit
        # stages stable down to the smallest droplets of ~2e8 drawn from these
        # ensembles (about 0.2 um, response time ~6e-7 s)
        log("[bench] serving stream, RK3 steps ...")
        serve_rk3 = serve_bench(args.seed, args.serve_seconds, args.serve_rate, mode=MODE_RK3, dt=1e-6)

    value = n * args.steps / t_mixed
    per_step_iters = (rows[:, 4] / np.maximum(rows[:, 6], 1)).tolist() if len(rows) else []
    hist = rows[-args.steps:, 8:8 + 64] if len(rows) else np.zeros((0, 64), dtype=np.int64)
    last = int(np.max(np.nonzero(hist.sum(axis=0))[0])) + 1 if hist.size and hist.sum() else 1
    out = {
        "metric": "droplet-updates/s per RK3 step, FP64 vs mixed; serve p50/p99 latency",
        "value": value,
        "unit": "droplet-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic (fields on the host, droplets drawn on the device)",
        "config": {"workload": wl.name, "droplets": n, "grid": f"{g.nnx}x{g.nny}x{g.nnz}",
                   "mode": "BE (ipart_method=2)", "precision": "mixed", "dt": wl.consts.dt,
                   "coupling": "on", "parallelism": "1 GPU"},
        "fp64": fp64,
        "mixed_over_fp64": (value / fp64["value"]) if fp64 else None,
        "accuracy": accuracy,
        "newton": {"iters_per_solve": cnt_mixed["gn_iters"] / max(cnt_mixed["solves"], 1),
                   "lm_fallbacks": cnt_mixed["num100"], "failed_solves": cnt_mixed["numimpos"],
                   "activations": cnt_mixed["actnum"], "deactivations": cnt_mixed["denum"],
                   "iters_per_solve_per_step": per_step_iters[-args.steps:],
                   "gn_iteration_histogram_per_step": {"columns": "solves whose Gauss-Newton ran k iterations, "
                                                                  "k = 0 (fell back to LM), 1, 2, ...",
                                                       "rows": hist[:, :last].tolist()}},
        "onestep_calls": onestep,
        "kernel_ms": {k: v[1] / v[0] for k, v in ktimes.items()},
        "roofline": roofline,
        "rk3": rk3,
        "config5": config5,
        "cpu_baseline": cpu,
        "serve": serve,
        "serve_rk3": serve_rk3,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
