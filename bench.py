#!/usr/bin/env python3
"""bench.py -- beamlet-dose throughput of the HIP CCCS engine on one MI355X, batched vs plain.

Headline workload (the batched beamlet configuration): 128^3 water phantom at 2.5 mm with an
ellipsoidal PTV (7.0 x 4.8 x 1.0 cm semi-axes), 50 non-coplanar seed-drawn beams, 40x40 fluence
maps of 5 mm beamlets open over the PTV projection (13689 beamlets, 90-458 per beam), 64 cones,
kernel_extent 1 cm, sparse columns thresholded at 1e-4 of each column's max.  One step = every
beamlet column of all 50 beams through the batched faithful pipeline (per-sample Siddon TERMA, as
the reference), columns built on the device; density and kernel tables resident in HBM.

The headline steps run with per-kernel profiling off.  A separate step with one cone direction in
flight (bl_streams = 1) and profiling on gives per-kernel times that add up to no more than the
step, and the dominant kernel's roofline from them: the sample stage priced by its algorithmic
bytes (16 B per stored REV sample: density and TERMA read, then written line-major) against HBM.

Beside the headline the JSON line carries: the plain path (the reference dataflow on HIP: one
batch at a time, per-cone launches over whole REV grids, dense per-beamlet unpack copied to the
host and sparsified there) on a bounded sample of beams, with the batched/plain ratio both
device-resident and like for like (columns delivered to a host sink on both sides); the batched
rate with columns copied to the host (PCIe-inclusive); p50/p99 latency of single-beam beamlet
requests; the plain full-beam path (200x200x160 thorax, 20 coplanar beams) with its cone
kernel's roofline; the serving stream; the 4pi precompute sample; and the CPU oracle timed on a
bounded sample (1 warm-up, median of 5; also the accuracy check).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
THRESH = 1e-4

# kernels whose algorithmic work is counted in FP32 operations (Siddon marches, line walks); the
# others are counted in bytes and priced against HBM
VALU_KERNELS = ("terma", "bl_siddon", "bl_walk", "plain_ray", "plain_conv")


def _percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def _pmc_traffic(kernel: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, if one exists."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def roofline_row(prof: dict, name: str, pmc_name: str | None = None):
    """Roofline position of kernel `name` from one-stream profiling rows: algorithmic bytes (or
    FP32 operations for the VALU-counted kernels) per launch over the average launch time."""
    row = prof[name]
    per_launch_ms = row["ms"] / max(1, row["launches"])
    if name in VALU_KERNELS:
        achieved = row["flops"] / (row["ms"] * 1e-3) / 1e12
        out = {"kernel": name, "bound": "valu_fp32", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
               "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS}
    else:
        achieved = row["bytes"] / (row["ms"] * 1e-3) / 1e9
        out = {"kernel": name, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": achieved / HBM_PEAK_GBS}
    out["traffic"] = _pmc_traffic(pmc_name or name)
    out["alg_bytes_per_launch"] = row["bytes"] / max(1, row["launches"])
    out["avg_launch_ms"] = per_launch_ms
    out["time_share"] = row["ms"] / sum(r["ms"] for r in prof.values())
    return out


def roofline(prof: dict):
    """Roofline row of the kernel with the largest one-stream time."""
    name = max(prof.items(), key=lambda kv: kv[1]["ms"])[0]
    return roofline_row(prof, name)


def _kernels(prof, steps: int = 1):
    return {k: {"launches": v["launches"], "ms_per_step": round(v["ms"] / steps, 3),
                "alg_GBps": round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1),
                "alg_TFLOPs": round(v["flops"] / max(v["ms"], 1e-9) / 1e9, 3)} for k, v in prof.items()}


def _cores():
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    return min(int(os.environ.get("OMP_NUM_THREADS", cores)), cores)


def cpu_baseline(w, eng, nbeamlets: int, runs: int = 5):
    """The oracle (the reference algorithm restated in C++/OpenMP) on a bounded sample: the first
    `nbeamlets` active beamlets of the first beam, dense columns; one warm-up run, then the median
    of `runs` timed runs.  The columns are compared with the GPU's."""
    from endtoend_amd import abi
    from oracle import oracle as orc  # test infrastructure: the checker, never the measured path
    from tests.dose_metrics import max_err_pct
    f = w.fmaps[0].copy()
    act = np.flatnonzero(f.ravel() > 0)
    keep = np.zeros(f.size, bool)
    keep[act[:nbeamlets]] = True
    f.ravel()[~keep] = 0.0
    ref, _ = orc.beamlet_dense(w.spec, w.beams[0], f)  # warm-up
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        orc.beamlet_dense(w.spec, w.beams[0], f)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    got = eng.beamlet_dense(w.beams[0], f, abi.CCCS_BEAMLET_FAITHFUL)
    err = max(max_err_pct(r, g) for r, g in zip(ref, got))
    n = ref.shape[0]
    return {"value": n / dt, "unit": "beamlet-doses/s", "cores": _cores(), "kind": "port",
            "sample": f"{n} beamlets of beam 0 of the batched workload, dense columns; 1 warm-up, median of "
                      f"{runs} runs: {dt:.2f} s (runs {min(times):.2f}-{max(times):.2f} s)"}, err


BEAM_KNOBS = ("cone_ramp_min", "cone_class_len0", "cone_class_len1", "cone_class_len2", "cone_class_threads1",
              "cone_class_threads2", "cone_class_threads3", "pvcs_pass_cones")
BEAMLET_KNOBS = ("bl_sample_waves", "bl_walk_waves", "bl_walk_split", "bl_streams", "bl_job_gb")


def plain_beam_path(steps: int = 2, autotune: bool = False):
    """The plain full-beam path on the 20-beam thorax plan: beamlet-doses/s and per-beam latency."""
    from endtoend_amd import abi
    from endtoend_amd import workloads as W
    from endtoend_amd.engine import Engine
    w = W.cfg2()
    nv = w.spec.density.size
    fl = np.concatenate([f.ravel() for f in w.fmaps]).astype(np.float32)
    offs = np.cumsum([0] + [f.size for f in w.fmaps])[:-1]
    eng = Engine(w.spec)
    # the cone-kernel schedule: the library's defaults (chosen by its autotuner on this kind of
    # plan), or a fresh autotune on two of the plan's beams
    tune = (eng.autotune(w.beams[:2], w.fmaps[:2], kind=abi.CCCS_TUNE_BEAM)["knobs"] if autotune else
            {k: eng.get_schedule(k) for k in BEAM_KNOBS})
    d_fl = eng.dev_alloc(fl.nbytes)
    eng.dev_upload(d_fl, fl)
    d_dose = eng.dev_alloc(nv * 4)
    eng.beam_dose_device(w.beams, d_fl, d_dose)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.dev_zero(d_dose, nv * 4)
        eng.beam_dose_device(w.beams, d_fl, d_dose)
    eng.sync()
    dt = (time.perf_counter() - t0) / steps
    lat = []
    for i, b in enumerate(w.beams):
        eng.sync()
        ta = time.perf_counter()
        eng.beam_dose_device([b], d_fl + 4 * int(offs[i]), d_dose)
        eng.sync()
        lat.append(1e3 * (time.perf_counter() - ta))
    # per-kernel times of one plan (one stream) and the cone kernel's roofline: 12 B per stored
    # REV sample (density and TERMA in, cone dose out)
    eng.profile(True)
    eng.profile_reset()
    eng.dev_zero(d_dose, nv * 4)
    eng.beam_dose_device(w.beams, d_fl, d_dose)
    eng.sync()
    prof = eng.profile_read()
    eng.profile(False)
    eng.dev_free(d_fl)
    eng.dev_free(d_dose)
    eng.close()
    return {"workload": "200x200x160 thorax @ 2 mm, 20 coplanar beams, 40x40 x 5 mm U[0,1] fluence masked "
                        "to the PTV, full-beam dose", "value": w.active_beamlets / dt, "unit": "beamlet-doses/s",
            "ms_per_plan": 1e3 * dt, "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99)},
            "kernels": _kernels(prof), "roofline_cone": roofline_row(prof, "cone", "cone_64"), "schedule": tune}


def full_4pi(stride: int = 12, rate: float = 20.0, requests: bool = True):
    """Full 4pi precompute on a bounded sample: every stride-th beam of the 1162-beam 4pi list on
    the batched workload's phantom, 64x64 maps of 2.5 mm beamlets over the PTV projection, 96 cones.
    First the batched rate with the columns left in HBM; then the same beams as beamlet requests
    through the serving loop (columns to the host, native counting sink) while single-beam dose
    requests arrive at `rate`/s on the same resident state (Poisson; 4pi-list beams, 20^2-64^2 x
    5 mm U[0,1] maps): the precompute's rate under that load and the requests' latency (skipped
    with requests=False)."""
    import threading

    from endtoend_amd import abi
    from endtoend_amd import workloads as W
    from endtoend_amd.engine import Engine, Server
    w = W.cfg5(stride=stride)
    nbl = w.active_beamlets
    FA = abi.CCCS_BEAMLET_FAITHFUL
    with Engine(w.spec) as eng:
        eng.beamlet_dose(w.beams[:2], w.fmaps[:2], threshold=THRESH, mode=FA, keep=None)  # warm
        eng.sync()
        t0 = time.perf_counter()
        eng.beamlet_dose(w.beams, w.fmaps, threshold=THRESH, mode=FA, keep=None)
        eng.sync()
        span_batched = time.perf_counter() - t0
        batched = nbl / span_batched
    which = "all of the 1162-beam 4pi list" if stride == 1 else f"every {stride}th of the 1162-beam 4pi list"
    desc = (f"{len(w.beams)} beams ({which}), 128^3 water @ 2.5 mm + "
            "ellipsoidal PTV, 64x64 x 2.5 mm fluence open over the PTV projection, 96 cones, sparse "
            "columns at 1e-4 of column max, faithful TERMA")
    out = {"workload": desc, "beamlets": nbl,
           "batched": {"value": batched, "unit": "beamlet-doses/s", "seconds": span_batched}}
    # the plain path (reference dataflow) on two beams of this sample, and the batched path on the
    # same two beams, columns delivered to a host sink on both sides
    with Engine(w.spec) as eng:
        pb = [0, len(w.beams) // 2]
        pbl = sum(int((w.fmaps[i] > 0).sum()) for i in pb)
        sb, sf = [w.beams[i] for i in pb], [w.fmaps[i] for i in pb]
        eng.beamlet_dose(sb[:1], sf[:1], threshold=THRESH, mode=abi.CCCS_BEAMLET_PLAIN, keep=False)  # warm
        t0 = time.perf_counter()
        eng.beamlet_dose(sb, sf, threshold=THRESH, mode=abi.CCCS_BEAMLET_PLAIN, keep=False)
        eng.sync()
        plain = pbl / (time.perf_counter() - t0)
        t0 = time.perf_counter()
        eng.beamlet_dose(sb, sf, threshold=THRESH, mode=FA, keep=False)
        eng.sync()
        same = pbl / (time.perf_counter() - t0)
    out["plain"] = {"value": plain, "unit": "beamlet-doses/s", "sample": f"beams {pb} ({pbl} beamlets)",
                    "batched_same_sample_pcie_inclusive": same, "batched_vs_plain_like_for_like": same / plain}
    if not requests:
        return out
    rng = np.random.default_rng(W.SEED_ARRIVALS)
    reqs = []
    for i in rng.integers(0, len(w.beams), 400):
        b = w.beams[int(i)].copy()
        n = int(rng.integers(20, 65))
        b.fmap_size[:] = (n, n)
        b.beamlet_size[:] = (0.5, 0.5)
        reqs.append((b, rng.uniform(0.0, 1.0, (n, n)).astype(np.float32)))
    gaps = rng.exponential(1.0 / rate, len(reqs))
    with Server(w.spec, slots=2) as srv:
        srv.wait(srv.submit(*reqs[0]))                                       # warm both kinds
        srv.wait(srv.submit(w.beams[0], w.fmaps[0], abi.CCCS_REQ_BEAMLET, THRESH), abi.CCCS_REQ_BEAMLET,
                 counts_only=True)
        lat, done = [], threading.Event()

        def requests():
            ids = []
            for k in range(1, len(reqs)):
                if done.wait(gaps[k]):
                    break
                ids.append(srv.submit(*reqs[k]))
            lat.extend(srv.wait(r)[1] for r in ids)

        th = threading.Thread(target=requests)
        t0 = time.perf_counter()
        th.start()
        pre = [srv.submit(b, f, abi.CCCS_REQ_BEAMLET, THRESH) for b, f in zip(w.beams, w.fmaps)]
        cols = sum(srv.wait(r, abi.CCCS_REQ_BEAMLET, counts_only=True)[0]["columns"] for r in pre)
        span = time.perf_counter() - t0
        done.set()
        th.join()
    out["with_requests"] = {"precompute_value": cols / span, "unit": "beamlet-doses/s", "columns": cols,
                            "request_rate": rate, "requests": len(lat),
                            "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99),
                                           "max": max(lat) if lat else None}}
    return out


def serving_breakdown(nreq: int = 24, cached: bool = True):
    """Where a serving request's time goes: `nreq` requests of the serving stream (REV grids within
    the limit) on one context, fluence already in HBM, with the angle cache of their sources (or
    without); per-kernel ms per request from the library's event timers, the device time per
    request, and the host-inclusive time of cccs_beam_dose (dose volume back in host memory)."""
    from endtoend_amd import workloads as W
    from endtoend_amd.engine import Engine
    w = W.cfg4(nreq=300)
    with Engine(w.spec) as eng:
        sel = []
        for i, (b, f) in enumerate(zip(w.beams, w.fmaps)):
            try:
                eng.geometry(b, f)
            except Exception:   # REV grid beyond the limit (the reference's ArraySizeError)
                continue
            sel.append(i)
            if len(sel) == nreq:
                break
        cache = eng.angle_cache([w.beams[i] for i in sel]) if cached else None
        if cache is not None:
            eng.use_angle_cache(cache)
        nf = max(int(w.fmaps[i].size) for i in sel)
        dfl = eng.dev_alloc(4 * nf)
        dd = eng.dev_alloc(4 * int(w.spec.density.size))
        for i in sel[:3]:
            eng.beam_dose([w.beams[i]], [w.fmaps[i]])
        dev = []
        for i in sel:
            eng.dev_upload(dfl, np.ascontiguousarray(w.fmaps[i], np.float32))
            eng.sync()
            t0 = time.perf_counter()
            eng.beam_dose_device([w.beams[i]], dfl, dd)
            eng.sync()
            dev.append(1e3 * (time.perf_counter() - t0))
        eng.profile(True)
        eng.profile_reset()
        for i in sel:
            eng.dev_upload(dfl, np.ascontiguousarray(w.fmaps[i], np.float32))
            eng.beam_dose_device([w.beams[i]], dfl, dd)
        eng.sync()
        prof = eng.profile_read()
        eng.profile(False)
        host = []
        for i in sel:
            t0 = time.perf_counter()
            eng.beam_dose([w.beams[i]], [w.fmaps[i]])
            host.append(1e3 * (time.perf_counter() - t0))
        eng.dev_free(dfl)
        eng.dev_free(dd)
        eng.use_angle_cache(None)
        if cache is not None:
            cache.close()
    return {"requests": len(sel), "angle_cache": cached,
            "kernel_ms_per_request": {k: v["ms"] / len(sel) for k, v in prof.items() if v["launches"]},
            "device_ms": {"p50": _percentile(dev, 50), "max": max(dev)},
            "host_inclusive_ms": {"p50": _percentile(host, 50), "max": max(host)}}


def serve_profile(nreq: int):
    """Profiling runs: `nreq` serving requests on one context with the angle cache, fluence in HBM."""
    print(json.dumps(serving_breakdown(nreq=nreq, cached=True)))


def serving(slots: int = 2, nreq: int = 300, rate: float = 200.0, preload: bool = True, sweep_only: bool = False):
    """Annealing-style request stream through the serving loop: Poisson arrivals of single-beam
    dose requests on the 256x256x180 thorax phantom (beams from the 1162-beam 4pi list, 20^2-64^2
    fluence maps).  With `preload` the server first keeps the unit TERMA of all 1162 candidate
    sources resident (angle cache).  Latency is submit-to-completion (result in host memory),
    measured inside the server.  Reported: the closed-loop capacity; one request at a time (the
    optimiser's sequential proposals); the stream at the offered rate; the same stream at 80 % of
    capacity; and a sweep of rates for the highest one whose p99 stays under 50 ms."""
    import threading

    from endtoend_amd import abi
    from endtoend_amd import engine as E
    from endtoend_amd import workloads as W
    from endtoend_amd.engine import Server
    w = W.cfg4(nreq=nreq, rate=rate)
    out = {"workload": "256x256x180 thorax @ 2 mm, single-beam dose requests, beams from the 1162-beam 4pi list, "
                       "20^2-64^2 x 5 mm U[0,1] fluence, Poisson arrivals",
           "slots": slots, "seed": w.seed}
    with Server(w.spec, slots=slots) as srv:
        if preload:
            cands = E.load_beam_list(os.path.join(W.DATA, "beamlist_4pi.txt"))
            t0 = time.perf_counter()
            n = srv.preload_angles(cands, max_bytes=96 << 30)
            out["angle_cache"] = {"sources": n, "GB": 4.0 * n * w.spec.density.size / 1e9,
                                  "build_s": time.perf_counter() - t0}
        for i in range(2 * slots):                       # warm every slot
            srv.wait(srv.submit(w.beams[i], w.fmaps[i]))
        # capacity: requests submitted at once (closed loop, the queue never empties)
        ncap = min(nreq, 60)
        t1 = time.perf_counter()
        cap_ids = [srv.submit(w.beams[i], w.fmaps[i]) for i in range(ncap)]
        ok = 0
        for r in cap_ids:
            try:
                srv.wait(r)
                ok += 1
            except abi.CCCSError:
                pass
        capacity = ok / (time.perf_counter() - t1)
        # the optimiser's own pattern: one proposal at a time, each submitted when the previous
        # result is back in host memory
        seq = []
        for i in range(ncap, min(nreq, ncap + 60)):
            try:
                seq.append(srv.wait(srv.submit(w.beams[i], w.fmaps[i]))[1])
            except abi.CCCSError:
                pass
        sequential = {"requests": len(seq), "latency_ms": {"p50": _percentile(seq, 50), "p99": _percentile(seq, 99),
                                                           "max": max(seq)}}

        def stream(arrivals):
            ids, lat, bs, failed = [], [], [], []
            ready = threading.Semaphore(0)

            def consume():
                for k in range(len(arrivals)):
                    ready.acquire()
                    try:
                        _, l, b = srv.wait(ids[k])
                    except abi.CCCSError as e:   # a REV grid beyond the limit (the reference's ArraySizeError)
                        failed.append(e.code)
                        continue
                    lat.append(l)
                    bs.append(b)

            th = threading.Thread(target=consume)
            th.start()
            t0 = time.perf_counter()
            for i in range(len(arrivals)):
                dt = arrivals[i] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids.append(srv.submit(w.beams[i % nreq], w.fmaps[i % nreq]))
                ready.release()
            th.join()
            span = time.perf_counter() - t0
            return {"offered_rate": len(arrivals) / float(arrivals[-1]), "requests": len(arrivals),
                    "failed": len(failed), "completed_per_s": len(lat) / span,
                    "latency_ms": {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99), "max": max(lat)},
                    "mean_batch": float(np.mean(bs))}

        offered = stream(w.arrivals)
        sustain = 0.8 * capacity
        n2 = min(nreq, max(60, int(sustain * 8)))
        sus = stream(w.arrivals[:n2] * (rate / sustain))
        # rate sweep: the same arrival pattern scaled to each rate, about 6 s of arrivals each
        sweep, best = [], None
        for frac in (0.2, 0.3, 0.4, 0.5, 0.65, 0.8):
            r = frac * capacity
            n3 = min(nreq, max(60, int(r * 6)))
            res = stream(w.arrivals[:n3] * (rate / r))
            sweep.append({"rate": res["offered_rate"], "p50_ms": res["latency_ms"]["p50"],
                          "p99_ms": res["latency_ms"]["p99"], "requests": n3})
            if res["latency_ms"]["p99"] < 50.0:
                best = res["offered_rate"] if best is None else max(best, res["offered_rate"])
    out.update({"capacity_per_s": capacity, "sequential": sequential, "at_offered_rate": offered,
                "at_80pct_capacity": sus, "rate_sweep": sweep, "max_rate_p99_under_50ms": best})
    return out


def copy_calibration(mib: int, reps: int = 5):
    """Streams mib MiB through the library's copy kernel `reps` times (reads = writes = mib MiB per
    launch), so rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE can be scaled to bytes."""
    from endtoend_amd import workloads as W
    from endtoend_amd.engine import Engine
    eng = Engine(W.small().spec)
    n = mib << 20
    a, b = eng.dev_alloc(n), eng.dev_alloc(n)
    eng.dev_zero(a, n)
    for _ in range(reps):
        eng.dev_copy(b, a, n)
    eng.sync()
    eng.dev_free(a)
    eng.dev_free(b)
    eng.close()
    print(json.dumps({"copy_calibration_mib": mib, "launches": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-beams", type=int, default=2, help="beams in the plain-path sample")
    ap.add_argument("--latency-beams", type=int, default=20, help="single-beam requests timed for p50/p99")
    ap.add_argument("--cpu-beamlets", type=int, default=60,
                    help="beamlets in the oracle sample; 1 warm-up + 5 timed runs, about 20 s of host work (0 = skip)")
    ap.add_argument("--no-extras", action="store_true", help="headline only (profiling runs)")
    ap.add_argument("--plain-beam-only", action="store_true",
                    help="only the plain full-beam path on the thorax plan (profiling runs)")
    ap.add_argument("--serving-only", action="store_true", help="only the serving stream")
    ap.add_argument("--serve-profile", type=int, default=0,
                    help="only this many serving requests on one context with the angle cache (profiling runs)")
    ap.add_argument("--full-4pi-only", action="store_true",
                    help="only the 4pi precompute sample with concurrent requests (profiling runs)")
    ap.add_argument("--full-4pi-stride", type=int, default=12,
                    help="with --full-4pi-only: every n-th beam of the 1162-beam list (1 = the whole list)")
    ap.add_argument("--full-4pi-no-requests", action="store_true",
                    help="with --full-4pi-only: the batched precompute alone")
    ap.add_argument("--autotune", action="store_true",
                    help="run the schedule autotuner on a sample of each workload before timing it")
    ap.add_argument("--bl-streams", type=int, default=0,
                    help="cone directions in flight for the whole run (profiling runs: 1); 0 = the library default")
    ap.add_argument("--bl-sample-waves", type=int, default=0,
                    help="sample-stage waves for the whole run (profiling runs: 262144, as in the one-stream "
                         "pass); 0 = the library default")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only run the streaming copy kernel on this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (beams shard across GPUs "
              f"with no exchange in beamlet mode)", file=sys.stderr)
        sys.exit(2)
    if args.copy_mib > 0:
        copy_calibration(args.copy_mib)
        return
    if args.plain_beam_only:
        print(json.dumps(plain_beam_path(steps=args.steps, autotune=args.autotune)))
        return
    if args.serving_only:
        print(json.dumps(serving()))
        return
    if args.serve_profile > 0:
        serve_profile(args.serve_profile)
        return
    if args.full_4pi_only:
        print(json.dumps(full_4pi(stride=args.full_4pi_stride, requests=not args.full_4pi_no_requests)))
        return

    from endtoend_amd import abi
    from endtoend_amd import workloads as W
    from endtoend_amd.engine import Engine

    w = W.cfg3()
    eng = Engine(w.spec)
    nbl = w.active_beamlets
    FA = abi.CCCS_BEAMLET_FAITHFUL
    # launch schedule: the library's defaults (chosen by its autotuner on this workload), or with
    # --autotune a fresh tuning on four of the workload's beams before the warmup
    if args.autotune:
        schedule = eng.autotune(w.beams[:4], w.fmaps[:4], kind=abi.CCCS_TUNE_BEAMLET)["knobs"]
    else:
        if args.bl_streams > 0:
            eng.set_schedule("bl_streams", args.bl_streams)
        if args.bl_sample_waves > 0:
            eng.set_schedule("bl_sample_waves", args.bl_sample_waves)
        schedule = {k: eng.get_schedule(k) for k in BEAMLET_KNOBS}

    def step():
        eng.beamlet_dose(w.beams, w.fmaps, threshold=THRESH, mode=FA, keep=None)

    for _ in range(args.warmup):
        step()
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    eng.sync()
    t1 = time.perf_counter()
    dt = t1 - t0
    # per-kernel times: one step with one cone direction in flight, so that no two kernels share
    # the device and the kernel times add up to no more than that step.  Alone on the device the
    # sample stage runs at its standalone best, 4 waves per SIMD (262144 waves); the headline
    # keeps it at 2 waves per SIMD so that the other directions' stages run beside it.
    nstreams = eng.get_schedule("bl_streams")
    swaves = eng.get_schedule("bl_sample_waves")
    alone = {"bl_streams": 1, "bl_sample_waves": 262144}
    for k, v in alone.items():
        eng.set_schedule(k, v)
    step()
    eng.sync()
    eng.profile(True)
    eng.profile_reset()
    t2 = time.perf_counter()
    step()
    eng.sync()
    one_stream_ms = 1e3 * (time.perf_counter() - t2)
    prof = eng.profile_read()
    eng.profile(False)
    eng.set_schedule("bl_streams", nstreams)
    eng.set_schedule("bl_sample_waves", swaves)
    ms_per_step = 1e3 * dt / args.steps
    value = args.steps * nbl / dt

    out = {
        "metric": "beamlet-doses/sec at <=0.5% Dmax err, batched vs plain; p50/p99 beam latency",
        "value": value,
        "unit": "beamlet-doses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched beamlet: 128^3 water @ 2.5 mm + ellipsoidal PTV (7.0x4.8x1.0 cm), 50 "
                               "non-coplanar beams, 40x40 x 5 mm fluence open over the PTV projection, 64 cones, "
                               "kernel_extent 1 cm, sparse columns at 1e-4 of column max, faithful per-sample Siddon "
                               "TERMA, columns built in HBM",
                   "path": "batched", "beamlet_doses_per_step": nbl, "beams_per_step": len(w.beams), "seed": w.seed,
                   "schedule": schedule},
        "roofline": roofline(prof),
        "kernels_one_stream": {"ms_per_step": one_stream_ms, "schedule": alone, "kernels": _kernels(prof)},
    }
    if not args.no_extras:
        # columns copied to the host through the native counting sink (PCIe-inclusive)
        t0 = time.perf_counter()
        tot = eng.beamlet_dose(w.beams, w.fmaps, threshold=THRESH, mode=FA, keep=False)
        eng.sync()
        out["pcie_inclusive"] = {"value": nbl / (time.perf_counter() - t0), "unit": "beamlet-doses/s",
                                 "columns": tot["columns"], "nnz": tot["nnz"]}
        # plain path (reference dataflow) on a bounded sample of beams
        pb = list(range(min(args.plain_beams, len(w.beams))))
        pbl = sum(int((w.fmaps[i] > 0).sum()) for i in pb)
        eng.beamlet_dose([w.beams[i] for i in pb], [w.fmaps[i] for i in pb], threshold=THRESH,
                         mode=abi.CCCS_BEAMLET_PLAIN, keep=False)  # warm
        t0 = time.perf_counter()
        eng.beamlet_dose([w.beams[i] for i in pb], [w.fmaps[i] for i in pb], threshold=THRESH,
                         mode=abi.CCCS_BEAMLET_PLAIN, keep=False)
        eng.sync()
        plain = pbl / (time.perf_counter() - t0)
        # batched on the same sample, device-resident like the headline
        t0 = time.perf_counter()
        eng.beamlet_dose([w.beams[i] for i in pb], [w.fmaps[i] for i in pb], threshold=THRESH, mode=FA, keep=None)
        eng.sync()
        same = pbl / (time.perf_counter() - t0)
        out["plain"] = {"value": plain, "unit": "beamlet-doses/s", "sample": f"beams {pb} ({pbl} beamlets)",
                        "batched_same_sample": same, "batched_vs_plain": value / plain,
                        # both sides deliver every column to a host sink (the plain path cannot
                        # leave its columns on the device)
                        "batched_vs_plain_like_for_like": out["pcie_inclusive"]["value"] / plain}
        # single-beam request latency (one beam per call, columns to the host)
        lat = []
        for i in range(min(args.latency_beams, len(w.beams))):
            eng.sync()
            ta = time.perf_counter()
            eng.beamlet_dose([w.beams[i]], [w.fmaps[i]], threshold=THRESH, mode=FA, keep=False)
            eng.sync()
            lat.append(1e3 * (time.perf_counter() - ta))
        out["latency_ms"] = {"p50": _percentile(lat, 50), "p99": _percentile(lat, 99),
                             "per": "single-beam beamlet request"}
        cpu, err = (None, None)
        if args.cpu_beamlets > 0:
            cpu, err = cpu_baseline(w, eng, args.cpu_beamlets)
        out["max_err_pct_dmax"] = err
        out["cpu_baseline"] = cpu
        eng.close()
        out["plain_beam_path"] = plain_beam_path(autotune=args.autotune)
        out["serving"] = serving()
        out["serving"]["breakdown"] = serving_breakdown()
        out["full_4pi"] = full_4pi()
    else:
        eng.close()
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
