"""Benchmark of the Wyner-Ziv decoding hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads
  wz_cif_q8     (default) the plain drop-in configuration at its headline cell:
                CIF 352x288, 33 frames, GOP 2, Qindex 8 (16 WZ frames, 63
                bitplanes each, N = 6336), synthetic sequence as described in
                dvc_amd.workload.  One step decodes `--streams` such sequences
                (16 WZ frames each, 64 sequences = 1024 frames by default) through
                one long-lived continuous-batching engine (2 lanes that step
                independently) with the speculative persistent LDPCA launch;
                value = WZ frames/s, timed without per-kernel events (steady
                steps replay captured HIP graphs), with per-frame latency.
                Beside it: one more step with HIP events around every launch
                (kernel durations, the roofline), the plain path (one frame in
                flight through a long-lived engine, sequential rate search) on
                --plain-frames frames, the same workload with ParallelMode 1
                BP semantics and with host buffers (PCIe-inclusive), a bounded
                run of the serving configuration, and the VALU issue-rate
                calibration that prices the BP roofline.
  ldpca_batch   512 seeded bitplanes (384 x N=6336, 128 x N=1584) decoded in
                one speculative persistent launch per step; one step = the
                whole set.  value = BP edge-updates/s, counting the
                iterations the sequential rate search runs (an edge-update is
                one edge's check-to-variable plus variable-to-check update in
                one iteration; E = 3N per code).

  serve, si_batch, hbm_copy, valu_peak: the serving loop, batched side
                information, a device copy (the HBM traffic calibration) and
                the VALU issue-rate calibration.

Prints one JSON line (see the end of main()).  The CPU baseline leg runs the
oracle (ParallelMode 1 over the host cores, and ParallelMode 0 on one core)
on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "WZ frames/sec and BP edge-updates/s (batched vs plain, 1 GPU); p50/p99 latency"
HBM_PEAK_GBS = 8000.0
LDS_PEAK_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B/clk x 2.4 GHz


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bench_ldpca(args):
    from dvc_amd import _lib, workload as W
    t0 = time.time()
    planes = W.ldpca_bitplane_set(n_cif=384, n_qcif=128, seed=2)
    log(f"[bench] generated {len(planes)} bitplanes in {time.time() - t0:.1f}s")
    lrs = [p["LR"] for p in planes]
    accs = [p["acc"] for p in planes]
    crcs = [p["crc"] for p in planes]
    ctx = _lib.Context(device=0, max_iter=args.max_iter, bp_threads=args.bp_threads)
    for i in range(args.warmup):
        decs, sts = ctx.ldpca_decode_batch(lrs, accs, crcs)
        log(f"[bench] warmup {i} done")
    bad = sum(int(not np.array_equal(d, p["bits"])) for d, p in zip(decs, planes)) if args.warmup else 0
    ctx.profile(True)
    times = []
    for i in range(args.steps):
        t = time.perf_counter()
        decs, sts = ctx.ldpca_decode_batch(lrs, accs, crcs)
        times.append(time.perf_counter() - t)
        log(f"[bench] step {i}: {times[-1] * 1e3:.1f} ms")
    bp_ms, bp_n = ctx.profile_get("bp_6336")
    bpq_ms, bpq_n = ctx.profile_get("bp_1584")
    ce_ms, ce_n = ctx.profile_get("ce_llr")
    ctx.profile(False)
    bad += sum(int(not np.array_equal(d, p["bits"])) for d, p in zip(decs, planes))
    edge_updates = sum(st["iterations"] * 3 * p["N"] for st, p in zip(sts, planes))
    t_step = float(np.mean(times))
    value = edge_updates / t_step

    # Roofline of the dominant kernel (bp_6336): algorithmic HBM bytes per
    # launch = per accepted (bitplane, code) item 4N (LLR) + ceil(m/8)
    # (syndrome) + ceil(N/8) (decisions) + 12N (the code's column table: one
    # word per edge), summed over every code the sequential search runs.
    ladders = {n: W.ladder(n) for n in (1584, 6336)}
    m_of = {n: (n // 66) * ladders[n].bits_per_period for n in ladders}
    hbm_bytes = 0
    lds_bytes = 0
    for st, p in zip(sts, planes):
        if p["N"] != 6336:
            continue
        n = p["N"]
        for lvl in range(st["start_level"], st["level"] + 1):
            if lvl == 64:
                continue
            hbm_bytes += 4 * n + (int(m_of[n][lvl]) + 7) // 8 + (n + 7) // 8 + 12 * n
        lds_bytes += 16 * st["iterations"] * 3 * n
    bp_avg_ms = bp_ms / max(bp_n, 1)
    per_step = max(bp_n // max(args.steps, 1), 1)  # bp_6336 launches per step
    res = {
        "metric": METRIC,
        "value": value,
        "unit": "BP edge-updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Bernoulli bitplanes, Laplacian soft inputs)",
        "config": {"workload": "ldpca_batch: 512 bitplanes (384 N=6336 + 128 N=1584), all codes speculative, "
                               f"max_iter={args.max_iter}", "parallelism": "single GPU"},
        "roofline": bp_roofline(bp_avg_ms, hbm_bytes / per_step, lds_bytes / 16.0 / per_step, "ldpca_batch",
                                valu_calibration(ctx)),
        "kernels_ms": {"bp_6336": bp_ms / max(bp_n, 1), "bp_1584": bpq_ms / max(bpq_n, 1),
                       "ce_llr": ce_ms / max(ce_n, 1)},
        "wrong_bitplanes": bad,
        "edge_updates_per_step": edge_updates,
        # BP iterations the batch ran (speculative and cancelled codes included)
        # over those the sequential search needs
        "bp_iterations_run_over_needed": sum(st["iterations_run"] for st in sts) / max(1, sum(st["iterations"] for st in sts)),
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_ldpca(planes, args)
    return res


def make_wz_inputs(streams, frames=33, q=8, gop=2, seed0=1100):
    from dvc_amd import workload as W
    jobs = []
    for st in range(streams):
        vid = W.synthetic_video(352, 288, frames, 6, seed0 + st)
        for k in range(0, frames - 1, gop):
            cur = k + gop // 2
            wz = W.encode_wz_frame(vid[cur], q)
            jobs.append(dict(past=vid[k], fut=vid[k + gop], orig=vid[cur], wz=wz, idx=(k, cur, k + gop), seq=st))
    return jobs


class DeviceInputs:
    """The workload's inputs made resident in HBM before the timed region:
    each sequence's key frames once, every frame's accumulated syndromes, and
    an output frame per WZ frame (library device buffers)."""

    def __init__(self, ctx, jobs):
        from dvc_amd import _lib
        self.keys = {}
        self.acc, self.rec = [], []
        for j in jobs:
            for f, img in ((j["idx"][0], j["past"]), (j["idx"][2], j["fut"])):
                k = ((j["seq"] + 1) << 32) | f
                if k not in self.keys:
                    self.keys[k] = _lib.DeviceBuffer.from_array(ctx, img)
            self.acc.append(_lib.DeviceBuffer.from_array(ctx, np.asarray(j["wz"].acc, np.uint8)))
            self.rec.append(_lib.DeviceBuffer(ctx, j["past"].size))
        self.shapes = [j["past"].shape for j in jobs]

    def submit_all(self, eng, jobs, gen=0, index=None):
        """Submits the frames (jobs[k] is the workload's frame index[k], all
        of them by default); key frames go in under keys of generation `gen`,
        so a later generation brings them in (and filters them) again rather
        than finding them resident."""
        index = range(len(jobs)) if index is None else index
        tickets = []
        for j, i in zip(jobs, index):
            acc, rec = self.acc[i], self.rec[i]
            kp = ((j["seq"] + 1) << 32) | j["idx"][0]
            kf = ((j["seq"] + 1) << 32) | j["idx"][2]
            h, w = j["past"].shape
            tickets.append(eng.submit_device(w, h, j["wz"], self.keys[kp].ptr, self.keys[kf].ptr,
                                             acc.ptr, rec.ptr, idx=j["idx"], past_key=(gen << 48) | kp,
                                             fut_key=(gen << 48) | kf))
        return tickets

    def release_keys(self, eng, jobs, gen):
        for k in {(gen << 48) | ((j["seq"] + 1) << 32) | f for j in jobs for f in (j["idx"][0], j["idx"][2])}:
            eng.release_key(k)

    def recons(self, i):
        return self.rec[i].download(self.shapes[i], np.uint8)


def run_engine(eng, tickets, t_submit=None, lat=None):
    """Steps the engine until every ticket is done; returns the frames'
    stats.  With t_submit (perf_counter time of the submissions) and a list
    `lat`, appends each frame's latency from submission to completion."""
    pending = set(tickets)
    while pending:
        eng.step()
        now = time.perf_counter()
        for t in list(pending):
            if eng.done(t):
                pending.discard(t)
                if lat is not None:
                    lat.append(now - t_submit)
    return [eng.result(t)[1] for t in tickets]


def psnr(a, b):
    """frame.cpp:115-124."""
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return 0.0 if mse == 0.0 else 10.0 * np.log10(255.0 * 255.0 / mse)


KERNELS = ["create_si", "keyframe_filters", "noise_param", "oct", "smf_init", "noise_dist", "condbit", "ce_llr",
           "bp_6336", "invh", "put_bitplane", "reconstruct", "inverse_transform", "update_si", "update_smf",
           "update_si_ml"]


def bench_wz(args):
    from dvc_amd import _lib
    t0 = time.time()
    jobs = make_wz_inputs(args.streams)
    log(f"[bench] encoded {len(jobs)} CIF Q8 WZ frames in {time.time() - t0:.1f}s")
    ctx = _lib.Context(device=0, max_iter=args.max_iter, bp_threads=args.bp_threads)
    tuning = ctx.autotune() if args.autotune else ctx.tuning()
    log(f"[bench] BP tuning {tuning}")
    dev = DeviceInputs(ctx, jobs)
    # one long-lived engine, as a server keeps it: warmup steps leave its
    # device-block pool and captured launch graphs in place; every step
    # brings its key frames in afresh (new keys) and decodes every frame
    eng = _lib.Engine(ctx, max_inflight=len(jobs), lanes=args.lanes, device_buffers=True, lane_steps=1 << 20)
    gen = 0
    for i in range(args.warmup):
        run_engine(eng, dev.submit_all(eng, jobs, gen))
        dev.release_keys(eng, jobs, gen)
        gen += 1
        log(f"[bench] warmup {i} done")
    # timed steps: no per-kernel timing (graph replay of steady steps on)
    times, frame_lat = [], []
    for i in range(args.steps):
        t = time.perf_counter()
        stats = run_engine(eng, dev.submit_all(eng, jobs, gen), t, frame_lat)
        times.append(time.perf_counter() - t)
        dev.release_keys(eng, jobs, gen)  # a server releases a stream's key frames once no frame needs them
        gen += 1
        log(f"[bench] step {i}: {times[-1] * 1e3:.1f} ms for {len(jobs)} frames")
    outs = [(dev.recons(i), dict(st, psnr=psnr(dev.recons(i), j["orig"]))) for i, (j, st) in enumerate(zip(jobs, stats))]
    # kernel durations: the same step again with HIP events around every
    # launch on the stream it runs on (this turns graph replay off)
    kern = {}
    ctx.profile(True)
    for i in range(args.kernel_steps):
        t = time.perf_counter()
        run_engine(eng, dev.submit_all(eng, jobs, gen))
        dev.release_keys(eng, jobs, gen)
        gen += 1
        log(f"[bench] kernel-timing step {i}: {(time.perf_counter() - t) * 1e3:.1f} ms")
    for nm in KERNELS:
        ms, cnt = ctx.profile_get(nm)
        if cnt:
            kern[nm] = {"total_ms": ms / args.kernel_steps, "launches": cnt // args.kernel_steps,
                        "avg_launch_ms": ms / cnt}
    ctx.profile(False)
    eng.close()
    mode1 = bench_wz_mode1(args, jobs, dev) if args.mode1_steps else None
    pcie = bench_wz_host_buffers(args, ctx, jobs) if args.pcie_steps else None
    plain = bench_wz_plain(args, ctx, jobs, dev, outs) if args.plain_frames else None
    serve_leg = bench_serve_leg(args, ctx) if args.serve_rate > 0 else None
    valu = valu_calibration(ctx)
    psnrs = [st["psnr"] for _, st in outs]
    t_step = float(np.mean(times))
    value = len(jobs) / t_step
    eu = sum(sum(st["iterations"]) * 3 * 6336 for _, st in outs)
    dom = max(kern.items(), key=lambda kv: kv[1]["total_ms"]) if kern else ("", {"total_ms": 0, "launches": 1})
    res = {
        "metric": METRIC, "value": value, "unit": "WZ frames/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded CIF sequences: 40 Gaussian blobs, global shift U{-6..6}, 4 moving patches, noise)",
        "config": {"workload": f"wz_cif_q8: CIF 352x288, 33 frames, GOP 2, Qindex 8, {args.streams} sequences "
                               f"({len(jobs)} WZ frames, 63 bitplanes each) per step, batched, max_iter={args.max_iter}",
                   "parallelism": "single GPU"},
        "plain_fps": plain["fps"] if plain else None,
        "batched_over_plain": value / plain["fps"] if plain else None,
        "plain": plain,
        "bp_edge_updates_per_s": eu / t_step,
        "latency_ms": {"frame_p50": float(np.percentile(frame_lat, 50) * 1e3),
                       "frame_p99": float(np.percentile(frame_lat, 99) * 1e3),
                       "frames_sampled": len(frame_lat),
                       "plain_frame_p50": plain["latency_ms"]["p50"] if plain else None,
                       "plain_frame_p99": plain["latency_ms"]["p99"] if plain else None,
                       "note": "batched: each frame from the submission of its step's batch to its completion"},
        "psnr_mean": float(np.mean(psnrs)),
        "recon_host_roundtrips": 0,  # reconstruction's exp runs on the device (dvc_exp)
        "inputs": "resident in HBM before the timed region (key frames, syndromes, output frames)",
        "timing": "timed steps run without per-kernel events (steady steps replay captured HIP graphs); "
                  f"kernel durations come from {args.kernel_steps} further step(s) of the same workload with HIP "
                  "events around each launch on its stream",
        "kernels_ms_per_step": kern, "dominant_kernel": dom[0], "bp_tuning": tuning,
        "parallel_mode_1": mode1,
        "pcie_inclusive": pcie,
        "serve": serve_leg,
        "valu_calibration": valu,
    }
    res["roofline"] = roofline_wz(dom, kern, args, len(jobs), outs, valu)
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_wz(jobs, args)
    return res


def bench_wz_plain(args, ctx, jobs, dev, outs):
    """The plain drop-in path: one frame at a time (an engine with one frame
    in flight) with the sequential rate search (one BP launch per code, the
    host choosing the next code), inputs resident in HBM as in the batched
    leg, through one long-lived engine.  Also one dvc_decode_wz_frames call
    (batched = 0), which sets up and tears down its own engine, timed
    separately as the per-call set-up cost."""
    from dvc_amd import _lib
    n = min(args.plain_frames, len(jobs))
    eng = _lib.Engine(ctx, max_inflight=1, lanes=1, sequential_bp=True, device_buffers=True, lane_steps=1 << 20)
    sub = jobs[:n]
    run_engine(eng, dev.submit_all(eng, sub[:1], 3000, [0]))  # warm the engine's pool
    dev.release_keys(eng, sub[:1], 3000)
    lat, mism = [], 0
    t_all = time.perf_counter()
    for i in range(n):
        t = time.perf_counter()
        (st,) = run_engine(eng, dev.submit_all(eng, [sub[i]], 3001 + i, [i]))
        lat.append(time.perf_counter() - t)
        dev.release_keys(eng, [sub[i]], 3001 + i)
        rec = dev.recons(i)
        mism += int(not (np.array_equal(rec, outs[i][0]) and st["levels"] == outs[i][1]["levels"]))
    dt = time.perf_counter() - t_all
    eng.close()
    t = time.perf_counter()
    (crec, cst), = ctx.decode_wz_frames([jobs[0]], batched=False)
    call_ms = (time.perf_counter() - t) * 1e3
    mism += int(not (np.array_equal(crec, outs[0][0]) and cst["psnr"] == outs[0][1]["psnr"]))
    log(f"[bench] plain: {n} frames, mean {np.mean(lat) * 1e3:.1f} ms/frame; one dvc_decode_wz_frames call "
        f"{call_ms:.1f} ms")
    return {"fps": n / dt, "frames": n, "ms_per_frame": float(np.mean(lat) * 1e3),
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
            "engine": "long-lived, max_inflight 1, sequential_bp 1, lane_steps 2^20, HBM-resident inputs",
            "decode_wz_frames_call_ms": call_ms,
            "mismatches_vs_batched": mism}


def bench_serve_leg(args, ctx):
    """A bounded run of the serving configuration beside the headline: 32
    streams (24 QCIF + 8 CIF, GOP U{2,4,8}, Qindex U{1..8}), Poisson arrivals
    at --serve-rate requests/s, with a sample of served frames checked
    against their decodes alone."""
    from dvc_amd import serve
    t0 = time.time()
    streams = serve.make_streams(frames=args.serve_leg_frames, seed=4001)
    log(f"[bench] serve leg: encoded {sum(len(s.order) for s in streams)} frames in {time.time() - t0:.1f}s")
    res, _ = serve.run(ctx, streams, args.serve_rate, max_inflight=args.serve_inflight, lanes=args.lanes,
                       verify=serve.verify_solo(ctx, sample=8))
    log(f"[bench] serve leg: {res['frames_per_s']:.1f} frames/s, p50 {res['latency_ms']['p50']:.0f} ms, "
        f"p99 {res['latency_ms']['p99']:.0f} ms, {res['verified']}")
    res["config"] = (f"32 streams (24 QCIF + 8 CIF), {args.serve_leg_frames} frames each, GOP U{{2,4,8}}, "
                     f"Qindex U{{1..8}}, seed 4001, Poisson arrivals at {args.serve_rate:.0f} requests/s")
    return res


def valu_calibration(ctx, waves_per_cu=16, iters=4000):
    """Measured VALU issue rates (wave64 instructions/s) of dependent-free
    chains built with the library's flags (packed FP32 off): the peaks the
    BP roofline is priced against."""
    from dvc_amd import _lib
    out = {}
    for kind, name in enumerate(["fp32_fma", "fp32_add", "fp32_mul", "fp64_fma", "int32_add", "sad_u8", "mqsad_u32_u8",
                                 "qsad_pk_u16_u8", "trans_rcp_f32"]):
        ms, wi = ctypes.c_double(), ctypes.c_double()
        rc = _lib.lib.dvc_selftest_valu_rate(ctx.h, kind, waves_per_cu, iters, ctypes.byref(ms), ctypes.byref(wi))
        if rc != _lib.DVC_OK and kind >= 7:  # a library without the later kinds
            continue
        _lib._check(rc, "dvc_selftest_valu_rate")
        out[name] = {"wave_instr_per_s": wi.value / (ms.value * 1e-3), "kernel_ms": ms.value,
                     "wave_instr": wi.value}
    # the BP roofline's peak: the highest rate of the arithmetic kinds (the SAD
    # kinds price the SI refresh: 4 and 16 absolute differences per lane)
    out["peak_wave_instr_per_s"] = max(out[k]["wave_instr_per_s"] for k in
                                       ("fp32_fma", "fp32_add", "fp32_mul", "fp64_fma", "int32_add"))
    out["sad_abs_diffs_per_s"] = {"sad_u8": 64 * 4 * out["sad_u8"]["wave_instr_per_s"],
                                  "mqsad_u32_u8": 64 * 16 * out["mqsad_u32_u8"]["wave_instr_per_s"],
                                  "qsad_pk_u16_u8": 64 * 16 * out["qsad_pk_u16_u8"]["wave_instr_per_s"]
                                  if "qsad_pk_u16_u8" in out else None}
    out["per_cu_cycle_at_2.4GHz"] = out["peak_wave_instr_per_s"] / (256 * 2.4e9)
    return out


def bench_wz_host_buffers(args, ctx, jobs):
    """The same workload with host (pageable) request buffers: key frames,
    syndromes and output frames cross PCIe inside the timed steps, and the
    PSNR is computed on the host.  Reported beside `value`, never as it."""
    from dvc_amd import _lib
    eng = _lib.Engine(ctx, max_inflight=len(jobs), lanes=args.lanes, device_buffers=False, lane_steps=1 << 20)

    def submit_all(gen):
        return [eng.submit(j, past_key=(gen << 48) | ((j["seq"] + 1) << 32) | j["idx"][0],
                           fut_key=(gen << 48) | ((j["seq"] + 1) << 32) | j["idx"][2]) for j in jobs]
    run_engine(eng, submit_all(2000))
    times = []
    for i in range(args.pcie_steps):
        t = time.perf_counter()
        run_engine(eng, submit_all(2001 + i))
        times.append(time.perf_counter() - t)
        log(f"[bench] host buffers step {i}: {times[-1] * 1e3:.1f} ms for {len(jobs)} frames")
    eng.close()
    t_step = float(np.mean(times))
    keys = {(j["seq"], j["idx"][0]) for j in jobs} | {(j["seq"], j["idx"][2]) for j in jobs}
    fs = jobs[0]["past"].size
    pcie = len(keys) * fs + sum(len(j["wz"].acc) + fs for j in jobs)  # key frames and syndromes up, frames down
    return {"value": len(jobs) / t_step, "unit": "WZ frames/s", "ms_per_step": t_step * 1e3,
            "bytes_over_pcie_per_frame": int(pcie / len(jobs))}


def bench_wz_mode1(args, jobs, dev):
    """The same workload decoded with ParallelMode 1 BP semantics (the
    reference's OpenMP decoder: early jump on repeated decisions), which is
    also the semantics of the CPU baseline leg."""
    from dvc_amd import _lib
    ctx = _lib.Context(device=0, max_iter=args.max_iter, bp_threads=args.bp_threads, parallel_mode=1)
    eng = _lib.Engine(ctx, max_inflight=len(jobs), lanes=args.lanes, device_buffers=True, lane_steps=1 << 20)
    run_engine(eng, dev.submit_all(eng, jobs, 1000))
    times = []
    for i in range(args.mode1_steps):
        t = time.perf_counter()
        stats = run_engine(eng, dev.submit_all(eng, jobs, 1001 + i))
        times.append(time.perf_counter() - t)
        log(f"[bench] ParallelMode 1 step {i}: {times[-1] * 1e3:.1f} ms for {len(jobs)} frames")
    eng.close()
    plain_t = None
    if args.plain_frames:
        t = time.perf_counter()
        ctx.decode_wz_frames([jobs[0]], batched=False)
        plain_t = time.perf_counter() - t
    t_step = float(np.mean(times))
    ctx.close()
    return {"value": len(jobs) / t_step, "unit": "WZ frames/s", "ms_per_step": t_step * 1e3,
            "plain_fps": (1.0 / plain_t) if plain_t else None,
            "bp_iterations_per_frame": float(np.mean([sum(st["iterations"]) for st in stats]))}


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE passes scaled by the
    copy calibration), or None when there is none."""
    f = ROOT / "profiles" / f"{kernel}_{workload}.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(f"traffic_bytes_per_launch {kernel} "):
            return float(line.split()[-1])
    return None


def profiled_value(kernel, workload, key):
    """A per-launch figure `key` of `kernel` from the committed rocprofv3
    summary of this workload, or None."""
    f = ROOT / "profiles" / f"{kernel}_{workload}.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(f"{key} {kernel} "):
            return float(line.split()[-1])
    return None


def profiled_mix(kernel, workload):
    """The VALU instruction classes per launch of `kernel` (the committed
    rocprofv3 SQ_INSTS_VALU_* passes of this workload): dict class -> count."""
    f = ROOT / "profiles" / f"{kernel}_{workload}.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(f"valu_mix_per_launch {kernel} "):
            return {k: float(v) for k, v in (t.split("=") for t in line.split()[2:])}
    return None


def mix_peak(mix, valu):
    """The issue rate the instruction mix allows: total / sum(n_k / rate_k),
    each class at the rate valu_calibration measured for it (fp32 for the
    classes it has no kernel for)."""
    k = valu
    fp32 = max(k["fp32_fma"]["wave_instr_per_s"], k["fp32_add"]["wave_instr_per_s"], k["fp32_mul"]["wave_instr_per_s"])
    rate = {"int32": k["int32_add"]["wave_instr_per_s"], "fp32": fp32, "f64": k["fp64_fma"]["wave_instr_per_s"],
            "trans": k["trans_rcp_f32"]["wave_instr_per_s"] if "trans_rcp_f32" in k else fp32 / 2,
            "cvt": fp32, "int64": fp32, "other": fp32}
    total = sum(mix.values())
    return total / sum(n / rate[c] for c, n in mix.items()), rate


def bp_roofline(kernel_ms_avg, hbm_bytes_per_launch, edge_updates_per_launch, workload, valu=None):
    """Roofline of the BP kernel.  Its bound is the VALU issue port: the exact
    glibc-identical transcendentals (two phi per edge-update) are about 200
    VALU instructions per edge-update, against 16 LDS bytes and a few HBM
    bytes.  achieved = the profiled VALU wave64 instructions per launch (the
    committed rocprofv3 SQ_INSTS_VALU pass of this workload) over the average
    launch duration measured here; peak = the issue rate measured by
    valu_calibration (dependent-free chains built with the library's flags).
    Also reported: HBM (per (bitplane, code) the sequential search runs, 4N
    LLR + ceil(m/8) syndrome + ceil(N/8) decisions + 12N column table, a word per edge)
    against 8 TB/s, with the profiled traffic, and LDS (16 bytes per
    edge-update) against 256 CUs x 128 B/clk x 2.4 GHz."""
    t = kernel_ms_avg * 1e-3
    hbm = hbm_bytes_per_launch / t / 1e9
    lds = 16.0 * edge_updates_per_launch / t / 1e9
    vi = profiled_value("bp_6336", workload, "valu_wave_instr_per_launch")
    res = {"kernel": "bp_6336", "kernel_ms_avg": kernel_ms_avg,
           "traffic": profiled_traffic("bp_6336", workload),
           "hbm": {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": hbm / HBM_PEAK_GBS,
                   "algorithmic_bytes_per_launch": hbm_bytes_per_launch},
           "lds": {"achieved": lds, "peak": LDS_PEAK_GBS, "unit": "GB/s", "frac": lds / LDS_PEAK_GBS,
                   "bytes_per_launch": 16.0 * edge_updates_per_launch},
           "profile": f"profiles/bp_6336_{workload}.md"}
    mix = profiled_mix("bp_6336", workload)
    if vi and valu:
        if mix:
            peak, rates = mix_peak(mix, valu)
            src = ("the instruction mix of the committed SQ_INSTS_VALU_* passes, each class at the issue rate "
                   "valu_calibration measured in this run")
        else:
            peak, rates = valu["peak_wave_instr_per_s"], None
            src = "valu_calibration (measured in this run), highest rate"
        res.update({"bound": "valu", "achieved": vi / t, "peak": peak, "unit": "VALU wave64-instr/s",
                    "frac": vi / t / peak, "wave_instr_per_launch": vi, "instruction_mix_per_launch": mix,
                    "class_rates": rates, "peak_source": src,
                    "peak_single_rate": valu["peak_wave_instr_per_s"],
                    "frac_of_single_rate": vi / t / valu["peak_wave_instr_per_s"]})
    else:
        res.update({"bound": "hbm", "achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm / HBM_PEAK_GBS})
    return res


def roofline_wz(dom, kern, args, nframes, outs, valu=None):
    """Roofline of the dominant kernel (the BP launch of each bitplane step)."""
    name, d = dom
    if name != "bp_6336":
        return {"bound": "hbm", "kernel": name, "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None}
    launches = max(d["launches"], 1)
    byts = 0
    eu = 0
    for _, st in outs:
        for b in range(st["nbitplanes"]):
            for lvl in range(st["start_levels"][b], st["levels"][b] + 1):
                if lvl < 64:
                    byts += 4 * 6336 + (96 * (lvl + 2) + 7) // 8 + 792 + 12 * 6336
            eu += st["iterations"][b] * 3 * 6336
    return bp_roofline(d["total_ms"] / launches, byts / launches, eu / launches, "wz_cif_q8", valu)


def cpu_baseline_wz(jobs, args):
    """Oracle (CPU restatement of the reference decoder) on a bounded sample of
    the same workload: ParallelMode 1 (OpenMP over the host cores the run is
    given) on consecutive frames until about `cpu_seconds` have passed, and
    ParallelMode 0 on one core for one frame."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
    t = time.perf_counter()
    n = 0
    split = {}
    for j in jobs:
        _, info = O.decode_wz_frame(j["past"], j["fut"], j["orig"], j["wz"], 8, mode=1, nthreads=cores,
                                    max_iter=args.max_iter, idx=j["idx"])
        for k, v in info["times_ms"].items():
            split[k] = split.get(k, 0.0) + v
        n += 1
        if time.perf_counter() - t > args.cpu_seconds:
            break
    dt = time.perf_counter() - t
    j = jobs[0]
    t0 = time.perf_counter()
    _, info0 = O.decode_wz_frame(j["past"], j["fut"], j["orig"], j["wz"], 8, mode=0, nthreads=1,
                                 max_iter=args.max_iter, idx=j["idx"])
    dt0 = time.perf_counter() - t0
    return {"value": n / dt, "unit": "WZ frames/s", "cores": cores, "kind": "port",
            "sample": f"{n} CIF Q8 WZ frames of the same workload, ParallelMode 1 (OpenMP, {cores} threads), "
                      f"{dt:.1f}s; ParallelMode 0 on 1 core: 1 frame in {dt0:.1f}s",
            "mode0_1core_fps": 1.0 / dt0,
            # the decodeGOP components the reference times under SHOW_TIME_INFO, ms per frame
            "split_ms_per_frame": {k: v / n for k, v in split.items()},
            "mode0_split_ms_per_frame": dict(info0["times_ms"])}


def cpu_baseline_ldpca(planes, args):
    """Oracle (CPU restatement, ParallelMode 0, one core) on a bounded sample."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib as O
    sample = [p for p in planes if p["N"] == 6336][:args.cpu_planes]
    t = time.perf_counter()
    eu = 0
    for p in sample:
        _, st, _ = O.ldpca_decode(p["LR"], p["acc"], p["crc"], max_iter=args.max_iter)
        eu += st["total_iters"] * 3 * p["N"]
        if time.perf_counter() - t > args.cpu_seconds:
            break
    dt = time.perf_counter() - t
    return {"value": eu / dt, "unit": "BP edge-updates/s", "cores": 1, "kind": "port",
            "sample": f"{len(sample)} N=6336 bitplanes of the same set (or {args.cpu_seconds}s cap), "
                      f"sequential rate search, {dt:.1f}s"}


def bench_serve(args):
    """Serving loop: 32 concurrent synthetic streams (24 QCIF + 8 CIF, GOP
    from {2,4,8}, Qindex 1..8), Poisson arrivals at each rate of --rates;
    value = frames/s at the highest rate, with p50/p99 latency per rate and a
    sample of served frames checked against their solo decodes."""
    from dvc_amd import _lib, serve
    ctx = _lib.Context(device=0, max_iter=args.max_iter, bp_threads=args.bp_threads)
    if args.warmup:
        serve.run(ctx, serve.make_streams(n_qcif=2, n_cif=1, frames=5), rate=1000.0, lanes=args.lanes)
    sweep = []
    for rate in [float(r) for r in args.rates.split(",")]:
        streams = serve.make_streams(frames=args.serve_frames)
        res, _ = serve.run(ctx, streams, rate, max_inflight=args.serve_inflight, lanes=args.lanes,
                           verify=serve.verify_solo(ctx, sample=8))
        log(f"[bench] serve rate {rate:.0f}/s: {res['frames_per_s']:.1f} frames/s, p50 "
            f"{res['latency_ms']['p50']:.0f} ms, p99 {res['latency_ms']['p99']:.0f} ms, {res['verified']}")
        sweep.append(res)
    top = sweep[-1]
    return {"metric": METRIC, "value": top["frames_per_s"], "unit": "WZ frames/s", "n_gpus": 1,
            "steps": len(sweep), "warmup": args.warmup, "ms_per_step": top["wall_s"] * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded QCIF/CIF sequences, Poisson arrivals)",
            "config": {"workload": f"serve: 32 streams (24 QCIF + 8 CIF), {args.serve_frames} frames each, GOP U{{2,4,8}}, "
                                   f"Qindex U{{1..8}}, Poisson arrivals at {args.rates} requests/s, max_iter={args.max_iter}"},
            "latency_ms": {"p50": top["latency_ms"]["p50"], "p99": top["latency_ms"]["p99"]},
            "sweep": sweep}


def bench_si(args):
    """Batched side-information and noise-model generation: 64 CIF key-frame
    pairs (16 consecutive pairs k, k+2 of 4 sequences of the CIF generator),
    WZ frame at the midpoint; value = pairs/s of dvc_side_info_batch (host
    buffers in and out), with a sample checked against the oracle."""
    from dvc_amd import _lib, workload as W
    pairs = []
    for st in range(4):
        vid = W.synthetic_video(352, 288, 33, 6, 1300 + st)
        for k in range(0, 32, 2):
            pairs.append(dict(past=vid[k], fut=vid[k + 2], idx=(k, k + 1, k + 2), qindex=8))
    ctx = _lib.Context(device=0, max_iter=args.max_iter)
    for _ in range(args.warmup):
        ctx.side_info_batch(pairs)
    ctx.profile(True)
    times = []
    for i in range(args.steps):
        t = time.perf_counter()
        outs = ctx.side_info_batch(pairs)
        times.append(time.perf_counter() - t)
        log(f"[bench] step {i}: {times[-1] * 1e3:.1f} ms for {len(pairs)} pairs")
    kern = {}
    for nm in ["keyframe_filters", "create_si", "noise_param"]:
        ms, cnt = ctx.profile_get(nm)
        if cnt:
            kern[nm] = {"total_ms": ms / args.steps, "launches": cnt // args.steps}
    ctx.profile(False)
    res = {"metric": METRIC, "value": len(pairs) / float(np.mean(times)), "unit": "SI+alpha key-frame pairs/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": float(np.mean(times)) * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32/f64/f32",
           "data": "synthetic (seeded CIF sequences)",
           "config": {"workload": "si_batch: 64 CIF key-frame pairs, WZ frame at the midpoint, Qindex 8 bands for alpha"},
           "kernels_ms_per_step": kern}
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        from test_gpu_si import oracle_si_alpha
        t = time.perf_counter()
        n = 0
        bad = 0
        for d, (si, r, al) in zip(pairs, outs):
            osi, ores, oal = oracle_si_alpha(d["past"], d["fut"], d["idx"], 8)
            bad += int(not (np.array_equal(si, osi) and np.array_equal(al.view(np.uint32), oal.view(np.uint32))))
            n += 1
            if time.perf_counter() - t > args.cpu_seconds:
                break
        dt = time.perf_counter() - t
        res["cpu_baseline"] = {"value": n / dt, "unit": "SI+alpha key-frame pairs/s", "cores": 1, "kind": "port",
                               "sample": f"{n} pairs of the same set on the oracle (sequential), {dt:.1f}s; "
                                         f"{bad} mismatching outputs"}
    return res


def bench_copy(args):
    """HBM calibration: a device-to-device copy of a known size (1 GiB read +
    1 GiB written per step).  Its FETCH_SIZE / WRITE_SIZE readings scale the
    same counters of the decoding kernels to bytes."""
    from dvc_amd import _lib
    nbytes = 1 << 30
    ctx = _lib.Context(device=0)
    src = _lib.DeviceBuffer(ctx, nbytes)
    dst = _lib.DeviceBuffer(ctx, nbytes)
    pattern = np.random.default_rng(0).integers(0, 256, 1 << 20, dtype=np.uint8)
    for off in range(0, nbytes, pattern.size):
        _lib._check(_lib.lib.dvc_buffer_copy(ctx.h, src.ptr + off, pattern.ctypes.data, pattern.size, 0), "copy")
    for _ in range(args.warmup):
        dst.copy_from(src)
    t = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_from(src)
    dt = (time.perf_counter() - t) / args.steps
    ok = np.array_equal(dst.download(pattern.size, np.uint8), pattern)
    gbs = 2 * nbytes / dt / 1e9
    return {"metric": "HBM copy bandwidth (calibration)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": "hbm_copy: 1 GiB device-to-device copy per step (hipMemcpyAsync)"},
            "bytes_read_per_step": nbytes, "bytes_written_per_step": nbytes, "copy_verified": bool(ok)}


def bench_valu(args):
    """VALU issue-rate calibration: one launch per step of each chain kind
    (fp32 fma/add/mul, fp64 fma, int32 add), 16 waves per CU; value = the
    highest rate in wave64 instructions/s.  Its rocprofv3 kernel trace and
    SQ_INSTS_VALU pass are the committed calibration of the BP roofline."""
    from dvc_amd import _lib
    ctx = _lib.Context(device=0)
    for _ in range(args.warmup):
        valu_calibration(ctx)
    runs = [valu_calibration(ctx) for _ in range(args.steps)]
    kinds = [k for k in runs[0] if isinstance(runs[0][k], dict) and "kernel_ms" in runs[0][k]]
    best = {k: max(r[k]["wave_instr_per_s"] for r in runs) for k in kinds}
    peak = max(best.values())
    return {"metric": "VALU issue rate (calibration)", "value": peak, "unit": "VALU wave64-instr/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": float(np.mean([sum(r[k]["kernel_ms"] for k in kinds) for r in runs])),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64/i32",
            "data": "synthetic",
            "config": {"workload": "valu_peak: 8 independent chains per lane, 128 VALU instructions per trip, "
                                   "4000 trips, 16 waves per CU, library compile flags (packed FP32 off)"},
            "per_kind_wave_instr_per_s": best, "per_cu_cycle_at_2.4GHz": peak / (256 * 2.4e9),
            "wave_instr_per_launch": runs[0]["fp32_fma"]["wave_instr"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="wz_cif_q8", choices=["wz_cif_q8", "ldpca_batch", "hbm_copy", "serve", "si_batch", "valu_peak"])
    ap.add_argument("--streams", type=int, default=64, help="CIF sequences of 16 WZ frames decoded per step")
    ap.add_argument("--plain-frames", type=int, default=24,
                    help="wz_cif_q8: frames of the plain leg (one frame in flight, sequential rate search)")
    ap.add_argument("--kernel-steps", type=int, default=1,
                    help="wz_cif_q8: untimed steps with per-kernel HIP events after the timed ones")
    ap.add_argument("--serve-rate", type=float, default=200.0,
                    help="wz_cif_q8: arrival rate of the bounded serving leg (0 skips it)")
    ap.add_argument("--serve-leg-frames", type=int, default=9, help="wz_cif_q8: frames per stream of the serving leg")
    ap.add_argument("--max-iter", type=int, default=100)
    ap.add_argument("--bp-threads", type=int, default=0, help="threads per CIF BP workgroup (0: library default)")
    ap.add_argument("--lanes", type=int, default=2, help="engine lanes (0: library default)")
    ap.add_argument("--autotune", action="store_true",
                    help="run dvc_autotune before timing (the library defaults are its choices on MI355X)")
    ap.add_argument("--rates", default="50,200,800", help="serve: Poisson arrival rates (requests/s)")
    ap.add_argument("--serve-frames", type=int, default=17, help="serve: frames per stream")
    ap.add_argument("--serve-inflight", type=int, default=128, help="serve: engine frames in flight")
    ap.add_argument("--mode1-steps", type=int, default=2,
                    help="wz_cif_q8: also time this many steps with ParallelMode 1 BP semantics (0: skip)")
    ap.add_argument("--pcie-steps", type=int, default=2,
                    help="timed steps of the host-buffer leg (PCIe-inclusive rate; 0 skips it)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-planes", type=int, default=150)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(streams would shard across GPUs without any exchange).", file=sys.stderr)
        sys.exit(2)
    res = {"wz_cif_q8": bench_wz, "ldpca_batch": bench_ldpca, "hbm_copy": bench_copy,
           "serve": bench_serve, "si_batch": bench_si, "valu_peak": bench_valu}[args.workload](args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
