#!/usr/bin/env python3
"""Benchmark of the Bloch scan hot path on one MI355X.

Workload (one "step" = one full scan of it, inputs resident in HBM):
  3-D balanced SSFP, 128 x 128 x 64 voxels x 4 isochromats (uniform +-pi
  intravoxel spread), T1/T2/PD/dB0 seeded as the single-slice case, 8
  Gaussian receive coils, 45 deg hard rf with alternating phase, TR 5 ms
  on a 10 us grid (500 steps), 128 x 64 phase encodes x 128 readout
  samples = 8192 TRs = 4.096e6 steps = 1.72e13 isochromat-steps.

Prints one JSON line with isochromat-steps/s, the forward kernel's roofline
and the fp32 CPU oracle timed on the first TRs of the same phantom on this
host's cores, then sections for the other paths (batched fingerprinting,
adjoint, multi-pool at its full size, recon, SAR, sequence compiler).

The forward kernel keeps each thread's isochromats in registers for a whole
launch (16 TRs), so HBM carries little (rocprofv3 FETCH_SIZE + WRITE_SIZE,
about 77 GB/s) and the kernel is bound by VALU instruction issue. The
roofline is that of issue: busy 4-cycle issue slots (VALU instructions less
the dual-issued pairs, from rocprofv3 counters of this kernel build in
profiles/bloch_fwd_bssfp3d_counters.json) over the slots of the live launch
time at the peak clock. Beside it: the counted fp32 flops against 157.3
TFLOP/s, the measured HBM rate against 8 TB/s, and, labelled as a
convention, the bytes a step-at-a-time engine would move (24 B state
read+write, 4 B dWRnd, 36/S B per-voxel maps, 8*Rx/S B of coil maps on ADC
steps per isochromat-step) and a flop model (10 per free step, +35 per rf
step, +8 per coil on ADC steps).

Profile: python bench.py --write-profile profiles/bloch_fwd_bssfp3d_counters.json
DB... turns rocprofv3 databases of this command into that file.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "isochromat-steps/s + % HBM roofline; batched speedup; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3           # MI355X fp32 vector peak (packed)
FP32_NONPACKED_PEAK_TFLOPS = 78.6  # non-packed fp32 FMA rate; this build disables packed fp32


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(spins: int, rx: int, adc_fraction: float) -> float:
    """Bytes per isochromat-step a step-at-a-time engine moves (see module doc)."""
    return 24 + 4 + 36.0 / spins + adc_fraction * 8.0 * rx / spins


def algorithmic_flops(rf_fraction: float, adc_fraction: float, rx: int) -> float:
    """fp32 flops per isochromat-step (see module doc)."""
    return 10.0 + 35.0 * rf_fraction + 8.0 * rx * adc_fraction


def cpu_baseline(n_tr: int, threads: int) -> dict:
    import oracle
    from mrilab_amd import workloads
    sc = workloads.bssfp3d(n_tr=n_tr)
    iso_steps = sc.n_iso * sc.n_steps()
    oracle.set_threads(threads, "f32")
    t0 = time.perf_counter()
    o = oracle.scan(sc, "f32")
    dt = time.perf_counter() - t0
    oracle.set_threads(0, "f32")
    # the GPU engine on the same sample: parity at the full phantom size
    import numpy as np
    from mrilab_amd import engine
    g = sc.copy()
    engine.DoScanAtGPU(*g.structs)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    sig = lambda x: x.VSig["Sx"] + 1j * x.VSig["Sy"]
    mag = lambda x: np.stack([x.VObj[k] for k in ("Mx", "My", "Mz")])
    parity = {"signal": rel(sig(g), sig(o)), "magnetization": rel(mag(g), mag(o))}
    log(f"cpu baseline sample: GPU vs oracle rel-L2 signal {parity['signal']:.2e} "
        f"magnetization {parity['magnetization']:.2e}")
    return {"value": iso_steps / dt, "unit": "isochromat-steps/s", "cores": threads, "kind": "port",
            "sample": f"first {n_tr} TRs of the same phantom and sequence ({iso_steps:.3g} isochromat-steps, "
                      f"{dt:.1f} s), fp32 restatement of DoScanAtCPU, OpenMP over slices",
            "gpu_rel_l2_on_sample": parity}


def batched_section(n_sets: int, plain_sample: int, reps: int) -> dict:
    """Batched uneven-length fingerprinting: n_sets single-voxel scans (64
    isochromats, 500-1500 TRs) as one persistent launch replayed from a
    captured graph, against plain per-scan DoScanAtGPU calls on a sample."""
    import numpy as np
    from mrilab_amd import engine, workloads
    t0 = time.perf_counter()
    scans = workloads.fingerprint_batch(n_sets=n_sets)
    log(f"batched: {n_sets} fingerprint scans generated ({time.perf_counter() - t0:.1f} s)")
    outs = [sc.copy() for sc in scans]  # prepared outside the timers, as for the plain calls
    t0 = time.perf_counter()
    b = engine.Batch([o.structs for o in outs], options=engine.make_options(timing=1))
    t_create = time.perf_counter() - t0
    info = b.info()
    log(f"batched: create {t_create:.2f} s (library {info['create_ms']:.0f} ms: decode {info['decode_ms']:.0f} ms, "
        f"tuning {info['tune_ms']:.0f} ms over {info['n_candidates']} shapes; best spt {info['spins_per_thread']} "
        f"block {info['block_size']})")
    b.run()
    b.sync()
    run_ms = []
    t0 = time.perf_counter()
    for _ in range(reps):
        b.run()
        b.sync()
        run_ms.append(b.info()["last_run_ms"])
    wall_run = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    b.fetch()  # into the structs given at create, in place as MRiLab's globals
    t_fetch = time.perf_counter() - t0
    b.close()
    # plain per-scan launches, end to end (what BatchSim.m does per scan): all
    # scans, or an evenly spaced sample whose mean time is scaled to all
    if plain_sample <= 0 or plain_sample >= n_sets:
        plain_sample = n_sets
    idx = np.linspace(0, n_sets - 1, plain_sample).astype(int)
    plain = []
    errs = []
    for i in idx:
        o = scans[i].copy()
        t0 = time.perf_counter()
        engine.DoScanAtGPU(*o.structs)
        plain.append(time.perf_counter() - t0)
        a = o.VSig["Sx"] + 1j * o.VSig["Sy"]
        c = outs[i].VSig["Sx"] + 1j * outs[i].VSig["Sy"]
        errs.append(float(np.linalg.norm(c - a) / max(np.linalg.norm(a), 1e-30)))
        if len(plain) % 1024 == 0:
            log(f"batched: {len(plain)}/{plain_sample} plain per-scan calls, {np.sum(plain):.1f} s")
    plain_total = float(np.sum(plain)) if plain_sample == n_sets else float(np.mean(plain)) * n_sets
    batched_e2e = t_create + wall_run + t_fetch
    iso_steps = info["n_iso_steps"]
    return {"workload": f"{n_sets} single-voxel fingerprinting scans x 64 isochromats, 500-1500 TRs of 5 steps "
                        "(hard pulse, TE 3 ms, TR 10-15 ms), flips 5-70 deg",
            "isochromat_steps": iso_steps, "value": iso_steps / wall_run, "unit": "isochromat-steps/s",
            "ms_per_batch_run": wall_run * 1e3, "device_ms_per_run": float(np.median(run_ms)),
            "end_to_end_s": batched_e2e, "create_s": t_create, "decode_ms": info["decode_ms"],
            "tune_ms": info["tune_ms"], "library_create_ms": info["create_ms"], "fetch_s": t_fetch,
            "plain_per_scan_ms": float(np.mean(plain)) * 1e3, "plain_sample": int(plain_sample),
            "plain_total_s": plain_total,
            "plain_total_kind": "measured, every scan" if plain_sample == n_sets else
                                f"extrapolated from {plain_sample} evenly spaced scans",
            "speedup": plain_total / wall_run, "speedup_end_to_end": plain_total / batched_e2e,
            "max_rel_l2_vs_plain": max(errs),
            "shape": {"spins_per_thread": info["spins_per_thread"], "block_size": info["block_size"],
                      "grid_blocks": info["grid_blocks"], "tiles": info["n_tiles"],
                      "tuned_candidates": info["n_candidates"], "tune_best_ms": info["tune_best_ms"],
                      "tune_worst_ms": info["tune_worst_ms"]}}


def adjoint_section(reps: int) -> dict:
    """Adjoint rf design: 128x128 slice x 32 isochromats, 2000 complex rf
    samples at 4 us under a spiral-in gradient, checkpoint every 64 steps;
    device time of the forward pass and of the backward (gradient) pass, the
    latter with segments rebuilt backwards (default) and recomputed into
    scratch memory."""
    import numpy as np
    from mrilab_amd import engine, workloads
    e = workloads.excitation()
    out = {}
    for mode in (0, 1):
        adj = engine.Adjoint(e, checkpoint=64, segment_mode=mode)
        adj.forward(e.rf, e.T1, e.T2, e.dB0)
        adj.backward()
        fwd, bwd = [], []
        for _ in range(reps):
            adj.forward(e.rf, e.T1, e.T2, e.dB0)
            fwd.append(adj.info()["forward_ms"])
            adj.backward()
            bwd.append(adj.info()["backward_ms"])
        info = adj.info()
        adj.close()
        out[mode] = (float(np.median(fwd)), float(np.median(bwd)), info)
    f, b, info = out[0]
    bs = out[1][1]
    iso_steps = info["n_iso"] * info["n_steps"]
    log(f"adjoint: forward {f:.2f} ms, backward {b:.2f} ms (rebuilt={info['segments_rebuilt']}), ratio {b / f:.2f}; "
        f"backward through scratch {bs:.2f} ms")
    return {"workload": "rf design, 128x128 x 32 isochromats, 2000 complex rf samples at 4 us, spiral-in gradient, "
                        "L2 loss on the final Mxy, gradients w.r.t. rf, T1, T2, dB0, checkpoint K=64",
            "forward_ms": f, "backward_ms": b, "adjoint_over_forward": b / f,
            "segments_rebuilt": info["segments_rebuilt"],
            "backward_scratch_ms": bs, "adjoint_over_forward_scratch": bs / f,
            "forward_isochromat_steps_per_s": iso_steps / (f / 1e3), "checkpoints": info["n_checkpoints"],
            "spins_per_thread": info["spins_per_thread"]}


def multipool_section(threads: int, sat_time: float = 2.0, extras: bool = True) -> dict:
    """Multi-pool exchange with MT saturation, the full configuration: the
    256x256 3-pool phantom (free water, myelin water, bound pool with a
    super-Lorentzian line) under 40 saturation offsets over +-50 ppm at 3 T,
    each a ``sat_time`` s train of 300 deg Gaussian pulses (20 ms on, 20 ms
    off) then readout, every offset from equilibrium. The 40 offsets run as
    one batched launch (offsets x voxels as the parallel axis), in the fp32
    operator split and in the fp64 exact scheme; device time of a whole run.
    Beside it: both schemes against the oracle on a 2x2-voxel crop at 4 of
    the offsets over the full train, and the oracle's split scheme timed on
    a crop (CPU baseline)."""
    import numpy as np
    import oracle
    from mrilab_amd import engine, workloads
    t0 = time.perf_counter()
    scans = workloads.mt_cest(sat_time=sat_time, per_offset=True)
    n_off = len(scans)
    log(f"multipool: {n_off} offsets x {scans[0].n_steps()} steps built ({time.perf_counter() - t0:.1f} s)")
    out = {"workload": f"256x256 voxels x 3 pools (Dirichlet(8,1,1) fractions, k~U[1,30]/s, T2b 10 us), {n_off} "
                       f"saturation offsets over +-50 ppm at 3 T, each {sat_time:g} s of 300 deg Gaussian pulses "
                       "(20 ms on, 20 ms off) then readout, 10 us steps, offsets batched in one launch"}
    # scheme 1 folds the repeated saturation periods into loop records (each
    # period's exact affine map, found by running the period once per basis
    # state, applied once per period); scheme 2 steps every record
    for scheme in (0, 1, 2):
        name = ("split", "exact", "exact_stepwise")[scheme]
        t0 = time.perf_counter()
        b = engine.MultiPoolBatch(scans, options=engine.make_options(multipool_scheme=scheme, timing=1))
        t_create = time.perf_counter() - t0
        t0 = time.perf_counter()
        b.run()
        b.sync()
        el = time.perf_counter() - t0
        info = b.info()
        b.close()
        vs = 256 * 256 * info["n_steps"]
        out[name] = {"precision": "fp32" if scheme == 0 else "fp64", "steps_all_offsets": info["n_steps"],
                     "seconds": el, "kernel_seconds": info["fwd_kernel_ms"] / 1e3, "create_s": t_create,
                     "launches": info["n_launches"], "voxel_steps_per_s": vs / el}
        if scheme == 1:
            out[name]["note"] = ("the 50 identical 40 ms periods of each train are one loop record: the period "
                                 "runs for the zero state and the unit states of the 7 state entries that are not "
                                 "the bound pool's (zero) transverse entries, and its exact affine map is applied "
                                 "once per period; voxel_steps_per_s counts the sequence's steps")
        log(f"multipool {name}: {info['n_steps']} steps over {n_off} offsets in {el:.2f} s "
            f"({vs / el:.3e} voxel-steps/s, {info['n_launches']} launches)")
    if not extras:
        return out
    # parity on a crop over the full train, 4 offsets
    idx = [0, 13, 26, 39]
    crop = workloads.mt_cest(sat_time=sat_time, per_offset=True, crop=(slice(100, 102), slice(60, 62)),
                             offset_index=idx)
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    sig = lambda x: x.VSig["Sx"] + 1j * x.VSig["Sy"]
    par = {}
    for scheme in (0, 1):
        outs = [c.copy() for c in crop]
        with engine.MultiPoolBatch(crop, options=engine.make_options(multipool_scheme=scheme)) as b:
            b.run()
            b.fetch(outs)
        r64 = [oracle.gm_scan(c, "f64", scheme) for c in crop]
        if scheme == 1:
            par["exact_vs_fp64_oracle"] = max(rel(sig(g), sig(r)) for g, r in zip(outs, r64))
        else:
            # an fp32 evaluation of the split drifts from its fp64 evaluation
            # over 200,000 steps; the fp32 oracle's drift is the yardstick
            r32 = [oracle.gm_scan(c, "f32", 0) for c in crop]
            par["split_vs_fp64_split"] = max(rel(sig(g), sig(r)) for g, r in zip(outs, r64))
            par["fp32_oracle_vs_fp64_split"] = max(rel(sig(a), sig(r)) for a, r in zip(r32, r64))
            par["split_vs_fp32_oracle"] = max(rel(sig(g), sig(a)) for g, a in zip(outs, r32))
    out["crop_parity"] = {"sample": f"2x2 voxels, offsets {idx}, full train", "max_signal_rel_l2": par}
    log(f"multipool crop parity: {par}")
    # CPU baseline: the oracle's split scheme on a 32x32 crop, first offset, full train
    cpu = workloads.mt_cest(sat_time=sat_time, per_offset=True, crop=(slice(0, 32), slice(0, 32)),
                            offset_index=[0])[0]
    oracle.set_threads(threads, "f32")
    t0 = time.perf_counter()
    oracle.gm_scan(cpu, precision="f32", scheme=0)
    el = time.perf_counter() - t0
    oracle.set_threads(0, "f32")
    vs = 32 * 32 * cpu.n_steps()
    out["cpu_split"] = {"value": vs / el, "unit": "voxel-steps/s", "cores": threads, "kind": "port",
                        "sample": f"32x32 crop, first offset ({cpu.n_steps()} steps, {el:.1f} s), fp32 "
                                  "restatement of BlochKernelGM.cuh's split, OpenMP over voxels"}
    log(f"multipool cpu split: {vs / el:.3e} voxel-steps/s on {threads} threads")
    return out


def recon_section(plan, sc, reps: int) -> dict:
    """Signal to image on the scan's own k-space: the plan's device-resident
    signal reconstructed per coil (fftshift(ifftn(fftshift)), fp64 hipFFT,
    mrs_plan_recon), against the oracle (numpy) on the fetched signal."""
    import numpy as np
    import torch
    import oracle.recon as orc
    from mrilab_amd import engine
    nro, npe, nsl, rx = 128, int(sc.VObj["Rho"].shape[0]), int(sc.VObj["Rho"].shape[2]), int(sc.VCoi["RxCoilNum"])
    d = engine.recon_desc(nro, npe, nsl, 1, rx, 1)
    img = plan.recon(d, device=True)
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        img = plan.recon(d, device=True)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    gpu_s = float(np.median(ts))
    vs = {k: np.zeros_like(v) for k, v in sc.VSig.items() if k in ("Sx", "Sy", "Kx", "Ky", "Kz")}
    vs.update({k: v for k, v in sc.VSig.items() if k not in vs})
    plan.fetch(VSig=vs)
    t0 = time.perf_counter()
    want = orc.cart_recon(vs["Sx"], vs["Sy"], nro, npe, nsl, 1, rx, 1)
    cpu_s = time.perf_counter() - t0
    got = img["Real"].cpu().numpy().transpose(range(4, -1, -1)) + 1j * img["Imag"].cpu().numpy().transpose(
        range(4, -1, -1))
    ref = want["Real"] + 1j * want["Imag"]
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    vox = int(np.prod(engine.recon_dims(d)))
    log(f"recon: {gpu_s * 1e3:.2f} ms per call ({vox} image values), CPU oracle {cpu_s:.2f} s, rel-L2 {err:.1e}")
    return {"workload": f"the scan's k-space: {nro} readout x {npe} x {nsl} phase encodes x {rx} coils -> "
                        f"{nro - 1}x{npe}x{nsl} complex fp64 images per coil",
            "ms_per_call": gpu_s * 1e3, "image_values_per_s": vox / gpu_s, "rel_l2_vs_oracle": err,
            "cpu": {"seconds": cpu_s, "kind": "port", "cores": 1,
                    "sample": "numpy restatement of DoCartRecon.m on the fetched signal (pocketfft)"}}


def sar_section(threads: int, cpu_tr: int = 64) -> dict:
    """Local SAR (mrs_DoCalSARAtGPU) of the 3-D bSSFP scan's sequence on its
    128x128x64 phantom (a tissue ellipsoid in air), transmitted on 2 coils
    in parallel (E fields per coil, ECon), 16 sample times; per call: host
    event loop, uploads, the voxel kernel, downloads. The oracle runs the
    same on the first ``cpu_tr`` TRs. Then the 10-gram average
    (mrs_DoSARAverageAtGPU) of 4 windows of that SAR."""
    import numpy as np
    import oracle
    from mrilab_amd import engine, workloads
    sc = workloads.sar_scan(workloads.bssfp3d(), n_tx=2, n_samples=16, body=True)
    a = sc.copy()
    engine.DoCalSARAtGPU(*a.structs)  # warm-up (module load)
    ts = []
    for _ in range(3):
        a = sc.copy()
        t0 = time.perf_counter()
        engine.DoCalSARAtGPU(*a.structs)
        ts.append(time.perf_counter() - t0)
    gpu_s = float(np.median(ts))
    steps = int(a.VSig["SAR_rf_steps"])
    n_tr = int(sc.VCtl["TRNum"])
    small = workloads.sar_scan(workloads.bssfp3d(n_tr=cpu_tr), n_tx=2, n_samples=4)
    oracle.set_threads(threads, "f32")
    t0 = time.perf_counter()
    o = oracle.sar(small)
    cpu_s = time.perf_counter() - t0
    oracle.set_threads(0, "f32")
    cpu_steps = int(o.VSig["SAR_rf_steps"])
    nvox = int(np.prod(sc.VObj["MassDen"].shape))
    log(f"sar: {gpu_s * 1e3:.1f} ms per call ({nvox} voxels, {steps} rf steps of {n_tr} TRs); "
        f"CPU oracle {cpu_s:.2f} s for {cpu_steps} rf steps")
    return {"workload": f"128x128x64 voxels, 2 Tx coils in parallel, the bSSFP sequence ({n_tr} TRs, {steps} "
                        "steps with rf on), 16 SAR samples",
            "ms_per_call": gpu_s * 1e3, "voxel_rf_steps_per_s": nvox * steps / gpu_s,
            "cpu": {"seconds": cpu_s, "rf_steps": cpu_steps, "voxel_rf_steps_per_s": nvox * cpu_steps / cpu_s,
                    "cores": threads, "kind": "port",
                    "sample": f"first {cpu_tr} TRs, C restatement of DoCalSARAtCPU, OpenMP over voxels"},
            "average": sar_average_section(a, threads)}


def sar_average_section(sc, threads: int, grams: float = 10.0, cpu_slices=(28, 36)) -> dict:
    """N-gram average of the SAR engine's output: 4 windows of the 16
    samples (sar_window: DoSARWatchExec.m's rate / (2 MassDen) and power),
    then mrs_DoSARAverageAtGPU on device-resident fp64 inputs; the oracle
    (DoSARAverageAtCPU restated) on ``cpu_slices`` of the same volume,
    compared bit for bit there."""
    import numpy as np
    import torch
    import oracle
    from mrilab_amd import engine
    ts = sc.VSig["tSample"]
    idx = [0, 4, 8, 12, ts.size - 1]
    SAR, Power = engine.sar_window(sc.VObj, sc.VSig, ts[idx[:-1]], ts[idx[1:]])
    M = np.asarray(sc.VObj["MassDen"], dtype=np.float64)
    d = tuple(float(sc.VObj[k]) for k in ("XDimRes", "YDimRes", "ZDimRes"))

    def dev(a):  # column-major device tensor
        t = torch.from_numpy(np.ascontiguousarray(a.transpose(*reversed(range(a.ndim))))).cuda()
        return t.permute(*reversed(range(a.ndim)))

    dS, dP, dM = dev(SAR), dev(Power), dev(M)
    engine.DoSARAverageAtGPU(dS, dP, dM, *d, grams)  # warm-up
    torch.cuda.synchronize()
    tt = []
    for _ in range(3):
        t0 = time.perf_counter()
        a, b = engine.DoSARAverageAtGPU(dS, dP, dM, *d, grams)
        torch.cuda.synchronize()
        tt.append(time.perf_counter() - t0)
    gpu_s = float(np.median(tt))
    t0 = time.perf_counter()
    engine.DoSARAverageAtGPU(SAR, Power, M, *d, grams)
    host_s = time.perf_counter() - t0
    oracle.set_threads(threads, "f64")
    t0 = time.perf_counter()
    oa, ob = oracle.sar_average(SAR, Power, M, *d, grams, slices=cpu_slices)
    cpu_s = time.perf_counter() - t0
    oracle.set_threads(0, "f64")
    s0, s1 = cpu_slices
    ga, gb = a.cpu().numpy(), b.cpu().numpy()
    exact = bool(np.array_equal(ga[:, :, s0:s1], oa[:, :, s0:s1]) and np.array_equal(gb[:, :, s0:s1], ob[:, :, s0:s1]))
    T = SAR.shape[3]
    tissue = int(np.count_nonzero(M))
    cpu_tissue = int(np.count_nonzero(M[:, :, s0:s1]))
    log(f"sar average: {gpu_s * 1e3:.2f} ms per call ({tissue} tissue voxels x {T} windows, {grams} g; host arrays "
        f"{host_s * 1e3:.1f} ms); CPU oracle {cpu_s:.2f} s for slices {s0}-{s1 - 1}; bit-exact there: {exact}")
    return {"workload": f"{grams:g}-gram average, 128x128x64 voxels ({tissue} with tissue, "
                        f"{d[0] * 1e3:g} mm voxels), {T} SAR windows",
            "ms_per_call": gpu_s * 1e3, "ms_per_call_host_arrays": host_s * 1e3,
            "tissue_voxel_windows_per_s": tissue * T / gpu_s, "bit_exact_vs_oracle": exact,
            "cpu": {"seconds": cpu_s, "tissue_voxel_windows_per_s": cpu_tissue * T / cpu_s, "cores": threads,
                    "kind": "port",
                    "sample": f"slices {s0}-{s1 - 1}, C restatement of DoSARAverageAtCPU, OpenMP over rows"}}


def seqgen_section(sample_tr: int) -> dict:
    """The sequence compiler (host code in the library) on the 3-D GRE PSD
    of tests/golden at the benchmark's encoding (128 readout, 128 x 64
    phase encodes: 8192 TRs), against the Python oracle of DoPulseGen on
    its first TRs, with the lines of that sample compared."""
    import numpy as np
    from mrilab_amd import engine
    from oracle import seqgen
    here = os.path.dirname(os.path.abspath(__file__))
    psd = open(os.path.join(here, "tests", "golden", "psd_gre3d_small.xml")).read()
    v = dict(TR=8e-3, TE=4e-3, FlipAng=20.0, FOVFreq=0.256, FOVPhase=0.256, FOVSlice=0.128, ResFreq=129,
             ResPhase=128, SliceNum=64, FirstPhNum=128, SecondPhNum=64, BandWidth=100e3, RFreq=0.256 / 128,
             MinUpdRate=1e-6, TEAnchor="Middle", MultiTransmit="off")
    engine.compile_psd(psd, dict(v, FirstPhNum=1, SecondPhNum=1))  # library loaded outside the timer
    t0 = time.perf_counter()
    q = engine.compile_psd(psd, v)
    lib_s = time.perf_counter() - t0
    n_tr, n_ts = q["TRNum"], q["tsLine"].size
    vs = dict(v, FirstPhNum=sample_tr, SecondPhNum=1)
    t0 = time.perf_counter()
    o = seqgen.compile_psd(psd, vs)
    cpu_s = time.perf_counter() - t0
    a = engine.compile_psd(psd, vs)
    equal = all(np.array_equal(a[k], o[k]) if k in ("utsLine", "tsLine", "flagsLine", "ADCLine", "ExtLine")
                else np.allclose(a[k], o[k], rtol=1e-14, atol=1e-30) for k in engine.SEQ_LINES)
    log(f"sequence compiler: {n_tr} TRs, {n_ts} time points in {lib_s:.2f} s; oracle {cpu_s / sample_tr * 1e3:.2f} "
        f"ms per TR; sample lines equal: {equal}")
    return {"workload": f"3-D GRE PSD (tests/golden/psd_gre3d_small.xml), 128 readout x 128 x 64 phase encodes: "
                        f"{n_tr} TRs, {n_ts} time points", "seconds": lib_s, "trs_per_s": n_tr / lib_s,
            "cores": 1, "sample_lines_equal_to_oracle": bool(equal),
            "cpu": {"seconds_per_tr": cpu_s / sample_tr, "kind": "port", "cores": 1,
                    "sample": f"first {sample_tr} TRs, Python restatement of DoPulseGen and the elements"}}


def coil_sum_section(n_tr: int) -> dict:
    """Forward kernel time per launch on the first ``n_tr`` TRs of the same
    scan with each coil sum: the default MFMA projection and the wavefront
    butterfly (mrs_options.coil_sum = 1), with the signals' rel-L2."""
    import numpy as np
    from mrilab_amd import engine, workloads
    sc = workloads.bssfp3d(n_tr=n_tr)
    out, sig = {}, {}
    for name, cs in (("mfma", 0), ("butterfly", 1)):
        o = sc.copy()
        with engine.Plan(*o.structs, options=engine.make_options(timing=1, coil_sum=cs)) as p:
            p.run()
            p.sync()
            ms = []
            for _ in range(3):
                p.run()
                p.sync()
                inf = p.info()
                ms.append(inf["fwd_kernel_ms"] / inf["n_launches"])
            p.fetch(o.VObj, o.VVar, o.VSig)
        out[f"{name}_ms_per_launch"] = float(np.median(ms))
        sig[name] = np.asarray(o.VSig["Sx"]) + 1j * np.asarray(o.VSig["Sy"])
    d = sig["mfma"] - sig["butterfly"]
    out["rel_l2_between"] = float(np.linalg.norm(d) / np.linalg.norm(sig["butterfly"]))
    out["tr"] = n_tr
    log(f"coil sum: MFMA {out['mfma_ms_per_launch']:.2f} ms, butterfly {out['butterfly_ms_per_launch']:.2f} ms "
        f"per launch; rel-L2 between {out['rel_l2_between']:.1e}")
    return out


# The forward kernel's per-launch hardware counters (rocprofv3 --pmc) live in
# profiles/bloch_fwd_bssfp3d_counters.json, written by --write-profile from
# the databases of the commands it lists. The kernel's instruction counts are
# fixed by its code and by the launch (16 TRs of this workload), so they hold
# for every launch of the scan; the file records a hash of the sources the
# kernel is compiled from and is ignored when they have changed since.
COUNTERS_JSON = os.path.join(ROOT, "profiles", "bloch_fwd_bssfp3d_counters.json")
KERNEL_SOURCES = ("mrilab_amd/csrc/engine.hip", "mrilab_amd/csrc/spin_step.hpp", "mrilab_amd/csrc/device_common.hpp",
                  "mrilab_amd/csrc/decode.hpp", "mrilab_amd/csrc/decode.cpp")
FWD_KERNEL = "bloch_fwd_kernel<4, 8, false, 0, true, true>"
SIMDS = 1024                # 256 CUs x 4 SIMDs
PEAK_CLOCK_HZ = 2.4e9       # MI355X peak engine clock


def kernel_sources_hash() -> str:
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def _pmc_per_dispatch(db: str) -> dict:
    """{dispatch: (kernel name, duration ms, {counter: value summed over instances})}."""
    import sqlite3
    con = sqlite3.connect(db)
    out = {}
    for disp, name, cname, val, t0, t1 in con.execute(
            "select dispatch_id, kernel_name, counter_name, value, start, end from counters_collection"):
        k, dur, vals = out.setdefault(disp, (name, (t1 - t0) * 1e-6, {}))
        vals[cname] = vals.get(cname, 0.0) + val
    return out


def write_profile(out_path: str, dbs: list) -> None:
    """Forward-kernel counter profile from rocprofv3 databases: --pmc runs
    (any counters), a FETCH_SIZE and a WRITE_SIZE run made with
    --calibrate-copy (the 256 MiB device copy scales the byte counters), and
    a --kernel-trace run; per launch, the last launch of the kernel in each
    database (the first carries the cold caches)."""
    import sqlite3
    import statistics
    prof = {"kernel": FWD_KERNEL, "sources_sha256_16": kernel_sources_hash(),
            "workload": "3-D bSSFP 128x128x64 x 4 isochromats, 8 Rx coils; one launch = 16 TRs = 8000 steps, "
                        "2048 ADC samples, 3.355e10 isochromat-steps", "counters": {}}
    for db in dbs:
        con = sqlite3.connect(db)
        tables = {r[0] for r in con.execute("select name from sqlite_master")}
        if "counters_collection" in tables and con.execute("select count(*) from counters_collection").fetchone()[0]:
            per = _pmc_per_dispatch(db)
            fwd = [d for d in sorted(per) if FWD_KERNEL in per[d][0]]
            last = per[fwd[-1]]
            prof["counters"].update(last[2])
            prof.setdefault("pmc_launch_ms", {})[os.path.basename(db)] = last[1]
            if "GRBM_GUI_ACTIVE" in last[2]:  # summed over the 8 XCDs
                prof.setdefault("clock_hz", {})[os.path.basename(db)] = last[2]["GRBM_GUI_ACTIVE"] / 8 / (last[1] / 1e3)
            copies = [per[d] for d in sorted(per) if "copyBuffer" in per[d][0] and
                      max(per[d][2].values()) > 100000]
            for cname in ("FETCH_SIZE", "WRITE_SIZE"):
                if cname in last[2] and copies:
                    prof[cname + "_copy_kib"] = statistics.median(c[2][cname] for c in copies)
        elif "kernels" in tables:
            d = [(t1 - t0) * 1e-6 for name, t0, t1 in con.execute("select name, start, end from kernels")
                 if FWD_KERNEL in name]
            prof["trace"] = {"launches": len(d), "avg_ms": sum(d) / len(d), "median_ms": statistics.median(d),
                             "min_ms": min(d), "max_ms": max(d)}
    c = prof["counters"]
    copy_bytes = 256 << 20
    if "FETCH_SIZE" in c and "WRITE_SIZE" in c:
        ff = copy_bytes / 1024 / prof["FETCH_SIZE_copy_kib"]
        wf = copy_bytes / 1024 / prof["WRITE_SIZE_copy_kib"]
        prof["traffic"] = {"fetch_factor": ff, "write_factor": wf, "read_bytes": c["FETCH_SIZE"] * 1024 * ff,
                           "write_bytes": c["WRITE_SIZE"] * 1024 * wf}
        prof["traffic"]["bytes_per_launch"] = prof["traffic"]["read_bytes"] + prof["traffic"]["write_bytes"]
    prof["commands"] = [a for a in os.environ.get("PROFILE_COMMANDS", "").split(";;") if a]
    with open(out_path, "w") as f:
        json.dump(prof, f, indent=1)
    print(json.dumps(prof, indent=1))


def load_profile() -> dict | None:
    if not os.path.exists(COUNTERS_JSON):
        return None
    with open(COUNTERS_JSON) as f:
        prof = json.load(f)
    if prof.get("sources_sha256_16") != kernel_sources_hash():
        log("profiles/bloch_fwd_bssfp3d_counters.json was measured on other kernel sources: roofline counters "
            "omitted")
        return None
    return prof


def roofline(prof: dict | None, avg_launch_s: float, ab: float, iso_steps_per_launch: float, flop: float) -> dict:
    """The forward kernel is bound by VALU issue. Each SIMD issues at most one
    VALU instruction per 4-cycle slot, or two (dual issue, counted by
    SQ_ACTIVE_INST_VALU2); busy slots per launch = SQ_INSTS_VALU -
    SQ_ACTIVE_INST_VALU2. frac = busy slots / (SIMDs x peak clock x live
    launch time / 4); achieved and peak are the same ratio in instructions:
    the kernel's issue rate, and the rate at which every slot issues with
    this kernel's dual-issue share."""
    out = {"bound": "valu", "achieved": None, "peak": None, "unit": "G VALU wave-instructions/s", "frac": None,
           "traffic": None, "kernel": FWD_KERNEL, "avg_launch_ms": avg_launch_s * 1e3}
    alg_gbs = ab * iso_steps_per_launch / avg_launch_s / 1e9
    out["algorithmic_bytes_convention"] = {
        "bytes_per_isochromat_step": ab, "GBs": alg_gbs, "x_hbm_peak": alg_gbs / HBM_PEAK_GBS,
        "note": "bytes a step-at-a-time engine would move; the state stays in registers, so this is not HBM work"}
    out["model_flops"] = {"flop_per_isochromat_step": flop,
                          "TFLOPs": flop * iso_steps_per_launch / avg_launch_s / 1e12}
    if prof is None:
        return out
    c = prof["counters"]
    if "SQ_INSTS_VALU" in c and "SQ_ACTIVE_INST_VALU2" in c:
        insts, dual = c["SQ_INSTS_VALU"], c["SQ_ACTIVE_INST_VALU2"]
        slots = SIMDS * PEAK_CLOCK_HZ * avg_launch_s / 4
        out["achieved"] = insts / avg_launch_s / 1e9
        out["peak"] = SIMDS * PEAK_CLOCK_HZ / 4 * insts / (insts - dual) / 1e9
        out["frac"] = (insts - dual) / slots
        out["valu"] = {"instructions_per_launch": insts, "dual_issue_quad_cycles": dual,
                       "dual_issued_share": 2 * dual / insts,
                       "instructions_over_single_issue_slots": insts / slots}
        if "SQ_VALU_MFMA_BUSY_CYCLES" in c:
            # fp32 MFMAs (the coil projection) do not overlap VALU issue on
            # gfx950 (SQ_VALU_MFMA_COEXEC_CYCLES = 0): SIMD cycles spent
            # issuing VALU work or executing MFMAs, over the launch's cycles
            mfma = c["SQ_VALU_MFMA_BUSY_CYCLES"] / SIMDS
            simd = PEAK_CLOCK_HZ * avg_launch_s
            out["simd_busy"] = {"valu_frac": (insts - dual) * 4 / SIMDS / simd, "mfma_frac": mfma / simd,
                                "frac": ((insts - dual) * 4 / SIMDS + mfma) / simd,
                                "note": "at the peak clock; VALU busy slots x 4 cycles plus MFMA busy cycles, "
                                        "per SIMD, over the launch's cycles"}
    if "SQ_INSTS_VALU_FLOPS_FP32" in c:
        tf = c["SQ_INSTS_VALU_FLOPS_FP32"] * 64 / avg_launch_s / 1e12
        out["counted_fp32"] = {"TFLOPs": tf, "peak": FP32_PEAK_TFLOPS, "frac": tf / FP32_PEAK_TFLOPS}
    if "traffic" in prof:
        t = prof["traffic"]["bytes_per_launch"]
        out["traffic"] = t
        out["hbm"] = {"GBs": t / avg_launch_s / 1e9, "peak": HBM_PEAK_GBS, "frac": t / avg_launch_s / 1e9 / HBM_PEAK_GBS}
    if "trace" in prof:
        out["rocprof_trace_avg_ms"] = prof["trace"]["avg_ms"]
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tr", type=int, default=None, help="truncate the scan to its first TRs (profiling)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-tr", type=int, default=2)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched fingerprinting section")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint section")
    ap.add_argument("--no-multipool", action="store_true", help="skip the multi-pool section")
    ap.add_argument("--no-recon", action="store_true", help="skip the image reconstruction section")
    ap.add_argument("--no-sar", action="store_true", help="skip the local SAR section")
    ap.add_argument("--multipool-sat", type=float, default=2.0,
                    help="saturation train per offset of the multi-pool section, seconds")
    ap.add_argument("--multipool-timing-only", action="store_true",
                    help="multi-pool section: only the batched runs of both schemes")
    ap.add_argument("--batch-sets", type=int, default=4096)
    ap.add_argument("--batch-plain-sample", type=int, default=0,
                    help="plain per-scan calls timed for the batched speed-up (0: every scan)")
    ap.add_argument("--write-profile", nargs="+", metavar=("OUT_JSON", "DB"),
                    help="write the forward kernel's counter profile from rocprofv3 databases and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 256 MiB device copy (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.write_profile:
        write_profile(args.write_profile[0], args.write_profile[1:])
        return 0
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this engine runs the scan on one MI355X and has no "
              "multi-GPU path; run with --gpus 1", file=sys.stderr)
        return 2

    from mrilab_amd import engine, workloads

    if args.calibrate_copy:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 256 MiB device-to-device")
        del src, dst

    log("building the workload")
    t0 = time.perf_counter()
    sc = workloads.bssfp3d(n_tr=args.tr)
    log(f"{sc.name}: {sc.n_iso} isochromats, {sc.n_steps()} steps ({time.perf_counter() - t0:.1f} s)")
    opts = engine.make_options(timing=1)
    plan = engine.Plan(*sc.structs, options=opts)
    info = plan.info()
    log(f"plan: {info['n_steps']} steps, {info['n_adc']} ADC samples, {info['n_launches']} launches, "
        f"{info['n_cached_steps']} cached free steps, {info['n_rf_steps']} rf steps")
    for w in range(args.warmup):
        t = time.perf_counter()
        plan.run()
        plan.sync()
        log(f"warmup {w + 1}/{args.warmup}: {time.perf_counter() - t:.2f} s")
    plan.sync()
    fwd_ms = red_ms = 0.0
    t_start = time.perf_counter()
    for k in range(args.steps):
        plan.run()
        plan.sync()
        inf = plan.info()
        fwd_ms += inf["fwd_kernel_ms"]
        red_ms += inf["reduce_kernel_ms"]
        log(f"step {k + 1}/{args.steps}: kernels {inf['fwd_kernel_ms']:.1f} + {inf['reduce_kernel_ms']:.1f} ms")
    plan.sync()
    elapsed = time.perf_counter() - t_start
    iso_steps = info["n_iso"] * info["n_steps"]
    value = iso_steps * args.steps / elapsed
    spins, rx = int(sc.VObj["SpinNum"]), int(sc.VCoi["RxCoilNum"])
    ab = algorithmic_bytes(spins, rx, info["n_adc"] / info["n_steps"])
    n_launch = info["n_launches"]
    avg_launch_s = fwd_ms / 1e3 / (args.steps * n_launch)
    flop = algorithmic_flops(info["n_rf_steps"] / info["n_steps"], info["n_adc"] / info["n_steps"], rx)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "isochromat-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "3-D bSSFP 128x128x64 voxels x 4 isochromats, 8 Rx coils, 45 deg hard rf, "
                               "TR 5 ms = 500 x 10 us steps, 128x64 phase encodes x 128 readout",
                   "isochromats": info["n_iso"], "steps": info["n_steps"], "adc_samples": info["n_adc"],
                   "launches_per_step": n_launch, "precision_mode": "reference",
                   "free_steps": "deferred" if info["deferred_free_steps"] else "stepwise"},
        "roofline": roofline(load_profile(), avg_launch_s, ab, iso_steps / n_launch, flop),
        "kernel_ms_per_step": {"bloch_fwd": fwd_ms / args.steps, "signal_reduce": red_ms / args.steps},
    }
    if not args.no_recon and args.tr is None:
        line["recon"] = recon_section(plan, sc, 5)
    plan.close()
    if args.tr is None:
        line["coil_sum"] = coil_sum_section(64)
        line["seqgen"] = seqgen_section(64)
    if not args.no_batched:
        line["batched"] = batched_section(args.batch_sets, args.batch_plain_sample, 3)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_section(3)
    if not args.no_multipool:
        mp_threads = min(64, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        line["multipool"] = multipool_section(mp_threads, args.multipool_sat, not args.multipool_timing_only)
    if not args.no_sar:
        sar_threads = min(64, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        line["sar"] = sar_section(sar_threads)
    if not args.no_cpu_baseline:
        threads = min(64, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        log(f"cpu baseline: fp32 oracle on the first {args.cpu_sample_tr} TRs, {threads} threads")
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample_tr, threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
