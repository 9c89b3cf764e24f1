#!/usr/bin/env python3
"""Benchmark of the Bloch scan hot path on one MI355X.

Workload (one "step" = one full scan of it, inputs resident in HBM):
  3-D balanced SSFP, 128 x 128 x 64 voxels x 4 isochromats (uniform +-pi
  intravoxel spread), T1/T2/PD/dB0 seeded as the single-slice case, 8
  Gaussian receive coils, 45 deg hard rf with alternating phase, TR 5 ms
  on a 10 us grid (500 steps), 128 x 64 phase encodes x 128 readout
  samples = 8192 TRs = 4.096e6 steps = 1.72e13 isochromat-steps.

Prints one JSON line with isochromat-steps/s, the forward kernel's
roofline under the algorithmic-bytes convention (per isochromat-step:
24 B state read+write, 4 B dWRnd, 36/S B per-voxel maps, 8*Rx/S B of coil
maps on ADC steps) and the fp32 CPU oracle timed on the first TRs of the
same phantom on this host's cores.

The kernel keeps the state in registers for a whole launch, so the bytes
it really moves (rocprofv3 FETCH_SIZE + WRITE_SIZE, stored in
profiles/bloch_fwd_bssfp3d_traffic.json) are a small fraction of the
algorithmic bytes and the algorithmic-byte rate can exceed HBM peak. The
line therefore also carries the measured HBM rate and the VALU roofline:
algorithmic fp32 flops per isochromat-step = 10 (cached free-precession
rotation 6, transverse and longitudinal relaxation 4) + 35 more on rf steps
(Rodrigues rotation; plus sqrt, reciprocal and sin/cos) + 8 per receive
coil on ADC steps (complex product with the coil map and accumulation).

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
    # plain per-scan launches on a sample, end to end (what BatchSim.m does per scan)
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
    plain_total = float(np.mean(plain)) * n_sets
    batched_e2e = t_create + wall_run + t_fetch
    iso_steps = info["n_iso_steps"]
    return {"workload": f"{n_sets} single-voxel fingerprinting scans x 64 isochromats, 500-1500 TRs of 5 steps "
                        "(hard pulse, TE 3 ms, TR 10-15 ms), flips 5-70 deg",
            "isochromat_steps": iso_steps, "value": iso_steps / wall_run, "unit": "isochromat-steps/s",
            "ms_per_batch_run": wall_run * 1e3, "device_ms_per_run": float(np.median(run_ms)),
            "end_to_end_s": batched_e2e, "create_s": t_create, "decode_ms": info["decode_ms"],
            "tune_ms": info["tune_ms"], "library_create_ms": info["create_ms"], "fetch_s": t_fetch,
            "plain_per_scan_ms": float(np.mean(plain)) * 1e3, "plain_sample": int(plain_sample),
            "plain_total_s_estimated": plain_total,
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


def multipool_section(sat_split: float, sat_exact: float, cpu_shape: int, threads: int) -> dict:
    """Multi-pool exchange with MT saturation on the 256x256 3-pool phantom
    (free water, myelin water, bound pool with a super-Lorentzian line):
    the first of the 40 saturation offsets (-50 ppm), its Gaussian pulse
    train cut to ``sat_split`` s (operator split, fp32) and ``sat_exact`` s
    (fp64 matrix exponential), then readout. Device time per scan, and the
    CPU oracle's split scheme on a ``cpu_shape``^2 crop for comparison."""
    import oracle
    from mrilab_amd import engine, workloads
    out = {"workload": "256x256 voxels x 3 pools (Dirichlet(8,1,1) fractions, k~U[1,30]/s, T2b 10 us), first "
                       "saturation offset (-50 ppm at 3 T) of 300 deg Gaussian pulses (20 ms on, 20 ms off), "
                       "10 us steps"}
    for scheme, sat in ((0, sat_split), (1, sat_exact)):
        sc = workloads.mt_cest(shape=(256, 256, 1), n_offsets=1, sat_time=sat)
        o = sc.copy()
        opt = engine.make_options(multipool_scheme=scheme, timing=1, model=1)
        with engine.Plan(*o.structs, options=opt) as p:
            p.run()
            p.sync()
            t0 = time.perf_counter()
            p.run()
            p.sync()
            el = time.perf_counter() - t0
            info = p.info()
        vs = 256 * 256 * info["n_steps"]
        name = "split" if scheme == 0 else "exact"
        out[name] = {"saturation_s": sat, "steps": info["n_steps"], "ms_per_scan": el * 1e3,
                     "kernel_ms": info["fwd_kernel_ms"], "voxel_steps_per_s": vs / el,
                     # 40 offsets x a 2 s train: time grows with the train length
                     "full_config_s_estimated": 40 * (2.0 / sat) * el}
        log(f"multipool {name}: {info['n_steps']} steps, {el * 1e3:.1f} ms, {vs / el:.3e} voxel-steps/s")
    sc = workloads.mt_cest(shape=(cpu_shape, cpu_shape, 1), n_offsets=1, sat_time=sat_split)
    oracle.set_threads(threads, "f32")
    t0 = time.perf_counter()
    oracle.gm_scan(sc, precision="f32", scheme=0)
    el = time.perf_counter() - t0
    oracle.set_threads(0, "f32")
    vs = cpu_shape * cpu_shape * sc.n_steps()
    out["cpu_split"] = {"value": vs / el, "unit": "voxel-steps/s", "cores": threads, "kind": "port",
                        "sample": f"{cpu_shape}x{cpu_shape} crop, {sc.n_steps()} steps ({el:.1f} s), fp32 "
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


def traffic_from_profile() -> float | None:
    """HBM bytes per forward launch measured by rocprofv3 (profiles/)."""
    path = os.path.join(ROOT, "profiles", "bloch_fwd_bssfp3d_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return float(json.load(f)["bytes_per_launch"])


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
    ap.add_argument("--batch-sets", type=int, default=4096)
    ap.add_argument("--batch-plain-sample", type=int, default=256)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a plain 256 MiB device copy (reference for HBM byte counters)")
    args = ap.parse_args()
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
    ab_per_launch = iso_steps / n_launch * ab
    achieved = ab_per_launch / avg_launch_s / 1e9
    traffic = traffic_from_profile()
    flop = algorithmic_flops(info["n_rf_steps"] / info["n_steps"], info["n_adc"] / info["n_steps"], rx)
    tflops = iso_steps / n_launch * flop / avg_launch_s / 1e12
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
                   "launches_per_step": n_launch, "precision_mode": "reference"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "convention": f"algorithmic bytes {ab:.2f} B per isochromat-step",
                     "kernel": "bloch_fwd_kernel<4,8>", "avg_launch_ms": avg_launch_s * 1e3,
                     "measured_hbm_GBs": None if traffic is None else traffic / avg_launch_s / 1e9,
                     "valu": {"achieved": tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": tflops / FP32_PEAK_TFLOPS,
                              "frac_nonpacked": tflops / FP32_NONPACKED_PEAK_TFLOPS,
                              "flop_per_isochromat_step": flop}},
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
        line["multipool"] = multipool_section(0.2, 0.04, 64, mp_threads)
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
