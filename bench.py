#!/usr/bin/env python3
"""RK4 steps/s at 256x256x129 on one MI355X: mixed precision (headline) and FP64.

Workload: constant stratification, Lx = Ly = 200 km, Lz = 1300 m,
N0 = 5.2e-3 s^-1, latitude 33 deg, Nj = 85 (2/3 rule), nonhydrostatic,
antialiased; seeded synthetic flow (|k|^-2 amplitudes, j <= 64, half wave /
half geostrophic energy, rms horizontal speed 5 cm/s); fixed dt = 0.5 CFL.
A "step" is one classic RK4 step (four nonlinear-flux evaluations plus the
stage combinations) with the state resident in HBM.

Prints one JSON line.  value = mixed-precision RK4 steps/s, the median of
`--runs` timed repetitions of exactly `--steps` steps each (after `--warmup`
untimed steps); the FP64 rate (same protocol), the mixed-vs-FP64 relative
total-energy error after the timed steps and its maximum over every step of
an untimed lockstep run (`--energy-steps`, 200 by default), the
reference-style relative error max|delta| / max|FP64| of Ap, Am, A0 and of
the APV field at the end of that run, the dominant kernel's roofline (HIP events, live) and the CPU
oracle baseline ride along.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wave-vortex-model_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

METRIC = "RK4 steps/s at 256×256×129: mixed vs FP64 GPU at ≤1e-6 rel energy error"
HBM_PEAK_GBS = 8000.0
# HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE passes
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_rk4_256x256x129_mixed.md")
# per-dispatch SQ counters (MFMA instructions among them) of the same kernels
COUNTER_SUMMARY = os.path.join(ROOT, "profiles", "sq_counters_rk4_256x256x129_mixed.md")
MFMA_PEAK_TFLOPS_F32 = 157.3
MFMA_PEAK_TFLOPS_F16 = 2500.0  # dense FP16
MIXED = "mixed-split16"        # the headline mixed-precision form: split-FP16 vertical GEMMs
VALU_SIMDS = 256 * 4          # CUs x SIMDs
PEAK_CLOCK_HZ = 2.4e9         # peak engine clock
# stage name -> kernel of the mixed-precision RHS at 256x256x129
STAGE_KERNEL = {"fields": "k_fields<float", "synth": "k_synth16", "yfft_inv": "k_yfft<float, 8, 4, true",
                "xfused": "k_xfused<float", "yfft_fwd": "k_yfft<float, 8, 4, false", "proj": "k_proj16",
                "stage": "k_stage<float"}


def per_kernel_rooflines(rec):
    """Every kernel of the mixed RHS against the chip's peaks: algorithmic
    bytes per launch over its HIP-event time against HBM, its measured HBM
    bytes per dispatch, and for the MFMA kernels the executed MFMA rate
    (FP16 and FP32 MFMA operations per dispatch from the counter summary,
    SQ_INSTS_VALU_MFMA_MOPS_* x 512) against the dense peak of its type."""
    from wvamd.rocprof_summary import hbm_traffic, kernel_counters
    out = []
    for r in rec:
        avg_ms = r["total_ms"] / max(r["launches"], 1)
        e = {"kernel": r["name"], "avg_ms": round(avg_ms, 4), "algorithmic_MB": round(r["bytes_per_launch"] / 1e6, 1),
             "hbm_GBs": round(r["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9, 1)}
        e["hbm_frac"] = round(e["hbm_GBs"] / HBM_PEAK_GBS, 4)
        prefix = STAGE_KERNEL.get(r["name"], "?")
        t = hbm_traffic(TRAFFIC_SUMMARY, prefix)
        e["traffic_MB"] = round(t["traffic_bytes"] / 1e6, 1) if t else None
        c = kernel_counters(COUNTER_SUMMARY, prefix)
        for key, peak in (("SQ_INSTS_VALU_MFMA_MOPS_F16", MFMA_PEAK_TFLOPS_F16),
                          ("SQ_INSTS_VALU_MFMA_MOPS_F32", MFMA_PEAK_TFLOPS_F32)):
            if c and c.get(key):
                tflops = c[key] * 512 / (avg_ms * 1e-3) / 1e12
                e["mfma_TFLOPs"] = round(tflops, 1)
                e["mfma_type"] = key.rsplit("_", 1)[1].lower()
                e["mfma_frac"] = round(tflops / peak, 4)
        out.append(e)
    return out


def step_model_bytes(Nx, Ny, Nz, M, b_spec, b_real):
    """Reference dataflow bytes of one nonhydrostatic RK4 step: 4 fluxes of
    72h + 47r + 71.5m plus 25 unfused passes over the 3-array state."""
    h = (Nx // 2 + 1) * Ny * Nz * b_spec
    r = Nx * Ny * Nz * b_real
    m = 16 * M
    return 4 * (72 * h + 47 * r + 71.5 * m) + 25 * 3 * m


def workload_config():
    import wvamd
    return wvamd.make_config(256, 256, 129, 85, Lx=200e3, Ly=200e3, Lz=1300.0, N0=5.2e-3, latitude=33.0)


def initial_state(k, seed=4001, jmax=64, urms=0.05):
    from wvamd import flows
    md = k.modes()
    Ap, Am, A0 = flows.random_flow_shape(md["Kh"], k.Nj, k.coefficient_table("ApmTE"), k.coefficient_table("A0TE"),
                                         seed=seed, jmax=jmax)
    f = k.transform_wave_vortex_to_uvw(*[k.to_device(a) for a in (Ap, Am, A0)]).cpu().numpy()
    s = flows.scale_factor_for_rms(f, k.config["Lz"], urms)
    Ap, Am, A0 = Ap * s, Am * s, A0 * s
    f = k.transform_wave_vortex_to_uvw(*[k.to_device(a) for a in (Ap, Am, A0)]).cpu().numpy()
    dt = flows.cfl_time_step(f, md["k"], md["l"], k.coefficient_table("omega"), k.config["Lz"], 0.5)
    return (Ap, Am, A0), dt


def alternating_steps(kernels, state, dt, steps, warmup, runs):
    """timed_steps for several kernels with their repetitions interleaved
    (kernel a run 1, kernel b run 1, ..., kernel a run 2, ...), so that a
    drift of the clocks over the benchmark affects each alike; returns the
    seconds of each run per kernel."""
    import torch
    devs, ts = [], []
    for k in kernels:
        dev = [k.to_device(a) for a in state]
        ts.append(k.rk4_advance(*dev, t=0.0, h=dt, nsteps=warmup) if warmup else 0.0)
        devs.append(dev)
    torch.cuda.synchronize()
    secs = [[] for _ in kernels]
    for _ in range(runs):
        for i, k in enumerate(kernels):
            t0 = time.perf_counter()
            ts[i] = k.rk4_advance(*devs[i], t=ts[i], h=dt, nsteps=steps)
            torch.cuda.synchronize()
            secs[i].append(time.perf_counter() - t0)
    return secs, devs


def timed_steps(k, state, dt, steps, warmup, runs=1):
    """Seconds of each of `runs` repetitions of exactly `steps` RK4 steps
    (the state carries on from one repetition to the next), after `warmup`
    untimed steps; returns (seconds per run, device state)."""
    import torch
    dev = [k.to_device(a) for a in state]
    t = k.rk4_advance(*dev, t=0.0, h=dt, nsteps=warmup) if warmup else 0.0
    torch.cuda.synchronize()
    secs = []
    for _ in range(runs):
        t0 = time.perf_counter()
        t = k.rk4_advance(*dev, t=t, h=dt, nsteps=steps)
        torch.cuda.synchronize()
        secs.append(time.perf_counter() - t0)
    return secs, dev


def lockstep_errors(k32, k64, state, dt, steps):
    """Mixed and FP64 runs advanced one step at a time from the same state:
    max over steps of |E_mixed - E_fp64| / E_fp64, and max|delta| / max|FP64|
    of Ap, Am, A0 and of the APV field (formed by the FP64 kernel from each
    final state) after the last step (the reference's relative-error form)."""
    import torch
    d32 = [k32.to_device(a) for a in state]
    d64 = [k64.to_device(a) for a in state]
    t32 = t64 = 0.0
    worst = 0.0
    for _ in range(steps):
        t32 = k32.rk4_advance(*d32, t=t32, h=dt, nsteps=1)
        t64 = k64.rk4_advance(*d64, t=t64, h=dt, nsteps=1)
        e32, e64 = k64.total_energy(*d32), k64.total_energy(*d64)
        worst = max(worst, abs(e32 - e64) / abs(e64))
    rel = {name: float((a - b).abs().max() / b.abs().max()) for name, a, b in zip(("Ap", "Am", "A0"), d32, d64)}
    a32, a64 = k64.apv(*d32, t=t32), k64.apv(*d64, t=t64)
    rel["APV"] = float((a32 - a64).abs().max() / a64.abs().max())
    torch.cuda.synchronize()
    return worst, rel


def stage_profile(k, state, dt, steps=2):
    """Per-kernel device time from HIP events recorded on the launch stream."""
    import torch
    dev = [k.to_device(a) for a in state]
    k.rk4_advance(*dev, t=0.0, h=dt, nsteps=1)
    k.reset_stage_timings()
    k.set_stage_instrumentation(True)
    k.rk4_advance(*dev, t=0.0, h=dt, nsteps=steps)
    torch.cuda.synchronize()
    rec = k.stage_timings()
    k.set_stage_instrumentation(False)
    return rec


def host_flux_rate(k, state, calls=5):
    """Flux evaluations/s through the host-memory entry point (copies in and
    out over PCIe each call, as the reference's host-memory contract does)."""
    k.nonlinear_flux_host(*state, t=0.0)
    t0 = time.perf_counter()
    for _ in range(calls):
        k.nonlinear_flux_host(*state, t=0.0)
    return calls / (time.perf_counter() - t0)


def cpu_baseline(cfg, state, dt, threads, steps=4):
    """The reference's CPU path where its kernel is built (oracle/_ref): one
    classic RK4 step of the reference kernel (CompiledKernel compiled
    unchanged, over the build-authored radix-2 OpenMP FFT engine; FFTW, its
    production engine, is not installed) -- four nonlinearFlux calls and
    WVFixedStepRK4's stage arithmetic (WVRungeKutta.cpp:335-462) in numpy.
    Otherwise the oracle restatement's RK4 steps."""
    import numpy as np
    from oracle import ReferenceFast
    if ReferenceFast.available():
        ReferenceFast.set_threads(threads)
        r = ReferenceFast(cfg)
        y = [np.asarray(a, np.complex128).reshape(r.Nkl, r.Nj) for a in state]
        r.nonlinear_flux(*y, t=0.0)  # warm-up: first touch of the scratch arena
        t0 = time.perf_counter()
        h, t = dt, 0.0
        k1 = r.nonlinear_flux(*y, t=t)
        k2 = r.nonlinear_flux(*[a + (h / 2) * b for a, b in zip(y, k1)], t=t + h / 2)
        k3 = r.nonlinear_flux(*[a + (h / 2) * b for a, b in zip(y, k2)], t=t + h / 2)
        k4 = r.nonlinear_flux(*[a + h * b for a, b in zip(y, k3)], t=t + h)
        y = [a + (h / 6) * (((b1 + 2 * b2) + 2 * b3) + b4) for a, b1, b2, b3, b4 in zip(y, k1, k2, k3, k4)]
        sec = time.perf_counter() - t0
        return {"value": round(1.0 / sec, 5), "unit": "steps/s", "cores": threads, "kind": "reference",
                "sample": f"1 FP64 RK4 step (4 nonlinearFlux calls) of the reference kernel at 256x256x129 "
                          f"from the benchmark state, {sec:.1f} s; engine: build-authored radix-2 FFTs on "
                          f"{threads} OpenMP threads (FFTW, the reference's production engine, is not installed)"}
    from oracle import Oracle
    Oracle.set_threads(threads)
    o = Oracle(cfg)
    t0 = time.perf_counter()
    o.rk4(*state, t=0.0, dt=dt, nsteps=steps)
    sec = time.perf_counter() - t0
    return {"value": round(steps / sec, 5), "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} FP64 RK4 steps (4 fluxes each) of the oracle at 256x256x129 from the benchmark "
                      f"state, {sec:.1f} s"}


def scalar_advection_leg(k64, calls=20, cpu=False):
    """advectFGridScalar (FP64, antialiased) at the workload's grid: GPU time
    per call with the inputs resident, against its algorithmic bytes (the
    scalar, the three velocities and the result, 8 B each per point); with
    `cpu`, one call of the numpy restatement on the host (the CPU reference
    is the same algorithm in numpy/FFTW-style transforms)."""
    import numpy as np
    import torch
    Nz, Ny, Nx = k64.spatial_shape
    rng = np.random.default_rng(5)
    scalar_h = rng.standard_normal((Nz, Ny, Nx))
    uvw_h = rng.standard_normal((3, Nz, Ny, Nx)) * 0.05
    scalar = torch.as_tensor(scalar_h).to(f"cuda:{k64.device}")
    uvw = torch.as_tensor(uvw_h).to(f"cuda:{k64.device}")
    out = torch.empty_like(scalar)
    k64.advect_f_grid_scalar(scalar, uvw, True, out=out)  # the first call allocates the path's scratch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(calls):
        k64.advect_f_grid_scalar(scalar, uvw, True, out=out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / calls * 1e3
    alg = 5 * 8 * Nx * Ny * Nz
    res = {"precision": "fp64", "antialias": True, "calls": calls, "ms_per_call": round(ms, 4),
           "algorithmic_MB": round(alg / 1e6, 1), "hbm_GBs": round(alg / (ms * 1e-3) / 1e9, 1)}
    if cpu:
        from oracle import Oracle
        from scalar_advection import advect_f_grid_scalar
        rows = Oracle(k64.config).modes()["row"]
        t0 = time.perf_counter()
        advect_f_grid_scalar(k64.config, scalar_h, uvw_h, True, rows)
        res["cpu_numpy_s_per_call"] = round(time.perf_counter() - t0, 3)
    return res


def c3_leg(wvamd, steps=10, runs=3):
    """FP64 and mixed RK4 steps/s on the arbitrary-stratification workload,
    the mixed-vs-FP64 total-energy difference after the timed steps, and the
    per-stage times of one mixed step."""
    import numpy as np
    from wvamd.workloads import arbitrary_stratification_512
    cfg, ops, k64, A, dt = arbitrary_stratification_512(wvamd)
    k32 = wvamd.WVTransformHydrostaticKernel(cfg, ops, "mixed")
    s64, d64 = timed_steps(k64, A, dt, steps, 1, runs)
    s32, d32 = timed_steps(k32, A, dt, steps, 1, runs)
    e64, e32 = k64.total_energy(*d64), k64.total_energy(*d32)
    rec = stage_profile(k32, A, dt, steps=1)
    sps64, sps32 = steps / float(np.median(s64)), steps / float(np.median(s32))
    return {"workload": "hydrostatic N^2 = N0^2 exp(2z/b), 512x512, 65 points, 64 modes, Lx 400 km, Lz 4000 m",
            "data": "synthetic (seeded |k|^-2 waves + |k|^-3 geostrophic, rms u 0.1 m/s, dt = 0.5 CFL)",
            "dt_s": dt, "Nkl": k64.Nkl, "steps": steps, "runs": runs,
            "fp64_steps_per_s": round(sps64, 2), "mixed_steps_per_s": round(sps32, 2),
            "mixed_over_fp64": round(sps32 / sps64, 3),
            "energy_rel_err_mixed_vs_fp64": abs(e32 - e64) / abs(e64),
            "stage_ms_per_launch_mixed": {r["name"]: round(r["total_ms"] / max(r["launches"], 1), 4) for r in rec}}


def runtime_length_leg(wvamd, n=240, steps=10, runs=3):
    """FP64 and mixed RK4 steps/s at a horizontal length that is not a power
    of two (240 = 15 x 16: runtime-length in-LDS FFTs), same vertical grid,
    flow and CFL step as the headline workload, beside the headline rate per
    grid point."""
    import numpy as np
    cfg = wvamd.make_config(n, n, 129, 85, Lx=200e3 * n / 256, Ly=200e3 * n / 256, Lz=1300.0, N0=5.2e-3,
                            latitude=33.0)
    out = {"workload": f"wave-vortex nonlinear flux + fixed RK4, {n}x{n}x129 nonhydrostatic constant N",
           "steps": steps, "runs": runs}
    for prec in ("fp64", "mixed-split16"):
        k = wvamd.WVTransformConstantStratificationKernel(cfg, prec, 0)
        state, dt = initial_state(k)
        secs, _ = timed_steps(k, state, dt, steps, 1, runs)
        sps = steps / float(np.median(secs))
        key = "fp64" if prec == "fp64" else "mixed"
        out[f"{key}_steps_per_s"] = round(sps, 2)
        out[f"{key}_Mpoint_steps_per_s"] = round(sps * n * n * 129 / 1e6, 1)
        del k
    out["mixed_over_fp64"] = round(out["mixed_steps_per_s"] / out["fp64_steps_per_s"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--runs", type=int, default=5, help="timed repetitions of --steps steps (median reported)")
    ap.add_argument("--energy-steps", type=int, default=200, help="steps of the untimed lockstep error run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--profile", choices=["mixed", "mixed-fp32", "fp64", "both"], default=None,
                    help="run only `steps` RK4 steps of the workload plus a 512 MiB copy kernel "
                         "(the byte-counter calibration), for rocprofv3")
    ap.add_argument("--no-c3", action="store_true", help="skip the arbitrary-stratification (512x512x65) and runtime-length (240x240x129) legs")
    ap.add_argument("--profile-c3", choices=["mixed", "fp64"], default=None,
                    help="run only `steps` RK4 steps of the 512x512x65 arbitrary-stratification workload")
    ap.add_argument("--profile-scalar", action="store_true",
                    help="run only five F-grid scalar advection calls at the workload's grid, for rocprofv3")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import wvamd

    cfg = workload_config()
    if args.profile_c3:
        from wvamd.workloads import arbitrary_stratification_512
        ccfg, ops, k64, A, cdt = arbitrary_stratification_512(wvamd)
        k = k64 if args.profile_c3 == "fp64" else wvamd.WVTransformHydrostaticKernel(ccfg, ops, "mixed")
        secs, _ = timed_steps(k, A, cdt, args.steps, args.warmup)
        print(f"c3 {args.profile_c3}: {args.steps / secs[0]:.2f} steps/s", flush=True)
        return
    if args.profile_scalar:
        res = scalar_advection_leg(wvamd.WVTransformConstantStratificationKernel(cfg, "fp64"), calls=5)
        print(f"scalar advection: {res['ms_per_call']} ms per call", flush=True)
        return
    if args.profile:
        kk = wvamd.WVTransformConstantStratificationKernel(cfg, "fp64")
        state, dt = initial_state(kk)
        names = {"mixed": MIXED, "mixed-fp32": "mixed", "fp64": "fp64"}
        for prec in (("mixed", "fp64") if args.profile == "both" else (args.profile,)):
            k = kk if prec == "fp64" else wvamd.WVTransformConstantStratificationKernel(cfg, names[prec])
            secs, _ = timed_steps(k, state, dt, args.steps, args.warmup)
            print(f"{prec}: {args.steps / secs[0]:.2f} steps/s", flush=True)
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.mul(a, 2.0, out=b)  # calibration kernel: reads and writes exactly 512 MiB
        torch.cuda.synchronize()
        print("copy calibration: 536870912 bytes read, 536870912 bytes written", flush=True)
        return
    k64 = wvamd.WVTransformConstantStratificationKernel(cfg, "fp64")
    k32 = wvamd.WVTransformConstantStratificationKernel(cfg, MIXED)
    kf32 = wvamd.WVTransformConstantStratificationKernel(cfg, "mixed")  # FP32-MFMA vertical GEMMs, for comparison
    state, dt = initial_state(k64)

    runs = max(1, args.runs)
    (secs32, secs64, secsf), (d32, d64, df) = alternating_steps((k32, k64, kf32), state, dt, args.steps,
                                                                args.warmup, runs)
    e32 = k64.total_energy(*d32)
    e64 = k64.total_energy(*d64)
    energy_err = abs(e32 - e64) / abs(e64)
    sec32 = float(np.median(secs32))
    sec64 = float(np.median(secs64))
    sps32 = args.steps / sec32
    sps64 = args.steps / sec64
    spsf = args.steps / float(np.median(secsf))
    ef = abs(k64.total_energy(*df) - e64) / abs(e64)
    del d32, d64, df, kf32
    worst_energy, coef_err = lockstep_errors(k32, k64, state, dt, args.energy_steps) if args.energy_steps > 0 \
        else (None, None)

    host32 = host_flux_rate(k32, state)
    host64 = host_flux_rate(k64, state)
    rec = stage_profile(k32, state, dt)
    dom = max(rec, key=lambda r: r["total_ms"])
    avg_ms = dom["total_ms"] / max(dom["launches"], 1)
    achieved = dom["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
    from wvamd.rocprof_summary import hbm_traffic
    measured = hbm_traffic(TRAFFIC_SUMMARY, STAGE_KERNEL.get(dom["name"], "?"))
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(measured["traffic_bytes"]) if measured else None,
                "kernel": dom["name"], "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes_per_launch": round(dom["bytes_per_launch"])}
    # What limits that kernel: its VALU issue rate, against one wave64 VALU
    # instruction per 4 cycles per SIMD (1024 SIMDs) at the 2.4 GHz peak
    # engine clock, from the per-dispatch instruction count of the counter
    # summary; the HBM fraction above is the roofline the contract names.
    from wvamd.rocprof_summary import kernel_counters
    c = kernel_counters(COUNTER_SUMMARY, STAGE_KERNEL.get(dom["name"], "?"))
    if c and c.get("SQ_INSTS_VALU"):
        roofline["limiter"] = "valu"
        roofline["valu_issue_frac"] = round(c["SQ_INSTS_VALU"] * 4 / (VALU_SIMDS * avg_ms * 1e-3 * PEAK_CLOCK_HZ), 4)
    b32 = step_model_bytes(cfg["Nx"], cfg["Ny"], cfg["Nz"], k32.Nkl * k32.Nj, 8, 4)
    b64 = step_model_bytes(cfg["Nx"], cfg["Ny"], cfg["Nz"], k64.Nkl * k64.Nj, 16, 8)
    # The reference's unfused dataflow (per-pass arrays read and written) at
    # this step rate: not a roofline fraction -- above 1 means fusion avoided
    # that many of the reference dataflow's bytes.
    step_model = {"mixed_bytes_per_step": b32, "mixed_equivalent_GBs": round(b32 * sps32 / 1e9, 1),
                  "mixed_over_hbm_peak": round(b32 * sps32 / 1e9 / HBM_PEAK_GBS, 4),
                  "fp64_bytes_per_step": b64, "fp64_equivalent_GBs": round(b64 * sps64 / 1e9, 1),
                  "fp64_over_hbm_peak": round(b64 * sps64 / 1e9 / HBM_PEAK_GBS, 4)}
    stages = {r["name"]: round(r["total_ms"] / max(r["launches"], 1), 4) for r in rec}
    kernels = per_kernel_rooflines(rec)

    scalar_leg = scalar_advection_leg(k64, cpu=not args.no_cpu_baseline)
    c3 = None
    if not args.no_c3:
        del k32
        c3 = c3_leg(wvamd)
    rt = None if args.no_c3 else runtime_length_leg(wvamd)
    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(cfg, state, dt, threads)

    out = {
        "metric": METRIC, "value": round(sps32, 3), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * sec32 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 fields/FFTs, split-f16 MFMA vertical GEMMs (f32 accumulate), f64 coefficients+RK4",
        "data": "synthetic (seeded |k|^-2 flow, rms u 5 cm/s, dt = 0.5 CFL)",
        "config": {"workload": "wave-vortex nonlinear flux + fixed RK4, 256x256x129 nonhydrostatic constant N",
                   "Lx_m": 200e3, "Lz_m": 1300.0, "Nj": 85, "Nkl": k64.Nkl, "dt_s": dt},
        "fp64_steps_per_s": round(sps64, 3), "fp64_ms_per_step": round(1e3 * sec64 / args.steps, 4),
        "mixed_over_fp64": round(sps32 / sps64, 3), "energy_rel_err_mixed_vs_fp64": energy_err,
        "runs": runs, "mixed_steps_per_s_runs": [round(args.steps / s, 2) for s in secs32],
        "fp64_steps_per_s_runs": [round(args.steps / s, 2) for s in secs64],
        "mixed_fp32_mfma": {"steps_per_s": round(spsf, 3), "over_fp64": round(spsf / sps64, 3),
                            "energy_rel_err_vs_fp64": ef,
                            "note": "the same mixed path with the vertical GEMMs on FP32 MFMA"},
        "lockstep": {"steps": args.energy_steps, "energy_rel_err_max_over_steps": worst_energy,
                     "coef_rel_err_max_abs": coef_err,
                     "note": "coef_rel_err_max_abs.APV: the APV field (eta_true = eta, constant-N background) "
                             "of the mixed and FP64 final states, both formed by the FP64 kernel"},
        "roofline": roofline, "reference_dataflow_bytes_saved": step_model,
        "host_memory_flux_per_s": {"mixed": round(host32, 2), "fp64": round(host64, 2),
                                   "note": "PCIe-inclusive: state copied in and flux copied out per call"},
        "stage_ms_per_launch_mixed": stages, "kernels_mixed": kernels, "scalar_advection": scalar_leg,
        "arbitrary_stratification_512x512x65": c3, "runtime_length_240x240x129": rt, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
