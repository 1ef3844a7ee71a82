#!/usr/bin/env python3
"""NEMOLite2D time step benchmark on one MI355X.

Workload (default): 4096 x 4096 interior T points, FP64, seeded smooth
bathymetry (100 m + 50 m relief, sigma 32 cells), open west edge forced by
1.0 m * sin(2 pi t / 44712 s), dt 20 s, cbfr 1.5e-3, visc 0.1. One step is
the reference PSy layer's invoke_time_step (continuity, momentum u/v, bc_ssh,
solid and Flather boundaries, next_sshu/v, the now<-after copies), executed
by libnemo_hip.so's fused kernel plus the Flather segment kernel, replayed as
one hipGraph; inputs are resident in HBM before the timed region.

Prints ONE JSON line (rank 0, N = 1):
  value            interior grid-point updates per second (whole job)
  roofline         the fused step kernel: algorithmic bytes (212 B per update,
                   every array of the reference kernel signatures touched once)
                   over its mean duration from HIP events, against 8 TB/s
  cpu_baseline     the reference's own OpenMP PSy layer (time_step_omp.cpp,
                   built into oracle/_ref) on this host's cores, on a bounded
                   sample of the same workload; "rows" adds its serial layer
                   (time_step.cpp) pinned to one core, and both at 128^2 for
                   2000 steps; "host" names the CPU (lscpu)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = ("grid-pt updates/s + % HBM peak at 4096²; "
          "FP32/learned vs plain FP64, same err")
HBM_PEAK_GBS = 8000.0
B_ALG_F64 = 212          # bytes per interior grid-point update, FP64
B_ALG_MIXED = 113        # FP32 fields, the byte mask and the ssh compensation word
# nemo_hip_fused_variant_shape's march kind -> description
_ALIGNED = "overlapped-wavefront march, 56 columns, 64-byte-aligned stores"
MARCH_KINDS = {
    0: "row march", 1: "row march, loads one row ahead",
    2: "overlapped-wavefront march", 3: "overlapped-wavefront march, >= 3 waves per SIMD",
    4: "overlapped-wavefront march, >= 4 waves per SIMD",
    5: _ALIGNED,
    6: "overlapped-wavefront march, 48 columns, 64-byte-aligned stores",
    7: _ALIGNED + ", loads two rows ahead",
    8: _ALIGNED + ", loads two rows ahead, >= 3 waves per SIMD",
    9: "persistent chunked march, 56 columns, 64-byte-aligned stores, loads two rows ahead",
    10: "persistent chunked march, 56 columns, 64-byte-aligned stores",
    11: _ALIGNED + ", loads three rows ahead",
}
# rocprofv3 summaries holding "[label] traffic_bytes_per_launch:" lines
PROFILE_TRAFFIC = (ROOT / "profiles" / "fused_step_4096_counters.md",
                   ROOT / "profiles" / "ensemble_step_256_domains_counters.md",
                   ROOT / "profiles" / "mixed_step_4096_prefetch2_counters.md",
                   ROOT / "profiles" / "ensemble_step_256_domains_prefetch2_counters.md",
                   ROOT / "profiles" / "step_kernels_4096_traffic.md")


def parse_args():
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--n", type=int, default=4096, help="interior points per side")
    p.add_argument("--kernel-steps", type=int, default=200,
                   help="steps timed per kernel with HIP events for the roofline")
    p.add_argument("--tune-steps", type=int, default=20,
                   help="steps per fused-kernel shape in the autotune pass")
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="target CPU time of the cpu_baseline sample")
    p.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg")
    p.add_argument("--no-mixed", action="store_true", help="skip the FP32 mixed leg")
    p.add_argument("--no-ensemble", action="store_true", help="skip the ensemble leg")
    p.add_argument("--ens-domains", type=int, default=256)
    p.add_argument("--ens-steps", type=int, default=2000)
    p.add_argument("--no-closure", action="store_true", help="skip the learned-closure leg")
    p.add_argument("--closure-seeds", default="64:66",
                   help="held-out bathymetry seeds of the closure error check")
    p.add_argument("--closure-steps", type=int, default=2000)
    return p.parse_args()


def traffic_from_profile(label: str):
    """HBM bytes per launch of the kernel recorded under `label` ("variant 13",
    "mixed rows 64", "ensemble f64 4x56p1"), measured with rocprofv3 PMC counters (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration) and kept in
    profiles/; None when that kernel shape was not profiled."""
    key = f"[{label}] traffic_bytes_per_launch:"
    for path in PROFILE_TRAFFIC:
        if not path.exists():
            continue
        for line in path.read_text().splitlines():
            if line.strip().startswith(key):
                return float(line.split(":", 1)[1])
    return None


def _cpu_row(args: list[str]) -> dict:
    """One CPU timing (oracle/cpu_bench.py) in a process of its own, so that
    the serial row's core pinning and the OpenMP settings stay apart."""
    import subprocess
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", *args], cwd=ROOT,
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"cpu_bench {args} failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline(n: int, seconds: float) -> dict:
    """The reference's CPU PSy layer on this host, in four rows: OpenMP
    (time_step_omp.cpp, the reported value) and serial (time_step.cpp,
    pinned to one core) on a bounded sample of the n^2 workload, then both
    on the 128^2 configuration for its full 2000 steps. Each row's process
    records the host CPU (lscpu) and the cores it used."""
    omp = _cpu_row(["--n", str(n), "--mode", "omp", "--seconds", str(seconds)])
    serial = _cpu_row(["--n", str(n), "--mode", "serial", "--seconds", str(seconds)])
    c1_serial = _cpu_row(["--mode", "serial", "--c1", "--steps", "2000"])
    c1_omp = _cpu_row(["--mode", "omp", "--c1", "--steps", "2000"])
    rows = []
    for name, r in ((f"{n}^2 OpenMP", omp), (f"{n}^2 serial", serial),
                    ("128^2 x 2000 steps serial", c1_serial),
                    ("128^2 x 2000 steps OpenMP", c1_omp)):
        rows.append({"row": name, "value": r["value"], "cores": r["cores"],
                     "steps_timed": r["steps_timed"], "seconds": r["seconds"],
                     "code": r["code"], "pinning": r["pinning"]})
    return {
        "value": omp["value"],
        "unit": "grid-pt updates/s",
        "cores": omp["cores"],
        "kind": omp["kind"],
        "sample": (f"{omp['steps_timed']} steps of the {n}x{n} FP64 workload after 1 "
                   f"untimed step; reference manual_versions/psykal_cpp/time_step_omp.cpp, "
                   f"OpenMP {omp['cores']} threads ({omp['pinning']}), {omp['flags']}"),
        "seconds": omp["seconds"],
        "host": omp["host"],
        "rows": rows,
    }


def rel_err(x, ref):
    import numpy as np
    d = x - ref
    return (float(np.max(np.abs(d)) / max(np.max(np.abs(ref)), 1e-300)),
            float(np.linalg.norm(d) / max(np.linalg.norm(ref), 1e-300)))


def mixed_leg(pristine, a, pts, f64_ms_per_step, f64_kernel_ms, f64_variant) -> dict:
    """The FP32 mixed form on the same workload: throughput (same timing
    protocol as the FP64 line), kernel roofline against its 113 B per
    update, and its error against the FP64 form (bitwise the oracle) after
    250 / 500 / 1000 steps."""
    from psyclonebench_amd import _native
    from psyclonebench_amd.psy import MixedStepper, TimeStepper

    lib = _native.hip_lib()
    st = MixedStepper(pristine.copy())
    st.run(1 + a.warmup)
    tune = {}
    for cols in (61, 56, 48):
        lib.nemo_mp_set_columns(st.handle, cols)
        for pf in (0, 1, 2):
            for rows in ((4, 8) if pf else (4, 8, 16, 32, 48, 64, 80, 96, 128)):
                lib.nemo_mp_set_prefetch(st.handle, 0)
                lib.nemo_mp_set_rows(st.handle, rows)
                lib.nemo_mp_set_prefetch(st.handle, pf)
                tune[(rows, cols, pf)] = st.timed_run(a.tune_steps)
    rows, cols, pf = min(tune, key=tune.get)
    lib.nemo_mp_set_prefetch(st.handle, 0)
    lib.nemo_mp_set_rows(st.handle, rows)
    lib.nemo_mp_set_columns(st.handle, cols)
    lib.nemo_mp_set_prefetch(st.handle, pf)
    st.run(a.steps)
    lib.nemo_hip_synchronize()
    t0 = time.perf_counter()
    st.run(a.steps)
    lib.nemo_hip_synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = st.timed_run(a.kernel_steps)
    st.release()

    # accuracy against FP64
    g = pristine.grid
    sl = (slice(g.ys, g.ye + 1), slice(g.xs, g.xe + 1))
    lib.nemo_hip_set_fused_variant(f64_variant)
    ref = TimeStepper(pristine.copy(), mode=2)
    mp = MixedStepper(pristine.copy(), rows=rows, cols=cols, prefetch=pf)
    errs, done = {}, 0
    for cp in (250, 500, 1000):
        ref.run(cp - done)
        mp.run(cp - done)
        done = cp
        errs[str(cp)] = {n: rel_err(mp.fetch(n)[sl], ref.fetch(n)[sl])
                         for n in ("sshn_t", "un", "vn")}
    ref.release()
    mp.release()
    b_alg = B_ALG_MIXED * pts
    achieved = b_alg / (kernel_ms * 1e-3) / 1e9
    ms_step = wall * 1e3 / a.steps
    return {
        "value": pts * a.steps / wall,
        "unit": "grid-pt updates/s",
        "dtype": "f32 fields, compensated f32 ssh with f64 accumulation, f64 tide",
        "ms_per_step": ms_step,
        "speedup_vs_f64": f64_ms_per_step / ms_step,
        "kernel_speedup_vs_f64": f64_kernel_ms / kernel_ms,
        "rows_per_band": rows,
        "columns_per_wavefront": cols,
        "loads_issued_rows_ahead": pf + 1,
        "rows_kernel_ms": {f"{r}x{c}" + (f"p{q}" if q else ""): round(v, 4)
                           for (r, c, q), v in tune.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profile(
                         f"mixed rows {rows}" + ("" if cols == 61 else f" cols {cols}")
                         + ("" if not pf else " prefetch" if pf == 1 else f" prefetch {pf}")),
                     "kernel": (f"k_mixed_wave_p<{rows}, {cols}, {pf}>" if pf else
                                f"k_mixed_wave<{rows}>" if cols == 61
                                else f"k_mixed_wave_a<{rows}, {cols}>"),
                     "kernel_ms": kernel_ms,
                     "bytes_per_launch_alg": b_alg},
        "error_vs_f64": {"metric": "(relmax, relrms) over the internal T region",
                         "after_steps": errs},
    }


def closure_leg(a) -> dict:
    """The learned coarse closure (trained weights shipped with the package):
    wall time of one coarse step (1024^2, dt 80 s) plus the network's
    correction, captured as one graph (and, for comparison, issued from the
    host), against the fine FP64 run over the same 80 s (4 steps at 4096^2),
    the network's MFMA rate, and its error against the box-filtered fine run
    next to the plain coarse step's, on held-out bathymetries."""
    from psyclonebench_amd import closure_run as cr
    from psyclonebench_amd.closure import FLOP_PER_POINT

    w = cr.DEFAULT_WEIGHTS
    if not w.exists():
        return {"skipped": f"no trained weights at {w.relative_to(ROOT)}"}
    sp = cr.throughput(a.n, 4, w)
    pts = (a.n // 4) ** 2
    flops = FLOP_PER_POINT * pts
    achieved = flops / (sp["forward_ms"] * 1e-3) / 1e12
    ev = cr.evaluate(a.n, 4, cr.seeds(a.closure_seeds), a.closure_steps, 500, w,
                     log_fn=lambda *x: None)
    return {
        "workload": (f"coarse {a.n // 4}^2 (dx 4 km, dt 80 s) with the learned correction "
                     f"vs fine {a.n}^2 FP64 (dx 1 km, dt 20 s); seeds {a.closure_seeds} "
                     f"held out, {a.closure_steps} coarse steps from rest"),
        "speedup_vs_fine": sp["speedup_vs_fine"],
        "ms_per_coarse_step": sp["learned_ms_per_coarse_step"],
        "ms_per_coarse_step_host_issued": sp["learned_stepwise_ms_per_coarse_step"],
        "plain_ms_per_coarse_step": sp["plain_ms_per_coarse_step"],
        "fine_ms_per_80s": sp["fine_ms_per_coarse_interval"],
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": achieved / 2500.0,
                     "kernel": "k_conv_in + k_conv_sh x2 + k_conv_out (4 layers, bf16 MFMA)",
                     "forward_ms": sp["forward_ms"], "flop_per_forward_alg": flops},
        "error_vs_fine": {"metric": "relative RMS over the interior vs the box-filtered "
                                    "fine state, mean over every 500 steps",
                          "ratio_learned_over_plain_mean": ev["ratio_mean"],
                          "ratio_learned_over_plain_final": ev["ratio_final"],
                          "per_seed": {str(k): {"plain_mean": v["mean"]["plain"],
                                                "learned_mean": v["mean"]["learned"],
                                                "guard_trips": v["guard_trips"]}
                                       for k, v in ev["seeds"].items()}},
    }


def ensemble_leg(a) -> dict:
    """The batched ensemble: --ens-domains domains (seed 0, sides
    log-uniform in [64, 1024], own bathymetry and tide), --ens-steps steps
    per precision in one graph replay after an autotune of the launch shape
    (the default shape unless another is more than 1.5% faster); kernel
    roofline from HIP events; FP32 (mixed form) error against FP64 after
    the same steps, worst domain per field."""
    from psyclonebench_amd.ensemble import EnsembleStepper, ensemble_configs
    from psyclonebench_amd.model import Model

    cfgs = ensemble_configs(a.ens_domains, seed=0, nsteps=a.ens_steps)
    models = [Model(c) for c in cfgs]
    out = {"workload": (f"{a.ens_domains} domains, seed 0, interior sides log-uniform in "
                        "[64,1024], mean depth U[50,500] m +-50% relief, tide amplitude "
                        "U[0.5,2] m, period U[40000,46000] s, open west edge"),
           "steps": a.ens_steps}
    fields = ("sshn_t", "un", "vn")
    finals = {}
    for prec, b_alg in ((64, B_ALG_F64), (32, B_ALG_MIXED)):
        ens = EnsembleStepper(models, precision=prec)
        pts = ens.interior_points
        # the library's default shape stays unless another beats it by more
        # than the 1.5% that 5-step timings vary by
        default_shape = ens.shape
        best, tune_ms = ens.autotune(5)
        if tune_ms[default_shape] <= 1.015 * tune_ms[best]:
            best = default_shape
            ens.set_shape(best)
        shape = ens.shapes()[best]
        ens.run(1 + a.warmup)
        ens.synchronize()
        t0 = time.perf_counter()
        ens.run(a.ens_steps)
        ens.synchronize()
        wall = time.perf_counter() - t0
        kernel_ms = ens.timed_run(min(a.kernel_steps, 100))
        achieved = b_alg * pts / (kernel_ms * 1e-3) / 1e9
        out[f"f{prec}"] = {
            "value": pts * a.ens_steps / wall,
            "unit": "grid-pt updates/s",
            "interior_points": pts,
            "ms_per_step": wall * 1e3 / a.ens_steps,
            "shape": {"rows": shape[0], "columns": shape[3], "prefetch": shape[4],
                      "launch": "persistent" if shape[2] else "one wavefront per tile",
                      "waves_per_workgroup": shape[1], "workgroups": shape[2] or None},
            "autotune_kernel_ms": [round(x, 4) for x in tune_ms],
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": (traffic_from_profile(
                             f"ensemble f{prec} {shape[0]}x{shape[3]}p{shape[4]}")
                             if shape[2] == 0 else None),
                         "kernel": (f"k_ens_{'f64' if prec == 64 else 'mixed'}"
                                    f"{'' if shape[2] else '_grid'}"
                                    f"<{shape[0]}, {shape[3]}, {shape[4]}>"),
                         "kernel_ms": kernel_ms, "bytes_per_launch_alg": b_alg * pts},
        }
        ens.release()
        # accuracy run: the same number of steps from the initial state
        ens = EnsembleStepper(models, precision=prec)
        ens.set_shape(best)
        ens.run(a.ens_steps)
        finals[prec] = [{n: ens.fetch(d, n)[m.grid.ys:m.grid.ye + 1, m.grid.xs:m.grid.xe + 1]
                         for n in fields} for d, m in enumerate(models)]
        ens.release()
    worst = {n: [0.0, 0.0] for n in fields}
    for d in range(len(models)):
        for n in fields:
            e = rel_err(finals[32][d][n], finals[64][d][n])
            worst[n] = [max(worst[n][0], e[0]), max(worst[n][1], e[1])]
    out["f32_error_vs_f64"] = {"metric": "(relmax, relrms) over the internal T region, "
                                         "worst domain", "after_steps": a.ens_steps,
                               "fields": worst}
    return out


def main():
    a = parse_args()
    if a.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {a.gpus} is not supported "
              "(the step would shard by 2-D domain decomposition with depth-1 halos; "
              "not built)", file=sys.stderr)
        sys.exit(2)

    from psyclonebench_amd import _native
    from psyclonebench_amd.model import Model, c2_config
    from psyclonebench_amd.psy import TimeStepper

    lib = _native.hip_lib()
    ndev, info = _native.device_info()
    if ndev < 1:
        raise SystemExit(f"bench.py: no GPU visible ({info})")

    cfg = c2_config(a.n, a.warmup + a.steps, seed=0)
    pristine = Model(cfg)          # initial state, copied for every leg
    model = pristine.copy()
    g = model.grid
    pts = (g.M - 2) * (g.N - 2)
    st = TimeStepper(model, mode=2)
    # first call uploads, pairs the ping-pong buffers; then warmup (captures
    # the graph of the timed length once, untimed)
    st.run(1 + a.warmup)
    st.synchronize()
    # autotune the fused kernel shape on this workload (HIP-event timed); the
    # library default stays unless a shape beats it by more than the 1.5%
    # that 20-step timings vary by
    default_variant = lib.nemo_hip_fused_variant()
    tune = {}
    for v in range(-1, lib.nemo_hip_fused_variant_count()):
        lib.nemo_hip_set_fused_variant(v)
        lib.nemo_hip_set_kernel_timing(1)
        st.run(a.tune_steps)
        tune[v] = lib.nemo_hip_fused_kernel_ms()
        lib.nemo_hip_set_kernel_timing(0)
    best = min(tune, key=tune.get)
    if tune[default_variant] <= 1.015 * tune[best]:
        best = default_variant
    lib.nemo_hip_set_fused_variant(best)
    bw, rows, pf = C.c_int(), C.c_int(), C.c_int()
    lib.nemo_hip_fused_variant_shape(best, C.byref(bw), C.byref(rows), C.byref(pf))
    st.run(a.steps)             # capture + first replay of a K-step graph
    st.synchronize()
    t0 = time.perf_counter()
    st.run(a.steps)
    st.synchronize()
    wall = time.perf_counter() - t0
    updates_per_s = pts * a.steps / wall

    # roofline: mean duration of the fused kernel from HIP events
    lib.nemo_hip_set_kernel_timing(1)
    st.run(a.kernel_steps)
    kernel_ms = lib.nemo_hip_fused_kernel_ms()
    lib.nemo_hip_set_kernel_timing(0)
    st.release()
    bytes_alg = B_ALG_F64 * pts
    achieved = bytes_alg / (kernel_ms * 1e-3) / 1e9
    traffic = traffic_from_profile(f"variant {best}")

    out = {
        "metric": METRIC,
        "value": updates_per_s,
        "unit": "grid-pt updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"NEMOLite2D FP64 step, {a.n}x{a.n} interior T points, seeded "
                         "bathymetry (seed 0), open west edge with 1 m tide, fused "
                         "kernel + Flather segments, hipGraph replay"),
            "interior": [g.M - 2, g.N - 2],
            "alloc": [g.W, g.H],
            "device": info,
            "hbm_frac_whole_step": B_ALG_F64 * updates_per_s / (HBM_PEAK_GBS * 1e9),
            "fused_variant": best,
            "fused_shape": {"block_columns": bw.value, "rows": rows.value,
                            "kind": MARCH_KINDS[pf.value]},
            "variant_kernel_ms": {str(k): round(v, 4) for k, v in tune.items()},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": lib.nemo_hip_fused_variant_kernel(best).decode(),
            "kernel_ms": kernel_ms,
            "bytes_per_launch_alg": bytes_alg,
        },
    }
    if not a.no_mixed:
        out["mixed"] = mixed_leg(pristine, a, pts, wall * 1e3 / a.steps, kernel_ms, best)
    if not a.no_closure:
        out["closure"] = closure_leg(a)
    if not a.no_ensemble:
        out["ensemble"] = ensemble_leg(a)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.n, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
