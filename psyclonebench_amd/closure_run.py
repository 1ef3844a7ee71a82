"""Training and evaluation of the learned coarse closure on the GPU.

    python -m psyclonebench_amd.closure_run train --n 4096 --bathy 0:64 --steps 500 \\
        --seed 1 --out psyclonebench_amd/data/closure_weights.npz
    python -m psyclonebench_amd.closure_run eval --n 4096 --bathy 64:80 --steps 2000 \\
        --weights psyclonebench_amd/data/closure_weights.npz

Training: for every bathymetry seed (the fine configuration is the 4096^2
FP64 workload with that seed: depth 100 m + 50 m relief, open west edge, 1 m
tide, dt 20 s, dx 1 km), the fine model first spins up for a random
0 .. --max-spin coarse intervals, so that the training windows cover the
tidal cycle; the coarse model starts from the box filter F of that fine
state. Then, per coarse step, the fine model runs `factor` steps, the coarse
model (dx 4 km, dt 80 s, box-filtered depth) takes one plain step to C_k+1,
and the network takes one Adam step (cosine learning rate, optional global
norm clipping) on the pair (input C_k+1, target F_k+1 - C_k+1). The coarse
model then continues from its own corrected state (as at inference, energy
guard on), and is reset to F_k+1 every r steps, r growing from 1 to
--rollout over the run, so that the network learns to pull drifted states
back toward the fine solution. A rollout whose step loss exceeds
--reset-spike times the seed's running mean is reset at once: a state that
far off would only feed the network gradients of its own drift. The saved
weights are a moving average of the weights over the second half of the
run (--ema decay per sample, one sample every --ema-every steps). Every
--val-every seeds the current weights run --val-steps coarse steps from rest
on a held-out bathymetry (--val-bathy) beside the plain coarse model, and the
weights with the lowest error ratio (plus the fraction of corrections the
energy guard rejected) are the ones saved. The
loss and the correction leave out the ssh of cells the step clamps to the
tide, and an activity gate 1 - exp(-(a / --gate-eps)^2), a = |u| + |v| +
|ssh| in input units, scales the output so that water at rest gets no
correction. Normalisation scales
come from a pilot run of the first seed.

Evaluation: for every held-out seed, the fine run, the plain coarse run and
the coarse run with the learned correction (and the energy guard) advance
side by side from rest; every `--every` coarse steps the coarse states are
compared with the box-filtered fine state (relative RMS error over the
interior, per field).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from .closure import ClosureNet, coarse_model, init_params
from .model import Model, c2_config
from .psy import TimeStepper

FIELDS = ("sshn_t", "un", "vn")
# error fields: ssh over the whole interior and over the cells the step does
# not clamp to the tide (neither coarse form can change the clamped ones)
EVAL_FIELDS = FIELDS + ("sshn_t_free",)
DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "data" / "closure_weights.npz"


def log(*a):
    print(*a, flush=True)


def seeds(spec: str) -> list[int]:
    a, b = spec.split(":")
    return list(range(int(a), int(b)))


def pair(n: int, factor: int, seed: int, coarse_steps: int):
    fine = Model(c2_config(n, factor * coarse_steps, seed=seed))
    coarse = coarse_model(fine, factor)
    return fine, coarse


def interior(a: np.ndarray, g) -> np.ndarray:
    return a[g.ys:g.ye + 1, g.xs:g.xe + 1]


def rel_rms(x: np.ndarray, ref: np.ndarray) -> float:
    return float(np.linalg.norm(x - ref) / max(np.linalg.norm(ref), 1e-300))


def guard_abs(coarse: Model) -> float:
    """Energy allowance of the guard per step: g a^2 times the cells of one
    grid column, the energy of one column of cells raised to the tidal
    amplitude a (about 40 times what the tide brings in per coarse step)."""
    g = coarse.grid
    return 9.80665 * coarse.cfg.tide_amp ** 2 * (g.ye - g.ys + 1)


def configure(net: ClosureNet, norm: dict) -> None:
    """Normalisation, activity gate and corrected ssh cells of a trained
    network (weights saved before the ssh-cell setting existed correct the
    free cells only)."""
    net.set_norm(norm["in_shift"], norm["in_scale"], norm["out_scale"])
    net.set_gate(norm.get("gate_eps", 0.0))
    net.set_ssh_cells(norm.get("ssh_cells", "free") == "wet")
    net.set_guard_limit(norm.get("guard_max_out", 0.0))


def free_cells(coarse: Model) -> np.ndarray:
    """Interior T cells whose ssh the step does not clamp to the tide (wet,
    no open 4-neighbour)."""
    tm = coarse.array("tmask")
    return interior((tm > 0) & (np.roll(tm, 1, 1) >= 0) & (np.roll(tm, -1, 1) >= 0)
                    & (np.roll(tm, 1, 0) >= 0) & (np.roll(tm, -1, 0) >= 0), coarse.grid)


def _active_rms(a: np.ndarray) -> float:
    m = np.abs(a) > 1e-6 * max(float(np.max(np.abs(a))), 1e-300)
    return float(np.sqrt(np.mean(a[m] ** 2))) if np.any(m) else 1.0


def wet_cells(coarse: Model) -> np.ndarray:
    """Interior T cells that are wet."""
    return interior(coarse.array("tmask") > 0, coarse.grid)


def pilot_norm(n: int, factor: int, seed: int, steps: int, every: int = 50,
               ssh_cells: str = "free") -> dict:
    """Input and target scales from a plain run of one seed (each coarse step
    started from the filtered fine state): RMS over the cells where the field
    is not (numerically) zero, sampled every `every` coarse steps; the ssh
    over the cells the closure corrects (ssh_cells "wet" or "free")."""
    fine, coarse = pair(n, factor, seed, steps)
    sf, sc = TimeStepper(fine, mode=2), TimeStepper(coarse, mode=2)
    net = ClosureNet(coarse)
    g = coarse.grid
    xs = {k: [] for k in FIELDS}
    ts = {k: [] for k in FIELDS}
    for k in range(steps):
        sf.run(factor)
        sc.run(1)
        net.filter_fine(sf, factor)
        if (k + 1) % every == 0:
            for f in FIELDS:
                c = interior(sc.fetch(f), g)
                F = interior(net.filtered(f, (g.H, g.W)), g)
                xs[f].append(c)
                ts[f].append(F - c)
        net.set_state(sc)
    net.release()
    sf.release()
    sc.release()
    # the ssh scales cover the cells whose ssh is corrected and learned
    cells = wet_cells(coarse) if ssh_cells == "wet" else free_cells(coarse)
    xs["sshn_t"] = [a[cells] for a in xs["sshn_t"]]
    ts["sshn_t"] = [a[cells] for a in ts["sshn_t"]]
    ht = interior(coarse.array("ht"), g)
    order = ("un", "vn", "sshn_t")
    return {"in_shift": [0.0, 0.0, 0.0, float(np.mean(ht))],
            "in_scale": [_active_rms(np.concatenate([a.ravel() for a in xs[f]])) for f in order]
            + [max(float(np.std(ht)), 1.0)],
            "out_scale": [_active_rms(np.concatenate([a.ravel() for a in ts[f]])) for f in order]}


def train(a) -> None:
    bathy = seeds(a.bathy)
    t0 = time.perf_counter()
    if a.resume:
        z = np.load(a.resume)
        params = z["params"]
        norm = json.loads(str(z["meta"]))["norm"]
        log("resumed from", a.resume)
    else:
        params = init_params(a.seed)
        log("pilot run for the normalisation, seed", bathy[0])
        norm = pilot_norm(a.n, a.factor, bathy[0], a.pilot_steps, ssh_cells=a.ssh_cells)
        norm["ssh_cells"] = a.ssh_cells
    norm["guard_max_out"] = a.guard_max_out
    if a.gate_eps is not None:
        norm["gate_eps"] = a.gate_eps
    log("norm", json.dumps(norm))
    net = None
    total = len(bathy) * a.steps
    done = 0
    history = []
    spin_rng = np.random.default_rng(a.seed + 1)
    ema = None                    # moving average of the weights (--ema)
    spikes = 0
    best = None                   # (score, params, after seed) of the validation
    validations = []
    for bi, b in enumerate(bathy):
        fine, coarse = pair(a.n, a.factor, b, a.steps)
        sf, sc = TimeStepper(fine, mode=2), TimeStepper(coarse, mode=2)
        # the training window of this seed starts after a random spin-up of
        # the fine run (0 .. --max-spin coarse intervals), so that the pairs
        # cover the tidal cycle, not only its first hours
        spin = int(spin_rng.integers(0, a.max_spin + 1)) if a.max_spin > 0 else 0
        if spin:
            sf.run(a.factor * spin)
        sc.run(1)                 # first call: device state and pairing
        sc.istp = coarse.cfg.nit000 + spin
        if net is None:
            net = ClosureNet(coarse, params=params)
            configure(net, norm)
            net.set_clip(a.clip)
            net.set_loss_weights(*a.loss_weights)
        g = coarse.grid
        net.set_guard(a.guard_rel, guard_abs(coarse))
        net.filter_fine(sf, a.factor)
        net.set_state(sc)         # the window starts from the filtered fine state
        losses = []
        since_reset = 0
        loss_avg = None
        for k in range(a.steps):
            frac = done / max(total - 1, 1)
            lr = a.lr_min + 0.5 * (a.lr - a.lr_min) * (1.0 + np.cos(np.pi * frac))
            net.set_optimizer("adam", lr=lr)
            sf.run(a.factor)
            sc.run(1)
            net.filter_fine(sf, a.factor)
            net.train_step(sc)
            rollout = 1 + int(round((a.rollout - 1) * frac))
            since_reset += 1
            spike = False
            if a.reset_spike > 0:
                # a rollout whose loss jumps far above this seed's running
                # mean has drifted: restart it from the filtered fine state
                # instead of training on where it went
                loss = net.last_loss()
                spike = loss_avg is not None and k >= 10 and loss > a.reset_spike * loss_avg
                if not spike:
                    loss_avg = loss if loss_avg is None else 0.95 * loss_avg + 0.05 * loss
                else:
                    spikes += 1
            if since_reset >= rollout or spike:
                net.set_state(sc)
                since_reset = 0
            else:
                net.correct(sc, guard=True)
            done += 1
            if a.ema > 0 and frac >= a.ema_start and done % a.ema_every == 0:
                p = net.params()
                ema = p.astype(np.float64) if ema is None else a.ema * ema + (1.0 - a.ema) * p
            if (k + 1) % a.log_every == 0:
                st = net.last_stats()
                losses.append([st["loss_u"], st["loss_v"], st["loss_ssh"], st["grad_norm"]])
        history.append({"seed": b, "loss_u_v_ssh_gradnorm": losses})
        first = losses[0] if losses else [float("nan")] * 4
        last = losses[-1] if losses else [float("nan")] * 4
        log(f"seed {b}: loss u/v/ssh {first[0]:.3e} {first[1]:.3e} {first[2]:.3e} -> "
            f"{last[0]:.3e} {last[1]:.3e} {last[2]:.3e}  |g| {last[3]:.2e}  lr {lr:.2e}  "
            f"{time.perf_counter() - t0:.0f} s")
        sf.release()
        sc.release()
        params = net.params() if ema is None else ema.astype(np.float32)
        save(a.out, params, norm, a, history)
        if a.val_every > 0 and ((bi + 1) % a.val_every == 0 or bi == len(bathy) - 1):
            score, ratios, trips = validation_score(a, params, norm, history)
            validations.append({"after_seed": b, "score": score, "ratio_mean": ratios,
                                "guard_trips": trips})
            log(f"validation after seed {b}: score {score:.3f}  ratios "
                f"{ {k: round(v, 3) for k, v in ratios.items()} }  trips {trips}")
            if not np.isfinite(score):
                score = float("inf")  # a run that went non-finite is the worst
            if best is None or score < best[0]:
                best = (score, params.copy(), b)
    net.release()
    if best is not None:
        params = best[1]
        save(a.out, params, norm, a, history,
             extra={"validation": validations, "selected_after_seed": best[2]})
        if best[0] >= 1.0:
            log("warning: no checkpoint beat the plain coarse model on the validation "
                f"bathymetry (best score {best[0]:.3f}); the saved network is not an "
                "improvement")
    log(f"trained {done} steps in {time.perf_counter() - t0:.0f} s -> {a.out}"
        + (f" (moving average of the weights, decay {a.ema})" if ema is not None else "")
        + (f", {spikes} rollouts restarted on a loss spike" if a.reset_spike > 0 else "")
        + (f"; kept the weights after seed {best[2]} (validation score {best[0]:.3f})"
           if best is not None else ""))


def validation_score(a, params, norm, history) -> tuple[float, dict, int]:
    """Model selection: the learned / plain error ratio of `params` on the
    held-out --val-bathy seeds, over --val-steps coarse steps from rest (mean
    over time and seeds, averaged over u, v and the ssh of the cells the
    closure corrects; the worst seed's average weighs in as much as the
    mean), plus the fraction of those steps whose correction the energy
    guard rejected."""
    tmp = Path(a.out).with_name(Path(a.out).stem + ".validation.npz")
    save(tmp, params, norm, a, history)
    vseeds = val_seeds(a.val_bathy)
    try:
        res = evaluate(a.n, a.factor, vseeds, a.val_steps,
                       max(a.val_steps // 10, 1), tmp, log_fn=lambda *x: None)
    finally:
        tmp.unlink()
    ssh = "sshn_t" if norm.get("ssh_cells", "free") == "wet" else "sshn_t_free"
    fields = ("un", "vn", ssh)
    ratios = {f: float(res["ratio_mean"][f]) for f in fields}
    worst = max(float(np.mean([res["seeds"][b]["ratio_mean"][f] for f in fields]))
                for b in vseeds)
    trips = sum(int(res["seeds"][b]["guard_trips"]) for b in vseeds)
    score = 0.5 * (float(np.mean(list(ratios.values()))) + worst)
    return score + trips / (a.val_steps * len(vseeds)), ratios, trips


def val_seeds(spec) -> list[int]:
    """--val-bathy: one seed ("1000") or a range ("1000:1002")."""
    spec = str(spec)
    return seeds(spec) if ":" in spec else [int(spec)]


def save(path, params, norm, a, history, extra=None) -> None:
    meta = {"norm": norm, "n_fine": a.n, "factor": a.factor, "bathy": a.bathy,
            "coarse_steps_per_seed": a.steps, "lr": a.lr, "lr_min": a.lr_min,
            "rollout": a.rollout, "guard_rel": a.guard_rel, "clip": a.clip,
            "max_spin": a.max_spin, "loss_weights": list(a.loss_weights),
            "ema": a.ema, "ema_every": a.ema_every, "ema_start": a.ema_start,
            "reset_spike": a.reset_spike, "val_bathy": a.val_bathy,
            "val_every": a.val_every, "val_steps": a.val_steps,
            "optimizer": "adam", "history": history}
    meta.update(extra or {})
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    np.savez(path, params=np.asarray(params, np.float32), meta=json.dumps(meta))


def load_weights(path) -> tuple[np.ndarray, dict]:
    z = np.load(path)
    return z["params"], json.loads(str(z["meta"]))


def evaluate(n: int, factor: int, bathy: list[int], steps: int, every: int, weights,
             guard_rel: float | None = None, log_fn=log, trace_energy: bool = False,
             guard_max_out: float | None = None) -> dict:
    """guard_rel None: the energy-guard fraction the weights were trained
    with; guard_max_out None: their bound on each applied component."""
    params, meta = load_weights(weights)
    norm = dict(meta["norm"])
    if guard_rel is None:
        guard_rel = float(meta.get("guard_rel", 0.25))
    if guard_max_out is not None:
        norm["guard_max_out"] = guard_max_out
    out = {"seeds": {}, "every": every, "steps": steps, "guard_rel": guard_rel,
           "guard_max_out": float(norm.get("guard_max_out", 0.0))}
    for b in bathy:
        fine, coarse = pair(n, factor, b, steps)
        plain = coarse.copy()
        sf, sp, sl = (TimeStepper(fine, mode=2), TimeStepper(plain, mode=2),
                      TimeStepper(coarse, mode=2))
        net = ClosureNet(coarse, params=params)
        configure(net, norm)
        g = coarse.grid
        net.set_guard(guard_rel, guard_abs(coarse))
        free = free_cells(coarse)
        errs = {"plain": {f: [] for f in EVAL_FIELDS}, "learned": {f: [] for f in EVAL_FIELDS}}
        ratios = []
        k = 0
        while k < steps:
            # advance every run to the next comparison; the learned steps go
            # as one graph replay of coarse step + correction (nemo_cl_steps)
            # once the coarse model has its device state
            nblk = min(every - k % every, steps - k)
            if trace_energy or k == 0:
                per_step = nblk if trace_energy else 1
                for _ in range(per_step):
                    sl.run(1)
                    net.correct(sl, guard=True)
                    if trace_energy:
                        st = net.last_stats()
                        ratios.append((st["energy_after"] - st["energy_before"])
                                      / max(st["energy_before"], 1e-300))
                if nblk > per_step:
                    net.steps(sl, nblk - per_step, guard=True)
            else:
                net.steps(sl, nblk, guard=True)
            sf.run(factor * nblk)
            sp.run(nblk)
            k += nblk
            if k % every == 0:
                net.filter_fine(sf, factor)
                for f in FIELDS:
                    F = interior(net.filtered(f, (g.H, g.W)), g)
                    for m, st in (("plain", sp), ("learned", sl)):
                        x = interior(st.fetch(f), g)
                        errs[m][f].append(rel_rms(x, F))
                        if f == "sshn_t":
                            errs[m]["sshn_t_free"].append(rel_rms(x[free], F[free]))
        trips = net.guard_trips()
        res = {"final": {m: {f: errs[m][f][-1] for f in EVAL_FIELDS} for m in errs},
               "mean": {m: {f: float(np.mean(errs[m][f])) for f in EVAL_FIELDS} for m in errs},
               "guard_trips": trips}
        res["ratio_final"] = {f: res["final"]["learned"][f] / res["final"]["plain"][f]
                              for f in EVAL_FIELDS}
        res["ratio_mean"] = {f: res["mean"]["learned"][f] / res["mean"]["plain"][f]
                             for f in EVAL_FIELDS}
        out["seeds"][b] = res
        log_fn(f"seed {b}: plain {res['final']['plain']}  learned {res['final']['learned']}"
               f"  ratio(final) {res['ratio_final']}  trips {trips}")
        if trace_energy:
            r = np.array(ratios)
            res["energy_change_percentiles"] = {
                str(q): float(np.percentile(r, q)) for q in (1, 10, 50, 90, 99, 100)}
            log_fn(f"seed {b}: relative energy change per correction, percentiles "
                   f"{res['energy_change_percentiles']}")
        net.release()
        for s in (sf, sp, sl):
            s.release()
    for key in ("ratio_final", "ratio_mean"):
        out[key] = {f: float(np.mean([out["seeds"][b][key][f] for b in bathy]))
                    for f in EVAL_FIELDS}
    return out


def throughput(n: int, factor: int, weights, steps: int = 200) -> dict:
    """Wall time per simulated coarse interval (factor fine steps) of the
    fine FP64 run and of one coarse step plus the learned correction."""
    from ._native import hip_lib
    params, meta = load_weights(weights)
    norm = meta["norm"]
    fine, coarse = pair(n, factor, 0, steps)
    sf, sl = TimeStepper(fine, mode=2), TimeStepper(coarse, mode=2)
    sf.run(factor * 10)
    sl.run(1)
    net = ClosureNet(coarse, params=params)
    configure(net, norm)
    hip_lib().nemo_hip_synchronize()
    t0 = time.perf_counter()
    sf.run(factor * steps)
    hip_lib().nemo_hip_synchronize()
    fine_ms = (time.perf_counter() - t0) * 1e3 / steps
    net.steps(sl, steps, guard=True)      # captures the graph of `steps` learned steps
    hip_lib().nemo_hip_synchronize()
    t0 = time.perf_counter()
    net.steps(sl, steps, guard=True)
    hip_lib().nemo_hip_synchronize()
    learned_ms = (time.perf_counter() - t0) * 1e3 / steps
    # the same steps as host-issued launches (c_invoke_time_step + nemo_cl_correct)
    t0 = time.perf_counter()
    for _ in range(steps):
        sl.run(1)
        net.correct(sl, guard=True)
    hip_lib().nemo_hip_synchronize()
    stepwise_ms = (time.perf_counter() - t0) * 1e3 / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        sl.run(1)
    hip_lib().nemo_hip_synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / steps
    net.set_timing(True)
    fwd = []
    for _ in range(20):
        net.correct(sl, guard=True)
        fwd.append(net.forward_ms())
    net.set_timing(False)
    net.release()
    sf.release()
    sl.release()
    return {"fine_ms_per_coarse_interval": fine_ms, "learned_ms_per_coarse_step": learned_ms,
            "learned_stepwise_ms_per_coarse_step": stepwise_ms,
            "plain_ms_per_coarse_step": plain_ms, "forward_ms": float(np.median(fwd)),
            "speedup_vs_fine": fine_ms / learned_ms}


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    sub = p.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--n", type=int, default=4096)
    t.add_argument("--factor", type=int, default=4)
    t.add_argument("--bathy", default="0:64")
    t.add_argument("--steps", type=int, default=500)
    # defaults: the recipe of the shipped weights
    t.add_argument("--lr", type=float, default=3e-4)
    t.add_argument("--lr-min", type=float, default=3e-5)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--log-every", type=int, default=100)
    t.add_argument("--rollout", type=int, default=15,
                   help="longest stretch of steps on the corrected state between resets")
    t.add_argument("--guard-rel", type=float, default=0.25)
    t.add_argument("--clip", type=float, default=1.0,
                   help="largest global gradient norm (0: no clipping)")
    t.add_argument("--loss-weights", type=lambda s: tuple(float(x) for x in s.split(",")),
                   default=(1.0, 3.0, 1.0), help="u,v,ssh weights of the squared errors")
    t.add_argument("--gate-eps", type=float, default=0.01,
                   help="activity gate width in normalised input units (0: no gate)")
    t.add_argument("--pilot-steps", type=int, default=2000,
                   help="coarse steps of the normalisation pilot run")
    t.add_argument("--max-spin", type=int, default=1500,
                   help="longest spin-up of a seed's fine run before its window, in "
                        "coarse steps")
    t.add_argument("--ema", type=float, default=0.99,
                   help="save a moving average of the weights with this decay per sample "
                        "(0: the last weights)")
    t.add_argument("--ema-every", type=int, default=10,
                   help="training steps between samples of the moving average")
    t.add_argument("--ema-start", type=float, default=0.5,
                   help="fraction of the run after which the average starts")
    t.add_argument("--reset-spike", type=float, default=4.0,
                   help="restart a rollout from the filtered fine state when a step's loss "
                        "exceeds this multiple of the seed's running mean (0: never)")
    t.add_argument("--val-bathy", default="1000:1002",
                   help="held-out bathymetry seed(s) of the model-selection runs: one seed "
                        "or a range a:b")
    t.add_argument("--val-every", type=int, default=8,
                   help="training seeds between model-selection runs (0: keep the last "
                        "weights)")
    t.add_argument("--val-steps", type=int, default=2000,
                   help="coarse steps of each model-selection run (as long as the "
                        "evaluation)")
    t.add_argument("--guard-max-out", type=float, default=20.0,
                   help="the energy guard also rejects corrections with any component "
                        "beyond this many normalised units (0: no bound)")
    t.add_argument("--ssh-cells", choices=("wet", "free"), default="free",
                   help="ssh cells the closure corrects and learns: every wet cell, or "
                        "only the cells the step does not clamp to the tide")
    t.add_argument("--resume", default=None)
    t.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    e = sub.add_parser("eval")
    e.add_argument("--n", type=int, default=4096)
    e.add_argument("--factor", type=int, default=4)
    e.add_argument("--bathy", default="64:80")
    e.add_argument("--steps", type=int, default=2000)
    e.add_argument("--every", type=int, default=100)
    e.add_argument("--weights", default=str(DEFAULT_WEIGHTS))
    e.add_argument("--json", default=None)
    e.add_argument("--guard-rel", type=float, default=None,
                   help="energy-guard fraction (default: the one used in training)")
    e.add_argument("--guard-max-out", type=float, default=None,
                   help="bound on each applied component in normalised units (0: none; "
                        "default: the one the weights were trained with)")
    e.add_argument("--trace-energy", action="store_true",
                   help="record the energy change of every correction (synchronises)")
    s = sub.add_parser("speed")
    s.add_argument("--n", type=int, default=4096)
    s.add_argument("--factor", type=int, default=4)
    s.add_argument("--weights", default=str(DEFAULT_WEIGHTS))
    s.add_argument("--steps", type=int, default=200,
                   help="coarse steps timed (the fine run covers the same model time)")
    s.add_argument("--conv-form", type=int, default=None,
                   help="launch form of the 32 -> 32 convolutions (0 or 1; default: the "
                        "library's)")
    return p


def main(argv=None) -> None:
    a = build_parser().parse_args(argv)
    if a.cmd == "train":
        train(a)
    elif a.cmd == "eval":
        res = evaluate(a.n, a.factor, seeds(a.bathy), a.steps, a.every, a.weights,
                       guard_rel=a.guard_rel, trace_energy=a.trace_energy,
                       guard_max_out=a.guard_max_out)
        log(json.dumps({k: res[k] for k in ("ratio_final", "ratio_mean")}))
        if a.json:
            Path(a.json).write_text(json.dumps(res, indent=1))
    else:
        if a.conv_form is not None:
            from .closure import set_conv_form
            set_conv_form(a.conv_form)
        log(json.dumps(throughput(a.n, a.factor, a.weights, steps=a.steps)))


if __name__ == "__main__":
    main(sys.argv[1:])
