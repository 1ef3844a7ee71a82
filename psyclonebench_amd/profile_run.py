"""Fixed workload for rocprofv3 runs (kernel trace and PMC counters).

    python -m psyclonebench_amd.profile_run --n 4096 --fused 10 --unfused 2

Runs `--fused` fused steps of the 4096^2 FP64 workload as plain launches
(one k_fused_step + one k_flather dispatch per step) and `--unfused` steps of
the unfused sequence, whose k_copy3 dispatch is a plain copy of known size
(3 arrays read + 3 written over the copy region) used to calibrate the
FETCH_SIZE / WRITE_SIZE counters. Prints the dispatch geometry and the
algorithmic byte counts the profiles are compared with. `--closure-train`
runs training steps of the learned closure (forward, loss, backward, Adam)
on synthetic input / target pairs over the 1024^2 coarse grid of the same
workload, uploaded from the host each step. `--ensemble` runs steps of the
256-domain ensemble (seed 0, sides log-uniform in [64, 1024]) as plain
launches, per precision and one-wavefront-per-tile shape.
"""
from __future__ import annotations

import argparse
import json

from . import _native
from .model import Model, c2_config
from .psy import TimeStepper

B_ALG_F64 = 212


def main() -> None:
    p = argparse.ArgumentParser()
    p.add_argument("--n", type=int, default=4096)
    p.add_argument("--fused", type=int, default=10)
    p.add_argument("--unfused", type=int, default=2)
    p.add_argument("--variant", type=str, default="0",
                   help="fused kernel shape(s), comma separated (-1 tile, 0.. march)")
    p.add_argument("--mixed", type=int, default=0,
                   help="steps of the FP32 mixed form (plain launches)")
    p.add_argument("--mixed-rows", type=str, default="64",
                   help="rows per band of the mixed kernel, comma separated")
    p.add_argument("--mixed-cols", type=str, default="61",
                   help="output columns per wavefront of the mixed kernel (61, 48), "
                        "comma separated")
    p.add_argument("--mixed-prefetch", type=int, default=0,
                   help="extra rows of loads the mixed kernel keeps in flight (0, 1 or 2)")
    p.add_argument("--ensemble", type=int, default=0,
                   help="ensemble steps (plain launches) per precision and shape")
    p.add_argument("--ens-domains", type=int, default=256)
    p.add_argument("--ens-precision", type=str, default="64,32")
    p.add_argument("--ens-shapes", type=str, default="4x56p1",
                   help="one-wavefront-per-tile ensemble shapes ROWSxCOLSpPREFETCH "
                        "(COLS 0: per domain), comma separated")
    p.add_argument("--closure-train", type=int, default=0,
                   help="closure training steps on synthetic coarse-grid pairs")
    a = p.parse_args()
    lib = _native.hip_lib()
    variants = [int(v) for v in a.variant.split(",")]
    cfg = c2_config(a.n, 1, seed=0)
    out = {"n": a.n, "variants": variants}
    if a.fused:
        out["fused_kernel_ms"] = {}
        for v in variants:
            lib.nemo_hip_set_fused_variant(v)
            m = Model(cfg)
            g = m.grid
            st = TimeStepper(m, mode=2)
            lib.nemo_hip_set_kernel_timing(1)      # plain launches, event-timed
            st.run(a.fused + 1)
            out["fused_kernel_ms"][v] = lib.nemo_hip_fused_kernel_ms()
            lib.nemo_hip_set_kernel_timing(0)
            st.release()
        pts = (g.M - 2) * (g.N - 2)
        out["fused_alg_bytes_per_launch"] = B_ALG_F64 * pts
    if a.unfused:
        m = Model(cfg)
        g = m.grid
        st = TimeStepper(m, mode=1)
        for _ in range(a.unfused):
            st.step()
        st.synchronize()
        st.release()
        rows = g.ye - g.ys + 2          # ys-1 .. ye
        cols = g.xe - g.xs + 3          # xs-1 .. xe+1
        out["copy3_bytes_read"] = 3 * 8 * rows * cols
        out["copy3_bytes_written"] = 3 * 8 * rows * cols
    if a.mixed:
        from .psy import MixedStepper
        out["mixed_kernel_ms"] = {}
        for cols in [int(c) for c in a.mixed_cols.split(",")]:
            for rows in [int(r) for r in a.mixed_rows.split(",")]:
                st = MixedStepper(Model(cfg), rows=rows, cols=cols, prefetch=a.mixed_prefetch)
                st.run(1)
                out["mixed_kernel_ms"][f"{rows}x{cols}"] = st.timed_run(a.mixed)
                st.release()
        out["mixed_alg_bytes_per_launch"] = 113 * (cfg.jpiglo - 2) * (cfg.jpjglo - 2)
    if a.ensemble:
        from .ensemble import EnsembleStepper, ensemble_configs
        models = [Model(c) for c in ensemble_configs(a.ens_domains, seed=0)]
        out["ensemble_kernel_ms"] = {}
        for prec in [int(q) for q in a.ens_precision.split(",")]:
            for spec in a.ens_shapes.split(","):
                rc, pf = spec.split("p")
                rows, cols = (int(x) for x in rc.split("x"))
                ens = EnsembleStepper(models, precision=prec)
                k = [i for i, sh in enumerate(ens.shapes())
                     if sh[2] == 0 and (sh[0], sh[3], sh[4]) == (rows, cols, int(pf))]
                if not k:
                    raise SystemExit(f"no one-wavefront-per-tile ensemble shape {spec}")
                ens.set_shape(k[0])
                ens.run(1)
                out["ensemble_kernel_ms"][f"f{prec} {spec}"] = ens.timed_run(a.ensemble)
                pts = ens.interior_points
                ens.release()
        out["ensemble_interior_points"] = pts
        out["ensemble_alg_bytes_per_launch"] = {"f64": B_ALG_F64 * pts, "f32": 113 * pts}
    if a.closure_train:
        import numpy as np

        from .closure import ClosureNet, coarse_model
        coarse = coarse_model(Model(cfg), 4)
        g = coarse.grid
        ny, nx = g.ye - g.ys + 1, g.xe - g.xs + 1
        net = ClosureNet(coarse, seed=0)
        net.set_optimizer("adam", lr=1e-4)
        rng = np.random.default_rng(0)
        x = rng.standard_normal((ny, nx, 4))
        t = 0.1 * rng.standard_normal((ny, nx, 3))
        losses = [net.train_host(x, t) for _ in range(a.closure_train)]
        net.release()
        out["closure_grid"] = [nx, ny]
        out["closure_train_loss_first_last"] = [losses[0], losses[-1]]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
