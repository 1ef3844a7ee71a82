"""Planted mistakes in the oracle must be caught by its pins.

Each case rewrites one expression of oracle/nemo_oracle.c the way a
plausible restatement error would (a sign, an index, a transposed operand, a
dropped term, a wrong loop bound), builds the mutant into a temporary
library and checks that the golden fixtures written by the reference reject
it. The oracle source itself is not touched."""
from __future__ import annotations

import ctypes as C
import importlib.util
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle
from helpers import GOLDEN, kernel_fixture, kernel_inputs, load_npz, manifest, model_for_case

SRC = Path(oracle.HERE) / "nemo_oracle.c"

_spec = importlib.util.spec_from_file_location("make_golden", GOLDEN / "make_golden.py")
make_golden = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_golden)

MUTANTS = {
    "continuity_sign": ("(f_w - f_e + f_s - f_n)", "(f_e - f_w + f_s - f_n)"),
    "continuity_index": ("double f_s = (sshn_v[s] + hv[s]) * vn[s];",
                         "double f_s = (sshn_v[s] + hv[s]) * vn[c];"),
    "mom_u_length_scale": ("double u_e = 0.5 * (un[c] + un[e]) * e2t[e];",
                           "double u_e = 0.5 * (un[c] + un[e]) * e2t[c];"),
    "mom_u_depth_operand": ("double dep_s = 0.5 * (hv[s] + sshn_v[s] + hv[se] + sshn_v[se]);",
                            "double dep_s = 0.5 * (hv[s] + sshn_v[s] + hv[se] + sshn_v[s]);"),
    "mom_u_upwind_swap": ("double q_w = up_minus(u_w) * un[c] + up_plus(u_w) * un[w];",
                          "double q_w = up_plus(u_w) * un[c] + up_minus(u_w) * un[w];"),
    "mom_u_no_viscosity": ("vis = visc * vis;\n\n    /* Coriolis and pressure gradient */",
                           "vis = 0.0 * vis;\n\n    /* Coriolis and pressure gradient */"),
    "mom_v_coriolis_sign": ("double cor = -0.5 * (2. * omega", "double cor = 0.5 * (2. * omega"),
    "mom_v_quirk_fixed": ("if (tm[c] + tm[e] <= 0) return;\n    if (tm[c] <= 0 || tm[n] <= 0) return;",
                          "if (tm[c] + tm[n] <= 0) return;\n    if (tm[c] <= 0 || tm[n] <= 0) return;"),
    "mom_v_slip_wall": ("int dry_e = (tm[e] <= 0 || tm[ne] <= 0);",
                        "int dry_e = (tm[e] <= 0);"),
    "friction_sign": ("(hu[c] + ssha_u[c]) / (1.0 + cbfr * rdt);",
                      "(hu[c] + ssha_u[c]) / (1.0 - cbfr * rdt);"),
    "flather_sign": ("vel[c] = vel[p] + sqrt(g / h[c]) * (ssh[c] - ssh[p]);",
                     "vel[c] = vel[p] + sqrt(g / h[c]) * (ssh[p] - ssh[c]);"),
    "next_ssh_area": ("ssh_face[c] = 0.5 * vol / e12f[c];", "ssh_face[c] = 0.5 * vol / e12t[c];"),
    "bc_ssh_strict": ("if (tm[c - W] < 0 || tm[c + W] < 0", "if (tm[c - W] <= 0 || tm[c + W] < 0"),
    "copy_region": ("PAR for (int jj = ys - 1; jj < ye + 1; jj++)",
                    "PAR for (int jj = ys; jj < ye + 1; jj++)"),
    "flather_v_parallel_order": (
        "    for (int jj = ys - 1; jj <= ye; jj++)                                      \\\n"
        "        for (int ji = xs - 1; ji <= xe + 1; ji++)                              \\\n"
        "            pt_flather(jj * W + ji, W, va, hv, sshn_v, tmask, g);",
        "    for (int jj = ye; jj >= ys - 1; jj--)                                      \\\n"
        "        for (int ji = xs - 1; ji <= xe + 1; ji++)                              \\\n"
        "            pt_flather(jj * W + ji, W, va, hv, sshn_v, tmask, g);"),
}


def build_mutant(name: str, tmp: Path) -> C.CDLL:
    src = SRC.read_text()
    old, new = MUTANTS[name]
    assert src.count(old) >= 1, f"mutation anchor for {name} not found"
    (tmp / "nemo_oracle.h").write_text((Path(oracle.HERE) / "nemo_oracle.h").read_text())
    (tmp / "m.c").write_text(src.replace(old, new))
    so = tmp / f"lib_{name}.so"
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp",
                    "-o", str(so), str(tmp / "m.c"), "-lm"], check=True)
    lib = C.CDLL(str(so))
    lib.oracle_invoke_time_step.argtypes = oracle._STEP_ARGTYPES
    return lib


def fixtures_reject(lib) -> list[str]:
    """Names of the fixture checks the library fails."""
    failed = []
    meta, z = kernel_fixture()
    sc = tuple(meta["scalars"][s] for s in make_golden.SCAL)
    for mname in meta["masks"]:
        inp = kernel_inputs(mname)
        arrs = [inp[n] for n in oracle.ARRAY_NAMES]
        lib.oracle_invoke_time_step(*oracle.pointers(arrs), meta["scalars"]["istep"],
                                    1, meta["M"] - 2, 1, meta["N"] - 2, meta["W"], *sc)
        for n in make_golden.ALL_OUT:
            if not np.array_equal(inp[n].view(np.uint64), z[f"{mname}/step/{n}"].view(np.uint64)):
                failed.append(f"kernels/{mname}/{n}")
                break
    for case in ("ref_s32_south", "ref_s32_east", "bathy_s64_west"):
        c = manifest()["steps"][case]
        m = model_for_case(case)
        g = m.grid
        arrays = m.arrays()
        cp = c["checkpoints"][2]
        for istp in range(1, cp + 1):
            lib.oracle_invoke_time_step(*oracle.pointers(arrays), istp, *g.bounds, g.W,
                                        *m.scalars)
        fx = load_npz(f"steps_{case}.npz")
        for n in c["fields"]:
            if not np.array_equal(m.array(n), fx[f"{n}@{cp}"]):
                failed.append(f"{case}/{n}@{cp}")
                break
    return failed


def test_unmutated_source_passes(tmp_path):
    src = SRC.read_text()
    (tmp_path / "nemo_oracle.h").write_text((Path(oracle.HERE) / "nemo_oracle.h").read_text())
    (tmp_path / "m.c").write_text(src)
    so = tmp_path / "lib_clean.so"
    subprocess.run(["gcc", "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp",
                    "-o", str(so), str(tmp_path / "m.c"), "-lm"], check=True)
    lib = C.CDLL(str(so))
    lib.oracle_invoke_time_step.argtypes = oracle._STEP_ARGTYPES
    assert fixtures_reject(lib) == []


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_is_rejected(name, tmp_path):
    lib = build_mutant(name, tmp_path)
    assert fixtures_reject(lib), f"mutant {name} passed every fixture"
