"""The Fortran driver (psyclonebench_amd/fortran/nemolite2d_hip.f90) steps the
model through the reference's Fortran binding of the C ABI with the
device-resident field protocol (device slots, data_on_device, a
read_from_device_c hook installed after the first call). Its checksums,
read back through that hook during the run and at its end, must be the
oracle's for the same steps, and digit for digit those of the Python
driver over the same library."""
import re
import shutil
import subprocess
from pathlib import Path

import pytest

import oracle
from psyclonebench_amd import gocean_io as gio
from psyclonebench_amd.model import Model, reference_config
from psyclonebench_amd.psy import TimeStepper

PKG = Path(__file__).resolve().parents[1] / "psyclonebench_amd"
EXE = PKG / "nemolite2d_hip"

NAMELIST = """&namctl
   jpiglo = {n2}, jpjglo = {n2}, jphgr_msh = 1,
   dx = 1000.0, dy = 1000.0, dep_const = 100.0,
   nit000 = 1, nitend = {steps}, irecord = {irecord},
   rdt = 20.0, cbfr = 0.00015, visc = 0.1
/
"""


def _digits18(x: float) -> float:
    """x rounded to the 18 significant digits of the driver's E26.18."""
    return float(f"{x:.17e}")


def _built():
    if not EXE.exists():
        pytest.skip("no Fortran compiler where the package was built")


def test_driver_links_the_in_tree_libraries():
    _built()
    if shutil.which("ldd") is None:
        pytest.skip("no ldd")
    out = subprocess.run(["ldd", str(EXE)], capture_output=True, text=True).stdout
    for lib in ("libnemo_hip.so", "libnemo_setup.so"):
        line = next(l for l in out.splitlines() if lib in l)
        assert str(PKG) in line and "not found" not in line, line


@pytest.mark.gpu
@pytest.mark.parametrize("n,steps,irecord", [(128, 200, 50), (250, 57, 19)])
def test_fortran_driver_matches_oracle_and_python_driver(tmp_path, n, steps, irecord):
    _built()
    (tmp_path / "namelist").write_text(NAMELIST.format(n2=n + 2, steps=steps, irecord=irecord))
    print(f"fortran driver n={n} steps={steps} irecord={irecord}", flush=True)
    r = subprocess.run([str(EXE)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = dict(re.findall(r"(\w+) checksum = +(\S+)", r.stdout))
    full = dict(re.findall(r"(\w+) full checksum = +(\S+)", r.stdout))
    mid = {int(k): float(v) for k, v in re.findall(r"step +(\d+) sshn checksum = +(\S+)",
                                                    r.stdout)}
    assert "Simulation finished!!" in r.stdout
    assert sorted(mid) == list(range(irecord, steps + 1, irecord))

    cfg, _ = gio.read_namelist(tmp_path / "namelist")
    assert cfg == reference_config(n, steps)

    # the oracle over the same steps: ssh at every record step, u and v at the end
    mo = Model(cfg)
    g = mo.grid
    arrays = mo.arrays()
    done = 0
    for k in sorted(mid):
        oracle.run_steps(arrays, done + 1, k - done, g.bounds, g.W, mo.scalars)
        done = k
        assert float(mid[k]) == _digits18(oracle.checksum(mo.array("sshn_t"), g.W, g.bounds)), k
    oracle.run_steps(arrays, done + 1, steps - done, g.bounds, g.W, mo.scalars)
    for name in ("ua", "va"):
        assert float(full[name]) == _digits18(oracle.checksum(mo.array(name), g.W, g.bounds))

    m = Model(cfg)
    st = TimeStepper(m)
    for _ in range(steps):
        st.step()
    for name, field in (("ua", "ua"), ("va", "va"), ("sshn", "sshn_t")):
        cs = m.checksum(field)    # read back from the device
        assert gio.fortran_e(cs, 16, 8).strip() == got[name], (name, cs, got[name])
    st.release()
