"""The C++ host API (include/hemepure_amd.hpp) compiled with g++ against the
library, checked against the Python binding of the same library."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import hemepure_amd as H

ROOT = Path(__file__).resolve().parents[1]
SRC = Path(__file__).resolve().parent / "cpp" / "host_api.cpp"
EXE = ROOT / "hemepure_amd" / "_build" / "host_api_test"


def build_exe() -> Path:
    if not EXE.exists() or EXE.stat().st_mtime < max(SRC.stat().st_mtime, H.LIB_PATH.stat().st_mtime,
                                                        (ROOT / "include" / "hemepure_amd.hpp").stat().st_mtime):
        EXE.parent.mkdir(parents=True, exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(SRC),
                        f"-L{H.LIB_PATH.parent}", "-lhemepure_amd", f"-Wl,-rpath,{H.LIB_PATH.parent}",
                        "-o", str(EXE)], check=True)
    return EXE


def run(mode, *extra):
    out = subprocess.run([str(build_exe()), mode, *map(str, extra)], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    return [l.split() for l in out.stdout.splitlines()]


def test_lattice_data_accessors():
    got = run("tables")
    g = H.Geometry.cylinder(6.0, 24, 16)
    lat = H.Lattice(g, None)
    facts = {l[0]: l[1:] for l in got if l[0] != "class"}
    assert int(facts["sites"][0]) == lat.n
    assert [int(l[2]) for l in got if l[0] == "class"] == list(lat.class_counts())
    push, _ = lat.neighbour_tables()
    assert int(facts["pushsum"][0]) == int(push.sum())
    coords = lat.site_coords()
    wall, iol, ident, types = lat.site_masks()
    assert [int(v) for v in facts["last"][:3]] == list(coords[-1])
    assert int(facts["last"][4]) == types[-1]
    assert int(facts["last"][6]) == max(c for c in range(6) if lat.class_counts()[c] > 0)


@pytest.mark.gpu
def test_simulation_master_matches_the_python_binding(tmp_path):
    """SimulationMaster<LBM<D3Q19, TRT, BFL, Nash>> over 150 steps with a
    property refresh at the end and stability checks: fields and
    distributions bit-for-bit those of the same run through the Python
    binding."""
    dump = tmp_path / "fields.bin"
    got = {l[0]: l[1:] for l in run("run", dump)}
    g = H.Geometry.cylinder(6.0, 24, 16)
    i0, o0 = g.inlets()[0], g.outlets()[0]
    lat = H.Lattice(g, H.params(kernel=H.KERNEL_TRT, wall_bc=H.WALL_BFL, tau=0.7),
                    [H.iolet_desc(H.COSINE, i0["position"], i0["normal"], density_mean=1.002)],
                    [H.iolet_desc(H.COSINE, o0["position"], o0["normal"], density_mean=1.0)])
    lat.init_equilibrium(1.0)
    lat.step(150, properties=True)
    rho, u, _ = lat.fields()
    f = lat.download_f()
    assert int(got["timestep"][0]) == lat.timestep() == 151
    raw = np.fromfile(dump, dtype=np.float64)
    n = lat.n
    assert np.array_equal(raw[:n], rho)
    assert np.array_equal(raw[n:4 * n], u.ravel())
    assert np.array_equal(raw[4 * n:], f.ravel())
    assert got["stable"] == ["1"]
