"""The .gmy reader (hlb_geom_read_gmy) and the case-config reader against
the reference's own files: fixture counts, the FourCube unit-test geometry,
tables identical to the oracle's, and the reader's error paths."""
import json
import os
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

import hemepure_amd as H
import oracle as O
from fixtures import four_cube

GOLDEN = Path(__file__).resolve().parent / "golden"
REF = Path("/root/reference")


def test_four_cube_file_is_the_four_cube_fixture():
    """src/unittests/resources/four_cube.gmy: 64 fluid sites in one 6^3 block;
    its link types are exactly those of the FourCube fixture (z faces iolets,
    x/y faces walls, edge links through z counted as iolets), cut distances
    0.5, normals on the 48 wall sites."""
    g = H.Geometry.read_gmy(GOLDEN / "four_cube.gmy")
    f = four_cube(6)
    assert g.n_sites == 64 and g.blocks == (1, 1, 1) and g.block_size == 6
    np.testing.assert_array_equal(g.coords, f.coords)
    np.testing.assert_array_equal(g.link_type, f.link_type)
    cut = g.link_type > 0
    assert (g.link_dist[cut] == 0.5).all() and (g.link_dist[~cut] == -1.0).all()
    np.testing.assert_array_equal(g.normal_avail, f.normal_avail)
    assert set(np.unique(g.link_iolet[g.link_type >= H.LINK_INLET])) == {0}


def test_bifurcation_case_geometry():
    g = H.Geometry.read_gmy(GOLDEN / "bifurcation_lores.gmy")
    assert g.n_sites == 197431 and g.blocks == (20, 5, 23) and g.block_size == 8
    assert sorted(np.unique(g.link_iolet[g.link_type == H.LINK_OUTLET])) == [0, 1]
    assert list(np.unique(g.link_iolet[g.link_type == H.LINK_INLET])) == [0]
    cfg = json.loads((GOLDEN / "bifurcation_lores_iolets.json").read_text())
    ins = [H.iolet_desc(H.COSINE, **d) for d in cfg["inlets"]]
    outs = [H.iolet_desc(H.COSINE, **d) for d in cfg["outlets"]]
    lat = H.Lattice(g, H.params(tau=cfg["tau"]), ins, outs)
    ol = O.OracleLattice(g)
    np.testing.assert_array_equal(lat.class_counts(), ol.class_counts())
    np.testing.assert_array_equal(lat.site_coords(), ol.site_coords())
    push, _ = lat.neighbour_tables()
    np.testing.assert_array_equal(push, ol.push_table())


@pytest.mark.skipif(not REF.exists(), reason="needs the reference's case files")
def test_reference_case_files():
    pipe = H.Geometry.read_gmy(REF / "cases/pipe/pipe.gmy")
    assert pipe.n_sites == 1987419 and pipe.blocks == (11, 11, 51)
    a = H.Geometry.read_gmy(REF / "cases/bifurcation/bifurcation_lores/bifurcation.gmy")
    b = H.Geometry.read_gmy(REF / "cases/bifurcation/bifurcation_lores/bifurcation.gmy+", gmy_plus=True)
    for name in ("coords", "link_type", "link_iolet", "link_dist", "normal_avail", "normal"):
        np.testing.assert_array_equal(getattr(a, name), getattr(b, name))


@pytest.mark.skipif(not REF.exists(), reason="needs the reference's case files")
def test_case_config_units():
    """UnitConverter: tau = 1/2 + (dt nu / rho) / (Cs2 dx^2); a pressure of p
    mmHg is the density 1 + p 133.3223874 / (1000 dx^2 / dt^2) / Cs2."""
    from hemepure_amd.config import load_case

    c = load_case(REF / "cases/pipe/input.xml")
    dt, dx = 1e-5, 5e-5
    assert abs(c.tau - (0.5 + (dt * 0.004 / 1000.0) / (dx * dx / 3.0))) < 1e-15
    lp = 1000.0 * dx * dx / (dt * dt)
    assert abs(c.inlets[0].density_mean - (1.0 + 0.01 * 133.3223874 / lp * 3.0)) < 1e-14
    assert c.outlets[0].density_mean == 1.0 and c.inlets[0].period == pytest.approx(1.0 / dt)
    assert c.geometry().n_sites == 1987419


def _write(tmp_path, data: bytes) -> Path:
    p = tmp_path / "x.gmy"
    p.write_bytes(data)
    return p


def _preamble(version=4, blocks=(1, 1, 1), bs=2, magic=0x686C6221):
    return struct.pack(">8I", magic, 0x676D7904, version, *blocks, bs, 0)


def test_reader_errors(tmp_path):
    with pytest.raises(H.HlbError, match="magic"):
        H.Geometry.read_gmy(_write(tmp_path, _preamble(magic=1) + b"\0" * 12))
    with pytest.raises(H.HlbError, match="version 3"):
        H.Geometry.read_gmy(_write(tmp_path, _preamble(version=3) + b"\0" * 12))
    with pytest.raises(H.HlbError, match="header runs past"):
        H.Geometry.read_gmy(_write(tmp_path, _preamble(blocks=(4, 4, 4))))
    # one fluid site claimed, block of 8 solid sites stored
    raw = struct.pack(">8I", *([0] * 8))
    comp = zlib.compress(raw)
    data = _preamble() + struct.pack(">3I", 1, len(comp), len(raw)) + comp
    with pytest.raises(H.HlbError, match="holds 0 fluid sites, its header says 1"):
        H.Geometry.read_gmy(_write(tmp_path, data))
    # a wall link without a normal
    site = [1] + [1] * 0
    recs = [1]
    for i in range(26):
        recs += [1] if i == 0 else [0]
        if i == 0:
            recs += [0x3F000000]  # float 0.5
    recs += [0]  # no normal: malformed for a wall site
    raw = struct.pack(f">{len(recs)}I", *recs) + struct.pack(">7I", *([0] * 7))
    comp = zlib.compress(raw)
    data = _preamble() + struct.pack(">3I", 1, len(comp), len(raw)) + comp
    with pytest.raises(H.HlbError, match="wall fluid site without"):
        H.Geometry.read_gmy(_write(tmp_path, data))
    with pytest.raises(H.HlbError, match="cannot open"):
        H.Geometry.read_gmy(tmp_path / "missing.gmy")
