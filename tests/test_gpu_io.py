"""Output files of the HIP library against the reference's layouts: legacy
binary VTK (fluid and DEM writers of functions_OUTPUT.cuh) and restart.txt
(save_restart), with the restart read back through the input reader."""
import json
from pathlib import Path

import numpy as np
import pytest

from fluidized_bed_master_amd import abi, configs
from tests.helpers import read_vtk

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
FLUID_FIELDS = [("porosity", "float"), ("conv_rate", "float"), ("pressure", "float"), ("cond_rate", "float"),
                ("dtemp", "float"), ("buffer_type", "int"), ("density", "float"), ("temperature", "float"),
                ("uz", "float")]


def _stepped():
    from fluidized_bed_master_amd.sophia import Sophia
    p, parts = configs.small_isph_dem()
    g = Sophia(p, parts, precision=abi.PREC_FP64)
    g.step(abi.ppe_opts(method=abi.PPE_JACOBI))
    return g


def test_dem_vtk_matches_reference_layout(tmp_path):
    g = _stepped()
    path = tmp_path / "dem.vtk"
    g.write_vtk(path, abi.VTK_DEM)
    v = read_vtk(path)
    lay = json.load(open(GOLD / "reference_dem_vtk_layout.json"))
    d1, _, d3 = g.get_particles(abi.DEM)
    sel = d1["p_type"] > 1000
    n = int((d1["dem_idx"] > 0).sum())
    assert v["header"][:4] == lay["header"][:4] and v["header"][4] == f"POINTS\t{n}\tfloat"
    assert v["point_data"] == f"POINT_DATA\t{n}" and v["field_line"] == lay["field_line"]
    assert [(f["name"], f["ncomp"], f["type"]) for f in v["fields"]] == \
        [(f["name"], f["ncomp"], f["type"]) for f in lay["fields"]]
    assert v["end"] == v["bytes"]
    pts = np.array(v["points"]).reshape(-1, 3)
    ref = np.stack([d1["x"][sel], d1["y"][sel], d1["z"][sel]], 1).astype(np.float32)
    assert np.array_equal(pts.astype(np.float32), ref)
    src = {"itype": d1["i_type"].astype(float), "mag_uijf": d1["test6"], "Euler": d1["eliz"], "porosity": d1["DEMpor"],
           "uz": d1["uz"], "Q_sd": d1["Q_sd"], "ftotal_z": d3["ftotalz"], "temperature_surface": d1["temp"],
           "rad": d1["rad"], "dtemp_dt": d3["dtemp"]}
    for f in v["fields"]:
        assert np.array_equal(np.array(f["values"], np.float32), src[f["name"]][sel].astype(np.float32)), f["name"]


def test_fluid_vtk_layout_and_values(tmp_path):
    g = _stepped()
    path = tmp_path / "fluid.vtk"
    g.write_vtk(path, abi.VTK_FLUID)
    v = read_vtk(path)
    s1, _, s3 = g.get_particles(abi.SPH)
    sel = s1["p_type"] == 1
    n = int(sel.sum())
    assert v["header"][4] == f"POINTS\t{n}\tfloat" and v["field_line"] == "FIELD FieldData\t9"
    assert [(f["name"], f["type"]) for f in v["fields"]] == FLUID_FIELDS
    assert v["end"] == v["bytes"]
    src = {"porosity": s1["DEMpor"], "conv_rate": s1["Q_sdf"], "pressure": s1["pres"], "cond_rate": s1["Q_f"],
           "dtemp": s3["dtemp"], "buffer_type": s1["buffer_type"], "density": s1["rho"], "temperature": s1["temp"],
           "uz": s1["uz"]}
    for f in v["fields"]:
        want = src[f["name"]][sel]
        got = np.array(f["values"])
        if f["type"] == "int":
            assert np.array_equal(got, want)
        else:
            assert np.array_equal(got.astype(np.float32), want.astype(np.float32)), f["name"]


def test_restart_text_and_read_back(tmp_path):
    from fluidized_bed_master_amd.sophia import Sophia
    g = _stepped()
    path = tmp_path / "restart.txt"
    g.write_restart(path)
    lines = open(path).read().splitlines()
    assert lines[0] == "1 2 3 4 5 6 7 8 9 10 11 12 13 26"
    rows = []
    for sp in (abi.SPH, abi.DEM):
        a1, a2, _ = g.get_particles(sp)
        for i in range(len(a1)):
            q = a1[i]
            rows.append("%f\t%f\t%f\t%f\t%f\t%f\t%e\t%d\t%e\t%f\t%f\t%f\t%f\t%f" % (
                q["x"], q["y"], q["z"], q["ux"], q["uy"], q["uz"], q["m"], q["p_type"], q["h"], q["temp"],
                q["pres"], q["rho"], a2[i]["rho_ref"], q["flt_s"]))
    assert lines[1:] == rows
    # read back as an input file: the values come back at the printed precision
    h = Sophia.from_files(str(GOLD / "reference_solv.txt"), str(path), precision=abi.PREC_FP64)
    n_sph = h.count(abi.SPH) if hasattr(h, "count") else None
    vals = np.array([[float(t) for t in r.split("\t")] for r in rows])
    got = np.concatenate([h.get_particles(sp)[0] for sp in (abi.SPH, abi.DEM)])
    # the reader orders SPH before DEM as written; compare as sorted tuples
    key = lambda a: np.lexsort((a[:, 2], a[:, 1], a[:, 0]))
    mine = np.stack([got["x"], got["y"], got["z"], got["m"], got["p_type"].astype(float), got["h"]], 1)
    theirs = vals[:, [0, 1, 2, 6, 7, 8]]
    assert np.array_equal(mine[key(mine)], theirs[key(theirs)])
    assert n_sph is None or n_sph == int((vals[:, 7] <= 1000).sum())
