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


# every writer of functions_OUTPUT.cuh: (kind, species, selection, count,
# [(array, ncomp, type, source)]) with the source taking (p1, p2, p3)
def _kinds():
    def f(name, get, typ="float"):
        return (name, 1, typ, get)

    def v3(name, a, b, c):
        return (name, 3, "float", lambda s1, s2, s3: np.stack([s3[a], s3[b], s3[c]], 1))

    one = lambda k: (lambda s1, s2, s3: s1[k])
    three = lambda k: (lambda s1, s2, s3: s3[k])
    dem_sel, dem_cnt = (lambda s1: s1["p_type"] > 1000), (lambda s1: s1["dem_idx"] > 0)
    return {
        abi.VTK_AIR: (abi.SPH, lambda s1: s1["p_type"] == 3, None, [
            f("rho_ref", lambda s1, s2, s3: s2["rho_ref"]), f("buffer_type", one("buffer_type")), f("uz", one("uz")),
            f("temperature", one("temp")), f("density", one("rho")), f("porosity", one("DEMpor")),
            f("pressure", one("pres")), f("drho_dt", three("drho")), f("dtemp_dt", three("dtemp")),
            f("ptype", one("p_type"))]),
        abi.VTK_SOLID: (abi.DEM, dem_sel, dem_cnt, [
            f("dem_idx", one("dem_idx")), f("mag_uijf", one("test6")), f("drag_r", one("test7")),
            f("porosity", one("DEMpor")), f("uz", one("uz")), f("Q_sd", one("Q_sd")), f("ftotal_z", three("ftotalz")),
            f("temperature_surface", one("temp")), f("rad", one("rad")), f("dtemp_dt", three("dtemp"))]),
        abi.VTK_MOVING: (abi.SPH, lambda s1: s1["p_type"] == 9, None,
                         [f("density", one("rho")), f("pressure", one("pres"))]),
        abi.VTK_FLUID2: (abi.SPH, lambda s1: s1["p_type"] == 2, None, [
            f("rho", one("rho")), f("pres", one("pres")), f("curvature", three("curv")),
            f("detection", three("lbl_surf")), v3("surface_tension", "fsx", "fsy", "fsz"),
            v3("normal_vector_c", "nx_c", "ny_c", "nz_c"), v3("normal_vector", "nx", "ny", "nz")]),
        abi.VTK_BOUNDARY: (abi.SPH, lambda s1: s1["p_type"] == 0, None, [
            f("uy", one("uy")), f("ux", one("ux")), f("uz", one("uz")), f("temperature", one("temp")),
            f("pressure", one("pres")), f("density", one("rho"))]),
        abi.VTK_FLUID_PGF: (abi.SPH, lambda s1: s1["p_type"] == 1, None, [
            f("porosity", one("DEMpor")), f("pressuregrad_x", one("pgf_x")), f("pressuregrad_y", one("pgf_y")),
            f("pressuregrad_z", one("pgf_z")), f("ux", one("ux")), f("pressure", one("pres")), f("uy", one("uy")),
            f("dtemp", three("dtemp")), f("buffer_type", one("buffer_type"), "int"), f("density", one("rho")),
            f("temperature", one("temp")), f("uz", one("uz"))]),
        abi.VTK_ACTIVE: (abi.SPH, lambda s1: (s1["i_type"] == 1) | (s1["i_type"] == 2), None, [
            f("rho", one("rho")), f("ptype", one("p_type"), "int"), f("pres", one("pres"))]),
        abi.VTK_MOVING_UY: (abi.SPH, lambda s1: s1["p_type"] == 9, None,
                            [f("rho", one("rho")), f("uy", one("uy"))]),
    }


def test_every_vtk_writer_layout_and_values(tmp_path):
    """The eight writers of functions_OUTPUT.cuh beyond the fluid and DEM
    ones: header, point count, array names / components / types in the
    reference's order, and every value, against the particle records (float
    arrays as the float rounding of the record, int arrays exact).  The
    small ISPH-DEM case holds walls, fluid and spheres; the air, fluid2 and
    moving writers select particle types it lacks and write their layout
    with a count of zero, as the reference does."""
    g = _stepped()
    counts = {}
    for kind, (sp, sel_fn, cnt_fn, arrays) in _kinds().items():
        path = tmp_path / f"k{kind}.vtk"
        g.write_vtk(path, kind)
        v = read_vtk(path)
        s1, s2, s3 = g.get_particles(sp)
        sel = sel_fn(s1)
        n = int((cnt_fn or sel_fn)(s1).sum())
        counts[kind] = n
        assert v["header"][:4] == ["# vtk DataFile Version 3.0", "Print out results in vtk format", "BINARY",
                                   "DATASET POLYDATA"]
        assert v["header"][4] == f"POINTS\t{n}\tfloat" and v["point_data"] == f"POINT_DATA\t{n}", kind
        assert v["field_line"] == f"FIELD FieldData\t{len(arrays)}", kind
        assert [(a["name"], a["ncomp"], a["type"]) for a in v["fields"]] == [a[:3] for a in arrays], kind
        assert v["end"] == v["bytes"], kind
        pts = np.array(v["points"], np.float32).reshape(-1, 3)
        ref = np.stack([s1["x"][sel], s1["y"][sel], s1["z"][sel]], 1).astype(np.float32)
        assert np.array_equal(pts, ref), kind
        for a, (name, ncomp, typ, src) in zip(v["fields"], arrays):
            want = np.asarray(src(s1, s2, s3))[sel].reshape(-1)
            got = np.array(a["values"])
            if typ == "int":
                assert np.array_equal(got, want), (kind, name)
            else:
                assert np.array_equal(got.astype(np.float32), want.astype(np.float32)), (kind, name)
    assert counts[abi.VTK_BOUNDARY] > 0 and counts[abi.VTK_SOLID] > 0 and counts[abi.VTK_FLUID_PGF] > 0
    with pytest.raises(RuntimeError):
        g.write_vtk(tmp_path / "bad.vtk", 10)


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
