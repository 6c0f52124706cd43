"""Shared helpers of the test suite."""
import numpy as np

from tests.conftest import rel_err  # noqa: F401

# fields compared after each stage, by species
SPH_STAGE_FIELDS = {
    "prep": ["DEMpor", "dDEMpor", "dDEMpor_prev", "rho", "m", "flt_s", "grad_rhox", "grad_rhoy", "grad_rhoz"],
    "interaction": ["pgf_x", "pgf_y", "pgf_z"],
    "ppe": ["pres", "PPE1", "PPE2", "PPE3", "PPE4", "rho"],
    "correct": ["ux", "uy", "uz", "x", "y", "z", "pres"],
    "xsph": ["ux", "uy", "uz", "temp", "XSPH_ux", "XSPH_uz"],
}
SPH_STAGE_P3 = {
    "prep": ["Cm"],
    "interaction": ["ftotalx", "ftotaly", "ftotalz", "ftotal"],
    "coupling": ["ftotalx", "ftotaly", "ftotalz", "ftotal"],
    "ppe": ["fpx", "fpy", "fpz"],
}
DEM_STAGE_FIELDS = {
    "prep": ["flt_s", "flt_sd", "flt_sd_2", "DEMpor"],
    "dem_interaction": ["ct_idx", "del_s"],
    "coupling": ["DEMpor", "Fdx_df", "Fdy_df", "Fdz_df", "Fdx_b", "Fdy_b", "Fdz_b", "test6"],
    "predict": ["ux", "uy", "uz"],
    "correct": ["x", "y", "z", "ux", "uy", "uz"],
}
DEM_STAGE_P3 = {
    "dem_interaction": ["ftotalx", "ftotaly", "ftotalz", "torqx", "torqy", "torqz"],
    "coupling": ["ftotalx", "ftotaly", "ftotalz"],
}


def compare_fields(a, b, fields, tol, where, floor=0.0, mask=None, report=None):
    worst = {}
    for f in fields:
        x, y = a[f], b[f]
        if mask is not None:
            x, y = x[mask], y[mask]
        e = rel_err(x, y, floor)
        worst[f] = e
        if report is not None:
            report[f"{where}:{f}"] = e
        assert e <= tol, f"{where}: field {f} rel err {e:.3e} > {tol:.1e}"
    return worst


def read_vtk(path):
    """Parse a legacy VTK 3.0 binary POLYDATA file of the reference's output
    layout (functions_OUTPUT.cuh): five header lines ending in "POINTS n
    float", 3n big-endian floats, "POINT_DATA n", "FIELD FieldData k", then
    k arrays "name ncomp count type" of big-endian 4-byte words."""
    import struct
    b = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        e = b.index(b"\n", pos)
        s = b[pos:e].decode("ascii")
        pos = e + 1
        return s

    header = [line() for _ in range(5)]
    n = int(header[4].split("\t")[1])
    points = struct.unpack(">%df" % (3 * n), b[pos:pos + 12 * n])
    pos += 12 * n
    point_data = line()
    field_line = line()
    k = int(field_line.split("\t")[1])
    fields = []
    for _ in range(k):
        l = line()
        name, ncomp, count, typ = l.split("\t")
        cnt = int(count) * int(ncomp)
        vals = struct.unpack(">%d%s" % (cnt, "f" if typ == "float" else "i"), b[pos:pos + 4 * cnt])
        pos += 4 * cnt
        fields.append({"line": l, "name": name, "ncomp": int(ncomp), "count": int(count), "type": typ,
                       "values": vals})
    return {"header": header, "points": points, "point_data": point_data, "field_line": field_line,
            "fields": fields, "bytes": len(b), "end": pos}
