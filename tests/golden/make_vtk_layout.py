"""Writes reference_dem_vtk_layout.json from one of the reference's own DEM
output files (plotdata/demDEM_*_0stp.vtk, written by
save_plot_fluid_vtk_bin_solid_dem, functions_OUTPUT.cuh:1135-1446): the
header lines, the field sequence with component counts and types, the point
count, the file length and the first three values of each array, so that the
layout check does not need the reference tree.

    python tests/golden/make_vtk_layout.py /root/reference/plotdata/demDEM_decouple1_dt5E-07_0stp.vtk
"""
import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[2]))
from tests.helpers import read_vtk  # noqa: E402


def main(src):
    v = read_vtk(src)
    out = {
        "source": Path(src).name,
        "header": v["header"],
        "point_data": v["point_data"],
        "field_line": v["field_line"],
        "n_points": len(v["points"]) // 3,
        "bytes": v["bytes"],
        "points_head": [float(x) for x in v["points"][:6]],
        "fields": [{"line": f["line"], "name": f["name"], "ncomp": f["ncomp"], "count": f["count"],
                    "type": f["type"], "head": [float(x) for x in f["values"][:3]]} for f in v["fields"]],
    }
    dst = Path(__file__).resolve().parent / "reference_dem_vtk_layout.json"
    dst.write_text(json.dumps(out, indent=1) + "\n")
    print(dst)


if __name__ == "__main__":
    main(sys.argv[1])
