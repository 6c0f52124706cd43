"""Host-side mirror of the SOPHIA per-stage call surface over the C ABI.

The reference advances a run through SOPHIA_single_ISPH
(ISPH_Calc_seperate.cuh:13-503), whose stages are NNPS, KNL, INTERACTION,
PPE, DEM_INTERACTION, SPH_DEM_COUPLING and TIME.  `Sophia` exposes the same
stages as methods over one device context; every call goes to the HIP
library (libsophia.so) and raises `SophiaError` on failure -- there is no CPU
fallback on this path.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import abi
from .abi import (CORRECT, DEM, PART1, PART2, PART3, PREC_MIXED, PREDICT, SPH, Domain, Params, PpeOpts,
                  PpeStats, check, ptr)


class Sophia:
    """One device context: the state of one SPH-DEM run on one GPU."""

    def __init__(self, params: Params, particles: np.ndarray, device: int = 0, precision: int = PREC_MIXED):
        if particles.dtype != PART1:
            raise TypeError("particles must be a part1 record array (fluidized_bed_master_amd.abi.PART1)")
        self.lib = abi.load_library()
        self._ctx = C.c_void_p()
        parts = np.ascontiguousarray(particles)
        check(self.lib.sophia_create_from_arrays(C.byref(params), ptr(parts), len(parts), device, precision,
                                                 C.byref(self._ctx)), "sophia_create_from_arrays")
        self.precision = precision

    @classmethod
    def from_files(cls, solv: str, input_txt: str, ptype: str | None = None, data: str | None = None,
                   device: int = 0, precision: int = PREC_MIXED) -> "Sophia":
        """The reference's start-up from solv.txt, input.txt and, when given,
        p_type.txt and data.txt."""
        self = cls.__new__(cls)
        self.lib = abi.load_library()
        self._ctx = C.c_void_p()
        enc = (lambda x: None if x is None else x.encode())
        check(self.lib.sophia_create(solv.encode(), input_txt.encode(), enc(ptype), enc(data), device, precision,
                                     C.byref(self._ctx)), "sophia_create")
        self.precision = precision
        return self

    def matvec_probe(self, variant: int, x: np.ndarray, reps: int = 20):
        """(y, mean microseconds per launch) of the inner FP32 matvec: variant 0
        stored operator, 1 matrix-free."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        us = C.c_double()
        check(self.lib.sophia_ppe_matvec_probe(self._ctx, variant, reps, ptr(x), ptr(y), len(x), C.byref(us)),
              "sophia_ppe_matvec_probe")
        return y, us.value

    def set_table(self, table: "Table") -> None:
        check(self.lib.sophia_set_table(self._ctx, C.byref(table)), "sophia_set_table")

    def close(self) -> None:
        if getattr(self, "_ctx", None):
            self.lib.sophia_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- metadata -----------------------------------------------------------
    @property
    def domain(self) -> Domain:
        d = Domain()
        check(self.lib.sophia_get_domain(self._ctx, C.byref(d)), "sophia_get_domain")
        return d

    @property
    def params(self) -> Params:
        p = Params()
        check(self.lib.sophia_get_params(self._ctx, C.byref(p)), "sophia_get_params")
        return p

    def count(self, species: int) -> int:
        n = C.c_int32()
        check(self.lib.sophia_count(self._ctx, species, C.byref(n)), "sophia_count")
        return n.value

    def clock(self):
        c, t = C.c_int64(), C.c_double()
        check(self.lib.sophia_get_clock(self._ctx, C.byref(c), C.byref(t)), "sophia_get_clock")
        return c.value, t.value

    def sync(self) -> None:
        check(self.lib.sophia_sync(self._ctx), "sophia_sync")

    # --- stages (reference order) -------------------------------------------
    def nnps(self) -> None:
        check(self.lib.sophia_nnps(self._ctx), "sophia_nnps")

    def prep(self) -> None:
        check(self.lib.sophia_prep(self._ctx), "sophia_prep")

    def interaction(self) -> None:
        check(self.lib.sophia_interaction(self._ctx), "sophia_interaction")

    def dem_interaction(self) -> None:
        check(self.lib.sophia_dem_interaction(self._ctx), "sophia_dem_interaction")

    def sph_dem_coupling(self) -> None:
        check(self.lib.sophia_sph_dem_coupling(self._ctx), "sophia_sph_dem_coupling")

    def time(self, phase: int) -> None:
        check(self.lib.sophia_time(self._ctx, phase), "sophia_time")

    def predict(self) -> None:
        self.time(PREDICT)

    def correct(self) -> None:
        self.time(CORRECT)

    def ppe(self, opts: Optional[PpeOpts] = None) -> dict:
        st = PpeStats()
        check(self.lib.sophia_ppe(self._ctx, C.byref(opts) if opts is not None else None, C.byref(st)),
              "sophia_ppe")
        return st.as_dict()

    def xsph(self) -> None:
        check(self.lib.sophia_xsph(self._ctx), "sophia_xsph")

    def step(self, opts: Optional[PpeOpts] = None) -> dict:
        st = PpeStats()
        check(self.lib.sophia_step(self._ctx, C.byref(opts) if opts is not None else None, C.byref(st)),
              "sophia_step")
        return st.as_dict()

    def run(self, n: int, opts: Optional[PpeOpts] = None) -> dict:
        st = PpeStats()
        check(self.lib.sophia_run(self._ctx, n, C.byref(opts) if opts is not None else None, C.byref(st)),
              "sophia_run")
        return st.as_dict()

    # --- inspection ----------------------------------------------------------
    def get_nnps(self, species: int = SPH) -> dict:
        n = self.count(species)
        nc = self.domain.num_cells
        keys = np.zeros(n, np.int32)
        order = np.zeros(n, np.int32)
        cs = np.zeros(nc, np.uint32)
        ce = np.zeros(nc, np.uint32)
        cnt = np.zeros(n, np.int32) if species == SPH else None
        check(self.lib.sophia_get_nnps(self._ctx, species, ptr(keys), ptr(order), ptr(cs), ptr(ce), ptr(cnt)),
              "sophia_get_nnps")
        return {"keys": keys, "order": order, "cell_start": cs, "cell_end": ce, "nbr_count": cnt}

    def get_particles(self, species: int = SPH):
        n = self.count(species)
        p1 = np.zeros(n, PART1)
        p2 = np.zeros(n, PART2)
        p3 = np.zeros(n, PART3)
        check(self.lib.sophia_get_particles(self._ctx, species, ptr(p1), ptr(p2), ptr(p3)), "sophia_get_particles")
        return p1, p2, p3

    def set_pressure(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float64)
        check(self.lib.sophia_set_pressure(self._ctx, ptr(p), len(p)), "sophia_set_pressure")

    def get_ppe_system(self):
        n = self.count(SPH)
        d = np.zeros(n)
        b = np.zeros(n)
        r = np.zeros(n, np.int32)
        check(self.lib.sophia_get_ppe_system(self._ctx, ptr(d), ptr(b), ptr(r)), "sophia_get_ppe_system")
        return d, b, r

    def ppe_apply(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        check(self.lib.sophia_ppe_apply(self._ctx, ptr(x), ptr(y), len(x)), "sophia_ppe_apply")
        return y

    def wcsph_stage(self, stage: int):
        """One stage (0-8) of the WCSPH step (solver_type 0)."""
        check(self.lib.sophia_wcsph_stage(self._ctx, stage), "sophia_wcsph_stage")

    def write_vtk(self, path, kind: int = 0) -> None:
        """Legacy binary VTK in the layout of one of the reference's writers
        (functions_OUTPUT.cuh): abi.VTK_FLUID (0) and abi.VTK_DEM (1) are the
        two its SPH-DEM drivers call; abi.VTK_AIR ... abi.VTK_MOVING_UY the
        others (air, solid, moving, fluid2, boundary, pressure-gradient fluid,
        i_type 1|2 and the second moving writer)."""
        check(self.lib.sophia_write_vtk(self._ctx, str(path).encode(), kind), "sophia_write_vtk")

    def write_restart(self, path) -> None:
        """restart.txt of save_restart (functions_OUTPUT.cuh:1926-1955)."""
        check(self.lib.sophia_write_restart(self._ctx, str(path).encode()), "sophia_write_restart")

    def monitors(self, species: int = SPH) -> np.ndarray:
        """max |u|, max rho, max ftotal of one species."""
        out = np.zeros(3)
        check(self.lib.sophia_monitors(self._ctx, species, ptr(out)), "sophia_monitors")
        return out

    def ppe_coarse(self):
        """(dims, E, X) of the last two-level solve: aggregates per axis and the
        row-major FP32 coarse operator and its approximate inverse."""
        dims = (C.c_int32 * 3)()
        check(self.lib.sophia_ppe_coarse(self._ctx, dims, None, None), "sophia_ppe_coarse")
        na = dims[0] * dims[1] * dims[2]
        E = np.zeros((na, na), np.float32)
        X = np.zeros((na, na), np.float32)
        check(self.lib.sophia_ppe_coarse(self._ctx, dims, ptr(E), ptr(X)), "sophia_ppe_coarse")
        return tuple(dims), E, X

    def ppe_layout(self) -> dict:
        nnz = C.c_int64()
        ib = C.c_int32()
        cb = C.c_int32()
        check(self.lib.sophia_ppe_layout(self._ctx, C.byref(nnz), C.byref(ib), C.byref(cb)), "sophia_ppe_layout")
        return {"nnz": nnz.value, "index_bytes": ib.value, "coef_bytes": cb.value}

    # --- learned warm start ----------------------------------------------------
    @staticmethod
    def mlp_param_count() -> int:
        n = C.c_int32()
        check(abi.load_library().sophia_mlp_param_count(C.byref(n)), "sophia_mlp_param_count")
        return n.value

    def mlp_set_params(self, w: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        check(self.lib.sophia_mlp_set_params(self._ctx, ptr(w), len(w)), "sophia_mlp_set_params")

    def mlp_get_params(self) -> np.ndarray:
        w = np.zeros(self.mlp_param_count(), np.float32)
        check(self.lib.sophia_mlp_get_params(self._ctx, ptr(w), len(w)), "sophia_mlp_get_params")
        return w

    def mlp_predict(self) -> np.ndarray:
        out = np.zeros(self.count(SPH))
        check(self.lib.sophia_mlp_predict(self._ctx, ptr(out), len(out)), "sophia_mlp_predict")
        return out

    def mlp_train_step(self, target: np.ndarray, lr: float = 1e-3) -> float:
        t = np.ascontiguousarray(target, np.float64)
        loss = C.c_double()
        check(self.lib.sophia_mlp_train_step(self._ctx, ptr(t), len(t), lr, C.byref(loss)), "sophia_mlp_train_step")
        return loss.value

    def mlp_loss_grad(self, target: np.ndarray):
        t = np.ascontiguousarray(target, np.float64)
        g = np.zeros(self.mlp_param_count(), np.float32)
        loss = C.c_double()
        check(self.lib.sophia_mlp_loss_grad(self._ctx, ptr(t), len(t), C.byref(loss), ptr(g), len(g)),
              "sophia_mlp_loss_grad")
        return loss.value, g

    def mlp_reset_adam(self) -> None:
        check(self.lib.sophia_mlp_reset_adam(self._ctx), "sophia_mlp_reset_adam")

    # --- device timers -----------------------------------------------------------
    def timers(self, on: bool = True) -> None:
        check(self.lib.sophia_timers_enable(self._ctx, 1 if on else 0), "sophia_timers_enable")

    def timers_report(self) -> dict:
        """{kernel group: (launches, total ms)} measured with HIP events."""
        buf = C.create_string_buffer(1 << 16)
        check(self.lib.sophia_timers_report(self._ctx, buf, len(buf)), "sophia_timers_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out


def knl_eval(kernel_type: int, dim: int, h: float, r: np.ndarray):
    """W and dW/dr on the device (functions_KNL.cuh:47-173)."""
    lib = abi.load_library()
    r = np.ascontiguousarray(r, np.float64)
    w = np.zeros_like(r)
    dw = np.zeros_like(r)
    check(lib.sophia_knl_eval(kernel_type, dim, h, ptr(r), len(r), ptr(w), ptr(dw)), "sophia_knl_eval")
    return w, dw


class Table(C.Structure):
    """sophia_table: data.txt property blocks by p_type."""
    _fields_ = [("n_tables", C.c_int32), ("n_values", C.c_int32),
                ("index", C.c_int32 * 10), ("size", C.c_int32 * 10),
                ("T", C.c_double * 50), ("h", C.c_double * 50), ("k", C.c_double * 50), ("mu", C.c_double * 50)]

    def block(self, p_type: int) -> dict:
        a, n = self.index[p_type], self.size[p_type]
        return {f: np.array(getattr(self, f)[a:a + n]) for f in ("T", "h", "k", "mu")}

    def eval(self, p_type: int, what: int, x: float) -> float:
        y = C.c_double()
        check(abi.load_library().sophia_table_eval(C.byref(self), p_type, what, x, C.byref(y)), "sophia_table_eval")
        return y.value


PROP_MU, PROP_K, PROP_H, PROP_T = 0, 1, 2, 3


def read_table(path: str) -> Table:
    t = Table()
    check(abi.load_library().sophia_read_table(path.encode(), C.byref(t)), "sophia_read_table")
    return t


def read_solv(path: str) -> Params:
    p = abi.default_params()
    check(abi.load_library().sophia_read_solv(path.encode(), C.byref(p)), "sophia_read_solv")
    return p


def read_input(path: str) -> np.ndarray:
    lib = abi.load_library()
    n = C.c_int32()
    check(lib.sophia_read_input(path.encode(), None, 0, C.byref(n)), "sophia_read_input")
    out = np.zeros(n.value, PART1)
    check(lib.sophia_read_input(path.encode(), ptr(out), n.value, C.byref(n)), "sophia_read_input")
    return out


def read_ptype(path: str) -> np.ndarray:
    lib = abi.load_library()
    n = C.c_int32()
    check(lib.sophia_read_ptype(path.encode(), None, 0, C.byref(n)), "sophia_read_ptype")
    out = np.zeros(n.value, np.int32)
    check(lib.sophia_read_ptype(path.encode(), ptr(out), n.value, C.byref(n)), "sophia_read_ptype")
    return out
