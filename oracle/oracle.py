"""ctypes binding of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
this module, and only as the checker (or the reported CPU baseline); the
product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from fluidized_bed_master_amd.abi import (CORRECT, PART1, PART2, PART3, PREDICT, SPH, Domain, Params, PpeOpts,
                                          PpeStats, ptr)

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

_P = C.POINTER
_VP = C.c_void_p
_SIG = {
    "orc_create": (_VP, [_P(Params), _VP, C.c_int32]),
    "orc_destroy": (None, [_VP]),
    "orc_set_table": (None, [_VP, _VP]),
    "orc_get_domain": (None, [_VP, _P(Domain)]),
    "orc_get_clock": (None, [_VP, _P(C.c_int64), _P(C.c_double)]),
    "orc_count": (C.c_int, [_VP, C.c_int32]),
    "orc_nnps": (None, [_VP]),
    "orc_prep": (None, [_VP]),
    "orc_interaction": (None, [_VP]),
    "orc_dem_interaction": (None, [_VP]),
    "orc_sph_dem_coupling": (None, [_VP]),
    "orc_time": (None, [_VP, C.c_int32]),
    "orc_ppe": (C.c_int, [_VP, _P(PpeOpts), _P(PpeStats)]),
    "orc_xsph": (None, [_VP]),
    "orc_step": (C.c_int, [_VP, _P(PpeOpts), _P(PpeStats)]),
    "orc_get_nnps": (None, [_VP, C.c_int32, _VP, _VP, _VP, _VP, _VP]),
    "orc_get_particles": (None, [_VP, C.c_int32, _VP, _VP, _VP]),
    "orc_wcsph_stage": (C.c_int, [_VP, C.c_int32]),
    "orc_set_pressure": (None, [_VP, _VP, C.c_int32]),
    "orc_get_ppe_system": (None, [_VP, _VP, _VP, _VP]),
    "orc_ppe_apply": (None, [_VP, _VP, _VP]),
    "orc_knl_tmpA": (C.c_double, [C.c_int32, C.c_int32, C.c_double]),
    "orc_knl_eval": (None, [C.c_int32, C.c_int32, C.c_double, _VP, C.c_int32, _VP, _VP]),
    "orc_cell_key": (C.c_int32, [_VP, C.c_double, C.c_double, C.c_double]),
    "orc_brute_neighbor_counts": (None, [_VP, _VP]),
    "orc_set_threads": (None, [C.c_int32]),
    "orc_get_threads": (C.c_int32, []),
}

_lib = None


def load(build_if_missing: bool = True) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists() and build_if_missing:
        subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    lib = C.CDLL(str(LIB_PATH))
    for k, (res, args) in _SIG.items():
        f = getattr(lib, k)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def set_threads(n: int) -> None:
    load().orc_set_threads(n)


def get_threads() -> int:
    return load().orc_get_threads()


class Oracle:
    """Same stage surface as fluidized_bed_master_amd.sophia.Sophia, on the CPU in FP64."""

    def __init__(self, params: Params, particles: np.ndarray):
        self.lib = load()
        parts = np.ascontiguousarray(particles)
        assert parts.dtype == PART1
        self._s = self.lib.orc_create(C.byref(params), ptr(parts), len(parts))
        if not self._s:
            raise RuntimeError("orc_create failed")

    def set_table(self, table):
        self.lib.orc_set_table(self._s, C.byref(table))

    def close(self):
        if self._s:
            self.lib.orc_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def domain(self) -> Domain:
        d = Domain()
        self.lib.orc_get_domain(self._s, C.byref(d))
        return d

    def count(self, species: int) -> int:
        return self.lib.orc_count(self._s, species)

    def clock(self):
        c, t = C.c_int64(), C.c_double()
        self.lib.orc_get_clock(self._s, C.byref(c), C.byref(t))
        return c.value, t.value

    def nnps(self):
        self.lib.orc_nnps(self._s)

    def prep(self):
        self.lib.orc_prep(self._s)

    def interaction(self):
        self.lib.orc_interaction(self._s)

    def dem_interaction(self):
        self.lib.orc_dem_interaction(self._s)

    def sph_dem_coupling(self):
        self.lib.orc_sph_dem_coupling(self._s)

    def time(self, phase):
        self.lib.orc_time(self._s, phase)

    def predict(self):
        self.time(PREDICT)

    def correct(self):
        self.time(CORRECT)

    def ppe(self, opts: PpeOpts | None = None) -> dict:
        st = PpeStats()
        self.lib.orc_ppe(self._s, C.byref(opts) if opts is not None else None, C.byref(st))
        return st.as_dict()

    def xsph(self):
        self.lib.orc_xsph(self._s)

    def step(self, opts: PpeOpts | None = None) -> dict:
        st = PpeStats()
        rc = self.lib.orc_step(self._s, C.byref(opts) if opts is not None else None, C.byref(st))
        if rc:
            raise RuntimeError("oracle step: configuration outside the supported set")
        return st.as_dict()

    def wcsph_stage(self, stage: int):
        if self.lib.orc_wcsph_stage(self._s, stage):
            raise RuntimeError("oracle WCSPH: configuration outside the supported set")

    def get_nnps(self, species: int = SPH, counts: bool = True) -> dict:
        n = self.count(species)
        nc = self.domain.num_cells
        keys = np.zeros(n, np.int32)
        order = np.zeros(n, np.int32)
        cs = np.zeros(nc, np.uint32)
        ce = np.zeros(nc, np.uint32)
        cnt = np.zeros(n, np.int32) if counts else None
        self.lib.orc_get_nnps(self._s, species, ptr(keys), ptr(order), ptr(cs), ptr(ce), ptr(cnt))
        return {"keys": keys, "order": order, "cell_start": cs, "cell_end": ce, "nbr_count": cnt}

    def get_particles(self, species: int = SPH):
        n = self.count(species)
        p1 = np.zeros(n, PART1)
        p2 = np.zeros(n, PART2)
        p3 = np.zeros(n, PART3)
        self.lib.orc_get_particles(self._s, species, ptr(p1), ptr(p2), ptr(p3))
        return p1, p2, p3

    def monitors(self, species: int = SPH) -> np.ndarray:
        """kernel_copy_max_sph/_dem + DeviceReduce::Max (functions_OUTPUT.cuh:35-65):
        max |u|, rho, P3.ftotal over p_type != 0 and i_type < i_type_crt (2);
        excluded particles count as 0."""
        p1, _, p3 = self.get_particles(species)
        keep = (p1["p_type"] != 0) & (p1["i_type"] < 2)
        u = np.sqrt(p1["ux"] ** 2 + p1["uy"] ** 2 + p1["uz"] ** 2)
        out = np.zeros(3)
        for k, a in enumerate((u, p1["rho"], p3["ftotal"])):
            v = np.where(keep, a, 0.0)
            out[k] = max(0.0, float(v.max())) if len(v) else 0.0
        return out

    def set_pressure(self, p):
        p = np.ascontiguousarray(p, np.float64)
        self.lib.orc_set_pressure(self._s, ptr(p), len(p))

    def get_ppe_system(self):
        n = self.count(SPH)
        d = np.zeros(n)
        b = np.zeros(n)
        r = np.zeros(n, np.int32)
        self.lib.orc_get_ppe_system(self._s, ptr(d), ptr(b), ptr(r))
        return d, b, r

    def ppe_apply(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros_like(x)
        self.lib.orc_ppe_apply(self._s, ptr(x), ptr(y))
        return y

    def cell_key(self, x, y, z) -> int:
        return self.lib.orc_cell_key(self._s, x, y, z)

    def brute_neighbor_counts(self) -> np.ndarray:
        c = np.zeros(self.count(SPH), np.int32)
        self.lib.orc_brute_neighbor_counts(self._s, ptr(c))
        return c


def knl_eval(kernel_type: int, dim: int, h: float, r: np.ndarray):
    lib = load()
    r = np.ascontiguousarray(r, np.float64)
    w = np.zeros_like(r)
    dw = np.zeros_like(r)
    lib.orc_knl_eval(kernel_type, dim, h, ptr(r), len(r), ptr(w), ptr(dw))
    return w, dw


def knl_tmpA(kernel_type: int, dim: int, h: float) -> float:
    return load().orc_knl_tmpA(kernel_type, dim, h)
