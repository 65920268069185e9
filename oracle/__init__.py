"""ctypes binding of the CPU restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg import this module, as the oracle the
library's results are compared with or as the timed CPU baseline.  The product (``glsm_fluid_3d_amd``) never imports it.

The classes mirror the reference's ``Gel`` / ``Coupler`` / ``Fluid``
(src/gel.h, src/coupling.h, src/fluid.h) and use the field numbering of
include/glsm_fluid.h, so a test can drive the oracle and the HIP library with
the same call sequence.  ``prec`` selects the fluid arithmetic: 64 (the FP64
oracle) or 32 (the reference's FP32 arithmetic).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# field numbering of include/glsm_fluid.h
GEL_FIELDS = [
    "um", "vm", "wm", "um_norm", "vm_norm", "wmp", "T0", "T1", "T2", "vol", "pm",
    "rm", "rm_loc", "nmSm", "rn", "Veln", "Vels", "Fn", "Fn_robin", "Fdrag_robin",
    "un_norm", "un_robin", "vn_norm", "map_node", "map_element", "bIndex",
    "center_vm", "center_wm", "center_rm", "center_Fn", "center_Veln", "filament",
]
GEL_FIELD = {n: i for i, n in enumerate(GEL_FIELDS)}
CP_FIELDS = ["lag", "Ul", "Vl", "Fdrag", "Frep", "Cl", "Dl", "Sl", "owner", "bIndex"]
CP_FIELD = {n: i for i, n in enumerate(CP_FIELDS)}
FL_FIELDS = ["u", "c", "f", "fpost"]
FL_FIELD = {n: i for i, n in enumerate(FL_FIELDS)}

_ELEM_D = {"um", "vm", "wm", "um_norm", "vm_norm", "wmp", "T0", "T1", "T2", "vol", "pm"}
_NODE_D3 = {"rn", "Veln", "Vels", "Fn", "Fn_robin", "Fdrag_robin"}
_NODE_D = {"un_norm", "un_robin", "vn_norm"}


def gel_field_shape(name: str, nN: int, nE: int, M: int, maxfil: int = 300):
    """(numpy dtype, shape) of a gel field; shared by the oracle and the GPU binding."""
    if name in _ELEM_D:
        return np.float64, (nE,)
    if name == "rm":
        return np.float64, (nE, 3)
    if name == "rm_loc":
        return np.float64, (nE * 8, 3)
    if name == "nmSm":
        return np.float64, (nE * 6, 3)
    if name in _NODE_D3:
        return np.float64, (nN, 3)
    if name in _NODE_D:
        return np.float64, (nN,)
    if name == "map_node":
        return np.int32, (nN,)
    if name == "map_element":
        return np.int32, (nE,)
    if name == "bIndex":
        return np.int32, (M,)
    if name in ("center_vm", "center_wm"):
        return np.float64, (1000,)
    if name in ("center_rm", "center_Fn", "center_Veln"):
        return np.float64, (1000, 3)
    if name == "filament":
        return np.float64, (1000 * maxfil, 3)
    raise KeyError(name)


_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (g++ -fopenmp)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        sig = {
            "oracle_set_threads": (i, [i]),
            "oracle_gel_create": (vp, [i, i, i, d, d, d, i, i]),
            "oracle_gel_destroy": (None, [vp]),
            "oracle_gel_step_elasticity": (None, [vp, i]),
            "oracle_gel_step_chemistry": (None, [vp, i]),
            "oracle_gel_record_center": (None, [vp, i]),
            "oracle_gel_reset_to_quiescent": (None, [vp]),
            "oracle_gel_fire_excitation_pulse": (None, [vp]),
            "oracle_gel_set_anchor_z": (None, [vp, d]),
            "oracle_gel_build_tube_mask": (None, [vp, i, i]),
            "oracle_gel_perturb_chemistry": (None, [vp, C.c_uint64, d]),
            "oracle_gel_set_chemistry_rates": (None, [vp, d, d]),
            "oracle_gel_dims": (None, [vp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
            "oracle_gel_params": (None, [vp, C.POINTER(d), C.POINTER(d), C.POINTER(d)]),
            "oracle_gel_read": (i, [vp, i, vp, C.c_size_t]),
            "oracle_gel_write": (i, [vp, i, vp, C.c_size_t]),
            "oracle_coupler_create": (vp, [C.POINTER(vp), i, i]),
            "oracle_coupler_destroy": (None, [vp]),
            "oracle_coupler_pack_from_gels": (None, [vp]),
            "oracle_coupler_scatter_to_gels": (None, [vp]),
            "oracle_coupler_apply_gel_repulsion": (None, [vp]),
            "oracle_coupler_marker_count": (i, [vp]),
            "oracle_coupler_read": (i, [vp, i, C.POINTER(d), C.c_size_t]),
            "oracle_coupler_write": (i, [vp, i, C.POINTER(d), C.c_size_t]),
            "oracle_fluid_create": (vp, [i, i, i, vp]),
            "oracle_fluid_destroy": (None, [vp]),
            "oracle_fluid_step_velocity": (None, [vp, i]),
            "oracle_fluid_sub_step": (None, [vp, d]),
            "oracle_fluid_step_concentration": (None, [vp, i]),
            "oracle_fluid_set_periodic": (None, [vp, i, i, i]),
            "oracle_fluid_set_ibm": (None, [vp, d, d]),
            "oracle_fluid_set_force": (None, [vp, d, d, d]),
            "oracle_fluid_dims": (None, [vp, C.POINTER(i), C.POINTER(i)]),
            "oracle_fluid_beta_eff": (d, [vp, i]),
            "oracle_fluid_read": (i, [vp, i, C.POINTER(d), C.c_size_t]),
            "oracle_fluid_write": (i, [vp, i, C.POINTER(d), C.c_size_t]),
            "oracle_system_step": (None, [C.POINTER(vp), i, vp, vp, i]),
            "oracle_delta4_f64": (d, [d]),
            "oracle_delta4_f32": (C.c_float, [C.c_float]),
            "oracle_lattice_tables": (None, [C.POINTER(i), C.POINTER(d), C.POINTER(i)]),
            "oracle_gel_phase": (i, [vp, i]),
            "oracle_solve_bilinear": (i, [C.POINTER(d), d, C.POINTER(d), C.POINTER(d)]),
            "oracle_fu": (d, [d, d, d, d]),
            "oracle_fv": (d, [d, d, d]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Gel:
    """Oracle counterpart of Gel (src/gel.h:33-165)."""

    def __init__(self, size, pos, gel_type=1, gel_id=1):
        self.h = lib().oracle_gel_create(int(size[0]), int(size[1]), int(size[2]),
                                         float(pos[0]), float(pos[1]), float(pos[2]),
                                         int(gel_type), int(gel_id))
        g3 = (C.c_int * 3)()
        nN, nE, M = C.c_int(), C.c_int(), C.c_int()
        lib().oracle_gel_dims(self.h, g3, C.byref(nN), C.byref(nE), C.byref(M))
        self.node_grid = tuple(g3)
        self.num_nodes, self.num_elements, self.boundary_count = nN.value, nE.value, M.value

    def close(self):
        if self.h:
            lib().oracle_gel_destroy(self.h)
            self.h = None

    __del__ = close

    def step_elasticity(self, it):
        lib().oracle_gel_step_elasticity(self.h, int(it))

    def step_chemistry(self, it):
        lib().oracle_gel_step_chemistry(self.h, int(it))

    def record_center(self, it):
        lib().oracle_gel_record_center(self.h, int(it))

    def reset_to_quiescent(self):
        lib().oracle_gel_reset_to_quiescent(self.h)

    def fire_excitation_pulse(self):
        lib().oracle_gel_fire_excitation_pulse(self.h)

    def set_anchor_z(self, z):
        lib().oracle_gel_set_anchor_z(self.h, float(z))

    def build_tube_mask(self, wall, circular=False):
        lib().oracle_gel_build_tube_mask(self.h, int(wall), int(bool(circular)))

    def perturb_chemistry(self, seed, amplitude):
        lib().oracle_gel_perturb_chemistry(self.h, int(seed), float(amplitude))

    def set_chemistry_rates(self, f, ep):
        lib().oracle_gel_set_chemistry_rates(self.h, float(f), float(ep))

    PHASES = ["service_nodes", "element_properties", "chem_boundary", "unnorm", "pressure", "nodes_velocity",
              "nodes_position", "terms", "chem", "set_zero"]

    def phase(self, name):
        """Run one phase of the gel step alone (calServiceNodesPositionD,
        calElementPropertiesD, calChemBoundaryD, calUnnormD, calPressureD,
        calNodesVelocityD, calInternalNodesPositionD, calTermsD, calChemD,
        setZero)."""
        rc = lib().oracle_gel_phase(self.h, self.PHASES.index(name))
        if rc != 0:
            raise RuntimeError(f"oracle_gel_phase({name}) -> {rc}")

    def steady_state(self):
        u, v, w = C.c_double(), C.c_double(), C.c_double()
        lib().oracle_gel_params(self.h, C.byref(u), C.byref(v), C.byref(w))
        return u.value, v.value, w.value

    def read(self, name):
        dt, shape = gel_field_shape(name, self.num_nodes, self.num_elements, self.boundary_count)
        a = np.empty(shape, dtype=dt)
        rc = lib().oracle_gel_read(self.h, GEL_FIELD[name], a.ctypes.data, a.nbytes)
        if rc != 0:
            raise RuntimeError(f"oracle_gel_read({name}) -> {rc}")
        return a

    def write(self, name, arr):
        dt, shape = gel_field_shape(name, self.num_nodes, self.num_elements, self.boundary_count)
        a = np.ascontiguousarray(np.asarray(arr, dtype=dt).reshape(shape))
        rc = lib().oracle_gel_write(self.h, GEL_FIELD[name], a.ctypes.data, a.nbytes)
        if rc != 0:
            raise RuntimeError(f"oracle_gel_write({name}) -> {rc}")


class Coupler:
    """Oracle counterpart of Coupler (src/coupling.h:14-58)."""

    def __init__(self, gels, prec=64):
        self.gels = list(gels)
        arr = (C.c_void_p * len(self.gels))(*[g.h for g in self.gels])
        self.prec = prec
        self.h = lib().oracle_coupler_create(arr, len(self.gels), int(prec))
        self.M = lib().oracle_coupler_marker_count(self.h)

    def close(self):
        if self.h:
            lib().oracle_coupler_destroy(self.h)
            self.h = None

    __del__ = close

    def pack_from_gels(self):
        lib().oracle_coupler_pack_from_gels(self.h)

    def scatter_to_gels(self):
        lib().oracle_coupler_scatter_to_gels(self.h)

    def apply_gel_repulsion(self):
        lib().oracle_coupler_apply_gel_repulsion(self.h)

    def read(self, name):
        n3 = name in ("lag", "Ul", "Vl", "Fdrag", "Frep")
        a = np.empty((self.M, 3) if n3 else (self.M,), dtype=np.float64)
        rc = lib().oracle_coupler_read(self.h, CP_FIELD[name], _dptr(a), a.size)
        if rc != 0:
            raise RuntimeError(f"oracle_coupler_read({name}) -> {rc}")
        return a

    def write(self, name, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64))
        rc = lib().oracle_coupler_write(self.h, CP_FIELD[name], _dptr(a), a.size)
        if rc != 0:
            raise RuntimeError(f"oracle_coupler_write({name}) -> {rc}")


class Fluid:
    """Oracle counterpart of Fluid (src/fluid.h:24-84)."""

    def __init__(self, size, coupler: Coupler):
        self.coupler = coupler
        self.h = lib().oracle_fluid_create(int(size[0]), int(size[1]), int(size[2]), coupler.h)
        g3 = (C.c_int * 3)()
        ns = C.c_int()
        lib().oracle_fluid_dims(self.h, g3, C.byref(ns))
        self.grid = tuple(g3)
        self.nsub = ns.value
        self.N = g3[0] * g3[1] * g3[2]

    def close(self):
        if self.h:
            lib().oracle_fluid_destroy(self.h)
            self.h = None

    __del__ = close

    def step_velocity(self, it):
        lib().oracle_fluid_step_velocity(self.h, int(it))

    def sub_step(self, beta_eff):
        lib().oracle_fluid_sub_step(self.h, float(beta_eff))

    def step_concentration(self, it):
        lib().oracle_fluid_step_concentration(self.h, int(it))

    def set_periodic(self, px, py, pz):
        lib().oracle_fluid_set_periodic(self.h, int(px), int(py), int(pz))

    def set_ibm(self, beta, ramp_steps):
        lib().oracle_fluid_set_ibm(self.h, float(beta), float(ramp_steps))

    def set_force(self, fx, fy, fz):
        lib().oracle_fluid_set_force(self.h, float(fx), float(fy), float(fz))

    def beta_eff(self, it):
        return lib().oracle_fluid_beta_eff(self.h, int(it))

    def read(self, name):
        if name == "rho":  # rho - 1 of the pre-collision populations, FP64
            return self.read("f").sum(axis=0) - 1.0
        shape = {"u": (self.N, 3), "c": (self.N,), "f": (19, self.N), "fpost": (19, self.N)}[name]
        a = np.empty(shape, dtype=np.float64)
        rc = lib().oracle_fluid_read(self.h, FL_FIELD[name], _dptr(a), a.size)
        if rc != 0:
            raise RuntimeError(f"oracle_fluid_read({name}) -> {rc}")
        return a

    def write(self, name, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float64))
        rc = lib().oracle_fluid_write(self.h, FL_FIELD[name], _dptr(a), a.size)
        if rc != 0:
            raise RuntimeError(f"oracle_fluid_write({name}) -> {rc}")


def system_step(gels, coupler: Coupler, fluid: Fluid, it: int):
    """One time step in sim.cpp's order (sim.cpp:66-81)."""
    arr = (C.c_void_p * len(gels))(*[g.h for g in gels])
    lib().oracle_system_step(arr, len(gels), coupler.h, fluid.h, int(it))


def delta4(r, prec=64):
    return lib().oracle_delta4_f64(float(r)) if prec == 64 else lib().oracle_delta4_f32(float(r))


def lattice_tables():
    c = (C.c_int * 57)()
    w = (C.c_double * 19)()
    o = (C.c_int * 19)()
    lib().oracle_lattice_tables(c, w, o)
    return np.array(c).reshape(19, 3), np.array(w), np.array(o)


def fu(u, v, w, f=0.9):
    return lib().oracle_fu(u, v, w, f)


def fv(u, v, w):
    return lib().oracle_fv(u, v, w)


def solve_bilinear(corners, v_iso):
    """solveBilinear (gel_kernels.cu:165-214): corners = (A, B, C, D) of the
    first field at (0,0), (1,0), (0,1), (1,1), then (E, F, G, H) of the
    second.  Returns (x, y) in the unit square where both bilinear
    interpolants equal v_iso, or None."""
    a = (C.c_double * 8)(*[float(v) for v in corners])
    x, y = C.c_double(), C.c_double()
    ok = lib().oracle_solve_bilinear(a, float(v_iso), C.byref(x), C.byref(y))
    return (x.value, y.value) if ok else None
