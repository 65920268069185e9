"""MI355X-native gLSM_fluid_3D time step -- Python binding of the C ABI.

The compute lives in ``libglsm_fluid.so`` (hand-written HIP for gfx950, built
in-tree by ``glsm_fluid_3d_amd.build``).  This module only loads it with
ctypes and mirrors the reference's ``Gel`` / ``Coupler`` / ``Fluid`` classes
(src/gel.h:33-165, src/coupling.h:14-58, src/fluid.h:24-84): same
constructor arguments, step methods and phase controls, plus whole-field
reads and writes.  There is no fallback: if the library is missing or no GPU
is present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libglsm_fluid.so")

GEL_FIELDS = [
    "um", "vm", "wm", "um_norm", "vm_norm", "wmp", "T0", "T1", "T2", "vol", "pm",
    "rm", "rm_loc", "nmSm", "rn", "Veln", "Vels", "Fn", "Fn_robin", "Fdrag_robin",
    "un_norm", "un_robin", "vn_norm", "map_node", "map_element", "bIndex",
    "center_vm", "center_wm", "center_rm", "center_Fn", "center_Veln", "filament",
]
GEL_FIELD = {n: i for i, n in enumerate(GEL_FIELDS)}
CP_FIELDS = ["lag", "Ul", "Vl", "Fdrag", "Frep", "Cl", "Dl", "Sl", "owner", "bIndex"]
CP_FIELD = {n: i for i, n in enumerate(CP_FIELDS)}
FL_FIELDS = ["u", "c", "f", "fpost", "rho"]
FL_FIELD = {n: i for i, n in enumerate(FL_FIELDS)}
STORAGE = {"f32full": 0, "f32": 1, "f16": 2, "bf16": 3}
STORAGE_BYTES = {"f32full": 4, "f32": 4, "f16": 2, "bf16": 2}

# every entry point of include/glsm_fluid.h: name -> (restype, argtypes)
_vp, _i, _d, _sz, _f = C.c_void_p, C.c_int, C.c_double, C.c_size_t, C.c_float
_pi = C.POINTER(C.c_int)
SIGNATURES = {
    "glsm_init": (_i, [_i]),
    "glsm_synchronize": (_i, []),
    "glsm_last_error": (C.c_char_p, []),
    "glsm_version": (C.c_char_p, []),
    "glsm_gel_create": (_i, [_i, _i, _i, _d, _d, _d, _i, _i, _i, C.POINTER(_vp)]),
    "glsm_gel_destroy": (_i, [_vp]),
    "glsm_gel_step_elasticity": (_i, [_vp, _i]),
    "glsm_gel_step_chemistry": (_i, [_vp, _i]),
    "glsm_gel_record_center": (_i, [_vp, _i]),
    "glsm_gel_write_files": (_i, [_vp, _i]),
    "glsm_gel_save_state": (_i, [_vp, C.c_char_p]),
    "glsm_gel_load_state": (_i, [_vp, C.c_char_p]),
    "glsm_gel_reset_to_quiescent": (_i, [_vp]),
    "glsm_gel_fire_excitation_pulse": (_i, [_vp]),
    "glsm_gel_set_anchor_z": (_i, [_vp, _d]),
    "glsm_gel_build_tube_mask": (_i, [_vp, _i, _i]),
    "glsm_gel_get_index": (_i, [_vp, _i, _i, _i, _i]),
    "glsm_gel_dims": (_i, [_vp, _pi, _pi, _pi, _pi]),
    "glsm_gel_perturb_chemistry": (_i, [_vp, C.c_uint64, _d]),
    "glsm_gel_set_chemistry_rates": (_i, [_vp, _d, _d]),
    "glsm_gel_check_finite": (_i, [_vp, C.POINTER(C.c_longlong)]),
    "glsm_gel_field_bytes": (_sz, [_vp, _i]),
    "glsm_gel_read": (_i, [_vp, _i, _vp, _sz]),
    "glsm_gel_write": (_i, [_vp, _i, _vp, _sz]),
    "glsm_coupler_create": (_i, [C.POINTER(_vp), _i, C.POINTER(_vp)]),
    "glsm_coupler_destroy": (_i, [_vp]),
    "glsm_coupler_pack_from_gels": (_i, [_vp]),
    "glsm_coupler_scatter_to_gels": (_i, [_vp]),
    "glsm_coupler_apply_gel_repulsion": (_i, [_vp]),
    "glsm_coupler_marker_count": (_i, [_vp]),
    "glsm_coupler_field_bytes": (_sz, [_vp, _i]),
    "glsm_coupler_read": (_i, [_vp, _i, _vp, _sz]),
    "glsm_fluid_create": (_i, [_i, _i, _i, _i, _vp, C.POINTER(_vp)]),
    "glsm_fluid_destroy": (_i, [_vp]),
    "glsm_fluid_step_velocity": (_i, [_vp, _i]),
    "glsm_fluid_step_concentration": (_i, [_vp, _i]),
    "glsm_fluid_sub_step": (_i, [_vp, _f]),
    "glsm_fluid_write_files": (_i, [_vp, _i]),
    "glsm_fluid_save_state": (_i, [_vp, C.c_char_p]),
    "glsm_fluid_conservation": (_i, [_vp, C.POINTER(_d)]),
    "glsm_fluid_conservation_static": (_i, [_vp, C.POINTER(_d)]),
    "glsm_fluid_load_state": (_i, [_vp, C.c_char_p]),
    "glsm_fluid_check_finite": (_i, [_vp, C.POINTER(C.c_longlong), C.POINTER(_d), C.POINTER(_d)]),
    "glsm_fluid_set_periodic": (_i, [_vp, _i, _i, _i]),
    "glsm_fluid_set_ibm": (_i, [_vp, _f, _f]),
    "glsm_fluid_set_storage": (_i, [_vp, _i, _i]),
    "glsm_fluid_enable_lbm_timing": (_i, [_vp, _i, _i]),
    "glsm_fluid_lbm_timing": (_i, [_vp, C.POINTER(_d), _pi]),
    "glsm_fluid_dims": (_i, [_vp, _pi, _pi]),
    "glsm_fluid_field_bytes": (_sz, [_vp, _i]),
    "glsm_fluid_read": (_i, [_vp, _i, _vp, _sz]),
    "glsm_fluid_write": (_i, [_vp, _i, _vp, _sz]),
    "glsm_set_interp_mode": (_i, [_i]),
    "glsm_set_repulsion_cull": (_i, [_i]),
    "glsm_set_lattice_xcd_run": (_i, [_i]),
    "glsm_set_lbm_lds_reserve": (_i, [_i]),
    "glsm_batch_create": (_i, [_i, _i, C.POINTER(_vp)]),
    "glsm_batch_destroy": (_i, [_vp]),
    "glsm_batch_add": (_i, [_vp, _vp, _vp, _i, _i, C.c_char_p]),
    "glsm_batch_remove": (_i, [_vp, _i]),
    "glsm_batch_size": (_i, [_vp]),
    "glsm_batch_iteration": (_i, [_vp, _i]),
    "glsm_batch_step": (_i, [_vp, _i]),
    "glsm_batch_use_graph": (_i, [_vp, _i]),
    "glsm_batch_set_launch": (_i, [_vp, _i, _i, _i, _i, _i, _i]),
    "glsm_batch_get_launch": (_i, [_vp] + [C.POINTER(_i)] * 6),
    "glsm_batch_autotune": (_i, [_vp, _i, C.POINTER(_d)]),
    "glsm_batch_enable_lbm_timing": (_i, [_vp, _i]),
    "glsm_batch_lbm_timing": (_i, [_vp, C.POINTER(_d), C.POINTER(_i)]),
    "glsm_slab_create": (_i, [_vp, _i, _i, C.POINTER(_vp)]),
    "glsm_slab_destroy": (_i, [_vp]),
    "glsm_slab_load": (_i, [_vp, _vp]),
    "glsm_slab_store": (_i, [_vp, _vp]),
    "glsm_slab_chain_step": (_i, [C.POINTER(_vp), _i, _i, _i]),
    "glsm_slab_face_bytes": (_i, [_vp, C.POINTER(_sz)]),
    "glsm_slab_enable_timing": (_i, [_vp, _i]),
    "glsm_slab_timing": (_i, [_vp, C.POINTER(_d), C.POINTER(_d)]),
    "glsm_output_create": (_i, [_i, C.POINTER(_vp)]),
    "glsm_output_destroy": (_i, [_vp]),
    "glsm_output_gel": (_i, [_vp, _vp, _i, C.c_char_p]),
    "glsm_output_fluid": (_i, [_vp, _vp, _i, C.c_char_p]),
    "glsm_output_release": (_i, [_vp, _vp]),
    "glsm_output_flush": (_i, [_vp]),
    "glsm_output_status": (_i, [_vp, C.POINTER(C.c_long), C.POINTER(C.c_long), C.c_char_p, _sz]),
    "glsm_batch_set_output": (_i, [_vp, _vp]),
    "glsm_device_memory": (_i, [C.POINTER(_sz), C.POINTER(_sz)]),
    "glsm_system_step": (_i, [C.POINTER(_vp), _i, _vp, _vp, _i]),
    "glsm_copy_bandwidth": (_i, [_sz, _i, C.POINTER(_d)]),
}

_lib = None


class GlsmError(RuntimeError):
    pass


def compile_library() -> str:
    """Compile libglsm_fluid.so in-tree (see glsm_fluid_3d_amd/build.py)."""
    import importlib

    return importlib.import_module(__name__ + ".build").build()


def load(path: str = LIB_PATH):
    """Load libglsm_fluid.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise GlsmError(f"{path} is missing: run `python -m glsm_fluid_3d_amd.build` first")
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc < 0:
        msg = load().glsm_last_error().decode(errors="replace")
        raise GlsmError(f"{what} failed ({rc}): {msg}")
    return rc


def init(device: int = 0):
    _check(load().glsm_init(int(device)), "glsm_init")


def set_interp_mode(mode: int):
    """0 = default (staged interpolation; sampling by marker count),
    1 = wavefront per marker, 2 = thread per marker, 3 = products staged in
    LDS (eight threads per marker)."""
    _check(load().glsm_set_interp_mode(int(mode)), "glsm_set_interp_mode")


def set_lbm_lds_reserve(nbytes: int):
    """LDS reserved per stream/collide work-group (caps the work-groups per
    CU): -1 automatic (default), else 0..65536 bytes; results are identical."""
    _check(load().glsm_set_lbm_lds_reserve(int(nbytes)), "glsm_set_lbm_lds_reserve")


def set_lattice_xcd_run(run: int):
    """XCD run length of the single-lattice stream/collide block order: -1
    automatic (default), 0, 4, 16 or 64; results are identical."""
    _check(load().glsm_set_lattice_xcd_run(int(run)), "glsm_set_lattice_xcd_run")


def set_repulsion_cull(min_markers: int):
    """Skip out-of-range gels in the repulsion from this many markers on
    (default 4096; 0 = never); results are identical either way."""
    _check(load().glsm_set_repulsion_cull(int(min_markers)), "glsm_set_repulsion_cull")


def synchronize():
    _check(load().glsm_synchronize(), "glsm_synchronize")


def gel_field_shape(name, nN, nE, M, maxfil=300):
    _ELEM_D = {"um", "vm", "wm", "um_norm", "vm_norm", "wmp", "T0", "T1", "T2", "vol", "pm"}
    _NODE_D3 = {"rn", "Veln", "Vels", "Fn", "Fn_robin", "Fdrag_robin"}
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
    if name in ("un_norm", "un_robin", "vn_norm"):
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


class Gel:
    """Gel(gelSize, gelPosition, gelType, gel_id, time) -- src/gel.h:35."""

    def __init__(self, size, pos, gel_type=1, gel_id=1, time=0):
        h = C.c_void_p()
        _check(load().glsm_gel_create(int(size[0]), int(size[1]), int(size[2]), float(pos[0]),
                                      float(pos[1]), float(pos[2]), int(gel_type), int(gel_id),
                                      int(time), C.byref(h)), "glsm_gel_create")
        self.h = h.value
        g3 = (C.c_int * 3)()
        nN, nE, M = C.c_int(), C.c_int(), C.c_int()
        _check(load().glsm_gel_dims(self.h, g3, C.byref(nN), C.byref(nE), C.byref(M)), "glsm_gel_dims")
        self.node_grid = tuple(g3)
        self.num_nodes, self.num_elements, self.boundary_count = nN.value, nE.value, M.value

    def close(self):
        if getattr(self, "h", None):
            load().glsm_gel_destroy(self.h)
            self.h = None

    def step_elasticity(self, it):
        _check(load().glsm_gel_step_elasticity(self.h, int(it)), "step_elasticity")

    def step_chemistry(self, it):
        _check(load().glsm_gel_step_chemistry(self.h, int(it)), "step_chemistry")

    def record_center(self, it):
        _check(load().glsm_gel_record_center(self.h, int(it)), "record_center")

    def write_files(self, it):
        _check(load().glsm_gel_write_files(self.h, int(it)), "write_files")

    def save_state(self, path):
        """Lossless binary snapshot (every device field, params, flags)."""
        _check(load().glsm_gel_save_state(self.h, os.fsencode(path)), "save_state")

    def load_state(self, path):
        _check(load().glsm_gel_load_state(self.h, os.fsencode(path)), "load_state")

    def reset_to_quiescent(self):
        _check(load().glsm_gel_reset_to_quiescent(self.h), "reset_to_quiescent")

    def fire_excitation_pulse(self):
        _check(load().glsm_gel_fire_excitation_pulse(self.h), "fire_excitation_pulse")

    def set_anchor_z(self, z):
        _check(load().glsm_gel_set_anchor_z(self.h, float(z)), "set_anchor_z")

    def build_tube_mask(self, wall, circular=False):
        _check(load().glsm_gel_build_tube_mask(self.h, int(wall), int(bool(circular))), "build_tube_mask")

    def perturb_chemistry(self, seed, amplitude):
        _check(load().glsm_gel_perturb_chemistry(self.h, int(seed), float(amplitude)), "perturb_chemistry")

    def set_chemistry_rates(self, f, ep):
        """Rate constants f, ep; steady state solved again, chemistry reset to it."""
        _check(load().glsm_gel_set_chemistry_rates(self.h, float(f), float(ep)), "set_chemistry_rates")

    def check_finite(self):
        """Number of non-finite values in rn, Veln, um, vm, wm."""
        n = C.c_longlong()
        _check(load().glsm_gel_check_finite(self.h, C.byref(n)), "gel check_finite")
        return n.value

    def get_index(self, xi, yi, zi, size):
        return _check(load().glsm_gel_get_index(self.h, xi, yi, zi, size), "get_index")

    def read(self, name):
        dt, shape = gel_field_shape(name, self.num_nodes, self.num_elements, self.boundary_count)
        a = np.empty(shape, dtype=dt)
        _check(load().glsm_gel_read(self.h, GEL_FIELD[name], a.ctypes.data, a.nbytes), f"gel read {name}")
        return a

    def write(self, name, arr):
        dt, shape = gel_field_shape(name, self.num_nodes, self.num_elements, self.boundary_count)
        a = np.ascontiguousarray(np.asarray(arr, dtype=dt).reshape(shape))
        _check(load().glsm_gel_write(self.h, GEL_FIELD[name], a.ctypes.data, a.nbytes), f"gel write {name}")


class Coupler:
    """Coupler(std::vector<Gel*>&) -- src/coupling.h:16."""

    def __init__(self, gels):
        self.gels = list(gels)
        arr = (C.c_void_p * len(self.gels))(*[g.h for g in self.gels])
        h = C.c_void_p()
        _check(load().glsm_coupler_create(arr, len(self.gels), C.byref(h)), "glsm_coupler_create")
        self.h = h.value
        self.M = _check(load().glsm_coupler_marker_count(self.h), "marker_count")

    def close(self):
        if getattr(self, "h", None):
            load().glsm_coupler_destroy(self.h)
            self.h = None

    def pack_from_gels(self):
        _check(load().glsm_coupler_pack_from_gels(self.h), "pack_from_gels")

    def scatter_to_gels(self):
        _check(load().glsm_coupler_scatter_to_gels(self.h), "scatter_to_gels")

    def apply_gel_repulsion(self):
        _check(load().glsm_coupler_apply_gel_repulsion(self.h), "apply_gel_repulsion")

    def read(self, name):
        if name in ("lag", "Ul", "Vl", "Fdrag", "Frep"):
            a = np.empty((self.M, 3), dtype=np.float32)
        elif name in ("owner", "bIndex"):
            a = np.empty((self.M,), dtype=np.int32)
        else:
            a = np.empty((self.M,), dtype=np.float32)
        _check(load().glsm_coupler_read(self.h, CP_FIELD[name], a.ctypes.data, a.nbytes), f"coupler read {name}")
        return a


class Fluid:
    """Fluid(int3 fluidSize, int time, Coupler*) -- src/fluid.h:26."""

    def __init__(self, size, coupler: Coupler, time=0):
        self.coupler = coupler
        h = C.c_void_p()
        _check(load().glsm_fluid_create(int(size[0]), int(size[1]), int(size[2]), int(time), coupler.h,
                                        C.byref(h)), "glsm_fluid_create")
        self.h = h.value
        g3 = (C.c_int * 3)()
        ns = C.c_int()
        _check(load().glsm_fluid_dims(self.h, g3, C.byref(ns)), "glsm_fluid_dims")
        self.grid = tuple(g3)
        self.nsub = ns.value
        self.N = g3[0] * g3[1] * g3[2]

    def close(self):
        if getattr(self, "h", None):
            load().glsm_fluid_destroy(self.h)
            self.h = None

    def step_velocity(self, it):
        _check(load().glsm_fluid_step_velocity(self.h, int(it)), "step_velocity")

    def step_concentration(self, it):
        _check(load().glsm_fluid_step_concentration(self.h, int(it)), "step_concentration")

    def sub_step(self, beta_eff):
        _check(load().glsm_fluid_sub_step(self.h, float(beta_eff)), "sub_step")

    def write_files(self, it):
        _check(load().glsm_fluid_write_files(self.h, int(it)), "write_files")

    def conservation(self):
        """FP64 (sum(rho - 1), sum(rho ux), sum(rho uy), sum(rho uz)) of the
        pre-collision populations of the next sub-step."""
        out = (_d * 4)()
        _check(load().glsm_fluid_conservation(self.h, out), "conservation")
        return np.array(out[:], dtype=np.float64)

    def check_finite(self):
        """{"bad_cells": cells with a non-finite stored population, velocity
        or concentration (an overflowed FP16 deviation is stored as Inf),
        "max_abs_dev": largest finite |f_q - w_q|, "limit": largest
        magnitude the storage represents}."""
        n, m, lim = C.c_longlong(), C.c_double(), C.c_double()
        _check(load().glsm_fluid_check_finite(self.h, C.byref(n), C.byref(m), C.byref(lim)), "fluid check_finite")
        return {"bad_cells": n.value, "max_abs_dev": m.value, "limit": lim.value}

    def save_state(self, path):
        """Lossless binary snapshot (stored populations, u, c, params)."""
        _check(load().glsm_fluid_save_state(self.h, os.fsencode(path)), "save_state")

    def load_state(self, path):
        _check(load().glsm_fluid_load_state(self.h, os.fsencode(path)), "load_state")

    def set_periodic(self, px, py, pz):
        _check(load().glsm_fluid_set_periodic(self.h, int(px), int(py), int(pz)), "set_periodic")

    def set_ibm(self, beta, ramp_steps):
        _check(load().glsm_fluid_set_ibm(self.h, float(beta), float(ramp_steps)), "set_ibm")

    def set_storage(self, storage, scale_exp=20):
        """storage: 'f32' (deviations, default), 'f32full', 'f16', 'bf16' or the enum value."""
        code = STORAGE[storage] if isinstance(storage, str) else int(storage)
        _check(load().glsm_fluid_set_storage(self.h, code, int(scale_exp)), "set_storage")

    def enable_lbm_timing(self, every=1, max_samples=4096):
        _check(load().glsm_fluid_enable_lbm_timing(self.h, int(every), int(max_samples)), "enable_lbm_timing")

    def lbm_timing(self):
        """(mean stream/collide launch duration in ms, number of samples)."""
        ms, n = C.c_double(), C.c_int()
        _check(load().glsm_fluid_lbm_timing(self.h, C.byref(ms), C.byref(n)), "lbm_timing")
        return ms.value, n.value

    def read(self, name):
        """u [N, 3], c [N], f / fpost [19, N] (float32); rho [N] (float64,
        rho - 1 of the f populations summed in FP64 from the stored values)."""
        shape = {"u": (self.N, 3), "c": (self.N,), "f": (19, self.N), "fpost": (19, self.N), "rho": (self.N,)}[name]
        a = np.empty(shape, dtype=np.float64 if name == "rho" else np.float32)
        _check(load().glsm_fluid_read(self.h, FL_FIELD[name], a.ctypes.data, a.nbytes), f"fluid read {name}")
        return a

    def write(self, name, arr):
        shape = {"u": (self.N, 3), "c": (self.N,), "f": (19, self.N)}[name]
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float32).reshape(shape))
        _check(load().glsm_fluid_write(self.h, FL_FIELD[name], a.ctypes.data, a.nbytes), f"fluid write {name}")


def copy_bandwidth(nbytes: int, reps: int = 20) -> float:
    """GB/s ((read + written) bytes / s) of a plain 16-byte-per-lane device copy."""
    g = C.c_double()
    _check(load().glsm_copy_bandwidth(int(nbytes), int(reps), C.byref(g)), "copy_bandwidth")
    return g.value


def system_step(gels, coupler: Coupler, fluid: Fluid, it: int):
    """One time step in sim.cpp's order (sim.cpp:66-81)."""
    arr = (C.c_void_p * len(gels))(*[g.h for g in gels])
    _check(load().glsm_system_step(arr, len(gels), coupler.h, fluid.h, int(it)), "glsm_system_step")


class Batch:
    """Many gel-fluid instances advanced together (include/batch.h).

    Each member is (gels, coupler, fluid); the Python objects must stay
    alive while they are members.  step() advances every member by one time
    step with one launch per phase (HIP-graph captured); each member gives
    the same bits as system_step on its own objects."""

    def __init__(self, storage="f32", scale_exp=20):
        h = C.c_void_p()
        _check(load().glsm_batch_create(STORAGE[storage], int(scale_exp), C.byref(h)), "glsm_batch_create")
        self.h = h.value
        self.storage = storage
        self.members = {}

    def close(self):
        if getattr(self, "h", None):
            load().glsm_batch_destroy(self.h)
            self.h = None
            self.members = {}

    def add(self, gels, coupler, fluid, iteration=0, write_files=False, out_dir=""):
        mid = _check(load().glsm_batch_add(self.h, coupler.h, fluid.h, int(iteration), int(bool(write_files)),
                                           os.fsencode(out_dir)), "batch add")
        self.members[mid] = (gels, coupler, fluid)
        return mid

    def remove(self, mid):
        _check(load().glsm_batch_remove(self.h, int(mid)), "batch remove")
        self.members.pop(mid, None)

    def __len__(self):
        return _check(load().glsm_batch_size(self.h), "batch size")

    def iteration(self, mid):
        return _check(load().glsm_batch_iteration(self.h, int(mid)), "batch iteration")

    def step(self, n=1):
        _check(load().glsm_batch_step(self.h, int(n)), "batch step")

    def use_graph(self, on):
        _check(load().glsm_batch_use_graph(self.h, int(bool(on))), "batch use_graph")

    def set_output(self, output):
        """Route the members' .dat output through an OutputStream (None: the
        objects' own synchronous writeFiles)."""
        self.output = output
        _check(load().glsm_batch_set_output(self.h, output.h if output else None), "batch set_output")

    def set_launch(self, nt, run, grid, wgb=1, chunk=1, lds=0):
        """Launch shape of the batched stream/collide kernel (see
        glsm_batch_set_launch)."""
        _check(load().glsm_batch_set_launch(self.h, int(nt), int(run), int(grid), int(wgb), int(chunk), int(lds)),
               "batch set_launch")

    def launch(self):
        v = [C.c_int() for _ in range(6)]
        _check(load().glsm_batch_get_launch(self.h, *[C.byref(x) for x in v]), "batch launch")
        return dict(zip(("nt", "run", "grid", "wgb", "chunk", "lds"), (x.value for x in v)))

    def autotune(self, reps=5, cache=None):
        """Time every launch shape on the current state and keep the fastest;
        returns (shape, mean launch ms).  With `cache` (a JSON file path, or
        True for the default tune_cache_path()), a shape stored for the same
        workload key -- storage, device, the members' lattices and marker
        counts -- is reused without timing, and a new pick is stored."""
        path = tune_cache_path() if cache is True else cache
        key = self.workload_key() if path else None
        if path:
            hit = _tune_cache_get(path, key)
            if hit is not None:
                self.set_launch(**hit["launch"])
                return self.launch(), hit["ms"]
        ms = C.c_double()
        _check(load().glsm_batch_autotune(self.h, int(reps), C.byref(ms)), "batch autotune")
        if path:
            _tune_cache_put(path, key, {"launch": self.launch(), "ms": ms.value})
        return self.launch(), ms.value

    def workload_key(self):
        """The key a tuned launch shape is cached under."""
        import hashlib

        members = sorted((f.grid, c.M) for _, c, f in self.members.values())
        digest = hashlib.sha1(repr(members).encode()).hexdigest()[:16]
        return f"{self.storage}|{device_name()}|{len(members)} members|{sum(f.N for _, _, f in self.members.values())} cells|{digest}"

    def enable_lbm_timing(self, every):
        _check(load().glsm_batch_enable_lbm_timing(self.h, int(every)), "batch enable_lbm_timing")

    def lbm_timing(self):
        ms, n = C.c_double(), C.c_int()
        _check(load().glsm_batch_lbm_timing(self.h, C.byref(ms), C.byref(n)), "batch lbm_timing")
        return ms.value, n.value


class Slab:
    """Planes [z0, z0 + nz) of `like`'s lattice with one halo plane on each
    side (include/slab.h): one GPU's share of a z-decomposed lattice."""

    def __init__(self, like: "Fluid", z0: int, nz: int):
        h = C.c_void_p()
        _check(load().glsm_slab_create(like.h, int(z0), int(nz), C.byref(h)), "glsm_slab_create")
        self.h = h
        self.z0, self.nz = int(z0), int(nz)

    def close(self):
        if self.h:
            _check(load().glsm_slab_destroy(self.h), "glsm_slab_destroy")
            self.h = None

    def load(self, src: "Fluid"):
        _check(load().glsm_slab_load(self.h, src.h), "glsm_slab_load")

    def store(self, dst: "Fluid"):
        _check(load().glsm_slab_store(self.h, dst.h), "glsm_slab_store")

    def face_bytes(self) -> int:
        n = C.c_size_t()
        _check(load().glsm_slab_face_bytes(self.h, C.byref(n)), "glsm_slab_face_bytes")
        return n.value

    def enable_timing(self, on=True):
        _check(load().glsm_slab_enable_timing(self.h, 1 if on else 0), "glsm_slab_enable_timing")

    def timing(self):
        """(stream/collide ms, pack -> transfer -> unpack ms), mean per sub-step."""
        a, b = C.c_double(), C.c_double()
        _check(load().glsm_slab_timing(self.h, C.byref(a), C.byref(b)), "glsm_slab_timing")
        return a.value, b.value


def slab_chain_step(chain, steps=1, write_u=False):
    """`steps` sub-steps of a chain of slabs covering the lattice in z order."""
    arr = (C.c_void_p * len(chain))(*[s.h for s in chain])
    _check(load().glsm_slab_chain_step(arr, len(chain), int(steps), 1 if write_u else 0), "glsm_slab_chain_step")


def split_z(nz_total: int, parts: int):
    """(z0, nz) of `parts` contiguous slabs, sizes differing by at most one."""
    base, extra = divmod(nz_total, parts)
    out, z = [], 0
    for k in range(parts):
        n = base + (1 if k < extra else 0)
        out.append((z, n))
        z += n
    return out


class OutputStream:
    """Asynchronous .dat output (include/output.h): device snapshot in stream
    order, copy to pinned memory on a copy stream, formatting on writer
    threads, in the reference's file layout."""

    def __init__(self, writers=4):
        h = C.c_void_p()
        _check(load().glsm_output_create(int(writers), C.byref(h)), "glsm_output_create")
        self.h = h.value

    def close(self):
        if getattr(self, "h", None):
            load().glsm_output_destroy(self.h)
            self.h = None

    def gel(self, gel, time, out_dir=""):
        _check(load().glsm_output_gel(self.h, gel.h, int(time), os.fsencode(out_dir)), "output gel")

    def fluid(self, fluid, time, out_dir=""):
        _check(load().glsm_output_fluid(self.h, fluid.h, int(time), os.fsencode(out_dir)), "output fluid")

    def release(self, obj):
        _check(load().glsm_output_release(self.h, obj.h), "output release")

    def flush(self):
        _check(load().glsm_output_flush(self.h), "output flush")

    def status(self):
        p, c = C.c_long(), C.c_long()
        err = C.create_string_buffer(512)
        _check(load().glsm_output_status(self.h, C.byref(p), C.byref(c), err, 512), "output status")
        return {"pending": p.value, "completed": c.value, "error": err.value.decode(errors="replace")}


def device_memory():
    """(free, total) bytes of the current device."""
    f, t = C.c_size_t(), C.c_size_t()
    _check(load().glsm_device_memory(C.byref(f), C.byref(t)), "glsm_device_memory")
    return f.value, t.value


def device_name():
    """The library build (target architecture) and the current device's
    memory size: the device part of the tuning cache key."""
    return f"{load().glsm_version().decode(errors='replace')} {device_memory()[1] >> 30} GiB"


def tune_cache_path():
    """$GLSM_TUNE_CACHE, else $XDG_CACHE_HOME (or ~/.cache)/glsm_fluid_3d_amd/launch_shapes.json."""
    p = os.environ.get("GLSM_TUNE_CACHE")
    if p:
        return p
    root = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(root, "glsm_fluid_3d_amd", "launch_shapes.json")


def _tune_cache_get(path, key):
    import json

    try:
        with open(path) as fh:
            return json.load(fh).get(key)
    except (OSError, ValueError):
        return None


def _tune_cache_put(path, key, value):
    import json

    try:
        with open(path) as fh:
            data = json.load(fh)
    except (OSError, ValueError):
        data = {}
    data[key] = value
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(data, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)
