"""ctypes binding of libhoomd_amd.so (the C ABI declared in include/hoomd_amd.h).

The library is built in-tree by ``hoomd-blue_amd/build.py``. Loading fails loudly when it is
missing: there is no CPU fallback for the GPU path.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhoomd_amd.so")
MAX_TYPES = 4

_LIB = None

_u = C.c_uint
_i = C.c_int
_d = C.c_double
_p = C.c_void_p
_u64 = C.c_uint64
_u16 = C.c_uint16


class HmBox(C.Structure):
    _fields_ = [("L", _d * 3)]


class HmCellArgs(C.Structure):
    _fields_ = [("N", _u), ("pos", _p), ("box", HmBox), ("dim", _u * 3), ("cmax", _u),
                ("cell_size", _p), ("entries", _p), ("last_pos", _p), ("conditions", _p),
                ("cell_idx", _p)]


class HmNlistArgs(C.Structure):
    _fields_ = [("N", _u), ("pos", _p), ("box", HmBox), ("dim", _u * 3), ("cmax", _u),
                ("cell_size", _p), ("entries", _p), ("cell_adj", _p), ("n_adj", _u),
                ("ntypes", _u), ("r_cut", _d * (MAX_TYPES * MAX_TYPES)), ("r_buff", _d),
                ("n_excl", _p), ("excl", _p), ("max_excl", _u), ("nmax", _u), ("nlist", _p),
                ("n_neigh", _p), ("conditions", _p), ("local_rows", _p)]


class HmLJParams(C.Structure):
    _fields_ = [("sigma", _d), ("epsilon", _d)]


class HmTableParams(C.Structure):
    _fields_ = [("r_min", _d), ("width", _u), ("U", _p), ("F", _p)]


class HmPairArgs(C.Structure):
    _fields_ = [("N", _u), ("pos", _p), ("box", HmBox), ("nlist", _p), ("nmax", _u),
                ("n_neigh", _p), ("ntypes", _u), ("r_cut", _d * (MAX_TYPES * MAX_TYPES)),
                ("r_on", _d * (MAX_TYPES * MAX_TYPES)), ("shift_mode", _i),
                ("compute_virial", _i), ("force", _p), ("virial", _p), ("tpp", _u),
                ("block_size", _u)]


class HmPairCellArgs(C.Structure):
    _fields_ = [("N", _u), ("pos", _p), ("box", HmBox), ("dim", _u * 3), ("cmax", _u),
                ("cell_size", _p), ("cell_idx", _p), ("cell_adj", _p), ("n_adj", _u),
                ("local_rows", _p), ("nmax", _u), ("n_neigh", _p), ("ntypes", _u),
                ("r_cut", _d * (MAX_TYPES * MAX_TYPES)), ("r_on", _d * (MAX_TYPES * MAX_TYPES)),
                ("shift_mode", _i), ("compute_virial", _i), ("force", _p), ("virial", _p),
                ("tpp", _u), ("workgroups", _u)]


class HmFeneParams(C.Structure):
    _fields_ = [("k", _d), ("r0", _d), ("epsilon", _d), ("sigma", _d), ("delta", _d)]


class HmHarmonicParams(C.Structure):
    _fields_ = [("k", _d), ("r0", _d)]


MLP_NK = 16
MLP_NPARAM = 5313
MLP_NBF16 = 10240
MLP_GRAD_BLOCKS = 256


class HmMlpModel(C.Structure):
    _fields_ = [("r_cut", _d), ("width", _d), ("mu0", _d), ("dmu", _d),
                ("gmean", _d * MLP_NK), ("gscale", _d * MLP_NK)]


class HmMlpArgs(C.Structure):
    _fields_ = [("N", _u), ("pos", _p), ("box", HmBox), ("nlist", _p), ("nmax", _u),
                ("n_neigh", _p), ("params", _p), ("params_bf16", _p), ("G", _p), ("A", _p),
                ("Esite", _p), ("force", _p), ("virial", _p), ("target", _p), ("err", _p),
                ("C", _p), ("loss_partial", _p), ("grad_partial", _p), ("A_bf16", _p),
                ("pos32", _p)]


class HmMlpTrainArgs(C.Structure):
    _fields_ = [("nframes", _u), ("frame_pos", _p), ("frame_force", _p), ("nsteps", _u),
                ("order", _p), ("lr", _d), ("beta1", _d), ("beta2", _d), ("eps", _d),
                ("loss", _p), ("frames_per_step", _u)]


class HmBondArgs(C.Structure):
    _fields_ = [("N", _u), ("pos", _p), ("box", HmBox), ("tag", _p), ("rtag", _p),
                ("bond_row", _p), ("bond_other", _p), ("bond_type", _p), ("nbond_types", _u),
                ("compute_virial", _i), ("force", _p), ("virial", _p), ("flag", _p)]


# every exported symbol of include/hoomd_amd.h with (restype, argtypes)
SIGNATURES = {
    "hm_version": (C.c_char_p, []),
    "hm_device_count": (_i, []),
    "hm_last_error": (C.c_char_p, []),
    "hm_malloc": (_i, [C.POINTER(_p), C.c_size_t]),
    "hm_free": (_i, [_p]),
    "hm_memcpy_htod": (_i, [_p, _p, C.c_size_t]),
    "hm_memcpy_dtoh": (_i, [_p, _p, C.c_size_t]),
    "hm_memset": (_i, [_p, _i, C.c_size_t]),
    "hm_memcpy_dtod": (_i, [_p, _p, C.c_size_t]),
    "hm_synchronize": (_i, []),
    "hm_cell_dims": (_i, [C.POINTER(HmBox), _d, _p]),
    "hm_cell_n_adj": (_u, [_p]),
    "hm_cell_adj": (_i, [_p, _p, _p]),
    "hm_cell_list_build": (_i, [C.POINTER(HmCellArgs), _p]),
    "hm_nlist_build": (_i, [C.POINTER(HmNlistArgs), _p]),
    "hm_nlist_check": (_i, [_u, _p, _p, HmBox, _d, _d, _p, _p]),
    "hm_pair_lj": (_i, [C.POINTER(HmPairArgs), _p, _p]),
    "hm_pair_table": (_i, [C.POINTER(HmPairArgs), _p, _p]),
    "hm_pair_lj_cells": (_i, [C.POINTER(HmPairCellArgs), _p, _p]),
    "hm_bond_fene": (_i, [C.POINTER(HmBondArgs), _p, _p]),
    "hm_bond_harmonic": (_i, [C.POINTER(HmBondArgs), _p, _p]),
    "hm_mlp_cg_loss_blocks": (_u, [_u]),
    "hm_mlp_cg_pack": (_i, [_p, _p, _p]),
    "hm_mlp_cg_descriptor": (_i, [C.POINTER(HmMlpArgs), C.POINTER(HmMlpModel), _p]),
    "hm_mlp_cg_forward": (_i, [C.POINTER(HmMlpArgs), C.POINTER(HmMlpModel), _p]),
    "hm_mlp_cg_forces": (_i, [C.POINTER(HmMlpArgs), C.POINTER(HmMlpModel), _i, _p]),
    "hm_mlp_cg_backward": (_i, [C.POINTER(HmMlpArgs), C.POINTER(HmMlpModel), _p]),
    "hm_mlp_cg_loss": (_i, [_p, _u, _p, _p]),
    "hm_mlp_cg_grad": (_i, [_p, _p, _p]),
    "hm_mlp_cg_adam": (_i, [_p, _p, _p, _p, _p, _d, _d, _d, _d, _u, _p]),
    "hm_nvt_step_one": (_i, [_u, _p, _p, _p, _p, HmBox, _d, _d, _p]),
    "hm_nvt_step_two": (_i, [_u, _p, _p, _p, _u, _d, _d, _p]),
    "hm_langevin_step_two": (_i, [_u, _p, _p, _p, _p, _p, _u, _p, _u, _d, _d, _u64, _u16, _p]),
    "hm_thermo": (_i, [_u, _p, _p, _p, _u, _p, _p]),
    "hm_thermo_noalloc": (_i, [_u, _p, _p, _p, _u, _p, _p]),
    "hm_bussi_alpha_host": (_d, [_d, _d, _d, _d, _d, _u64, _u16, C.c_uint32]),
    "hm_thermalize": (_i, [_u, _p, _p, _d, _u64, _u16, _p]),
    "hm_sfc_order": (_i, [_u, _p, HmBox, _u, _p, _p]),
    "hm_sfc_traversal": (_i, [_u, _p, _p]),
    "hm_sys_create": (_p, [_u, HmBox, _u, _p, _p, _p, _p, _p, _u16]),
    "hm_sys_destroy": (None, [_p]),
    "hm_sys_set_nlist": (_i, [_p, _d, _i]),
    "hm_sys_set_nlist_policy": (_i, [_p, _u, _i]),
    "hm_sys_add_pair_lj": (_i, [_p, _p, _p, _p, _i]),
    "hm_sys_add_pair_table": (_i, [_p, _p, _p, _p, _p, _p]),
    "hm_sys_add_bond_fene": (_i, [_p, _u, _p, _p, _u, _p]),
    "hm_sys_set_method": (_i, [_p, _i, _d, _d, _p]),
    "hm_sys_set_sorter": (_i, [_p, _u, _u]),
    "hm_sys_set_compute_virial": (_i, [_p, _i]),
    "hm_sys_set_pair_launch": (_i, [_p, _u, _u]),
    "hm_sys_set_pair_kernel": (_i, [_p, _i, _u, _u]),
    "hm_sys_set_autotune": (_i, [_p, _i]),
    "hm_sys_pair_config": (_i, [_p, _p]),
    "hm_sys_tune_log": (_i, [_p, _p, _u]),
    "hm_sys_set_chunk": (_i, [_p, _u, _i]),
    "hm_sys_run": (_i, [_p, _u64]),
    "hm_sys_prime": (_i, [_p, _u64]),
    "hm_sys_timestep": (_u64, [_p]),
    "hm_sys_set_timestep": (_i, [_p, _u64]),
    "hm_sys_get_state": (_i, [_p, _p, _p, _p, _p]),
    "hm_sys_set_state": (_i, [_p, _p, _p, _p]),
    "hm_sys_get_forces": (_i, [_p, _i, _p, _p]),
    "hm_sys_get_pairs": (C.c_longlong, [_p, _p, C.c_longlong]),
    "hm_sys_thermo": (_i, [_p, _p]),
    "hm_sys_stats": (_i, [_p, _p]),
    "hm_sys_thermalize": (_i, [_p, _d]),
    "hm_sys_stream": (_p, [_p]),
    "hm_sys_time_pair": (_i, [_p, _i, _u, C.POINTER(_d), C.POINTER(_u64)]),
    "hm_sys_add_bond_harmonic": (_i, [_p, _u, _p, _p, _u, _p]),
    "hm_sys_update_pair_table": (_i, [_p, _i, _p, _p]),
    "hm_sys_add_mlp_cg": (_i, [_p, C.POINTER(HmMlpModel), _p]),
    "hm_sys_mlp_set_params": (_i, [_p, _i, _p]),
    "hm_sys_mlp_get_params": (_i, [_p, _i, _p]),
    "hm_sys_mlp_set_model": (_i, [_p, _i, C.POINTER(HmMlpModel)]),
    "hm_sys_mlp_set_pair_table": (_i, [_p, _i, _u, _d, _p, _p]),
    "hm_sys_mlp_descriptors": (_i, [_p, _i, _p, _p]),
    "hm_sys_mlp_train": (_i, [_p, _i, C.POINTER(HmMlpTrainArgs)]),
    "hm_sys_mlp_loss_grad": (_i, [_p, _i, _p, _p, C.POINTER(_d), _p]),
    "hm_sys_set_site_map": (_i, [_p, _u, _p, _p]),
    "hm_sys_map_frame": (_i, [_p, _p, _p]),
    "hm_sys_time_mlp": (_i, [_p, _i, _u, _p]),
    "hm_sys_time_build": (_i, [_p, _u, _p]),
    "hm_sys_set_list_capacity": (_i, [_p, _u, _u]),
    "hm_sys_nlist_launch": (_i, [_p, _p, _p, _u]),
    "hm_tune_cache_clear": (None, []),
    "hm_sys_frame_subtract_force": (_i, [_p, _i, _p, _p]),
    "hm_frames_sub_harmonic": (_i, [_u, _u, _p, _p, _p, _p, _p, _d, _d, HmBox, _p]),
    # reference-layout pair driver (include/hoomd_amd_pair.h)
    "hm_compute_pair_forces_lj": (_i, [_p, _p, _p]),
    # batched replicas
    "hm_batch_create": (_p, [_u, _p, _p, _p, _p, _p, _p, _p, _d, _d, _d, _d, _d, _d]),
    "hm_batch_destroy": (None, [_p]),
    "hm_batch_thermalize": (_i, [_p, _u64]),
    "hm_batch_set_timestep": (_i, [_p, _u64]),
    "hm_batch_run": (_i, [_p, _u64]),
    "hm_batch_prime": (_i, [_p, _u64]),
    "hm_batch_set_chunk": (_i, [_p, _u]),
    "hm_batch_set_sort_period": (_i, [_p, _u]),
    "hm_batch_tune": (_i, [_p, _p, _u]),
    "hm_batch_set_pair_launch": (_i, [_p, _u, _u]),
    "hm_batch_config": (_i, [_p, _p]),
    "hm_batch_get_state": (_i, [_p, _p, _p, _p]),
    "hm_batch_get_forces": (_i, [_p, _p]),
    "hm_batch_replica_stats": (_i, [_p, _p]),
    "hm_batch_stats": (_i, [_p, _p]),
    "hm_batch_time_pair": (_i, [_p, _u, C.POINTER(_d)]),
    "hm_batch_set_list_sizes": (_i, [_p, _u, _u]),
}


class HoomdAmdError(RuntimeError):
    """Error returned by the HIP library (message from hm_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load libhoomd_amd.so; raise if it was not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python hoomd-blue_amd/build.py` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(code, what=""):
    if code is None or code >= 0:
        return code
    msg = lib().hm_last_error().decode()
    raise HoomdAmdError(code, f"{what}: {msg}" if what else msg)


def box(L):
    Ls = np.broadcast_to(np.asarray(L, np.float64), (3,))
    b = HmBox()
    for k in range(3):
        b.L[k] = float(Ls[k])
    return b


class DeviceArray:
    """A device buffer with numpy upload/download through the C ABI."""

    def __init__(self, shape, dtype, fill=None):
        self.shape = tuple(np.atleast_1d(shape)) if not isinstance(shape, tuple) else shape
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        ptr = C.c_void_p()
        check(lib().hm_malloc(C.byref(ptr), max(self.nbytes, 16)), "hm_malloc")
        self.ptr = ptr.value
        if fill is not None:
            check(lib().hm_memset(self.ptr, int(fill), max(self.nbytes, 16)), "hm_memset")

    @classmethod
    def from_numpy(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            check(lib().hm_memcpy_htod(d.ptr, a.ctypes.data, a.nbytes), "hm_memcpy_htod")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().hm_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes), "hm_memcpy_dtoh")
        return out

    def zero(self):
        check(lib().hm_memset(self.ptr, 0, max(self.nbytes, 16)), "hm_memset")

    def __del__(self):
        if getattr(self, "ptr", None) and _LIB is not None:
            _LIB.hm_free(self.ptr)
            self.ptr = None


def pointer_array(ptrs):
    """A host array of device pointers (const float* const*)."""
    arr = (C.c_void_p * max(1, len(ptrs)))(*[p for p in ptrs])
    return arr
