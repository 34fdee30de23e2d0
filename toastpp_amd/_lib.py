"""ctypes binding of libtoastpp_amd.so (the C-ABI declared in include/toastpp_amd.h).

The library is required: there is no CPU fallback.  If the shared object is
missing, loading fails with an error that says how to build it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtoastpp_amd.so")

TMI_OK = 0
TMI_ERR_INVALID_ARG = 1
TMI_ERR_NOT_CONVERGED = 2
TMI_ERR_OOM = 3
TMI_ERR_DEVICE = 4
TMI_ERR_IO = 5
TMI_ERR_STATE = 6
TMI_ERR_UNSUPPORTED = 7
TMI_ERR_TIMEOUT = 8

LSOLVER_DIRECT, LSOLVER_CG, LSOLVER_BICGSTAB, LSOLVER_GMRES, LSOLVER_COCG = 0, 1, 2, 3, 4
PRECON_NONE, PRECON_DIAG = 0, 1
MODE_FP64_SINGLE, MODE_MIXED_BLOCK, MODE_FP64_BLOCK = 0, 1, 2
DATA_LIN, DATA_LOG = 0, 1
ORDER_NATURAL, ORDER_MORTON, ORDER_CLUSTER = 0, 1, 2
ELID_TET4, ELID_TRI3 = 3, 15
BASIS_NODAL, BASIS_ELEMENT, BASIS_RASTER = 0, 1, 2

_ERRNAMES = {1: "invalid argument", 2: "not converged", 3: "out of memory", 4: "device error", 5: "I/O error",
             6: "invalid state", 7: "unsupported", 8: "timeout"}


class TmiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%s)" % (msg, _ERRNAMES.get(code, str(code))))
        self.code = code


class NotConverged(TmiError):
    pass


class ColStat(C.Structure):
    _fields_ = [("iters", C.c_int), ("sweeps", C.c_int), ("relres", C.c_double), ("converged", C.c_int)]


D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
P = C.c_void_p
PP = C.POINTER(C.c_void_p)
IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)

# name -> argtypes (all entry points return int status unless noted)
_SIGS = {
    "tmi_device_info": [C.c_char_p, C.c_int, DP],
    "tmi_synchronize": [],
    "tmi_device_alloc": [C.c_ulonglong, PP],
    "tmi_device_free": [P],
    "tmi_memcpy_h2d": [P, P, C.c_ulonglong],
    "tmi_memcpy_d2h": [P, P, C.c_ulonglong],
    "tmi_memcpy_d2d": [P, P, C.c_ulonglong],
    "tmi_read_mesh_file": [C.c_char_p, IP, IP, IP, IP, P, P, P],
    "tmi_mesh_make": [C.c_int, C.c_int, D, C.c_int, I, PP],
    "tmi_mesh_read": [C.c_char_p, PP],
    "tmi_mesh_free": [P],
    "tmi_mesh_dims": [P, IP, IP, IP, IP],
    "tmi_mesh_get_nodes": [P, D],
    "tmi_mesh_get_elements": [P, I],
    "tmi_mesh_get_bnd": [P, I],
    "tmi_mesh_element_size": [P, D],
    "tmi_mesh_sparse_structure": [P, P, P, IP],
    "tmi_mesh_read_qm": [P, C.c_char_p],
    "tmi_mesh_set_qm": [P, C.c_int, P, C.c_int, P, I, P, C.c_int],
    "tmi_mesh_qm_dims": [P, IP, IP, IP],
    "tmi_mesh_get_qm": [P, P, P, P, P],
    "tmi_qvec_gaussian": [P, C.c_int, D, C.c_double, C.c_int, P, D],
    "tmi_qvec_point": [P, C.c_int, D, C.c_int, D],
    "tmi_optical_to_coeffs": [C.c_int, D, D, D, D, D, D],
    "tmi_solver_create": [P, C.c_int, C.c_int, C.c_double, PP],
    "tmi_solver_free": [P],
    "tmi_solver_set_linsolver": [P, C.c_char_p, C.c_double],
    "tmi_solver_set_precon": [P, C.c_int],
    "tmi_solver_set_maxit": [P, C.c_int],
    "tmi_solver_set_mode": [P, C.c_int],
    "tmi_solver_set_tile": [P, C.c_int],
    "tmi_solver_set_order": [P, C.c_int],
    "tmi_node_order": [C.c_int, C.c_int, C.c_int, D, I, I, I],
    "tmi_solver_set_rows_per_wave": [P, C.c_int],
    "tmi_solver_set_spmm_variant": [P, C.c_int],
    "tmi_solver_work_bytes": [P, C.c_int, DP],
    "tmi_solver_set_compaction": [P, C.c_int],
    "tmi_solver_set_fallback": [P, C.c_int, C.c_int],
    "tmi_solver_autotune": [P, C.c_int, P, P, C.c_int, IP, IP, IP, DP],
    "tmi_solver_allocate": [P],
    "tmi_solver_assemble": [P, D, D, D, C.c_double],
    "tmi_solver_reset": [P, D, D, D, C.c_double],
    "tmi_solver_assemble_basis": [P, D, D, D, C.c_double, C.c_int],
    "tmi_solver_reset_basis": [P, D, D, D, C.c_double, C.c_int],
    "tmi_solver_nnz": [P, IP],
    "tmi_solver_get_sysmat": [P, P, P, P],
    "tmi_solver_calc_fields": [P, C.c_int, I, I, D, D, P],
    "tmi_solver_calc_fields_device": [P, C.c_int, P, P, P],
    "tmi_solver_calc_fields_freqs_device": [P, C.c_int, P, P, P, P],
    "tmi_solver_project_all": [P, C.c_int, I, I, D, D, C.c_int, D],
    "tmi_solver_project_all_device": [P, C.c_int, I, I, D, P, C.c_int, C.c_int, D],
    "tmi_solver_project_all_device_noalloc": [P, C.c_int, I, I, D, P, C.c_int, C.c_int, D],
    "tmi_solver_last_timing": [P, DP, DP],
    "tmi_solver_last_counts": [P, P, P, P],
    "tmi_solver_set_graphs": [P, C.c_int],
    "tmi_solver_time_kernel": [P, C.c_int, C.c_int, DP, DP],
    "tmi_jacobian_nodal": [P, C.c_int, C.c_int, D, D, D, C.c_int, D],
    "tmi_jacobian_nodal_device": [P, C.c_int, C.c_int, P, C.c_int, P, C.c_int, P, C.c_int, P],
    "tmi_jacobian_cw": [P, C.c_int, C.c_int, D, D, P, C.c_int, P, P],
    "tmi_jacobian_element": [P, C.c_int, C.c_int, D, D, D, C.c_int, D],
    "tmi_jacobian_element_device": [P, C.c_int, C.c_int, P, C.c_int, P, C.c_int, P, C.c_int, P],
    "tmi_jacobian_cw_element": [P, C.c_int, C.c_int, D, D, P, C.c_int, P, P],
    "tmi_jacobian_mvec": [P, P, C.c_int, C.c_int, C.c_int, I, I, D, D, D, C.c_int, D],
    "tmi_jacobian_cw_mvec": [P, P, C.c_int, C.c_int, C.c_int, I, I, D, D, D, C.c_int, P, P],
    "tmi_raster_data_gradient": [P, C.c_int, D, D, D],
    "tmi_raster_get_nodemap": [P, P, P],
    "tmi_raster_map_grid_to_mesh": [P, C.c_int, D, D],
    "tmi_raster_data_gradient_device": [P, C.c_int, P, C.c_int, P, C.c_int, P],
    "tmi_jacobian_raster_cw": [P, C.c_int, C.c_int, D, D, P, C.c_int, P, P],
    "tmi_raster_create": [P, I, P, PP],
    "tmi_raster_free": [P],
    "tmi_raster_dims": [P, IP, IP],
    "tmi_raster_bbox": [P, D, D],
    "tmi_raster_get_elref": [P, P, P],
    "tmi_raster_get_bmap": [P, P, P],
    "tmi_raster_map_mesh_to_grid": [P, C.c_int, D, D],
    "tmi_jacobian_raster": [P, C.c_int, C.c_int, D, D, D, C.c_int, D],
    "tmi_jacobian_raster_device": [P, C.c_int, C.c_int, P, C.c_int, P, C.c_int, P, C.c_int, P],
    "tmi_server_create": [C.c_int, C.c_double, PP],
    "tmi_server_free": [P],
    "tmi_server_submit": [P, P, C.c_int, P, P, C.c_int, C.POINTER(C.c_longlong)],
    "tmi_server_wait": [P, C.c_longlong, C.c_double, P, DP, IP],
    "tmi_server_counts": [P, IP, IP, DP],
    "tmi_raster_create_basis": [P, I, I, P, PP],
    "tmi_raster_basis_len": [P, IP, IP],
    "tmi_raster_get_gmap": [P, P, P, P, IP],
    "tmi_raster_get_gimap": [P, P, P, P, IP],
    "tmi_raster_map_grid_to_basis": [P, C.c_int, D, D],
    "tmi_raster_map_basis_to_grid": [P, C.c_int, D, D],
    "tmi_read_nim": [C.c_char_p, C.c_int, IP, IP, P],
    "tmi_write_nim": [C.c_char_p, C.c_char_p, C.c_int, C.c_int, D, C.c_int],
    "tmi_read_vector": [C.c_char_p, IP, P],
    "tmi_write_vector": [C.c_char_p, C.c_int, D],
}

TMI_NIM_FWDFEM = 0
TMI_NIM_TOASTMOD = 1

_lib = None


def lib():
    """Load libtoastpp_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libtoastpp_amd.so not found at %s: build it with `python -m toastpp_amd.build` "
                              "(there is no CPU fallback)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        L.tmi_last_error.restype = C.c_char_p
        L.tmi_last_error.argtypes = []
        L.tmi_server_clock.restype = C.c_double
        L.tmi_server_clock.argtypes = [C.c_void_p]
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return ["tmi_last_error", "tmi_server_clock"] + list(_SIGS)


def check(status, allow_notconv=False):
    if status == TMI_OK:
        return status
    msg = lib().tmi_last_error().decode(errors="replace")
    if status == TMI_ERR_NOT_CONVERGED:
        if allow_notconv:
            return status
        raise NotConverged(status, msg)
    raise TmiError(status, msg)


def call(name, *args, allow_notconv=False):
    return check(getattr(lib(), name)(*args), allow_notconv=allow_notconv)


def ptr(a):
    """Raw data pointer of a numpy array (or None)."""
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class DeviceArray:
    """A GPU buffer of the library's device holding a numpy-shaped array."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self._p = C.c_void_p()
        call("tmi_device_alloc", self.nbytes, C.byref(self._p))

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        call("tmi_memcpy_h2d", d._p, a.ctypes.data_as(C.c_void_p), d.nbytes)
        return d

    @property
    def ptr(self):
        return self._p.value

    def to_host(self):
        a = np.empty(self.shape, self.dtype)
        call("tmi_memcpy_d2h", a.ctypes.data_as(C.c_void_p), self._p, self.nbytes)
        return a

    def free(self):
        if self._p.value and _lib is not None:
            _lib.tmi_device_free(self._p)
        self._p = C.c_void_p()

    def __del__(self):
        self.free()


def synchronize():
    call("tmi_synchronize")


def device_info():
    buf = C.create_string_buffer(256)
    gib = C.c_double()
    call("tmi_device_info", buf, 256, C.byref(gib))
    return buf.value.decode(), gib.value
