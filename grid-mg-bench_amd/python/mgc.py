"""ctypes binding of libmgc.so (the C ABI declared in include/mgc.h).

This is the Python face of the C ABI that the tests and the benchmark drive:
every call goes straight to an ``extern "C"`` entry point of the HIP library.
Device memory and streams come from the library's own helpers (no torch), so
one HIP runtime serves the whole process.  A missing library raises
``MgcError``: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

import numpy as np

F64, F32, F16, BF16 = 0, 1, 2, 3
SOLVER_MR, SOLVER_GMRES, SOLVER_FGMRES, SOLVER_BICGSTAB = 1, 2, 3, 4
BASIS_NULL_VECTORS, BASIS_TEST_VECTORS = 1, 2
LAYOUT_FLAT, LAYOUT_SPLIT = 0, 1
EVEN, ODD, FULL = 0, 1, -1

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libmgc.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "mgc.h")

_c = ctypes
_vp, _i, _sz, _u64 = _c.c_void_p, _c.c_int, _c.c_size_t, _c.c_uint64
_ip = _c.POINTER(_c.c_int)


class MGParams(ctypes.Structure):
    """mgc_mg_params of include/mgc.h."""
    _fields_ = [("nlevels", _c.c_int), ("block", (_c.c_int * 4) * 3), ("nb", _c.c_int * 3),
                ("precision", _c.c_int), ("coarsen_mode", _c.c_int),
                ("setup_iters", _c.c_int), ("setup_solve_iters", _c.c_int),
                ("setup_tol", _c.c_double), ("smoother_tol", _c.c_double),
                ("smoother_iters", _c.c_int), ("smoother_restart", _c.c_int),
                ("wrapper_tol", _c.c_double), ("wrapper_iters", _c.c_int),
                ("wrapper_restart", _c.c_int), ("coarsest_tol", _c.c_double),
                ("coarsest_iters", _c.c_int), ("coarsest_restart", _c.c_int),
                ("seed", _c.c_ulonglong), ("setup_solver", _c.c_int),
                ("setup_restart", _c.c_int), ("setup_basis_type", _c.c_int),
                ("setup_pre_orthonormalise", _c.c_int), ("setup_post_orthonormalise", _c.c_int),
                ("setup_refine_iters", _c.c_int), ("setup_recursive_updates", _c.c_int),
                ("smoother_solver", _c.c_int), ("coarsest_solver", _c.c_int),
                ("graph", _c.c_int), ("setup_nrhs", _c.c_int)]


LINOP_APPLY = ctypes.CFUNCTYPE(_c.c_int, _c.c_void_p, _c.c_int, _c.c_int, _c.c_void_p,
                               _c.c_void_p, _c.c_int, _c.c_void_p)


class FineLinop(ctypes.Structure):
    """mgc_fine_linop of include/mgc.h: a fine operator given as a callback."""
    _fields_ = [("apply", LINOP_APPLY), ("ctx", _c.c_void_p), ("dims", _c.c_int * 4),
                ("precisions", _c.c_int), ("max_nrhs", _c.c_int),
                ("gamma5_hermitian", _c.c_int)]


class TestParams(ctypes.Structure):
    """mgc_test_params of include/mgc.h (MGTestOtherParams)."""
    _fields_ = [("outer_tol", _c.c_double), ("outer_maxit", _c.c_int),
                ("outer_restart", _c.c_int), ("mass_setup", _c.c_double),
                ("mass_solve", _c.c_double), ("csw", _c.c_double),
                ("config", _c.c_char * 1024), ("source_type", _c.c_char * 32),
                ("antiperiodic_t", _c.c_int)]


class SolveStats(ctypes.Structure):
    """mgc_solve_stats of include/mgc.h."""
    _fields_ = [("iterations", _c.c_int), ("converged", _c.c_int),
                ("precond_applies", _c.c_int), ("true_residual", _c.c_double),
                ("seconds", _c.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}

# argument types of every entry point (return type is int unless noted)
_SIGS = {
    "mgc_last_error": ([], _c.c_char_p),
    "mgc_version": ([], _i),
    "mgc_device_count": ([_ip], _i),
    "mgc_device_name": ([_c.c_char_p, _i], _i),
    "mgc_malloc": ([_c.POINTER(_vp), _sz], _i),
    "mgc_free": ([_vp], _i),
    "mgc_memcpy_h2d": ([_vp, _vp, _sz, _vp], _i),
    "mgc_memcpy_d2h": ([_vp, _vp, _sz, _vp], _i),
    "mgc_memcpy_d2d": ([_vp, _vp, _sz, _vp], _i),
    "mgc_fill_zero": ([_vp, _sz, _vp], _i),
    "mgc_device_copy": ([_vp, _vp, _sz, _vp], _i),
    "mgc_stream_create": ([_c.POINTER(_vp)], _i),
    "mgc_stream_destroy": ([_vp], _i),
    "mgc_stream_synchronize": ([_vp], _i),
    "mgc_device_synchronize": ([], _i),
    "mgc_event_create": ([_c.POINTER(_vp)], _i),
    "mgc_event_destroy": ([_vp], _i),
    "mgc_event_record": ([_vp, _vp], _i),
    "mgc_event_synchronize": ([_vp], _i),
    "mgc_event_elapsed_ms": ([_c.POINTER(_c.c_float), _vp, _vp], _i),
    "mgc_random_gaussian": ([_vp, _sz, _i, _u64, _u64, _vp], _i),
    "mgc_coarse_op_create": ([_c.POINTER(_vp), _ip, _i, _i, _i], _i),
    "mgc_coarse_op_destroy": ([_vp], _i),
    "mgc_coarse_op_info": ([_vp, _ip, _ip, _ip, _ip], _i),
    "mgc_coarse_op_set_links": ([_vp, _i, _vp, _i, _vp], _i),
    "mgc_coarse_op_get_links": ([_vp, _i, _vp, _vp], _i),
    "mgc_coarse_op_get_self_inverse": ([_vp, _vp, _vp], _i),
    "mgc_coarse_op_M": ([_vp, _vp, _vp, _vp], _i),
    "mgc_coarse_op_Mdag": ([_vp, _vp, _vp, _vp], _i),
    "mgc_coarse_op_Mdiag": ([_vp, _vp, _vp, _vp], _i),
    "mgc_coarse_op_Mdir": ([_vp, _vp, _vp, _i, _i, _vp], _i),
    "mgc_coarse_op_MdirAll": ([_vp, _vp, _vp, _vp], _i),
    "mgc_coarse_op_Dhop": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarse_op_DhopOE": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarse_op_DhopEO": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarse_op_Meooe": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarse_op_MeooeDag": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarse_op_Mooee": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarse_op_MooeeInv": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarse_op_MooeeDag": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarse_op_MooeeInvDag": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarse_op_prepare_mrhs": ([_vp, _i, _vp], _i),
    "mgc_coarse_op_M_mrhs": ([_vp, _vp, _vp, _i, _vp, _i, _i, _i, _vp], _i),
    "mgc_coarse_op_M_mrhs_batched": ([_vp, _i, _vp, _vp, _i, _vp, _i, _i, _i, _vp], _i),
    "mgc_coarse_solve_batched": ([_vp, _i, _vp, _vp, _i, _i, _c.c_double, _i, _i, _vp, _vp], _i),
    "mgc_coarse_batch_create": ([_vp, _vp, _i, _i, _i, _i, _vp], _i),
    "mgc_coarse_batch_solve": ([_vp, _vp, _vp, _c.c_double, _i, _vp, _vp], _i),
    "mgc_coarse_batch_destroy": ([_vp], _i),
    "mgc_coarse_batch_create_fixed": ([_vp, _vp, _i, _i, _i, _i, _i, _vp], _i),
    "mgc_coarse_op_M_mrhs_noalloc": ([_vp, _vp, _vp, _i, _i, _i, _vp], _i),
    "mgc_quantize": ([_vp, _sz, _i, _i, _vp, _vp, _vp], _i),
    "mgc_mrhs_set_variant": ([_i], _i),
    "mgc_mrhs_get_variant": ([_i, _i, _i, _ip], _i),
    "mgc_fine_op_create": ([_c.POINTER(_vp), _ip, _c.c_double, _c.c_double, _i], _i),
    "mgc_fine_op_destroy": ([_vp], _i),
    "mgc_fine_op_set_gauge": ([_vp, _vp, _vp], _i),
    "mgc_fine_op_get_diag": ([_vp, _vp, _vp], _i),
    "mgc_fine_op_M": ([_vp, _vp, _vp, _i, _i, _vp], _i),
    "mgc_fine_op_Mdag": ([_vp, _vp, _vp, _i, _i, _vp], _i),
    "mgc_fine_op_Mdiag": ([_vp, _vp, _vp, _i, _i, _vp], _i),
    "mgc_fine_op_Mdir": ([_vp, _vp, _vp, _i, _i, _i, _i, _vp], _i),
    "mgc_fine_op_set_variant": ([_i], _i),
    "mgc_random_gauge": ([_vp, _ip, _c.c_double, _u64, _u64, _vp], _i),
    "mgc_mg_params_default": ([_c.POINTER(MGParams)], _i),
    "mgc_mg_create": ([_c.POINTER(_vp), _vp, _c.POINTER(MGParams)], _i),
    "mgc_mg_destroy": ([_vp], _i),
    "mgc_mg_setup": ([_vp, _vp], _i),
    "mgc_mg_precondition": ([_vp, _vp, _vp, _vp], _i),
    "mgc_mg_vcycle": ([_vp, _vp, _vp, _i, _c.c_double, _vp], _i),
    "mgc_mg_solve": ([_vp, _vp, _vp, _c.c_double, _i, _i, _c.POINTER(SolveStats), _vp], _i),
    "mgc_mg_solve_mrhs": ([_vp, _vp, _vp, _i, _i, _c.c_double, _i, _i, _vp, _vp], _i),
    "mgc_mg_coarse_op": ([_vp, _i, _c.POINTER(_vp)], _i),
    "mgc_mg_aggregation": ([_vp, _i, _c.POINTER(_vp)], _i),
    "mgc_fine_gmres_solve": ([_vp, _vp, _vp, _c.c_double, _i, _i, _c.POINTER(SolveStats),
                              _vp], _i),
    "mgc_coarsen_operator": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_coarsen_set_variant": ([_i], _i),
    "mgc_test_params_default": ([_c.POINTER(TestParams)], _i),
    "mgc_params_read_xml": ([_c.c_char_p, _c.c_void_p, _c.c_void_p], _i),
    "mgc_params_write_xml": ([_c.c_char_p, _c.c_void_p, _c.c_void_p], _i),
    "mgc_gauge_read": ([_c.c_char_p, _ip, _vp, _sz, _c.POINTER(_c.c_double)], _i),
    "mgc_gauge_write": ([_c.c_char_p, _i, _ip, _vp, _i], _i),
    "mgc_coarsen_operator_linop": ([_vp, _c.POINTER(FineLinop), _vp, _i, _vp], _i),
    "mgc_fine_op_linop": ([_vp, _c.POINTER(FineLinop)], _i),
    "mgc_mg_create_linop": ([_c.POINTER(_vp), _c.POINTER(FineLinop), _c.POINTER(MGParams)], _i),
    "mgc_coarsen_coarse_operator": ([_vp, _vp, _vp, _i, _vp], _i),
    "mgc_aggregation_create_ex": ([_c.POINTER(_vp), _ip, _ip, _i, _i], _i),
    "mgc_aggregation_project_masked": ([_vp, _vp, _vp, _i, _i, _i, _i, _vp], _i),
    "mgc_aggregation_set_ortho_method": ([_i], _i),
    "mgc_coarse_op_construct_remaining_links": ([_vp, _i, _vp], _i),
    "mgc_aggregation_create": ([_c.POINTER(_vp), _ip, _ip, _i], _i),
    "mgc_aggregation_destroy": ([_vp], _i),
    "mgc_aggregation_info": ([_vp, _ip, _ip, _ip, _ip], _i),
    "mgc_aggregation_set_basis": ([_vp, _vp, _vp], _i),
    "mgc_aggregation_get_basis": ([_vp, _vp, _vp], _i),
    "mgc_aggregation_orthogonalise": ([_vp, _i, _vp], _i),
    "mgc_aggregation_check_orthogonal": ([_vp, _c.POINTER(_c.c_double), _vp], _i),
    "mgc_aggregation_project": ([_vp, _vp, _vp, _i, _i, _vp], _i),
    "mgc_aggregation_promote": ([_vp, _vp, _vp, _i, _i, _vp], _i),
}


class MgcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mgc error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libmgc.so once and attach argument types; raise when it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MgcError(-1, f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(code: int) -> None:
    if code != 0:
        raise MgcError(code, lib().mgc_last_error().decode())


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def declared_symbols() -> list:
    """Names of every function declared in include/mgc.h."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(mgc_[a-zA-Z0-9_]+)\s*\(", text)))


# ----------------------------------------------------------------------------
# runtime helpers
# ----------------------------------------------------------------------------
def device_count() -> int:
    n = _c.c_int(0)
    call("mgc_device_count", _c.byref(n))
    return n.value


def device_name() -> str:
    buf = _c.create_string_buffer(256)
    call("mgc_device_name", buf, 256)
    return buf.value.decode()


def synchronize() -> None:
    call("mgc_device_synchronize")


class Stream:
    def __init__(self):
        h = _vp()
        call("mgc_stream_create", _c.byref(h))
        self.handle = h.value

    def sync(self) -> None:
        call("mgc_stream_synchronize", self.handle)

    def close(self) -> None:
        if self.handle:
            call("mgc_stream_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = _vp()
        call("mgc_event_create", _c.byref(h))
        self.handle = h.value

    def record(self, stream: Optional[Stream] = None) -> None:
        call("mgc_event_record", self.handle, stream.handle if stream else None)

    def sync(self) -> None:
        call("mgc_event_synchronize", self.handle)

    def elapsed_ms(self, later: "Event") -> float:
        ms = _c.c_float(0.0)
        call("mgc_event_elapsed_ms", _c.byref(ms), self.handle, later.handle)
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                call("mgc_event_destroy", self.handle)
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation owned by Python; freed on close()/garbage collection."""

    def __init__(self, nbytes: int):
        h = _vp()
        call("mgc_malloc", _c.byref(h), _sz(int(nbytes)))
        self.ptr = h.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, arr: np.ndarray, stream: Optional[Stream] = None) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(arr.nbytes)
        call("mgc_memcpy_h2d", buf.ptr, arr.ctypes.data, _sz(arr.nbytes),
             stream.handle if stream else None)
        return buf

    def upload(self, arr: np.ndarray, stream: Optional[Stream] = None) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than the device buffer")
        call("mgc_memcpy_h2d", self.ptr, arr.ctypes.data, _sz(arr.nbytes),
             stream.handle if stream else None)

    def to_numpy(self, dtype, shape, stream: Optional[Stream] = None) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested more bytes than the buffer holds")
        call("mgc_memcpy_d2h", out.ctypes.data, self.ptr, _sz(out.nbytes),
             stream.handle if stream else None)
        return out

    def zero(self, stream: Optional[Stream] = None) -> None:
        call("mgc_fill_zero", self.ptr, _sz(self.nbytes), stream.handle if stream else None)

    def close(self) -> None:
        if self.ptr:
            call("mgc_free", self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def random_gaussian(buf: DeviceBuffer, n_complex: int, prec: int, seed: int, stream_id: int,
                    stream: Optional[Stream] = None) -> None:
    call("mgc_random_gaussian", buf.ptr, _sz(n_complex), prec, _u64(seed), _u64(stream_id),
         stream.handle if stream else None)


def complex_dtype(prec: int):
    return {F64: np.complex128, F32: np.complex64}[prec]


# ----------------------------------------------------------------------------
# coarse operator
# ----------------------------------------------------------------------------
def _p(x):
    """Pointer of a DeviceBuffer or raw int."""
    if x is None:
        return None
    return x.ptr if isinstance(x, DeviceBuffer) else int(x)


class CoarseOp:
    """Handle of a coarse operator (the reference's CoarsenedMatrix)."""

    def __init__(self, dims: Sequence[int], nbasis: int, prec: int = F64, hermitian: int = 0):
        h = _vp()
        d = (_c.c_int * 4)(*[int(v) for v in dims])
        call("mgc_coarse_op_create", _c.byref(h), d, int(nbasis), int(prec), int(hermitian))
        self.handle = h.value
        self.owned = True
        self.dims = tuple(int(v) for v in dims)
        self.N = int(nbasis)
        self.prec = int(prec)
        self.V = int(np.prod(self.dims))

    @classmethod
    def borrow(cls, handle: int, dims, nbasis: int, prec: int) -> "CoarseOp":
        """Wrap a handle owned elsewhere (e.g. by a multigrid hierarchy)."""
        self = cls.__new__(cls)
        self.handle, self.owned = handle, False
        self.dims, self.N, self.prec = tuple(dims), int(nbasis), int(prec)
        self.V = int(np.prod(self.dims))
        return self

    def close(self) -> None:
        if self.handle and self.owned:
            call("mgc_coarse_op_destroy", self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_links(self, links: DeviceBuffer, layout: int = LAYOUT_FLAT, invert: bool = True,
                  stream: Optional[Stream] = None) -> None:
        call("mgc_coarse_op_set_links", self.handle, layout, _p(links), int(bool(invert)),
             stream.handle if stream else None)

    def get_links(self, out: DeviceBuffer, layout: int = LAYOUT_FLAT, stream=None) -> None:
        call("mgc_coarse_op_get_links", self.handle, layout, _p(out),
             stream.handle if stream else None)

    def get_self_inverse(self, out: DeviceBuffer, stream=None) -> None:
        call("mgc_coarse_op_get_self_inverse", self.handle, _p(out),
             stream.handle if stream else None)

    def _s(self, stream):
        return stream.handle if stream else None

    def M(self, i, o, stream=None):
        call("mgc_coarse_op_M", self.handle, _p(i), _p(o), self._s(stream))

    def Mdag(self, i, o, stream=None):
        call("mgc_coarse_op_Mdag", self.handle, _p(i), _p(o), self._s(stream))

    def Mdiag(self, i, o, stream=None):
        call("mgc_coarse_op_Mdiag", self.handle, _p(i), _p(o), self._s(stream))

    def Mdir(self, i, o, dir_, disp, stream=None):
        call("mgc_coarse_op_Mdir", self.handle, _p(i), _p(o), int(dir_), int(disp),
             self._s(stream))

    def MdirAll(self, i, o8, stream=None):
        call("mgc_coarse_op_MdirAll", self.handle, _p(i), _p(o8), self._s(stream))

    def Dhop(self, i, o, dag=0, stream=None):
        call("mgc_coarse_op_Dhop", self.handle, _p(i), _p(o), int(dag), self._s(stream))

    def DhopOE(self, i, o, dag=0, stream=None):
        call("mgc_coarse_op_DhopOE", self.handle, _p(i), _p(o), int(dag), self._s(stream))

    def DhopEO(self, i, o, dag=0, stream=None):
        call("mgc_coarse_op_DhopEO", self.handle, _p(i), _p(o), int(dag), self._s(stream))

    def Meooe(self, i, o, in_cb, stream=None):
        call("mgc_coarse_op_Meooe", self.handle, _p(i), _p(o), int(in_cb), self._s(stream))

    def MeooeDag(self, i, o, in_cb, stream=None):
        call("mgc_coarse_op_MeooeDag", self.handle, _p(i), _p(o), int(in_cb), self._s(stream))

    def Mooee(self, i, o, cb=FULL, stream=None):
        call("mgc_coarse_op_Mooee", self.handle, _p(i), _p(o), int(cb), self._s(stream))

    def MooeeInv(self, i, o, cb=FULL, stream=None):
        call("mgc_coarse_op_MooeeInv", self.handle, _p(i), _p(o), int(cb), self._s(stream))

    def MooeeDag(self, i, o, cb=FULL, stream=None):
        call("mgc_coarse_op_MooeeDag", self.handle, _p(i), _p(o), int(cb), self._s(stream))

    def MooeeInvDag(self, i, o, cb=FULL, stream=None):
        call("mgc_coarse_op_MooeeInvDag", self.handle, _p(i), _p(o), int(cb), self._s(stream))

    def prepare_mrhs(self, mrhs_prec, stream=None):
        call("mgc_coarse_op_prepare_mrhs", self.handle, int(mrhs_prec), self._s(stream))

    def M_mrhs(self, i, o, nrhs, mrhs_prec=F32, in_prec=F32, in_scales=None, dag=0,
               stream=None):
        call("mgc_coarse_op_M_mrhs", self.handle, _p(i), _p(in_scales), int(in_prec), _p(o),
             int(nrhs), int(mrhs_prec), int(dag), self._s(stream))


    def M_mrhs_noalloc(self, i, o, nrhs, mrhs_prec=F16, dag=0, stream=None):
        """FP32 input, no device allocation (operand prepared beforehand)."""
        call("mgc_coarse_op_M_mrhs_noalloc", self.handle, _p(i), _p(o), int(nrhs),
             int(mrhs_prec), int(dag), self._s(stream))


def M_mrhs_batched(ops, ins, outs, nrhs, mrhs_prec=F32, in_scales=None, dag=0, stream=None):
    """One launch of the multi-rhs M over several coarse lattices (same nbasis)."""
    n = len(ops)
    if not (len(ins) == n and len(outs) == n):
        raise ValueError("ops, ins and outs must have the same length")
    arr = ctypes.c_void_p * n
    sc = arr(*[_p(x) for x in in_scales]) if in_scales is not None else None
    call("mgc_coarse_op_M_mrhs_batched", arr(*[o.handle for o in ops]), n,
         arr(*[_p(x) for x in ins]), sc, int(mrhs_prec), arr(*[_p(x) for x in outs]),
         int(nrhs), int(mrhs_prec), int(dag), stream.handle if stream else None)


def coarse_solve_batched(ops, bs, xs, nrhs, mrhs_prec=F32, tol=1e-6, maxit=200, restart=20,
                         stream=None) -> list:
    """Lockstep GMRES on several coarse lattices, one batched stencil launch per
    application; b, x per lattice are FP32 [V][nrhs][N].  Returns the stats of
    every (lattice, rhs) column, lattice-major."""
    n = len(ops)
    if not (len(bs) == n and len(xs) == n):
        raise ValueError("ops, bs and xs must have the same length")
    arr = ctypes.c_void_p * n
    st = (SolveStats * (n * int(nrhs)))()
    call("mgc_coarse_solve_batched", arr(*[o.handle for o in ops]), n, arr(*[_p(b) for b in bs]),
         arr(*[_p(x) for x in xs]), int(nrhs), int(mrhs_prec), float(tol), int(maxit),
         int(restart), st, stream.handle if stream else None)
    return [s.as_dict() for s in st]


class CoarseBatch:
    """A prepared lockstep coarse solve over several lattices (mgc_coarse_batch):
    the workspaces and Krylov basis are kept between solves."""

    def __init__(self, ops, nrhs, mrhs_prec=F32, restart=20, stream=None, fixed_steps=0,
                 graph=2):
        """fixed_steps > 0: mgc_coarse_batch_create_fixed with that step count
        and graph mode (1 eager, 2 captured HIP graph)."""
        self.ops, self.nrhs, self.stream = list(ops), int(nrhs), stream
        arr = ctypes.c_void_p * len(self.ops)
        h = _vp()
        if fixed_steps:
            call("mgc_coarse_batch_create_fixed", _c.byref(h),
                 arr(*[o.handle for o in self.ops]), len(self.ops), self.nrhs, int(mrhs_prec),
                 int(fixed_steps), int(graph), stream.handle if stream else None)
        else:
            call("mgc_coarse_batch_create", _c.byref(h), arr(*[o.handle for o in self.ops]),
                 len(self.ops), self.nrhs, int(mrhs_prec), int(restart),
                 stream.handle if stream else None)
        self.handle = h.value

    def solve(self, bs, xs, tol=1e-6, maxit=200) -> list:
        n = len(self.ops)
        if not (len(bs) == n and len(xs) == n):
            raise ValueError("one b and one x per lattice")
        arr = ctypes.c_void_p * n
        st = (SolveStats * (n * self.nrhs))()
        call("mgc_coarse_batch_solve", self.handle, arr(*[_p(b) for b in bs]),
             arr(*[_p(x) for x in xs]), float(tol), int(maxit), st,
             self.stream.handle if self.stream else None)
        return [s.as_dict() for s in st]

    def close(self) -> None:
        if self.handle:
            call("mgc_coarse_batch_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def quantize(inp, rows, n, prec, out, scales, stream=None):
    call("mgc_quantize", _p(inp), _sz(int(rows)), int(n), int(prec), _p(out), _p(scales),
         stream.handle if stream else None)


# ----------------------------------------------------------------------------
# aggregation
# ----------------------------------------------------------------------------
class Aggregation:
    """Handle of a block-transfer operator (the reference's Aggregation)."""

    def __init__(self, fine_dims: Sequence[int], block: Sequence[int], nb: int,
                 site_elems: int = 12):
        h = _vp()
        fd = (_c.c_int * 4)(*[int(v) for v in fine_dims])
        bd = (_c.c_int * 4)(*[int(v) for v in block])
        call("mgc_aggregation_create_ex", _c.byref(h), fd, bd, int(nb), int(site_elems))
        self.handle = h.value
        self.owned = True
        self._geometry(fine_dims, block, nb, site_elems)

    def _geometry(self, fine_dims, block, nb, site_elems):
        self.Ef = int(site_elems)
        self.fine_dims = tuple(int(v) for v in fine_dims)
        self.block = tuple(int(v) for v in block)
        self.coarse_dims = tuple(f // b for f, b in zip(self.fine_dims, self.block))
        self.nb = int(nb)
        self.N = 2 * self.nb
        self.Vf = int(np.prod(self.fine_dims))
        self.Vc = int(np.prod(self.coarse_dims))

    @classmethod
    def borrow(cls, handle: int, fine_dims, block, nb: int, site_elems: int = 12):
        """Wrap a handle owned elsewhere (e.g. by a multigrid hierarchy)."""
        self = cls.__new__(cls)
        self.handle, self.owned = handle, False
        self._geometry(fine_dims, block, nb, site_elems)
        return self

    def close(self) -> None:
        if self.handle and self.owned:
            call("mgc_aggregation_destroy", self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _s(self, stream):
        return stream.handle if stream else None

    def set_basis(self, basis, stream=None) -> None:
        call("mgc_aggregation_set_basis", self.handle, _p(basis), self._s(stream))

    def get_basis(self, out, stream=None) -> None:
        call("mgc_aggregation_get_basis", self.handle, _p(out), self._s(stream))

    def orthogonalise(self, passes: int = 2, stream=None) -> None:
        call("mgc_aggregation_orthogonalise", self.handle, int(passes), self._s(stream))

    def check_orthogonal(self, stream=None) -> np.ndarray:
        errs = np.zeros(self.nb, dtype=np.float64)
        call("mgc_aggregation_check_orthogonal", self.handle,
             errs.ctypes.data_as(_c.POINTER(_c.c_double)), self._s(stream))
        return errs

    def project(self, fine, coarse, prec=F64, nrhs=1, stream=None) -> None:
        call("mgc_aggregation_project", self.handle, _p(fine), _p(coarse), int(prec), int(nrhs),
             self._s(stream))

    def project_masked(self, fine, coarse, kind, point, prec=F64, nrhs=1, stream=None) -> None:
        call("mgc_aggregation_project_masked", self.handle, _p(fine), _p(coarse), int(prec),
             int(nrhs), int(kind), int(point), self._s(stream))

    def promote(self, coarse, fine, prec=F64, nrhs=1, stream=None) -> None:
        call("mgc_aggregation_promote", self.handle, _p(coarse), _p(fine), int(prec), int(nrhs),
             self._s(stream))


# ----------------------------------------------------------------------------
# fine Wilson-clover operator
# ----------------------------------------------------------------------------
class FineOp:
    """Wilson-clover fine operator (the fine LinearOperator of CoarsenOperator)."""

    def __init__(self, dims: Sequence[int], mass: float = -0.1, csw: float = 1.0,
                 antiperiodic_t: bool = False):
        h = _vp()
        d = (_c.c_int * 4)(*[int(v) for v in dims])
        call("mgc_fine_op_create", _c.byref(h), d, float(mass), float(csw),
             int(bool(antiperiodic_t)))
        self.handle = h.value
        self.dims = tuple(int(v) for v in dims)
        self.V = int(np.prod(self.dims))

    def close(self) -> None:
        if self.handle:
            call("mgc_fine_op_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _s(self, stream):
        return stream.handle if stream else None

    def set_gauge(self, gauge, stream=None):
        call("mgc_fine_op_set_gauge", self.handle, _p(gauge), self._s(stream))

    def get_diag(self, out, stream=None):
        call("mgc_fine_op_get_diag", self.handle, _p(out), self._s(stream))

    def M(self, i, o, prec=F64, nrhs=1, stream=None):
        call("mgc_fine_op_M", self.handle, _p(i), _p(o), int(prec), int(nrhs), self._s(stream))

    def Mdag(self, i, o, prec=F64, nrhs=1, stream=None):
        call("mgc_fine_op_Mdag", self.handle, _p(i), _p(o), int(prec), int(nrhs),
             self._s(stream))

    def Mdiag(self, i, o, prec=F64, nrhs=1, stream=None):
        call("mgc_fine_op_Mdiag", self.handle, _p(i), _p(o), int(prec), int(nrhs),
             self._s(stream))

    def Mdir(self, i, o, dir_, disp, prec=F64, nrhs=1, stream=None):
        call("mgc_fine_op_Mdir", self.handle, _p(i), _p(o), int(dir_), int(disp), int(prec),
             int(nrhs), self._s(stream))


def random_gauge(buf: DeviceBuffer, dims, eps=0.4, seed=1, stream_id=0, stream=None):
    d = (_c.c_int * 4)(*[int(v) for v in dims])
    call("mgc_random_gauge", buf.ptr, d, float(eps), _u64(seed), _u64(stream_id),
         stream.handle if stream else None)


COARSEN_FULL, COARSEN_DIRSAVE = 0, 1


def coarsen_operator(cop: CoarseOp, fop: FineOp, agg: "Aggregation", mode=COARSEN_DIRSAVE,
                     stream=None):
    """CoarsenOperator: build cop's links from the fine operator and the basis."""
    call("mgc_coarsen_operator", cop.handle, fop.handle, agg.handle, int(mode),
         stream.handle if stream else None)


def fine_op_linop(fop: "FineOp") -> FineLinop:
    """The library's own fine operator behind the callback interface."""
    lop = FineLinop()
    call("mgc_fine_op_linop", fop.handle, _c.byref(lop))
    return lop


def make_linop(fn, dims, precisions=(F64,), max_nrhs=1, gamma5_hermitian=True) -> FineLinop:
    """A callback operator from a Python function
    fn(point, prec, in_ptr, out_ptr, nrhs, stream) -> int (0 on success).
    Keep the returned object alive while the library may call it."""
    lop = FineLinop()
    lop._py_fn = LINOP_APPLY(lambda ctx, point, prec, i, o, nrhs, st: fn(point, prec, i, o, nrhs,
                                                                        st))
    lop.apply = lop._py_fn
    lop.ctx = None
    for d in range(4):
        lop.dims[d] = int(dims[d])
    lop.precisions = sum(1 << p for p in precisions)
    lop.max_nrhs = int(max_nrhs)
    lop.gamma5_hermitian = int(bool(gamma5_hermitian))
    return lop


def coarsen_operator_linop(cop: CoarseOp, lop: FineLinop, agg: "Aggregation",
                           mode=COARSEN_DIRSAVE, stream=None):
    """CoarsenOperator of a callback operator."""
    call("mgc_coarsen_operator_linop", cop.handle, _c.byref(lop), agg.handle, int(mode),
         stream.handle if stream else None)


def coarsen_coarse_operator(cop2: CoarseOp, cop1: CoarseOp, agg: "Aggregation",
                            mode=COARSEN_DIRSAVE, stream=None):
    """CoarsenOperator of a coarse operator onto the next level."""
    call("mgc_coarsen_coarse_operator", cop2.handle, cop1.handle, agg.handle, int(mode),
         stream.handle if stream else None)


def mg_params(**kw) -> MGParams:
    """Reference defaults (include/MultigridParams.h:95-133) with overrides.
    block / nb take per-coarsening lists."""
    p = MGParams()
    call("mgc_mg_params_default", _c.byref(p))
    for k, v in kw.items():
        if k == "block":
            for l, b in enumerate(v):
                for d in range(4):
                    p.block[l][d] = int(b[d])
        elif k == "nb":
            for l, b in enumerate(v):
                p.nb[l] = int(b)
        else:
            setattr(p, k, v)
    return p


class Multigrid:
    """Handle of the multigrid hierarchy (MultiGridPreconditioner)."""

    def __init__(self, fop, params: MGParams):
        """fop: a FineOp, or a FineLinop (callback operator, kept alive here)."""
        h = _vp()
        if isinstance(fop, FineLinop):
            call("mgc_mg_create_linop", _c.byref(h), _c.byref(fop), _c.byref(params))
        else:
            call("mgc_mg_create", _c.byref(h), fop.handle, _c.byref(params))
        self.handle = h.value
        self.fop = fop
        self.params = params
        self.stream = None

    def close(self):
        if self.handle:
            call("mgc_mg_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _s(self):
        return self.stream.handle if self.stream else None

    def setup(self, stream=None):
        self.stream = stream
        call("mgc_mg_setup", self.handle, self._s())

    def precondition(self, i, o):
        call("mgc_mg_precondition", self.handle, _p(i), _p(o), self._s())

    def vcycle(self, i, o, max_iter=1, tol=0.0):
        call("mgc_mg_vcycle", self.handle, _p(i), _p(o), int(max_iter), float(tol), self._s())

    def solve(self, b, x, tol=1e-10, maxit=200, restart=20) -> dict:
        st = SolveStats()
        call("mgc_mg_solve", self.handle, _p(b), _p(x), float(tol), int(maxit), int(restart),
             _c.byref(st), self._s())
        return st.as_dict()

    def solve_mrhs(self, b, x, nrhs, coarse_prec=F32, tol=1e-10, maxit=200, restart=20) -> list:
        """Batched solve of nrhs systems; b, x are [site][rhs][12] FP64."""
        st = (SolveStats * int(nrhs))()
        call("mgc_mg_solve_mrhs", self.handle, _p(b), _p(x), int(nrhs), int(coarse_prec),
             float(tol), int(maxit), int(restart), st, self._s())
        return [s.as_dict() for s in st]

    def coarse_op_handle(self, level: int) -> int:
        h = _vp()
        call("mgc_mg_coarse_op", self.handle, int(level), _c.byref(h))
        return h.value

    def aggregation_handle(self, level: int) -> int:
        h = _vp()
        call("mgc_mg_aggregation", self.handle, int(level), _c.byref(h))
        return h.value


def fine_gmres_solve(fop, b, x, tol=1e-10, maxit=1000, restart=20, stream=None) -> dict:
    st = SolveStats()
    call("mgc_fine_gmres_solve", fop.handle, _p(b), _p(x), float(tol), int(maxit), int(restart),
         _c.byref(st), stream.handle if stream else None)
    return st.as_dict()


def construct_remaining_links(cop: CoarseOp, disp_have=-1, stream=None):
    call("mgc_coarse_op_construct_remaining_links", cop.handle, int(disp_have),
         stream.handle if stream else None)


# ----------------------------------------------------------------------------
# on-disk formats (host only)
# ----------------------------------------------------------------------------
GAUGE_NERSC, GAUGE_ILDG = 0, 1


def test_params() -> TestParams:
    t = TestParams()
    call("mgc_test_params_default", _c.byref(t))
    return t


def read_params_xml(path: str):
    """(MGParams, TestParams) from the reference drivers' XML parameter file."""
    mg, t = MGParams(), TestParams()
    call("mgc_params_read_xml", os.fsencode(path), _c.byref(mg), _c.byref(t))
    return mg, t


def write_params_xml(path: str, mg: MGParams, test: Optional[TestParams] = None) -> None:
    call("mgc_params_write_xml", os.fsencode(path), _c.byref(mg),
         _c.byref(test) if test is not None else None)


def read_gauge(path: str):
    """(dims, gauge [4][V][3][3] complex128 on the host, plaquette) of a NERSC or ILDG file."""
    dims = (_c.c_int * 4)()
    call("mgc_gauge_read", os.fsencode(path), dims, None, _sz(0), None)
    V = int(np.prod(list(dims)))
    U = np.empty((4, V, 3, 3), dtype=np.complex128)
    plaq = _c.c_double()
    call("mgc_gauge_read", os.fsencode(path), dims, U.ctypes.data, _sz(U.nbytes), _c.byref(plaq))
    return tuple(dims), U, plaq.value


def write_gauge(path: str, dims, U: np.ndarray, fmt: int = GAUGE_NERSC, bits: int = 64) -> None:
    U = np.ascontiguousarray(U, dtype=np.complex128)
    d = (_c.c_int * 4)(*[int(v) for v in dims])
    call("mgc_gauge_write", os.fsencode(path), int(fmt), d, U.ctypes.data, int(bits))
