"""Python mirror of the reference's SWIG module ``zefr`` (include/zefr.i, include/zefr.hpp,
include/zefr_interface.hpp, include/inputstruct.hpp), over the C ABI of libzefr_amd.so.

Same names and call sequence as the reference's Python harness
(run/zefrInterface.py:43-417)::

    from zefr_amd import zefr
    zefr.initialize("input_taylorgreen")
    z = zefr.get_zefr_object()
    inp = z.get_input(); inp.dt = 2e-4; z.init_inputs()
    z.setup_solver()
    for it in range(n): z.do_step()
    res = z.write_residual()

Errors raised inside the library come back as ``RuntimeError`` with the library's
message.  Overset / TIOGA callbacks and moving-grid members of the reference's API are
absent from this build (single grid, fixed mesh).
"""

from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._lib import ZefrInput, check, last_error, lib

__all__ = ["InputStruct", "DataStruct", "Zefr", "initialize", "get_zefr_object", "set_zefr_object",
           "finalize", "use_gpus", "get_q_spts", "get_dq_spts", "get_q_spts_d", "get_dq_spts_d",
           "get_q_spt"]

_STR_FIELDS = {"dt_scheme", "meshfile", "output_prefix", "spt_type", "mg_cycle", "restart_case"}
_LEVEL_FIELDS = ("mg_levels", "mg_steps")


class InputStruct:
    """Attribute view of the library's ``zefr_input_t`` (InputStruct field names)."""

    def __init__(self, ptr):
        object.__setattr__(self, "_p", ptr)

    def __getattr__(self, name):
        s = self._p.contents
        if name == "meshBounds":
            return {s.meshBoundName[i].value.decode(): s.meshBoundBC[i].value.decode()
                    for i in range(s.nMeshBounds)}
        v = getattr(s, name)
        if name in _STR_FIELDS:
            return v.decode()
        if name in ("V_fs", "norm_fs"):
            return list(v)
        if name in _LEVEL_FIELDS:
            return list(v)[:s.nLevels]
        return v

    def __setattr__(self, name, value):
        s = self._p.contents
        if name in _STR_FIELDS:
            value = value.encode()
        elif name in ("V_fs", "norm_fs"):
            arr = getattr(s, name)
            for i, x in enumerate(value):
                arr[i] = x
            return
        elif name in _LEVEL_FIELDS:
            arr = getattr(s, name)
            for i, x in enumerate(value):
                arr[i] = x
            s.nLevels = len(value)
            return
        elif name == "meshBounds":
            s.nMeshBounds = len(value)
            for i, (k, b) in enumerate(value.items()):
                s.meshBoundName[i].value = k.lower().encode()
                s.meshBoundBC[i].value = b.lower().encode()
            return
        setattr(s, name, value)


class DataStruct:
    """``DataStruct`` (include/inputstruct.hpp:99-111): field count, spts per element type
    and device pointers of the solution / gradient of element type 0."""

    def __init__(self, z):
        self._z = z

    @property
    def nfields(self):
        return self._z.dims()["nVars"]

    @property
    def nspts(self):
        return [self._z.dims()["nSpts"]]

    @property
    def u_spts_d(self):
        return [self._z.get_u_spts_d()[0]]


class Zefr:
    """The reference's ``Zefr`` object (include/zefr.hpp:24-209)."""

    def __init__(self, handle=None, owned=True):
        self._h = handle if handle is not None else lib().zefr_create()
        if not self._h:
            raise RuntimeError(last_error())
        self._owned = owned and handle is None
        self._t_start = time.perf_counter()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and getattr(self, "_owned", False):
            lib().zefr_destroy(h)
            self._h = None

    # -- setup (src/zefr.cpp:286-363)
    def read_input(self, inputfile):
        check(lib().zefr_read_input(self._h, inputfile.encode()))

    def get_input(self):
        return InputStruct(lib().zefr_get_input(self._h))

    def get_data(self):
        return DataStruct(self)

    def init_inputs(self):
        check(lib().zefr_init_inputs(self._h))

    def setup_solver(self):
        check(lib().zefr_setup_solver(self._h))

    def setup_host(self):
        """Everything of setup_solver that runs on the host (no GPU touched)."""
        check(lib().zefr_setup_host(self._h))

    # -- stepping (src/zefr.cpp:371-430)
    def do_step(self):
        check(lib().zefr_do_step(self._h))

    def do_rk_stage(self, iter, stage):
        check(lib().zefr_do_rk_stage(self._h, iter, stage))

    def do_rk_stage_start(self, iter, stage):
        check(lib().zefr_do_rk_stage_start(self._h, iter, stage))

    def do_rk_stage_mid(self, iter, stage):
        check(lib().zefr_do_rk_stage_mid(self._h, iter, stage))

    def do_rk_stage_finish(self, iter, stage):
        check(lib().zefr_do_rk_stage_finish(self._h, iter, stage))

    def write_residual(self):
        r = np.zeros(self.dims()["nVars"])
        check(lib().zefr_write_residual(self._h, r.ctypes.data_as(C.POINTER(C.c_double))))
        return r

    # -- Zefr methods outside the residual path (src/zefr.cpp:432-484)
    def adapt_dt(self):
        """Zefr::adapt_dt (src/zefr.cpp:432-438): without adapt_dt (the only mode here;
        adaptive steps are rejected at input) it returns the supplied dt."""
        return self.get_input().dt

    def update_averages(self):
        """Zefr::update_averages (src/zefr.cpp:455-459): a no-op without tavg (time averaging
        is not supported)."""

    def write_averages(self):
        """Zefr::write_averages (src/zefr.cpp:461-465): a no-op without tavg."""

    def write_wall_time(self):
        """Zefr::write_wall_time (src/zefr.cpp:477-484): the wall time since the object was
        created."""
        print(f"Elapsed time: {time.perf_counter() - self._t_start} s")

    def write_forces(self):
        """Zefr::write_forces (src/zefr.cpp:467-470): wall forces need wall boundaries; only
        periodic meshes are supported."""
        raise NotImplementedError("write_forces: wall boundaries are not supported (periodic meshes only)")

    def write_error(self):
        """Zefr::write_error (src/zefr.cpp:472-475): error norms against analytic solutions
        are not computed."""
        raise NotImplementedError("write_error: error norms against analytic solutions are not supported")

    # -- FRSolver-level calls
    def write_solution(self):
        """Zefr::write_solution: with write_pyfr, <output_prefix>/<output_prefix>-<iter>.pyfrs."""
        check(lib().zefr_write_solution(self._h))

    def restart_solution(self):
        """Zefr::restart_solution: with restart (restart_type 2), read
        <restart_case>/<restart_case>-<restart_iter>.pyfrs into the solver."""
        check(lib().zefr_restart_solution(self._h))

    def compute_residual(self, stage=0):
        check(lib().zefr_compute_residual(self._h, stage))

    def compute_residual_start(self, stage=0):
        check(lib().zefr_compute_residual_start(self._h, stage))

    def compute_residual_mid(self, stage=0):
        check(lib().zefr_compute_residual_mid(self._h, stage))

    def compute_residual_finish(self, stage=0):
        check(lib().zefr_compute_residual_finish(self._h, stage))

    def synchronize(self):
        check(lib().zefr_synchronize(self._h))

    def autotune(self, reps=3):
        eb = lib().zefr_autotune(self._h, reps)
        if eb < 0:
            raise RuntimeError(last_error())
        return eb

    def set_tile(self, eles_per_block):
        check(lib().zefr_set_tile(self._h, eles_per_block))

    def set_tiles(self, tiles):
        """Per-launch workgroup tiles (trace, gradient/one-sided, divergence/face correction)."""
        arr = (C.c_int * 3)(*tiles)
        check(lib().zefr_set_tiles(self._h, arr))

    def tiles(self):
        arr = (C.c_int * 3)()
        check(lib().zefr_get_tiles(self._h, arr))
        return list(arr)

    def set_step_mode(self, fuse_rk=True, use_graph=False):
        """RK update fused into each residual's last kernel; whole step as a HIP graph."""
        check(lib().zefr_set_step_mode(self._h, int(fuse_rk), int(use_graph)))

    def set_pmg_persistent(self, on=True):
        """p-multigrid V-cycles as one persistent (cooperative) launch; inviscid runs."""
        check(lib().zefr_set_pmg_persistent(self._h, int(on)))

    # -- functional residual and its vector-Jacobian product on device tensors
    def state_tensor(self):
        """A device tensor copy of the current state in the device layout."""
        import torch
        return torch.from_numpy(self.state("U", layout="device")).to("cuda")

    def _tensor_ok(self, t):
        d = self.dims()
        shape = (d["nEles"], d["nVars"], d["nSpts"])
        if tuple(t.shape) != shape or not t.is_contiguous() or str(t.dtype) != "torch.float64" or not t.is_cuda:
            raise ValueError(f"expected a contiguous float64 device tensor of shape {shape} "
                             "(the device layout [nEles][nVars][nSpts])")

    # Both run on torch's current stream (torch and the library share one HIP runtime, see
    # _lib._init_torch_first), ordered with the surrounding torch work.
    def residual_at(self, U, out=None):
        """R(U) = divF for the device state U (torch float64 [nEles][nVars][nSpts]),
        without touching the solver's own state."""
        import torch
        self._tensor_ok(U)
        out = torch.empty_like(U) if out is None else out
        self._tensor_ok(out)
        self.synchronize()  # the solver's own stream is non-blocking; it shares the scratch
        st = torch.cuda.current_stream().cuda_stream
        check(lib().zefr_residual_at(self._h, U.data_ptr(), out.data_ptr(), st))
        return out

    def residual_vjp(self, U, w, out=None):
        """v = (dR/dU)^T w at U (device tensors in the device layout)."""
        import torch
        self._tensor_ok(U)
        self._tensor_ok(w)
        out = torch.empty_like(U) if out is None else out
        self._tensor_ok(out)
        self.synchronize()  # the solver's own stream is non-blocking; it shares the scratch
        st = torch.cuda.current_stream().cuda_stream
        check(lib().zefr_residual_vjp(self._h, U.data_ptr(), w.data_ptr(), out.data_ptr(), st))
        return out

    # -- data access (src/zefr.cpp:560-650)
    def _strides(self, fn, with_dim, etype):
        s = [C.c_int() for _ in range(4 if with_dim else 3)]
        ptr = fn(self._h, *[C.byref(x) for x in s], etype)
        if not ptr:
            raise RuntimeError(last_error())
        return (ptr,) + tuple(x.value for x in s)

    def get_u_spts_d(self, etype=0):
        """(device pointer, ele_stride, spt_stride, var_stride)."""
        return self._strides(lib().zefr_get_u_spts_d, False, etype)

    def get_du_spts_d(self, etype=0):
        """(device pointer, ele_stride, spt_stride, var_stride, dim_stride)."""
        return self._strides(lib().zefr_get_du_spts_d, True, etype)

    def get_u_spts(self, etype=0):
        """Host copy as a numpy array [nEles, nVars, nSpts] plus the strides."""
        ptr, es, ss, vs = self._strides(lib().zefr_get_u_spts, False, etype)
        d = self.dims()
        a = np.ctypeslib.as_array(ptr, shape=(d["nEles"] * d["nVars"] * d["nSpts"],))
        return a.reshape(d["nEles"], d["nVars"], d["nSpts"]).copy(), es, ss, vs

    def get_du_spts(self, etype=0):
        ptr, es, ss, vs, ds = self._strides(lib().zefr_get_du_spts, True, etype)
        d = self.dims()
        n = d["nEles"] * d["nDims"] * d["nVars"] * d["nSpts"]
        a = np.ctypeslib.as_array(ptr, shape=(n,))
        return a.reshape(d["nEles"], d["nDims"], d["nVars"], d["nSpts"]).copy(), es, ss, vs, ds

    def get_u_spt(self, ele, spt, var):
        return lib().zefr_get_u_spt(self._h, ele, spt, var)

    def get_grad_spt(self, ele, spt, dim, var):
        return lib().zefr_get_grad_spt(self._h, ele, spt, dim, var)

    # -- conveniences for tests and benchmarks
    def dims(self):
        d = (C.c_int * 10)()
        check(lib().zefr_get_dims(self._h, d))
        keys = ["nEles", "nSpts", "nFpts", "nVars", "nDims", "nGfpts", "nGfpts_int", "nStages", "order",
                "etype"]
        return dict(zip(keys, list(d)))

    def _copy(self, which, stage, arr, to_device):
        check(lib().zefr_copy_state(self._h, which, stage, arr.ctypes.data_as(C.POINTER(C.c_double)),
                                    arr.size, int(to_device)))

    def state(self, which="U", stage=0, layout="ref"):
        """Copy U / U_ini / divF(stage) to the host.  layout 'ref' returns the reference's
        [nSpts][nVars][nEles]; 'device' returns [nEles][nVars][nSpts]."""
        d = self.dims()
        a = np.empty((d["nEles"], d["nVars"], d["nSpts"]))
        self._copy({"U": 0, "U_ini": 1, "divF": 2}[which], stage, a, False)
        return np.ascontiguousarray(a.transpose(2, 1, 0)) if layout == "ref" else a

    def set_state(self, U, layout="ref"):
        """Overwrite U on the device from the reference layout [nSpts][nVars][nEles]."""
        d = self.dims()
        U = np.asarray(U, dtype=np.float64)
        if layout == "ref":
            U = U.transpose(2, 1, 0)
        a = np.ascontiguousarray(U)
        assert a.shape == (d["nEles"], d["nVars"], d["nSpts"])
        self._copy(0, 0, a, True)

    def pmg_state(self, level, which="U"):
        """A p-multigrid level's array in the device layout [nEles][nVars][nSpts(level)]:
        which = 'U', 'divF' (stage 0), 'source', 'solution' or 'correction'."""
        d = self.dims()
        inp = self.get_input()
        nS = (inp.mg_levels[level] + 1) ** d["nDims"]
        a = np.empty((d["nEles"], d["nVars"], nS))
        code = {"U": 0, "divF": 2, "source": 3, "solution": 4, "correction": 5}[which]
        check(lib().zefr_pmg_copy_state(self._h, level, code, a.ctypes.data_as(C.POINTER(C.c_double)), a.size, 0))
        return a

    def set_pmg_state(self, level, U):
        """Overwrite a p-multigrid level's U from the device layout [nEles][nVars][nSpts(level)]."""
        a = np.ascontiguousarray(np.asarray(U, dtype=np.float64))
        check(lib().zefr_pmg_copy_state(self._h, level, 0, a.ctypes.data_as(C.POINTER(C.c_double)), a.size, 1))

    def fpt_maps(self):
        d = self.dims()
        g = np.zeros((d["nFpts"], d["nEles"]), np.int32)
        s = np.zeros((d["nFpts"], d["nEles"]), np.int8)
        check(lib().zefr_get_fpt_maps(self._h, g.ctypes.data_as(C.POINTER(C.c_int)),
                                      s.ctypes.data_as(C.POINTER(C.c_byte))))
        return g, s

    def host_array(self, name):
        n = lib().zefr_get_host_array(self._h, name.encode(), None, 0)
        if n < 0:
            raise RuntimeError(last_error())
        buf = np.zeros(n)
        if lib().zefr_get_host_array(self._h, name.encode(), buf.ctypes.data_as(C.POINTER(C.c_double)), n) != n:
            raise RuntimeError(last_error())
        return buf


class Batch:
    """Residuals of several set-up solvers as one HIP graph launch (zefr_batch_*): the
    members' compute_residual(0) as a multi-case plan, one kernel launch per kernel, shape
    and flux branch over all members (``launches``)."""

    def __init__(self, solvers):
        self._solvers = list(solvers)  # kept alive as long as the graph refers to them
        arr = (C.c_void_p * len(self._solvers))(*[z._h for z in self._solvers])
        self._h = lib().zefr_batch_create(arr, len(self._solvers))
        if not self._h:
            raise RuntimeError(last_error())

    def residual(self, sync=True):
        for z in self._solvers:
            z.synchronize()
        check(lib().zefr_batch_residual(self._h))
        if sync:
            self.synchronize()

    def synchronize(self):
        check(lib().zefr_batch_synchronize(self._h))

    @property
    def launches(self):
        return lib().zefr_batch_launches(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().zefr_batch_destroy(h)
            self._h = None


# ---------------------------------------------------------------- module-level (zefr namespace)
_singleton = None


def initialize(input_file):
    """zefr::initialize (include/zefr_interface.hpp:94)."""
    check(lib().zefr_initialize(input_file.encode()))


def get_zefr_object():
    global _singleton
    h = lib().zefr_get_zefr_object()
    if not h:
        return None
    if _singleton is None or _singleton._h != h:
        _singleton = Zefr(handle=h, owned=False)
    return _singleton


def set_zefr_object(z):
    global _singleton
    lib().zefr_set_zefr_object(z._h if z is not None else None)
    if z is not None:
        z._owned = False
    _singleton = z


def finalize():
    global _singleton
    lib().zefr_finalize()
    _singleton = None


def use_gpus():
    return bool(lib().zefr_use_gpus())


def _single(fn, ndim):
    s = [C.c_int() for _ in range(ndim)]
    p = fn(*[C.byref(x) for x in s], 0)
    if not p:
        raise RuntimeError(last_error())
    return (p,) + tuple(x.value for x in s)


def get_q_spts(etype=0):
    return get_zefr_object().get_u_spts(etype)


def get_dq_spts(etype=0):
    return get_zefr_object().get_du_spts(etype)


def get_q_spts_d(etype=0):
    return get_zefr_object().get_u_spts_d(etype)


def get_dq_spts_d(etype=0):
    return get_zefr_object().get_du_spts_d(etype)


def get_q_spt(ele, spt, var):
    return lib().zefr_get_q_spt(ele, spt, var)


# TIOGA/overset data exchange (include/zefr_interface.hpp:98-112): no overset grids here.
def _no_overset(name):
    def fn(*args, **kwargs):
        raise NotImplementedError(f"{name}: overset (TIOGA) grids are not supported")
    fn.__name__ = name
    fn.__doc__ = f"zefr::{name} (include/zefr_interface.hpp:98-112): overset grids are not supported."
    return fn


get_basic_geo_data = _no_overset("get_basic_geo_data")
get_extra_geo_data = _no_overset("get_extra_geo_data")
get_gpu_geo_data = _no_overset("get_gpu_geo_data")
get_callback_funcs = _no_overset("get_callback_funcs")
