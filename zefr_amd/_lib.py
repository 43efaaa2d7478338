"""ctypes binding of ``libzefr_amd.so`` (declarations in include/zefr_amd.h).

The shared library is built in-tree by ``zefr_amd/csrc/Makefile`` (``build()`` below or
``__graft_entry__.build()``).  Importing this module never builds anything and never
falls back to another implementation: if the library is missing, ``lib()`` raises.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzefr_amd.so")
CSRC = os.path.join(HERE, "csrc")

MAX_BOUNDS = 16
NAME_LEN = 64
MAX_LEVELS = 8


class ZefrInput(C.Structure):
    """``zefr_input_t``: the residual-path subset of the reference's InputStruct."""

    _fields_ = [
        ("equation", C.c_uint), ("nDims", C.c_uint), ("order", C.c_uint), ("viscous", C.c_int),
        ("fconv_type", C.c_uint), ("fvisc_type", C.c_uint),
        ("rus_k", C.c_double), ("ldg_b", C.c_double), ("ldg_tau", C.c_double),
        ("dt_scheme", C.c_char * 16), ("nStages", C.c_uint), ("dt_type", C.c_uint), ("dt", C.c_double),
        ("n_steps", C.c_uint), ("report_freq", C.c_uint), ("write_freq", C.c_uint),
        ("res_type", C.c_uint), ("res_field", C.c_uint), ("ic_type", C.c_uint), ("test_case", C.c_uint),
        ("gamma", C.c_double), ("prandtl", C.c_double), ("mu", C.c_double), ("R", C.c_double),
        ("S", C.c_double), ("T_gas", C.c_double), ("rt", C.c_double), ("c_sth", C.c_double),
        ("rho_fs", C.c_double), ("P_fs", C.c_double), ("v_mag_fs", C.c_double), ("mach_fs", C.c_double),
        ("Re_fs", C.c_double), ("L_fs", C.c_double), ("T_fs", C.c_double), ("T_tot_fs", C.c_double),
        ("P_tot_fs", C.c_double), ("V_fs", C.c_double * 3), ("norm_fs", C.c_double * 3),
        ("fix_vis", C.c_int), ("disable_nondim", C.c_int),
        ("meshfile", C.c_char * 512), ("output_prefix", C.c_char * 256), ("spt_type", C.c_char * 32),
        ("nMeshBounds", C.c_int),
        ("meshBoundName", (C.c_char * NAME_LEN) * MAX_BOUNDS),
        ("meshBoundBC", (C.c_char * NAME_LEN) * MAX_BOUNDS),
        ("iter", C.c_uint), ("initIter", C.c_uint), ("time", C.c_double),
        ("p_multi", C.c_int), ("nLevels", C.c_uint), ("mg_levels", C.c_uint * MAX_LEVELS),
        ("mg_steps", C.c_uint * MAX_LEVELS), ("mg_cycle", C.c_char * 8), ("FMG_vcycles", C.c_uint),
        ("rel_fac", C.c_double),
        ("restart", C.c_int), ("restart_type", C.c_uint), ("restart_iter", C.c_uint),
        ("restart_case", C.c_char * 256), ("write_pyfr", C.c_int), ("write_paraview", C.c_int),
    ]


class EleDesc(C.Structure):
    """``zfr_ele_desc_t``: one element block for the kernel-level entry points."""

    _fields_ = [
        ("nDims", C.c_int), ("order", C.c_int), ("nVars", C.c_int), ("nEles", C.c_int),
        ("nSpts", C.c_int), ("nFpts", C.c_int), ("viscous", C.c_int), ("eles_per_block", C.c_int),
        ("gamma", C.c_double), ("prandtl", C.c_double), ("mu", C.c_double), ("rt", C.c_double),
        ("c_sth", C.c_double), ("fix_vis", C.c_int), ("rus_k", C.c_double), ("ldg_b", C.c_double),
        ("ldg_tau", C.c_double), ("ops1d", C.c_void_p), ("metric", C.c_void_p), ("fgeo", C.c_void_p),
        ("nbr", C.c_void_p), ("Utr", C.c_void_p), ("Gtr", C.c_void_p), ("Ftr", C.c_void_p),
        ("fconv_type", C.c_int),
    ]


class ZfrPhys(C.Structure):
    """``zfr_phys_t``: physical parameters of the pointwise fluxes."""

    _fields_ = [("gamma", C.c_double), ("prandtl", C.c_double), ("mu", C.c_double), ("rt", C.c_double),
                ("c_sth", C.c_double), ("fix_vis", C.c_int), ("rus_k", C.c_double), ("ldg_b", C.c_double),
                ("ldg_tau", C.c_double), ("fconv_type", C.c_int)]


# name -> (restype, argtypes); every symbol declared in include/zefr_amd.h
_P = C.c_void_p
_I = C.c_int
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)
_CP = C.c_char_p
SIGNATURES = {
    "zefr_last_error": (C.c_char_p, []),
    "zefr_read_input_file": (_I, [_CP, C.POINTER(ZefrInput)]),
    "zefr_initialize_inputs": (_I, [C.POINTER(ZefrInput)]),
    "zefr_create": (_P, []),
    "zefr_destroy": (None, [_P]),
    "zefr_read_input": (_I, [_P, _CP]),
    "zefr_get_input": (C.POINTER(ZefrInput), [_P]),
    "zefr_init_inputs": (_I, [_P]),
    "zefr_setup_solver": (_I, [_P]),
    "zefr_setup_host": (_I, [_P]),
    "zefr_do_step": (_I, [_P]),
    "zefr_do_rk_stage": (_I, [_P, _I, _I]),
    "zefr_do_rk_stage_start": (_I, [_P, _I, _I]),
    "zefr_do_rk_stage_mid": (_I, [_P, _I, _I]),
    "zefr_do_rk_stage_finish": (_I, [_P, _I, _I]),
    "zefr_write_residual": (_I, [_P, _DP]),
    "zefr_compute_residual": (_I, [_P, C.c_uint]),
    "zefr_compute_residual_start": (_I, [_P, C.c_uint]),
    "zefr_compute_residual_mid": (_I, [_P, C.c_uint]),
    "zefr_compute_residual_finish": (_I, [_P, C.c_uint]),
    "zefr_update": (_I, [_P]),
    "zefr_synchronize": (_I, [_P]),
    "zefr_get_u_spts_d": (_P, [_P, _IP, _IP, _IP, _I]),
    "zefr_get_du_spts_d": (_P, [_P, _IP, _IP, _IP, _IP, _I]),
    "zefr_get_u_spts": (_DP, [_P, _IP, _IP, _IP, _I]),
    "zefr_get_du_spts": (_DP, [_P, _IP, _IP, _IP, _IP, _I]),
    "zefr_get_u_spt": (C.c_double, [_P, _I, _I, _I]),
    "zefr_get_grad_spt": (C.c_double, [_P, _I, _I, _I, _I]),
    "zefr_get_dims": (_I, [_P, _IP]),
    "zefr_copy_state": (_I, [_P, _I, _I, _DP, C.c_long, _I]),
    "zefr_get_fpt_maps": (_I, [_P, _IP, C.POINTER(C.c_byte)]),
    "zefr_get_host_array": (C.c_long, [_P, _CP, _DP, C.c_long]),
    "zefr_autotune": (_I, [_P, _I]),
    "zefr_set_tile": (_I, [_P, _I]),
    "zefr_set_tiles": (_I, [_P, _IP]),
    "zefr_get_tiles": (_I, [_P, _IP]),
    "zefr_set_step_mode": (_I, [_P, _I, _I]),
    "zefr_set_pmg_persistent": (_I, [_P, _I]),
    "zefr_time_residual_kernels": (_I, [_P, _I, _DP]),
    "zefr_residual_at": (_I, [_P, _P, _P, _P]),
    "zefr_residual_vjp": (_I, [_P, _P, _P, _P, _P]),
    "zefr_pmg_copy_state": (_I, [_P, _I, _I, _DP, C.c_long, _I]),
    "zefr_batch_create": (_P, [C.POINTER(_P), _I]),
    "zefr_batch_residual": (_I, [_P]),
    "zefr_batch_synchronize": (_I, [_P]),
    "zefr_batch_launches": (_I, [_P]),
    "zefr_batch_destroy": (None, [_P]),
    "zefr_write_solution": (_I, [_P]),
    "zefr_restart_solution": (_I, [_P]),
    "zefr_initialize": (_I, [_CP]),
    "zefr_get_zefr_object": (_P, []),
    "zefr_set_zefr_object": (None, [_P]),
    "zefr_finalize": (_I, []),
    "zefr_use_gpus": (_I, []),
    "zefr_get_q_spts": (_DP, [_IP, _IP, _IP, _I]),
    "zefr_get_dq_spts": (_DP, [_IP, _IP, _IP, _IP, _I]),
    "zefr_get_q_spts_d": (_P, [_IP, _IP, _IP, _I]),
    "zefr_get_dq_spts_d": (_P, [_IP, _IP, _IP, _IP, _I]),
    "zefr_get_q_spt": (C.c_double, [_I, _I, _I]),
    "zfr_extrapolate": (_I, [C.POINTER(EleDesc), _P, _P]),
    "zfr_extrapolate_partner": (_I, [C.POINTER(EleDesc), _P, _P]),
    "zfr_gradient": (_I, [C.POINTER(EleDesc), _P, _P, _P]),
    "zfr_divergence": (_I, [C.POINTER(EleDesc), _P, _P, _P]),
    "zfr_gradient_flux": (_I, [C.POINTER(EleDesc), _P, _P, _P]),
    "zfr_gradient_flux_traced": (_I, [C.POINTER(EleDesc), _P, _P, _P]),
    "zfr_face_correction": (_I, [C.POINTER(EleDesc), _P, _P]),
    "zfr_rk_stage": (_I, [C.POINTER(EleDesc), _P, _P, _P, C.c_double, C.c_double, _P]),
    "zfr_rk_final": (_I, [C.POINTER(EleDesc), _P, _P, _P, _DP, _I, C.c_double, _P]),
    "zfr_residual_norm": (_I, [C.POINTER(EleDesc), _P, _I, _P, _P, _P]),
    "zfr_residual_norm_static": (_I, [C.POINTER(EleDesc), _P, _I, _P, _P]),
    "zfr_residual_vjp": (_I, [C.POINTER(EleDesc), _P, _P, _P, _P, _P, _P]),
    "zfr_face_correction_rk": (_I, [C.POINTER(EleDesc), _P, _P, _P]),
    "zfr_divergence_rk": (_I, [C.POINTER(EleDesc), _P, _P, _P, _P]),
    "zfr_stream_copy": (_I, [_P, _P, C.c_long, _P]),
    "zfr_multi_create": (_I, [_P, C.POINTER(_P), C.POINTER(_P), _I, C.POINTER(_P)]),
    "zfr_multi_residual": (_I, [_P, _P]),
    "zfr_multi_launches": (_I, [_P]),
    "zfr_multi_destroy": (None, [_P]),
    "zfr_persist_record_begin": (_I, []),
    "zfr_persist_record_end": (_I, [C.POINTER(_P)]),
    "zfr_persist_launch": (_I, [_P, _P]),
    "zfr_persist_info": (_I, [_P, _IP, _IP]),
    "zfr_persist_set_workgroups": (_I, [_P, _I]),
    "zfr_persist_destroy": (None, [_P]),
    "zfr_rk_stage_source": (_I, [C.POINTER(EleDesc), _P, _P, _P, _P, C.c_double, C.c_double, _P]),
    "zfr_rk_final_source": (_I, [C.POINTER(EleDesc), _P, _P, _P, _P, _DP, _I, C.c_double, _P]),
    "zfr_pmg_transfer": (_I, [_I, _I, C.c_long, _I, _I, _P, C.c_double, _P, C.c_double, _P, _P]),
    "zfr_axpby": (_I, [C.c_long, C.c_double, _P, C.c_double, _P, _P, _P]),
    "zfr_host_point_flux": (_I, [_I, C.POINTER(ZfrPhys), _DP, _DP, _DP]),
    "zfr_host_point_flux_vjp": (_I, [_I, C.POINTER(ZfrPhys), _DP, _DP, _DP, _DP, _DP]),
    "zfr_host_common_flux": (_I, [_I, C.POINTER(ZfrPhys), _I, _DP, _DP, _DP, _DP, _DP, C.c_double, _DP]),
    "zfr_host_common_flux_vjp": (_I, [_I, C.POINTER(ZfrPhys), _I, _DP, _DP, _DP, _DP, _DP, C.c_double, _DP,
                                      _DP, _DP, _DP, _DP]),
}


def build(verbose=False):
    """Compile libzefr_amd.so for gfx950 (hipcc cross-compiles without a GPU)."""
    jobs = str(min(16, os.cpu_count() or 4))
    out = None if verbose else subprocess.DEVNULL
    subprocess.run(["make", "-j", jobs, "-C", CSRC], check=True, stdout=out)


_lib = None


def _init_torch_first():
    """PyTorch-ROCm ships its own HIP runtime (torch/lib/libamdhip64.so, loaded into the
    global symbol scope).  Imported first, it also serves this library: the library's HIP
    calls bind to it, so there is one runtime in the process and torch's streams, events
    and allocations are valid here.  Loaded the other way round, two runtimes open the
    device and torch then reports no GPU."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        _init_torch_first()
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libzefr_amd.so is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                "or `make -C zefr_amd/csrc`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().zefr_last_error().decode()


def check(rc):
    if rc != 0:
        raise RuntimeError(last_error())
    return rc
