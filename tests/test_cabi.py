"""The C ABI library: it loads without a GPU and exports every function that
include/zefr_amd.h declares, and the ctypes table covers exactly that set.  No compute
entry point is called here (that needs a HIP device)."""

from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

from zefr_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zefr_amd.h")


def declared_functions():
    with open(HEADER) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", " ", txt)
    names = re.findall(r"\b((?:zefr|zfr)_[A-Za-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["zefr_compute_residual", "zefr_compute_residual_start", "zefr_compute_residual_mid",
                 "zefr_compute_residual_finish", "zefr_do_step", "zefr_do_rk_stage", "zefr_write_residual",
                 "zefr_get_q_spts_d", "zefr_get_dq_spts_d", "zfr_extrapolate", "zfr_gradient_flux",
                 "zfr_face_correction", "zfr_rk_stage", "zfr_residual_norm", "zefr_last_error"]:
        assert must in names


def test_library_loads_and_exports_every_declared_symbol():
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert set(_lib.SIGNATURES) == set(declared_functions())


def test_exported_dynamic_symbols_are_c_names():
    """nm -D lists the declared names unmangled (extern "C")."""
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    for n in declared_functions():
        assert n in syms, n


def test_struct_sizes_match_header_layout():
    """zefr_input_t / zfr_ele_desc_t as seen by ctypes vs. a C compile of the header."""
    code =('#include "zefr_amd.h"\n#include <stdio.h>\n#include <stddef.h>\n'
            'int main(void){printf("%zu %zu %zu %zu\\n", sizeof(zefr_input_t), sizeof(zfr_ele_desc_t),'
            ' offsetof(zefr_input_t, meshBoundBC), offsetof(zfr_ele_desc_t, Ftr)); return 0;}\n')
    import tempfile
    d = tempfile.mkdtemp(prefix="zefr_sizes_")
    c = os.path.join(d, "sizes.c")
    exe = os.path.join(d, "sizes")
    with open(c, "w") as f:
        f.write(code)
    r = subprocess.run(["gcc", "-std=c11", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                        "-I", "/opt/rocm/include", c, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("the header needs hip_runtime_api.h from /opt/rocm/include: " + r.stderr[-200:])
    got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [C.sizeof(_lib.ZefrInput), C.sizeof(_lib.EleDesc),
                   _lib.ZefrInput.meshBoundBC.offset, _lib.EleDesc.Ftr.offset]


def test_errors_without_a_device_are_reported_not_raised_through_c():
    """Host-side validation fails with a status and a message, not a crash."""
    L = _lib.lib()
    z = L.zefr_create()
    assert z
    try:
        rc = L.zefr_read_input(z, b"/nonexistent/deck")
        assert rc != 0
        assert _lib.last_error()
        assert L.zefr_compute_residual(z, 0) != 0  # no solver set up yet
        assert _lib.last_error()
    finally:
        L.zefr_destroy(z)


def test_kernel_entries_reject_missing_buffers_before_any_launch():
    """A descriptor without the arrays a launch would touch is an error with a message,
    returned before any HIP call (so this runs without a GPU), not a device fault."""
    L = _lib.lib()
    d = _lib.EleDesc(nDims=3, order=3, nVars=5, nEles=8, nSpts=64, nFpts=96, viscous=1, ldg_b=0.5,
                     gamma=1.4, prandtl=0.72, mu=1e-3, fix_vis=1, ldg_tau=0.1)
    buf = C.c_void_p(0x1000)  # never dereferenced: every call below fails its checks first
    cases = [
        ("zfr_gradient_flux", (C.byref(d), buf, buf, None), "ops1d"),
        ("zfr_face_correction", (C.byref(d), buf, None), "ops1d"),
        ("zfr_divergence", (C.byref(d), buf, buf, None), "ops1d"),
    ]
    for name, args, msg in cases:
        assert getattr(L, name)(*args) != 0
        assert msg in _lib.last_error()
    d.ops1d = d.metric = d.fgeo = d.nbr = buf  # geometry present, traces missing
    assert L.zfr_gradient_flux(C.byref(d), buf, buf, None) != 0 and "Utr" in _lib.last_error()
    assert L.zfr_gradient_flux_traced(C.byref(d), buf, buf, None) != 0 and "Utr" in _lib.last_error()
    d.Utr = buf
    assert L.zfr_gradient_flux(C.byref(d), buf, buf, None) != 0 and "Ftr" in _lib.last_error()
    assert L.zfr_face_correction(C.byref(d), buf, None) != 0 and "Ftr" in _lib.last_error()
    d.ldg_b = 0.2
    assert L.zfr_divergence(C.byref(d), buf, buf, None) != 0 and "Gtr" in _lib.last_error()


def test_kernel_entries_reject_null_arguments():
    """A NULL descriptor or a NULL state / residual argument is an error with a message,
    returned before any HIP call."""
    L = _lib.lib()
    buf = C.c_void_p(0x1000)  # never dereferenced
    d = _lib.EleDesc(nDims=3, order=3, nVars=5, nEles=8, nSpts=64, nFpts=96, viscous=1, ldg_b=0.5,
                     gamma=1.4, prandtl=0.72, mu=1e-3, fix_vis=1, ldg_tau=0.1,
                     ops1d=buf, metric=buf, fgeo=buf, nbr=buf, Utr=buf, Ftr=buf, Gtr=buf)
    rk = C.c_void_p(0x2000)
    null_desc = [("zfr_extrapolate", (None, buf, None)), ("zfr_gradient", (None, buf, buf, None)),
                 ("zfr_divergence", (None, buf, buf, None)), ("zfr_gradient_flux", (None, buf, buf, None)),
                 ("zfr_gradient_flux_traced", (None, buf, buf, None)), ("zfr_face_correction", (None, buf, None)),
                 ("zfr_face_correction_rk", (None, buf, rk, None)), ("zfr_divergence_rk", (None, buf, buf, rk, None))]
    for name, args in null_desc:
        assert getattr(L, name)(*args) != 0, name
        assert "NULL descriptor" in _lib.last_error(), (name, _lib.last_error())
    null_arg = [("zfr_extrapolate", (C.byref(d), None, None), "U"),
                ("zfr_gradient", (C.byref(d), None, buf, None), "U"),
                ("zfr_divergence", (C.byref(d), None, buf, None), "U"),
                ("zfr_divergence", (C.byref(d), buf, None, None), "divF"),
                ("zfr_gradient_flux", (C.byref(d), None, buf, None), "U"),
                ("zfr_gradient_flux", (C.byref(d), buf, None, None), "divF"),
                ("zfr_gradient_flux_traced", (C.byref(d), buf, None, None), "divF"),
                ("zfr_face_correction", (C.byref(d), None, None), "divF"),
                ("zfr_face_correction_rk", (C.byref(d), None, rk, None), "divF"),
                ("zfr_divergence_rk", (C.byref(d), None, buf, rk, None), "U")]
    for name, args, what in null_arg:
        assert getattr(L, name)(*args) != 0, name
        assert f"{what} is NULL" in _lib.last_error(), (name, _lib.last_error())


def test_persistent_plan_recording_rules():
    """zfr_persist_record_begin / _end: one recording per thread, an empty recording is an
    error, a NULL plan discards, and while a thread records, launches other than the Euler
    divergence fail with a message before any HIP call (so this runs without a GPU)."""
    L = _lib.lib()
    assert L.zfr_persist_record_end(None) != 0 and "not recording" in _lib.last_error()
    assert L.zfr_persist_record_begin() == 0
    assert L.zfr_persist_record_begin() != 0 and "already recording" in _lib.last_error()
    p = C.c_void_p()
    assert L.zfr_persist_record_end(C.byref(p)) != 0 and "no launch was recorded" in _lib.last_error()
    assert not p.value
    buf = C.c_void_p(0x1000)  # never dereferenced
    d = _lib.EleDesc(nDims=3, order=3, nVars=5, nEles=8, nSpts=64, nFpts=96, viscous=1, ldg_b=0.5,
                     gamma=1.4, prandtl=0.72, mu=1e-3, fix_vis=1, ldg_tau=0.1,
                     ops1d=buf, metric=buf, fgeo=buf, nbr=buf, Utr=buf, Ftr=buf, Gtr=buf)
    assert L.zfr_persist_record_begin() == 0
    try:
        assert L.zfr_gradient_flux(C.byref(d), buf, buf, None) != 0
        assert "cannot be recorded" in _lib.last_error()
        assert L.zfr_divergence(C.byref(d), buf, buf, None) != 0  # viscous: not an Euler divergence
        assert "only Euler divergence launches" in _lib.last_error()
        assert L.zfr_axpby(4, 1.0, buf, 1.0, buf, buf, None) != 0
        assert "zfr_axpby cannot be recorded" in _lib.last_error()
        assert L.zfr_pmg_transfer(3, 5, 8, 4, 3, buf, 1.0, buf, 0.0, buf, None) != 0
        assert "zfr_pmg_transfer cannot be recorded" in _lib.last_error()
        d.viscous = 0
        d.order = 6  # a quad-only order: no hex kernel for it
        d.nSpts, d.nFpts = 343, 294
        assert L.zfr_divergence(C.byref(d), buf, buf, None) != 0
    finally:
        assert L.zfr_persist_record_end(None) == 0  # discarded
    assert L.zfr_persist_record_end(None) != 0  # no longer recording


class _RkArgs(C.Structure):
    """zfr_rk_args_t of include/zefr_amd.h (layout checked against a C compile below)."""
    _P = C.c_void_p
    _fields_ = [("U", _P), ("U_ini", _P), ("divF0", _P), ("mode", C.c_int), ("stage", C.c_int),
                ("nStages", C.c_int), ("alpha", C.c_double), ("dt", C.c_double), ("beta", C.c_double * 4),
                ("source", _P), ("trace_out", _P), ("U_ini_out", _P), ("corr_out", _P), ("corr_base", _P),
                ("source_out", _P), ("restrict_A", _P), ("restrict_src", _P), ("restrict_U", _P),
                ("restrict_divF", _P), ("restrict_n", C.c_int), ("prolong_n", C.c_int), ("prolong_A", _P),
                ("prolong_alpha", C.c_double), ("prolong_out", _P)]


def test_rk_source_out_rules():
    """source_out goes with mode 3 (source_out -= divF), or with mode 1 and a source (the
    stage forms the p-multigrid source term, then updates with it); anything else is an
    error before any HIP call."""
    code = ('#include "zefr_amd.h"\n#include <stdio.h>\n#include <stddef.h>\n'
            'int main(void){printf("%zu %zu %zu\\n", sizeof(zfr_rk_args_t), offsetof(zfr_rk_args_t, source_out),'
            ' offsetof(zfr_rk_args_t, prolong_out)); return 0;}\n')
    import tempfile
    tmp = tempfile.mkdtemp(prefix="zefr_rk_")
    c, exe = os.path.join(tmp, "rk.c"), os.path.join(tmp, "rk")
    with open(c, "w") as f:
        f.write(code)
    r = subprocess.run(["gcc", "-std=c11", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"),
                        "-I", "/opt/rocm/include", c, "-o", exe], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip("the header needs hip_runtime_api.h from /opt/rocm/include: " + r.stderr[-200:])
    got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [C.sizeof(_RkArgs), _RkArgs.source_out.offset, _RkArgs.prolong_out.offset]
    L = _lib.lib()
    buf = C.c_void_p(0x1000)  # never dereferenced: every call below fails its checks first
    d = _lib.EleDesc(nDims=2, order=3, nVars=4, nEles=8, nSpts=16, nFpts=16, viscous=0, gamma=1.4,
                     ops1d=buf, metric=buf, fgeo=buf, nbr=buf, Utr=buf)
    for mode, source, msg in [(1, None, "source_out is for mode 3"), (2, buf, "source_out is for mode 3"),
                              (4, buf, "source_out is for mode 3")]:
        a = _RkArgs(U=buf, U_ini=buf, divF0=buf, mode=mode, stage=0, nStages=4, source=source, source_out=buf)
        assert L.zfr_divergence_rk(C.byref(d), buf, buf, C.byref(a), None) != 0
        assert msg in _lib.last_error(), (mode, _lib.last_error())
    # mode 1 with a source passes the rule and fails a later check (the stage index), still
    # before any launch
    a = _RkArgs(U=buf, U_ini=buf, divF0=buf, mode=1, stage=7, nStages=4, source=buf, source_out=buf)
    assert L.zfr_divergence_rk(C.byref(d), buf, buf, C.byref(a), None) != 0
    assert "stage out of range" in _lib.last_error(), _lib.last_error()
