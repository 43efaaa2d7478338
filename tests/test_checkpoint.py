"""PyFR-format solution files (.pyfrs) on the CPU: FRSolver::write_solution_pyfr and
restart_pyfr (src/solver.cpp:4275-4837) through the product's C ABI on a host-only solver.

The files are read back with tests/h5read.py, an HDF5 reader independent of the product's
writer, against the layout the reference's reader expects: scalar variable-length strings
"config", "stats", "mesh_uuid" (djb2 of the mesh file, src/funcs.cpp:1536-1555) and
"soln_<quad|hex>_p0" of shape [nSpts][nVars][nEles].  A restart reproduces the written
state bit for bit; a restart at another order equals the Lagrange interpolant of the
written points, evaluated here with numpy.
"""

from __future__ import annotations

import os

import numpy as np
import pytest

import h5read
from common import TWO_PI, make_mesh, tmp, write_deck

pytestmark = pytest.mark.skipif(not h5read.available(), reason="no HDF5 C library in this environment")


def _djb2(path):
    h = 5381
    for b in open(path, "rb").read():
        if b == 0:
            break
        c = b - 256 if b > 127 else b
        h = (h * 33 + c) % (1 << 64)
    return str(h)


def _solver(name, D, order, extra=None, n=None, ic=None, mesh=None):
    from zefr_amd import zefr
    n = n or ((3,) * D)
    if mesh is None:
        if D == 2:
            mesh = make_mesh(name + ".msh", n, (-5.0, -5.0), (5.0, 5.0))
        else:
            mesh = make_mesh(name + ".msh", n, (0.0,) * 3, (TWO_PI,) * 3)
    ic = ic or ({"ic_type": 1, "test_case": 1} if D == 2 else {"ic_type": 1, "test_case": 4})
    e = {"write_pyfr": 1, "output_prefix": "ckpt"}
    e.update(extra or {})
    deck = write_deck(tmp(name + ".deck"), mesh, D, order, D == 3, extra=e, **ic)
    z = zefr.Zefr()
    z.read_input(deck)
    z.setup_host()
    return z, mesh


def _state(z):
    d = z.dims()
    return z.host_array("U_init").reshape(d["nEles"], d["nVars"], d["nSpts"])


@pytest.mark.parametrize("D,order", [(2, 3), (3, 2)])
def test_pyfrs_layout(D, order, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    z, mesh = _solver(f"ck_layout{D}", D, order)
    z.write_solution()
    path = tmp_path / "ckpt" / "ckpt-0.pyfrs"
    assert path.exists()
    et = "quad" if D == 2 else "hex"
    with h5read.File(str(path)) as f:
        assert f.names() == ["config", "mesh_uuid", f"soln_{et}_p0", "stats"]
        uuid, var = f.string("mesh_uuid")
        assert var and uuid == _djb2(mesh)
        cfg, _ = f.string("config")
        stats, _ = f.string("stats")
        soln = f.doubles(f"soln_{et}_p0")
    assert f"system = {'navier-stokes' if D == 3 else 'euler'}" in cfg and f"order = {order}" in cfg
    assert f"[solver-elements-{et}]\nsoln-pts = gauss-legendre" in cfg
    assert "fields = rho,rhou,rhov," + ("rhoW," if D == 3 else "") + "E" in stats
    assert "tcurr = 0\n" in stats and "nsteps = 0\n" in stats
    U = _state(z)
    assert soln.shape == (U.shape[2], U.shape[1], U.shape[0])
    assert np.array_equal(soln, U.transpose(2, 1, 0))


def test_restart_round_trip(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    z, mesh = _solver("ck_rt", 3, 2)
    z.write_solution()
    z2, _ = _solver("ck_rt2", 3, 2, mesh=mesh, ic={"ic_type": 0},
                    extra={"restart": 1, "restart_type": 2, "restart_case": "ckpt", "restart_iter": 0})
    assert not np.array_equal(_state(z2), _state(z))
    z2.restart_solution()
    assert np.array_equal(_state(z2), _state(z))
    assert z2.get_input().iter == 0 and z2.get_input().time == 0.0


def _interp_matrix(n_from, n_to):
    xf, _ = np.polynomial.legendre.leggauss(n_from)
    xt, _ = np.polynomial.legendre.leggauss(n_to)
    A = np.zeros((n_to, n_from))
    for j in range(n_from):
        c = np.zeros(n_from)
        c[j] = 1.0
        A[:, j] = np.polynomial.polynomial.polyval(xt, np.polynomial.polynomial.polyfit(xf, c, n_from - 1))
    return A


@pytest.mark.parametrize("D,p_from,p_to", [(2, 2, 4), (2, 4, 3), (3, 1, 2)])
def test_restart_at_another_order(D, p_from, p_to, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    z, mesh = _solver(f"ck_ord{D}{p_from}", D, p_from)
    z.write_solution()
    z2, _ = _solver(f"ck_ord{D}{p_to}", D, p_to, mesh=mesh, ic={"ic_type": 0},
                    extra={"restart": 1, "restart_type": 2, "restart_case": "ckpt", "restart_iter": 0})
    z2.restart_solution()
    A = _interp_matrix(p_from + 1, p_to + 1)
    K = A
    for _ in range(D - 1):
        K = np.kron(A, K)  # points x fastest
    ref = np.einsum("ts,evs->evt", K, _state(z))
    assert np.abs(_state(z2) - ref).max() <= 1e-13 * np.abs(ref).max()


def test_restart_rejects_another_mesh(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    z, mesh = _solver("ck_m1", 2, 2)
    z.write_solution()
    other = make_mesh("ck_m2.msh", (3, 3), (-5.0, -5.0), (5.0, 5.1))
    z2, _ = _solver("ck_m2", 2, 2, mesh=other,
                    extra={"restart": 1, "restart_type": 2, "restart_case": "ckpt", "restart_iter": 0})
    with pytest.raises(RuntimeError, match="mesh_uuid"):
        z2.restart_solution()


@pytest.mark.parametrize("rtype,msg", [(1, "ParaView"), (0, "restart_type")])
def test_restart_type_must_be_pyfr(rtype, msg, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    z, _ = _solver("ck_rtype", 2, 2, extra={"restart": 1, "restart_type": rtype, "restart_case": "ckpt"})
    with pytest.raises(RuntimeError, match=msg):
        z.restart_solution()


def test_no_restart_flag_is_a_no_op(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    z, _ = _solver("ck_noop", 2, 2, extra={"write_pyfr": 0})
    before = _state(z).copy()
    z.restart_solution()
    z.write_solution()
    assert not os.path.exists(tmp_path / "ckpt")
    assert np.array_equal(_state(z), before)
