"""Host-side setup of the product (no GPU): input deck, Gmsh reader, periodic coupling,
global flux-point numbering, operators, metrics and face geometry, compared with the
oracle on identical meshes.

The integer maps (fpt2gfpt, slot) must match bit-exactly; operators and geometry to a
few ulps (the product builds dense operators from its 1D tables, the oracle from dense
Lagrange products).  Runs through the C ABI (libzefr_amd.so) with setup_host(), which
touches no HIP device.
"""

from __future__ import annotations

import math
import os
import shutil

import numpy as np
import pytest

from common import TWO_PI, make_mesh, tmp, write_deck, product
from oracle.oracle import Oracle, make_phys

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def _pair(mesh, D, order, viscous=True, bounds=(("periodic", "periodic"), ("fluid", "fluid")),
          oracle_mesh=None, name="hs"):
    deck = write_deck(tmp(name + ".deck"), mesh, D, order, viscous, bounds=bounds)
    z = product(deck, device=False)
    inp = z.get_input()
    o = Oracle(oracle_mesh or mesh, D, order, make_phys(viscous=viscous, mu=inp.mu))
    return z, o


def _compare(z, o, D, tol=4e-15):
    d = z.dims()
    assert (d["nEles"], d["nSpts"], d["nFpts"], d["nVars"], d["nGfpts"]) == \
        (o.nEles, o.nSpts, o.nFpts, o.nVars, o.nGfpts)
    g, s = z.fpt_maps()
    go, so = o.maps()
    np.testing.assert_array_equal(g, go)
    np.testing.assert_array_equal(s, so)
    for name in ["oppE", "oppD", "oppD_fpts", "oppDiv", "oppDiv_fpts", "weights_spts"]:
        a, b = z.host_array(name), o.array(name)
        assert a.shape == b.shape, name
        assert np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max()), name
    for name in ["coord_spts", "coord_fpts", "jaco_det_spts", "inv_jaco_spts", "inv_jaco_fpts",
                 "norm", "dA"]:
        a, b = z.host_array(name), o.array(name)
        assert a.shape == b.shape, name
        assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max()), name


CASES = [
    dict(D=2, n=(4, 4), lo=(-5.0, -5.0), hi=(5.0, 5.0), order=3, perturb=0.0),
    dict(D=2, n=(5, 3), lo=(0.0, 0.0), hi=(1.0, 2.0), order=2, perturb=0.04),
    dict(D=2, n=(6, 6), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=4, perturb=0.0),
    dict(D=2, n=(3, 3), lo=(0.0, 0.0), hi=(1.0, 1.0), order=0, perturb=0.0),
    dict(D=2, n=(3, 4), lo=(0.0, 0.0), hi=(1.0, 1.5), order=7, perturb=0.03),
    dict(D=2, n=(3, 3), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=9, perturb=0.0),
    dict(D=3, n=(3, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, perturb=0.0),
    dict(D=3, n=(2, 3, 4), lo=(0.0, -1.0, 0.5), hi=(1.0, 1.0, 2.0), order=2, perturb=0.03),
    dict(D=3, n=(2, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=4, perturb=0.02),
    dict(D=3, n=(3, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=1, perturb=0.0),
    dict(D=3, n=(2, 2, 3), lo=(0.0,) * 3, hi=(1.0,) * 3, order=6, perturb=0.02),
    dict(D=3, n=(2, 2, 2), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=7, perturb=0.0),
]


@pytest.mark.parametrize("c", CASES, ids=lambda c: f"{c['D']}d_{'x'.join(map(str, c['n']))}_p{c['order']}"
                                                   f"{'_curved' if c['perturb'] else ''}")
def test_host_setup_matches_oracle(c):
    mesh = make_mesh(f"hs_{c['D']}_{c['order']}_{c['perturb']}.msh", c["n"], c["lo"], c["hi"], c["perturb"])
    z, o = _pair(mesh, c["D"], c["order"])
    _compare(z, o, c["D"])


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_reference_hexbox_mesh(order):
    """meshes/HexBox.msh of the reference: Gmsh's own node numbering (not a structured
    order), PERIODIC boundary quads and a FLUID volume."""
    mesh = os.path.join(GOLDEN, "HexBox.msh")
    z, o = _pair(mesh, 3, order, name=f"hexbox{order}")
    _compare(z, o, 3)
    assert z.dims()["nEles"] == 8


def test_reference_quadbox_mesh_with_bound_mapping():
    """meshes/QuadBox_4x4.msh names its boundary FARFIELD; the deck maps it to periodic
    (mesh_bound farfield periodic).  The oracle reads a copy whose group is renamed."""
    src = os.path.join(GOLDEN, "QuadBox_4x4.msh")
    ren = tmp("QuadBox_4x4_periodic.msh")
    with open(src) as f, open(ren, "w") as g:
        g.write(f.read().replace('"FARFIELD"', '"PERIODIC"'))
    z, o = _pair(src, 2, 3, bounds=(("farfield", "periodic"), ("fluid", "fluid")), oracle_mesh=ren,
                 name="quadbox")
    _compare(z, o, 2)
    assert z.dims()["nEles"] == 16


def test_operators_are_sum_factorised_tables():
    """The dense operators exported by the product are rebuilt from the 1D tables the
    kernels read (ops1d): lE[2][N], Dss[N][N], Dsf[N][2]."""
    mesh = make_mesh("ops.msh", (2, 2, 2), (0.0,) * 3, (1.0,) * 3)
    z, o = _pair(mesh, 3, 3, name="ops")
    N = 4
    ops = z.host_array("ops1d")
    assert ops.size >= 2 * N + N * N + 2 * N
    lE = ops[:2 * N].reshape(2, N)
    np.testing.assert_allclose(lE.sum(axis=1), 1.0, rtol=1e-15)  # partition of unity
    Dss = ops[2 * N:2 * N + N * N].reshape(N, N)
    Dsf = ops[2 * N + N * N:2 * N + N * N + 2 * N].reshape(N, 2)
    np.testing.assert_allclose(Dss.sum(axis=1) + Dsf.sum(axis=1), 0.0, atol=1e-13)  # d/dx 1 = 0


# ---------------------------------------------------------------- input deck
def _deck_lines(path, lines):
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


BASE = ["nDims 2", "meshfile {mesh}", "mesh_bound periodic periodic", "mesh_bound fluid fluid",
        "order 2", "viscous 1", "n_steps 3", "dt_scheme RK44", "dt 1e-3"]


def _read(extra, name):
    from zefr_amd import zefr
    mesh = make_mesh("deck.msh", (2, 2), (0.0, 0.0), (1.0, 1.0))
    path = _deck_lines(tmp(name), [l.format(mesh=mesh) for l in BASE] + extra)
    z = zefr.Zefr()
    z.read_input(path)
    return z.get_input()


def test_deck_disable_nondim_closed_form():
    """src/input.cpp:431-478 with disable_nondim: U = Ma sqrt(gamma P / rho),
    mu = rho U L / Re, T = P / (rho R)."""
    inp = _read(["disable_nondim 1", "mach_fs 0.2", "Re_fs 400", "rho_fs 1.3", "P_fs 2.0", "L_fs 0.5",
                 "gamma 1.4", "R 286.9"], "d1.deck")
    U = 0.2 * math.sqrt(1.4 * 2.0 / 1.3)
    assert math.isclose(inp.v_mag_fs, U, rel_tol=1e-15)
    assert math.isclose(inp.mu, 1.3 * U * 0.5 / 400, rel_tol=1e-15)
    assert math.isclose(inp.T_fs, 2.0 / (1.3 * 286.9), rel_tol=1e-15)
    assert inp.fix_vis == 1
    assert inp.nStages == 4


def test_deck_nondim_closed_form():
    """With nondimensionalisation (src/input.cpp:478-520) and constant viscosity the
    freestream becomes rho = 1, |V| = 1, T = 1, P = 1 / (gamma Ma^2), mu = L / Re."""
    inp = _read(["disable_nondim 0", "fix_vis 1", "mach_fs 0.3", "Re_fs 1000", "rho_fs 1.1",
                 "P_fs 1.0e5", "L_fs 2.0", "T_fs 300", "gamma 1.4", "R 286.9", "T_gas 291.15",
                 "S 120"], "d2.deck")
    assert math.isclose(inp.rho_fs, 1.0, rel_tol=1e-14)
    assert math.isclose(inp.P_fs, 1.0 / (1.4 * 0.09), rel_tol=1e-13)
    assert math.isclose(inp.mu, 2.0 / 1000, rel_tol=1e-13)
    assert math.isclose(inp.T_fs, 1.0) and math.isclose(inp.v_mag_fs, 1.0)
    assert math.isclose(inp.c_sth, 120 / 291.15, rel_tol=1e-15)
    V = 0.3 * math.sqrt(1.4 * 286.9 * 300)
    assert math.isclose(inp.rt, 291.15 * 286.9 / V ** 2, rel_tol=1e-14)


@pytest.mark.parametrize("extra,msg", [
    (["equation AdvDiff"], "AdvDiff"),
    (["dt_type 1"], "dt_type"),
    (["motion 1"], "moving"),
    (["fconv_type HLLC"], "Convective interface flux"),  # the reference's message
])
def test_deck_rejects_unsupported_options(extra, msg):
    with pytest.raises(RuntimeError, match=msg):
        _read(extra, "bad.deck")


def test_triangle_mesh_is_rejected_with_the_reason():
    """Two triangles splitting the unit square: the setup refuses them and names the
    reference's missing triangle volume divergence (src/elements.cpp:591-730)."""
    from zefr_amd import zefr
    mesh = _deck_lines(tmp("tri.msh"), [
        "$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$PhysicalNames", "2", '1 1 "PERIODIC"', '2 2 "FLUID"',
        "$EndPhysicalNames", "$Nodes", "4", "1 0 0 0", "2 1 0 0", "3 1 1 0", "4 0 1 0", "$EndNodes",
        "$Elements", "2", "1 2 2 2 2 1 2 3", "2 2 2 2 2 1 3 4", "$EndElements"])
    path = _deck_lines(tmp("tri.deck"), [l.format(mesh=mesh) for l in BASE])
    z = zefr.Zefr()
    z.read_input(path)
    with pytest.raises(RuntimeError, match="triangular elements.*volume divergence"):
        z.setup_host()


@pytest.mark.parametrize("D,order,ok", [(2, 9, True), (2, 10, False), (3, 8, True), (3, 9, False)])
def test_order_limits(D, order, ok):
    """Quads run to order 9 (the reference's Gauss-Legendre tables end at 10 points); hexes
    to order 8, with a message naming the limit."""
    from zefr_amd import zefr
    n = (2,) * D
    mesh = make_mesh(f"ordlim{D}.msh", n, (0.0,) * D, (1.0,) * D)
    lines = [l.format(mesh=mesh) for l in BASE if not l.startswith(("nDims", "order"))]
    path = _deck_lines(tmp(f"ordlim{D}_{order}.deck"), lines + [f"nDims {D}", f"order {order}"])
    z = zefr.Zefr()
    z.read_input(path)
    if ok:
        z.setup_host()
        assert z.dims()["nSpts"] == (order + 1) ** D
    else:
        with pytest.raises(RuntimeError, match="order must be in 0.." + ("9 for quads" if D == 2 else "8 for hexes")):
            z.setup_host()


def test_deck_rejects_missing_required_key():
    from zefr_amd import zefr
    path = _deck_lines(tmp("missing.deck"), ["nDims 2", "order 2"])
    z = zefr.Zefr()
    with pytest.raises(RuntimeError):
        z.read_input(path)


# ---------------------------------------------------------------- mesh writer
def test_mesh_writer_format_and_roundtrip():
    from zefr_amd.mesh import write_box
    path = tmp("fmt.msh")
    lo, hi = (0.1, -0.3, 0.0), (1.0 / 3.0, 0.7, math.pi)
    write_box(path, (2, 3, 1), lo, hi)
    with open(path) as f:
        txt = f.read().split("\n")
    assert txt[:3] == ["$MeshFormat", "2.2 0 8", "$EndMeshFormat"]
    i = txt.index("$Nodes")
    nn = int(txt[i + 1])
    assert nn == 3 * 4 * 2
    first = txt[i + 2].split()
    assert float(first[1]) == lo[0] and float(first[3]) == lo[2]
    last = txt[i + 1 + nn].split()
    assert float(last[1]) == hi[0] and float(last[2]) == hi[1] and float(last[3]) == hi[2]
    j = txt.index("$Elements")
    types = [int(l.split()[1]) for l in txt[j + 2:j + 2 + int(txt[j + 1])]]
    assert types.count(5) == 6                               # hexes
    assert types.count(3) == 2 * (2 * 3 + 3 * 1 + 2 * 1)     # boundary quads
    assert '"PERIODIC"' in "\n".join(txt) and '"FLUID"' in "\n".join(txt)


def test_setup_rejects_unknown_boundary_group():
    mesh = make_mesh("unk.msh", (2, 2), (0.0, 0.0), (1.0, 1.0))
    deck = write_deck(tmp("unk.deck"), mesh, 2, 2, False, bounds=(("fluid", "fluid"),))
    with pytest.raises(RuntimeError):
        product(deck, device=False)


def test_library_surface_outside_the_residual_path():
    """The reference module's remaining names exist: adapt_dt returns the supplied dt, the
    averaging calls are no-ops without tavg, write_wall_time prints the elapsed time, and
    the force, error and overset (TIOGA) calls raise with the reason."""
    from zefr_amd import zefr
    mesh = make_mesh("surface.msh", (2, 2), (0.0, 0.0), (1.0, 1.0))
    path = _deck_lines(tmp("surface.deck"), [l.format(mesh=mesh) for l in BASE])
    z = zefr.Zefr()
    z.read_input(path)
    assert z.adapt_dt() == z.get_input().dt == 1e-3
    assert z.update_averages() is None and z.write_averages() is None
    z.write_wall_time()
    with pytest.raises(NotImplementedError, match="wall boundaries"):
        z.write_forces()
    with pytest.raises(NotImplementedError, match="analytic"):
        z.write_error()
    for name in ("get_basic_geo_data", "get_extra_geo_data", "get_gpu_geo_data", "get_callback_funcs"):
        with pytest.raises(NotImplementedError, match="overset"):
            getattr(zefr, name)()
