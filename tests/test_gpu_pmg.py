"""p-multigrid on the GPU against the oracle (PMGrid, src/multigrid.cpp:18-518).

* The transfer kernel (zfr_pmg_transfer) for every pair of orders 0..5 in 2D and 3D against
  the oracle's dense tensor operators (oppRes / oppPro as Quads/Hexas::setup_PMG build them).
* V-cycles on quads (the 4 -> 3 -> 2 -> 1 -> 0 level list of the 128^2 steady case, on a
  16^2 mesh and on the stated 128^2 mesh, with the seeded noisy freestream), on quads with
  skipped orders and rel_fac < 1,
  on hexes (Navier-Stokes), and one FMG cycle: the fine solution after each cycle against
  the oracle (1e-12 norm-wise per variable), and every coarse level's source term, stored
  solution, correction and state.
* The RK smoothing under a source term: fused and separate update kernels bit-identical.
* The V-cycle as one persistent (cooperative) launch: bit-identical to the eager cycle.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from common import EPS, TWO_PI, divergence_scale, make_mesh, paired, tmp

pytestmark = pytest.mark.gpu


def _pmg_case(name, n, lo, hi, order, viscous, state, levels, steps, cycle="V", rel_fac=1.0, fmg_vcycles=1,
              dt=1e-3):
    extra = {"p_multi": 1, "mg_levels": " ".join(map(str, levels)), "mg_steps": " ".join(map(str, steps)),
             "mg_cycle": cycle, "rel_fac": repr(rel_fac), "FMG_vcycles": fmg_vcycles}
    return dict(name=name, n=n, lo=lo, hi=hi, order=order, viscous=viscous, state=state, extra=extra, dt=dt,
                levels=levels, steps=steps, cycle=cycle, rel_fac=rel_fac, fmg_vcycles=fmg_vcycles)


def _paired_pmg(case):
    from oracle.oracle import OraclePMG
    z, o, U0 = paired(case)
    p = OraclePMG(o, tmp(case["name"] + ".msh"), case["levels"], case["steps"], rel_fac=case["rel_fac"],
                  cycle=case["cycle"], fmg_vcycles=case["fmg_vcycles"])
    return z, o, p


def _close(got, ref, scale, tol=1e-12):
    return np.abs(got - ref).max() <= tol * scale


def _check_cycle(z, o, p, case, it):
    ref = o.U()
    got = z.state("U")
    for v in range(ref.shape[1]):
        scale = np.abs(ref[:, v]).max()
        assert _close(got[:, v], ref[:, v], scale), (case["name"], it, v, np.abs(got[:, v] - ref[:, v]).max())
    Umax = np.abs(ref).max()
    # a source term is the difference of two residuals: its rounding follows the magnitude
    # of the terms summed into divF (the fine level's, from the oracle's last residual)
    fres = max(np.abs(o.divF(0)).max(), 64 * EPS / 1e-12 * np.abs(divergence_scale(o)).max())
    for lv in range(1, len(case["levels"])):
        for which, scale in (("U", Umax), ("solution", Umax), ("correction", Umax), ("source", fres)):
            a = z.pmg_state(lv, which).transpose(2, 1, 0)
            b = p.array(lv, which)
            assert _close(a, b, scale), (case["name"], it, lv, which, np.abs(a - b).max(), scale)


CASES = [
    _pmg_case("pmg_quad_p4_16", (16, 16), (0.0, 0.0), (1.0, 1.0), 4, False, "noisy:7", [4, 3, 2, 1, 0],
              [1, 1, 1, 1, 2], dt=2e-4),
    _pmg_case("pmg_quad_ns_skip", (6, 6), (0.0, 0.0), (TWO_PI, TWO_PI), 5, True, "smooth", [5, 3, 1],
              [2, 2, 3], rel_fac=0.8),
    _pmg_case("pmg_hex_ns", (3, 3, 3), (0.0,) * 3, (TWO_PI,) * 3, 3, True, "tgv3d", [3, 2, 0], [1, 2, 2]),
    _pmg_case("pmg_hex_euler", (3, 3, 3), (0.0,) * 3, (1.0,) * 3, 2, False, "noisy:3", [2, 1], [1, 3],
              dt=2e-4),
    # the quad half of the steady multigrid configuration at its stated size: 128^2 quads,
    # p = 4, levels 4 -> 3 -> 2 -> 1 -> 0, one smoothing step per level
    _pmg_case("pmg_quad_p4_128", (128, 128), (0.0, 0.0), (1.0, 1.0), 4, False, "noisy:7", [4, 3, 2, 1, 0],
              [1, 1, 1, 1, 1], dt=1e-4),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_v_cycles_vs_oracle(case):
    print("[pmg]", case["name"], flush=True)
    z, o, p = _paired_pmg(case)
    for it in range(3):
        z.do_step()
        p.cycle(1)
        _check_cycle(z, o, p, case, it)


def test_fmg_cycle_vs_oracle():
    case = _pmg_case("pmg_fmg", (5, 5), (0.0, 0.0), (TWO_PI, TWO_PI), 3, True, "smooth", [3, 2, 1],
                     [1, 1, 2], cycle="FMG", fmg_vcycles=2)
    print("[pmg]", case["name"], flush=True)
    z, o, p = _paired_pmg(case)
    # FMG starts from the coarsest solver's own state: the deck's initial condition in the
    # product (the oracle has none).  Both start from that state plus seeded noise, so that
    # no variable is identically zero.
    rng = np.random.default_rng(4)
    for lv in range(1, len(case["levels"])):
        Uc = z.pmg_state(lv, "U")
        Uc = Uc + 0.02 * rng.standard_normal(Uc.shape)
        z.set_pmg_state(lv, Uc)
        p.set_U(lv, Uc.transpose(2, 1, 0))
    z.do_step()
    p.cycle(1)
    assert z.get_input().mg_cycle == "V"  # FMG hands over to V-cycles, as the reference does
    _check_cycle(z, o, p, case, 0)
    z.do_step()
    p.cycle(1)
    _check_cycle(z, o, p, case, 1)


FUSE_CASES = [
    _pmg_case("pmg_fuse", (4, 4), (0.0, 0.0), (TWO_PI, TWO_PI), 3, True, "smooth", [3, 1, 0], [1, 2, 2]),
    _pmg_case("pmg_fuse_euler", (5, 4), (0.0, 0.0), (1.0, 1.0), 4, False, "noisy:5", [4, 2, 1, 0], [1, 2, 1, 3],
              dt=2e-4),
    _pmg_case("pmg_fuse_beta", (4, 4), (0.0, 0.0), (TWO_PI, TWO_PI), 3, True, "smooth", [3, 2, 0], [2, 1, 2],
              rel_fac=0.9),
]
FUSE_CASES[2]["extra"]["ldg_b"] = 0.3  # the general-beta LDG path
# a coarse level without smoothing steps, and forward-Euler steps: their source terms are
# formed by a residual launch of their own, not by a first step's stage 0
FUSE_CASES += [
    _pmg_case("pmg_fuse_zero_steps", (5, 4), (0.0, 0.0), (1.0, 1.0), 3, False, "noisy:5", [3, 2, 0], [1, 0, 2],
              dt=2e-4),
    dict(_pmg_case("pmg_fuse_forward_euler", (4, 5), (0.0, 0.0), (1.0, 1.0), 3, False, "noisy:5", [3, 1, 0],
                   [1, 2, 1], dt=5e-5), dt_scheme="Euler"),
]


@pytest.mark.parametrize("case", FUSE_CASES, ids=[c["name"] for c in FUSE_CASES])
def test_fused_and_separate_source_updates_agree(case):
    """Fused: the stored solution, the correction, the source term and the residual plus
    source term are written by the smoothing steps' and residuals' own kernels.  Separate:
    copies, axpby launches and the plain update kernels.  Every level bit-identical."""
    print("[pmg]", case["name"], flush=True)
    out = {}
    for fuse in (True, False):
        z, o, U0 = paired(case)
        z.set_step_mode(fuse, False)
        for _ in range(2):
            z.do_step()
        out[fuse] = [z.state("U")] + [z.pmg_state(lv, w) for lv in range(1, len(case["levels"]))
                                      for w in ("U", "source", "solution", "correction")]
    for a, b in zip(out[True], out[False]):
        assert np.array_equal(a, b)


def test_transfer_kernel_all_orders():
    import torch
    from oracle.oracle import Oracle, OraclePMG, make_phys
    from zefr_amd import _lib
    from zefr_amd.zefr import check
    L = _lib.lib()
    rng = np.random.default_rng(11)
    for D, n in ((2, (2, 2)), (3, (2, 2, 2))):
        mesh = make_mesh(f"pmg_ops_{D}.msh", n, (0.0,) * D, (1.0,) * D)
        o = Oracle(mesh, D, 5, make_phys())
        p = OraclePMG(o, mesh, [5, 4, 3, 2, 1, 0], [1] * 6)
        for lv in range(6):
            for kind in ("res", "pro"):
                if (kind == "res" and lv == 5) or (kind == "pro" and lv == 0):
                    continue
                A = p.operator(lv, kind)  # [nS_to][nS_from]
                pf, pt = 5 - lv, (4 - lv if kind == "res" else 6 - lv)
                A1 = _chain_1d(pf, pt)
                E, V = 37, D + 2
                x = rng.standard_normal((E, V, (pf + 1) ** D))
                y0 = rng.standard_normal((E, V, (pt + 1) ** D))
                for alpha, beta in ((1.0, 0.0), (0.7, 1.0)):
                    xd = torch.from_numpy(x).cuda()
                    yd = torch.from_numpy(y0).cuda()
                    Ad = torch.from_numpy(A1).cuda()
                    check(L.zfr_pmg_transfer(D, V, E, pf + 1, pt + 1, C.c_void_p(Ad.data_ptr()), alpha,
                                             C.c_void_p(xd.data_ptr()), beta, C.c_void_p(yd.data_ptr()), None))
                    torch.cuda.synchronize()
                    ref = alpha * np.einsum("ts,evs->evt", A, x) + beta * y0
                    got = yd.cpu().numpy()
                    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), (D, lv, kind, alpha)


def _chain_1d(p, q):
    # single-order steps between Gauss-Legendre sets, chained in the reference's association
    def step(nf, nt):
        xf, _ = np.polynomial.legendre.leggauss(nf)
        xt, _ = np.polynomial.legendre.leggauss(nt)
        A = np.zeros((nt, nf))
        for j in range(nf):
            c = np.zeros(nf)
            c[j] = 1.0
            coef = np.polynomial.polynomial.polyfit(xf, c, nf - 1)
            A[:, j] = np.polynomial.polynomial.polyval(xt, coef)
        return A
    if q < p:
        R = step(q + 2, q + 1)
        for P in range(q + 1, p):
            R = R @ step(P + 2, P + 1)
        return R
    Pm = step(q, q + 1)
    for P in range(q - 1, p, -1):
        Pm = Pm @ step(P, P + 1)
    return Pm


def test_axpby_kernel():
    import torch
    from zefr_amd import _lib
    from zefr_amd.zefr import check
    L = _lib.lib()
    rng = np.random.default_rng(2)
    n = 100003
    x = torch.from_numpy(rng.standard_normal(n)).cuda()
    y = torch.from_numpy(rng.standard_normal(n)).cuda()
    z = torch.empty_like(x)
    check(L.zfr_axpby(n, 1.0, C.c_void_p(x.data_ptr()), -1.0, C.c_void_p(y.data_ptr()), C.c_void_p(z.data_ptr()),
                      None))
    torch.cuda.synchronize()
    assert torch.equal(z, x - y)
    # in place, z aliasing x
    x0 = x.clone()
    check(L.zfr_axpby(n, 1.0, C.c_void_p(x.data_ptr()), 1.0, C.c_void_p(y.data_ptr()), C.c_void_p(x.data_ptr()),
                      None))
    torch.cuda.synchronize()
    assert torch.equal(x, x0 + y)


def test_vcycle_graph_matches_eager():
    """A V-cycle captured as one HIP graph and replayed gives the eager cycle's results bit
    for bit, on every level, with eager work and synchronisation between the replays."""
    case = _pmg_case("pmg_graph", (6, 6), (0.0, 0.0), (TWO_PI, TWO_PI), 4, True, "smooth", [4, 2, 1, 0],
                     [1, 1, 2, 2])
    zg, _, _ = paired(case)
    ze, _, _ = paired(case)
    zg.set_step_mode(True, True)
    for it in range(3):
        zg.do_step()
        ze.do_step()
        assert np.array_equal(zg.state("U"), ze.state("U")), it
        for lv in range(1, 4):
            for which in ("U", "source", "correction"):
                assert np.array_equal(zg.pmg_state(lv, which), ze.pmg_state(lv, which)), (it, lv, which)
        assert zg.get_input().iter == ze.get_input().iter


PERSIST_CASES = [
    _pmg_case("pmg_persist_quad", (12, 10), (0.0, 0.0), (1.0, 1.0), 4, False, "noisy:7", [4, 3, 2, 1, 0],
              [1, 1, 1, 1, 2], dt=2e-4),
    _pmg_case("pmg_persist_hex_roe", (3, 3, 3), (0.0,) * 3, (1.0,) * 3, 3, False, "noisy:3", [3, 1, 0], [2, 1, 2],
              dt=2e-4, rel_fac=0.9),
]
PERSIST_CASES[1]["extra"]["fconv_type"] = "Roe"


@pytest.mark.parametrize("case", PERSIST_CASES, ids=[c["name"] for c in PERSIST_CASES])
def test_persistent_vcycle_matches_eager(case):
    """The V-cycle recorded once and replayed as one cooperative launch (every level's
    residuals, RK stages, source terms and fused transfers, a grid-wide barrier between
    them) gives the eager cycle's results bit for bit on every level, cycle after cycle,
    and advances the iteration counters as the eager cycle does; it also matches the
    oracle."""
    print("[pmg]", case["name"], flush=True)
    zp, o, p = _paired_pmg(case)
    ze, _, _ = paired(dict(case, name=case["name"] + "_eager"))
    zp.set_pmg_persistent(True)
    for it in range(3):
        zp.do_step()
        ze.do_step()
        p.cycle(1)
        assert np.array_equal(zp.state("U"), ze.state("U")), it
        for lv in range(1, len(case["levels"])):
            for which in ("U", "source", "solution", "correction"):
                assert np.array_equal(zp.pmg_state(lv, which), ze.pmg_state(lv, which)), (it, lv, which)
        assert zp.get_input().iter == ze.get_input().iter
        _check_cycle(zp, o, p, case, it)


def test_persistent_vcycle_refuses_what_it_cannot_hold():
    """A viscous V-cycle has launches other than the Euler divergence: do_step fails with
    the reason before anything runs, and the eager cycle still works afterwards."""
    case = _pmg_case("pmg_persist_ns", (4, 4), (0.0, 0.0), (TWO_PI, TWO_PI), 3, True, "smooth", [3, 1, 0],
                     [1, 1, 1])
    z, _, _ = paired(case)
    z.set_pmg_persistent(True)
    with pytest.raises(RuntimeError, match="inviscid runs only"):
        z.do_step()
    z.set_pmg_persistent(False)
    z.do_step()
    assert np.isfinite(z.state("U")).all()
