"""Time-step execution modes on the GPU: the RK update fused into each residual's last
kernel and the whole step replayed from a captured HIP graph must give exactly the
results of the plain launch sequence (FRSolver::step_RK, src/solver.cpp:2865-3097), and
match the oracle."""

from __future__ import annotations

import numpy as np
import pytest

from common import TWO_PI, paired

pytestmark = pytest.mark.gpu

CASES = [
    dict(name="step_h_ns_p3", n=(3, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True, state="tgv3d",
         dt=2e-3),
    dict(name="step_q_euler_p3", n=(6, 5), lo=(-5.0, -5.0), hi=(5.0, 5.0), order=3, viscous=False, state="vortex"),
    dict(name="step_q_ns_general_beta", n=(4, 4), lo=(0.0, 0.0), hi=(1.0, 1.0), order=2, viscous=True,
         state="smooth", extra={"ldg_b": 0.2}),
    dict(name="step_q_ns_forward_euler", n=(4, 3), lo=(0.0, 0.0), hi=(1.0, 1.0), order=2, viscous=True,
         state="smooth", dt_scheme="Euler"),
]


def run(case, fuse, graph, steps):
    z, o, _ = paired(case)
    z.set_step_mode(fuse, graph)
    for _ in range(steps):
        z.do_step()
    return z, o, z.state("U")


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_fused_rk_update_is_bit_identical(case):
    """The fused step keeps its starting state in U and its stage states in U_ini until the
    final combination: U, U_ini (the step's starting state afterwards, for several stages)
    and every stage's residual must still equal the plain sequence's."""
    print(f"[step] {case['name']}", flush=True)
    zp, o, plain = run(case, False, False, 4)
    zf, _, fused = run(case, True, False, 4)
    np.testing.assert_array_equal(fused, plain)
    if zp.get_input().nStages > 1:  # a one-stage step does not write U_ini
        np.testing.assert_array_equal(zf.state("U_ini"), zp.state("U_ini"))
    for s in range(zp.get_input().nStages):
        np.testing.assert_array_equal(zf.state("divF", s), zp.state("divF", s))
    o.step_rk(4)
    ref = o.U()
    assert np.abs(fused - ref).max() <= 1e-10 * np.abs(ref).max()


@pytest.mark.parametrize("case", CASES[:2], ids=lambda c: c["name"])
@pytest.mark.parametrize("fuse", [True, False])
def test_graph_replay_is_bit_identical(case, fuse):
    """Several replays with eager work and host reads in between; a dt change recaptures."""
    print(f"[graph] {case['name']} fuse={fuse}", flush=True)
    z_e, _, _ = paired(case)
    z_g, _, _ = paired(case)
    z_e.set_step_mode(fuse, False)
    z_g.set_step_mode(fuse, True)
    for rep in range(3):
        z_e.do_step()
        z_g.do_step()
        z_g.compute_residual(0)  # eager work on the same stream between replays
        z_e.compute_residual(0)
        np.testing.assert_array_equal(z_g.state("U"), z_e.state("U"))
        np.testing.assert_array_equal(z_g.state("divF", 0), z_e.state("divF", 0))
    for z in (z_e, z_g):
        z.get_input().dt = z.get_input().dt * 0.5
    for rep in range(2):
        z_e.do_step()
        z_g.do_step()
    np.testing.assert_array_equal(z_g.state("U"), z_e.state("U"))
