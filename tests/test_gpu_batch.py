"""Batched residuals (zefr_batch_*): the residuals of independent solvers as one multi-case
plan (zfr_multi_*: one launch per kernel and flux branch over all members, whatever their
order and element type) captured as one HIP graph.  The members cover Euler, the one-sided
LDG path, general beta and Roe on quads and hexes of several orders.  Each replay must reproduce every member's eager residual to
rounding (1e-15 relative: the multi-case kernels are separate compilations of the same
arithmetic, and the hex p=3 one-sided member differs from its eager residual in one ulp of an
entry); the graph is replayed several times with eager work and synchronisation in between."""

from __future__ import annotations

import numpy as np
import pytest

from common import TWO_PI, paired
from test_gpu_sweep import sweep_cases

pytestmark = pytest.mark.gpu


def _members():
    cases = sweep_cases(count=10, seed=77)
    cases += [
        dict(name="batch_h_ns_p3", n=(3, 3, 3), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=3, viscous=True,
             state="tgv3d"),
        dict(name="batch_q_ns_central", n=(5, 4), lo=(0.0, 0.0), hi=(1.0, 1.0), order=2, viscous=True,
             state="smooth", extra={"ldg_b": 0.0}),
        dict(name="batch_h_roe", n=(2, 2, 2), lo=(0.0,) * 3, hi=(TWO_PI,) * 3, order=2, viscous=True,
             state="tgv3d", extra={"fconv_type": "Roe"}),
        # quads and hexes in the same groups: a quad one-sided member beside the hex p=3 one,
        # a hex Euler member beside the quad Euler ones
        dict(name="batch_q_ns_p4", n=(6, 5), lo=(0.0, 0.0), hi=(TWO_PI, TWO_PI), order=4, viscous=True,
             state="tgv2d"),
        dict(name="batch_h_euler_p2", n=(3, 2, 2), lo=(0.0,) * 3, hi=(1.0,) * 3, order=2, viscous=False,
             state="noisy:9"),
    ]
    return cases


def test_batch_graph_matches_eager_residuals():
    import torch
    from zefr_amd.zefr import Batch
    zs, refs, names = [], [], []
    for case in _members():
        names.append(case["name"])
        print("[batch] member", case["name"], flush=True)
        z, _, _ = paired(case)
        z.compute_residual(0)
        refs.append(z.state("divF", 0, layout="device").copy())
        zs.append(z)
    b = Batch(zs)
    # one launch per kernel and flux branch over every order and element type: the Euler
    # divergence (quads p 1..5 and a hex), the Rusanov one-sided kernel (quad p=4, hex p=3),
    # the Roe one-sided kernel, the face correction of all three one-sided members, and the
    # general-beta member's extrapolate / gradient / divergence
    assert b.launches == 7, b.launches
    for rep in range(3):
        for z in zs:  # clear every residual eagerly between replays
            d = z.dims()
            zero = np.zeros((d["nEles"], d["nVars"], d["nSpts"]))
            z._copy(2, 0, zero, True)
        x = torch.randn(1 << 16, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert torch.isfinite(x.sum())
        b.residual()
        bad = []
        for case, z, ref in zip(names, zs, refs):
            got = z.state("divF", 0, layout="device")
            err = float(np.abs(got - ref).max() / np.abs(ref).max())
            if not err <= 1e-15:
                bad.append((case, err))
        assert not bad, (rep, bad)
    del b
