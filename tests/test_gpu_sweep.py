"""The 64 independent periodic quad cases of the batched configuration: seeded sizes
n x n with n in 8..64, orders 1..5, Euler with Rusanov, each on a seeded perturbed freestream
(rho = 1, u = 1, v = 0, p = 1 with 5% Gaussian point noise).  One residual of each against
the oracle, to the residual bar of the parity tests."""

from __future__ import annotations

import numpy as np
import pytest

from common import paired, residual_parity

pytestmark = pytest.mark.gpu


def sweep_cases(count=64, seed=2024):
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(count):
        n = int(rng.integers(8, 65))
        p = int(rng.integers(1, 6))
        cases.append(dict(name=f"sweep{i:02d}_n{n}_p{p}", n=(n, n), lo=(0.0, 0.0), hi=(1.0, 1.0), order=p,
                          viscous=False, state=f"noisy:{1000 + i}", dt=1e-4))
    return cases


def test_sweep_residuals_vs_oracle():
    worst = 0.0
    for case in sweep_cases():
        print("[sweep]", case["name"], flush=True)
        z, o, _ = paired(case)
        z.compute_residual(0)
        o.compute_residual(0)
        err, kappa, ok = residual_parity(z.state("divF", 0), o)
        assert ok, f"{case['name']}: relative L-inf {err:.3e} (cancellation factor {kappa:.1e})"
        worst = max(worst, err)
        del z, o
    print(f"[sweep] worst relative L-inf {worst:.3e}", flush=True)


def test_sweep_batch_graph_vs_oracle():
    """The same 64 cases through the batch graph (zefr.Batch: the members' residuals as a
    multi-case plan, one launch for all five orders): every member against its oracle."""
    from zefr_amd import zefr
    zs, os_ = [], []
    for case in sweep_cases():
        z, o, _ = paired(case)
        o.compute_residual(0)
        zs.append(z)
        os_.append(o)
    b = zefr.Batch(zs)
    print(f"[sweep] batch of {len(zs)} cases in {b.launches} launch(es)", flush=True)
    assert b.launches == 1
    b.residual()
    for case, z, o in zip(sweep_cases(), zs, os_):
        err, kappa, ok = residual_parity(z.state("divF", 0), o)
        assert ok, f"batch {case['name']}: relative L-inf {err:.3e} (cancellation factor {kappa:.1e})"
