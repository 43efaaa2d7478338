"""Checkpoint and resume on the GPU: a run that writes a .pyfrs file mid-way and a second
solver restarted from that file continue bit-identically (FRSolver::write_solution_pyfr /
restart_pyfr, src/solver.cpp:4275-4837)."""

from __future__ import annotations

import numpy as np
import pytest

import h5read
from common import TWO_PI, make_mesh, tmp, write_deck

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not h5read.available(), reason="no HDF5 C library in this environment")]


@pytest.mark.parametrize("D,order", [(3, 3), (2, 4)])
def test_resume_is_bit_identical(D, order, tmp_path, monkeypatch):
    from zefr_amd import zefr
    monkeypatch.chdir(tmp_path)
    print(f"[ckpt] D={D} p={order}", flush=True)
    if D == 3:
        mesh = make_mesh("ck_gpu3.msh", (4, 4, 4), (0.0,) * 3, (TWO_PI,) * 3)
        ic = {"ic_type": 1, "test_case": 4}
    else:
        mesh = make_mesh("ck_gpu2.msh", (8, 8), (-5.0, -5.0), (5.0, 5.0))
        ic = {"ic_type": 1, "test_case": 1}
    base = {"write_pyfr": 1, "output_prefix": "run"}
    deck = write_deck(tmp(f"ck_gpu{D}.deck"), mesh, D, order, D == 3, dt=2e-4, extra=base, **ic)
    z = zefr.Zefr()
    z.read_input(deck)
    z.setup_solver()
    for _ in range(3):
        z.do_step()
    z.write_solution()
    assert (tmp_path / "run" / "run-3.pyfrs").exists()
    for _ in range(3):
        z.do_step()
    Ua = z.state("U", layout="device")
    deck2 = write_deck(tmp(f"ck_gpu{D}_r.deck"), mesh, D, order, D == 3, dt=2e-4,
                       extra=dict(base, restart=1, restart_type=2, restart_case="run", restart_iter=3), **ic)
    z2 = zefr.Zefr()
    z2.read_input(deck2)
    z2.setup_solver()
    z2.restart_solution()
    assert z2.get_input().iter == 3
    assert z2.get_input().time == pytest.approx(3 * 2e-4, rel=1e-15)
    for _ in range(3):
        z2.do_step()
    assert np.array_equal(z2.state("U", layout="device"), Ua)
