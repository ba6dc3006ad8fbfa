"""Uneven batches on the GPU: several instances of different mesh sizes, physics and NSTEP (water with a pressure
source, viscoelastic solids with a force, CPML frames and one with Stacey edges) advanced together as one union mesh,
one launch per kernel for all of them.  Each instance's seismograms, over its own NSTEP, against the FP32 oracle run
on that instance alone, within rel-L2 max(1e-5, 2 x the FP32 oracle's own distance from the FP64 one), and against
the same instance run alone on the GPU."""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config
from specfem2d_amd.driver import Simulation, UnevenBatch

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_uneven_batch_against_oracle_and_single_runs():
    cfgs = config.uneven_batch_instances(n=4, seed=3, size_range=((10, 30), (8, 20)), nstep_range=(600, 1000),
                                         nelem_pml=3, nrec=3)
    cfgs[3] = cfgs[3].replace(pml=False, stacey=True)
    print("case: uneven batch of " + ", ".join(f"{c.name} ({c.nstep} steps)" for c in cfgs), flush=True)
    u = UnevenBatch(cfgs)
    seis, t = u.forward()
    again, _ = u.forward()
    u.close()
    for k, c in enumerate(cfgs):
        assert seis[k].shape == (c.nrec, 2, c.nstep)
        single = Simulation(c)
        alone, _ = single.forward()
        single.close()
        ref = OracleSim(c, "sp").forward()
        gap = _rel(ref, OracleSim(c, "dp").forward())
        r, ra = _rel(seis[k], ref), _rel(seis[k], alone)
        print(f"{c.name}: rel-L2 vs FP32 oracle {r:.3e} (FP32 oracle vs FP64 {gap:.3e}), vs the instance alone on "
              f"the GPU {ra:.3e}")
        assert np.abs(ref).max() > 0
        assert r < max(TOL, 2 * gap) and ra < max(TOL, 2 * gap)
    # a second run from rest repeats the first up to the order of FP32 atomic additions
    worst = max(_rel(a, b) for a, b in zip(again, seis))
    print(f"second run vs first: {worst:.2e}; union loop {t['loop_ms']:.1f} ms")
    assert worst < 5e-6


def test_uneven_runs_in_nstep_groups():
    """The batch split into unions of similar NSTEP returns each instance's traces in input order, equal to the one
    union's up to the order of FP32 atomic additions."""
    from specfem2d_amd.driver import UnevenRuns
    cfgs = config.uneven_batch_instances(n=5, seed=4, size_range=((10, 30), (8, 20)), nstep_range=(300, 900),
                                         nelem_pml=3, nrec=3)
    print("case: 5 instances in NSTEP groups, NSTEP " + ", ".join(str(c.nstep) for c in cfgs), flush=True)
    runs = UnevenRuns(cfgs, max_groups=3, step_cost=0.0)  # these meshes are small: split them all the same
    split, _ = runs.forward()
    padded = 1 - runs.useful_updates() / runs.union_updates()
    groups = runs.groups
    runs.close()
    one = UnevenBatch(cfgs)
    whole, _ = one.forward()
    one.close()
    worst = max(_rel(a, b) for a, b in zip(split, whole))
    print(f"groups {groups}, padded fraction {padded:.3f}; grouped vs one union: {worst:.2e}")
    assert len(groups) > 1 and all(s.shape == (c.nrec, 2, c.nstep) for s, c in zip(split, cfgs))
    assert worst < 5e-6
