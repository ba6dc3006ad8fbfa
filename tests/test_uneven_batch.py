"""The host side of uneven batches (sf2d_sim_create_union): the union of several instances -- different mesh sizes,
fluid and viscoelastic solids, CPML and Stacey edges -- carries each instance's numbering, mass, source and receiver
set-up at its offsets.  CPU only."""
import numpy as np
import pytest

from specfem2d_amd import config
from specfem2d_amd.driver import Simulation, UnevenBatch


def _instances():
    cfgs = config.uneven_batch_instances(n=4, seed=3, size_range=((10, 30), (8, 20)), nstep_range=(200, 400),
                                         nelem_pml=3, nrec=3)
    cfgs[3] = cfgs[3].replace(pml=False, stacey=True)  # one viscoelastic instance with Stacey edges
    return cfgs


def test_union_concatenates_the_instances():
    cfgs = _instances()
    u = UnevenBatch(cfgs, device=False)
    nstep = max(c.nstep for c in cfgs)
    singles = [Simulation(c.replace(nstep=nstep), device=False) for c in cfgs]
    assert u.nspec == sum(s.nspec for s in singles) and u.nglob == sum(s.nglob for s in singles)
    assert u.nspec_acoustic == sum(s.nspec_acoustic for s in singles)
    ib, mass = u.ibool(), u.rmass_inverse()
    e0 = g0 = 0
    for s in singles:
        np.testing.assert_array_equal(ib[e0:e0 + s.nspec], s.ibool() + g0)
        np.testing.assert_array_equal(mass[g0:g0 + s.nglob], s.rmass_inverse())
        e0 += s.nspec
        g0 += s.nglob
    si, sx, sg, sa, stf = u.sources()
    ri, rx, rg, _, _ = u.receivers()
    e0 = r0 = 0
    for k, s in enumerate(singles):
        ki, kx, kg, ka, kstf = s.sources()
        np.testing.assert_array_equal(si[k], ki[0] + e0)
        np.testing.assert_array_equal(sa[k], ka[0])
        np.testing.assert_array_equal(stf[k], kstf[0])
        qi, qx, qg, _, _ = s.receivers()
        np.testing.assert_array_equal(ri[r0:r0 + len(qi)], qi + e0)
        np.testing.assert_array_equal(rx[r0:r0 + len(qi)], qx)
        e0 += s.nspec
        r0 += len(qi)
    # only the Stacey instance's absorbing sides are Stacey faces of the union
    assert u.nabs == singles[3].nabs > 0


def _brute_force_cost(cfgs, max_groups, step_cost):
    """Least modelled cost over every split of the NSTEP-sorted instances into at most max_groups runs."""
    import itertools
    from specfem2d_amd.driver import VISCO_POINT_COST
    order = sorted(range(len(cfgs)), key=lambda k: cfgs[k].nstep)
    n, best = len(order), float("inf")
    for g in range(1, max_groups + 1):
        for cuts in itertools.combinations(range(1, n), g - 1):
            bounds = (0,) + cuts + (n,)
            cost = 0.0
            for a, b in zip(bounds[:-1], bounds[1:]):
                idx = order[a:b]
                w = sum(cfgs[k].nglob * (VISCO_POINT_COST if cfgs[k].attenuation else 1.0) for k in idx)
                cost += max(cfgs[k].nstep for k in idx) * (step_cost + w)
            best = min(best, cost)
    return best


def test_nstep_groups_minimise_the_modelled_work():
    from specfem2d_amd.driver import VISCO_POINT_COST, plan_nstep_groups
    cfgs = config.uneven_batch_instances(n=9, seed=11)
    for max_groups, step_cost in ((1, 2e6), (3, 2e6), (4, 0.0), (9, 5e7)):
        groups = plan_nstep_groups(cfgs, max_groups, step_cost)
        assert sorted(k for g in groups for k in g) == list(range(len(cfgs))) and len(groups) <= max_groups
        # each group a contiguous run in NSTEP order
        ns = [[cfgs[k].nstep for k in g] for g in groups]
        assert all(max(a) <= min(b) for a, b in zip(ns[:-1], ns[1:]))
        cost = sum(max(cfgs[k].nstep for k in g) * (step_cost + sum(
            cfgs[k].nglob * (VISCO_POINT_COST if cfgs[k].attenuation else 1.0) for k in g)) for g in groups)
        assert cost == pytest.approx(_brute_force_cost(cfgs, max_groups, step_cost), rel=1e-12)
    # no fixed cost and enough groups: every distinct NSTEP its own group, no padding
    assert len(plan_nstep_groups(cfgs, 9, 0.0)) == len({c.nstep for c in cfgs})
