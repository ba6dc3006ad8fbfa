"""The host driver's setup (what xspecfem2D hands to prepare_constants_device_ & co.) against the
oracle: integer/index data bit-exact, FP32 arrays to a few ulp.  Runs on the CPU (no device call)."""
import numpy as np
import pytest

from oracle.oracle import OracleSim
from specfem2d_amd import config
from specfem2d_amd.driver import Simulation

CASES = {
    "tromp2005": lambda: config.tromp2005_kernel(nstep=50),
    "c1_80x60": lambda: config.plain_elastic_c1(nstep=50),
    "small_gll": lambda: config.small_test_mesh(nx=9, nz=7, nstep=50, gll=True),
    "layers": lambda: config.small_test_mesh(nx=6, nz=5, nstep=50).replace(
        interfaces=[-500.0, -320.0, -90.0, 0.0], nz_layers=[2, 2, 1], xmax=730.0, absorbtop=True),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def pair(request):
    cfg = CASES[request.param]()
    return cfg, Simulation(cfg, device=False), OracleSim(cfg, "sp")


def test_ibool_bit_exact(pair):
    cfg, sim, orc = pair
    np.testing.assert_array_equal(sim.ibool(), orc.ibool())
    assert sim.nglob == orc.nglob == cfg.nglob


def test_geometry_and_materials(pair):
    _, sim, orc = pair
    for name in ("xix", "xiz", "gammax", "gammaz", "jacobian", "mu", "kappa", "rho_vp", "rho_vs"):
        a, b = sim.elem_array(name), orc.elem_array(name).astype(np.float32)
        scale = np.abs(b).max()
        np.testing.assert_allclose(a, b, rtol=4e-7, atol=1e-6 * scale, err_msg=name)


def test_mass_matrix(pair):
    _, sim, orc = pair
    np.testing.assert_allclose(sim.rmass_inverse(), orc.rmass_inverse(), rtol=2e-6)


def test_sources_and_receivers_bit_exact_indices(pair):
    cfg, sim, orc = pair
    si, sx, sg, sa, stf = sim.sources()
    oi, ox, og, osa, ostf = orc.sources()
    np.testing.assert_array_equal(si, oi)
    np.testing.assert_allclose(sx, ox, atol=1e-12)
    np.testing.assert_allclose(sg, og, atol=1e-12)
    np.testing.assert_allclose(sa, osa, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(stf, ostf, rtol=1e-6, atol=1e-6 * np.abs(ostf).max())
    assert sim.t0 == orc.t0
    ri, rx, rg, rxi, rga = sim.receivers()
    qi, qx, qg, qxi, qga = orc.receivers()
    np.testing.assert_array_equal(ri, qi)
    np.testing.assert_allclose(rx, qx, atol=1e-12)
    np.testing.assert_allclose(rg, qg, atol=1e-12)
    np.testing.assert_allclose(rxi, qxi, atol=1e-7)
    np.testing.assert_allclose(rga, qga, atol=1e-7)


def test_stacey_arrays(pair):
    cfg, sim, orc = pair
    isp, ij, nrm, jw, edge, ib = sim.stacey()
    oisp, ocode, ocorner, oib = orc.absorbing()
    np.testing.assert_array_equal(isp, oisp)
    np.testing.assert_array_equal(edge, ocode)
    np.testing.assert_array_equal(ib, oib)
    # skipped corner points carry NGLL+1, exactly where the oracle's corner flags say
    for f in range(len(isp)):
        skip = ij[f, :, 0] == 6
        exp = np.zeros(5, bool)
        if edge[f] == 1:
            exp[0], exp[4] = ocorner[f, 0], ocorner[f, 1]
        if edge[f] == 3:
            exp[0], exp[4] = ocorner[f, 2], ocorner[f, 3]
        np.testing.assert_array_equal(skip, exp)
    # axis-aligned edges: outward unit normals
    expect = {1: (0, -1), 2: (1, 0), 3: (0, 1), 4: (-1, 0)}
    for f in range(len(isp)):
        np.testing.assert_allclose(nrm[f], np.tile(expect[edge[f]], (5, 1)), atol=1e-6)


def test_set_model_equals_a_fresh_setup():
    """sf2d_sim_set_model recomputes moduli, impedances and the mass matrix exactly as a new simulation would."""
    base = config.small_test_mesh(nx=9, nz=7, nstep=50, gll=True, seed=1).replace(absorbtop=True)
    other = config.small_test_mesh(nx=9, nz=7, nstep=50, gll=True, seed=2).replace(absorbtop=True)
    sim = Simulation(base, device=False)
    sim.set_model(other.rho, other.vp, other.vs)
    fresh = Simulation(other, device=False)
    for name in ("mu", "kappa", "rho_vp", "rho_vs", "rho"):
        np.testing.assert_array_equal(sim.elem_array(name), fresh.elem_array(name), err_msg=name)
    np.testing.assert_array_equal(sim.rmass_inverse(), fresh.rmass_inverse())
    with pytest.raises(ValueError):
        sim.set_model(-np.asarray(other.rho), other.vp, other.vs)


@pytest.mark.parametrize("case", ["pressure_source", "coupled"])
def test_fluid_setup_bit_exact(case):
    """Fluid elements: the host driver's fluid mass (with the fluid Stacey term and the sea-floor exclusions),
    fluid-solid edges, pressure source arrays, time functions (Gaussian with USE_TRICK_FOR_BETTER_PRESSURE) and
    t0 equal the oracle's bit for bit."""
    from specfem2d_amd.config import Source
    if case == "pressure_source":
        cfg = config.pressure_source_acoustic_44x44(nstep=50)
    else:
        cfg = config.layered_fluid_solid(
            12, 9, 3, 100.0, 50, 1e-3,
            sources=[Source(x=530.0, z=-130.0, f0=12.0, factor=1e3, time_function_type=3),
                     Source(x=420.0, z=-600.0, f0=12.0, factor=1e9, anglesource=20.0)],
            receivers=np.array([[300.0, -150.0], [700.0, -700.0]]), seismotype=4).replace(use_trick=True)
    sim, orc = Simulation(cfg, device=False), OracleSim(cfg, "sp")
    assert sim.nspec_acoustic == orc.nspec_acoustic > 0
    assert sim.t0 == orc.t0
    np.testing.assert_array_equal(sim.rmass_acoustic().astype(np.float64), orc.rmass_acoustic())
    np.testing.assert_array_equal(sim.rmass_inverse().astype(np.float64), orc.rmass_inverse())
    assert sim.ncoupling == orc.ncoupling
    if sim.ncoupling:
        np.testing.assert_array_equal(sim.coupling(), orc.coupling())
    si, _, _, sa, st = sim.sources()
    oi, _, _, osa, ost = orc.sources()
    np.testing.assert_array_equal(si, oi)
    np.testing.assert_array_equal(sa.astype(np.float64), osa)
    np.testing.assert_array_equal(st.astype(np.float64), ost)
    sim.close()


def test_attenuation_coefficients_match_oracle():
    """The host driver's Zener fit (Cholesky on the normal equations) and single-precision convolution
    coefficients against the oracle's (Gaussian elimination), per point, on a heterogeneous-Q mesh."""
    rng = np.random.default_rng(3)
    cfg = config.small_test_mesh(nx=8, nz=6, nstep=10)
    n = cfg.nspec * 25
    qk = np.repeat(rng.uniform(10.0, 200.0, cfg.nspec), 25)
    qm = np.repeat(rng.uniform(8.0, 150.0, cfg.nspec), 25)
    qk[:50], qm[:50] = 9999.0, 9999.0
    # the last elements: quality factors that vary from point to point inside the element, and mixed pairs (one
    # of Qkappa, Qmu at 9999, the other finite: the reference still fits both, prepare_attenuation.f90:347)
    tail = slice(n - 250, n)
    qk[tail] = rng.uniform(12.0, 90.0, 250)
    qm[tail] = rng.uniform(9.0, 70.0, 250)
    mixed_k = n - 250 + np.arange(0, 250, 7)
    mixed_m = n - 250 + np.arange(3, 250, 7)
    qk[mixed_k] = 9999.0
    qm[mixed_m] = 9999.0
    cfg = cfg.replace(attenuation=True, qkappa=qk, qmu=qm)
    sim = Simulation(cfg, device=False)
    mine = sim.attenuation_coefficients()
    sim.close()
    ref = OracleSim(cfg, "sp").attenuation_coefficients()
    for a, b in zip(mine, ref):
        b = b.astype(np.float32)
        assert np.all(np.abs(a - b) <= 2 * np.spacing(np.abs(b).astype(np.float32)))
    assert np.all(mine[0].reshape(n, 3)[:50] == 0)  # Q = 9999: no attenuation
    # a mixed pair is fitted: non-zero coefficients for the finite factor (A_mu, B_mu where only Qkappa is 9999)
    a_mu = mine[2].reshape(n, 3)
    assert np.all(a_mu[mixed_k] != 0)
    # per-point variation inside an element reaches the coefficients
    assert len(np.unique(mine[3].reshape(n, 3)[tail][:25, 0])) == 22  # 4 of the 25 points have Qmu = 9999
