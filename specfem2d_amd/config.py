"""Simulation description for the SPECFEM2D elastic time-stepping path.

A :class:`SimulationConfig` carries the subset of the reference's ``DATA/Par_file``,
``DATA/SOURCE`` and ``DATA/STATIONS`` that the hot path needs: a structured internal
mesh with flat interfaces (``src/meshfem2D/read_mesh_files.F90:169-338``), per-GLL-point
or table materials (``src/specfem2D/read_materials.f90:138-141``,
``src/specfem2D/setup_mesh.F90:680-792``), Stacey absorbing edges, point force sources
and receivers.  Parameter names follow the Par_file keys.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

NGLL = 5
NGLL2 = NGLL * NGLL

# GLL nodes for 5 points (roots of (1-x^2) P'_4)
_R37 = float(np.sqrt(3.0 / 7.0))
XIGLL = np.array([-1.0, -_R37, 0.0, _R37, 1.0])


@dataclass
class Source:
    """A source of ``DATA/SOURCE``: a collocated force (a pressure source in a fluid element), ``source_type = 1``,
    or a moment tensor (Mxx, Mzz, Mxz), ``source_type = 2``."""

    x: float
    z: float
    f0: float
    factor: float = 1.0e10
    anglesource: float = 0.0  # degrees
    tshift: float = 0.0
    time_function_type: int = 1  # 1 Ricker, 2 first derivative of a Gaussian, 3 Gaussian
    source_type: int = 1
    Mxx: float = 0.0
    Mzz: float = 0.0
    Mxz: float = 0.0


# corner nodes of each side of a 4-node element (1 bottom, 2 right, 3 top, 4 left), counter-clockwise nodes
SIDE_NODES = {1: (0, 1), 2: (1, 2), 3: (2, 3), 4: (3, 0)}


@dataclass
class ExternalMesh:
    """A mesh read from the files of ``read_external_mesh`` (src/meshfem2D/read_external_mesh_files.F90): 4-node
    elements of any shape, their material numbers, and the element sides on the free surface and on the absorbing
    boundary (each as element and side code, 1 bottom, 2 right, 3 top, 4 left, found from the side's two nodes)."""

    knods: np.ndarray         # [nspec, 4] 0-based control nodes, counter-clockwise from the first
    coorg: np.ndarray         # [npgeo, 2] control-node coordinates (x, z)
    material: np.ndarray      # [nspec] 1-based material numbers
    free_elem: np.ndarray     # [nfree] 0-based elements of the free-surface sides
    free_code: np.ndarray     # [nfree] their side codes
    abs_elem: np.ndarray      # [nabs] 0-based elements of the absorbing sides
    abs_code: np.ndarray      # [nabs] their side codes

    @property
    def nspec(self) -> int:
        return int(self.knods.shape[0])

    def side_code(self, e: int, n1: int, n2: int) -> int:
        nodes = [int(x) for x in self.knods[e]]
        for code, (a, b) in SIDE_NODES.items():
            if {nodes[a], nodes[b]} == {int(n1), int(n2)}:
                return code
        raise ValueError(f"nodes {n1 + 1}, {n2 + 1} are not a side of element {e + 1}")

    def gll_point_coords(self):
        """(x, z) of every element GLL point under the bilinear map, [nspec, 25] each."""
        xi = np.tile(XIGLL, NGLL)
        ga = np.repeat(XIGLL, NGLL)
        sh = np.stack([0.25 * (xi - 1) * (ga - 1), -0.25 * (xi + 1) * (ga - 1), 0.25 * (xi + 1) * (ga + 1),
                       -0.25 * (xi - 1) * (ga + 1)])  # [4, 25]
        cx = self.coorg[self.knods, 0]  # [nspec, 4]
        cz = self.coorg[self.knods, 1]
        return cx @ sh, cz @ sh


@dataclass
class SimulationConfig:
    nx: int
    nz: int
    xmin: float
    xmax: float
    interfaces: Sequence[float]  # bottom first
    nz_layers: Sequence[int]
    nstep: int
    dt: float
    rho: object = 2700.0  # scalar or array [nspec, 25] (i fastest)
    vp: object = 3000.0
    vs: object = 1732.0
    material_mode: int = 0  # 0: Par_file material table, 1: MODEL=gll per-point values
    stacey: bool = True
    absorbbottom: bool = True
    absorbright: bool = True
    absorbtop: bool = False
    absorbleft: bool = True
    sources: List[Source] = field(default_factory=list)
    receivers: Optional[np.ndarray] = None  # [nrec, 2] (x, z)
    user_t0: float = 0.0
    anglerec: float = 0.0
    seismotype: int = 1  # 1 displacement, 2 velocity, 3 acceleration, 4 pressure, 6 potential (in fluid elements)
    use_trick: bool = False  # USE_TRICK_FOR_BETTER_PRESSURE
    # ATTENUATION_VISCOELASTIC: Qkappa, Qmu (scalars or per GLL point; 9999 = no attenuation) and N_SLS; the
    # reference frequency is the first source's f0
    attenuation: bool = False
    qkappa: object = 9999.0
    qmu: object = 9999.0
    n_sls: int = 3
    # PML_BOUNDARY_CONDITIONS: a convolutional PML on the absorbing sides (instead of Stacey edges), NELEM_PML_THICKNESS
    # elements thick, with the Par_file's K_MIN_PML, K_MAX_PML and damping_change_factor_{acoustic,elastic}
    pml: bool = False
    nelem_pml: int = 3
    pml_k_min: float = 1.0
    pml_k_max: float = 1.0
    pml_damping_acoustic: float = 0.5
    pml_damping_elastic: float = 1.0
    name: str = "custom"
    # read_external_mesh: the mesh comes from files (nx, nz, xmin, xmax, interfaces and the absorb* flags of the
    # internal mesher are then unused; Stacey edges are the mesh's absorbing sides)
    mesh: Optional[ExternalMesh] = None
    # anisotropic elements (material table icodemat 2): a flag per element [nspec] and, per GLL point, the
    # stiffness (c11, c13, c15, c33, c35, c55) [nspec*25, 6]; rho, vp = sqrt(c33/rho) and vs = sqrt(c55/rho) stay in
    # rho/vp/vs for the mass matrix and the Stacey impedances (read_materials.f90:160-200)
    aniso_elem: Optional[np.ndarray] = None
    aniso_c: Optional[np.ndarray] = None

    # ------------------------------------------------------------------
    @property
    def nspec(self) -> int:
        return self.mesh.nspec if self.mesh is not None else self.nx * self.nz

    @property
    def nglob(self) -> int:
        if self.mesh is not None:
            raise ValueError("an external mesh's point count comes from its numbering (Simulation.nglob)")
        return (4 * self.nx + 1) * (4 * self.nz + 1)

    @property
    def nrec(self) -> int:
        return 0 if self.receivers is None else int(len(self.receivers))

    def replace(self, **kw) -> "SimulationConfig":
        return dataclasses.replace(self, **kw)

    def material_arrays(self):
        """Per-point (rho, vp, vs) arrays of shape [nspec*25], float64."""
        n = self.nspec * NGLL2
        out = []
        for v in (self.rho, self.vp, self.vs):
            a = np.asarray(v, dtype=np.float64)
            if a.ndim == 0:
                a = np.full(n, float(a))
            a = np.ascontiguousarray(a.reshape(n))
            out.append(a)
        return out

    def attenuation_arrays(self):
        """Per-point (Qkappa, Qmu) arrays of shape [nspec*25], float64."""
        n = self.nspec * NGLL2
        return [np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64).reshape(-1)
                                                     if np.ndim(v) else float(v), (n,)))
                for v in (self.qkappa, self.qmu)]

    def attenuation_f0(self) -> float:
        """ATTENUATION_f0_REFERENCE: the f0 of the first source (prepare_attenuation.f90:319-322)."""
        if not self.sources:
            raise ValueError("attenuation needs a source (its f0 is the reference frequency)")
        return float(self.sources[0].f0)

    def gll_point_coords(self):
        """Coordinates (x, z) of every element GLL point, shape [nspec, 25] each."""
        if self.mesh is not None:
            return self.mesh.gll_point_coords()
        xs = self.xmin + (self.xmax - self.xmin) * np.arange(self.nx + 1) / self.nx
        zrows = []
        for il, nzl in enumerate(self.nz_layers):
            bot, top = self.interfaces[il], self.interfaces[il + 1]
            g = np.arange(nzl) / nzl
            zrows.append((1.0 - g) * bot + g * top)
        zrows.append(np.array([self.interfaces[-1]]))
        zs = np.concatenate(zrows)
        ex = np.tile(np.arange(self.nx), self.nz)
        ez = np.repeat(np.arange(self.nz), self.nx)
        x0, x1 = xs[ex], xs[ex + 1]
        z0, z1 = zs[ez], zs[ez + 1]
        xi = np.tile(XIGLL, NGLL)  # i fastest
        ga = np.repeat(XIGLL, NGLL)
        X = x0[:, None] + (x1 - x0)[:, None] * (xi[None, :] + 1.0) * 0.5
        Z = z0[:, None] + (z1 - z0)[:, None] * (ga[None, :] + 1.0) * 0.5
        return X, Z


# ----------------------------------------------------------------------
# seeded Gaussian random fields (correlation length, +-amplitude)
# ----------------------------------------------------------------------
def gaussian_random_field(cfg: SimulationConfig, corr_len: float, amplitude: float, seed: int,
                          spacing: Optional[float] = None) -> np.ndarray:
    """Zero-mean field in [-amplitude, amplitude] sampled at the GLL points of ``cfg``.

    White noise on a regular grid is filtered with a Gaussian kernel of the given
    correlation length (FFT convolution), scaled so that max |f| = amplitude, and read
    at each GLL point by nearest grid node."""
    zmin, zmax = float(cfg.interfaces[0]), float(cfg.interfaces[-1])
    if spacing is None:
        spacing = min((cfg.xmax - cfg.xmin) / cfg.nx, (zmax - zmin) / cfg.nz) / 2.0
    nxg = int(np.ceil((cfg.xmax - cfg.xmin) / spacing)) + 1
    nzg = int(np.ceil((zmax - zmin) / spacing)) + 1
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((nzg, nxg)).astype(np.float32)
    kx = np.fft.rfftfreq(nxg, d=spacing)
    kz = np.fft.fftfreq(nzg, d=spacing)
    # Fourier transform of exp(-r^2/a^2) is proportional to exp(-pi^2 a^2 k^2)
    filt = np.exp(-(np.pi * corr_len) ** 2 * (kz[:, None] ** 2 + kx[None, :] ** 2)).astype(np.float32)
    f = np.fft.irfft2(np.fft.rfft2(noise) * filt, s=(nzg, nxg))
    f *= amplitude / max(float(np.abs(f).max()), 1e-30)
    X, Z = cfg.gll_point_coords()
    ix = np.clip(np.rint((X - cfg.xmin) / spacing).astype(np.int64), 0, nxg - 1)
    iz = np.clip(np.rint((Z - zmin) / spacing).astype(np.int64), 0, nzg - 1)
    return f[iz, ix].reshape(-1).astype(np.float64)


# ----------------------------------------------------------------------
# presets
# ----------------------------------------------------------------------
def tromp2005_kernel(nstep: int = 3000) -> SimulationConfig:
    """``EXAMPLES/Tromp2005_kernel/DATA/Par_file`` + ``SOURCE`` (forward run)."""
    return SimulationConfig(
        name="Tromp2005_kernel", nx=80, nz=32, xmin=0.0, xmax=200000.0, interfaces=[0.0, 80000.0],
        nz_layers=[32], nstep=nstep, dt=2.0e-2, rho=2600.0, vp=5800.0, vs=3198.6, material_mode=0,
        stacey=True, absorbbottom=True, absorbright=True, absorbtop=False, absorbleft=True,
        sources=[Source(x=50000.0, z=40000.0, f0=0.42, factor=0.75e10, anglesource=270.0, time_function_type=2)],
        receivers=np.array([[150000.0, 40000.0]]), user_t0=8.0)


def plain_elastic_c1(nstep: int = 1600) -> SimulationConfig:
    """Plain forward elastic half-space: 80x60 elements over 4000 m x 3000 m,
    vp 3000, vs 1732, rho 2700, Stacey on three sides, Ricker 10 Hz at (2000, -2000),
    dt 1.1e-3, 20 surface receivers at x = 100 + 200 k."""
    rec = np.stack([100.0 + 200.0 * np.arange(20), np.zeros(20)], axis=1)
    return SimulationConfig(
        name="plain_elastic_80x60", nx=80, nz=60, xmin=0.0, xmax=4000.0, interfaces=[-3000.0, 0.0],
        nz_layers=[60], nstep=nstep, dt=1.1e-3, rho=2700.0, vp=3000.0, vs=1732.0, material_mode=0,
        stacey=True, absorbbottom=True, absorbright=True, absorbtop=False, absorbleft=True,
        sources=[Source(x=2000.0, z=-2000.0, f0=10.0, factor=1.0e10, anglesource=0.0, time_function_type=1)],
        receivers=rec, user_t0=0.0)


def random_elastic(nx: int, nz: int, elem_size: float, nstep: int, dt: float, nrec: int, seed: int = 1,
                   f0: float = 8.0, corr_len: float = 200.0, amplitude: float = 0.10,
                   vp0: float = 3000.0, vs0: float = 1732.0, rho0: float = 2500.0,
                   name: Optional[str] = None) -> SimulationConfig:
    """Elastic half-space over a seeded Gaussian random field (+-amplitude around
    vp0/vs0/rho0), Stacey on three sides, free surface on top, a Ricker source at
    mid-width one third down, and ``nrec`` receivers evenly spread along the surface."""
    xmax = nx * elem_size
    depth = nz * elem_size
    cfg = SimulationConfig(
        name=name or f"random_elastic_{nx}x{nz}", nx=nx, nz=nz, xmin=0.0, xmax=xmax,
        interfaces=[-depth, 0.0], nz_layers=[nz], nstep=nstep, dt=dt, material_mode=1,
        stacey=True, absorbbottom=True, absorbright=True, absorbtop=False, absorbleft=True,
        sources=[Source(x=xmax / 2, z=-depth / 3, f0=f0, factor=1.0e10, anglesource=0.0, time_function_type=1)],
        receivers=np.stack([np.linspace(elem_size, xmax - elem_size, nrec), np.zeros(nrec)], axis=1))
    pert = gaussian_random_field(cfg, corr_len, amplitude, seed)
    return cfg.replace(rho=rho0 * (1.0 + pert), vp=vp0 * (1.0 + pert), vs=vs0 * (1.0 + pert))


def adjoint_gradient_800x400(nstep: int = 2000, anomaly: bool = False) -> SimulationConfig:
    """Single-shot adjoint-gradient workload: 800 x 400 elements of 20 m (16 km x 8 km), dt 5e-4 s (Courant
    number 0.43 at vp 3000 over the closest GLL points), 2000 steps, homogeneous background vp 3000, vs 1732,
    rho 2500 (MODEL=gll arrays), Stacey on three sides and a free surface, a vertical Ricker force of 10 Hz at
    1200 m depth under the centre, 200 surface receivers.  With ``anomaly`` the model carries a +5 % Gaussian vs
    anomaly of radius 500 m half-way between the source and the surface (the 'observed' model)."""
    nx, nz, h = 800, 400, 20.0
    xmax, depth = nx * h, nz * h
    cfg = SimulationConfig(
        name="adjoint_gradient_800x400", nx=nx, nz=nz, xmin=0.0, xmax=xmax, interfaces=[-depth, 0.0],
        nz_layers=[nz], nstep=nstep, dt=5.0e-4, material_mode=1, stacey=True, absorbbottom=True,
        absorbright=True, absorbtop=False, absorbleft=True,
        sources=[Source(x=xmax / 2, z=-1200.0, f0=10.0, factor=1.0e10, anglesource=0.0, time_function_type=1)],
        receivers=np.stack([np.linspace(h, xmax - h, 200), np.zeros(200)], axis=1))
    n = cfg.nspec * NGLL2
    rho, vp, vs = np.full(n, 2500.0), np.full(n, 3000.0), np.full(n, 1732.0)
    if anomaly:
        X, Z = cfg.gll_point_coords()
        r2 = (X.reshape(-1) - xmax / 2) ** 2 + (Z.reshape(-1) + 600.0) ** 2
        vs = vs * (1.0 + 0.05 * np.exp(-r2 / 500.0 ** 2))
    return cfg.replace(rho=rho, vp=vp, vs=vs)


def batched_fwi_600x300(nstep: int = 1500, true_model: bool = False, nrec: int = 150) -> SimulationConfig:
    """Multi-shot FWI workload: 600 x 300 elements of 20 m (12 km x 6 km), dt 4e-4 s (Courant number 0.38 at
    the true model's highest vp over the closest GLL points), 1500 steps, Stacey on three sides and a free
    surface, ``nrec`` receivers along the surface.  The synthetic model is homogeneous (vp 3000, vs 1732,
    rho 2500, MODEL=gll arrays); ``true_model`` gives the 'observed' model, a seeded Gaussian random field
    (correlation length 200 m, +-10 %) around it.  The shots come from ``surface_shots``."""
    nx, nz, h = 600, 300, 20.0
    xmax, depth = nx * h, nz * h
    cfg = SimulationConfig(
        name="batched_fwi_600x300", nx=nx, nz=nz, xmin=0.0, xmax=xmax, interfaces=[-depth, 0.0],
        nz_layers=[nz], nstep=nstep, dt=4.0e-4, material_mode=1, stacey=True, absorbbottom=True,
        absorbright=True, absorbtop=False, absorbleft=True,
        sources=[Source(x=xmax / 2, z=-10.0, f0=10.0, factor=1.0e10)],
        receivers=np.stack([np.linspace(h, xmax - h, nrec), np.zeros(nrec)], axis=1))
    n = cfg.nspec * NGLL2
    if true_model:
        pert = gaussian_random_field(cfg, 200.0, 0.10, 4242)
    else:
        pert = np.zeros(n)
    return cfg.replace(rho=2500.0 * (1.0 + pert), vp=3000.0 * (1.0 + pert), vs=1732.0 * (1.0 + pert))


def surface_shots(cfg: SimulationConfig, nshot: int, f0: float = 10.0, depth: float = 10.0) -> List[List[Source]]:
    """``nshot`` vertical Ricker force sources of frequency f0 evenly spaced along the free surface (from 2 % to
    98 % of the width, ``depth`` metres down), one source per shot."""
    w = cfg.xmax - cfg.xmin
    xs = cfg.xmin + w * (0.02 + 0.96 * (np.arange(nshot) + 0.5) / nshot)
    top = cfg.interfaces[-1]
    return [[Source(x=float(x), z=top - depth, f0=f0, factor=1.0e10)] for x in xs]


def small_test_mesh(nx: int = 8, nz: int = 6, nstep: int = 200, seed: int = 0, stacey: bool = True,
                    gll: bool = False) -> SimulationConfig:
    """A small mesh for fast parity tests: 100 m elements, Ricker 15 Hz inside."""
    elem = 100.0
    cfg = SimulationConfig(
        name=f"small_{nx}x{nz}", nx=nx, nz=nz, xmin=0.0, xmax=nx * elem, interfaces=[-nz * elem, 0.0],
        nz_layers=[nz], nstep=nstep, dt=2.0e-3, rho=2500.0, vp=3000.0, vs=1700.0, material_mode=0,
        stacey=stacey,
        sources=[Source(x=nx * elem * 0.37, z=-nz * elem * 0.45, f0=15.0, factor=1.0e9, anglesource=30.0)],
        receivers=np.array([[nx * elem * 0.8, 0.0], [nx * elem * 0.25, -nz * elem * 0.6],
                            [nx * elem * 0.5 + 13.0, -nz * elem * 0.3 - 7.0]]))
    if gll:
        rng = np.random.default_rng(seed)
        n = cfg.nspec * NGLL2
        p = 0.05 * rng.standard_normal(n)
        cfg = cfg.replace(material_mode=1, rho=2500.0 * (1 + p), vp=3000.0 * (1 + p), vs=1700.0 * (1 - p))
    return cfg


def pressure_source_acoustic_44x44(nstep: int = 1400) -> SimulationConfig:
    """``EXAMPLES/check_absolute_amplitude_of_pressure_source_seismograms_acoustic/REF_SEIS/Par_file`` +
    ``SOURCE``: a homogeneous fluid (rho 2000, vp 3297.849, vs 0) of 44 x 44 elements over 2000 m x 2000 m, a
    Gaussian pressure source of 18 Hz at (1000, 1000), one pressure receiver at (1500, 1500), dt 4e-4, 1400
    steps, USE_TRICK_FOR_BETTER_PRESSURE.  The reference absorbs with a 3-element CPML on every side; here the
    sides are Stacey edges, which changes the trace only after the waves reach that layer (about 0.35 s)."""
    return SimulationConfig(
        name="pressure_source_acoustic_44x44", nx=44, nz=44, xmin=0.0, xmax=2000.0, interfaces=[0.0, 2000.0],
        nz_layers=[44], nstep=nstep, dt=4.0e-4, rho=2000.0, vp=3297.849, vs=0.0, material_mode=0, stacey=True,
        absorbbottom=True, absorbright=True, absorbtop=True, absorbleft=True,
        sources=[Source(x=1000.0, z=1000.0, f0=18.0, factor=1.0, time_function_type=3)],
        receivers=np.array([[1500.0, 1500.0]]), seismotype=4, use_trick=True)


def layered_fluid_solid(nx: int, nz: int, nz_fluid: int, elem_size: float, nstep: int, dt: float,
                        rho_f: float = 1020.0, vp_f: float = 1500.0, rho_s: float = 2500.0, vp_s: float = 3000.0,
                        vs_s: float = 1732.0, absorbtop: bool = False, pert=None, name: Optional[str] = None,
                        sources: Optional[List[Source]] = None, receivers=None,
                        seismotype: int = 1) -> SimulationConfig:
    """A fluid layer of ``nz_fluid`` elements (vs = 0) over a solid: ``nx`` x ``nz`` elements of ``elem_size``,
    the free surface (or, with ``absorbtop``, a Stacey edge) on top of the fluid and Stacey edges elsewhere,
    MODEL=gll arrays; ``pert`` (an array over the GLL points) scales the solid's rho, vp and vs by 1 + pert."""
    xmax, depth = nx * elem_size, nz * elem_size
    cfg = SimulationConfig(
        name=name or f"fluid_solid_{nx}x{nz}", nx=nx, nz=nz, xmin=0.0, xmax=xmax, interfaces=[-depth, 0.0],
        nz_layers=[nz], nstep=nstep, dt=dt, material_mode=1, stacey=True, absorbbottom=True, absorbright=True,
        absorbtop=absorbtop, absorbleft=True, sources=sources or [], receivers=receivers, seismotype=seismotype)
    ez = np.repeat(np.arange(nz), nx)
    fluid = np.repeat((ez >= nz - nz_fluid)[:, None], NGLL2, axis=1).reshape(-1)
    p = np.zeros(fluid.size) if pert is None else np.asarray(pert, np.float64).reshape(-1)
    rho = np.where(fluid, rho_f, rho_s * (1.0 + p))
    vp = np.where(fluid, vp_f, vp_s * (1.0 + p))
    vs = np.where(fluid, 0.0, vs_s * (1.0 + p))
    return cfg.replace(rho=rho, vp=vp, vs=vs)


def fluid_solid_plane_wave(nstep: int = 1700, solid: bool = True, absorbtop: bool = True,
                           orientation: str = "above") -> SimulationConfig:
    """Plane-wave reflection at a flat fluid-solid interface, 50 m elements, fluid rho 1020, vp 1500 and solid
    rho 2500, vp 3000, vs 1732 (fluid too when not ``solid``), Ricker pressure sources of 5 Hz every 12.5 m along a
    line 1000 m from the interface in the fluid, pressure receivers 500 m from the interface (three) and 500 m
    beyond the source line (one), dt 1e-3, USE_TRICK_FOR_BETTER_PRESSURE.

    ``orientation``: ``"above"`` -- 120 x 80 elements (6 km x 4 km), the fluid the top 40 rows (the free surface
    on top unless ``absorbtop``); ``"below"`` -- the fluid the bottom 40 rows under the solid; ``"beside"`` --
    80 x 120 elements, the fluid the right 40 columns beside the solid.  Stacey edges elsewhere."""
    h = 50.0
    if orientation == "beside":
        nx, nz = 80, 120
        fluid = np.tile(np.arange(nx), nz) >= nx // 2
        depth = nz * h
        srcs = [Source(x=3000.0, z=float(z), f0=5.0, factor=1.0) for z in np.arange(-depth + 6.25, 0.0, 12.5)]
        rec = np.array([[2500.0, -3000.0], [2500.0, -2500.0], [2500.0, -3500.0], [3500.0, -3000.0]])
    else:
        nx, nz = 120, 80
        ez = np.repeat(np.arange(nz), nx)
        fluid = ez >= nz // 2 if orientation == "above" else ez < nz // 2
        z_src, z_rec, z_far = (-1000.0, -1500.0, -500.0) if orientation == "above" else (-3000.0, -2500.0, -3500.0)
        srcs = [Source(x=float(x), z=z_src, f0=5.0, factor=1.0) for x in np.arange(6.25, nx * h, 12.5)]
        rec = np.array([[2500.0, z_rec], [3000.0, z_rec], [3500.0, z_rec], [3000.0, z_far]])
    if not solid:
        fluid = np.ones_like(fluid)
    f25 = np.repeat(fluid, NGLL2)
    cfg = SimulationConfig(
        name=f"fluid_solid_plane_wave_{orientation}", nx=nx, nz=nz, xmin=0.0, xmax=nx * h,
        interfaces=[-nz * h, 0.0], nz_layers=[nz], nstep=nstep, dt=1.0e-3, material_mode=1, stacey=True,
        absorbbottom=True, absorbright=True, absorbtop=absorbtop or orientation != "above", absorbleft=True,
        sources=srcs, receivers=rec, seismotype=4, use_trick=True,
        rho=np.where(f25, 1020.0, 2500.0), vp=np.where(f25, 1500.0, 3000.0), vs=np.where(f25, 0.0, 1732.0))
    return cfg


def c2_mesh_coupled(nstep: int = 3000, nx: int = 2000, nz: int = 1000, elem_size: float = 10.0,
                    nrec: int = 200, seed: int = 2, nelem_pml: int = 0) -> SimulationConfig:
    """The large coupled forward workload on one GPU: 2000 x 1000 elements of 10 m (20 km x 10 km, about 32 M
    GLL points), the top 1 km an acoustic water layer (vp 1500, rho 1020) under a free surface, the solid below a
    seeded Gaussian random field (correlation length 200 m, +-10 %) around vp 3000, vs 1732, rho 2500; a Ricker
    force of 8 Hz at a third of the depth, ``nrec`` displacement receivers 20 m below the sea floor, dt 2e-4.
    The other three sides absorb: with a CPML layer ``nelem_pml`` elements thick (K 1, the Par_file's default
    damping factors) when ``nelem_pml`` > 0, else with Stacey edges."""
    xmax, depth = nx * elem_size, nz * elem_size
    nzf = nz // 10
    sea_floor = -nzf * elem_size
    rec = np.stack([np.linspace(elem_size, xmax - elem_size, nrec), np.full(nrec, sea_floor - 20.0)], axis=1)
    src = [Source(x=xmax / 2, z=-depth / 3, f0=8.0, factor=1.0e10)]
    base = SimulationConfig(name="c2", nx=nx, nz=nz, xmin=0.0, xmax=xmax, interfaces=[-depth, 0.0],
                            nz_layers=[nz], nstep=nstep, dt=2.0e-4)
    pert = gaussian_random_field(base, 200.0, 0.10, seed)
    cfg = layered_fluid_solid(nx, nz, nzf, elem_size, nstep, 2.0e-4, pert=pert,
                              name=f"c2_mesh_coupled_{nx}x{nz}", sources=src, receivers=rec)
    if nelem_pml > 0:
        cfg = cfg.replace(stacey=False, pml=True, nelem_pml=nelem_pml, name=f"c2_mesh_coupled_cpml_{nx}x{nz}")
    return cfg


def uneven_batch_instances(n: int = 48, seed: int = 5, elem_size: float = 20.0, dt: float = 4.0e-4,
                           nelem_pml: int = 10, size_range=((100, 1000), (100, 500)), nstep_range=(1000, 4000),
                           q_range=(30.0, 100.0), nrec: int = 20) -> List[SimulationConfig]:
    """The uneven-batch workload: ``n`` instances, mesh sizes drawn log-uniformly (nx in 100..1000, nz in 100..500
    elements of ``elem_size``), half acoustic (water, vp 1500, rho 1020) and half viscoelastic (a seeded random-field
    solid, +-10 % around vp 3000, vs 1732, rho 2500, with Qkappa and Qmu drawn uniformly in 30..100, 3 solids), NSTEP
    uniform in 1000..4000, a CPML frame ``nelem_pml`` elements thick on the left, right and bottom and a free surface on
    top, one Ricker source (a pressure source in the water, a force in the solid; f0 8-12 Hz) and ``nrec`` displacement
    receivers along a line 40 m down; FP32, Newmark, one time step for all.  Seeded."""
    rng = np.random.default_rng(seed)
    (nx0, nx1), (nz0, nz1) = size_range
    out = []
    for k in range(n):
        nx = int(round(np.exp(rng.uniform(np.log(nx0), np.log(nx1)))))
        nz = int(round(np.exp(rng.uniform(np.log(nz0), np.log(nz1)))))
        nstep = int(rng.integers(nstep_range[0], nstep_range[1] + 1))
        f0 = float(rng.uniform(8.0, 12.0))
        xmax, depth = nx * elem_size, nz * elem_size
        fluid = k % 2 == 0
        src = [Source(x=xmax * rng.uniform(0.3, 0.7), z=-depth * rng.uniform(0.2, 0.5), f0=f0,
                      factor=1.0e4 if fluid else 1.0e10)]
        rec = np.stack([np.linspace(xmax * 0.1, xmax * 0.9, nrec), np.full(nrec, -40.0)], axis=1)
        cfg = SimulationConfig(name=f"uneven_{k}_{'acoustic' if fluid else 'viscoelastic'}_{nx}x{nz}", nx=nx, nz=nz,
                               xmin=0.0, xmax=xmax, interfaces=[-depth, 0.0], nz_layers=[nz], nstep=nstep, dt=dt,
                               material_mode=1, stacey=False, pml=True, nelem_pml=nelem_pml, absorbtop=False,
                               sources=src, receivers=rec)
        n25 = cfg.nspec * NGLL2
        if fluid:
            cfg = cfg.replace(rho=np.full(n25, 1020.0), vp=np.full(n25, 1500.0), vs=np.zeros(n25))
        else:
            pert = gaussian_random_field(cfg, 200.0, 0.10, seed * 1000 + k)
            cfg = cfg.replace(rho=2500.0 * (1 + pert), vp=3000.0 * (1 + pert), vs=1732.0 * (1 + pert),
                              attenuation=True, qkappa=float(rng.uniform(*q_range)), qmu=float(rng.uniform(*q_range)))
        out.append(cfg)
    return out


def luoyang_fluid_solid(nstep: int = 1600) -> SimulationConfig:
    """``EXAMPLES/LuoYang_fluid_solid_kernel/DATA/Par_file`` + ``SOURCE`` (forward run): 100 x 60 elements of 50 m
    (5 km x 3 km, three layers of 20 rows), fluid (rho 2500, vp 2700) in columns 1-40 and 61-100 around a solid
    strip (rho 2700, vp 3000, vs 1732.051) in columns 41-60, Stacey edges on all sides, a Ricker pressure source of
    10 Hz (factor 1e10) at (4000, 1500) in the fluid, one potential receiver (seismotype 6) at (1000, 1500),
    dt 1.1e-3.  The reference meshes it with 9-node elements; on these rectangles the geometry is the same."""
    nx, nz = 100, 60
    ex = np.tile(np.arange(nx), nz)
    solid = np.repeat((ex >= 40) & (ex < 60), NGLL2)
    return SimulationConfig(
        name="LuoYang_fluid_solid_kernel", nx=nx, nz=nz, xmin=0.0, xmax=5000.0, interfaces=[0.0, 1000.0, 2000.0, 3000.0],
        nz_layers=[20, 20, 20], nstep=nstep, dt=1.1e-3, rho=np.where(solid, 2700.0, 2500.0),
        vp=np.where(solid, 3000.0, 2700.0), vs=np.where(solid, 1732.051, 0.0), material_mode=0, stacey=True,
        absorbbottom=True, absorbright=True, absorbtop=True, absorbleft=True,
        sources=[Source(x=4000.0, z=1500.0, f0=10.0, factor=1.0e10, time_function_type=1)],
        receivers=np.array([[1000.0, 1500.0]]), seismotype=6)
