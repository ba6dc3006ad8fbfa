"""Python handle on the host driver and the HIP library (ctypes, no torch types).

``Simulation(cfg)`` builds the solver setup in the C++ driver (``libspecfem2d_host.so``) and,
on a GPU, hands it to the Fortran-style C-ABI of ``libspecfem2d_gpu.so``.  There is no CPU
fallback: if the libraries are missing or no HIP device is present the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from .config import NGLL2, SimulationConfig

_HERE = os.path.dirname(os.path.abspath(__file__))
GPU_LIB = os.path.join(_HERE, "libspecfem2d_gpu.so")
HOST_LIB = os.path.join(_HERE, "libspecfem2d_host.so")
MAXLAYERS = 32


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("nz", C.c_int), ("nlayers", C.c_int), ("nz_layer", C.c_int * MAXLAYERS),
        ("xmin", C.c_double), ("xmax", C.c_double), ("zinterface", C.c_double * (MAXLAYERS + 1)),
        ("absorbbottom", C.c_int), ("absorbright", C.c_int), ("absorbtop", C.c_int), ("absorbleft", C.c_int),
        ("stacey", C.c_int), ("nstep", C.c_int), ("dt", C.c_double), ("user_t0", C.c_double),
        ("anglerec", C.c_double), ("material_mode", C.c_int), ("seismotype", C.c_int), ("use_trick", C.c_int),
        ("pml", C.c_int), ("nelem_pml", C.c_int), ("k_min_pml", C.c_double), ("k_max_pml", C.c_double),
        ("damping_ac", C.c_double), ("damping_el", C.c_double),
    ]


class SourceC(C.Structure):
    _fields_ = [
        ("x", C.c_double), ("z", C.c_double), ("f0", C.c_double), ("factor", C.c_double),
        ("anglesource", C.c_double), ("tshift", C.c_double), ("time_function_type", C.c_int),
        ("source_type", C.c_int), ("Mxx", C.c_double), ("Mzz", C.c_double), ("Mxz", C.c_double),
    ]


_lib_cache = {}


def load_libraries():
    """Load the HIP C-ABI library and the host driver; raise if they are not built."""
    if "host" in _lib_cache:
        return _lib_cache["gpu"], _lib_cache["host"]
    for path in (GPU_LIB, HOST_LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                               " (make -C specfem2d_amd)")
    gpu = C.CDLL(GPU_LIB, mode=C.RTLD_GLOBAL)
    host = C.CDLL(HOST_LIB)
    P = C.c_void_p
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    host.sf2d_sim_create.restype = P
    host.sf2d_sim_create.argtypes = [C.POINTER(Params), dp, dp, dp, C.c_int, C.POINTER(SourceC), C.c_int, dp,
                                     C.c_int, C.c_int]
    host.sf2d_sim_prepare_device.argtypes = [P]
    host.sf2d_sim_set_attenuation.argtypes = [P, dp, dp, C.c_int]
    host.sf2d_sim_set_attenuation.restype = C.c_int
    host.sf2d_sim_get_attenuation.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    host.sf2d_sim_prepare_device.restype = C.c_int
    host.sf2d_sim_destroy.argtypes = [P]
    for f in ("sf2d_sim_nspec", "sf2d_sim_nglob", "sf2d_sim_nabs", "sf2d_sim_nspec_acoustic", "sf2d_sim_ncoupling"):
        getattr(host, f).argtypes = [P]
        getattr(host, f).restype = C.c_int
    host.sf2d_sim_t0.argtypes = [P]
    host.sf2d_sim_t0.restype = C.c_double
    host.sf2d_sim_mesh_pointer.argtypes = [P]
    host.sf2d_sim_mesh_pointer.restype = C.c_long
    for f in ("sf2d_sim_get_ibool", "sf2d_sim_get_rmass_inverse"):
        getattr(host, f).argtypes = [P, P]
    host.sf2d_sim_get_elem_array.argtypes = [P, C.c_int, P]
    host.sf2d_sim_get_sources.argtypes = [P, P, P, P, P, P]
    host.sf2d_sim_get_receivers.argtypes = [P, P, P, P, P, P]
    host.sf2d_sim_get_stacey.argtypes = [P, P, P, P, P, P, P]
    host.sf2d_sim_forward.argtypes = [P, C.c_int, P, P, P]
    host.sf2d_sim_forward.restype = C.c_int
    host.sf2d_sim_adjoint.argtypes = [P, P, P, P, P, P, P, P]
    host.sf2d_sim_adjoint.restype = C.c_int
    host.sf2d_sim_set_chain_elems.argtypes = [P, C.c_int]
    host.sf2d_sim_set_chain_elems.restype = C.c_int
    host.sf2d_sim_set_anisotropy.argtypes = [P, P, P]
    host.sf2d_sim_set_anisotropy.restype = C.c_int
    host.sf2d_sim_adjoint_full.argtypes = [P] * 12
    host.sf2d_sim_adjoint_full.restype = C.c_int
    host.sf2d_sim_set_model.argtypes = [P, P, P, P]
    host.sf2d_sim_set_model.restype = C.c_int
    host.sf2d_sim_gradient_forward.argtypes = [P, P, C.c_int, P, P]
    host.sf2d_sim_gradient_forward.restype = C.c_int
    host.sf2d_sim_gradient_backward.argtypes = [P, P, P]
    host.sf2d_sim_gradient_backward.restype = C.c_int
    host.sf2d_sim_create_external.restype = P
    host.sf2d_sim_create_external.argtypes = [C.POINTER(Params), C.c_int, C.c_int, P, P, C.c_int, P, P, C.c_int, P, P,
                                              dp, dp, dp, C.c_int, C.POINTER(SourceC), C.c_int, dp, C.c_int, C.c_int]
    host.sf2d_sim_create_union.restype = P
    host.sf2d_sim_create_union.argtypes = [C.c_int, P]
    host.sf2d_sim_create_shots.restype = P
    host.sf2d_sim_create_shots.argtypes = [C.POINTER(Params), dp, dp, dp, C.c_int, P, C.POINTER(SourceC), C.c_int,
                                           dp]
    host.sf2d_sim_batch_forward.argtypes = [P, P, P]
    host.sf2d_sim_batch_forward.restype = C.c_int
    host.sf2d_sim_nshot.argtypes = [P]
    host.sf2d_sim_nshot.restype = C.c_int
    host.sf2d_sim_shot_t0.argtypes = [P, P]
    host.sf2d_sim_shot_misfits.argtypes = [P, P]
    host.sf2d_sim_batch_get_fields.argtypes = [P, C.c_int, P, P, P]
    host.sf2d_sim_reset.argtypes = [P]
    host.sf2d_sim_steps_done.argtypes = [P]
    host.sf2d_sim_steps_done.restype = C.c_int
    for f in ("sf2d_sim_get_fields", "sf2d_sim_set_fields", "sf2d_sim_get_b_fields", "sf2d_sim_set_b_fields",
              "sf2d_sim_get_acoustic_fields", "sf2d_sim_set_acoustic_fields"):
        getattr(host, f).argtypes = [P, P, P, P]
    host.sf2d_sim_get_rmass_acoustic.argtypes = [P, P]
    host.sf2d_sim_pml_counts.argtypes = [P, P]
    host.sf2d_sim_get_pml.argtypes = [P, P, P, P, P]
    gpu.sf2d_pml_stats.argtypes = [C.POINTER(C.c_long), P]
    host.sf2d_sim_get_coupling.argtypes = [P, P, P, P, P]
    gpu.sf2d_mesh_stream.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_mesh_stream.restype = P
    gpu.sf2d_mesh_affine_elements.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_mesh_affine_elements.restype = C.c_int
    gpu.sf2d_batch_shot_minor.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_batch_shot_minor.restype = C.c_int
    gpu.sf2d_mesh_chain_elems.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_mesh_chain_elems.restype = C.c_int
    gpu.sf2d_mesh_set_shots_per_wave.argtypes = [C.POINTER(C.c_long), C.c_int]
    gpu.sf2d_mesh_rect.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_mesh_rect.restype = C.c_int
    gpu.sf2d_mesh_col_rows.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_mesh_col_rows.restype = C.c_int
    gpu.sf2d_mesh_fluid_chain.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_mesh_fluid_chain.restype = C.c_int
    gpu.sf2d_mesh_chain.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_mesh_chain.restype = C.c_int
    gpu.sf2d_mesh_colours.argtypes = [C.POINTER(C.c_long), C.c_int]
    gpu.sf2d_mesh_colours.restype = C.c_int
    gpu.sf2d_mesh_synchronize.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_mesh_flush.argtypes = [C.POINTER(C.c_long)]
    gpu.sf2d_mesh_field.argtypes = [C.POINTER(C.c_long), C.c_int]
    gpu.sf2d_mesh_field.restype = P
    _lib_cache["gpu"], _lib_cache["host"] = gpu, host
    return gpu, host


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def params_from_config(cfg: SimulationConfig) -> Params:
    p = Params()
    p.nx, p.nz = cfg.nx, cfg.nz
    p.nlayers = len(cfg.nz_layers)
    if p.nlayers > MAXLAYERS:
        raise ValueError("too many layers")
    for i, n in enumerate(cfg.nz_layers):
        p.nz_layer[i] = int(n)
    for i, z in enumerate(cfg.interfaces):
        p.zinterface[i] = float(z)
    p.xmin, p.xmax = float(cfg.xmin), float(cfg.xmax)
    p.absorbbottom, p.absorbright = int(cfg.absorbbottom), int(cfg.absorbright)
    p.absorbtop, p.absorbleft = int(cfg.absorbtop), int(cfg.absorbleft)
    p.stacey = int(cfg.stacey)
    p.nstep, p.dt = int(cfg.nstep), float(cfg.dt)
    p.user_t0, p.anglerec = float(cfg.user_t0), float(cfg.anglerec)
    p.material_mode = int(cfg.material_mode)
    p.seismotype, p.use_trick = int(cfg.seismotype), int(cfg.use_trick)
    p.pml, p.nelem_pml = int(cfg.pml), int(cfg.nelem_pml)
    p.k_min_pml, p.k_max_pml = float(cfg.pml_k_min), float(cfg.pml_k_max)
    p.damping_ac, p.damping_el = float(cfg.pml_damping_acoustic), float(cfg.pml_damping_elastic)
    return p


def _sources_c(sources):
    srcs = (SourceC * max(1, len(sources)))()
    for k, s in enumerate(sources):
        srcs[k].x, srcs[k].z, srcs[k].f0, srcs[k].factor = s.x, s.z, s.f0, s.factor
        srcs[k].anglesource, srcs[k].tshift, srcs[k].time_function_type = s.anglesource, s.tshift, s.time_function_type
        srcs[k].source_type, srcs[k].Mxx, srcs[k].Mzz, srcs[k].Mxz = s.source_type, s.Mxx, s.Mzz, s.Mxz
    return srcs


class Simulation:
    """An elastic simulation driven through the GPU C-ABI.

    ``device=False`` builds only the host setup (usable without a GPU, for parity checks of
    the setup); ``device=True`` also prepares the GPU and enables the time loop."""

    ELEM = {"xix": 0, "xiz": 1, "gammax": 2, "gammaz": 3, "jacobian": 4, "kappa": 5, "mu": 6,
            "rho_vp": 7, "rho_vs": 8, "rho": 9}

    def __init__(self, cfg: SimulationConfig, simulation_type: int = 1, save_forward: bool = False,
                 device: bool = True, chain_elems: int = 0, shots_per_wave: int = 0):
        self.cfg = cfg
        self.gpu, self.host = load_libraries()
        rho, vp, vs = cfg.material_arrays()
        rec = np.ascontiguousarray(np.asarray(cfg.receivers if cfg.nrec else np.zeros((1, 2)), dtype=np.float64))
        p = params_from_config(cfg)
        self.h = self._create(p, rho, vp, vs, rec, int(simulation_type), int(save_forward))
        if not self.h:
            raise ValueError("invalid simulation parameters")
        self.simulation_type = simulation_type
        self.save_forward = save_forward
        self.nspec = self.host.sf2d_sim_nspec(self.h)
        self.nglob = self.host.sf2d_sim_nglob(self.h)
        self.nabs = self.host.sf2d_sim_nabs(self.h)
        self.t0 = self.host.sf2d_sim_t0(self.h)
        self.nspec_acoustic = self.host.sf2d_sim_nspec_acoustic(self.h)
        self.ncoupling = self.host.sf2d_sim_ncoupling(self.h)
        if cfg.attenuation:
            qk, qm = cfg.attenuation_arrays()
            if self.host.sf2d_sim_set_attenuation(self.h, qk, qm, int(cfg.n_sls)) != 0:
                raise NotImplementedError("viscoelastic attenuation: 3 solids, forward runs, quality factors 9999 "
                                          "in fluid elements")
        if cfg.aniso_elem is not None:
            ae = np.ascontiguousarray(np.asarray(cfg.aniso_elem, np.int32).reshape(-1))
            ac = np.ascontiguousarray(np.asarray(cfg.aniso_c, np.float64).T.reshape(-1))  # [6][nspec*25]
            if self.host.sf2d_sim_set_anisotropy(self.h, _ptr(ae), _ptr(ac)) != 0:
                raise NotImplementedError("anisotropic elements: forward runs of solids without attenuation, CPML "
                                          "or shot batches")
        if chain_elems and self.host.sf2d_sim_set_chain_elems(self.h, int(chain_elems)) != 0:
            raise ValueError(f"chain_elems {chain_elems} outside 0..12")
        self.on_device = False
        if device:
            self.prepare_device()
            if shots_per_wave:
                self.set_shots_per_wave(shots_per_wave)

    def _create(self, p, rho, vp, vs, rec, simulation_type, save_forward):
        srcs = _sources_c(self.cfg.sources)
        self._keep = (p, rho, vp, vs, srcs, rec)
        m = self.cfg.mesh
        if m is not None:
            i32 = lambda a: np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=np.int32)  # noqa: E731
            arrs = (i32(m.knods), np.ascontiguousarray(m.coorg, np.float64).reshape(-1), i32(m.free_elem),
                    i32(m.free_code), i32(m.abs_elem), i32(m.abs_code))
            self._keep_mesh = arrs
            return self.host.sf2d_sim_create_external(
                C.byref(p), m.nspec, int(m.coorg.shape[0]), _ptr(arrs[0]), _ptr(arrs[1]), len(m.free_elem),
                _ptr(arrs[2]), _ptr(arrs[3]), len(m.abs_elem), _ptr(arrs[4]), _ptr(arrs[5]), rho, vp, vs,
                len(self.cfg.sources), srcs, self.cfg.nrec, rec, simulation_type, save_forward)
        return self.host.sf2d_sim_create(C.byref(p), rho, vp, vs, len(self.cfg.sources), srcs, self.cfg.nrec, rec,
                                         simulation_type, save_forward)

    def prepare_device(self):
        if self.host.sf2d_sim_prepare_device(self.h) != 0:
            raise RuntimeError("device preparation failed")
        self.on_device = True

    def close(self):
        if getattr(self, "h", None):
            self.host.sf2d_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------
    @property
    def mesh_pointer(self) -> int:
        return self.host.sf2d_sim_mesh_pointer(self.h)

    def stream(self) -> int:
        mp = C.c_long(self.mesh_pointer)
        return self.gpu.sf2d_mesh_stream(C.byref(mp))

    def affine_elements(self) -> int:
        """Elements whose geometry is read as one record (patch assembly)."""
        self._need_device()
        mp = C.c_long(self.mesh_pointer)
        return self.gpu.sf2d_mesh_affine_elements(C.byref(mp))

    def chain_kernel(self) -> bool:
        """Whether the element forces run on a chain kernel (column lanes by default) for chain patches."""
        self._need_device()
        mp = C.c_long(self.mesh_pointer)
        return bool(self.gpu.sf2d_mesh_chain(C.byref(mp)))

    def chain_elems(self) -> int:
        """Elements per chain patch of this handle's plan."""
        self._need_device()
        mp = C.c_long(self.mesh_pointer)
        return self.gpu.sf2d_mesh_chain_elems(C.byref(mp))

    def set_shots_per_wave(self, shots: int):
        """Shots per wave of the batched chain kernel (0: the built-in default)."""
        self._need_device()
        mp = C.c_long(self.mesh_pointer)
        self.gpu.sf2d_mesh_set_shots_per_wave(C.byref(mp), int(shots))

    def rect_kernel(self) -> bool:
        """Whether the chain kernel uses its rectangle form (every element affine and axis-aligned)."""
        self._need_device()
        mp = C.c_long(self.mesh_pointer)
        return bool(self.gpu.sf2d_mesh_rect(C.byref(mp)))

    def fluid_chain(self) -> bool:
        """Whether the fluid's element forces run on the fluid chain kernel (known after the first step)."""
        self._need_device()
        mp = C.c_long(self.mesh_pointer)
        return bool(self.gpu.sf2d_mesh_fluid_chain(C.byref(mp)))

    def col_rows(self) -> bool:
        """Whether the column-lane kernel reads its point ids as one record per chain row."""
        self._need_device()
        mp = C.c_long(self.mesh_pointer)
        return bool(self.gpu.sf2d_mesh_col_rows(C.byref(mp)))

    def colours(self, iphase: int = 2) -> int:
        mp = C.c_long(self.mesh_pointer)
        return self.gpu.sf2d_mesh_colours(C.byref(mp), int(iphase))

    def device_field(self, which: int) -> int:
        mp = C.c_long(self.mesh_pointer)
        return self.gpu.sf2d_mesh_field(C.byref(mp), int(which)) or 0

    # setup products -----------------------------------------------------
    def ibool(self):
        out = np.zeros(self.nspec * NGLL2, np.int32)
        self.host.sf2d_sim_get_ibool(self.h, _ptr(out))
        return out.reshape(self.nspec, NGLL2)

    def rmass_inverse(self):
        out = np.zeros(2 * self.nglob, np.float32)
        self.host.sf2d_sim_get_rmass_inverse(self.h, _ptr(out))
        return out.reshape(self.nglob, 2)

    def elem_array(self, name):
        out = np.zeros(self.nspec * NGLL2, np.float32)
        self.host.sf2d_sim_get_elem_array(self.h, self.ELEM[name], _ptr(out))
        return out.reshape(self.nspec, NGLL2)

    def _source_list(self):
        return self.cfg.sources

    def sources(self):
        """Located sources: (ispec 1-based, xi, gamma, source arrays [nsrc,25,2], time functions [nsrc,nstep])."""
        n = max(1, len(self._source_list()))
        isp, xi, ga = np.zeros(n, np.int32), np.zeros(n), np.zeros(n)
        sa = np.zeros(n * NGLL2 * 2, np.float32)
        stf = np.zeros(self.cfg.nstep * n, np.float32)
        self.host.sf2d_sim_get_sources(self.h, _ptr(isp), _ptr(xi), _ptr(ga), _ptr(sa), _ptr(stf))
        k = len(self._source_list())
        return isp[:k], xi[:k], ga[:k], sa.reshape(n, NGLL2, 2)[:k], stf.reshape(self.cfg.nstep, n)[:, :k].T

    def receivers(self):
        n = max(1, self.cfg.nrec)
        isp, xi, ga = np.zeros(n, np.int32), np.zeros(n), np.zeros(n)
        xir, gar = np.zeros(5 * n, np.float32), np.zeros(5 * n, np.float32)
        self.host.sf2d_sim_get_receivers(self.h, _ptr(isp), _ptr(xi), _ptr(ga), _ptr(xir), _ptr(gar))
        k = self.cfg.nrec
        return isp[:k], xi[:k], ga[:k], xir.reshape(n, 5)[:k], gar.reshape(n, 5)[:k]

    def stacey(self):
        n = max(1, self.nabs)
        isp, ij = np.zeros(n, np.int32), np.zeros(n * 10, np.int32)
        nrm, jw = np.zeros(n * 10, np.float32), np.zeros(n * 5, np.float32)
        edge, ib = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.host.sf2d_sim_get_stacey(self.h, _ptr(isp), _ptr(ij), _ptr(nrm), _ptr(jw), _ptr(edge), _ptr(ib))
        k = self.nabs
        return isp[:k], ij.reshape(n, 5, 2)[:k], nrm.reshape(n, 5, 2)[:k], jw.reshape(n, 5)[:k], edge[:k], ib[:k]

    def attenuation_coefficients(self):
        """(A_kappa, B_kappa, A_mu, B_mu), each [nspec, 25, 3] float32, as handed to the device."""
        out = [np.zeros((self.nspec, 25, 3), np.float32) for _ in range(4)]
        self.host.sf2d_sim_get_attenuation(self.h, *[o.ctypes.data for o in out])
        return tuple(out)

    # time loop ----------------------------------------------------------
    def _need_device(self):
        if not self.on_device:
            raise RuntimeError("simulation is not prepared on a device")

    def reset(self):
        """Zero the wavefields and restart at step 1."""
        self._need_device()
        self.host.sf2d_sim_reset(self.h)

    @property
    def steps_done(self) -> int:
        return self.host.sf2d_sim_steps_done(self.h)

    def forward(self, nstep: Optional[int] = None, save_stacey: bool = False, phase_timing: bool = False):
        """Runs the next ``nstep`` steps of the GPU time loop (default: the remaining ones);
        returns (seismograms [nrec,2,nstep], timing dict[, stacey storage])."""
        self._need_device()
        n = (self.cfg.nstep - self.steps_done) if nstep is None else int(nstep)
        seis = np.zeros((max(1, self.cfg.nrec), 2, n))
        store = np.zeros((n, max(1, self.nabs), 5, 2), np.float32) if save_stacey else None
        timing = np.zeros(6)
        if phase_timing:
            timing[0] = -1.0
        rc = self.host.sf2d_sim_forward(self.h, n, _ptr(seis), _ptr(store), _ptr(timing))
        if rc != 0:
            raise RuntimeError("sf2d_sim_forward failed")
        t = dict(zip(("loop_ms", "predictor_ms", "forces_ms", "boundary_sources_ms", "corrector_ms",
                      "seismograms_ms"), timing.tolist()))
        seis = seis[: self.cfg.nrec]
        return (seis, t, store) if save_stacey else (seis, t)

    def adjoint(self, adj, absorb, lastframe):
        """SIMULATION_TYPE=3 run on the GPU (the simulation must be created with simulation_type=3).
        ``adj`` [nrec,2,nstep] adjoint source traces, ``absorb`` the Stacey storage and ``lastframe``
        (displ, veloc, accel) of a save_forward run.  Returns (adjoint seismograms [nrec,2,nstep],
        kernels dict rho/mu/kappa of [nspec,25] arrays)."""
        self._need_device()
        self._solid_only("adjoint")
        if self.simulation_type != 3:
            raise RuntimeError("create the simulation with simulation_type=3 for an adjoint run")
        n = self.cfg.nstep
        a = np.ascontiguousarray(np.asarray(adj, np.float32).reshape(-1))
        st = np.ascontiguousarray(np.asarray(absorb, np.float32).reshape(-1)) if self.nabs else np.zeros(1, np.float32)
        d, v, ac = (np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1)) for x in lastframe)
        seis = np.zeros((max(1, self.cfg.nrec), 2, n))
        ker = np.zeros(3 * self.nspec * NGLL2, np.float32)
        rc = self.host.sf2d_sim_adjoint(self.h, _ptr(a), _ptr(st), _ptr(d), _ptr(v), _ptr(ac), _ptr(seis), _ptr(ker))
        if rc != 0:
            raise RuntimeError("sf2d_sim_adjoint failed")
        ker = ker.reshape(3, self.nspec, NGLL2)
        return seis[: self.cfg.nrec], {"rho": ker[0], "mu": ker[1], "kappa": ker[2]}

    def adjoint_full(self, adj, absorb, lastframe_el=None, lastframe_ac=None):
        """SIMULATION_TYPE=3 run over fluid, solid or coupled meshes (``sf2d_sim_adjoint_full``).  ``adj``
        [nrec,2,nstep] adjoint source traces (a receiver in a fluid element uses component 0), ``absorb`` the
        Stacey storage of a save_forward run, ``lastframe_el`` (displ, veloc, accel) and ``lastframe_ac``
        (potential and its two time derivatives) at its end.  Returns (adjoint seismograms [nrec,2,nstep], solid
        kernels dict rho/mu/kappa, fluid kernels dict rho/kappa/rhop/alpha), kernels [nspec,25]."""
        self._need_device()
        if self.simulation_type != 3:
            raise RuntimeError("create the simulation with simulation_type=3 for an adjoint run")
        n = self.cfg.nstep
        f32 = lambda x: np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1))  # noqa: E731
        a = f32(adj)
        st = f32(absorb) if self.nabs else np.zeros(1, np.float32)
        el = [None] * 3 if lastframe_el is None else [f32(x) for x in lastframe_el]
        ac = [None] * 3 if lastframe_ac is None else [f32(x) for x in lastframe_ac]
        seis = np.zeros((max(1, self.cfg.nrec), 2, n))
        kel = np.zeros(3 * self.nspec * NGLL2, np.float32)
        kac = np.zeros(4 * self.nspec * NGLL2, np.float32)
        rc = self.host.sf2d_sim_adjoint_full(self.h, _ptr(a), _ptr(st), *[_ptr(x) for x in el + ac], _ptr(seis),
                                             _ptr(kel), _ptr(kac))
        if rc != 0:
            raise RuntimeError("sf2d_sim_adjoint_full failed (CPML, or a missing last frame)")
        kel = kel.reshape(3, self.nspec, NGLL2)
        kac = kac.reshape(4, self.nspec, NGLL2)
        return (seis[: self.cfg.nrec], dict(zip(("rho", "mu", "kappa"), kel)),
                dict(zip(("rho", "kappa", "rhop", "alpha"), kac)))

    # CPML layer ----------------------------------------------------------
    def pml(self):
        """The driver's CPML set-up: dict with is_pml, region [nspec]; kx, kz, dx, dz, ax, az [nspec_pml, 25]
        (double, layer elements in element order); the outer-boundary points of elastic and acoustic layer elements
        (0-based, increasing)."""
        counts = np.zeros(3, np.int32)
        self.host.sf2d_sim_pml_counts(self.h, _ptr(counts))
        npml, nel, nac = (int(c) for c in counts)
        reg = np.zeros(self.nspec, np.int32)
        prof = np.zeros(6 * max(1, npml) * NGLL2)
        pe, pa = np.zeros(max(1, nel), np.int32), np.zeros(max(1, nac), np.int32)
        self.host.sf2d_sim_get_pml(self.h, _ptr(reg), _ptr(prof), _ptr(pe), _ptr(pa))
        prof = prof[: 6 * npml * NGLL2].reshape(6, npml, NGLL2)
        out = {"region": reg, "points_elastic": pe[:nel] - 1, "points_acoustic": pa[:nac] - 1, "nspec_pml": npml}
        out.update(zip(("kx", "kz", "dx", "dz", "ax", "az"), prof))
        return out

    def pml_stats(self):
        """The GPU library's CPML layer: (solid layer elements, fluid layer elements, distinct coefficient sets,
        solid outer-boundary points, fluid outer-boundary points)."""
        self._need_device()
        out = np.zeros(5, np.int32)
        mp = C.c_long(self.mesh_pointer)
        self.gpu.sf2d_pml_stats(C.byref(mp), _ptr(out))
        return tuple(int(v) for v in out)

    # fluid elements ------------------------------------------------------
    def rmass_acoustic(self):
        out = np.zeros(self.nglob, np.float32)
        self.host.sf2d_sim_get_rmass_acoustic(self.h, _ptr(out))
        return out

    def coupling(self):
        """Fluid-solid edges: (fluid element, side, solid element, side), 0-based elements, sides 1 bottom,
        2 right, 3 top, 4 left."""
        n = max(1, self.ncoupling)
        arrs = [np.zeros(n, np.int32) for _ in range(4)]
        self.host.sf2d_sim_get_coupling(self.h, *[_ptr(a) for a in arrs])
        out = np.stack(arrs, axis=1)[: self.ncoupling]
        out[:, 0] -= 1
        out[:, 2] -= 1
        return out

    def acoustic_fields(self):
        """(potential, its first and second time derivatives) [nglob] of the fluid elements."""
        self._need_device()
        if not self.nspec_acoustic:
            raise RuntimeError("no fluid element in this mesh")
        p, v, a = (np.zeros(self.nglob, np.float32) for _ in range(3))
        self.host.sf2d_sim_get_acoustic_fields(self.h, _ptr(p), _ptr(v), _ptr(a))
        return p, v, a

    def set_acoustic_fields(self, pot=None, pot_dot=None, pot_dot_dot=None):
        self._need_device()
        arrs = [None if x is None else np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1))
                for x in (pot, pot_dot, pot_dot_dot)]
        self.host.sf2d_sim_set_acoustic_fields(self.h, *[_ptr(x) for x in arrs])

    def _solid_only(self, what):
        if self.nspec_acoustic:
            raise NotImplementedError(f"{what}: meshes with fluid elements support forward runs only")

    # model and gradient ------------------------------------------------
    def set_model(self, rho, vp, vs):
        """Replace the MODEL=gll parameters (per GLL point, [nspec,25] or flat); moduli, impedances and the
        mass matrix are recomputed and, on a prepared device, uploaded."""
        arrs = [np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1)) for x in (rho, vp, vs)]
        if any(a.size != self.nspec * NGLL2 for a in arrs):
            raise ValueError("model arrays must hold nspec*25 values")
        rc = self.host.sf2d_sim_set_model(self.h, *[_ptr(a) for a in arrs])
        if rc != 0:
            raise ValueError("invalid model (rho, vp must be > 0, vs >= 0)")
        self._keep_model = arrs

    def gradient_forward(self, obs, chunk: int = 100):
        """Forward run from rest over all NSTEP steps with a checkpoint every ``chunk`` steps kept on the
        device; returns (chi, seismograms) with chi = 1/2 dt sum |s - obs|^2 (obs [nrec,2,nstep])."""
        self._need_device()
        self._solid_only("gradient")
        n, shape = self.cfg.nstep, self._seis_shape()
        o = np.ascontiguousarray(np.asarray(obs, np.float64).reshape(-1))
        if o.size != int(np.prod(shape)):
            raise ValueError(f"obs must be {list(shape)} ([nshot,] nrec, 2, nstep)")
        chi = C.c_double(0.0)
        seis = np.zeros(int(np.prod(shape)) + 2 * n)
        rc = self.host.sf2d_sim_gradient_forward(self.h, _ptr(o), int(chunk), C.byref(chi), _ptr(seis))
        if rc != 0:
            raise RuntimeError("sf2d_sim_gradient_forward failed")
        return chi.value, seis[: int(np.prod(shape))].reshape(shape)

    def _seis_shape(self):
        return (self.cfg.nrec, 2, self.cfg.nstep)

    def gradient_backward(self, parts: bool = False):
        """Reverse sweep after gradient_forward: d chi / d (rho, vp, vs) as [nspec,25] arrays (dict), and with
        ``parts`` also d chi / d (kappa, mu, rho via mass, rho vp, rho vs)."""
        self._need_device()
        ne = self.nspec * NGLL2
        g = np.zeros(3 * ne)
        pp = np.zeros(5 * ne) if parts else None
        rc = self.host.sf2d_sim_gradient_backward(self.h, _ptr(g), _ptr(pp))
        if rc != 0:
            raise RuntimeError("sf2d_sim_gradient_backward failed (call gradient_forward first)")
        g = g.reshape(3, self.nspec, NGLL2)
        out = {"rho": g[0], "vp": g[1], "vs": g[2]}
        if parts:
            pp = pp.reshape(5, self.nspec, NGLL2)
            return out, dict(zip(("kappa", "mu", "rho_mass", "rho_vp", "rho_vs"), pp))
        return out

    def fields(self):
        self._need_device()
        d, v, a = (np.zeros(2 * self.nglob, np.float32) for _ in range(3))
        self.host.sf2d_sim_get_fields(self.h, _ptr(d), _ptr(v), _ptr(a))
        return d.reshape(-1, 2), v.reshape(-1, 2), a.reshape(-1, 2)

    def set_fields(self, displ=None, veloc=None, accel=None):
        self._need_device()
        arrs = [None if x is None else np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1))
                for x in (displ, veloc, accel)]
        self.host.sf2d_sim_set_fields(self.h, *[_ptr(x) for x in arrs])


class ShotBatch(Simulation):
    """``len(shots)`` forward simulations of one mesh, model and receiver set advanced together on the GPU:
    shot k is driven by the sources ``shots[k]`` (a list of config.Source), timed as a simulation of its own
    (``cfg.sources`` is ignored).  ``forward()`` returns seismograms [nshot, nrec, 2, nstep]; the gradient
    calls take observations of that shape and return the misfit summed over the shots and its gradient."""

    def __init__(self, cfg: SimulationConfig, shots, device: bool = True, chain_elems: int = 0,
                 shots_per_wave: int = 0):
        self.shots = [list(s) for s in shots]
        if not self.shots or any(len(s) == 0 for s in self.shots):
            raise ValueError("every shot needs at least one source")
        if np.any(np.asarray(cfg.material_arrays()[2]) == 0.0):
            raise NotImplementedError("shot batches of meshes with fluid elements (vs = 0) are not supported")
        if cfg.mesh is not None:
            raise NotImplementedError("shot batches of external meshes are not supported")
        self.nshot = len(self.shots)
        super().__init__(cfg, simulation_type=1, save_forward=False, device=device, chain_elems=chain_elems,
                         shots_per_wave=shots_per_wave)

    def _create(self, p, rho, vp, vs, rec, simulation_type, save_forward):
        flat = [s for shot in self.shots for s in shot]
        srcs = _sources_c(flat)
        counts = np.asarray([len(s) for s in self.shots], np.int32)
        self._keep = (p, rho, vp, vs, srcs, rec, counts)
        return self.host.sf2d_sim_create_shots(C.byref(p), rho, vp, vs, len(self.shots), _ptr(counts), srcs,
                                               self.cfg.nrec, rec)

    def _seis_shape(self):
        return (self.nshot, self.cfg.nrec, 2, self.cfg.nstep)

    def _source_list(self):
        return [s for shot in self.shots for s in shot]

    def shot_minor(self) -> bool:
        """Whether the batch fields are shot-minor (the lane-per-shot force kernel)."""
        self._need_device()
        mp = C.c_long(self.mesh_pointer)
        return self.gpu.sf2d_batch_shot_minor(C.byref(mp)) == 1

    def shot_t0(self):
        out = np.zeros(self.nshot)
        self.host.sf2d_sim_shot_t0(self.h, _ptr(out))
        return out

    def forward(self, timing: bool = False):
        """All NSTEP steps of every shot from rest: seismograms [nshot, nrec, 2, nstep] (and device ms)."""
        self._need_device()
        seis = np.zeros(int(np.prod(self._seis_shape())) + 1)
        t = np.zeros(1)
        if self.host.sf2d_sim_batch_forward(self.h, _ptr(seis), _ptr(t) if timing else None) != 0:
            raise RuntimeError("sf2d_sim_batch_forward failed")
        seis = seis[:-1].reshape(self._seis_shape())
        return (seis, float(t[0])) if timing else seis

    def shot_misfits(self):
        """Per-shot misfits of the last gradient_forward."""
        out = np.zeros(self.nshot)
        self.host.sf2d_sim_shot_misfits(self.h, _ptr(out))
        return out

    def fields(self, shot: int = 0):
        self._need_device()
        if not 0 <= shot < self.nshot:
            raise IndexError(shot)
        d, v, a = (np.zeros(2 * self.nglob, np.float32) for _ in range(3))
        self.host.sf2d_sim_batch_get_fields(self.h, int(shot), _ptr(d), _ptr(v), _ptr(a))
        return d.reshape(-1, 2), v.reshape(-1, 2), a.reshape(-1, 2)

    def adjoint(self, *a, **k):
        raise RuntimeError("SIMULATION_TYPE=3 runs take one simulation; use Simulation(simulation_type=3)")


class UnevenBatch(Simulation):
    """An uneven batch: ``len(cfgs)`` simulations of different meshes, sizes and physics (fluid or viscoelastic
    solids, CPML or Stacey edges, different NSTEP) advanced together on the GPU, one launch per kernel covering every
    instance.  Their time step and recording kind must agree.  The union runs the longest NSTEP; ``forward()``
    returns every instance's seismograms over its own NSTEP (a list of [nrec_k, 2, nstep_k] arrays).  The union
    numbers the points in instance order and the solid's Newmark updates stop at the last point of a solid element,
    so instances with solid elements placed first spare those updates over the points of all-fluid instances after
    them (``UnevenRuns`` orders them so).  No counterpart in the reference, which runs one mesh per simulation."""

    def __init__(self, cfgs, device: bool = True):
        cfgs = list(cfgs)
        if not cfgs:
            raise ValueError("an uneven batch needs at least one instance")
        nstep = max(int(c.nstep) for c in cfgs)
        if len({float(c.dt) for c in cfgs}) != 1 or len({int(c.seismotype) for c in cfgs}) != 1:
            raise ValueError("the instances of an uneven batch share DT and seismotype")
        self.instance_cfgs = cfgs
        self.instance_nstep = [int(c.nstep) for c in cfgs]
        # the instances are set up with the union's NSTEP (their time functions cover it), on the host only
        self.parts = [Simulation(c.replace(nstep=nstep), device=False) for c in cfgs]
        self.instance_nrec = [c.nrec for c in cfgs]
        self.instance_nglob = [p.nglob for p in self.parts]
        self.instance_nspec = [p.nspec for p in self.parts]
        # the union's own configuration carries the sources and receivers of all instances; its materials and
        # quality factors come from the instances' set-ups, which the union handle copies
        self._union_cfg = cfgs[0].replace(
            nstep=nstep, sources=[s for c in cfgs for s in c.sources], rho=1.0, vp=1.0, vs=0.0, attenuation=False,
            mesh=None, receivers=np.concatenate(
                [np.asarray(c.receivers, np.float64).reshape(-1, 2) for c in cfgs if c.nrec] or [np.zeros((0, 2))]))
        super().__init__(self._union_cfg, simulation_type=1, save_forward=False, device=False)
        for q in self.parts:  # the union holds copies of everything it needs
            q.close()
        self.parts = []
        if device:
            self.prepare_device()

    def _create(self, p, rho, vp, vs, rec, simulation_type, save_forward):
        arr = (C.c_void_p * len(self.parts))(*[q.h for q in self.parts])
        self._keep = (arr,)
        return self.host.sf2d_sim_create_union(len(self.parts), arr)

    def forward(self, nstep: Optional[int] = None, save_stacey: bool = False, phase_timing: bool = False):
        """All NSTEP steps of the union from rest: (per-instance seismograms, timing)."""
        if nstep is not None or save_stacey:
            raise ValueError("an uneven batch runs whole: forward() from rest")
        self.reset()
        seis, t = super().forward(phase_timing=phase_timing)
        out, r0 = [], 0
        for nr, ns in zip(self.instance_nrec, self.instance_nstep):
            out.append(seis[r0:r0 + nr, :, :ns])
            r0 += nr
        return out, t

    def useful_updates(self) -> int:
        """Sum over the instances of GLL points x their own NSTEP."""
        return int(sum(g * n for g, n in zip(self.instance_nglob, self.instance_nstep)))

    def close(self):
        super().close()
        for q in getattr(self, "parts", []):
            q.close()
        self.parts = []

    def attenuation_coefficients(self):
        raise NotImplementedError("per instance: use the instances' own simulations")


# relative cost per GLL point of one step: a viscoelastic solid point (forces with 3 solids' memory variables) against
# a fluid or plain elastic one, measured on the 48-instance uneven batch (11.8 vs 21.5 G points/s)
VISCO_POINT_COST = 1.8
# fixed cost of one step of one union (its kernel launches), in units of one fluid point
UNION_STEP_COST = 2.0e6


def _all_fluid(cfg) -> bool:
    """Whether every GLL point of an internal-mesher configuration has vs = 0 (a mesh of fluid elements only)."""
    return bool(np.all(np.asarray(cfg.vs, dtype=np.float64) == 0.0))


def plan_nstep_groups(cfgs, max_groups: int = 6, step_cost: float = UNION_STEP_COST):
    """Split uneven-batch instances into at most ``max_groups`` unions of similar NSTEP: contiguous runs of the
    instances in NSTEP order, chosen by dynamic programming to minimise the modelled work
    sum over groups of NSTEP_g x (step_cost + sum of the group's point costs).  A union steps every instance until
    its longest one ends, so one union of instances with NSTEP 1000..4000 spends about a third of its updates on
    finished instances; a few groups cut that to a few percent.  Returns lists of indices into ``cfgs``."""
    order = sorted(range(len(cfgs)), key=lambda k: int(cfgs[k].nstep))
    w = [cfgs[k].nglob * (VISCO_POINT_COST if cfgs[k].attenuation else 1.0) for k in order]
    ns = [int(cfgs[k].nstep) for k in order]
    n = len(order)
    pre = np.concatenate([[0.0], np.cumsum(w)])
    inf = float("inf")
    # best[g][j]: least cost of the first j instances in g groups
    best = np.full((max_groups + 1, n + 1), inf)
    cut = np.zeros((max_groups + 1, n + 1), np.int64)
    best[0][0] = 0.0
    for g in range(1, max_groups + 1):
        for j in range(1, n + 1):
            for i in range(g - 1, j):
                c = best[g - 1][i] + ns[j - 1] * (step_cost + pre[j] - pre[i])
                if c < best[g][j]:
                    best[g][j], cut[g][j] = c, i
    g = int(np.argmin(best[1:, n])) + 1
    groups, j = [], n
    while g > 0:
        i = int(cut[g][j])
        groups.append([order[k] for k in range(i, j)])
        j, g = i, g - 1
    return groups[::-1]


class UnevenRuns:
    """An uneven batch run as a few unions of instances of similar NSTEP (``plan_nstep_groups``), one after the
    other: ``forward()`` returns every instance's seismograms, in the order of ``cfgs``, over its own NSTEP.  With
    ``groups=1`` it is one ``UnevenBatch``."""

    def __init__(self, cfgs, max_groups: int = 6, device: bool = True, step_cost: float = UNION_STEP_COST):
        self.cfgs = list(cfgs)
        # inside each union the instances with solid elements first (see UnevenBatch)
        self.groups = [sorted(g, key=lambda k: _all_fluid(self.cfgs[k]))
                       for g in plan_nstep_groups(self.cfgs, max_groups, step_cost)]
        self.batches = []
        try:
            for idx in self.groups:
                self.batches.append(UnevenBatch([self.cfgs[k] for k in idx], device=device))
        except BaseException:
            self.close()
            raise

    def forward(self, phase_timing: bool = False):
        out = [None] * len(self.cfgs)
        timing = {"loop_ms": 0.0, "groups": []}
        for idx, b in zip(self.groups, self.batches):
            seis, t = b.forward(phase_timing=phase_timing)
            for k, s in zip(idx, seis):
                out[k] = s
            timing["loop_ms"] += t["loop_ms"]
            timing["groups"].append(t)
        return out, timing

    def useful_updates(self) -> int:
        return int(sum(b.useful_updates() for b in self.batches))

    def union_updates(self) -> int:
        """Updates the unions make, finished instances included."""
        return int(sum(b.nglob * b.cfg.nstep for b in self.batches))

    def close(self):
        for b in self.batches:
            b.close()
        self.batches = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
