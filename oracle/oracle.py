"""ctypes wrapper of the CPU oracle (test infrastructure only).

Imported by ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py``; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MAXLAYERS = 32
NGLL = 5
NGLL2 = 25


class OrcParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("nz", C.c_int), ("nlayers", C.c_int), ("nz_layer", C.c_int * MAXLAYERS),
        ("xmin", C.c_double), ("xmax", C.c_double), ("zinterface", C.c_double * (MAXLAYERS + 1)),
        ("absorb_bottom", C.c_int), ("absorb_right", C.c_int), ("absorb_top", C.c_int),
        ("absorb_left", C.c_int), ("stacey", C.c_int), ("nstep", C.c_int), ("dt", C.c_double),
        ("user_t0", C.c_double), ("anglerec_deg", C.c_double), ("material_mode", C.c_int),
        ("seismotype", C.c_int), ("use_trick", C.c_int), ("pml", C.c_int), ("nelem_pml", C.c_int),
        ("k_min_pml", C.c_double), ("k_max_pml", C.c_double), ("damping_ac", C.c_double),
        ("damping_el", C.c_double),
    ]


class OrcSource(C.Structure):
    _fields_ = [
        ("x", C.c_double), ("z", C.c_double), ("f0", C.c_double), ("factor", C.c_double),
        ("angle_deg", C.c_double), ("tshift", C.c_double), ("time_function_type", C.c_int),
        ("source_type", C.c_int), ("Mxx", C.c_double), ("Mzz", C.c_double), ("Mxz", C.c_double),
    ]


_LIBS = {}


def build():
    """Compile the oracle shared libraries (gcc, no GPU)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(precision: str):
    if precision not in _LIBS:
        path = os.path.join(_HERE, f"liboracle_{precision}.so")
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        P = C.c_void_p
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        lib.orc_new.restype = P
        lib.orc_new.argtypes = [C.POINTER(OrcParams), dp, dp, dp, C.c_int, C.POINTER(OrcSource), C.c_int, dp]
        lib.orc_new_external.restype = P
        lib.orc_new_external.argtypes = [C.POINTER(OrcParams), C.c_int, C.c_int, ip, dp, C.c_int, ip, ip, C.c_int,
                                         ip, ip, dp, dp, dp, C.c_int, C.POINTER(OrcSource), C.c_int, dp]
        lib.orc_free.argtypes = [P]
        for f in ("orc_nspec", "orc_nglob", "orc_nabs"):
            getattr(lib, f).argtypes = [P]
            getattr(lib, f).restype = C.c_int
        lib.orc_t0.argtypes = [P]
        lib.orc_t0.restype = C.c_double
        lib.orc_get_ibool.argtypes = [P, ip]
        lib.orc_get_coord.argtypes = [P, dp]
        lib.orc_get_elem_array.argtypes = [P, C.c_int, dp]
        lib.orc_get_rmass_inverse.argtypes = [P, dp]
        lib.orc_get_basis.argtypes = [P, dp, dp, dp, dp]
        lib.orc_get_absorbing.argtypes = [P, ip, ip, ip, ip]
        lib.orc_get_sources.argtypes = [P, ip, dp, dp, dp, dp]
        lib.orc_get_receivers.argtypes = [P, ip, dp, dp, dp, dp]
        lib.orc_forward.argtypes = [P, C.c_int, C.c_void_p, C.c_void_p]
        lib.orc_get_fields.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_set_fields.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_apply_stiffness.argtypes = [P, dp, dp]
        lib.orc_real_size.restype = C.c_int
        lib.orc_set_ftz.argtypes = [P, C.c_int]
        lib.orc_set_threads.restype = C.c_int
        lib.orc_set_threads.argtypes = [C.c_int]
        lib.orc_adjoint.argtypes = [P, dp, dp, dp, dp, dp, C.c_void_p, dp, C.c_int, C.c_void_p, C.c_void_p]
        lib.orc_set_anisotropy.argtypes = [P, ip, dp]
        lib.orc_set_anisotropy.restype = C.c_int
        lib.orc_adjoint_full.argtypes = [P, dp, dp, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_adjoint_full.restype = C.c_int
        lib.orc_kernel_increment.argtypes = [P, dp, dp, dp, dp]
        lib.orc_kernel_ac_increment.argtypes = [P, dp, dp, dp, C.c_double, dp]
        lib.orc_gradient.argtypes = [P, dp, dp, dp]
        lib.orc_gradient.restype = C.c_double
        lib.orc_set_attenuation.argtypes = [P, dp, dp, C.c_int, C.c_double]
        lib.orc_set_attenuation.restype = C.c_int
        lib.orc_get_attenuation.argtypes = [P, dp, dp, dp, dp]
        for f in ("orc_nspec_acoustic", "orc_ncoupling", "orc_nfree"):
            getattr(lib, f).argtypes = [P]
            getattr(lib, f).restype = C.c_int
        lib.orc_get_is_acoustic.argtypes = [P, ip]
        lib.orc_get_rmass_acoustic.argtypes = [P, dp]
        lib.orc_get_coupling.argtypes = [P, ip, ip, ip, ip]
        lib.orc_get_free_surface.argtypes = [P, ip, ip]
        lib.orc_get_acoustic_stacey_range.argtypes = [P, ip, ip]
        lib.orc_get_acoustic_fields.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_set_acoustic_fields.argtypes = [P, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_apply_acoustic_stiffness.argtypes = [P, dp, dp]
        lib.orc_nspec_pml.argtypes = [P]
        lib.orc_nspec_pml.restype = C.c_int
        lib.orc_npml_points.argtypes = [P, C.c_int]
        lib.orc_npml_points.restype = C.c_int
        lib.orc_get_pml.argtypes = [P, ip, ip, dp, ip, ip]
        lib.orc_set_pml_mistake.argtypes = [C.c_int]
        lib.orc_pml_coefficients.argtypes = [C.c_double] * 7 + [C.c_int, dp]
        _LIBS[precision] = lib
    return _LIBS[precision]


def set_threads(n: int, precision: str = "sp_omp") -> int:
    """Threads of the OpenMP build (the CPU baseline); returns the count in effect (1 for serial builds)."""
    return int(_lib(precision).orc_set_threads(int(n)))


def pml_coefficients(dt, kx, dx, ax, kz, dz, az, region, precision="dp"):
    """The oracle's CPML coefficients of one point: dict with lik31 / lik13 (A0..A2, (c0, c1, c2) for bb_1 and bb_2)
    and l (A0..A4, (c0, c1, c2) for alpha_x and alpha_z)."""
    out = np.zeros(29)
    _lib(precision).orc_pml_coefficients(dt, kx, dx, ax, kz, dz, az, int(region), out)
    return {"lik31": (out[0:3], out[3:6], out[6:9]), "lik13": (out[9:12], out[12:15], out[15:18]),
            "l": (out[18:23], out[23:26], out[26:29])}


def set_pml_mistake(which, precision="sp"):
    """Plant a mistake in the oracle's CPML terms (test infrastructure; 0 removes it)."""
    _lib(precision).orc_set_pml_mistake(int(which))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleSim:
    """One simulation set up by the oracle from a :class:`SimulationConfig`."""

    ELEM_ARRAYS = ("xix", "xiz", "gammax", "gammaz", "jacobian", "rho", "mu", "kappa", "rho_vp", "rho_vs")

    def __init__(self, cfg, precision: str = "sp", ftz: bool = True):
        self.cfg = cfg
        self.lib = _lib(precision)
        p = OrcParams()
        p.nx, p.nz = cfg.nx, cfg.nz
        p.nlayers = len(cfg.nz_layers)
        for i, n in enumerate(cfg.nz_layers):
            p.nz_layer[i] = int(n)
        for i, z in enumerate(cfg.interfaces):
            p.zinterface[i] = float(z)
        p.xmin, p.xmax = cfg.xmin, cfg.xmax
        p.absorb_bottom, p.absorb_right = int(cfg.absorbbottom), int(cfg.absorbright)
        p.absorb_top, p.absorb_left = int(cfg.absorbtop), int(cfg.absorbleft)
        p.stacey = int(cfg.stacey)
        p.nstep, p.dt = cfg.nstep, cfg.dt
        p.user_t0, p.anglerec_deg = cfg.user_t0, cfg.anglerec
        p.material_mode = cfg.material_mode
        p.seismotype, p.use_trick = int(cfg.seismotype), int(cfg.use_trick)
        p.pml, p.nelem_pml = int(cfg.pml), int(cfg.nelem_pml)
        p.k_min_pml, p.k_max_pml = float(cfg.pml_k_min), float(cfg.pml_k_max)
        p.damping_ac, p.damping_el = float(cfg.pml_damping_acoustic), float(cfg.pml_damping_elastic)
        rho, vp, vs = cfg.material_arrays()
        srcs = (OrcSource * max(1, len(cfg.sources)))()
        for k, s in enumerate(cfg.sources):
            srcs[k].x, srcs[k].z, srcs[k].f0, srcs[k].factor = s.x, s.z, s.f0, s.factor
            srcs[k].angle_deg, srcs[k].tshift, srcs[k].time_function_type = s.anglesource, s.tshift, s.time_function_type
            srcs[k].source_type, srcs[k].Mxx, srcs[k].Mzz, srcs[k].Mxz = s.source_type, s.Mxx, s.Mzz, s.Mxz
        rec = np.ascontiguousarray(np.asarray(cfg.receivers if cfg.nrec else np.zeros((1, 2)), dtype=np.float64))
        self._keep = (p, rho, vp, vs, srcs, rec)
        if cfg.mesh is None:
            self.h = self.lib.orc_new(C.byref(p), rho, vp, vs, len(cfg.sources), srcs, cfg.nrec, rec)
        else:
            m = cfg.mesh
            i32 = lambda a: np.ascontiguousarray(np.asarray(a).reshape(-1), dtype=np.int32)  # noqa: E731
            arrs = (i32(m.knods), np.ascontiguousarray(m.coorg, np.float64).reshape(-1), i32(m.free_elem),
                    i32(m.free_code), i32(m.abs_elem), i32(m.abs_code))
            self._keep_mesh = arrs
            self.h = self.lib.orc_new_external(C.byref(p), m.nspec, int(m.coorg.shape[0]), arrs[0], arrs[1],
                                               len(m.free_elem), arrs[2], arrs[3], len(m.abs_elem), arrs[4], arrs[5],
                                               rho, vp, vs, len(cfg.sources), srcs, cfg.nrec, rec)
            if not self.h:
                raise NotImplementedError("oracle: external meshes with a CPML layer")
        self.nspec = self.lib.orc_nspec(self.h)
        self.nglob = self.lib.orc_nglob(self.h)
        self.nabs = self.lib.orc_nabs(self.h)
        self.t0 = self.lib.orc_t0(self.h)
        self.nspec_acoustic = self.lib.orc_nspec_acoustic(self.h)
        self.ncoupling = self.lib.orc_ncoupling(self.h)
        self.lib.orc_set_ftz(self.h, int(ftz))
        if cfg.aniso_elem is not None:
            ae = np.ascontiguousarray(np.asarray(cfg.aniso_elem, np.int32).reshape(-1))
            ac = np.ascontiguousarray(np.asarray(cfg.aniso_c, np.float64).T.reshape(-1))  # [6][nspec*25]
            self._keep_aniso = (ae, ac)
            if self.lib.orc_set_anisotropy(self.h, ae, ac) != 0:
                raise NotImplementedError("oracle: anisotropic elements with attenuation or CPML")
        if cfg.attenuation:
            qk, qm = cfg.attenuation_arrays()
            self._keep_q = (qk, qm)
            if self.lib.orc_set_attenuation(self.h, qk, qm, int(cfg.n_sls), cfg.attenuation_f0()) != 0:
                raise ValueError("oracle: invalid attenuation set-up")

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.orc_free(self.h)
                self.h = None
        except Exception:
            pass

    # setup products --------------------------------------------------
    def ibool(self):
        out = np.zeros(self.nspec * NGLL2, dtype=np.int32)
        self.lib.orc_get_ibool(self.h, out)
        return out.reshape(self.nspec, NGLL2)

    def coord(self):
        out = np.zeros(2 * self.nglob)
        self.lib.orc_get_coord(self.h, out)
        return out.reshape(self.nglob, 2)

    def elem_array(self, name):
        out = np.zeros(self.nspec * NGLL2)
        self.lib.orc_get_elem_array(self.h, self.ELEM_ARRAYS.index(name), out)
        return out.reshape(self.nspec, NGLL2)

    def rmass_inverse(self):
        out = np.zeros(2 * self.nglob)
        self.lib.orc_get_rmass_inverse(self.h, out)
        return out.reshape(self.nglob, 2)

    def basis(self):
        xi, w, hp, hpw = np.zeros(5), np.zeros(5), np.zeros(25), np.zeros(25)
        self.lib.orc_get_basis(self.h, xi, w, hp, hpw)
        return xi, w, hp.reshape(5, 5), hpw.reshape(5, 5)

    def absorbing(self):
        n = max(1, self.nabs)
        isp, code, corner, ib = (np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(4 * n, np.int32),
                                 np.zeros(n, np.int32))
        self.lib.orc_get_absorbing(self.h, isp, code, corner, ib)
        k = self.nabs
        return isp[:k], code[:k], corner.reshape(n, 4)[:k], ib[:k]

    def sources(self):
        n = max(1, len(self.cfg.sources))
        isp, xi, ga = np.zeros(n, np.int32), np.zeros(n), np.zeros(n)
        sa = np.zeros(n * NGLL2 * 2)
        stf = np.zeros(n * self.cfg.nstep)
        self.lib.orc_get_sources(self.h, isp, xi, ga, sa, stf)
        k = len(self.cfg.sources)
        return isp[:k], xi[:k], ga[:k], sa.reshape(n, NGLL2, 2)[:k], stf.reshape(n, self.cfg.nstep)[:k]

    def receivers(self):
        n = max(1, self.cfg.nrec)
        isp, xi, ga = np.zeros(n, np.int32), np.zeros(n), np.zeros(n)
        xir, gar = np.zeros(5 * n), np.zeros(5 * n)
        self.lib.orc_get_receivers(self.h, isp, xi, ga, xir, gar)
        k = self.cfg.nrec
        return isp[:k], xi[:k], ga[:k], xir.reshape(n, 5)[:k], gar.reshape(n, 5)[:k]

    # time stepping ----------------------------------------------------
    def forward(self, nstep=None, save_stacey=False):
        """Run the forward loop; returns seismograms [nrec, 2, nstep] (and the Stacey
        storage [nstep, nabs, 5, 2] when ``save_stacey``)."""
        n = self.cfg.nstep if nstep is None else int(nstep)
        seis = np.zeros((max(1, self.cfg.nrec), 2, n))
        store = np.zeros((n, max(1, self.nabs), NGLL, 2)) if save_stacey else None
        self.lib.orc_forward(self.h, n, _ptr(seis), _ptr(store))
        seis = seis[: self.cfg.nrec]
        return (seis, store) if save_stacey else seis

    def fields(self):
        d, v, a = (np.zeros(2 * self.nglob) for _ in range(3))
        self.lib.orc_get_fields(self.h, _ptr(d), _ptr(v), _ptr(a))
        return d.reshape(-1, 2), v.reshape(-1, 2), a.reshape(-1, 2)

    def set_fields(self, displ=None, veloc=None, accel=None):
        arrs = [None if x is None else np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))
                for x in (displ, veloc, accel)]
        self.lib.orc_set_fields(self.h, *[_ptr(x) for x in arrs])

    def apply_stiffness(self, u):
        u = np.ascontiguousarray(np.asarray(u, np.float64).reshape(-1))
        out = np.zeros_like(u)
        self.lib.orc_apply_stiffness(self.h, u, out)
        return out.reshape(-1, 2)

    # acoustic domain ----------------------------------------------------
    def is_acoustic(self):
        out = np.zeros(self.nspec, np.int32)
        self.lib.orc_get_is_acoustic(self.h, out)
        return out.astype(bool)

    def rmass_acoustic(self):
        out = np.zeros(self.nglob)
        self.lib.orc_get_rmass_acoustic(self.h, out)
        return out

    def coupling(self):
        """Fluid-solid edges: (acoustic element, side, elastic element, side), 0-based elements, sides 1 bottom,
        2 right, 3 top, 4 left."""
        n = max(1, self.ncoupling)
        arrs = [np.zeros(n, np.int32) for _ in range(4)]
        self.lib.orc_get_coupling(self.h, *arrs)
        return np.stack(arrs, axis=1)[: self.ncoupling]

    def free_surface(self):
        n = max(1, self.lib.orc_nfree(self.h))
        isp, code = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.lib.orc_get_free_surface(self.h, isp, code)
        k = self.lib.orc_nfree(self.h)
        return isp[:k], code[:k]

    def acoustic_stacey_range(self):
        n = max(1, self.nabs)
        b, e = np.zeros(n, np.int32), np.zeros(n, np.int32)
        self.lib.orc_get_acoustic_stacey_range(self.h, b, e)
        return b[: self.nabs], e[: self.nabs]

    def acoustic_fields(self):
        p, v, a = (np.zeros(self.nglob) for _ in range(3))
        self.lib.orc_get_acoustic_fields(self.h, _ptr(p), _ptr(v), _ptr(a))
        return p, v, a

    def set_acoustic_fields(self, pot=None, pot_dot=None, pot_dd=None):
        arrs = [None if x is None else np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1))
                for x in (pot, pot_dot, pot_dd)]
        self.lib.orc_set_acoustic_fields(self.h, *[_ptr(x) for x in arrs])

    def apply_acoustic_stiffness(self, pot):
        pot = np.ascontiguousarray(np.asarray(pot, np.float64).reshape(-1))
        out = np.zeros_like(pot)
        self.lib.orc_apply_acoustic_stiffness(self.h, pot, out)
        return out

    def pml(self):
        """CPML set-up: dict with is_pml, region [nspec]; kx, kz, dx, dz, ax, az [nspec_pml, 25] (double); the
        outer-boundary points of elastic and acoustic PML elements (0-based, increasing)."""
        npml = self.lib.orc_nspec_pml(self.h)
        nel, nac = self.lib.orc_npml_points(self.h, 1), self.lib.orc_npml_points(self.h, 0)
        isp, reg = np.zeros(self.nspec, np.int32), np.zeros(self.nspec, np.int32)
        prof = np.zeros(6 * max(1, npml) * NGLL2)
        pe, pa = np.zeros(max(1, nel), np.int32), np.zeros(max(1, nac), np.int32)
        self.lib.orc_get_pml(self.h, isp, reg, prof, pe, pa)
        prof = prof[: 6 * npml * NGLL2].reshape(6, npml, NGLL2)
        out = {"is_pml": isp.astype(bool), "region": reg, "points_elastic": pe[:nel] - 1,
               "points_acoustic": pa[:nac] - 1}
        out.update(zip(("kx", "kz", "dx", "dz", "ax", "az"), prof))
        return out

    def attenuation_coefficients(self):
        """(A_kappa, B_kappa, A_mu, B_mu), each [nspec, 25, n_sls]."""
        n = self.nspec * 25 * self.cfg.n_sls
        out = [np.zeros(n) for _ in range(4)]
        self.lib.orc_get_attenuation(self.h, *out)
        return tuple(o.reshape(self.nspec, 25, self.cfg.n_sls) for o in out)

    def _elastic_only(self, what):
        if self.nspec_acoustic or self.cfg.attenuation:
            raise NotImplementedError(f"{what}: the oracle's adjoint path covers elastic meshes without attenuation only")

    def adjoint(self, adj, stacey_store, lastframe, dump_steps=()):
        """SIMULATION_TYPE=3 run with backward reconstruction.  ``adj`` [nrec,2,nstep] adjoint
        source traces, ``stacey_store`` from a SAVE_FORWARD run, ``lastframe`` = (displ, veloc, accel)
        at the end of that run.  Returns (adjoint seismograms [nrec,2,nstep], kernels dict of
        [nspec,25] arrays rho/mu/kappa, scaled as save_adjoint_kernels.f90 writes them)."""
        self._elastic_only("adjoint")
        n = self.cfg.nstep
        adj = np.ascontiguousarray(np.asarray(adj, np.float64).reshape(-1))
        st = np.ascontiguousarray(np.asarray(stacey_store, np.float64).reshape(-1))
        if st.size == 0:
            st = np.zeros(1)
        d, v, a = (np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1)) for x in lastframe)
        seis = np.zeros((max(1, self.cfg.nrec), 2, n))
        ker = np.zeros(3 * self.nspec * NGLL2)
        steps = np.ascontiguousarray(np.asarray(dump_steps, np.int32))
        dumps = np.zeros((max(1, len(steps)), 2 * self.nglob))
        self.lib.orc_adjoint(self.h, adj, st, d, v, a, _ptr(seis), ker, len(steps), _ptr(steps), _ptr(dumps))
        ker = ker.reshape(3, self.nspec, NGLL2)
        out = (seis[: self.cfg.nrec], {"rho": ker[0], "mu": ker[1], "kappa": ker[2]})
        if len(steps):
            return out + (dumps[: len(steps)].reshape(len(steps), -1, 2),)
        return out

    def adjoint_full(self, adj, stacey_store, lastframe_el=None, lastframe_ac=None):
        """SIMULATION_TYPE=3 run over fluid, solid or coupled meshes (no CPML, no attenuation).  ``adj``
        [nrec,2,nstep] adjoint source traces (a receiver in a fluid element uses component 0),
        ``stacey_store`` from a SAVE_FORWARD run (``forward(save_stacey=True)``), ``lastframe_el`` =
        (displ, veloc, accel) and ``lastframe_ac`` = (potential, its first and second time derivative) at the
        end of that run.  Returns (adjoint seismograms [nrec,2,nstep], solid kernels dict rho/mu/kappa scaled
        as save_adjoint_kernels.f90 writes them, fluid kernels dict rho/kappa/rhop/alpha), kernels [nspec,25]."""
        n = self.cfg.nstep
        adj = np.ascontiguousarray(np.asarray(adj, np.float64).reshape(-1))
        if adj.size == 0:
            adj = np.zeros(1)
        st = np.ascontiguousarray(np.asarray(stacey_store, np.float64).reshape(-1))
        if st.size == 0:
            st = np.zeros(1)
        lel = None if lastframe_el is None else np.ascontiguousarray(
            np.concatenate([np.asarray(x, np.float64).reshape(-1) for x in lastframe_el]))
        lac = None if lastframe_ac is None else np.ascontiguousarray(
            np.concatenate([np.asarray(x, np.float64).reshape(-1) for x in lastframe_ac]))
        seis = np.zeros((max(1, self.cfg.nrec), 2, n))
        kel = np.zeros(3 * self.nspec * NGLL2)
        kac = np.zeros(4 * self.nspec * NGLL2)
        rc = self.lib.orc_adjoint_full(self.h, adj, st, _ptr(lel), _ptr(lac), _ptr(seis), _ptr(kel), _ptr(kac))
        if rc != 0:
            raise NotImplementedError("oracle: adjoint runs with CPML or attenuation, or without the last frames")
        kel = kel.reshape(3, self.nspec, NGLL2)
        kac = kac.reshape(4, self.nspec, NGLL2)
        return (seis[: self.cfg.nrec], dict(zip(("rho", "mu", "kappa"), kel)),
                dict(zip(("rho", "kappa", "rhop", "alpha"), kac)))

    def gradient(self, obs):
        """Exact gradient of chi = 1/2 dt sum |s - obs|^2 over a full forward run from rest (seismotype
        displacement).  Returns (chi, grad dict rho/vp/vs, parts dict kappa/mu/rho_mass/rho_vp/rho_vs), all
        [nspec,25] in the MODEL=gll per-point parameterisation."""
        self._elastic_only("gradient")
        o = np.ascontiguousarray(np.asarray(obs, np.float64).reshape(-1))
        if o.size != max(1, self.cfg.nrec) * 2 * self.cfg.nstep and self.cfg.nrec > 0:
            raise ValueError("obs must be [nrec, 2, nstep]")
        parts = np.zeros(5 * self.nspec * NGLL2)
        grad = np.zeros(3 * self.nspec * NGLL2)
        chi = self.lib.orc_gradient(self.h, o, parts, grad)
        parts = parts.reshape(5, self.nspec, NGLL2)
        grad = grad.reshape(3, self.nspec, NGLL2)
        return chi, {"rho": grad[0], "vp": grad[1], "vs": grad[2]}, dict(
            zip(("kappa", "mu", "rho_mass", "rho_vp", "rho_vs"), parts))

    def kernel_ac_increment(self, pot, pot_dd, b_pot, dt):
        """One fluid kernel accumulation (compute_kernels_ac) for given fields: dict rho/kappa of [nspec,25]."""
        args = [np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1)) for x in (pot, pot_dd, b_pot)]
        out = np.zeros(2 * self.nspec * NGLL2)
        self.lib.orc_kernel_ac_increment(self.h, *args, float(dt), out)
        out = out.reshape(2, self.nspec, NGLL2)
        return {"rho": out[0], "kappa": out[1]}

    def kernel_increment(self, displ, accel, b_displ):
        """One unscaled kernel accumulation (compute_kernels_el) for given fields."""
        args = [np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1)) for x in (displ, accel, b_displ)]
        out = np.zeros(3 * self.nspec * NGLL2)
        self.lib.orc_kernel_increment(self.h, *args, out)
        out = out.reshape(3, self.nspec, NGLL2)
        return {"rho": out[0], "mu": out[1], "kappa": out[2]}
