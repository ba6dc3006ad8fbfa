"""The reference's Claerbout kernel benchmark as a workflow over an engine (test infrastructure only).

Restates what ``EXAMPLES/BENCHMARK_CLAERBOUT_ADJOINT/*/run_this_example_kernel.sh`` runs, so that a simulation engine
(the oracle, or the HIP library) can be checked against the figures the reference's run log prints
(``REF_KERNEL/output_run.log``):

1. data: a forward run of the target model (the Par_file's material table), saving the model on the GLL points
   (``save_model_files.f90:60-75``: rho, vp = rho_vp / rho, vs = rho_vs / rho, x, z in single precision);
2. a Gaussian perturbation of vp (``model_add_Gaussian_perturbation.py``: centre (500, 500), sigma = 1/4 of a
   270 m wavelength over sqrt(8), vp * (1 + percent * exp(-dx^2/2s^2 - dz^2/2s^2)), -2 %), stored in single
   precision: the initial model;
3. synthetics: a forward run of the initial model (MODEL = gll) with SAVE_FORWARD;
4. adjoint sources (``adj_seismogram.py``): syn - dat of the single-precision SU traces; for pressure records the
   negative second difference over DT^2 (first and last samples kept as the difference), for displacement records
   the difference itself;
5. the kernel run (SIMULATION_TYPE = 3): fluid kernels rhop, c (``compute_kernels.f90:426-439``), solid kernels
   rhop, alpha, beta (``save_adjoint_kernels.f90:160-166``);
6. the model update (``model_update.py``: step = percent / max |kernel|, m * exp(-step K)) and a forward run of
   the updated model;
7. the model-space and data-space inner products (``kernel_evaluation_postprocessing.py``, weights
   w_i w_j jacobian, ``save_adjoint_kernels.f90:706-735``).

NumPy 1.x evaluated a Python float against a float32 scalar in float64; the second difference of step 4 is
evaluated that way here.
"""
from __future__ import annotations

import dataclasses

import numpy as np

f32 = np.float32
CENTER_X = CENTER_Z = 500.0
WAVELENGTH = 1.0 / 10.0 * 2700.0
PERCENT = -0.02
STEP_PERCENT = 0.02


def gaussian(x, z):
    """model_add_Gaussian_perturbation.py: the perturbation's shape on the GLL points (float64)."""
    s = 0.25 * WAVELENGTH / np.sqrt(8.0)
    s2 = 2.0 * s * s
    return np.exp(-((CENTER_X - x) ** 2) / s2 - ((CENTER_Z - z) ** 2) / s2)


def adjoint_source(syn, dat, dt_us, pressure):
    """adj_seismogram.py: syn, dat [nrec, nstep] float32 traces; returns (adjoint source float32, max |syn - dat|,
    misfit sum (syn - dat)^2 DT)."""
    diff = (np.asarray(syn, f32) - np.asarray(dat, f32)).astype(f32)
    dt = dt_us * 1.0e-6
    misfit = float(sum(float(np.sum(d * d)) * dt for d in diff))
    adj = diff.copy()
    if pressure:
        fac = 1.0 / dt ** 2
        a = diff.astype(np.float64)
        adj[:, 1:-1] = (-((a[:, 2:] - 2.0 * a[:, 1:-1] + a[:, :-2]) * fac)).astype(f32)
    return adj, float(np.abs(diff).max()), misfit


def solid_kernels(kel, mu, kappa):
    """save_adjoint_kernels.f90:160-166 from the scaled rho, mu, kappa kernels (TWO and FOUR_THIRDS are double
    precision constants there)."""
    rho_kl, mu_kl, kappa_kl = (np.asarray(kel[k], f32) for k in ("rho", "mu", "kappa"))
    mul, kap = np.asarray(mu, f32), np.asarray(kappa, f32)
    ratio = (4.0 / 3.0) * mul.astype(np.float64) / kap.astype(np.float64)
    rhop = (rho_kl + kappa_kl + mu_kl).astype(f32)
    beta = (2.0 * (mu_kl.astype(np.float64) - ratio * kappa_kl.astype(np.float64))).astype(f32)
    alpha = (2.0 * (1.0 + ratio) * kappa_kl.astype(np.float64)).astype(f32)
    return rhop, alpha, beta


def run(cfg, engine, coupled):
    """The whole benchmark for ``cfg`` (from the example's Par_file) over ``engine``; returns the figures the run
    log prints, plus the kernels."""
    pressure = cfg.seismotype == 4
    dt_us = round(cfg.dt * 1.0e6)
    # 1. data, and the model as SAVE_MODEL = gll writes it
    dat = engine.forward(cfg)["seis"]
    geo = engine.setup(cfg)
    rho0, vp_t, vs_t = geo["rho"], geo["vp"], geo["vs"]
    # 2. initial model
    pert = gaussian(geo["x"].astype(np.float64), geo["z"].astype(np.float64))
    vp0 = (vp_t * (1.0 + PERCENT * pert)).astype(f32)
    cfg0 = with_model(cfg, rho0, vp0, vs_t)
    # 3. synthetics with the forward wavefield saved
    fw = engine.forward(cfg0, save=True)
    syn = fw["seis"]
    # 4. adjoint sources, one per recorded component
    comps = [0] if pressure else [0, 1]
    adj = np.zeros((syn.shape[0], 2, syn.shape[2]), f32)
    out = {"max_waveform_difference": [], "total_misfit": 0.0, "max_adjoint_source": []}
    for c in comps:
        a, dmax, mis = adjoint_source(syn[:, c].astype(f32), dat[:, c].astype(f32), dt_us, pressure)
        adj[:, c] = a
        out["max_waveform_difference"].append(dmax)
        out["max_adjoint_source"].append(float(np.abs(a).max()))
        out["total_misfit"] += mis
    # 5. kernels
    kel, kac = engine.adjoint(cfg0, adj, fw)
    krhop = np.asarray(kac["rhop"], f32).reshape(-1)
    kalpha = np.asarray(kac["alpha"], f32).reshape(-1)
    kbeta = np.zeros_like(kalpha)
    if coupled:
        rhop, alpha, beta = solid_kernels(kel, engine.setup(cfg0)["mu"], engine.setup(cfg0)["kappa"])
        krhop = (krhop + rhop.reshape(-1)).astype(f32)
        kalpha = (kalpha + alpha.reshape(-1)).astype(f32)
        kbeta = beta.reshape(-1)
    out.update(kernel_rhop_minmax=[float(krhop.min()), float(krhop.max())],
               kernel_alpha_minmax=[float(kalpha.min()), float(kalpha.max())],
               kernel_beta_minmax=[float(kbeta.min()), float(kbeta.max())],
               norm_rhop=float(np.sum(krhop * krhop)), norm_alpha=float(np.sum(kalpha * kalpha)),
               norm_beta=float(np.sum(kbeta * kbeta)))
    # 6. model update and the updated synthetics
    kmax = max(float(np.abs(krhop).max()), float(np.abs(kalpha).max()), float(np.abs(kbeta).max()))
    step = STEP_PERCENT / kmax
    rho1 = (rho0.reshape(-1) * np.exp(-step * krhop)).astype(f32)
    vp1 = (vp0.reshape(-1) * np.exp(-step * kalpha)).astype(f32)
    vs1 = (vs_t.reshape(-1) * np.exp(-step * kbeta)).astype(f32)
    syn1 = engine.forward(with_model(cfg, rho1, vp1, vs1))["seis"]
    # 7. inner products
    # (rho and vs are not perturbed: their terms vanish)
    w = geo["weights"].reshape(-1)
    dlnvp = ((vp_t - vp0) / vp0).reshape(-1)
    out["modelspace"] = float(np.sum(w * (-kalpha) * dlnvp))
    ds_new = 0.0
    for c in comps:
        d = syn1[:, c].astype(f32) - dat[:, c].astype(f32)
        ds_new += float(sum(float(np.sum(t * t)) * dt_us * 1.0e-6 for t in d))
    out["dataspace"] = out["total_misfit"]
    out["dataspace_new"] = ds_new
    out["kernels"] = {"rhop": krhop, "alpha": kalpha, "beta": kbeta}
    return out


def with_model(cfg, rho, vp, vs):
    """``cfg`` with per-GLL-point MODEL = gll values (single precision, as read from the model files)."""
    as64 = lambda a: np.asarray(a, f32).astype(np.float64).reshape(-1)  # noqa: E731
    return dataclasses.replace(cfg, rho=as64(rho), vp=as64(vp), vs=as64(vs), material_mode=1)


class OracleEngine:
    """The FP32 oracle as the benchmark's engine."""

    def __init__(self, precision="sp"):
        from oracle.oracle import OracleSim
        self.Sim = OracleSim
        self.precision = precision
        self._sims = {}

    def _sim(self, cfg):
        return self.Sim(cfg, self.precision)

    def setup(self, cfg):
        o = self._sim(cfg)
        rho = o.elem_array("rho").astype(f32)
        ib = o.ibool() - 1
        xz = o.coord()
        _, w, _, _ = o.basis()
        jac = o.elem_array("jacobian").astype(f32).astype(np.float64).reshape(-1, 5, 5)
        weights = (w[None, None, :] * w[None, :, None] * jac).astype(f32)  # [e][j][i]: w_i w_j jacobian
        return {"rho": rho, "vp": (o.elem_array("rho_vp").astype(f32) / rho).astype(f32),
                "vs": (o.elem_array("rho_vs").astype(f32) / rho).astype(f32),
                "mu": o.elem_array("mu").astype(f32), "kappa": o.elem_array("kappa").astype(f32),
                "x": xz[ib, 0].astype(f32), "z": xz[ib, 1].astype(f32), "weights": weights.reshape(-1, 25)}

    def forward(self, cfg, save=False):
        o = self._sim(cfg)
        if not save:
            return {"seis": o.forward()}
        seis, store = o.forward(save_stacey=True)
        out = {"seis": seis, "store": store, "sim": o}
        out["last_el"] = o.fields() if o.nspec_acoustic < o.nspec else None
        out["last_ac"] = o.acoustic_fields() if o.nspec_acoustic > 0 else None
        return out

    def adjoint(self, cfg, adj, fw):
        o = fw["sim"]
        _, kel, kac = o.adjoint_full(adj.astype(np.float64), fw["store"], fw["last_el"], fw["last_ac"])
        return kel, kac
