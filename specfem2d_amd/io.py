"""The reference's on-disk formats for this path: ``DATA/Par_file``, ``DATA/SOURCE``, the interfaces file and
``DATA/STATIONS`` in; seismograms out in the reference's ASCII and binary layouts.

Readers (``src/shared/param_reader.c``, ``read_parameter_file.F90``, ``read_material_table.f90``,
``read_regions.f90``, ``src/meshfem2D/read_interfaces_file.f90``, ``read_source_file.f90``,
``save_stations_file.f90``) build a :class:`~specfem2d_amd.config.SimulationConfig` for what this library runs:
the internal mesher with flat interfaces (NGNOD = 4, or 9, which on flat layers gives the same rectangles),
isotropic material tables (solid, or fluid when vs = 0), seismotypes 1-4 and 6,
Stacey or free boundaries, collocated force / pressure sources with time functions 1-3, Newmark, P-SV.
Anything else in the files is refused with a message naming the parameter.

Writers follow ``write_seismograms.F90:252-584``: ASCII traces ``NET.STA.CMP.semX`` with the gfortran
list-directed layout of ``write(11,*) time, ' ', value`` (byte-compatible), the binary
``U{x,z,p}_file_{single,double}_X.bin`` streams ([receiver][sample], native endianness), and the Seismic Unix
``U{x,z,p}_file_single_X.su`` files of SU_FORMAT (write_output_SU.f90: a 240-byte header per trace).
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .config import NGLL2, ExternalMesh, SimulationConfig, Source


# ----------------------------------------------------------------------------------------------------------
# Par_file
# ----------------------------------------------------------------------------------------------------------
@dataclass
class ParFile:
    """``name = value`` pairs in file order (names repeat for receiver sets), plus the material table and the
    regions, which follow ``nbmodels`` and ``nbregions`` as bare lines."""

    entries: List[Tuple[str, str]]
    materials: Dict[int, List[float]]  # model number -> [icodemat, val0 ... val12]
    regions: List[Tuple[int, int, int, int, int]]  # ix_start, ix_end, iz_start, iz_end, model (1-based)

    def get(self, name: str, default: Optional[str] = None) -> str:
        for k, v in self.entries:
            if k == name:
                return v
        if default is None:
            raise KeyError(f"Par_file: parameter {name} is missing")
        return default

    def all(self, name: str) -> List[str]:
        return [v for k, v in self.entries if k == name]


def _strip_comment(line: str) -> str:
    return line.split("#", 1)[0].strip()


def fortran_float(s: str) -> float:
    """A Fortran real literal: 1.d0, 2.5D-3, 4e-4, 3000."""
    return float(s.strip().replace("d", "e").replace("D", "e"))


def fortran_bool(s: str) -> bool:
    t = s.strip().lower()
    if t in (".true.", "true", "t", ".t."):
        return True
    if t in (".false.", "false", "f", ".f."):
        return False
    raise ValueError(f"not a Fortran logical: {s!r}")


def read_par_file(path: str) -> ParFile:
    entries: List[Tuple[str, str]] = []
    materials: Dict[int, List[float]] = {}
    regions: List[Tuple[int, int, int, int, int]] = []
    pending_models = pending_regions = 0
    with open(path) as f:
        for raw in f:
            line = _strip_comment(raw)
            if not line:
                continue
            if "=" in line:
                k, v = (x.strip() for x in line.split("=", 1))
                entries.append((k, v))
                if k == "nbmodels":
                    pending_models = int(v)
                elif k == "nbregions":
                    pending_regions = int(v)
                continue
            toks = line.split()
            if pending_models > 0:
                vals = [fortran_float(t) for t in toks]
                materials[int(vals[0])] = vals[1:]
                pending_models -= 1
            elif pending_regions > 0:
                ix0, ix1, iz0, iz1, m = (int(t) for t in toks[:5])
                regions.append((ix0, ix1, iz0, iz1, m))
                pending_regions -= 1
            else:
                raise ValueError(f"{path}: unexpected line {raw.rstrip()!r}")
    if pending_models or pending_regions:
        raise ValueError(f"{path}: material table or regions incomplete")
    return ParFile(entries, materials, regions)


def read_interfaces(path: str) -> Tuple[List[float], List[int], List[Tuple[np.ndarray, np.ndarray]]]:
    """Interfaces file of the internal mesher: returns (interface heights bottom first, elements per layer,
    the (x, z) points of each interface).  Only flat interfaces are supported by this library's mesher."""
    toks: List[str] = []
    with open(path) as f:
        for raw in f:
            line = _strip_comment(raw)
            if line:
                toks.extend(line.split())
    pos = 0

    def nxt():
        nonlocal pos
        pos += 1
        return toks[pos - 1]

    nint = int(nxt())
    curves = []
    for _ in range(nint):
        npts = int(nxt())
        xs = np.array([fortran_float(nxt()) for _ in range(2 * npts)]).reshape(npts, 2)
        curves.append((xs[:, 0], xs[:, 1]))
    nz_layers = [int(nxt()) for _ in range(nint - 1)]
    heights = []
    for k, (x, z) in enumerate(curves):
        if np.ptp(z) != 0.0:
            raise NotImplementedError(f"{path}: interface {k + 1} is not flat (topography is not supported)")
        heights.append(float(z[0]))
    return heights, nz_layers, curves


def read_source_file(path: str, nsources: int) -> List[Source]:
    """``DATA/SOURCE``: ``nsources`` blocks of ``name = value`` lines."""
    blocks: List[Dict[str, str]] = []
    cur: Dict[str, str] = {}
    with open(path) as f:
        for raw in f:
            line = _strip_comment(raw)
            if not line or "=" not in line:
                continue
            k, v = (x.strip() for x in line.split("=", 1))
            if k == "source_surf" and cur:
                blocks.append(cur)
                cur = {}
            cur[k] = v
    if cur:
        blocks.append(cur)
    if len(blocks) < nsources:
        raise ValueError(f"{path}: {len(blocks)} sources for NSOURCES = {nsources}")
    out = []
    for b in blocks[:nsources]:
        if fortran_bool(b.get("source_surf", ".false.")):
            raise NotImplementedError("SOURCE: source_surf = .true. is not supported")
        st = int(b["source_type"])
        if st not in (1, 2):
            raise NotImplementedError(f"SOURCE: source_type {st} (only force / pressure and moment-tensor sources)")
        tf = int(b["time_function_type"])
        if tf not in (1, 2, 3):
            raise NotImplementedError(f"SOURCE: time_function_type {tf} (only 1, 2, 3)")
        out.append(Source(x=fortran_float(b["xs"]), z=fortran_float(b["zs"]), f0=fortran_float(b["f0"]),
                          factor=fortran_float(b["factor"]), anglesource=fortran_float(b.get("anglesource", "0")),
                          tshift=fortran_float(b.get("tshift", "0")), time_function_type=tf, source_type=st,
                          Mxx=fortran_float(b.get("Mxx", "0")) if st == 2 else 0.0,
                          Mzz=fortran_float(b.get("Mzz", "0")) if st == 2 else 0.0,
                          Mxz=fortran_float(b.get("Mxz", "0")) if st == 2 else 0.0))
    return out


@dataclass
class Station:
    name: str
    network: str
    x: float
    z: float


def stations_from_receiver_sets(par: ParFile) -> List[Station]:
    """The receiver lines of the Par_file as ``xmeshfem2D`` writes them to ``DATA/STATIONS``
    (save_stations_file.f90:93-110): evenly spaced along each line, positions rounded to 7 decimals (f20.7)."""
    nsets = int(par.get("nreceiversets"))
    nrecs, xd, zd, xf, zf = (par.all(k) for k in ("nrec", "xdeb", "zdeb", "xfin", "zfin"))
    surf = par.all("record_at_surface_same_vertical")
    if min(len(nrecs), len(xd), len(zd), len(xf), len(zf)) < nsets:
        raise ValueError("Par_file: receiver sets incomplete")
    out: List[Station] = []
    k = 0
    for s in range(nsets):
        if s < len(surf) and fortran_bool(surf[s]):
            raise NotImplementedError("record_at_surface_same_vertical = .true. is not supported")
        n = int(nrecs[s])
        x0, z0, x1, z1 = (fortran_float(v[s]) for v in (xd, zd, xf, zf))
        for i in range(1, n + 1):
            if n > 1:
                x = x0 + float(i - 1) * (x1 - x0) / float(n - 1)
                z = z0 + float(i - 1) * (z1 - z0) / float(n - 1)
            else:
                x, z = x0, z0
            k += 1
            out.append(Station(f"S{k:04d}", "AA", float(f"{x:20.7f}"), float(f"{z:20.7f}")))
    return out


def read_stations(path: str) -> List[Station]:
    """``DATA/STATIONS``: ``station network x z elevation burial`` per line."""
    out = []
    with open(path) as f:
        for raw in f:
            t = raw.split()
            if len(t) >= 4:
                out.append(Station(t[0], t[1], fortran_float(t[2]), fortran_float(t[3])))
    return out


def config_from_par_file(data_dir: str, cpml: str = "cpml", nstep: Optional[int] = None) -> Tuple[SimulationConfig, List[Station]]:
    """Build the configuration of ``data_dir/Par_file`` (+ its SOURCE, interfaces file and STATIONS).

    ``cpml``: what to do with ``PML_BOUNDARY_CONDITIONS = .true.``: ``"cpml"`` (default) sets up the CPML layer
    as the Par_file describes it (read_parameter_file.F90:843-900); ``"stacey"`` puts Stacey edges on the absorbing
    sides instead (same results until the waves reach the absorbing layer); ``"error"`` refuses it."""
    par = read_par_file(os.path.join(data_dir, "Par_file"))
    g = par.get

    def refuse(cond, what):
        if cond:
            raise NotImplementedError(f"Par_file: {what} is not supported by this library")

    # 9-node elements on the internal mesher's flat layers are rectangles: the same geometry as 4-node ones
    refuse(int(g("NGNOD")) not in (4, 9), f"NGNOD = {g('NGNOD')}")
    refuse(not fortran_bool(g("P_SV")), "P_SV = .false. (SH waves)")
    refuse(fortran_bool(g("AXISYM", ".false.")), "AXISYM")
    refuse(int(g("time_stepping_scheme", "1")) != 1, "time_stepping_scheme other than Newmark")
    refuse(g("MODEL", "default").strip().lower() != "default", f"MODEL = {g('MODEL')}")
    external = fortran_bool(g("read_external_mesh", ".false."))
    attenuation = fortran_bool(g("ATTENUATION_VISCOELASTIC", ".false."))
    if attenuation:
        refuse(fortran_bool(g("USE_SOLVOPT", ".false.")), "USE_SOLVOPT (nonlinear fit of the relaxation times)")
        refuse(fortran_bool(g("READ_VELOCITIES_AT_f0", ".false.")), "READ_VELOCITIES_AT_f0")
    refuse(fortran_bool(g("ATTENUATION_VISCOACOUSTIC", ".false.")), "ATTENUATION_VISCOACOUSTIC")
    refuse(fortran_bool(g("initialfield", ".false.")), "initialfield (plane-wave initial fields)")
    refuse(fortran_bool(g("ADD_PERIODIC_CONDITIONS", ".false.")), "ADD_PERIODIC_CONDITIONS")
    refuse(fortran_bool(g("ACOUSTIC_FORCING", ".false.")), "ACOUSTIC_FORCING")
    refuse(int(g("NTSTEP_BETWEEN_OUTPUT_SAMPLE", "1")) != 1, "NTSTEP_BETWEEN_OUTPUT_SAMPLE > 1")
    pml = fortran_bool(g("PML_BOUNDARY_CONDITIONS", ".false."))
    stacey = fortran_bool(g("STACEY_ABSORBING_CONDITIONS", ".false."))
    pml_kw = {}
    if pml:
        if cpml == "stacey":
            stacey, pml = True, False
        elif cpml == "cpml":
            refuse(stacey, "PML_BOUNDARY_CONDITIONS together with STACEY_ABSORBING_CONDITIONS")
            refuse(fortran_bool(g("ROTATE_PML_ACTIVATE", ".false.")), "ROTATE_PML_ACTIVATE")
            refuse(fortran_bool(g("PML_PARAMETER_ADJUSTMENT", ".false.")), "PML_PARAMETER_ADJUSTMENT")
            pml_kw = dict(nelem_pml=int(g("NELEM_PML_THICKNESS", "3")),
                          pml_k_min=fortran_float(g("K_MIN_PML", "1.d0")),
                          pml_k_max=fortran_float(g("K_MAX_PML", "1.d0")),
                          pml_damping_acoustic=fortran_float(g("damping_change_factor_acoustic", "0.5d0")),
                          pml_damping_elastic=fortran_float(g("damping_change_factor_elastic", "1.d0")))
        else:
            raise NotImplementedError("Par_file: PML_BOUNDARY_CONDITIONS refused (cpml='error')")
    seismotypes = [int(t) for t in re.split(r"[ ,]+", g("seismotype").strip()) if t]
    refuse(len(seismotypes) != 1 or seismotypes[0] not in (1, 2, 3, 4, 6), f"seismotype = {g('seismotype')}")
    mesh = None
    if external:
        refuse(int(g("NGNOD")) != 4, "read_external_mesh with NGNOD = 9")
        refuse(pml, "read_external_mesh with PML_BOUNDARY_CONDITIONS")
        mesh = read_external_mesh(par, data_dir)
        nx, nz, heights, nz_layers = mesh.nspec, 1, [float(mesh.coorg[:, 1].min()), float(mesh.coorg[:, 1].max())], [1]
        xmin, xmax = float(mesh.coorg[:, 0].min()), float(mesh.coorg[:, 0].max())
        mat = mesh.material.astype(np.int64)
    else:
        heights, nz_layers, _ = read_interfaces(os.path.join(data_dir, os.path.basename(g("interfacesfile"))))
        nx = int(g("nx"))
        nz = sum(nz_layers)
        xmin, xmax = fortran_float(g("xmin")), fortran_float(g("xmax"))
        # materials of the elements (read_regions.f90: later regions overwrite earlier ones), per GLL point
        mat = np.zeros(nx * nz, np.int64)
        for ix0, ix1, iz0, iz1, m in par.regions:
            if m not in par.materials:
                raise ValueError(f"Par_file: region uses undefined material {m}")
            for iz in range(iz0, iz1 + 1):
                mat[(iz - 1) * nx + ix0 - 1:(iz - 1) * nx + ix1] = m
    if np.any(mat == 0) or np.any([m not in par.materials for m in np.unique(mat)]):
        raise ValueError("Par_file: elements without a material, or with an undefined one")
    ne = len(mat)
    rho = np.zeros(ne)
    vp = np.zeros(ne)
    vs = np.zeros(ne)
    qk = np.full(ne, 9999.0)
    qm = np.full(ne, 9999.0)
    aniso_elem, aniso_c = None, None
    for m, vals in par.materials.items():
        icode = int(vals[0])
        refuse(icode not in (1, 2), f"material {m} with icodemat {icode} (isotropic and anisotropic tables only)")
        sel = mat == m
        if icode == 2:
            # read_materials.f90:160-200: rho c11 c13 c15 c33 c35 c55 c12 c23 c25 c22 QKappa Qmu; cp = sqrt(c33/rho),
            # cs = sqrt(c55/rho) for the mass and the Stacey impedances
            refuse(attenuation, "anisotropic materials with ATTENUATION_VISCOELASTIC")
            refuse(pml, "anisotropic materials with PML_BOUNDARY_CONDITIONS")
            r, c11, c13, c15, c33, c35, c55 = (float(v) for v in vals[1:8])
            if aniso_elem is None:
                aniso_elem = np.zeros(len(mat), bool)
                aniso_c = np.zeros((len(mat) * NGLL2, 6))
            aniso_elem[sel] = True
            aniso_c[np.repeat(sel, NGLL2)] = (c11, c13, c15, c33, c35, c55)
            rho[sel], vp[sel], vs[sel] = r, np.sqrt(c33 / r), np.sqrt(c55 / r)
            continue
        rho[sel], vp[sel], vs[sel] = vals[1], vals[2], vals[3]
        if attenuation:  # read_material_table.f90: QKappa, Qmu are the 7th and 8th values
            if len(vals) < 8:
                raise ValueError(f"Par_file: material {m} has {len(vals)} values; ATTENUATION_VISCOELASTIC needs "
                                 "QKappa and Qmu (7th and 8th)")
            qk[sel], qm[sel] = vals[6], vals[7]
    rep = lambda a: np.repeat(a, NGLL2)  # noqa: E731
    nsrc = int(g("NSOURCES"))
    sources = read_source_file(os.path.join(data_dir, "SOURCE"), nsrc)
    if fortran_bool(g("use_existing_STATIONS", ".false.")):
        stations = read_stations(os.path.join(data_dir, "STATIONS"))
    else:
        stations = stations_from_receiver_sets(par)
    cfg = SimulationConfig(
        name=g("title", "Par_file").strip(), nx=nx, nz=nz, xmin=xmin, xmax=xmax, interfaces=heights,
        nz_layers=nz_layers, nstep=nstep or int(g("NSTEP")), dt=fortran_float(g("DT")), rho=rep(rho), vp=rep(vp),
        vs=rep(vs), material_mode=0, stacey=stacey, absorbbottom=fortran_bool(g("absorbbottom")),
        absorbright=fortran_bool(g("absorbright")), absorbtop=fortran_bool(g("absorbtop")),
        absorbleft=fortran_bool(g("absorbleft")), sources=sources,
        receivers=np.array([[s.x, s.z] for s in stations]) if stations else None,
        user_t0=fortran_float(g("USER_T0", "0")), anglerec=fortran_float(g("anglerec", "0")),
        seismotype=seismotypes[0], use_trick=fortran_bool(g("USE_TRICK_FOR_BETTER_PRESSURE", ".false.")),
        attenuation=attenuation, qkappa=rep(qk), qmu=rep(qm), n_sls=int(g("N_SLS", "3")), pml=pml, mesh=mesh,
        aniso_elem=aniso_elem, aniso_c=aniso_c, **pml_kw)
    return cfg, stations


def _mesh_file(par: ParFile, data_dir: str, key: str) -> str:
    """A file named by the Par_file, relative to the run directory (the parent of DATA/) or to the Par_file's."""
    rel = par.get(key).strip()
    for base in (os.path.dirname(os.path.abspath(data_dir)), data_dir):
        p = os.path.normpath(os.path.join(base, rel))
        if os.path.exists(p):
            return p
    raise FileNotFoundError(f"Par_file: {key} = {rel} not found next to {data_dir}")


def _ints(path: str) -> List[int]:
    with open(path) as f:
        return [int(t) for t in f.read().split()]


def read_external_mesh(par: ParFile, data_dir: str) -> ExternalMesh:
    """The files of ``read_external_mesh = .true.`` (src/meshfem2D/read_external_mesh_files.F90): ``mesh_file``
    (element count, then 4 control nodes per element, 1-based), ``nodes_coords_file`` (node count, then x z per node),
    ``materials_file`` (a material number per element), ``free_surface_file`` and ``absorbing_surface_file`` (a side
    count, then per side: element, node count 2, the two nodes).  Free-surface sides count on fluid elements only
    (read_external_acoustic_surface, :398-416, keeps the sides of fluid materials); absorbing sides on all."""
    mesh_path = _mesh_file(par, data_dir, "mesh_file")
    with open(mesh_path) as f:
        head = f.readline().split()
        nspec = int(head[0])
        knods = np.array([int(t) for t in f.read().split()], np.int64)
    if knods.size != 4 * nspec:
        raise ValueError(f"{mesh_path}: {knods.size} node numbers for {nspec} 4-node elements")
    with open(_mesh_file(par, data_dir, "nodes_coords_file")) as f:
        npgeo = int(f.readline().split()[0])
        coorg = np.array([fortran_float(t) for t in f.read().split()], np.float64).reshape(npgeo, 2)
    mat = np.array(_ints(_mesh_file(par, data_dir, "materials_file")), np.int64)
    if mat.size != nspec:
        raise ValueError(f"materials_file: {mat.size} materials for {nspec} elements")
    knods = knods.reshape(nspec, 4) - 1
    if knods.min() < 0 or knods.max() >= npgeo:
        raise ValueError("mesh_file: node numbers outside nodes_coords_file")
    mesh = ExternalMesh(knods=knods, coorg=coorg, material=mat, free_elem=np.zeros(0, np.int64),
                        free_code=np.zeros(0, np.int64), abs_elem=np.zeros(0, np.int64),
                        abs_code=np.zeros(0, np.int64))

    def sides(key):
        try:
            vals = _ints(_mesh_file(par, data_dir, key))
        except (KeyError, FileNotFoundError):
            return np.zeros(0, np.int64), np.zeros(0, np.int64)
        n, rows = vals[0], vals[1:]
        width = len(rows) // n if n else 4
        if n and (len(rows) % n or width not in (4, 5)):
            raise ValueError(f"{key}: {len(rows)} values for {n} sides")
        el, code = [], []
        for k in range(n):
            r = rows[k * width:(k + 1) * width]
            if r[1] != 2:
                raise NotImplementedError(f"{key}: a side with {r[1]} nodes (only 2-node sides)")
            el.append(r[0] - 1)
            code.append(mesh.side_code(r[0] - 1, r[2] - 1, r[3] - 1))
        return np.array(el, np.int64), np.array(code, np.int64)

    mesh.free_elem, mesh.free_code = sides("free_surface_file")
    mesh.abs_elem, mesh.abs_code = sides("absorbing_surface_file")
    return mesh


# ----------------------------------------------------------------------------------------------------------
# seismograms out
# ----------------------------------------------------------------------------------------------------------
def gfortran_real8(x: float) -> str:
    """One REAL(8) item of gfortran list-directed output: 17 significant digits in a 25-character field, F form
    for 0.1 <= |x| < 1e17 (the exponent's 5 columns left blank), E form with a 3-digit exponent otherwise."""
    if x != x or math.isinf(x):
        body = "NaN" if x != x else "Infinity"
        return f"{('-' if x < 0 else '') + body:>25}"
    sign = "-" if (x < 0 or (x == 0.0 and math.copysign(1.0, x) < 0)) else " "
    a = abs(x)
    if a == 0.0:
        body = "0." + "0" * 16 + " " * 5
    else:
        e = f"{a:.16E}"  # mantissa with 17 significant digits, rounded
        mant, exp = e.split("E")
        ex = int(exp)
        if -1 <= ex < 17:  # 0.1 <= |x| (after rounding) < 1e17: F form with 17 significant digits
            digits = 16 - ex  # decimals
            body = f"{a:.{digits}f}" + " " * 5
        else:
            body = f"{mant}E{'-' if ex < 0 else '+'}{abs(ex):03d}"
    return f"{sign + body:>25}"


def ascii_trace_lines(times: Sequence[float], values: Sequence[float]) -> str:
    """The text of one ASCII trace as ``write(11,*) time_t,' ',value`` writes it (write_seismograms.F90:521)."""
    return "".join(" " + gfortran_real8(float(t)) + "   " + gfortran_real8(float(v)) + "\n"
                   for t, v in zip(times, values))


_CHANNEL_SUFFIX = {1: "d", 2: "v", 3: "a", 4: "p", 6: "x"}


def write_ascii_seismograms(out_dir: str, seis: np.ndarray, stations: Sequence[Station], dt: float, t0: float,
                            seismotype: int = 1) -> List[str]:
    """``seis`` [nrec, 2, nstep] (component 0 only for pressure and potential); writes ``NET.STA.BX{X,Z}.semX``
    (``.PRE.semp`` for pressure, ``.POT.semx`` for the potential) with time = (it - 1) dt - t0; returns the paths."""
    os.makedirs(out_dir, exist_ok=True)
    nstep = seis.shape[-1]
    times = [float(i) * dt - t0 for i in range(nstep)]
    comp = _CHANNEL_SUFFIX[seismotype]
    channels = {4: ["PRE"], 6: ["POT"]}.get(seismotype, ["BXX", "BXZ"])
    paths = []
    for r, st in enumerate(stations):
        for c, ch in enumerate(channels):
            p = os.path.join(out_dir, f"{st.network}.{st.name}.{ch}.sem{comp}")
            with open(p, "w") as f:
                f.write(ascii_trace_lines(times, seis[r, c]))
            paths.append(p)
    return paths


def write_binary_seismograms(out_dir: str, seis: np.ndarray, seismotype: int = 1, double: bool = False) -> List[str]:
    """``U{x,z}_file_{single,double}_X.bin`` (``Up_`` for pressure): [receiver][sample] streams
    (write_seismograms.F90:526-560)."""
    os.makedirs(out_dir, exist_ok=True)
    comp = _CHANNEL_SUFFIX[seismotype]
    kind = "double" if double else "single"
    dtype = np.float64 if double else np.float32
    names = [("Up", 0)] if seismotype in (4, 6) else [("Ux", 0), ("Uz", 1)]
    paths = []
    for pre, c in names:
        p = os.path.join(out_dir, f"{pre}_file_{kind}_{comp}.bin")
        np.ascontiguousarray(seis[:, c, :], dtype=dtype).tofile(p)
        paths.append(p)
    return paths


def _nint(x: float) -> int:
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(np.sign(x) * np.floor(abs(x) + 0.5))


SU_HEADER_WORDS = 60  # 240 bytes: 28 int32, 2 + 2 int16, 30 float32


def su_header(irec: int, xr: float, zr: float, xs: float, zs: float, nsamples: int, sampling_dt: float,
              receiver_interval: float) -> bytes:
    """The 240-byte trace header of write_output_SU.f90:55-96 (little-endian, as the x86-64 build writes it):
    int32 words 1 receiver number, 10 offset NINT(xr - xs), 19-22 NINT(xs, zs, xr, zr); int16 pair (0, samples,
    -9999 past 32767); int16 pair (sampling interval in us, else ms, else s; 0); float32 word 18 the receiver
    interval x2 - x1 (0 for one receiver)."""
    h1 = np.zeros(28, "<i4")
    h1[0] = irec
    h1[9] = _nint(xr - xs)
    h1[18], h1[19], h1[20], h1[21] = _nint(xs), _nint(zs), _nint(xr), _nint(zr)
    h2 = np.array([0, nsamples if nsamples < 32768 else -9999], "<i2")
    us, ms = _nint(sampling_dt * 1.0e6), _nint(sampling_dt * 1.0e3)
    h3 = np.array([us if us < 32768 else (ms if ms < 32768 else _nint(sampling_dt)), 0], "<i2")
    h4 = np.zeros(30, "<f4")
    h4[17] = np.float32(receiver_interval)
    return h1.tobytes() + h2.tobytes() + h3.tobytes() + h4.tobytes()


def write_su_seismograms(out_dir: str, seis: np.ndarray, receivers, source_xz, dt: float, seismotype: int = 1,
                         subsamp: int = 1) -> List[str]:
    """Seismic Unix files as SU_FORMAT writes them (write_seismograms.F90:337-392, write_output_SU.f90:34-125):
    ``U{x,z}_file_single_X.su`` (``Up_`` for pressure and potential), per receiver a 240-byte header and its
    samples in float32.  ``seis`` [nrec, 2, nsamples], ``receivers`` [nrec, 2] (x, z), ``source_xz`` the first
    source's position, ``subsamp`` NTSTEP_BETWEEN_OUTPUT_SAMPLE."""
    os.makedirs(out_dir, exist_ok=True)
    seis = np.asarray(seis, np.float64)
    rec = np.asarray(receivers, np.float64).reshape(-1, 2)
    nrec, nsamp = seis.shape[0], seis.shape[-1]
    if rec.shape[0] != nrec:
        raise ValueError("one (x, z) per receiver")
    xs, zs = float(source_xz[0]), float(source_xz[1])
    interval = float(np.float32(rec[1, 0] - rec[0, 0])) if nrec > 1 else 0.0
    comp = _CHANNEL_SUFFIX[seismotype]
    names = [("Up", 0)] if seismotype in (4, 6) else [("Ux", 0), ("Uz", 1)]
    paths = []
    for pre, c in names:
        p = os.path.join(out_dir, f"{pre}_file_single_{comp}.su")
        with open(p, "wb") as f:
            for r in range(nrec):
                f.write(su_header(r + 1, rec[r, 0], rec[r, 1], xs, zs, nsamp, dt * subsamp, interval))
                f.write(np.ascontiguousarray(seis[r, c], "<f4").tobytes())
        paths.append(p)
    return paths


def read_su(path: str):
    """Traces of an SU file written as above: (headers [ntrace] as dicts, samples [ntrace, nsamples] float32),
    read the way the reference's own helper does (EXAMPLES/BENCHMARK_CLAERBOUT_ADJOINT/helper_functions.py:367)."""
    raw = open(path, "rb").read()
    if len(raw) < 240:
        raise ValueError(f"{path}: no SU trace header")
    nsamp = int(np.frombuffer(raw[114:116], "<i2")[0])
    size = 240 + 4 * nsamp
    if nsamp <= 0 or len(raw) % size:
        raise ValueError(f"{path}: not a sequence of SU traces of {nsamp} samples")
    heads, data = [], []
    for k in range(len(raw) // size):
        b = raw[k * size:(k + 1) * size]
        h1 = np.frombuffer(b[:112], "<i4")
        h2 = np.frombuffer(b[112:116], "<i2")
        h3 = np.frombuffer(b[116:120], "<i2")
        h4 = np.frombuffer(b[120:240], "<f4")
        heads.append({"irec": int(h1[0]), "offset": int(h1[9]), "xs": int(h1[18]), "zs": int(h1[19]),
                      "xr": int(h1[20]), "zr": int(h1[21]), "nsamples": int(h2[1]), "dt_code": int(h3[0]),
                      "receiver_interval": float(h4[17])})
        data.append(np.frombuffer(b[240:], "<f4"))
    return heads, np.array(data)


def read_ascii_trace(path: str) -> Tuple[np.ndarray, np.ndarray]:
    t, v = np.loadtxt(path, unpack=True)
    return t, v
