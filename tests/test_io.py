"""The reference's on-disk formats (specfem2d_amd/io.py): Par_file / SOURCE / interfaces / STATIONS in,
ASCII and binary seismograms out.

* ASCII writer: byte-for-byte equal to the reference's own committed traces (sha256 of 21 files from
  ``EXAMPLES/*/REF_SEIS``, one per example and recording kind, regenerated from their parsed values);
* readers: the reference's Tromp2005, pressure-source and LuoYang fluid-solid example inputs give the
  configurations written by hand from their Par_files, and the oracle run from the parsed Tromp2005 files
  reproduces REF_SEIS bit-exactly;
* receiver lines: the STATIONS rounding (f20.7) of save_stations_file.f90.
"""
import hashlib
import os

import numpy as np
import pytest

from specfem2d_amd import config, io


def test_ascii_writer_is_byte_compatible(golden_dir):
    g = np.load(f"{golden_dir}/ref_seis_ascii_digests.npz")
    names = g["names"]
    assert len(names) >= 15
    for k, name in enumerate(names):
        text = io.ascii_trace_lines(g[f"t{k}"], g[f"v{k}"])
        assert hashlib.sha256(text.encode()).digest() == g[f"sha{k}"].tobytes(), name


def test_gfortran_real8_fields():
    f = io.gfortran_real8
    assert f(-6.6666666666666666e-2) == " -6.6666666666666666E-002"
    assert f(0.11293333333333334) == " 0.11293333333333334     "
    assert f(0.0) == "  0.0000000000000000     "
    assert f(-8.0) == " -8.0000000000000000     "
    assert f(1.0e17) == "  1.0000000000000000E+017"
    assert f(0.09999999999999999) == "  9.9999999999999992E-002"
    assert all(len(f(x)) == 25 for x in (1e-300, -1e300, 3.5, 123456.789))


def _same(cfg, ref, keys):
    for k in keys:
        assert np.all(np.asarray(getattr(cfg, k)) == np.asarray(getattr(ref, k))), k


KEYS = ("nx", "nz", "xmin", "xmax", "interfaces", "nz_layers", "nstep", "dt", "stacey", "absorbbottom",
        "absorbright", "absorbtop", "absorbleft", "user_t0", "anglerec", "seismotype", "use_trick")


def test_tromp2005_inputs_give_the_configuration(golden_dir):
    cfg, stations = io.config_from_par_file(os.path.join(golden_dir, "examples", "Tromp2005_kernel"))
    ref = config.tromp2005_kernel()
    _same(cfg, ref, KEYS + ("material_mode",))
    assert cfg.sources == ref.sources
    np.testing.assert_array_equal(cfg.receivers, ref.receivers)
    assert [(s.network, s.name) for s in stations] == [("AA", "S0001")]
    for a, b in zip(cfg.material_arrays(), ref.material_arrays()):
        np.testing.assert_array_equal(a, b)


def test_tromp2005_from_files_reproduces_ref_seis(golden_dir, tmp_path):
    from oracle.oracle import OracleSim
    cfg, stations = io.config_from_par_file(os.path.join(golden_dir, "examples", "Tromp2005_kernel"))
    orc = OracleSim(cfg, "sp")
    seis = orc.forward()
    g = np.load(f"{golden_dir}/tromp2005_kernel_ref_seis.npz")
    np.testing.assert_array_equal(seis[0, 0], g["AA.S0001.BXX"])
    np.testing.assert_array_equal(seis[0, 1], g["AA.S0001.BXZ"])
    paths = io.write_ascii_seismograms(str(tmp_path), seis, stations, cfg.dt, orc.t0, seismotype=1)
    assert [os.path.basename(p) for p in paths] == ["AA.S0001.BXX.semd", "AA.S0001.BXZ.semd"]
    t, v = io.read_ascii_trace(paths[0])
    np.testing.assert_array_equal(t, g["time"])
    np.testing.assert_array_equal(v, g["AA.S0001.BXX"])


def test_pressure_inputs_give_the_configuration(golden_dir):
    d = os.path.join(golden_dir, "examples", "pressure_source_acoustic")
    with pytest.raises(NotImplementedError):
        io.config_from_par_file(d, cpml="error")  # the example asks for CPML
    cfg, stations = io.config_from_par_file(d, cpml="stacey")
    ref = config.pressure_source_acoustic_44x44()
    _same(cfg, ref, KEYS)
    assert cfg.sources == ref.sources
    np.testing.assert_array_equal(cfg.receivers, ref.receivers)
    for a, b in zip(cfg.material_arrays(), ref.material_arrays()):
        np.testing.assert_array_equal(a, b)


def test_receiver_lines_round_like_stations_file(tmp_path):
    par = tmp_path / "Par_file"
    par.write_text("nreceiversets = 2\nnrec = 4\nxdeb = 0.d0\nzdeb = -1.d0\nxfin = 1000.d0\nzfin = -1.d0\n"
                   "record_at_surface_same_vertical = .false.\nnrec = 1\nxdeb = 12.5\nzdeb = -7.\nxfin = 0\n"
                   "zfin = 0\nrecord_at_surface_same_vertical = .false.\n")
    st = io.stations_from_receiver_sets(io.read_par_file(str(par)))
    assert [s.name for s in st] == ["S0001", "S0002", "S0003", "S0004", "S0005"]
    assert st[1].x == float("333.3333333") and st[2].x == float("666.6666667")
    assert (st[4].x, st[4].z) == (12.5, -7.0)


def test_binary_writer_layout(tmp_path):
    seis = np.arange(3 * 2 * 5, dtype=np.float64).reshape(3, 2, 5)
    px, pz = io.write_binary_seismograms(str(tmp_path), seis, seismotype=1)
    assert os.path.basename(px) == "Ux_file_single_d.bin"
    np.testing.assert_array_equal(np.fromfile(pz, np.float32).reshape(3, 5), seis[:, 1, :])
    (pp,) = io.write_binary_seismograms(str(tmp_path), seis, seismotype=4, double=True)
    assert os.path.basename(pp) == "Up_file_double_p.bin"
    np.testing.assert_array_equal(np.fromfile(pp, np.float64).reshape(3, 5), seis[:, 0, :])


def test_luoyang_inputs_give_the_configuration(golden_dir):
    """A fluid-solid example with 9-node elements, three regions and a potential receiver (seismotype 6)."""
    cfg, stations = io.config_from_par_file(os.path.join(golden_dir, "examples", "LuoYang_fluid_solid_kernel"))
    ref = config.luoyang_fluid_solid()
    _same(cfg, ref, KEYS)
    assert cfg.sources == ref.sources
    np.testing.assert_array_equal(cfg.receivers, ref.receivers)
    for a, b in zip(cfg.material_arrays(), ref.material_arrays()):
        np.testing.assert_array_equal(a, b)


def test_su_writer_layout(tmp_path):
    """SU files as write_output_SU.f90 lays them out: per receiver 60 header words then the float32 samples; header
    fields at the words the reference's own reader takes them from (irec 1, offset 10, xs 19, zs 20, xr 21, zr 22,
    samples in the second int16 of word 29, the sampling interval in us in the first int16 of word 30, the receiver
    interval as float word 18 of the last 30), Fortran NINT rounding (halves away from zero)."""
    rng = np.random.default_rng(4)
    nrec, nstep = 5, 333
    seis = rng.standard_normal((nrec, 2, nstep))
    rec = np.array([[100.5 + 200.0 * k, -2.5] for k in range(nrec)])
    src = (1000.0, -500.5)
    paths = io.write_su_seismograms(str(tmp_path), seis, rec, src, dt=1.1e-3, seismotype=1)
    assert [os.path.basename(p) for p in paths] == ["Ux_file_single_d.su", "Uz_file_single_d.su"]
    for c, p in enumerate(paths):
        raw = open(p, "rb").read()
        assert len(raw) == nrec * (240 + 4 * nstep)
        words = np.frombuffer(raw, "<i4").reshape(nrec, 60 + nstep)
        for r in range(nrec):
            assert words[r, 0] == r + 1
            assert words[r, 9] == io._nint(rec[r, 0] - src[0])
            assert tuple(words[r, 18:22]) == (1000, -501, io._nint(rec[r, 0]), -3)
        heads, data = io.read_su(p)
        assert [h["nsamples"] for h in heads] == [nstep] * nrec
        assert all(h["dt_code"] == 1100 for h in heads)
        assert all(h["receiver_interval"] == 200.0 for h in heads)
        np.testing.assert_array_equal(data, seis[:, c, :].astype(np.float32))
    # the samples are the single-precision binary streams with headers interleaved
    b = io.write_binary_seismograms(str(tmp_path), seis, seismotype=1)
    np.testing.assert_array_equal(np.fromfile(b[0], np.float32).reshape(nrec, nstep), io.read_su(paths[0])[1])
    # pressure: one file; intervals of 40 ms and longer are written in ms
    p = io.write_su_seismograms(str(tmp_path), seis[:, :1], rec, src, dt=0.04, seismotype=4)
    assert [os.path.basename(x) for x in p] == ["Up_file_single_p.su"]
    assert io.read_su(p[0])[0][0]["dt_code"] == 40
    assert io._nint(2.5) == 3 and io._nint(-2.5) == -3 and io._nint(-0.4) == 0
