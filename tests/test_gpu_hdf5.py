"""FeVariable checkpoint files of a GPU solution (FeVariable_SaveToFile / FeVariable_ReadFromFile,
StgFEM/Discretisation/src/FeVariable.c:2136-2389, :2452-2730): the solved fields written and read back, their
structure equal to the reference's files of the same mesh (the lowres Rayleigh-Taylor run's fields, 10 x 10),
the coordinate columns equal to those of the reference's version-1 files (PeriodicBenchmark, 32 x 32 on the
unit square), a version-1 file loaded into the solution, and the reader's refusals."""
import os
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h5")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def _ust():
    import earthbyte_underworld_amd as ust
    try:
        ust.read_checkpoint(os.path.join(GOLD, "RayleighTaylorBenchmark_lowres_PressureField.00005.h5"))
    except ust.UstError as e:
        pytest.skip(f"no usable HDF5 C library: {e}")
    return ust


def _header(path):
    return subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout.split("\n", 1)[1]


def test_save_load_rt_lowres(tmp_path):
    ust = _ust()
    from earthbyte_underworld_amd.configs import rayleigh_taylor_2d
    s = rayleigh_taylor_2d(nel=(10, 10))
    s.assemble(gravity=1.0)
    st = s.solve(outer_rtol=1e-9, inner_rtol=1e-10)
    assert st["stokes_residual"] < 1e-7
    u, p = s.solution()
    T = np.linspace(0.0, 1.0, s.nnodes)
    s.set_temperature(T)
    files = {}
    for name in ("VelocityField", "PressureField", "TemperatureField"):
        for coords in (False, True):
            files[name, coords] = str(tmp_path / f"{name}.{int(coords)}.h5")
            s.save_field(name, files[name, coords], saveCoords=coords)
    vals = {"VelocityField": u.reshape(-1, 2), "PressureField": p.reshape(-1, 1), "TemperatureField": T.reshape(-1, 1)}
    xy, cc = s.node_coords(), s.cell_centres()
    for (name, coords), f in files.items():
        r = ust.read_checkpoint(f)
        assert r["version"] == 2 and r["dimensions"] == 2
        assert r["mesh resolution"] == (None if name == "PressureField" else [10, 10])
        d = r["data"]
        assert np.array_equal(d[:, 2 if coords else 0:], vals[name])
        if coords:
            assert np.allclose(d[:, :2], cc if name == "PressureField" else xy, rtol=0, atol=1e-15)
    if H5DUMP:  # the reference's files of the same 10 x 10 mesh have the same structure
        assert _header(files["VelocityField", False]) == \
            _header(os.path.join(GOLD, "RayleighTaylorBenchmark_lowres_VelocityField.00005.h5"))
        assert _header(files["PressureField", False]) == \
            _header(os.path.join(GOLD, "RayleighTaylorBenchmark_lowres_PressureField.00005.h5"))
    # read back into a fresh context of the same mesh
    s2 = rayleigh_taylor_2d(nel=(10, 10))
    s2.load_field("VelocityField", files["VelocityField", True])
    s2.load_field("PressureField", files["PressureField", False])
    u2, p2 = s2.solution()
    assert np.array_equal(u2, u) and np.array_equal(p2, p)
    # the reference's own fields load into the mesh they were written on
    s2.load_field("VelocityField", os.path.join(GOLD, "RayleighTaylorBenchmark_lowres_VelocityField.00005.h5"))
    ref = ust.read_checkpoint(os.path.join(GOLD, "RayleighTaylorBenchmark_lowres_VelocityField.00005.h5"))["data"]
    assert np.array_equal(s2.solution()[0], ref.reshape(-1))
    # refusals: another resolution, another dimension, a pressure file for the velocity
    s3 = rayleigh_taylor_2d(nel=(12, 10))
    with pytest.raises(ust.UstError, match="mesh resolution"):
        s3.load_field("VelocityField", files["VelocityField", False])
    with pytest.raises(ust.UstError, match="dimensions"):
        s3.load_field("VelocityField", os.path.join(GOLD, "CylinderRiseThermal_lowres_VelocityField.00005.h5"))
    with pytest.raises(ust.UstError, match="no mesh resolution"):
        s2.load_field("VelocityField", files["PressureField", False])
    with pytest.raises(ust.UstError, match="nodes"):  # version 1: no attributes, the node count decides
        s2.load_field("VelocityField", os.path.join(GOLD, "PeriodicBenchmark_VelocityField.00010.h5"))
    for x in (s, s2, s3):
        x.close()


def test_version1_files_and_coordinates(tmp_path):
    """The reference's version-1 files (rows: node index, coordinates, values) load into a 32 x 32 unit-square
    mesh; the coordinates this library writes equal the reference's coordinate columns exactly."""
    ust = _ust()
    s = ust.StokesSystem(2, (32, 32), (0.0, 0.0), (1.0, 1.0))
    v1 = ust.read_checkpoint(os.path.join(GOLD, "PeriodicBenchmark_VelocityField.00010.h5"))["data"]
    p1 = ust.read_checkpoint(os.path.join(GOLD, "PeriodicBenchmark_PressureField.00010.h5"))["data"]
    s.load_field("VelocityField", os.path.join(GOLD, "PeriodicBenchmark_VelocityField.00010.h5"))
    s.load_field("PressureField", os.path.join(GOLD, "PeriodicBenchmark_PressureField.00010.h5"))
    u, p = s.solution()
    assert np.array_equal(u.reshape(-1, 2), v1[:, 3:5]) and np.array_equal(p, p1[:, 3])
    for name, ref in (("VelocityField", v1), ("PressureField", p1)):
        f = str(tmp_path / f"{name}.h5")
        s.save_field(name, f, saveCoords=True)
        d = ust.read_checkpoint(f)["data"]
        assert np.array_equal(d[:, :2], ref[:, 1:3]) and np.array_equal(d[:, 2:], ref[:, 3:])
    s.close()
