"""Field-data ingest: shp_field_mesh_load_csv reads the reference's
Fields_*.csv / Triangles_*.csv files as InterpolationScheme does
(interpolation_scheme.cpp:66-172, line counts as returnSize
auxiliary_functions.cpp:39-53).  CPU only: the loader does no device work.

* Round trip: a mesh written in the reference's layout loads back bit for bit
  (17 significant digits and atof are exact for doubles).
* The field on the loaded arrays equals the reference's own InterpolationScheme,
  built from its sources into oracle/_ref, on the same files (bit for bit).
* Record rules the reference's parser fixes: fields split on ',' only, extra
  columns ignored, missing ones left at zero, atoi truncation of the triangle
  ids, a line longer than the 1024-byte buffer counted once per chunk.
"""
import ctypes

import numpy as np
import pytest

import oracle
import solveit_hip
from solveit_hip import synth


@pytest.fixture(scope="module")
def mesh():
    return synth.tip_field_mesh(nr=60, nz=50)


def _load(tmp_path, fields, tris, r0=1e-8, E0=2e9):
    return solveit_hip.load_field_mesh_csv(tmp_path / fields, tmp_path / tris, r0, E0)


def test_round_trip_is_exact(mesh, tmp_path):
    fields, tris = synth.write_mesh_csv(mesh, str(tmp_path))
    got = _load(tmp_path, fields, tris, mesh["r0"], mesh["E0"])
    for k in ("xy", "er", "ez"):
        assert got[k].dtype == np.float64
        np.testing.assert_array_equal(got[k], mesh[k])
    for k in ("tri", "nbr"):
        np.testing.assert_array_equal(got[k], np.asarray(mesh[k], np.int32))
    assert got["r0"] == mesh["r0"] and got["E0"] == mesh["E0"]


def test_loaded_mesh_gives_reference_field(mesh, tmp_path):
    L = oracle.ref_efield_lib()
    if L is None:
        pytest.skip("reference InterpolationScheme build (oracle/_ref) not present")
    fields, tris = synth.write_mesh_csv(mesh, str(tmp_path))
    r0, E0 = np.float32(mesh["r0"]), np.float32(mesh["E0"])
    got = _load(tmp_path, fields, tris, float(r0), float(E0))
    L.ref_efield_init((str(tmp_path) + "/").encode(), fields.encode(), tris.encode(), r0, E0)
    xy = np.ascontiguousarray(got["xy"])
    er, ez = np.ascontiguousarray(got["er"]), np.ascontiguousarray(got["ez"])
    tri, nbr = np.ascontiguousarray(got["tri"]), np.ascontiguousarray(got["nbr"])
    # prime both warm starts with the same interior point
    p0 = np.array([1e-6, 0.5e-6, 2e-6], np.float32)
    out = np.zeros(3, np.float32)
    L.ref_efield(p0[0], p0[1], p0[2], out.ctypes.data)
    e, t, _, _ = oracle.tri_search(got, np.array([np.hypot(p0[0] / r0, p0[1] / r0), p0[2] / r0]), start=-1)
    assert e == 0
    ti = ctypes.c_int(t)
    d = synth.cone(300, seed=5)
    pos = np.stack([d["x"], d["y"], d["z"]], 1).astype(np.float32)
    pos[:4, 2] *= -1
    mine = np.zeros(3, np.float32)
    for p in pos:
        L.ref_efield(p[0], p[1], p[2], out.ctypes.data)
        oracle.lib.orc_efield_ref(p[0], p[1], p[2], r0, E0, xy.ctypes.data, tri.shape[0], tri.ctypes.data,
                                  nbr.ctypes.data, er.ctypes.data, ez.ctypes.data, ctypes.addressof(ti),
                                  mine.ctypes.data)
        np.testing.assert_array_equal(mine, out)


def test_record_rules(tmp_path):
    (tmp_path / "F.csv").write_text("1.5,2.25,3,4,99\n"      # extra column ignored
                                    "5e-1, 6 ,7\n"           # missing Ez stays 0; atof skips leading blanks
                                    "-1,-2,-3,-4\n")
    (tmp_path / "T.csv").write_text("0,1,2,-1,2.9,1\n"       # atoi truncates 2.9
                                    "2,1,0,0\n")             # missing neighbours stay 0
    got = _load(tmp_path, "F.csv", "T.csv")
    np.testing.assert_array_equal(got["xy"], [[1.5, 2.25], [0.5, 6.0], [-1.0, -2.0]])
    np.testing.assert_array_equal(got["er"], [3.0, 7.0, -3.0])
    np.testing.assert_array_equal(got["ez"], [4.0, 0.0, -4.0])
    np.testing.assert_array_equal(got["tri"], [[0, 1, 2], [2, 1, 0]])
    np.testing.assert_array_equal(got["nbr"], [[-1, 2, 1], [0, 0, 0]])


def test_long_line_counts_per_buffer_chunk(tmp_path):
    # 1023 characters fill one fgets call; the rest of the line is a second record
    long_line = "1,2,3,4," + "0" * 1100 + "\n"
    (tmp_path / "F.csv").write_text(long_line + "5,6,7,8\n")
    (tmp_path / "T.csv").write_text("")
    got = _load(tmp_path, "F.csv", "T.csv")
    assert got["xy"].shape == (3, 2) and got["tri"].shape == (0, 3)
    np.testing.assert_array_equal(got["xy"][0], [1.0, 2.0])
    np.testing.assert_array_equal(got["xy"][2], [5.0, 6.0])


def test_missing_file_is_an_error(tmp_path):
    (tmp_path / "T.csv").write_text("0,1,2,-1,-1,-1\n")
    with pytest.raises(solveit_hip.ShpError, match="cannot open"):
        _load(tmp_path, "nope.csv", "T.csv")
    with pytest.raises(solveit_hip.ShpError, match="cannot open"):
        solveit_hip.load_field_mesh_csv(tmp_path / "T.csv", tmp_path / "nope.csv", 1.0, 1.0)
