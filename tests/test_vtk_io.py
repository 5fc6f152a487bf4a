"""VTK XML reader/writer (vtkxml) and the checkpoint format of fs5PlotUtils
(save_grid_and_particles_vti_vtp / load_restart_from_vtp, reference
fs5PlotUtils.py:253-646), on the host.

The decoder is pinned by a file that VTK itself wrote (tests/golden): every
array must reproduce the RangeMin/RangeMax that VTK stored beside it (the
component range of a scalar, the L2-norm range of a vector).  The writer is
checked by round trips, bit for bit."""
import base64
import os
import re
import zlib

import numpy as np
import pytest

from mpmwarpydw_amd import fs5PlotUtils, vtkxml

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "coupling_column_same.vtp")


def test_reads_a_file_written_by_vtk():
    d = vtkxml.read(GOLDEN)
    text = open(GOLDEN, "rb").read().decode("latin1")
    assert d["points"].shape == (32000, 3) and d["points"].dtype == np.float64
    ranges = re.findall(r'<DataArray type="(\w+)" Name="([^"]+)"[^>]*RangeMin="([^"]+)" RangeMax="([^"]+)"', text)
    checked = 0
    for typ, name, lo, hi in ranges:
        if name in ("connectivity", "offsets"):
            continue
        v = d["points"] if name == "Points" else d["point_data"][name]
        assert v.dtype == np.dtype(vtkxml._DTYPES[typ])
        m = np.sqrt((v.astype(np.float64) ** 2).sum(axis=1)) if v.ndim == 2 else v.astype(np.float64)
        assert m.min() == pytest.approx(float(lo), rel=1e-14) and m.max() == pytest.approx(float(hi), rel=1e-14), name
        checked += 1
    assert checked == 13  # the points and twelve point arrays
    conn, off = d["verts"]
    assert np.array_equal(conn, np.arange(32000)) and np.array_equal(off, np.arange(1, 32001))


def test_domain_from_vtp():
    dom = fs5PlotUtils.load_domain_vtp(GOLDEN)
    assert dom["x"].shape == (32000, 3) and dom["particle_volume"].shape == (32000,)
    assert {"density", "E", "nu", "ys", "alpha", "strainCriteria"} <= set(dom)


@pytest.mark.parametrize("header", ["UInt32", "UInt64"])
@pytest.mark.parametrize("compress", [True, False])
def test_round_trip_all_types(tmp_path, header, compress):
    rng = np.random.default_rng(1)
    n = 20000  # several 32 KiB blocks for the wide types
    pts = rng.standard_normal((n, 3)).astype(np.float32)
    arrays = {"f32": rng.standard_normal(n).astype(np.float32), "f64": rng.standard_normal(n),
              "i32": rng.integers(-5, 5, n).astype(np.int32), "i64": rng.integers(0, 1 << 40, n),
              "u8": rng.integers(0, 255, n).astype(np.uint8), "v3": rng.standard_normal((n, 3)).astype(np.float32),
              "m9": rng.standard_normal((n, 9)).astype(np.float32),
              "special": np.array([np.nan, np.inf, -0.0] + [1.0] * (n - 3), np.float32)}
    field = {"t": np.array([0.123456789012345678]), "k": np.array([3], np.int32)}
    p = str(tmp_path / "a.vtp")
    vtkxml.write_vtp(p, pts, arrays, field, verts=True, header_type=header, compress=compress)
    d = vtkxml.read(p)
    assert np.array_equal(d["points"], pts)
    for k, v in arrays.items():
        got = d["point_data"][k]
        assert got.dtype == v.dtype and got.shape == v.shape
        assert np.array_equal(got.view(np.uint8), v.view(np.uint8)), k  # bit for bit, NaN included
    assert d["field_data"]["t"][0] == field["t"][0] and d["field_data"]["k"][0] == 3


def test_empty_and_exact_block_sizes(tmp_path):
    for n in (0, 1, 8192, 8193):  # 8192 float32 = exactly one 32 KiB block
        pts = np.zeros((n, 3), np.float32)
        a = np.arange(n, dtype=np.float32)
        p = str(tmp_path / f"e{n}.vtp")
        vtkxml.write_vtp(p, pts, {"a": a})
        d = vtkxml.read(p)
        assert d["points"].shape == (n, 3) and np.array_equal(d["point_data"]["a"], a)


def _appended_file(values: np.ndarray, raw: bool, compress: bool) -> bytes:
    """A PolyData with its arrays in an AppendedData section, laid out as the
    VTK XML format defines it (offsets from the byte after '_')."""
    hdt = np.dtype("<u4")
    pts = np.zeros((values.shape[0], 3), np.float32)

    def payload(arr):
        b = arr.tobytes()
        if compress:
            c = zlib.compress(b)
            head = np.array([1, len(b), len(b), len(c)], hdt).tobytes()
            return (head + c) if raw else (base64.b64encode(head) + base64.b64encode(c))
        head = np.array([len(b)], hdt).tobytes()
        return (head + b) if raw else base64.b64encode(head + b)

    pa, pb = payload(values), payload(pts)
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    enc = "raw" if raw else "base64"
    n = values.shape[0]
    xml = (f'<?xml version="1.0"?>\n<VTKFile type="PolyData" version="0.1" byte_order="LittleEndian" '
           f'header_type="UInt32"{comp}>\n<PolyData>\n<Piece NumberOfPoints="{n}" NumberOfVerts="0" '
           f'NumberOfLines="0" NumberOfStrips="0" NumberOfPolys="0">\n<PointData>\n'
           f'<DataArray type="Float64" Name="a" format="appended" offset="0"/>\n</PointData>\n<Points>\n'
           f'<DataArray type="Float32" Name="Points" NumberOfComponents="3" format="appended" offset="{len(pa)}"/>\n'
           f'</Points>\n</Piece>\n</PolyData>\n<AppendedData encoding="{enc}">\n_').encode()
    return xml + pa + pb + b"\n</AppendedData>\n</VTKFile>\n"


@pytest.mark.parametrize("raw", [True, False])
@pytest.mark.parametrize("compress", [True, False])
def test_appended_data(tmp_path, raw, compress):
    v = np.random.default_rng(2).standard_normal(1000)
    p = tmp_path / "app.vtp"
    p.write_bytes(_appended_file(v, raw, compress))
    d = vtkxml.read(str(p))
    assert np.array_equal(d["point_data"]["a"], v) and d["points"].shape == (1000, 3)


def test_checkpoint_round_trip_is_bit_exact(tmp_path):
    """save_grid_and_particles_vti_vtp with restart data, then
    load_restart_from_vtp: every restart array comes back bit for bit, the
    stress from its 6 symmetric components, time and dx in FP64."""
    rng = np.random.default_rng(3)
    n = 5000
    f = lambda *s: rng.standard_normal(s).astype(np.float32)  # noqa: E731
    S = f(n, 3, 3)
    S = (S + S.transpose(0, 2, 1)) / 2
    restart = {"sim_time": 1.2345678901234567, "bigStep": 7, "dx": 0.4999999999, "particle_F": f(n, 3, 3),
               "particle_C": f(n, 3, 3), "particle_accumulated_strain": f(n), "ys_base": f(n),
               "particle_vol": f(n), "particle_mass": f(n), "particle_density": f(n),
               "particle_x_initial_xpbd": f(n, 3), "particle_cumDist_xpbd": f(n)}
    for prop in fs5PlotUtils.MATERIAL_PROPS:
        restart[prop] = f(n)
    x, v, r, ys, dmg = f(n, 3), f(n, 3), f(n), f(n), f(n)
    act = rng.integers(0, 2, n).astype(np.int32)
    lab = rng.integers(0, 3, n).astype(np.int32)
    grid = rng.random((7, 5, 6)).astype(np.float32)
    prefix = str(tmp_path / "out" / "sim_step_x")
    fs5PlotUtils.save_grid_and_particles_vti_vtp(prefix, grid, (1.0, 2.0, 3.0), 0.5, x, particle_radius=r,
                                                 particle_velocity=v, particle_ys=ys, particle_damage=dmg,
                                                 particle_stress_tensor=S, particle_active_label=act,
                                                 particle_material_label=lab, restart_data=restart)
    got = fs5PlotUtils.load_restart_from_vtp(prefix + "_particles.vtp")
    same = lambda a, b: np.array_equal(np.asarray(a).ravel().view(np.uint8), np.asarray(b).ravel().view(np.uint8))  # noqa
    for k in ("particle_F", "particle_C", "particle_accumulated_strain", "ys_base", "particle_vol", "particle_mass",
              "particle_density", "particle_x_initial_xpbd", "particle_cumDist_xpbd", *fs5PlotUtils.MATERIAL_PROPS):
        assert same(got[k], restart[k]), k
    for k, want in (("particle_x", x), ("particle_v", v), ("particle_radius", r), ("ys", ys),
                    ("particle_damage", dmg), ("activeLabel", act), ("materialLabel", lab), ("particle_stress", S)):
        assert same(got[k], want), k
    assert got["sim_time"] == restart["sim_time"] and got["dx"] == restart["dx"] and got["bigStep"] == 7
    g = vtkxml.read(prefix + "_grid.vti")
    assert g["extent"] == [0, 6, 0, 4, 0, 5] and g["origin"] == [1.0, 2.0, 3.0] and g["spacing"] == [0.5] * 3
    assert np.array_equal(g["point_data"]["grid_mass"].reshape(6, 5, 7).transpose(2, 1, 0), grid)


def test_restart_defaults_of_a_snapshot_without_restart_data(tmp_path):
    n = 10
    x = np.zeros((n, 3), np.float32)
    prefix = str(tmp_path / "s")
    fs5PlotUtils.save_grid_and_particles_vti_vtp(prefix, np.zeros((3, 3, 3)), (0, 0, 0), 1.0, x,
                                                 particle_ys=np.full(n, 5.0, np.float32))
    got = fs5PlotUtils.load_restart_from_vtp(prefix + "_particles.vtp")
    assert np.array_equal(got["particle_F"], np.tile(np.eye(3, dtype=np.float32), (n, 1, 1)))
    assert np.array_equal(got["particle_C"], np.zeros((n, 3, 3)))
    assert np.array_equal(got["ys_base"], got["ys"]) and got["bigStep"] == 0 and got["sim_time"] == 0.0
    assert np.array_equal(got["particle_radius"], np.full(n, 0.5, np.float32))


def test_hdf5_domain_matches_its_vtp_twin(tmp_path):
    """The reference's HDF5 domain (h5py-written) and its .vtp twin (VTK-written)
    hold the same particles: the two readers agree bit for bit."""
    import gzip

    from mpmwarpydw_amd import hdf5
    h5 = tmp_path / "d.h5"
    h5.write_bytes(gzip.decompress(open(GOLDEN.replace(".vtp", ".h5.gz"), "rb").read()))
    ds = hdf5.read_datasets(str(h5))
    assert ds["x"].shape == (3, 32000) and ds["x"].dtype == np.float64
    assert set(ds) == {"x", "particle_volume", "density", "E", "nu", "ys", "alpha", "hardening", "softening",
                       "eta_shear", "eta_bulk", "strainCriteria"}
    a = hdf5.load_domain_h5(str(h5))
    b = fs5PlotUtils.load_domain_vtp(GOLDEN)
    assert set(a) == set(b)
    for k in a:
        assert a[k].dtype == b[k].dtype and np.array_equal(a[k], b[k]), k


def test_hdf5_rejects_what_it_does_not_read(tmp_path):
    from mpmwarpydw_amd import hdf5
    p = tmp_path / "x.h5"
    p.write_bytes(b"not an hdf5 file" * 100)
    with pytest.raises(hdf5.HDF5FormatError):
        hdf5.read_datasets(str(p))
