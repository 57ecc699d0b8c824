"""HDF5 field and trajectory files (the reference's on-disk format,
field.py:144-151, field.py:409-431, field.py:718-736, equation.py:362-372).

Pinned against h5py itself: fixtures written by h5py in the reference's calls
(tests/golden/make_h5_fixtures.py) read back exactly, and files written here
are read by h5py (libhdf5) in a separate interpreter when one with h5py is
installed (this image: /opt/conda/bin/python3.9), else those checks skip."""

import os
import shutil
import subprocess

import numpy as np
import pytest
import torch

from jax_spectral_dns_amd import hdf5
from jax_spectral_dns_amd.domain import PhysicalDomain
from jax_spectral_dns_amd.field import Field, FourierField, PhysicalField, VectorField

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
H5PY_PYTHON = "/opt/conda/bin/python3.9"


def _h5py_available() -> bool:
    if not os.path.exists(H5PY_PYTHON):
        return False
    try:
        r = subprocess.run([H5PY_PYTHON, "-c", "import h5py"], capture_output=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired):
        return False
    return r.returncode == 0


def _h5py_dump(path: str, out_npz: str) -> None:
    """Every dataset of the file (paths joined by '|') and every attribute
    ('path@name'), as read by h5py, into an .npz."""
    script = r"""
import sys, h5py, numpy as np
out = {}
def visit(name, obj):
    if isinstance(obj, h5py.Dataset):
        out[name.replace('/', '|')] = obj[()]
        for k, v in obj.attrs.items():
            out[name.replace('/', '|') + '@' + k] = np.asarray(v)
with h5py.File(sys.argv[1], 'r') as f:
    f.visititems(visit)
    out['__groups__'] = np.array(sorted(k for k in f.keys()))
np.savez(sys.argv[2], **out)
"""
    r = subprocess.run([H5PY_PYTHON, "-c", script, path, out_npz], capture_output=True, timeout=300)
    assert r.returncode == 0, r.stderr.decode()[-2000:]


needs_h5py = pytest.mark.skipif(not _h5py_available(), reason="no interpreter with h5py installed")


def test_reads_fixtures_written_by_h5py():
    exp = np.load(os.path.join(GOLD, "h5_fixtures_expected.npz"))
    t = hdf5.read(os.path.join(GOLD, "h5_vector_field.h5"))
    assert hdf5.names(t) == ["12"]
    d = t["12"]["velocity_hat"]
    assert d.data.dtype == np.complex128 and np.array_equal(d.data, exp["vec"]) and d.attrs == {}
    t = hdf5.read(os.path.join(GOLD, "h5_physical_field.h5"))
    d = t["3"]["velocity_x"]
    assert d.data.dtype == np.float64 and np.array_equal(d.data, exp["phys"])
    assert int(d.attrs["time_step"]) == 3 and np.asarray(d.attrs["time_step"]).shape == ()
    t = hdf5.read(os.path.join(GOLD, "h5_trajectory.h5"))
    assert np.array_equal(t["trajectory"].data, exp["traj"])
    assert tuple(exp["traj_chunks"]) != exp["traj"].shape  # several chunks were read


def test_round_trip_dtypes_chunks_and_groups(tmp_path):
    rng = np.random.default_rng(5)
    big = rng.standard_normal((33, 20, 17))
    tree = {
        "c16": hdf5.Dataset(rng.standard_normal((4, 7)) + 1j * rng.standard_normal((4, 7)), {"time_step": np.int64(-4)}),
        "c8": (rng.standard_normal(9) + 1j).astype(np.complex64),
        "f4": rng.standard_normal((3, 3)).astype(np.float32),
        "i4": np.arange(-5, 5, dtype=np.int32),
        "u8": np.arange(7, dtype=np.uint64),
        "scalar": np.float64(2.5),
        "empty_group": {},
        # > 64 chunks: a two-level chunk B-tree
        "many_chunks": hdf5.Dataset(big, {"v": np.arange(3.0)}),
        # > 8 entries: several symbol-table nodes; > 256: a two-level group B-tree
        "wide": {"e%03d" % i: np.float64(i) for i in range(300)},
        "deep": {"a": {"b": {"c": np.int64(1)}}},
    }
    fn = str(tmp_path / "rt.h5")
    hdf5.write(fn, tree, compression=6, chunk_bytes=1000)
    t = hdf5.read(fn)
    assert hdf5.names(t) == sorted(tree.keys())
    assert np.array_equal(t["c16"].data, tree["c16"].data) and int(t["c16"].attrs["time_step"]) == -4
    for k in ("c8", "f4", "i4", "u8"):
        assert t[k].data.dtype == tree[k].dtype and np.array_equal(t[k].data, tree[k])
    assert t["scalar"].data.shape == () and float(t["scalar"].data) == 2.5
    assert t["empty_group"] == {}
    assert np.array_equal(t["many_chunks"].data, big)
    assert np.array_equal(t["many_chunks"].attrs["v"], np.arange(3.0))
    assert len(t["wide"]) == 300 and all(float(t["wide"]["e%03d" % i].data) == i for i in range(300))
    assert int(t["deep"]["a"]["b"]["c"].data) == 1
    # uncompressed: contiguous datasets
    hdf5.write(fn, {"x": big}, compression=None)
    assert np.array_equal(hdf5.read(fn)["x"].data, big)


@needs_h5py
def test_h5py_reads_files_written_here(tmp_path):
    rng = np.random.default_rng(7)
    vec = rng.standard_normal((3, 5, 9, 5)) + 1j * rng.standard_normal((3, 5, 9, 5))
    big = rng.standard_normal((33, 20, 17))
    fn = str(tmp_path / "ours.h5")
    hdf5.write(
        fn,
        {
            "12": {"velocity_hat": hdf5.Dataset(vec, {"time_step": np.int64(12)})},
            "many": big,
            "wide": {"e%03d" % i: np.int64(i) for i in range(300)},
            "f4": big[0].astype(np.float32),
        },
        chunk_bytes=1000,
    )
    out = str(tmp_path / "read.npz")
    _h5py_dump(fn, out)
    r = np.load(out)
    assert list(r["__groups__"]) == ["12", "f4", "many", "wide"]
    assert np.array_equal(r["12|velocity_hat"], vec) and r["12|velocity_hat"].dtype == np.complex128
    assert int(r["12|velocity_hat@time_step"]) == 12 and r["12|velocity_hat@time_step"].shape == ()
    assert np.array_equal(r["many"], big)
    assert np.array_equal(r["f4"], big[0].astype(np.float32))
    assert all(int(r["wide|e%03d" % i]) == i for i in range(300))


@needs_h5py
def test_h5py_round_trip_matches_fixture_bytes_semantics(tmp_path):
    """Our reading of a fixture and h5py's reading of our rewrite of it agree."""
    t = hdf5.read(os.path.join(GOLD, "h5_trajectory.h5"))
    fn = str(tmp_path / "traj.h5")
    hdf5.write(fn, t)
    out = str(tmp_path / "traj.npz")
    _h5py_dump(fn, out)
    assert np.array_equal(np.load(out)["trajectory"], t["trajectory"].data)


def test_field_save_to_file_and_from_file(tmp_path, monkeypatch):
    """save_to_file / FromFile of physical, spectral and vector fields, names
    relative to Field.field_dir, and negative time steps counting from the
    last group."""
    monkeypatch.setattr(Field, "field_dir", str(tmp_path / "fields"))
    dom = PhysicalDomain.create((8, 17, 8), (True, False, True), scale_factors=(2 * np.pi, 1.0, np.pi))
    f = PhysicalField.FromFunc(dom, lambda X: np.sin(X[0]) * (1 - X[1] ** 2) * np.cos(2 * X[2]), name="velocity_x")
    f.set_time_step(5)
    f.save_to_file("u_field")
    t = hdf5.read(str(tmp_path / "fields" / "u_field"))
    assert hdf5.names(t) == ["5"] and int(t["5"]["velocity_x"].attrs["time_step"]) == 5
    g = PhysicalField.FromFile(dom, "u_field", name="velocity_x")
    assert torch.equal(g.get_data(), f.get_data())
    h = f.hat()
    h.set_time_step(2)
    h.save_to_file(str(tmp_path / "u_hat.h5"))
    h2 = FourierField.FromFile(dom, str(tmp_path / "u_hat.h5"), name=h.name, time_step=2)
    assert torch.equal(h2.get_data(), h.get_data())
    v = VectorField([h, h * 2.0, h * -1.0], "velocity_hat")
    v.set_time_step(11)
    v.save_to_file("v_hat")
    w = VectorField.FromFile(dom, "v_hat", name="velocity_hat")
    assert w.get_time_step() == 11 and isinstance(w[0], FourierField)
    assert torch.equal(w.get_data(), v.get_data())
    data, ts = VectorField.read_hdf(str(tmp_path / "fields" / "v_hat"), "velocity_hat", -1)
    assert ts == 11 and torch.equal(data, v.get_data())
    with pytest.raises(AssertionError):
        VectorField.read_hdf(str(tmp_path / "fields" / "v_hat"), "velocity_hat", 3)
    # a reference-written vector field file opens through the same entry point
    shutil.copy(os.path.join(GOLD, "h5_vector_field.h5"), str(tmp_path / "fields" / "ref_vec"))
    data, ts = VectorField.read_hdf(str(tmp_path / "fields" / "ref_vec"), "velocity_hat", -1)
    assert ts == 12 and np.array_equal(data.numpy(), np.load(os.path.join(GOLD, "h5_fixtures_expected.npz"))["vec"])
