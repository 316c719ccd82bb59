"""HDF5 files without h5py (wakis_amd.h5io): round trips, and files that the
reference's own reader and writer (h5py, when an interpreter that has it is
installed) and HDF5's h5dump agree on.

The reference writes Ez.h5 with ``hf[name] = array`` (routines.py:378-389,
437-440) and the state file with ``create_dataset`` (solverFIT3D.py:914-928),
and reads them back with ``np.array(hf[name])`` and ``hf[name][:]``
(wakeSolver.py:1344-1357, solverFIT3D.py:977-980).
"""

import glob
import os
import shutil
import subprocess

import numpy as np
import pytest

from wakis_amd import h5io


def _h5py_python():
    """An interpreter other than this one whose h5py can read our files."""
    for cand in sorted(glob.glob("/opt/conda/bin/python3.[0-9]*")) + ["/opt/conda/bin/python3"]:
        if not os.path.exists(cand) or cand.endswith("-config"):
            continue
        try:
            r = subprocess.run([cand, "-c", "import h5py, numpy"], capture_output=True,
                               timeout=120)
        except (OSError, subprocess.TimeoutExpired):
            continue
        if r.returncode == 0:
            return cand
    return None


def _binding():
    """The ctypes binding's file class, tested whether or not h5py imports."""
    try:
        h5io._lib()
    except ImportError:
        return None
    return h5io._File


File = _binding()
pytestmark = pytest.mark.skipif(File is None, reason="no HDF5 C library installed")
_PY = _h5py_python()
_H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump"
                                     if os.path.exists("/opt/conda/bin/h5dump") else None)


def _sample_items():
    rng = np.random.default_rng(5)
    return {
        "x": np.linspace(-1e-3, 1e-3, 3),
        "dx": np.float64(2.5e-4),
        "t": np.arange(0, 7 * 1.3e-12, 1.3e-12),
        "#00000": rng.standard_normal((3, 3, 11)),
        "#00001": rng.standard_normal((3, 3, 11)).astype(np.float32),
        "#00002": rng.standard_normal(11),
        "n": np.int64(-7),
        "codes": np.arange(10, dtype=np.uint8),
        "empty": np.zeros(0),
    }


def test_round_trip(tmp_path):
    path = str(tmp_path / "a.h5")
    items = _sample_items()
    with File(path, "w") as f:
        for k, v in items.items():
            f[k] = v
        f.create_dataset("E", data=np.arange(6, dtype=np.float32))
    assert h5io.is_hdf5(path)
    with File(path, "r") as f:
        assert f.keys() == sorted(list(items) + ["E"])  # name order, as h5py lists them
        assert "t" in f and "missing" not in f
        for k, v in items.items():
            d = f[k]
            assert d.shape == np.shape(v) and d.dtype == np.asarray(v).dtype, k
            assert np.array_equal(np.array(d), v), k
        assert f["dx"][()] == 2.5e-4 and np.array(f["dx"]).shape == ()
        assert np.array_equal(f["#00000"][1, 1, 2:-2], items["#00000"][1, 1, 2:-2])
        assert np.array_equal(f["E"][:], np.arange(6, dtype=np.float32))
        with pytest.raises(KeyError):
            f["missing"]
    with pytest.raises(ValueError):
        f.keys()  # closed


def test_errors(tmp_path):
    path = str(tmp_path / "b.h5")
    with File(path, "w") as f:
        f["a"] = np.ones(3)
        with pytest.raises(ValueError):
            f["a"] = np.ones(3)  # h5py refuses an existing name too
    with pytest.raises(OSError):
        File(str(tmp_path / "nope.h5"), "r")
    with pytest.raises(ValueError):
        File(path, "x+")
    with File(path, "a") as f:  # append keeps what is there
        f["b"] = 2.0
    with File(path, "r") as f:
        assert f.keys() == ["a", "b"]


@pytest.mark.skipif(_PY is None, reason="no interpreter with h5py installed")
def test_h5py_reads_our_file_and_we_read_its(tmp_path):
    ours, theirs = str(tmp_path / "ours.h5"), str(tmp_path / "theirs.h5")
    items = _sample_items()
    with File(ours, "w") as f:
        for k, v in items.items():
            f[k] = v
    np.savez(str(tmp_path / "want.npz"), **{k.replace("#", "H"): v for k, v in items.items()})
    script = f"""
import h5py, numpy as np
want = np.load({str(tmp_path / 'want.npz')!r})
with h5py.File({ours!r}, 'r') as f:
    assert sorted(f.keys()) == sorted(k.replace('H', '#', 1) if k.startswith('H') else k
                                      for k in want.files), list(f.keys())
    for k in want.files:
        name = '#' + k[1:] if k.startswith('H') else k
        a = np.array(f[name])
        assert a.dtype == want[k].dtype and a.shape == want[k].shape, (name, a.dtype, a.shape)
        assert np.array_equal(a, want[k]), name
with h5py.File({theirs!r}, 'w') as g:
    g['E'] = np.arange(12, dtype=np.float32).reshape(3, 4)
    g.create_dataset('H', data=np.linspace(0, 1, 5))
    g['dz'] = 0.5
    g['k'] = np.arange(3, dtype=np.int32)
print('ok')
"""
    r = subprocess.run([_PY, "-c", script], capture_output=True, text=True, timeout=300,
                       cwd=str(tmp_path))
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr
    with File(theirs, "r") as f:
        assert f.keys() == ["E", "H", "dz", "k"]
        assert np.array_equal(f["E"][:], np.arange(12, dtype=np.float32).reshape(3, 4))
        assert f["E"].dtype == np.float32
        assert np.array_equal(f["H"][:], np.linspace(0, 1, 5))
        assert f["dz"].shape == () and f["dz"][()] == 0.5
        assert f["k"].dtype == np.int32 and np.array_equal(f["k"][:], [0, 1, 2])


@pytest.mark.skipif(_H5DUMP is None, reason="h5dump not installed")
def test_h5dump_lists_datasets(tmp_path):
    path = str(tmp_path / "c.h5")
    with File(path, "w") as f:
        f["dx"] = 0.25
        f["#00003"] = np.zeros((3, 3, 4), dtype=np.float32)
    out = subprocess.run([_H5DUMP, "-H", path], capture_output=True, text=True,
                         timeout=60).stdout
    assert 'DATASET "#00003"' in out and "H5T_IEEE_F32LE" in out
    assert "SIMPLE { ( 3, 3, 4 ) / ( 3, 3, 4 ) }" in out
    assert 'DATASET "dx"' in out and "SCALAR" in out
