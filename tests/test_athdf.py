"""The native .athdf reader (libmk_athdf.so) against files written by h5py.

The reference loads snapshots with h5py (athenak.py:79-101).  The image's
main interpreter has no h5py, so the test writes AthenaK-shaped files with
the image's other interpreter that has it (/opt/conda/bin/python3.9), in the
dtypes AthenaK stores (float32 fields and coordinates, int32 locations and
levels), once with fixed-length and once with variable-length VariableNames.
"""
import os
import subprocess

import numpy as np
import pytest

from mahakala_amd import _athdf
from tests import helpers

PY39 = "/opt/conda/bin/python3.9"

_WRITER = r"""
import sys, numpy as np, h5py
src, dst, vlen = sys.argv[1], sys.argv[2], sys.argv[3] == "1"
d = np.load(src)
with h5py.File(dst, "w") as f:
    for k in ("x1v", "x2v", "x3v", "x1f", "x2f", "x3f", "uov", "B", "LogicalLocations", "Levels"):
        f.create_dataset(k, data=d[k])
    names = [n.encode() for n in d["names"]]
    if vlen:
        f.attrs.create("VariableNames", names, dtype=h5py.special_dtype(vlen=bytes))
    else:
        f.attrs["VariableNames"] = np.array(names, dtype="S")
"""


def _have_writer():
    if not os.path.exists(PY39):
        return False
    return subprocess.run([PY39, "-c", "import h5py"], capture_output=True).returncode == 0


pytestmark = pytest.mark.skipif(not (os.path.exists(_athdf.LIB_PATH) and _have_writer()),
                                reason="needs libmk_athdf.so and h5py under /opt/conda/bin/python3.9")


def _datasets():
    ds = helpers.small_two_level(n=4, seed=5)
    out = {k: np.asarray(ds[k], dtype=np.float32) for k in ("x1v", "x2v", "x3v", "x1f", "x2f", "x3f", "uov", "B")}
    out["LogicalLocations"] = np.asarray(ds["LogicalLocations"], dtype=np.int32)
    out["Levels"] = np.asarray(ds["Levels"], dtype=np.int32)
    out["names"] = np.array(list(ds["VariableNames"]))
    return out


def _write(tmp_path, vlen):
    d = _datasets()
    src = tmp_path / "src.npz"
    np.savez(src, **d)
    dst = tmp_path / ("vlen.athdf" if vlen else "fixed.athdf")
    subprocess.run([PY39, "-c", _WRITER, str(src), str(dst), "1" if vlen else "0"], check=True, timeout=120)
    return d, dst


@pytest.mark.parametrize("vlen", [False, True])
def test_native_reader_returns_the_stored_arrays(tmp_path, vlen):
    d, path = _write(tmp_path, vlen)
    got = _athdf.read_athdf_native(str(path))
    for k in _athdf.DATASETS:
        assert got[k].dtype == d[k].dtype, k
        np.testing.assert_array_equal(got[k], d[k])
    assert got["VariableNames"] == [str(v) for v in d["names"]]


def test_read_athdf_uses_the_native_reader_without_h5py(tmp_path):
    from mahakala_amd.grmhd import athenak
    d, path = _write(tmp_path, False)
    got = athenak.read_athdf(str(path))
    np.testing.assert_array_equal(got["uov"], d["uov"])
    assert got["VariableNames"][0] == "dens"


def test_reader_errors(tmp_path):
    with pytest.raises(_athdf.AthdfError, match="cannot open"):
        _athdf.read_athdf_native(str(tmp_path / "missing.athdf"))
    bad = tmp_path / "bad.athdf"
    bad.write_bytes(b"not an hdf5 file")
    with pytest.raises(_athdf.AthdfError):
        _athdf.read_athdf_native(str(bad))
    d, path = _write(tmp_path, False)
    L = _athdf.lib()
    import ctypes as C
    h = C.c_void_p()
    assert L.mk_athdf_open(str(path).encode(), C.byref(h)) == 0
    t, nd = C.c_int32(), C.c_int32()
    dims = (C.c_int64 * 8)()
    assert L.mk_athdf_info(h, b"no_such_dataset", C.byref(t), C.byref(nd), dims) != 0
    small = C.create_string_buffer(4)
    n = C.c_int32()
    assert L.mk_athdf_varnames(h, small, 4, C.byref(n)) == -1
    L.mk_athdf_close(h)


@pytest.mark.gpu
def test_model_from_file_equals_model_from_datasets(tmp_path):
    """AthenakFluidModel(filename) (athenak.py:48-53) builds the same resident
    snapshot as from_datasets on the same arrays."""
    from mahakala_amd.grmhd.athenak import AthenakFluidModel
    d, path = _write(tmp_path, False)
    m_file = AthenakFluidModel(str(path), 0.5, fluid_gamma=13. / 9)
    ds = {k: d[k] for k in _athdf.DATASETS}
    ds["VariableNames"] = [str(v) for v in d["names"]]
    m_mem = AthenakFluidModel.from_datasets(ds, 0.5, fluid_gamma=13. / 9)
    np.testing.assert_array_equal(m_file.all_meshblocks, m_mem.all_meshblocks)
    assert list(m_file.variable_names) == list(m_mem.variable_names)
