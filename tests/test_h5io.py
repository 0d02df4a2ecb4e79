"""The HDF5 field-file layer (libflusi_h5.so) writes the reference's layout
(hdf5_wrapper.f90:183-330 write_field_hdf5 through the Fortran API: a real
(nx, ny, nz) array, x fastest, is the C dataspace {nz, ny, nx}; file type
single or double; rank-1 double / int attributes), checked against the
image's own h5dump, and reads it back.  CPU only."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import flusi_amd

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def _write(path, dset, f, single, truncate, attrs):
    lib = flusi_amd._h5lib()
    lib.fh5_write_field.argtypes = None
    nz, ny, nx = f.shape
    names = (ctypes.c_char_p * len(attrs))(*[a[0].encode() for a in attrs])
    arrs = [np.asarray(a[1], dtype=np.float64) for a in attrs]
    vals = (ctypes.POINTER(ctypes.c_double) * len(attrs))(
        *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in arrs])
    lens = (ctypes.c_int * len(attrs))(*[len(a) for a in arrs])
    kinds = (ctypes.c_int * len(attrs))(*[a[2] for a in attrs])
    err = ctypes.create_string_buffer(512)
    f = np.ascontiguousarray(f, dtype=np.float64)
    rc = lib.fh5_write_field(os.fsencode(path), dset.encode(), nx, ny, nz, f.ctypes.data_as(ctypes.c_void_p),
                             int(single), int(truncate), len(attrs), names, vals, lens, kinds, err, 512)
    return rc, err.value.decode()


@pytest.fixture(scope="module")
def h5lib():
    if not os.path.exists(flusi_amd.H5_LIB_PATH):
        pytest.fail("libflusi_h5.so is not built (python -m flusi_amd.build)")
    return flusi_amd._h5lib()


def test_snapshot_layout_matches_reference_writer(h5lib, tmp_path):
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(7)
    f = rng.standard_normal((nz, ny, nx))
    path = str(tmp_path / "ux_000250.h5")
    rc, msg = _write(path, "ux", f, True, True,
                     [("time", [0.25], 0), ("viscosity", [1e-2], 0), ("nxyz", [nx, ny, nz], 1)])
    assert rc == 0, msg
    if not os.path.exists(H5DUMP):
        pytest.skip("no h5dump in this image")
    out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "ux"' in out
    assert "H5T_IEEE_F32LE" in out
    assert f"SIMPLE {{ ( {nz}, {ny}, {nx} ) / ( {nz}, {ny}, {nx} ) }}" in out
    attrs = subprocess.run([H5DUMP, "-a", "/ux/nxyz", path], capture_output=True, text=True, check=True).stdout
    assert "H5T_STD_I32LE" in attrs and f"{nx}, {ny}, {nz}" in attrs
    # values in file order: x fastest (h5dump prints the last index fastest)
    vals = subprocess.run([H5DUMP, "-d", "/ux", "-s", "1,2,0", "-c", "1,1,3", path], capture_output=True,
                          text=True, check=True).stdout
    data = vals[vals.index("DATA {"):]
    nums = [float(v) for v in re.findall(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?", data.split(":", 1)[1])]
    np.testing.assert_allclose(nums[:3], f[1, 2, 0:3], rtol=1e-4)  # h5dump prints ~6 digits
    back = flusi_amd.read_h5_field(path)
    np.testing.assert_array_equal(back, f.astype(np.float32).astype(np.float64))
    assert flusi_amd.read_h5_attr(path, "ux", "time", 1)[0] == 0.25
    np.testing.assert_array_equal(flusi_amd.read_h5_attr(path, "ux", "nxyz", 3), [nx, ny, nz])


def test_backup_file_holds_many_double_datasets(h5lib, tmp_path):
    nx, ny, nz = 4, 4, 8
    path = str(tmp_path / "runtime_backup0.h5")
    fields = {}
    bckp = [1.5, 1e-3, 2e-3, 1, 42, nx, ny, nz]
    for k, name in enumerate(["ux", "uy", "uz", "nlkx0"]):
        fields[name] = np.random.default_rng(k).standard_normal((nz, ny, nx))
        rc, msg = _write(path, name, fields[name], False, k == 0, [("bckp", bckp, 0)])
        assert rc == 0, msg
    for name, f in fields.items():
        np.testing.assert_array_equal(flusi_amd.read_h5_field(path, name), f)
        np.testing.assert_array_equal(flusi_amd.read_h5_attr(path, name, "bckp", 8), bckp)
    if os.path.exists(H5DUMP):
        out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
        assert out.count("H5T_IEEE_F64LE") == 4 + 4  # datasets and their bckp attributes


def test_reference_error_codes(h5lib, tmp_path):
    path = str(tmp_path / "p_000000.h5")
    f = np.zeros((2, 3, 4))
    assert _write(path, "p", f, True, True, [])[0] == 0
    rc, msg = _write(path, "p", f, True, False, [])
    assert rc == 4441 and "exists" in msg           # hdf5_wrapper.f90 write to an existing dataset
    err = ctypes.create_string_buffer(512)
    out = np.empty((2, 3, 5))
    rc = h5lib.fh5_read_field(os.fsencode(path), b"p", 5, 3, 2, out.ctypes.data, err, 512)
    assert rc == 77776                              # runtime_backuping.f90 resolution mismatch
    with pytest.raises(flusi_amd.FlusiError):
        flusi_amd.read_h5_field(str(tmp_path / "missing_000.h5"))
