"""The C++ driver (the reference's Run / param.h / plugin roles on the C-ABI):
it builds, documents its command line, and fails loudly without a GPU."""
import os
import subprocess

import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
RUN = os.path.join(os.path.dirname(cfs.LIB_PATH), "cfs_run")


def test_driver_is_built_and_prints_usage():
    assert os.path.exists(RUN), "build the package first (__graft_entry__.build())"
    r = subprocess.run([RUN, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--case" in r.stderr
    r = subprocess.run([RUN, "--no-such-flag"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


@pytest.mark.skipif(cfs.device_count() > 0, reason="checks the no-GPU failure path")
def test_driver_fails_loudly_without_a_gpu():
    r = subprocess.run([RUN, "--case", "lax-liu", "--nx", "40", "--ny", "40", "--stages", "4"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "CFS_E_HIP" in r.stderr


def test_hdf5_writer_round_trip(tmp_path):
    # the driver's run-time-loaded HDF5 writer: datasets, shapes, type and
    # values read back through the library's own reader
    import shutil
    import _h5
    if _h5.lib() is None or shutil.which("g++") is None:
        pytest.skip("no HDF5 C library or no g++ in this environment")
    here = os.path.dirname(os.path.abspath(__file__))
    exe = str(tmp_path / "h5w")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(os.path.dirname(here), cfs.__name__, "driver"),
                    "-o", exe, os.path.join(here, "native", "h5_write_check.cpp"), "-ldl"], check=True)
    path = str(tmp_path / "snap.h5")
    r = subprocess.run([exe, path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _h5.read(path, ["A", "B", "C", "D"])
    i, j = np.meshgrid(np.arange(3), np.arange(5), indexing="ij")
    A = i + j / 1000.0
    assert got["D"] is None
    assert got["A"].shape == (3, 5) and np.array_equal(got["A"], A)
    assert np.array_equal(got["B"], -A) and np.array_equal(got["C"], (i * j).astype(float))
