"""The C++ driver (the reference's Run / param.h / plugin roles on the C-ABI):
it builds, documents its command line, and fails loudly without a GPU."""
import os
import subprocess

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
