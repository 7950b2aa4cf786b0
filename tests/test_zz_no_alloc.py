"""The no-allocation checksum entry point (cfs_checksums_device_noalloc)
against cfs_checksums: the same bits, on two calls in a row, the second with a
smaller input (the shared device scratch keeps the first call's contents),
and inputs beyond its fixed scratch refused before anything runs."""
import ctypes

import numpy as np
import pytest

import _oracle as O

cfs = O.cfs
pytestmark = pytest.mark.gpu

_HIP = None


def _hip():
    """The HIP runtime through ctypes: device buffers for the caller's side."""
    global _HIP
    if _HIP is None:
        cfs.library()   # loads libamdhip64 with the library
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _HIP = h
    return _HIP


class DevBuf:
    def __init__(self, nbytes):
        self.p = ctypes.c_void_p()
        assert _hip().hipMalloc(ctypes.byref(self.p), max(nbytes, 8)) == 0
        self.nbytes = nbytes

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert _hip().hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0   # host to device

    def download(self, a):
        assert _hip().hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.p, a.nbytes, 2) == 0   # device to host
        return a

    def free(self):
        if self.p:
            _hip().hipFree(self.p)
            self.p = ctypes.c_void_p()


def _noalloc_sums(U_dev, cells, type_dev):
    out = DevBuf(32)
    try:
        rc = cfs.library().cfs_checksums_device_noalloc(U_dev.p, cells, type_dev.p if type_dev else None, out.p,
                                                        None)
        assert rc == cfs.CFS_OK, cfs.library().cfs_status_string(rc)
        assert _hip().hipDeviceSynchronize() == 0
        return out.download(np.zeros(4))
    finally:
        out.free()


def _case(cfg, U, ib, steps):
    """cfs_checksums of a solver after `steps` CFL steps, its state copied
    into a caller's device buffer, and the cell types likewise."""
    s = cfs.Solver(cfg)
    if ib is not None:
        s.set_immersed(ib)
    s.set_state(U)
    s.advance(steps, dt=0.0, cfl=0.5)
    ref = s.checksums()
    cells = cfg.nx * cfg.ny
    U_dev = DevBuf(32 * cells)
    s.get_state_device(U_dev.p.value)
    s.close()
    type_dev = None
    if ib is not None:
        type_dev = DevBuf(cells)
        type_dev.upload(ib["cell_type"])
    return ref, U_dev, type_dev, cells


def test_noalloc_checksums_equal_cfs_checksums():
    # call 1: the Mach-3 cylinder at 2048 x 1024 (fluid cells only), after 3 steps
    cfg = cfs.config_cylinder(2048, 1024, 4.0, 2.0, positivity=2, hll_bounds=0)
    ib = cfs.circle_setup(cfg, 1.0, 1.0, 0.125, interp=cfs.IB_BILINEAR_POSITIVE)
    ref1, U1, T1, n1 = _case(cfg, cfs.cylinder_freestream(cfg, ib["cell_type"]), ib, 3)
    got1 = _noalloc_sums(U1, n1, T1)
    print("cylinder 2048x1024: %s vs %s" % (got1, ref1))
    assert np.array_equal(got1.view(np.uint64), ref1.view(np.uint64))
    # call 2: a smaller input, no cell types (the Lax-Liu problem at 400^2)
    cfg = cfs.config_riemann(400)
    ref2, U2, T2, n2 = _case(cfg, cfs.riemann_quadrants(cfg), None, 5)
    got2 = _noalloc_sums(U2, n2, None)
    print("riemann 400^2: %s vs %s" % (got2, ref2))
    assert np.array_equal(got2.view(np.uint64), ref2.view(np.uint64))
    # an input beyond the fixed scratch is refused before anything is queued
    rc = cfs.library().cfs_checksums_device_noalloc(U2.p, cfs.NOALLOC_MAX_CELLS + 1, None, U2.p, None)
    assert rc == cfs.CFS_E_LIMIT
    assert cfs.library().cfs_checksums_device_noalloc(U2.p, 0, None, U2.p, None) == cfs.CFS_E_ARG
    for b in (U1, T1, U2):
        b.free()
