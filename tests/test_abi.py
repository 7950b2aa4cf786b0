"""The C-ABI library loads without a GPU, exports every entry point that
include/cfs.h declares, and rejects bad arguments before touching HIP."""
import ctypes
import re

import numpy as np
import pytest

import _oracle as O

cfs = O.cfs


def declared_functions():
    src = open(cfs.HEADER_PATH).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cfs_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = cfs.library()
    names = declared_functions()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_layout():
    lib = cfs.library()
    assert lib.cfs_abi_version() == int(re.search(r"#define CFS_ABI_VERSION (\d+)", open(cfs.HEADER_PATH).read()).group(1))
    lib.cfs_config_size.restype = ctypes.c_int32
    assert lib.cfs_config_size() == ctypes.sizeof(cfs.Config)
    for st in range(-8, 1):
        assert b"unknown" not in lib.cfs_status_string(st)
    assert cfs.device_count() >= 0


def test_create_rejects_bad_configs_without_a_gpu():
    lib = cfs.library()
    h = ctypes.c_void_p()
    bad = [
        cfs.make_config(2, 10, 1.0, 1.0),                                  # too small
        cfs.make_config(10, 10, 0.0, 1.0),                                 # zero length
        cfs.make_config(10, 10, 1.0, 1.0, rk=4),                           # no such RK
        cfs.make_config(10, 10, 1.0, 1.0, bc=(2, 0, 0, 0)),                # periodic on one side only
        cfs.make_config(10, 10, 1.0, 1.0, precision=7),
        cfs.make_config(10, 10, 1.0, 1.0, viscous=2),
        cfs.make_config(10, 10, 1.0, 1.0, viscous=1, mu=-1e-3),             # negative viscosity
        cfs.make_config(10, 10, 1.0, 1.0, viscous=1, mu=1e-3, prandtl=0.0),
        cfs.make_config(10, 10, 1.0, 1.0, viscous=1, mu=1e-3, ma_ref=0.0),
    ]
    for c in bad:
        assert lib.cfs_create(ctypes.byref(c), None, ctypes.byref(h)) == cfs.CFS_E_ARG
        assert not h.value


def test_circle_setup_argument_errors():
    cfg = cfs.config_cylinder(64, 32, 4.0, 2.0)
    with pytest.raises(cfs.CfsError):
        cfs.circle_setup(cfg, 1.0, 1.0, -0.1)
    with pytest.raises(cfs.CfsError):
        cfs.circle_setup(cfg, 1.0, 1.0, 0.25, interp=5)
    # capacity too small reports the count and an error
    lib = cfs.library()
    n = ctypes.c_int32(0)
    ct = np.zeros(64 * 32, np.uint8)
    one_i = np.zeros(4, np.int32)
    one_d = np.zeros(10)
    rc = lib.cfs_circle_setup(ctypes.byref(cfg), 1.0, 1.0, 0.25, 400, 3, 1, 1, ct.ctypes.data, ctypes.byref(n),
                              one_i.ctypes.data, one_i.ctypes.data, one_d.ctypes.data, one_d.ctypes.data,
                              one_i.ctypes.data)
    assert rc == cfs.CFS_E_ARG and n.value > 1


def test_batch_footprint_and_bad_batch_configs_without_a_gpu():
    lib = cfs.library()
    cfg = cfs.make_config(100, 60, 1.0, 1.0, rk=3)
    # 3 FP64 plane-sets, rounded up to 256 B
    assert cfs.Batch.footprint(cfg) == (96 * 100 * 60 + 255) // 256 * 256
    with_ib = cfs.Batch.footprint(cfg, n_ig=50)
    assert with_ib > cfs.Batch.footprint(cfg) + 5 * 100 * 60
    assert lib.cfs_batch_footprint(ctypes.byref(cfs.make_config(2, 2, 1.0, 1.0)), 0) == -1
    h = ctypes.c_void_p()
    for bc in [cfs.BatchConfig(4, 0, 8, 0, 1 << 20), cfs.BatchConfig(3, 5, 8, 0, 1 << 20),
               cfs.BatchConfig(3, 0, 0, 0, 1 << 20), cfs.BatchConfig(3, 0, 8, -1, 1 << 20),
               cfs.BatchConfig(3, 0, 8, 0, 0)]:
        assert lib.cfs_batch_create(ctypes.byref(bc), None, ctypes.byref(h)) == cfs.CFS_E_ARG
        assert not h.value


def test_riemann_batch_problems_are_seeded():
    a = cfs.riemann_batch_problems(7, count=5)
    b = cfs.riemann_batch_problems(7, count=5)
    for (ca, ua), (cb, ub) in zip(a, b):
        assert ca.nx == cb.nx and np.array_equal(ua, ub)
    sizes = [c.nx for c, _ in cfs.riemann_batch_problems(1, count=256)]
    assert min(sizes) >= 64 and max(sizes) <= 512 and all(n % 32 == 0 for n in sizes)
    assert len(set(sizes)) > 10


def test_tile_shapes_without_a_gpu():
    lib = cfs.library()
    shapes = [cfs.tile_shape_info(k) for k in range(cfs.NUM_TILE_SHAPES)]
    assert shapes[0] == (15, 16, 256)
    for ti, tj, nt in shapes:
        # one cell per thread, the faces in two rounds of the workgroup
        assert ti * tj <= nt and (ti + 1) * tj + ti * (tj + 1) <= 2 * nt and nt % 64 == 0
    assert "#define CFS_NUM_TILE_SHAPES %d" % cfs.NUM_TILE_SHAPES in open(cfs.HEADER_PATH).read()
    assert lib.cfs_tile_shape_info(-1, None, None, None) == cfs.CFS_E_ARG
    assert lib.cfs_set_tile_shape(None, 0) == cfs.CFS_E_ARG
    assert lib.cfs_autotune(None, 5, None, None) == cfs.CFS_E_ARG
    h = ctypes.c_void_p()
    bc = cfs.BatchConfig(3, cfs.PREC_FP64, 4, 0, 1 << 20, 0, cfs.NUM_TILE_SHAPES)
    assert lib.cfs_batch_create(ctypes.byref(bc), None, ctypes.byref(h)) == cfs.CFS_E_ARG and not h.value
    assert lib.cfs_batch_autotune(ctypes.byref(bc), None, 0, 5, None, None) == cfs.CFS_E_ARG


def test_noalloc_checksums_refuse_bad_sizes_without_a_gpu():
    # both refusals come before any HIP call; the pointers are never read
    lib = cfs.library()
    fake = ctypes.c_void_p(256)
    assert lib.cfs_checksums_device_noalloc(fake, cfs.NOALLOC_MAX_CELLS + 1, None, fake, None) == cfs.CFS_E_LIMIT
    assert lib.cfs_checksums_device_noalloc(fake, 0, None, fake, None) == cfs.CFS_E_ARG
    assert lib.cfs_checksums_device_noalloc(None, 16, None, fake, None) == cfs.CFS_E_ARG
    assert b"CFS_E_LIMIT" in lib.cfs_status_string(cfs.CFS_E_LIMIT)
