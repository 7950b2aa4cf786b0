"""The figures bench.py reads from the committed rocprofv3 summaries under
profiles/: every one is present and consistent with the summary's own
arithmetic, so a summary made again in another layout cannot silently turn a
roofline figure into null.  CPU only (no kernels run)."""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_headline_counters(bench):
    traffic = bench.profile_figure("traffic_bytes_per_launch")
    lane_ops = bench.profile_figure("valu_lane_ops_per_cell_stage")
    assert traffic is not None and lane_ops is not None
    # 97 B per cell-stage on 2048 x 1024 cells; the measured traffic is within
    # a few percent of it (the tile order keeps halo re-reads in L2)
    alg = bench.BYTES_PER_CELL_STAGE * bench.NX * bench.NY
    assert 0.9 < traffic / alg < 1.5
    assert 2000 < lane_ops < 6000


def test_headline_counters_of_the_dual_face_shape(bench):
    # the autotuner may pick the dual-face shape for the headline workload:
    # its kernel has its own counters and its own name on the bench line
    traffic = bench.profile_figure("traffic_bytes_per_launch", bench.DUAL_FACE_SHAPE)
    lane_ops = bench.profile_figure("valu_lane_ops_per_cell_stage", bench.DUAL_FACE_SHAPE)
    assert traffic is not None and lane_ops is not None
    alg = bench.BYTES_PER_CELL_STAGE * bench.NX * bench.NY
    assert 0.9 < traffic / alg < 1.5
    assert 2000 < lane_ops < 6000
    assert lane_ops != bench.profile_figure("valu_lane_ops_per_cell_stage")
    assert bench.profile_figure("valu_lane_ops_per_cell_stage", 1) is None


def test_kernel_names_distinguish_the_tile_shapes(bench):
    names = {bench.kernel_name(bench.cfs.PREC_FP64, True, k) for k in range(bench.cfs.NUM_TILE_SHAPES)}
    assert len(names) == bench.cfs.NUM_TILE_SHAPES
    assert bench.kernel_name(bench.cfs.PREC_FP64, True, bench.DUAL_FACE_SHAPE).endswith("256, true>>")


def test_batch_profile_rows(bench):
    ratio = bench.profile_row(bench.BATCH_PROFILE, "traffic / algorithmic")
    lane_ops = bench.profile_row(bench.BATCH_PROFILE, "VALU lane-ops per cell-stage")
    cell_stages = bench.profile_row(bench.BATCH_PROFILE, "cell-stages")
    alg = bench.profile_row(bench.BATCH_PROFILE, "algorithmic bytes")
    assert None not in (ratio, lane_ops, cell_stages, alg)
    assert alg == pytest.approx(bench.BATCH_BYTES_PER_CELL_STAGE * cell_stages, rel=1e-3)
    assert 0.9 < ratio < 3.0
    assert 1000 < lane_ops < 4000


def test_shear_layer_profile_rows(bench):
    traffic = bench.profile_row(bench.KH_PROFILE, "traffic (sum)")
    alg = bench.profile_row(bench.KH_PROFILE, "algorithmic")
    lane_ops = bench.profile_row(bench.KH_PROFILE, "SQ_INSTS_VALU", col=3)
    lane_ops_fp32 = bench.profile_row(bench.KH_PROFILE, "SQ_INSTS_VALU", nth=1, col=3)
    assert None not in (traffic, alg, lane_ops, lane_ops_fp32)
    assert alg == pytest.approx(bench.BYTES_PER_CELL_STAGE_RK3 * 4096 * 4096, rel=1e-3)
    assert 0.9 < traffic / alg < 1.5
    # the FP64 row comes first: FP32-bulk needs fewer instructions
    assert 2000 < lane_ops < 6000 and lane_ops_fp32 < lane_ops


def test_laxliu_profile_rows(bench):
    lane_ops = bench.profile_row(bench.RIEMANN_PROFILE, "SQ_INSTS_VALU", col=3)
    fma = bench.profile_row(bench.RIEMANN_PROFILE, "SQ_INSTS_VALU_FMA_F64", col=3)
    assert None not in (lane_ops, fma)
    assert 2000 < lane_ops < 6000 and fma < lane_ops


def test_missing_rows_read_as_none(bench):
    assert bench.profile_row(bench.BATCH_PROFILE, "no such row") is None
    assert bench.profile_row("no_such_profile.md", "traffic") is None
