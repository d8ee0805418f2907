"""Tuning table: parsing and nearest-shape lookup through the C ABI (no GPU
calls), and the autotuner's table writer."""
import os

import pytest

import gmm_amd
from gmm_amd import autotune

TABLE = """# comment line
estep fp32 60 512 256 4096 30.1
estep fp32 60 2048 128 8192 250.0   # trailing comment
estep fp16x3 60 2048 128 2048 80.0 256
estep fp32 39 64 64 1024 1.0
batch fp32 60 2048 128 512 120.0 128
estep fp32 60 999 96 4096 1.0       # bad tile: skipped
estep fp64 60 512 128 4096 1.0      # bad precision: skipped
estep bf16x3 60 512 128 4096 1.0 64 # bad component tile: skipped
garbage
"""


@pytest.fixture
def table(tmp_path):
    p = tmp_path / "t.tune"
    p.write_text(TABLE)
    assert gmm_amd.tune_load(str(p)) == 5
    yield str(p)
    gmm_amd.tune_load(None)


def test_lookup_exact_and_nearest(table):
    E, B = gmm_amd.TUNE_ESTEP, gmm_amd.TUNE_BATCH
    fp32, fp16, bf16 = gmm_amd.PREC_FP32, gmm_amd.PREC_FP16X3, gmm_amd.PREC_BF16X3
    # (ll_frames, chunk_frames, comp_tile); comp_tile 128 where the line has none
    assert gmm_amd.tune_lookup(E, fp32, 60, 512) == (256, 4096, 128)
    assert gmm_amd.tune_lookup(E, fp32, 60, 2048) == (128, 8192, 128)
    # same D, nearest M on a log scale: 900 is nearer 512 than 2048
    assert gmm_amd.tune_lookup(E, fp32, 60, 900) == (256, 4096, 128)
    assert gmm_amd.tune_lookup(E, fp32, 60, 1200) == (128, 8192, 128)
    # other D: an entry with the same D is preferred over a nearer M
    assert gmm_amd.tune_lookup(E, fp32, 39, 2048) == (64, 1024, 128)
    assert gmm_amd.tune_lookup(E, fp32, 40, 64) == (64, 1024, 128)
    assert gmm_amd.tune_lookup(E, fp16, 20, 100) == (128, 2048, 256)
    assert gmm_amd.tune_lookup(B, fp32, 60, 2048) == (128, 512, 128)
    assert gmm_amd.tune_lookup(E, bf16, 60, 512) is None
    assert gmm_amd.tune_lookup(B, fp16, 60, 512) is None


def test_missing_file_empties_table(tmp_path):
    try:
        assert gmm_amd.tune_load(str(tmp_path / "none.tune")) == -1
        assert gmm_amd.tune_lookup(gmm_amd.TUNE_ESTEP, gmm_amd.PREC_FP32, 60, 512) is None
    finally:
        gmm_amd.tune_load(None)


def test_autotune_table_roundtrip(tmp_path):
    p = str(tmp_path / "sub" / "x.tune")
    rows = {("estep", "fp32", 60, 512): (128, 8192, 35.2, 128),
            ("estep", "fp16x3", 60, 2048): (128, 4096, 12.5, 256),
            ("batch", "fp32", 60, 2048): (256, 1024, 120.5, 128)}
    autotune.write_table(p, rows)
    assert autotune.read_table(p) == rows
    try:
        assert gmm_amd.tune_load(p) == 3
        assert gmm_amd.tune_lookup(gmm_amd.TUNE_ESTEP, gmm_amd.PREC_FP16X3, 60, 2048) == (
            128, 4096, 256)
    finally:
        gmm_amd.tune_load(None)


def test_committed_table_parses():
    path = autotune.DEFAULT_TABLE
    if not os.path.exists(path):
        pytest.skip("no committed tuning table")
    rows = autotune.read_table(path)
    assert gmm_amd.tune_load(path) == len(rows) > 0
    for (kind, prec, D, M), (ll, cf, ms, ct) in rows.items():
        assert kind in ("estep", "batch") and prec in autotune.PREC
        assert ll in (64, 128, 256) and cf >= 64 and ms > 0
        assert ct in autotune.comp_tiles(autotune.PREC[prec], D)


def test_autotune_frame_tile_candidates():
    """FP32: 64/128/256 (256 only up to depth 128); mixed: 128/256 (256 only
    where 2D+1 rounded up to 32 is a multiple of 64)."""
    fp32, fp16 = autotune.PREC["fp32"], autotune.PREC["fp16x3"]
    assert autotune.frame_tiles(fp32, 60) == [64, 128, 256]
    assert autotune.frame_tiles(fp32, 64) == [64, 128]
    assert autotune.frame_tiles(fp16, 60) == [128, 256]   # depth 128
    assert autotune.frame_tiles(fp16, 39) == [128]        # depth 96
    assert autotune.frame_tiles(fp16, 20) == [128, 256]   # depth 64


def test_autotune_comp_tile_candidates():
    """Mixed forms: 128 / 256 components per statistics block where the
    split feature rows hold 64 or 128 columns (contraction depths 64, 96 and
    128; at 256 columns eight waves would spill); other depths and FP32: 128
    only."""
    fp32, fp16 = autotune.PREC["fp32"], autotune.PREC["fp16x3"]
    assert autotune.comp_tiles(fp32, 60) == [128]
    assert autotune.comp_tiles(fp16, 60) == [128, 256]   # depth 128
    assert autotune.comp_tiles(fp16, 20) == [128, 256]   # depth 64
    assert autotune.comp_tiles(fp16, 39) == [128, 256]   # depth 96, rows of 128
    assert autotune.comp_tiles(fp16, 70) == [128]        # depth 160, rows of 256
    assert autotune.comp_tiles(fp16, 10) == [128]        # depth 32
    assert autotune.comp_tiles(fp16, 127) == [128]       # depth 256
