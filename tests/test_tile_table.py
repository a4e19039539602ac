"""Tile table ABI on the host (no GPU): validation, the default rule, and the
environment variable that applies a saved autotuner result at load."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture
def lib():
    from curegot_backward_amd import _lib
    _lib.set_tile_table([])
    yield _lib
    _lib.set_tile_table([])


def test_default_rule_matches_tuner_helper(lib):
    from curegot_backward_amd import autotune
    for m in (1, 2, 63, 64, 65, 128, 129, 200, 256, 257, 1000, 8192, 100000):
        assert lib.tile_cols(m) == autotune.default_cols(m), m
    assert autotune.candidates(128) == [1, 2] and autotune.candidates(4096) == [1, 2, 4]
    with pytest.raises(ValueError):
        lib.tile_cols(0)


def test_table_rows_cover_bands(lib):
    lib.set_tile_table([(128, 1), (4096, 2)])
    assert lib.get_tile_table() == [(128, 1), (4096, 2)]
    assert [lib.tile_cols(m) for m in (1, 128, 129, 4096, 4097)] == [1, 1, 2, 2, 4]
    lib.set_tile_table([])
    assert lib.get_tile_table() == [] and lib.tile_cols(129) == 4


@pytest.mark.parametrize("rows", [[(128, 3)], [(256, 1), (128, 2)], [(0, 1)], [(64, 1)] * 2])
def test_bad_rows_rejected_and_table_kept(lib, rows):
    lib.set_tile_table([(512, 2)])
    with pytest.raises(ValueError):
        lib.set_tile_table(rows)
    assert lib.get_tile_table() == [(512, 2)]


def test_env_applies_saved_table(tmp_path):
    path = tmp_path / "tiles.json"
    path.write_text(json.dumps({"table": [[256, 1], [2048, 2]]}))
    code = "from curegot_backward_amd import _lib; print(_lib.get_tile_table())"
    env = dict(os.environ, CUREGOT_AMD_TILE_TABLE=str(path))
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "[(256, 1), (2048, 2)]"
