"""The GEMM tile-shape table of the autotuner (host side; no device work)."""
import pytest

from wavebit_amd import _native


def test_configurations_and_table_round_trip():
    names = _native.gemm_tuning_configs()
    assert names[0] == '64x64' and len(names) == 4
    _native.gemm_tuning_clear()
    try:
        _native.gemm_tuning_set(0, 0, 722, 722, 722, 256, 3)
        _native.gemm_tuning_set(1, 0, 64, 2, 64, 8, 1)
        text = _native.gemm_tuning_table()
        assert text.splitlines() == ['0 0 722 722 722 256 3', '1 0 64 2 64 8 1']
        _native.gemm_tuning_clear()
        assert _native.gemm_tuning_table() == ''
        assert _native.gemm_tuning_load(text) == 2
        assert _native.gemm_tuning_table() == text
        _native.gemm_tuning_set(1, 0, 64, 2, 64, 8, -1)          # remove one entry
        assert _native.gemm_tuning_table().splitlines() == ['0 0 722 722 722 256 3']
        # malformed lines stop the parse; out-of-range configurations are skipped
        assert _native.gemm_tuning_load('0 0 1 2 3 4 9\nnot a line\n0 0 5 5 5 1 0\n') == 0
    finally:
        _native.gemm_tuning_clear()


@pytest.mark.parametrize('args', [(2, 0, 1, 1, 1, 1, 0), (0, 0, -1, 1, 1, 1, 0), (0, 0, 1, 1, 1, 1, 4)])
def test_bad_table_entries_are_refused(args):
    with pytest.raises(ValueError):
        _native.gemm_tuning_set(*args)
