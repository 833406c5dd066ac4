"""The QR iteration's rotation generator (``wavebit_amd/csrc/wb_givens.h``), host-compiled: on
edge cases (|f| down to 1e-200 of |g|, subnormal |f|^2 at the common scale, huge and tiny pairs)
every rotation is finite, unitary to round-off and maps [f; g] to [r; 0] to round-off."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_givens_edge_cases(tmp_path):
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not Path(hipcc).exists():
        pytest.skip('hipcc is not available')
    exe = tmp_path / 'givens'
    subprocess.run([hipcc, '-O2', '-x', 'hip', '--offload-arch=gfx950', str(ROOT / 'tests' / 'givens_host.cpp'),
                    '-o', str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split('\n')
    rows = [line.split() for line in out if line.strip()]
    assert len(rows) == 8
    for unit, res, finite in rows:
        assert finite == '1'
        assert float(unit) <= 4.5e-16 and float(res) <= 4.5e-16
