"""GEMM tile-shape autotuning for a problem's solve (an extension: the reference has no
counterpart, its BLAS picks tiles internally).

``autotune`` runs the problem once with shape recording on, then times every tile
configuration of the batched complex GEMM on each recorded shape and keeps the fastest. The
choice never changes a result (every configuration accumulates in the same order); later runs
of problems with the same shapes use it. The table lives in the process and can be saved and
restored as text (``wavebit_amd._native.gemm_tuning_table`` / ``gemm_tuning_load``).
"""
from .. import _native


def autotune(problem, *, numerics, backend='torch', device='cuda:0', precision='double', reps=3):
    """Tune the GEMM tile shapes of ``problem``'s solve; returns the number of shapes tuned."""
    import torch
    from ._lifecycle import prepare
    prepared = prepare(problem, numerics=numerics, backend=backend, device=device, precision=precision)
    dev = torch.device(device)
    _native.gemm_tuning_record(True)
    try:
        prepared.run(problem)
    finally:
        _native.gemm_tuning_record(False)
    return _native.gemm_autotune(dev, reps=reps)
