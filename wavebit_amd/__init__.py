"""wavebit_amd: an MI355X-native RCWA engine behind wavebit's ``rcwa`` lifecycle.

``wavebit_amd.rcwa`` mirrors ``wavebit.rcwa`` (Problem / Stack / Layer / Excitation /
PlaneWave / Numerics, check / prepare / run, Result); the numerics run as hand-written HIP
kernels for gfx950 in ``libwavebit_hip.so`` (build: ``python -m wavebit_amd.build``).
"""
from .errors import (IncompatibleProblemError, NumericalError, UnsupportedFeatureError,
                     ValidationError, WavebitError)
from .materials import Material
from .units import Units
from . import rcwa

__version__ = '0.1.0'
__all__ = ['rcwa', 'Material', 'Units', 'WavebitError', 'ValidationError', 'UnsupportedFeatureError',
           'IncompatibleProblemError', 'NumericalError', '__version__']
