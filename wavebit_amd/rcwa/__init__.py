"""RCWA on one MI355X: the reference's ``wavebit.rcwa`` lifecycle and declarations over a
HIP engine (raster layers; Li's or Laurent's rule, sampled or integrated coefficients; SMM;
complex FP64)."""
from .problem import Excitation, Layer, Outputs, Problem, Region, Stack
from .sources import PlaneWave
from .numerics import Numerics
from .results import DiffractionOrders, Result
from ._lifecycle import CheckReport, Execution, Prepared, check, prepare, run
from .tuning import autotune     # extension: GEMM tile-shape autotuning of a problem's solve

__all__ = ['Excitation', 'Layer', 'Outputs', 'Problem', 'Region', 'Stack', 'PlaneWave', 'Numerics',
           'DiffractionOrders', 'Result', 'CheckReport', 'Execution', 'Prepared', 'check',
           'prepare', 'run', 'autotune']
