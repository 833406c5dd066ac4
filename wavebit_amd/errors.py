"""Structured failures at the public boundary, with the reference's classes and codes.

Mirrors reference ``src/wavebit/errors.py:1-67``: every failure carries ``code`` (a dotted
``area.item`` name), ``path`` and an ``Issue``; ``ValidationError`` is a bad input,
``NumericalError`` a valid input the numerics could not answer.
"""
from dataclasses import dataclass
from typing import Any


@dataclass(frozen=True)
class Issue:
    code: str
    path: str
    message: str
    expected: Any = None
    actual: Any = None
    severity: str = 'error'


class WavebitError(Exception):
    default_code = 'input.invalid'

    def __init__(self, message, *, code=None, path='', expected=None, actual=None):
        if code is None:
            code = type(self).default_code
        self.code = code
        self.path = path
        self.message = message
        self.expected = expected
        self.actual = actual
        self.issue = Issue(code=code, path=path, message=message, expected=expected, actual=actual)
        super().__init__(f'{path}: {message}' if path else message)


class ValidationError(WavebitError, ValueError):
    """An input was refused: malformed, out of range, or not supported here."""
    default_code = 'input.invalid'


class UnsupportedFeatureError(ValidationError):
    """A meaningful request this engine does not implement."""
    default_code = 'feature.unsupported'


class IncompatibleProblemError(ValidationError):
    """A live problem violates a prepared calculation's reuse contract."""
    default_code = 'problem.incompatible'


class NumericalError(WavebitError):
    """A valid input whose calculation produced no valid numerical output."""
    default_code = 'numerical.failure'
