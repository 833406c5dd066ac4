"""Host logic of ``Result.field`` on the CPU: sample heights (counts, faces, snapping, range),
the transverse grid on rectangular and skewed cells, the coordinate checks and the refusals that
need no device. The heights rule is the oracle's (``oracle/rcwa_fields.heights``), which the
reference's frozen field cells pin."""
import numpy as np
import pytest

from oracle import rcwa_fields as orf
from wavebit_amd.errors import UnsupportedFeatureError, ValidationError
from wavebit_amd.rcwa import fields
from wavebit_amd.rcwa.results import pack


@pytest.mark.parametrize('z', [4, [0.0, 0.1, 0.25, 0.25 + 1e-17, 0.6, 0.25 - 1e-17, 0.1 + 0.15]])
def test_heights_follow_the_oracles_rule(z):
    d = [0.1, 0.15, 0.35]
    got = fields.heights(d, z)
    want = orf.heights(d, z)
    for g, w in zip(got, want):
        np.testing.assert_array_equal(g, w)


def test_heights_count_repeats_internal_faces_and_explicit_faces_take_the_upper_layer():
    zz, owner, edges = fields.heights([0.2, 0.3], 3)
    np.testing.assert_allclose(zz, [0, 0.1, 0.2, 0.2, 0.35, 0.5])
    assert list(owner) == [0, 0, 0, 1, 1, 1]
    zz, owner, _ = fields.heights([0.2, 0.3], [0.5, 0.2, 0.0])
    assert list(owner) == [1, 1, 0]
    with pytest.raises(ValidationError) as e:
        fields.heights([0.2, 0.3], [-1e-6])
    assert e.value.code == 'field.coordinates'


def test_xy_grid_rectangular_and_skewed():
    X, Y, xr, yr = fields.xy_grid((0.8, 0.5, 0.0, 0.0), 5, 3)
    np.testing.assert_allclose(xr, [-0.4, -0.2, 0.0, 0.2, 0.4])
    np.testing.assert_allclose(yr, [-0.25, 0.0, 0.25])
    assert X.shape == (3, 5) and np.all(X == xr[None, :]) and np.all(Y == yr[:, None])
    X, Y, _, _ = fields.xy_grid((0.8, 0.5, 0.0, 0.0), [0.1, 0.3], 2)
    np.testing.assert_allclose(X, [[0.1, 0.3], [0.1, 0.3]])
    # skewed: counts span the lattice vectors a1 = (1, 0), a2 = (0.5, 0.8)
    X, Y, xr, yr = fields.xy_grid((1.0, 0.8, 0.0, 0.5), 3, 3)
    assert xr.shape == (3, 3)
    np.testing.assert_allclose([X[0, 0], Y[0, 0]], [-0.5 - 0.25, -0.4])
    np.testing.assert_allclose([X[2, 2], Y[2, 2]], [0.5 + 0.25, 0.4])
    with pytest.raises(ValidationError):
        fields.xy_grid((1.0, 0.8, 0.0, 0.5), 3, [0.1])


@pytest.mark.parametrize('value,ok', [(3, True), (0, False), (True, False), (2.0, False),
                                      ([0.1, 0.2], True), ([], False), ([[0.1]], False),
                                      ([np.inf], False), ([1j], False)])
def test_coordinate_checks(value, ok):
    if ok:
        fields.check_coordinates('x', value)
    else:
        with pytest.raises(ValidationError) as e:
            fields.check_coordinates('x', value)
        assert e.value.code == 'field.coordinates'


def test_a_result_without_field_state_refuses_by_name():
    per = {'incident': {k: np.zeros((1, 1)) for k in ('R', 'T', 'R_s', 'R_p', 'T_s', 'T_p', 'r_s', 'r_p',
                                                      't_s', 't_p')}}
    per['incident']['power_entering'] = 0.0
    res = pack([per], 0, 0, [0.5], {}, {}, {'length_units': 'um'})
    with pytest.raises(UnsupportedFeatureError) as e:
        res.field(x=3, y=3, z=3)
    assert e.value.code == 'outputs.unavailable'
