"""Grazing layers on the GPU: the cluster rebuild (wb_solve_rebuild) against the oracle's
restatement of the reference's rebuild, and the correct-or-loud contract of the reference's own
graze stacks (tests/rcwa/test_rayleigh_guard.py:221-256).

Scenes as tests/test_oracle_graze.py: a 2-row pillar (12.25 in air) over a uniform spacer,
period 1000 x 1000 nm, thicknesses 300 (spacer) and 200 (pillar), air above a 2.25 substrate,
normal s incidence, orders (5, 1).
"""
import warnings

import numpy as np
import pytest
import torch

import wavebit_amd
from wavebit_amd import _native, rcwa
from wavebit_amd.errors import NumericalError
from wavebit_amd.rcwa import _engine as eng
from oracle import rcwa_oracle as orc

pytestmark = pytest.mark.gpu
RUN = dict(backend='torch', device='cuda:0')
DEV = torch.device('cuda:0')
PERIOD = (1000.0, 1000.0)
PILLAR_GRAZES = (616.5157327418264, 793.192493734249, 538.6449010842978)


def _stack(n_spacer):
    pillar = np.ones((1, 2, 20))
    pillar[0, :, 6:14] = 12.25
    return np.concatenate([np.full((1, 2, 20), n_spacer ** 2), pillar], axis=0).astype(np.complex128)


def _problem(stack, wls):
    layers = [rcwa.Layer(thickness=t, epsilon=e) for t, e in zip((300.0, 200.0), stack)]
    return rcwa.Problem(
        stack=rcwa.Stack(layers=layers, top=wavebit_amd.Material(1.0),
                         bottom=wavebit_amd.Material(2.25), period=PERIOD),
        excitation=rcwa.Excitation(wavelength=list(wls),
                                   sources=(rcwa.PlaneWave(polarization='s', theta=0.0, phi=0.0),)),
        units=wavebit_amd.Units(length='nm'))


def _oracle(stack, wl):
    with np.errstate(divide='ignore', invalid='ignore'):
        return orc.solve_raster(stack, [300.0, 200.0], 1.0, 2.25, PERIOD, wl, 0.0, 0.0, 's', (5, 1))


def _run(stack, wls, **numerics):
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter('always')
        res = rcwa.run(_problem(stack, wls), numerics=rcwa.Numerics(orders=(5, 1), **numerics), **RUN)
    return res, [r for r in rec if issubclass(r.category, UserWarning)]


@pytest.mark.parametrize('n_spacer', [2.0, 1.5])
def test_reference_graze_stacks_are_correct_or_loud(n_spacer):
    """The exact graze either refuses (NumericalError) or matches the 1e-8-detuned solve within
    1e-4 with R + T - 1 within 1e-4: never quietly wrong."""
    stack = _stack(n_spacer)
    far = _oracle(stack, 500.0 * (1 + 1e-8))['incident']
    try:
        res, _ = _run(stack, [500.0])
    except NumericalError as exc:
        # refused as a defective cluster ('rcwa.singular'), or, where the two products disagree
        # on the cluster and its modes are kept, by the non-finite check ('rcwa.nonfinite'): the
        # reference's loud outcomes are FloatingPointError and LinAlgError alike
        assert exc.code in ('rcwa.singular', 'rcwa.nonfinite')
        return
    R, T = res.R[0].sum(), res.T[0].sum()
    assert abs(R - far['R'].sum()) < 1e-4 and abs(R + T - 1.0) < 1e-4


@pytest.mark.parametrize('wl', PILLAR_GRAZES)
def test_pillar_graze_is_rebuilt_like_the_oracle(wl):
    """At an exact graze of the pillar layer (min|w| / max|w| ~ 1e-16) both the device and the
    oracle rebuild the cluster; per-order R and T agree within 1e-10, energy within 1e-10, and
    the run says nothing."""
    stack = _stack(1.5)
    res, said = _run(stack, [wl])
    ep = res.diagnostics['exceptional_points']
    assert ep['grazing_rebuilt_layer_solves'] == 1 and ep['grazing_kept_layer_solves'] == 0, ep
    assert not said, [str(r.message) for r in said]
    ref = _oracle(stack, wl)
    assert ref['layer_status'][1] == 'rebuilt'
    assert np.abs(res.R[0] - ref['incident']['R']).max() < 1e-10
    assert np.abs(res.T[0] - ref['incident']['T']).max() < 1e-10
    assert abs(res.R[0].sum() + res.T[0].sum() - 1.0) < 1e-10


@pytest.mark.parametrize('n_spacer', [2.0, 1.5])
def test_one_ulp_off_the_reference_graze_is_rebuilt_like_the_oracle(n_spacer):
    """One ULP above 500 nm the spacer still grazes; its cluster is representable and rebuilt on
    both sides. Its coupled rates (~1e-8) are set by round-off and carry an O(1) relative error
    on either side, which reaches R and T at the 1e-10 level: per-order agreement within 1e-9,
    the 1e-8-detuned solve within the reference's 1e-4."""
    stack = _stack(n_spacer)
    wl = float(np.nextafter(500.0, 501.0))
    res, said = _run(stack, [wl])
    assert res.diagnostics['exceptional_points']['grazing_rebuilt_layer_solves'] == 1
    assert not said
    ref = _oracle(stack, wl)['incident']
    assert np.abs(res.R[0] - ref['R']).max() < 1e-9
    assert np.abs(res.T[0] - ref['T']).max() < 1e-9
    far = _oracle(stack, 500.0 * (1 + 1e-8))['incident']
    assert abs(res.R[0].sum() - far['R'].sum()) < 1e-4


def test_native_rebuild_flags_only_the_grazing_pairs():
    """Through the C ABI: wb_solve flags the grazing layer, wb_solve_rebuild marks it rebuilt
    (pair flags cleared) and leaves a point without a graze bit for bit as wb_solve solved it."""
    stack = _stack(1.5)
    hu, hv = 5, 1
    conv = [_native.assemble_li(torch.from_numpy(stack).to(DEV), hu, hv)[:3]]
    b = eng.Bucket(hu=hu, hv=hv, path=eng.PATH_2D, n_layers=2)
    for wl in (PILLAR_GRAZES[0], 700.0):
        b.points.append(eng.Point(stack=0, wavelength=wl, theta=0.0, phi=0.0, polarization='s',
                                  eps_top=1.0, eps_bottom=2.25, period=PERIOD,
                                  thickness=(300.0, 200.0)))
    inputs = eng.bucket_inputs(b, conv, DEV)
    eff, amp, pe, flags = (t.cpu().numpy() for t in _native.solve(*inputs, hu, hv, _native.MODE_2D))
    assert flags[0, 1] & _native.FLAG_GRAZING and not flags[1].any()
    eff2, amp2, pe2, flags2 = (t.cpu().numpy() for t in _native.solve_rebuild(*inputs, hu, hv))
    assert flags2[0, 1] & _native.FLAG_REBUILT
    assert not flags2[0, 1] & (_native.FLAG_EP | _native.FLAG_EP_LOUD | _native.FLAG_DEFECTIVE)
    assert flags2[0, 0] == 0 and not flags2[1].any()
    assert np.array_equal(eff2[1], eff[1]) and np.array_equal(amp2[1], amp[1])
    # the host helper replaces only the flagged point
    e3, a3, p3, f3 = eff.copy(), amp.copy(), pe.copy(), flags.copy()
    assert eng.rebuild_grazing(b, conv, DEV, e3, a3, p3, f3) == 1
    assert np.array_equal(e3, eff2) and np.array_equal(f3, flags2)


def test_detuning_ladder_moves_a_refused_graze_off_the_condition():
    """With Numerics(detuning=step) a graze the device refuses is solved one rung away; the
    answer then matches the 1e-8-detuned solve within 1e-4 either way."""
    stack = _stack(2.0)
    far = _oracle(stack, 500.0 * (1 + 1e-8))['incident']
    res, said = _run(stack, [500.0, 650.0], detuning=1e-9)
    assert abs(res.R[0].sum() - far['R'].sum()) < 1e-4
    assert abs(res.R[0].sum() + res.T[0].sum() - 1.0) < 1e-9
    moved = res.diagnostics['effective_wavelengths'][0] != 500.0
    assert moved == any('detuning' in str(r.message) for r in said)
    ref = _oracle(stack, 650.0)['incident']
    assert np.abs(res.R[1] - ref['R']).max() < 1e-10


def test_scattering_output_at_a_graze_is_rebuilt():
    """Outputs(scattering=True) at a pillar graze: the rebuilt solve's specular columns are the
    reported amplitudes, and the efficiencies equal the plain rebuilt run's (within rounding: the
    plain solve's reflection-matrix cascade orders the same algebra differently)."""
    stack = _stack(1.5)
    wl = PILLAR_GRAZES[0]
    plain, _ = _run(stack, [wl])
    base = _problem(stack, [wl])
    prob = rcwa.Problem(stack=base.stack, excitation=base.excitation, units=base.units,
                        outputs=rcwa.Outputs(scattering=True))
    res = rcwa.run(prob, numerics=rcwa.Numerics(orders=(5, 1)), **RUN)
    assert res.diagnostics['exceptional_points']['grazing_rebuilt_layer_solves'] == 1
    # at the graze itself the two orderings part at ~3e-11: the parity bar, not the bit
    assert np.abs(res.R - plain.R).max() < 1e-10 and np.abs(res.T - plain.T).max() < 1e-10
    S = res.scattering[0]
    N = 33
    s_inc = res.diagnostics['incidence_variants']['s_inc']
    assert np.abs(S.S11[:N, N // 2] - s_inc['r_s'][0].reshape(-1)).max() < 1e-12
    assert np.abs(S.S21[:N, N // 2] - s_inc['t_s'][0].reshape(-1)).max() < 1e-12
