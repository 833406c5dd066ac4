"""Lifecycle and admission on the CPU: declarations, error classes and codes, the C-ABI exports.

No GPU work here: these mirror the reference's API tests (tests/api/test_unified_rcwa.py,
tests/api/test_error_codes.py) for the raster path this engine implements.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import wavebit_amd
from wavebit_amd import rcwa, _native
from wavebit_amd.errors import IncompatibleProblemError, UnsupportedFeatureError, ValidationError

ROOT = Path(__file__).resolve().parents[1]


def _problem(eps=None, wavelength=0.8, pol='p', theta=0.2, phi=0.3, layers=1):
    eps = np.full((layers, 4, 4), 4.0) if eps is None else eps
    return rcwa.Problem(
        stack=rcwa.Stack(layers=[rcwa.Layer(thickness=0.2, epsilon=e) for e in eps],
                         top=wavebit_amd.Material(1.0), bottom=wavebit_amd.Material(2.25),
                         period=(1.0, 1.0)),
        excitation=rcwa.Excitation(wavelength=wavelength,
                                   sources=(rcwa.PlaneWave(polarization=pol, theta=theta, phi=phi),)))


def test_header_declares_exactly_the_library_exports():
    header = (ROOT / 'include' / 'wavebit_amd.h').read_text()
    declared = set(re.findall(r'\b(wb_[a-z_]+)\s*\(', header))
    assert declared == set(_native.EXPORTS)
    lib = _native.load_library()
    for name in _native.EXPORTS:
        assert hasattr(lib, name), name
    assert lib.wb_version() == 1


def test_workspace_queries_need_no_gpu():
    lib = _native.load_library()
    assert lib.wb_zgeev_workspace_size(100, 4) > 2 * 100 * 100 * 16 * 4
    assert lib.wb_solve_workspace_size(2, 3, 2, 2, 0) > 0
    assert lib.wb_solve_workspace_size(0, 3, 2, 2, 0) == 0
    assert lib.wb_assemble_li_workspace_size(1, 8, 8, 3, 3) > 0


def test_flags_status_reads_a_batch_on_the_host():
    """wb_flags_status needs no device: non-convergence before singularity before notices."""
    F = _native
    assert F.flags_status([0, 0, 0]) == 'WB_OK'
    assert F.flags_status([0, F.FLAG_GRAZING], notice_mask=F.FLAG_GRAZING) == 'WB_FLAGGED'
    assert F.flags_status([0, F.FLAG_GRAZING]) == 'WB_OK'
    assert F.flags_status([F.FLAG_SINGULAR, F.FLAG_EP_LOUD], notice_mask=F.FLAG_EP_LOUD) == 'WB_SINGULAR'
    assert F.flags_status([[F.FLAG_SINGULAR], [F.FLAG_NOCONVERGE]]) == 'WB_NOCONVERGE'


def test_check_reports_estimates_without_solving():
    rep = rcwa.check(_problem(), numerics=rcwa.Numerics(orders=(3, 3)), backend='torch', device='cuda')
    assert rep.ok
    assert rep.estimates['retained_spatial_modes'] == 49
    assert rep.resolved['solve_path'] == '2d'


def test_one_row_raster_with_a_polarization_word_routes_to_the_pure_1d_path():
    p = rcwa.Problem(
        stack=rcwa.Stack(layers=[rcwa.Layer(thickness=0.3, epsilon=[[1, 1, 12.25, 12.25]])],
                         bottom=wavebit_amd.Material(2.25), period=(0.7, 0.7)),
        excitation=rcwa.Excitation(wavelength=0.9, sources=(rcwa.PlaneWave(polarization='s', theta=0.0),)))
    rep = rcwa.check(p, numerics=rcwa.Numerics(orders=(10, 0)), backend='torch', device='cuda')
    assert rep.ok and rep.resolved['solve_path'] == '1d_te'


@pytest.mark.parametrize('kwargs,code', [
    (dict(backend='numpy', device='cpu'), 'rcwa.backend'),
    (dict(backend='torch', device='cpu'), 'rcwa.device'),
    (dict(backend='torch', device='cuda', precision='single'), 'rcwa.precision'),
])
def test_cpu_backends_are_refused_not_rerouted(kwargs, code):
    rep = rcwa.check(_problem(), numerics=rcwa.Numerics(orders=(1, 1)), **kwargs)
    assert not rep.ok and rep.issues[0].code == code
    with pytest.raises(UnsupportedFeatureError) as e:
        rcwa.prepare(_problem(), numerics=rcwa.Numerics(orders=(1, 1)), **kwargs)
    assert e.value.code == code


@pytest.mark.parametrize('numerics', [rcwa.Numerics(orders=(1, 1), algorithm='etm'),
                                      rcwa.Numerics(orders=(1, 1), factorization='nvf',
                                                    raster_normal_method='contour'),
                                      rcwa.Numerics(orders=(1, 1), factorization='nvf',
                                                    nvf_projector='circular'),
                                      rcwa.Numerics(orders=(1, 1), li_nesting='symmetric'),
                                      rcwa.Numerics(orders=(1, 1), formulation='first_order')])
def test_non_default_numerics_are_unsupported(numerics):
    with pytest.raises(UnsupportedFeatureError) as e:
        rcwa.prepare(_problem(), numerics=numerics, backend='torch', device='cuda')
    assert e.value.code == 'rcwa.unsupported'


def test_laurent_is_accepted_and_reported():
    numerics = rcwa.Numerics(orders=(2, 2), factorization='laurent')
    rep = rcwa.check(_problem(), numerics=numerics, backend='torch', device='cuda')
    assert rep.ok
    assert rep.resolved['factorization'] == 'laurent' and rep.resolved['li_nesting'] is None
    rep = rcwa.check(_problem(), numerics=rcwa.Numerics(orders=(2, 2)), backend='torch', device='cuda')
    assert rep.resolved['factorization'] == 'li' and rep.resolved['li_nesting'] == 'fixed'


@pytest.mark.parametrize('kw', [dict(), dict(raster_normal_method='distance'),
                                dict(nvf_projector='linear'), dict(fourier_coefficients='integrated')])
def test_nvf_is_accepted_with_its_default_construction(kw):
    numerics = rcwa.Numerics(orders=(2, 2), factorization='nvf', **kw)
    rep = rcwa.check(_problem(), numerics=numerics, backend='torch', device='cuda')
    assert rep.ok and rep.resolved['factorization'] == 'nvf' and rep.resolved['li_nesting'] is None


@pytest.mark.parametrize('kw, path', [(dict(li_nesting='fixed'), 'numerics.li_nesting'),
                                      (dict(raster_normal_method='distance', factorization='li'),
                                       'numerics.raster_normal_method'),
                                      (dict(nvf_projector='linear', factorization='laurent'),
                                       'numerics.nvf_projector')])
def test_options_without_a_reader_are_refused(kw, path):
    """Reference engine.py:5807-5850, 5887-5899: a normal-field option without NVF, or a nesting
    under NVF, is refused rather than ignored."""
    kw = dict(kw)
    numerics = rcwa.Numerics(orders=(1, 1), factorization=kw.pop('factorization', 'nvf'), **kw)
    with pytest.raises(UnsupportedFeatureError) as e:
        rcwa.prepare(_problem(), numerics=numerics, backend='torch', device='cuda')
    assert e.value.code == 'rcwa.unsupported' and e.value.path == path


def test_nvf_takes_metal_layers():
    """Reference factorization.py:1968-2160: a layer that is not a passive dielectric is solved
    with [[1/eps]]'s singular values floored (on the device); it is not refused."""
    eps = np.ones((6, 6), dtype=complex)
    eps[2:4, 2:4] = -20.0 + 1.0j
    metal = rcwa.Problem(
        stack=rcwa.Stack(layers=[rcwa.Layer(thickness=0.1, epsilon=eps)], top=wavebit_amd.Material(1.0),
                         bottom=wavebit_amd.Material(2.25), period=(0.5, 0.5)),
        excitation=rcwa.Excitation(wavelength=0.7, sources=(rcwa.PlaneWave(polarization='s'),)))
    assert rcwa.check(metal, numerics=rcwa.Numerics(orders=(1, 1), factorization='nvf'),
                      backend='torch', device='cuda').ok


def test_nvf_workspace_queries_need_no_gpu():
    lib = _native.load_library()
    assert lib.wb_assemble_nvf_workspace_size(2, 8, 8, 3, 3) > lib.wb_assemble_laurent_workspace_size(2, 8, 8, 3, 3)
    assert lib.wb_assemble_nvf_workspace_size(0, 8, 8, 3, 3) == 0
    assert lib.wb_assemble_nvf_workspace_size(1, 8, 8, 16, 3) == 0


def test_integrated_coefficients_are_accepted_and_reported():
    numerics = rcwa.Numerics(orders=(2, 2), fourier_coefficients='integrated')
    rep = rcwa.check(_problem(), numerics=numerics, backend='torch', device='cuda')
    assert rep.ok and rep.resolved['fourier_coefficients'] == 'integrated'


@pytest.mark.parametrize('nesting', ['fixed', 'symmetric'])
def test_an_explicit_nesting_is_refused_under_laurent(nesting):
    """Reference engine.py:5887-5899: Laurent's rule nests nothing, so any explicit value is
    refused rather than ignored."""
    numerics = rcwa.Numerics(orders=(1, 1), factorization='laurent', li_nesting=nesting)
    with pytest.raises(UnsupportedFeatureError) as e:
        rcwa.prepare(_problem(), numerics=numerics, backend='torch', device='cuda')
    assert e.value.code == 'rcwa.unsupported' and 'li_nesting' in str(e.value)


def test_laurent_workspace_queries_need_no_gpu():
    lib = _native.load_library()
    assert 0 < lib.wb_assemble_laurent_workspace_size(2, 8, 8, 3, 3) < lib.wb_assemble_li_workspace_size(2, 8, 8, 3, 3)
    assert lib.wb_assemble_laurent_workspace_size(0, 8, 8, 3, 3) == 0
    assert lib.wb_assemble_laurent_adjoint_workspace_size(2, 8, 8, 3, 3) > 0


def test_declaration_errors_carry_codes():
    with pytest.raises(ValidationError) as e:
        rcwa.Numerics(orders=(1, -1))
    assert e.value.code == 'numerics.orders'
    with pytest.raises(ValidationError) as e:
        rcwa.PlaneWave(polarization='x')
    assert e.value.code == 'source.polarization'
    with pytest.raises(ValidationError) as e:
        rcwa.Layer(thickness=1.0, epsilon=[[1.0]], background=wavebit_amd.Material(2.0))
    assert e.value.code == 'layer.ambiguous'
    with pytest.raises(ValidationError) as e:
        rcwa.Stack(period=(1, 1), lattice=((1, 0), (0, 1)))
    assert e.value.code == 'stack.cell_ambiguous'


def test_y_orders_on_a_one_row_raster_are_refused():
    p = _problem(eps=np.full((1, 1, 4), 3.0))
    rep = rcwa.check(p, numerics=rcwa.Numerics(orders=(2, 2)), backend='torch', device='cuda')
    assert not rep.ok and rep.issues[0].code == 'rcwa.invalid'


def test_mixed_raster_and_uniform_layers_are_refused():
    p = rcwa.Problem(
        stack=rcwa.Stack(layers=[rcwa.Layer(thickness=0.1, epsilon=[[2.0, 3.0]]),
                                 rcwa.Layer(thickness=0.1, background=wavebit_amd.Material(2.0))]),
        excitation=rcwa.Excitation(wavelength=0.5, sources=(rcwa.PlaneWave(),)))
    rep = rcwa.check(p, numerics=rcwa.Numerics(orders=(1, 0)), backend='torch', device='cuda')
    assert not rep.ok and rep.issues[0].code == 'rcwa.mixed_layer_encoding'


def test_runtime_guards_refuse_nonfinite_and_negative():
    prep_kw = dict(numerics=rcwa.Numerics(orders=(1, 1)), backend='torch', device='cuda')
    bad = _problem(eps=np.array([[[np.nan, 1.0], [1.0, 1.0]]]))
    with pytest.raises(ValidationError) as e:
        rcwa.prepare(bad, **prep_kw)
    assert e.value.code == 'prepared.finite_inputs'


def test_prepared_refuses_a_structural_change():
    prep = rcwa.prepare(_problem(), numerics=rcwa.Numerics(orders=(1, 1)), backend='torch', device='cuda')
    other = _problem(layers=2)
    with pytest.raises(IncompatibleProblemError) as e:
        prep.run(other)
    assert e.value.code == 'prepared.structure'
    rep = prep.validate(other)
    assert not rep.ok and rep.issues[0].code == 'prepared.structure'


def test_engine_fails_loudly_without_a_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip('a GPU is visible')
    prep = rcwa.prepare(_problem(), numerics=rcwa.Numerics(orders=(1, 1)), backend='torch', device='cuda')
    with pytest.raises(_native.EngineUnavailable):
        prep.run(_problem())


def test_incident_kxy_follows_the_reference_rules():
    from wavebit_amd.rcwa import _engine as eng
    kx, ky = eng.incident_kxy(2.25, 0.3, 0.4)
    assert abs(kx - 1.5 * np.sin(0.3) * np.cos(0.4)) < 1e-15
    assert abs(ky - 1.5 * np.sin(0.3) * np.sin(0.4)) < 1e-15
    with pytest.raises(ValueError):
        eng.incident_kxy(-1.0, 0.1, 0.0)
    # grazing is held just below pi/2 (single-precision bound)
    kx, _ = eng.incident_kxy(1.0, float(np.float32(np.pi / 2)), 0.0)
    assert 0.999999 < kx < 1.0


@pytest.mark.parametrize('bits,ep,error,notice', [
    (0, 'rebuild', None, False), (1, 'rebuild', 'rcwa.singular', False), (2, 'off', 'rcwa.singular', False),
    # grazing: kept (no rebuild took) warns; rebuilt is quiet; a defective cluster refuses
    (8, 'off', None, True), (8, 'rebuild', None, True), (8 | 16, 'rebuild', None, False),
    (8 | 32, 'rebuild', 'rcwa.singular', False), (8 | 32, 'off', 'rcwa.singular', False),
    # pairs: counted; below the loud sine they refuse under 'rebuild' and stand under 'off'
    (4, 'rebuild', None, False), (4, 'off', None, False),
    (4 | 64, 'rebuild', 'rcwa.singular', False), (4 | 64, 'off', None, False)])
def test_solve_flags_map_to_errors_and_notices(bits, ep, error, notice):
    """The correct-or-loud contract of the Rayleigh/EP set: failed eigensolves and factorizations,
    defective grazing clusters and (under exceptional_points='rebuild') pairs whose modes may miss
    1e-10 raise; a rebuilt grazing layer and a pair above the loud sine stand quietly; a grazing
    cluster the rebuild could not restate is kept and announced."""
    from wavebit_amd.rcwa._lifecycle import flag_error, flag_notice
    flags = np.zeros((3, 2), dtype=np.int32)
    flags[1, 1] = bits
    numerics = rcwa.Numerics(orders=(1, 1), exceptional_points=ep)
    got = flag_error(flags, numerics)
    assert (got is None) if error is None else (got.code == error)
    msg, counts = flag_notice(flags, numerics)
    assert (msg is not None) == notice
    assert counts == {'grazing_layer_solves': int(bool(bits & 8)),
                      'grazing_rebuilt_layer_solves': int(bool(bits & 16)),
                      'grazing_kept_layer_solves': int(bits & (8 | 16 | 32) == 8),
                      'defective_pair_layer_solves': int(bool(bits & 4))}


def test_detuning_is_accepted_and_validated():
    rep = rcwa.check(_problem(), numerics=rcwa.Numerics(orders=(1, 1), detuning=1e-9), backend='torch',
                     device='cuda')
    assert rep.ok
    with pytest.raises(ValidationError) as e:
        rcwa.Numerics(orders=(1, 1), detuning=-1.0)
    assert e.value.code == 'numerics.detuning'


def test_detuning_ladder_moves_wavelength_then_angle_once_each(monkeypatch):
    """The ladder on the host (reference engine.py:6382-6483), with the device solve replaced
    by a stand-in that grazes at the requested point and at the nudged wavelength alone:
    the point ends at wavelength x (1+step) and theta x (1+step) -- each knob applied once."""
    import warnings
    from dataclasses import replace
    from wavebit_amd.rcwa import _lifecycle
    from wavebit_amd.rcwa._lifecycle import _detune, _resolve, Execution
    from wavebit_amd.rcwa.diagnostics import AutoDetuneWarning
    step, theta = 1e-6, 0.2
    p = _problem(wavelength=[0.8, 0.9], theta=theta)
    r = _resolve(p, rcwa.Numerics(orders=(1, 1), detuning=step), Execution(backend='torch', device='cuda'))
    calls = []

    def fake_solve(sub, cache):
        calls.append((tuple(sub.wavelengths), sub.source.theta))
        moved_theta = sub.source.theta != theta
        flags = np.array([[0 if moved_theta else _native.FLAG_GRAZING] for _ in sub.wavelengths])
        return [('solved', float(w), sub.source.theta) for w in sub.wavelengths], flags, None, None

    class Cache:
        def sharing_conv(self):
            return self
    monkeypatch.setattr(_lifecycle, '_solve_resolved', fake_solve)
    flags = np.array([[_native.FLAG_GRAZING], [0]])
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter('always')
        r2, per, fl, _ = _detune(r, Cache(), ['first0', 'first1'], flags, None,
                                 rcwa.Numerics(orders=(1, 1), detuning=step))
    assert any(issubclass(w.category, AutoDetuneWarning) for w in caught)
    assert calls == [((0.8 * (1 + step),), theta), ((0.8 * (1 + step),), theta * (1 + step))]
    assert tuple(r2.wavelengths) == (0.8 * (1 + step), 0.9)
    assert r2.thetas == (theta * (1 + step), theta)
    assert per[1] == 'first1' and per[0] == ('solved', 0.8 * (1 + step), theta * (1 + step))
    assert int(fl[0, 0]) == 0


def test_y_orders_above_fifteen_are_refused():
    with pytest.raises(UnsupportedFeatureError) as e:
        rcwa.prepare(_problem(), numerics=rcwa.Numerics(orders=(1, 16)), backend='torch', device='cuda')
    assert e.value.code == 'rcwa.unsupported' and e.value.path == 'numerics.orders'


def test_diffraction_orders_are_read_by_order_label_and_wavelength_index():
    from wavebit_amd.rcwa.results import DiffractionOrders
    eff = np.arange(2 * 3 * 5, dtype=float).reshape(2, 3, 5)
    amp = eff * (1 + 1j)
    d = DiffractionOrders(wavelengths=np.array([0.8, 0.9]), orders_u=np.arange(-2, 3), orders_v=np.arange(-1, 2),
                          amplitude_ratio_s=amp, amplitude_ratio_p=-amp, efficiency=eff, efficiency_s=eff / 2,
                          efficiency_p=eff / 2)
    o = d.order(1, -1)
    assert o.order_u == 1 and o.order_v == -1
    np.testing.assert_array_equal(o.efficiency, eff[:, 0, 3])
    np.testing.assert_array_equal(o.amplitude_ratio_p, -amp[:, 0, 3])
    assert d.order(0).efficiency.tolist() == eff[:, 1, 2].tolist()
    with pytest.raises(KeyError):
        d.order(3, 0)
    w = d.at_wavelength_index(1)
    assert w.wavelengths.tolist() == [0.9] and w.efficiency.shape == (1, 3, 5)
    with pytest.raises(IndexError):
        d.at_wavelength_index(2)
