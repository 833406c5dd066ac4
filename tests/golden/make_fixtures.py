"""Extract the reference's own golden data for the raster RCWA path into small fixtures.

Run once in a container that holds the reference checkout; the outputs are committed, and no
test or GPU run reads the reference at run time. Nothing is imported from the reference: its
``.npz`` files are read with ``numpy.load(allow_pickle=False)`` and its Li-nesting totals are
constants printed in its test source (``tests/references/test_reticolo_li_nesting.py:34-39``),
restated below as data.

Sources (paths relative to the reference root):
* ``tests/references/golden_reticolo.npz`` -- RETICOLO 1D per-order efficiencies; the inputs
  are the case table at ``tests/references/test_reticolo_golden.py:57-93``.
* ``tests/fixtures/legacy_rcwa/golden_legacy.npz`` -- legacy per-order efficiencies, boundary
  amplitudes, ``epz_i`` and ``q`` per layer; inputs at
  ``tests/rcwa/test_golden_legacy_parity.py:250-264`` (layers listed top-to-bottom there).
* ``tests/references/golden_fmmax_laurent.npz`` -- fmmax 1.7.1 ``Formulation.FFT`` (Laurent's
  rule) totals and per-order efficiencies at four samplings per pixel, double-precision-phase
  flavour, for the two rectangular cells and the hexagonal one; the scene table is
  ``tests/references/_fmmax_laurent_cases.py:39-71``.
* ``tests/fixtures/legacy_rcwa/golden_field.npz`` -- the reference's own in-layer field cells
  (its drift detector), downward frame, corner-unit incidence; inputs at
  ``tests/fixtures/legacy_rcwa/gen_goldens_field.py:84-108`` (layers listed top-to-bottom),
  sampling ``sample_x=5, sample_y=4, sample_z=6`` (``:117``).
* ``tests/references/golden_fmmax_field.npz`` -- fmmax 1.7.1 in-layer E/H at cell-centred
  points: the x-only grating and uniform slab (25 orders) and the crossed patch and its
  uniform control at 49 orders; scene table ``tests/references/_fmmax_field_cases.py``.
"""
import json
import sys
from pathlib import Path

import numpy as np

REF = Path(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
HERE = Path(__file__).resolve().parent

RETICOLO_CASES = ['te_normal', 'te_oblique', 'tm_normal', 'tm_oblique', 'te_quarter',
                  'tm_quarter', 'te_two_layer', 'tm_two_layer', 'te_high_order', 'tm_high_order']

LEGACY_CASES = ['te_1layer', 'tm_1layer', 'te_oblique', 'tm_oblique', 'te_2layer',
                'te_same_mat', '2d_normal', '2d_oblique', '2d_2layer', '2d_cfs']

FMMAX_LAURENT_CASES = ['rod_eps12_rect', 'hole_metal_rect', 'cross_eps6_hex']

FIELD_CASES = ['te_1d', 'tm_1d', '1d_2layer', 'conical', '2d', '2d_2layer', '2d_first_order']

# tests/references/test_reticolo_li_nesting.py:34-39 -- RETICOLO V10 total T, (TE, TM).
LI_NESTING_FIXED = {'3': [0.9822061471, 0.8913451356], '5': [0.9816211625, 0.8830613359]}


def main():
    g = np.load(REF / 'tests/references/golden_reticolo.npz', allow_pickle=False)
    out = {k: g[k] for k in g.files if k.split('__')[0] in RETICOLO_CASES}
    np.savez_compressed(HERE / 'reticolo_1d.npz', **out)

    lg = np.load(REF / 'tests/fixtures/legacy_rcwa/golden_legacy.npz', allow_pickle=False)
    keep = {}
    for k in lg.files:
        case = k.split('__')[0]
        if case not in LEGACY_CASES:
            continue
        parts = k.split('__')
        if parts[1] in ('res', 'res_te_inc', 'res_tm_inc') or parts[1] == 'grating_type':
            keep[k] = lg[k]
        elif parts[1].startswith('layer') and parts[2] in ('epz_i', 'q', 'd'):
            keep[k] = lg[k]
    np.savez_compressed(HERE / 'legacy_subset.npz', **keep)

    fl = np.load(REF / 'tests/references/golden_fmmax_laurent.npz', allow_pickle=False)
    laurent = {k: fl[k] for k in fl.files
               if k == '__levels' or (k.split('__')[0] in FMMAX_LAURENT_CASES
                                      and (k.endswith('__orders') or '__phase64__' in k))}
    np.savez_compressed(HERE / 'fmmax_laurent.npz', **laurent)

    gf = np.load(REF / 'tests/fixtures/legacy_rcwa/golden_field.npz', allow_pickle=False)
    fields = {k: gf[k] for k in gf.files
              if k.split('__')[0] in FIELD_CASES and k.split('__')[1] in ('default', 'all', 'grating_type')}
    np.savez_compressed(HERE / 'legacy_field.npz', **fields)

    ff = np.load(REF / 'tests/references/golden_fmmax_field.npz', allow_pickle=False)
    fmmax_field = {k: ff[k] for k in ff.files if '__9' not in k}
    np.savez_compressed(HERE / 'fmmax_field.npz', **fmmax_field)

    (HERE / 'li_nesting_fixed.json').write_text(json.dumps(
        {'source': 'tests/references/test_reticolo_li_nesting.py:34-39 (RETICOLO V10, nesting fixed)',
         'tolerance': 1e-9, 'total_T_te_tm': LI_NESTING_FIXED}, indent=1))
    print('wrote', len(out), 'reticolo arrays,', len(keep), 'legacy arrays,', len(laurent),
          'fmmax Laurent arrays,', len(fields), 'legacy field arrays,', len(fmmax_field),
          'fmmax field arrays')


if __name__ == '__main__':
    main()
