"""Sky-model catalog reader (montblanc_amd/sky_model.py) against the
reference's own known answers (montblanc/tests/test_utils.py:164-203) and
its example catalog (montblanc/examples/sky_model.txt:1-6), both restated
here as data."""

import numpy as np
import pytest

from montblanc_amd import sky_model, util

# test_utils.py:167-173
KAT_FILE = ('# format npsrc: l, m, I, Q, U, V\n'
            '11, 12, 13, 14, 15, 16\n'
            '21, 22, 23, 24, 25, 26\n'
            '# format ngsrc: l, m, I, Q, U, V, el, em, eR\n'
            '31, 32, 33, 34, 35, 36, 37, 38, 39\n'
            '41, 42, 43, 44, 45, 46, 47, 48, 49\n')

# examples/sky_model.txt:1-6
EXAMPLE_FILE = ('# format npsrc: l, m, I, Q, U, V, alpha\n'
                '10e-4, 20e-4, 0.8, 0.3, 0.2, 0.4, -0.8\n'
                '30e-4, 10e-4, 0.9, 0.33, 0.21, 0.42, -0.8\n'
                '# format ngsrc: l, m, I, Q, U, V, alpha, el, em, eR\n'
                '10e-4, 20e-4, 0.8, 0.3, 0.2, 0.4, -0.8, 0.1, 0.2, 0.9\n'
                '30e-4, 10e-4, 0.9, 0.33, 0.21, 0.42, -0.8, 0.1, 0.2, 0.9\n')


def write(tmp_path, text, name='sky.txt'):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_reference_known_answers(tmp_path):
    """test_utils.py:175-201: columns, counts and shape_arrays."""
    res = util.parse_sky_model(write(tmp_path, KAT_FILE))
    A, S = res.arrays, res.src_counts
    assert S == {'npsrc': 2, 'ngsrc': 2}
    assert A['l'] == ['11', '21', '31', '41']
    assert A['m'] == ['12', '22', '32', '42']
    assert A['I'] == ['13', '23', '33', '43']
    assert A['Q'] == ['14', '24', '34', '44']
    assert A['U'] == ['15', '25', '35', '45']
    assert A['V'] == ['16', '26', '36', '46']
    assert A['el'] == ['37', '47']
    assert A['em'] == ['38', '48']
    assert A['eR'] == ['39', '49']
    shape = (2, S['npsrc'] + S['ngsrc'])
    ary = res.shape_arrays(['l', 'm'], shape, np.float32)
    assert ary.shape == shape and ary.dtype == np.float32
    np.testing.assert_array_equal(ary[0], ((np.arange(4) + 1) * 10 + 1).astype(np.float32))
    np.testing.assert_array_equal(ary[1], ((np.arange(4) + 1) * 10 + 2).astype(np.float32))


def test_example_catalog_to_source_arrays(tmp_path):
    res = sky_model.parse_sky_model(write(tmp_path, EXAMPLE_FILE))
    a = sky_model.source_arrays(res, ntime=3)
    assert a['counts'] == (2, 2, 0)
    np.testing.assert_array_equal(a['lm'], [[10e-4, 20e-4], [30e-4, 10e-4], [10e-4, 20e-4], [30e-4, 10e-4]])
    iquv = np.array([[0.8, 0.3, 0.2, 0.4], [0.9, 0.33, 0.21, 0.42]] * 2)
    assert a['stokes'].shape == (4, 3, 4)
    for t in range(3):   # brightness repeated over time
        np.testing.assert_array_equal(a['stokes'][:, t], iquv)
    np.testing.assert_array_equal(a['alpha'], np.full((4, 3), -0.8))
    np.testing.assert_array_equal(a['gauss_shape'], [[0.1, 0.1], [0.2, 0.2], [0.9, 0.9]])
    assert a['sersic_shape'].shape == (3, 0)


def test_sources_ordered_by_type_and_defaults(tmp_path):
    """Sections in any order give point, gaussian, sersic (file order within
    a type); Q, U, V and alpha default to 0; blank and comment lines are
    skipped."""
    text = ('# a catalog\n'
            '# format nssrc: l, m, I, e1, e2, eS\n'
            '1, 2, 3, 0.1, 0.2, 0.3\n'
            '\n'
            '# format ngsrc: l, m, I, el, em, eR\n'
            '4, 5, 6, 0.4, 0.5, 0.6\n'
            '# format npsrc: l, m, I, alpha\n'
            '7, 8, 9, -0.7\n'
            '10, 11, 12, -0.5\n')
    a = sky_model.source_arrays(sky_model.parse_sky_model(write(tmp_path, text)), ntime=2, dtype=np.float32)
    assert a['counts'] == (2, 1, 1)
    assert a['lm'].dtype == np.float32
    np.testing.assert_array_equal(a['lm'], [[7, 8], [10, 11], [4, 5], [1, 2]])
    np.testing.assert_array_equal(a['stokes'][:, 0], [[9, 0, 0, 0], [12, 0, 0, 0], [6, 0, 0, 0], [3, 0, 0, 0]])
    np.testing.assert_allclose(a['alpha'][:, 1], [-0.7, -0.5, 0, 0])
    np.testing.assert_allclose(a['gauss_shape'][:, 0], [0.4, 0.5, 0.6])
    np.testing.assert_allclose(a['sersic_shape'][:, 0], [0.1, 0.2, 0.3])


def test_incomplete_sources_are_errors(tmp_path):
    no_shape = '# format ngsrc: l, m, I\n1, 2, 3\n'
    with pytest.raises(ValueError, match='el'):
        sky_model.source_arrays(sky_model.parse_sky_model(write(tmp_path, no_shape)), ntime=1)
    bad_type = '# format nxsrc: l, m, I\n1, 2, 3\n'
    with pytest.raises(ValueError, match='nxsrc'):
        sky_model.source_arrays(sky_model.parse_sky_model(write(tmp_path, bad_type, 'b.txt')), ntime=1)
