"""Pins the CPU oracle against the reference's own golden feature vectors:
010067_301x300-l_precalculated.sig (tests/pywndcharm_tests/test_FeatureVector.py:86-102),
t1_s01_c05_ij-l_precalculated.sig, the lymphoma tile .sig
(test_FeatureVector.py:152-217) and unbalanced.fit (tests/wndchrm_tests/test04-sigs.pl:16-31),
compared with the reference's comparator (wndcharm/utils.py:276-366).

The goldens were written by an older f14 (maximal correlation coefficient)
implementation than the reference's current CVIPtexture.cpp:993-1023, so the
f14 columns ("Haralick Textures (...) [16]/[17]") are excluded here and pinned
instead against the reference's own CVIPtexture.cpp + Eigen, compiled from
source (tests/test_oracle_ref.py)."""
import numpy as np
import pytest

from helpers import load_npz, orc_features, orc_names, wndcharm_compare


def f14_column(name):
    return name.startswith('Haralick Textures') and (name.endswith('[16]') or name.endswith('[17]'))


def check_vector(golden_names, golden_vals, mine, names):
    col = {n: i for i, n in enumerate(names)}
    bad = []
    for n, v in zip(golden_names, golden_vals):
        if f14_column(n):
            continue
        m = mine[col[n]]
        if not wndcharm_compare(m, v):
            bad.append((n, v, m))
    return bad


@pytest.mark.parametrize('case', ['img_t1_s01_c05_ij', 'img_010067_301x300'])
def test_long_set_sig(case):
    d = load_npz(case + '.npz')
    mine = orc_features(d['pixels'].astype(np.float64), long_set=True)
    bad = check_vector(list(d['names']), d['values'], mine, orc_names(True))
    assert not bad, bad[:10]


def test_short_set_lymphoma_tile():
    d = load_npz('img_lymphoma_tile.npz')
    assert len(d['values']) == 1059
    mine = orc_features(d['pixels'].astype(np.float64), long_set=False)
    names = orc_names(False)
    assert sorted(names) == sorted(d['names'])
    bad = check_vector(list(d['names']), d['values'], mine, names)
    assert not bad, bad[:10]


def test_plan_layout_matches_sig_names():
    d = load_npz('img_010067_301x300.npz')
    # the .sig lists the long set in plan-column order (Tasks.cpp:354-400)
    assert list(d['names']) == orc_names(True)


@pytest.mark.parametrize('idx', [0, 3, 5, 7])
def test_unbalanced_fit(idx):
    d = load_npz('unbalanced_fit.npz')
    img = d['pixels'][idx].astype(np.float64)
    mine = orc_features(img, long_set=True)
    bad = check_vector(list(d['names']), d['values'][idx], mine, orc_names(True))
    assert not bad, bad[:10]
