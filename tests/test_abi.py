"""The C-ABI library loads without a GPU and exports every function that
include/wndgpu.h declares; plan construction (host-only) matches the
reference's standard plans column for column."""
import ctypes
import os
import re

import pytest

from helpers import ROOT, orc_names

LIB = os.path.join(ROOT, 'wnd-charm_amd', 'lib', 'libwndgpu.so')


def declared():
    src = open(os.path.join(ROOT, 'include', 'wndgpu.h')).read()
    return sorted(set(re.findall(r'\b(wndgpu_[a-z0-9_]+)\s*\(', src)))


def test_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip('library not built')
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n


def test_standard_plans_match_reference_layout():
    if not os.path.exists(LIB):
        pytest.skip('library not built')
    import wndcharm
    longp = wndcharm.StdFeatureComputationPlans.getFeatureSetLong()
    shortp = wndcharm.StdFeatureComputationPlans.getFeatureSet()
    assert longp.n_features == 2919 and shortp.n_features == 1059
    assert longp.feature_names() == orc_names(True)
    assert shortp.feature_names() == orc_names(False)
    shortc = wndcharm.StdFeatureComputationPlans.getFeatureSetColor()
    longc = wndcharm.StdFeatureComputationPlans.getFeatureSetLongColor()
    assert shortc.n_features == 2199 and longc.n_features == 4059
    assert shortc.feature_names() == orc_names(2)
    assert longc.feature_names() == orc_names(3)
    # group offsets of the long set (Tasks.cpp:354-400): root group A/B/C, then node blocks
    for g, off in [('Edge Features ()', 0), ('Otsu Object Features ()', 28), ('Gabor Textures ()', 96),
                   ('Chebyshev-Fourier Coefficients (Fourier ())', 383), ('Haralick Textures (Wavelet (Edge ()))', 2639 + 136 + 48)]:
        assert longp.getFGoffset(g) == off, g


def test_custom_plan_parsing():
    if not os.path.exists(LIB):
        pytest.skip('library not built')
    import wndcharm
    p = wndcharm.FeatureComputationPlan('x')
    p.add('Zernike Coefficients (Wavelet (Edge ()))')
    p.add('  Gini Coefficient ( Fourier ( ) ) ')
    assert p.n_features == 73
    assert p.getFeatureNameByIndex(72) == 'Gini Coefficient (Fourier ()) [0]'
    bad = wndcharm.FeatureComputationPlan('y')
    bad.add('No Such Algorithm ()')
    with pytest.raises(wndcharm.WndGpuError):
        bad.finalize()


def test_cpp_facade_fails_loudly_without_gpu(tmp_path):
    """The compiled C++ caller has no host fallback: with no GPU it reports
    the engine error and exits non-zero."""
    exe = os.path.join(ROOT, 'wnd-charm_amd', 'lib', 'facade_check')
    if not os.path.exists(exe):
        pytest.skip('facade_check not built')
    if os.path.exists('/dev/kfd'):
        pytest.skip('a GPU is visible here; covered by tests/test_gpu_facade.py')
    import subprocess
    import numpy as np
    src = str(tmp_path / 'img.f64')
    np.zeros((32, 32)).tofile(src)
    r = subprocess.run([exe, src, '32', '32', str(tmp_path / 'out.f64')], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert 'no usable MI355X GPU' in r.stderr
