"""The C++ host API (wnd-charm_amd/cpp/wndcharm.hpp: ImageMatrix,
FeatureComputationPlanExecutor::run, FeatureAlgorithm::execute,
ImageTransform::execute, FeatureComputationPlan::add/finalize) driven by a
compiled caller (lib/facade_check) against the oracle on the same image."""
import os
import subprocess

import numpy as np
import pytest

from wndcharm import synthetic
from helpers import ROOT, ALG_NAMES, EXACT, exact_mask_for_names, orc_algorithm, orc_features, orc_names, orc_transform
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, 'wnd-charm_amd', 'lib', 'facade_check')


def test_cpp_facade_matches_oracle(tmp_path):
    img = synthetic.image(31, 10, (96, 80)).astype(np.float64)
    src, dst = str(tmp_path / 'img.f64'), str(tmp_path / 'out.f64')
    img.tofile(src)
    print('facade_check 96x80', flush=True)
    r = subprocess.run([BIN, src, '96', '80', dst], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(dst)
    long_want = orc_features(img, True)
    zi = ALG_NAMES.index('Zernike Coefficients')
    zern_want = orc_algorithm(zi, img)
    four_want = orc_transform(0, img)
    short_want = orc_features(img, False)
    sizes = [len(long_want), len(zern_want), four_want.size, len(short_want)]
    assert out.size == sum(sizes)
    parts = np.split(out, np.cumsum(sizes)[:-1])
    compare(parts[0], long_want, exact_mask_for_names(orc_names(True)), 'long set via executor')
    zmask = np.zeros(len(zern_want), bool)
    zmask[EXACT[zi]] = True
    compare(parts[1], zern_want, zmask, 'Zernike via FeatureAlgorithm')
    assert np.array_equal(parts[2].reshape(four_want.shape), four_want)
    compare(parts[3], short_want, exact_mask_for_names(orc_names(False)), 'short set built group by group')
