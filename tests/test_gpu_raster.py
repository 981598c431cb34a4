"""The HIP rasteriser against the oracle on the same packed splats: counts,
projected records, sorted pairs and ranges bit-exact; images within tolerance
of the oracle's restatement of the reference blend."""
import numpy as np
import pytest

import harness as H
from lucabrtrender_amd import raster, scenes

pytestmark = pytest.mark.gpu


def _cfg1(count=65536):
    cfg = scenes.config(1, count=count)
    s = raster.RenderSettings(width=cfg.width, height=cfg.height)
    return cfg, s


def test_config1_bit_exact_and_image():
    cfg, s = _cfg1()
    inst = [(cfg.scene, np.eye(4), raster.SplatEdit())]
    gpu = H.run_gpu(cfg.camera, s, inst)
    ora = H.run_oracle(cfg.camera, s, inst)
    print("pairs", ora["pairs"], "visible", ora["visible"])
    H.assert_bit_exact(gpu, ora)
    d = np.abs(gpu["colour"].astype(np.float64) - ora["colour"])
    print("image vs oracle: max", d.max(), "p99", np.quantile(d, 0.99))
    assert d.max() <= 1e-2 and np.quantile(d, 0.99) <= 1e-4
