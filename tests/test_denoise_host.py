"""Denoiser host logic without a GPU: the model table and loader errors
(src/oidn.ts:22-29, oidnNetwork.ts:181-215) and the archive table reader
against the oracle's."""
import os

import numpy as np
import pytest

from mutracer_amd import denoise
from oracle import unet


def test_model_table():
    assert denoise.MODEL_FILENAMES == {"small": "rt_hdr_alb_nrm_small.tza", "standard": "rt_hdr_alb_nrm.tza",
                                       "cleanAuxLarge": "rt_hdr_calb_cnrm_large.tza"}
    for m in denoise.BUNDLED_MODELS:
        assert os.path.isfile(os.path.join(denoise.WEIGHTS_DIR, denoise.MODEL_FILENAMES[m]))
    assert [n for n, _ in denoise.UNET] == [n for n, _, _ in unet.UNET]
    assert [(n, lv) for n, lv in denoise.UNET_LARGE] == [(n, lv) for n, lv, _ in unet.UNET_LARGE]


def test_loader_refuses_missing_weights():
    with pytest.raises(FileNotFoundError, match="not bundled"):
        denoise.Denoiser.load("cleanAuxLarge")
    with pytest.raises(ValueError, match="unknown model"):
        denoise.Denoiser.load("huge")


@pytest.mark.parametrize("path", [denoise.WEIGHTS, os.path.join(denoise.WEIGHTS_DIR, "rt_hdr_alb_nrm_small.tza")])
def test_archive_table(path):
    data = open(path, "rb").read()
    assert denoise.tza_shapes(data) == {k: a.shape for k, (a, _) in unet.read_tza(data).items()}
    synth = unet.synthetic_weights(unet.network_shapes(unet.WIDTHS["large"], 3, True))
    assert denoise.tza_shapes(unet.write_tza(synth)) == {k: a.shape for k, (a, _) in synth.items()}
    assert np.all([len(s) in (1, 4) for s in denoise.tza_shapes(data).values()])
