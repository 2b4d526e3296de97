"""GPU: what the trained denoiser gains over the method without it.

Held-out scenes (seeds never used in training) are traced at 4 spp and at
64 spp; against the 64 spp frame, the shipped network's output must have a
lower relative MSE (mean of (x - ref)^2 / (ref^2 + 0.01)) than the raw
4 spp frame and than fixed 5x5 filters (box, and a joint bilateral filter
guided by the G-buffer), i.e. the network is what brings the gain.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_trained_network_beats_no_network_and_fixed_filters(gpu_ctx):
    from vulkanhybridraytracer_amd import denoise
    assert os.path.exists(denoise.WEIGHTS_PATH), "trained weights are shipped with the package"
    dn = denoise.Denoiser.load(gpu_ctx)
    print("rendering held-out pairs", flush=True)
    samples = denoise.make_dataset(gpu_ctx, [9100, 9101], views=2)
    ev = denoise.evaluate(dn, samples)
    print("rMSE vs 64 spp:", ev, flush=True)
    assert ev["network"] < 0.5 * ev["noisy"]
    assert ev["network"] < ev["bilateral"]
    assert ev["network"] < ev["box"]
    # the untrained network (same architecture, initial weights) does not
    # have the gain: it is the training that brings it
    dn0 = denoise.Denoiser(gpu_ctx, seed=0)
    ev0 = denoise.evaluate(dn0, samples, with_filters=False)
    print("untrained:", ev0, flush=True)
    assert ev["network"] < 0.5 * ev0["network"]
    dn.close()
    dn0.close()
