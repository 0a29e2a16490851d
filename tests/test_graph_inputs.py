"""Captured renders and training steps read the Gaussian tensors in place at
every replay, so they reject tensors they would have to copy (a copy would
not see the caller's later in-place updates). The checks run before any
device work, so this runs without a GPU."""
import numpy as np
import pytest
import torch

from fastergs_amd import scenes
from fastergs_amd.graph import CapturedRenderer, CapturedTrainingStep
from tests.helpers import settings_for


def _inputs():
    sc = scenes.uniform_scene(64, seed=1)
    return {k: torch.from_numpy(np.ascontiguousarray(v)) for k, v in sc.items()}


@pytest.mark.parametrize("bad", ["means3D", "scales", "dc", "shs"])
@pytest.mark.parametrize("how", ["strided", "float64"])
def test_captured_classes_reject_copies(bad, how):
    cam = scenes.look_at_camera((0.0, 0.0, -4.0), 64, 48)
    st = settings_for(cam, device="cpu")
    d = _inputs()
    t = d[bad]
    if how == "strided":
        d[bad] = torch.cat([t, t], dim=-1)[..., ::2]  # same values, not contiguous
        assert not d[bad].is_contiguous()
    else:
        d[bad] = t.double()
    with pytest.raises(ValueError, match=bad):
        CapturedRenderer(st, d["means3D"], d["opacities"], d["scales"], d["rotations"],
                         dc=d["dc"], shs=d["shs"])
    with pytest.raises(ValueError, match=bad):
        CapturedTrainingStep([st], d["means3D"], d["opacities"], d["scales"], d["rotations"],
                             dc=d["dc"], shs=d["shs"])
