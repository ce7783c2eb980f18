"""The batched render readback (the fields scene.updateParticles reads,
js/scene.js:1595-1700) returns the same bytes as field-by-field downloads,
and only the picked fields."""
import numpy as np
import pytest

import golden
from gpu_scenes import gpu_from_golden

pytestmark = pytest.mark.gpu

FIELDS = ["px", "py", "pz", "pflag", "pT", "pLight", "pDepth", "pPh", "pWx", "pWy", "pWz"]


def test_render_readback_matches_downloads():
    g = golden.load("thermal_32")
    s = gpu_from_golden(g)
    for _ in range(3):
        s.substep(g.meta["stage_dt"])
    view = s.read_render()
    for k in FIELDS:
        want = s.download(k)
        assert np.array_equal(np.asarray(view[k]).astype(want.dtype), want), k
    part = s.read_render(("px", "pflag"))
    assert set(part) == {"px", "pflag"}
    assert np.array_equal(part["px"], s.download("px"))
    s.close()
