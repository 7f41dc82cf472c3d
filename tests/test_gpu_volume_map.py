"""GPU volume-map / SDF generator against the CPU oracle's FP64 restatement of
volume_map.cpp:9-54 (same float node positions and quadrature points)."""
import numpy as np
import pytest

import harness as H
import vkfluidsim_amd as V

pytestmark = pytest.mark.gpu

CASES = [
    (V.VFS_SHAPE_BOX_WALLS, (0.8, 0.4, 0.4, 0.8, 0.4, 0.4), (24, 14, 14), 0.04),
    (V.VFS_SHAPE_BOX_WALLS, (1.0, 0.5, 0.5, 1.0, 0.5, 0.5), (40, 22, 22), 0.03),
    (V.VFS_SHAPE_SPHERE, (1.0, 0.5, 0.5, 0.2), (20, 20, 20), 0.012),
    (V.VFS_SHAPE_BOX, (0.0, 0.0, 0.0, 0.1, 0.05, 0.08), (16, 12, 18), 0.04),
]


@pytest.mark.parametrize("kind,shape,res,h", CASES)
def test_volume_map_matches_oracle(kind, shape, res, h):
    g = V.volume_map(kind, shape, res, h)
    o = H.oracle_volume_map(kind, shape, res, h)
    assert g.box_size == o.box_size and g.box_pos == o.box_pos
    np.testing.assert_array_equal(g.sdf, o.sdf)
    # FP64 sums over 4096 quadrature points; the device pow/sqrt may differ in
    # the last bit of a double, which can move the FP32 result by one ulp
    np.testing.assert_allclose(g.vm, o.vm, rtol=2e-7, atol=1e-12)
    assert np.count_nonzero(g.vm) > 0
