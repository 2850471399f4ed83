"""Host-side logic of the C-ABI library (no GPU needed): workload generators
and voxel-op counting agree with the oracle's definitions."""
import numpy as np

from oracle import oracle as O


def test_brush_generator_matches_oracle():
    from voxel_amd.dag import gen_brushes
    for seed, n, world, mix in ((1, 64, 512, 0), (3, 4096, 16384, 1), (17, 100, 1 << 15, 1)):
        assert np.array_equal(gen_brushes(seed, O.STREAM_BRUSH, n, world, mix),
                              O.gen_brushes(seed, O.STREAM_BRUSH, n, world, mix))


def test_voxel_op_count_matches_brute_force():
    from voxel_amd.dag import brush_voxel_ops
    br = O.gen_brushes(3, O.STREAM_BRUSH, 300, 1024, 1)
    assert brush_voxel_ops(br, 1024) == O.brush_voxels(br, 1024)
    br = O.gen_brushes(1, O.STREAM_BRUSH, 64, 512, 0)
    assert brush_voxel_ops(br, 512) == O.brush_voxels(br, 512)


def test_voxel_op_count_known_answers():
    from voxel_amd.dag import BRUSH_DTYPE, brush_voxel_ops
    b = np.zeros(3, BRUSH_DTYPE)
    b[0] = (10, 10, 10, 1, 0, 0, 0, 5)      # sphere r=1: 7 voxels
    b[1] = (0, 0, 0, 2, 3, 4, 1, 0)         # box clipped at the corner: 3*4*5
    b[2] = (100, 100, 100, 2, 0, 0, 0, 1)   # r=2: 33 voxels
    assert brush_voxel_ops(b, 64) == 7 + 60 + 0
    assert brush_voxel_ops(b, 128) == 7 + 60 + 33
