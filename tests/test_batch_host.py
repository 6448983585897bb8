"""Host-side batch assembly: the concatenated mesh keeps every instance's groups and boundaries."""
import numpy as np

from nonlocal_amd import batch, configs


def test_concatenate_offsets():
    ps = [configs.thermal_q4(n, 0.5, 0.2, 1.0) for n in (3, 5, 4)]
    mesh, models, mats, bcs, nb = batch.concatenate(ps)
    assert nb.tolist() == [0, 16, 52, 77]
    assert mesh.n_nodes == 77 and mesh.n_elems == 9 + 25 + 16
    assert mesh.group_begin.tolist() == [0, 9, 34, 50]
    assert len(models) == len(mats) == 3 and len(bcs) == 12
    assert mesh.bgroup_begin.tolist() == [0, 3, 6, 9, 12, 17, 22, 27, 32, 36, 40, 44, 48]
    # element node ids of instance 1 are offset by 16 and stay within its node range
    e1 = mesh.elems[9:34]
    assert e1.min() == 16 and e1.max() == 51
    np.testing.assert_array_equal(mesh.elems[:9], ps[0].mesh.elems)
