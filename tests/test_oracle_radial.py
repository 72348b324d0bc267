"""The oracle's add_radial_force against its FP64 closed form (both signs,
clipped at the walls)."""
import numpy as np
import pytest

from radial_closed_form import radial_force


@pytest.mark.parametrize("args", [(0.2, 0.3, 0.1, 50.0), (0.3, 0.2, 0.08, -40.0), (0.01, 0.5, 0.2, 5.0)])
def test_radial_force_closed_form(args):
    from oracle import pyoracle as po
    o = po.Oracle(128, 128, 1.0 / 128, po.bench_props(), xy=np.zeros((0, 2), np.float32))
    o.add_radial_force(*args)
    fu, fv = radial_force(128, 128, 1.0 / 128, *args)
    repel = args[3] > 0
    gu = o.read("POST_FORCE_U" if repel else "EXT_FORCE_U")
    gv = o.read("POST_FORCE_V" if repel else "EXT_FORCE_V")
    assert not o.read("EXT_FORCE_U" if repel else "POST_FORCE_U").any()
    tol = 1e-5 * abs(args[3])
    assert np.max(np.abs(gu - fu)) <= tol and np.max(np.abs(gv - fv)) <= tol
    assert np.count_nonzero(gu) == np.count_nonzero(fu) and np.count_nonzero(gv) == np.count_nonzero(fv)
