"""Host side of batches of independent problems (pbfmm_batch_create): the
parts' octrees and plans are built and merged into one forest, and every
index of the merged plans is checked against what it indexes before any
device work.  Without a GPU the call must get through the merge and the
checks and stop at device initialisation; with one it must return the part
offsets.  Also the merge of constraint CSRs."""
import numpy as np
import pytest

import pibem_amd as P
from pibem_amd import mesh as M

PBFMM_ERR_HIP = -3


def _parts():
    return [M.config_sphere(8, 4), M.config_dcube(3, 4), M.config_spheroid(3, 4)]


def test_batch_merge_and_checks():
    probs = _parts()
    probs[2] = M.Problem(probs[2].mesh, probs[2].dirichlet, probs[2].exact, probs[2].tmp_rhs, 4)
    # the spheroid is Q2: the batch must refuse parts of different elements
    with pytest.raises(P.PbfmmError) as ei:
        P.setup_batch(probs)
    assert "differ" in str(ei.value)
    probs = probs[:2]
    try:
        b = P.setup_batch(probs)
    except P.PbfmmError as ex:
        # no device here: the merge and the index checks passed, the device
        # initialisation failed
        assert ex.code == PBFMM_ERR_HIP, ex
        assert "plan check" not in str(ex) and "forest" not in str(ex)
        return
    off = np.cumsum([0] + [p.mesh.n_dofs for p in probs])
    assert np.array_equal(b.offsets, off)


def test_merge_constraints():
    a = (np.array([0, 1, 1]), np.array([1]), np.array([1.0]), np.array([0.5, 0.0]), np.array([1, 0]))
    b = (np.array([0, 0, 2, 2]), np.array([0, 2]), np.array([0.25, 0.75]), np.array([0.0, 1.0, 0.0]),
         np.array([0, 1, 0]))
    ptr, idx, val, inh, isc = M.merge_constraints([a, b], [0, 2])
    assert ptr.tolist() == [0, 1, 1, 1, 3, 3]
    assert idx.tolist() == [1, 2, 4] and val.tolist() == [1.0, 0.25, 0.75]
    assert inh.tolist() == [0.5, 0.0, 0.0, 1.0, 0.0] and isc.tolist() == [1, 0, 0, 1, 0]
