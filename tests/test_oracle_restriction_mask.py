"""The oracle's ElementRestriction::BooleanMask (fem/restriction.cpp:255-290)
checked by what fixes it: every entry is 0 or 1, the E->L sum of the mask
(MultTranspose, restriction.cpp:159-193) is 1 at every L-DOF, so each L-DOF
is counted once; and the one marked entry of each L-DOF is, by brute force
over the E-vector, the E-index that comes first in that DOF's row of the
transpose table."""
import numpy as np
import pytest


@pytest.mark.parametrize("n,p", [(1, 1), (2, 3), (3, 2)])
def test_boolean_mask(orc, n, p):
    P = orc.Problem(n, p, perturb=0.1, seed=1)
    nE = P.ne * P.D ** 3
    m = orc.boolean_mask(P.offsets, P.indices, nE)
    assert set(np.unique(m)) <= {0.0, 1.0}
    assert int(m.sum()) == P.ndofs
    ones = orc.restriction_mult_transpose(P.offsets, P.indices, P.D ** 3, m, 0)
    assert np.array_equal(ones, np.ones(P.ndofs))
    # brute force: for each E-entry, its row (the L-DOF it gathers from) and
    # its position within the row; marked iff it has the smallest position
    pos = np.empty(nE, np.int64)
    pos[P.indices] = np.arange(nE)
    first = {}
    for ie in range(nE):
        row = P.gather_map[ie]
        if row not in first or pos[ie] < pos[first[row]]:
            first[row] = ie
    want = np.zeros(nE)
    want[list(first.values())] = 1.0
    assert np.array_equal(m, want)
