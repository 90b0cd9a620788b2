"""Pins of the oracle's TPZEquationFilter restatement (SetupMatrixData and
the element loop with an active filter, StrMatrix/TPZSSpStructMatrix.cpp:
49-193, StrMatrix/pzstrmatrixor.cpp:41-71,226-233): the condensed system is
R K R^T of the unfiltered one (R selects the active equations in order),
entry for entry, its pattern is the unfiltered pattern restricted to the
active rows and columns, the load is the unfiltered load on the active
equations and 0 elsewhere, and the all-equations filter changes nothing."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from neopz_amd import problems as P


def _dense_upper(ia, ja, a, n):
    rows = np.repeat(np.arange(n), np.diff(ia))
    return sp.csr_matrix((a, (rows, ja)), shape=(n, n))


CASES = {
    "elast_p2": lambda: P.elasticity3d(n=3, p=2, seed=1, load="traction"),
    "darcy_q2": lambda: P.darcy2d(n=6, p=2, seed=2),
    "poisson_tet_p3": lambda: P.Problem(material=P.POISSON, dim=3, p=3, nx=2, ny=1, nz=1, simplex=1,
                                        force_kind=P.FORCE_CONST, force_const=(1.0, 0, 0),
                                        bc_type=[0, 1, -1, -1, -1, -1]),
}


def _filters(o, pb, rng):
    n = o.neq
    econn, pos, size = o.connects()
    out = {"random": np.sort(rng.choice(n, (2 * n) // 3, replace=False)),
           "all": np.arange(n)}
    # whole connects dropped: the equations of the first third of the blocks
    keep = np.ones(n, bool)
    for c in range(len(pos) // 3):
        keep[pos[c]:pos[c] + size[c]] = False
    out["blocks"] = np.flatnonzero(keep)
    if pb.material == P.ELAST3D:  # one component of every node (directional constraint)
        out["component"] = np.flatnonzero(np.arange(n) % 3 != 1)
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_filtered_system_is_restriction(name):
    pb = CASES[name]()
    o = O.Oracle(pb)
    ia, ja = o.pattern()
    a, rhs = o.assemble()
    n = o.neq
    U = _dense_upper(ia, ja, a, n)
    S = _dense_upper(ia, ja, np.ones_like(a), n)  # structural pattern
    for fname, act in _filters(o, pb, np.random.default_rng(3)).items():
        fia, fja, fa, frhs = o.filtered(act)
        m = len(act)
        assert len(fia) == m + 1
        Uf = _dense_upper(fia, fja, fa, m)
        R = U[act][:, act]
        assert abs(Uf - R).max() == 0.0, fname
        Sf = _dense_upper(fia, fja, np.ones_like(fa), m)
        assert abs(Sf - S[act][:, act]).max() == 0.0, fname
        # rows sorted, diagonal first
        for r in range(m):
            row = fja[fia[r]:fia[r + 1]]
            assert row[0] == r and np.all(np.diff(row) > 0)
        assert np.array_equal(frhs[act], rhs[act]), fname
        inactive = np.setdiff1d(np.arange(n), act)
        assert np.all(frhs[inactive] == 0.0)
        if fname == "all":
            assert np.array_equal(fia, ia) and np.array_equal(fja, ja) and np.array_equal(fa, a)
