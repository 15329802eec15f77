"""Host-side checks of libtsw.so that need no GPU: the library loads, exports every function the C-ABI
header declares, and its level-symmetric ordinate sets satisfy the properties that define them."""
import math

import numpy as np
import pytest

import tsw


def test_library_exports_every_declared_symbol():
    L = tsw.lib()
    names = tsw.declared_symbols()
    expected = {"tsw_create", "tsw_destroy", "tsw_last_error", "tsw_problem_set", "tsw_set_pattern",
                "tsw_sweep_apply", "tsw_pmv", "tsw_smv", "tsw_lmv", "tsw_si_solve", "tsw_gmres_solve",
                "tsw_ordinates_level_symmetric", "tsw_sweep_timing"}
    assert expected <= set(names), expected - set(names)
    for name in names:
        assert hasattr(L, name), f"libtsw.so does not export {name}"


ORDERS = [2, 4, 6, 8, 10, 12, 16]


@pytest.mark.parametrize("order", ORDERS)
def test_level_symmetric_structure(order):
    xi, eta, w, off = tsw.level_symmetric(order)
    nq = order * (order + 2) // 2
    assert xi.size == nq and off[0] == 0 and off[4] == nq
    assert np.all(np.diff(off) == nq // 4)
    signs = [(1, 1), (1, -1), (-1, 1), (-1, -1)]  # OrdinateSet::SetSigns quadrant convention
    for k, (sx, sy) in enumerate(signs):
        sl = slice(off[k], off[k + 1])
        assert np.all(np.sign(xi[sl]) == sx) and np.all(np.sign(eta[sl]) == sy)
    assert np.all(w > 0)
    np.testing.assert_allclose(w.sum(), 4 * math.pi, rtol=1e-14)
    # directions lie inside the unit disc (mu_z > 0)
    assert np.all(xi ** 2 + eta ** 2 < 1.0)
    # level symmetry: the set of |xi| values equals the set of |eta| values
    np.testing.assert_allclose(np.sort(np.abs(xi)), np.sort(np.abs(eta)), rtol=1e-14)


@pytest.mark.parametrize("order", ORDERS)
def test_level_symmetric_moments(order):
    """Even moments of the direction cosines integrate exactly: sum w xi^2k = 4 pi/(2k+1), k < order/2,
    and odd moments vanish (the conditions the weights are built from, checked on the final set)."""
    xi, eta, w, _ = tsw.level_symmetric(order)
    mz2 = 1.0 - xi ** 2 - eta ** 2
    for k in range(order // 2):
        for c in (xi, eta):
            np.testing.assert_allclose(np.sum(w * c ** (2 * k)), 4 * math.pi / (2 * k + 1), rtol=1e-12)
        np.testing.assert_allclose(np.sum(w * mz2 ** k), 4 * math.pi / (2 * k + 1), rtol=1e-12)
    assert abs(np.sum(w * xi)) < 1e-13 and abs(np.sum(w * eta)) < 1e-13
    if order >= 6:  # mu^2 eta^2 follows from the mu^4 condition
        np.testing.assert_allclose(np.sum(w * xi ** 2 * eta ** 2), 4 * math.pi / 15, rtol=1e-12)


# Published LQ_N point weights (Lewis & Miller 1984, table 4-1; octant weights summing to 1).
PUBLISHED = {
    4: [0.3333333],
    6: [0.1761263, 0.1572071],
    8: [0.1209877, 0.0907407, 0.0925926],
    12: [0.0707626, 0.0558811, 0.0373377, 0.0502819, 0.0258513],
}


@pytest.mark.parametrize("order", sorted(PUBLISHED))
def test_level_symmetric_matches_published_weights(order):
    _, _, w, off = tsw.level_symmetric(order)
    octant = np.unique(np.round(w[off[0]:off[1]] / math.pi, 9))
    np.testing.assert_allclose(np.sort(octant), np.sort(PUBLISHED[order]), atol=2e-7)


def test_bad_order_rejected():
    for order in (0, 3, 14, 18):
        with pytest.raises(tsw.TswError):
            tsw.level_symmetric(order)
