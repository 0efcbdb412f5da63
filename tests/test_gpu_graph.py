"""Device graph construction vs the oracle (bit-exact CSR)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
SEED = 27491095


def same_csr(dev, host, weights=False):
    off, nb, w = dev.read_csr(weights=weights)
    assert np.array_equal(off, host.offsets)
    assert np.array_equal(nb, host.nbrs)
    if weights:
        assert np.array_equal(w, host.weights)


@pytest.mark.parametrize("scale,sym", [(6, True), (10, True), (10, False), (16, True)])
def test_rmat_csr_bit_exact(gi, scale, sym):
    s, d = O.rmat_edges(scale, 16, SEED)
    host = O.build_csr(1 << scale, s, d, sym)
    same_csr(gi.Graph.rmat(scale, 16, SEED, sym), host)


def test_uniform_and_grid_bit_exact(gi):
    s, d = O.uniform_edges(5000, 20000, 77)
    same_csr(gi.Graph.uniform(5000, 20000, 77, True), O.build_csr(5000, s, d, True))
    s, d = O.grid_edges(64, 40, 5)
    same_csr(gi.Graph.grid(64, 40, 5), O.build_csr(64 * 64, s, d, True))


def test_from_edges_squish(gi):
    e = np.array([[0, 0], [0, 1], [0, 1], [1, 0], [2, 2], [3, 1]], np.int32)
    for sym in (True, False):
        host = O.build_csr(4, e[:, 0], e[:, 1], sym)
        same_csr(gi.Graph.from_edges(4, e[:, 0], e[:, 1], sym), host)


def test_from_csr_roundtrip_and_validation(gi):
    s, d = O.rmat_edges(8, 8, 3)
    host = O.fill_weights(O.build_csr(256, s, d, True), 4, 1, 255)
    g = gi.Graph.from_csr(host.offsets, host.nbrs, host.weights)
    same_csr(g, host, weights=True)
    with pytest.raises(gi.GIError):
        gi.Graph.from_csr(np.array([0, 1], np.int32), np.array([5], np.int32))


def test_weights_bit_exact(gi):
    g = gi.Graph.rmat(10, 8, SEED, True).set_weights(11, 1, 255)
    s, d = O.rmat_edges(10, 8, SEED)
    host = O.fill_weights(O.build_csr(1 << 10, s, d, True), 11, 1, 255)
    same_csr(g, host, weights=True)


def test_transpose_bit_exact(gi):
    g = gi.Graph.rmat(10, 16, 9, False).set_weights(2, 1, 100)
    s, d = O.rmat_edges(10, 16, 9)
    host = O.fill_weights(O.build_csr(1 << 10, s, d, False), 2, 1, 100)
    t = g.transpose()
    th = O.transpose(host)
    same_csr(t, th, weights=True)
    sym = gi.Graph.rmat(8, 4, 9, True)
    assert sym.transpose().h == sym.h  # symmetric CSR is its own transpose


def test_set_weights_on_views_fails_and_base_stays_usable(gi):
    # a transpose or blocked view belongs to its base graph: new weights go on
    # the base, the view refuses them, and the base keeps working
    g = gi.Graph.rmat(10, 8, 3, False).set_weights(1, 1, 50)
    t = g.transpose()
    b = g.blocked(256)
    for view in (t, b):
        with pytest.raises(gi.GIError):
            view.set_weights(2, 1, 50)
    off, nb, w = g.read_csr(weights=True)
    assert len(nb) == g.num_edges and w.min() >= 1
    # new weights on the base drop its views: the old handles are refused
    g.set_weights(3, 1, 50)
    with pytest.raises(gi.GIError):
        t.set_weights(4, 1, 50)
    t2 = g.transpose()
    toff, tnb, tw = t2.read_csr(weights=True)
    host = O.transpose(O.CSR(*g.read_csr(weights=True)))
    assert np.array_equal(toff, host.offsets) and np.array_equal(tnb, host.nbrs)
    # a view outlives a base the caller no longer holds
    v = gi.Graph.rmat(9, 8, 5, False).transpose()
    assert v.num_vertices == 512 and v.read_csr()[0][-1] == v.num_edges
