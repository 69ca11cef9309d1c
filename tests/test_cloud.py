"""Point cloud, ghost band, tags and Morton order (host set-up): product == oracle bit for bit, and both against
brute force / the reference's unit-test expectations."""
import numpy as np
import pytest

from hcc_hip import _lib as L
from hcc_hip import workloads as W
from oracle import oracle as O


CASES = [
    W.domain(2, (0, 0), (1, 1), (True, True), 1 / 64, 2.5, 0.25, 0),
    W.domain(2, (0, 0), (1, 1), (False, False), 1 / 100, 2.5, 0.25, 0),
    W.domain(2, (0, 0), (1, 1), (False, False), 1 / 50, 2.5, 0.25, 11),
    W.domain(2, (0, 0), (1, 1), (True, False), 1 / 40, 2.5, 0.1, 2),
    W.domain(3, (0, 0, 0), (1, 1, 1), (True, True, True), 1 / 16, 2.2, 0.25, 0),
    W.domain(3, (0, 0, 0), (1, 1, 1), (False, True, False), 1 / 12, 2.2, 0.2, 5),
    W.domain(1, (-1,), (1,), (False,), 2 / 40, 2.5, 0.0, 0),
    W.domain(2, (-1, -1), (1, 1), (False, False), 2 / 30, 2.5, 0.0, 0),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_cloud_product_equals_oracle(k):
    dom = CASES[k]
    a = L.cloud_generate(W.domain_desc(dom))
    b = O.cloud_generate(dom)
    assert a["N"] == b["N"]
    assert np.array_equal(a["pos"], b["pos"])
    assert np.array_equal(a["tags"], b["tags"]) and np.array_equal(a["is_boundary"], b["is_boundary"])
    assert np.array_equal(a["canvas"], b["canvas"])
    assert a["morton_ties"] == 0


def test_periodic_cloud_counts_and_range():
    """unit_core.jl:320-328: a periodic box has exactly n^D particles, none on the boundary, all in [0, 1)."""
    for D in (1, 2):
        dom = W.domain(D, (0,) * D, (1,) * D, (True,) * D, 0.05, 2.5, 0.2, 3)
        c = L.cloud_generate(W.domain_desc(dom))
        assert c["N"] == 20 ** D and not c["is_boundary"].any()
        assert np.all((c["pos"] >= 0) & (c["pos"] < 1))


def test_ghost_band_is_brute_force_selection():
    """unit_core.jl:343-355: the narrow ghost band equals the brute-force selection within (R + h)."""
    dom = W.domain(2, (0, 0), (1, 1), (False, False), 1 / 20, 2.5, 0.0, 0)
    c = L.cloud_generate(W.domain_desc(dom))
    h = 1 / 20
    canvas = c["canvas"]
    n = [int(round((canvas[2 + d] - canvas[d]) / h)) + 1 for d in range(2)]
    xs = [canvas[d] + h * np.arange(n[d]) for d in range(2)]
    X, Y = np.meshgrid(xs[0], xs[1], indexing="ij")
    pts = np.column_stack([X.ravel(), Y.ravel()])
    inner = pts[(pts >= 0).all(1) & (pts <= 1).all(1)]
    outer = pts[~((pts >= 0).all(1) & (pts <= 1).all(1))]
    cut = (2.5 * h + h) ** 2
    d2 = ((outer[:, None, :] - inner[None, :, :]) ** 2).sum(-1)
    expected = int((d2 <= cut).any(1).sum())
    assert int(c["is_boundary"].sum()) == expected
    assert int((c["is_boundary"] == 0).sum()) == len(inner)


def test_box_tags_face_rule():
    """Domains.jl:68-94: 1 left, 2 right, 3 bottom, 4 top; corners go to the face the point is farthest outside."""
    dom = W.domain(2, (0, 0), (1, 1), (False, False), 1 / 20, 2.5, 0.0, 0)
    c = L.cloud_generate(W.domain_desc(dom))
    x, t = c["pos"], c["tags"]
    for i in np.nonzero(c["is_boundary"])[0]:
        outs = [-x[i, 0], x[i, 0] - 1, -x[i, 1], x[i, 1] - 1]
        best = int(np.argmax(outs))   # argmax keeps the first of equal values
        assert t[i] == best + 1
    assert set(np.unique(t)) == {0, 1, 2, 3, 4}


def test_walls_snap_to_lattice():
    """The lattice points on a non-periodic wall sit exactly on it (Domains.jl:243-253)."""
    dom = W.domain(1, (-1,), (1,), (False,), 2 / 30, 2.5, 0.0, 0)
    c = L.cloud_generate(W.domain_desc(dom))
    x = np.sort(c["pos"][:, 0])
    assert (x == -1.0).sum() == 1 and (x == 1.0).sum() == 1
    assert int((c["is_boundary"] == 0).sum()) == 31


def test_morton_order_is_sorted_by_key():
    dom = W.domain(2, (0, 0), (1, 1), (False, False), 1 / 50, 2.5, 0.25, 4)
    c = L.cloud_generate(W.domain_desc(dom))
    cm, ext = c["canvas"][:2], c["canvas"][2:] - c["canvas"][:2]

    def spread(v):
        v &= 0xFFFF
        v = (v | (v << 8)) & 0x00FF00FF
        v = (v | (v << 4)) & 0x0F0F0F0F
        v = (v | (v << 2)) & 0x33333333
        return (v | (v << 1)) & 0x55555555

    keys = []
    for p in c["pos"]:
        k = [int(min(max(np.floor((p[d] - cm[d]) / ext[d] * 65535.0), 0), 65535)) for d in range(2)]
        keys.append(spread(k[0]) | (spread(k[1]) << 1))
    assert np.all(np.diff(np.array(keys, np.int64)) >= 0)
