"""Host-side setup of the Laplace time field (sttopt/virtual_heat.py:465-543): the
domain's node graph, the wall's distance to the plate through the domain, and the cone
filter along the wall arc with its odd reflection at the plate corners -- against
closed forms (distances on rectangles and an L-shaped domain, ramps) and an explicit
extended-signal convolution."""

import numpy as np
import pytest


def _vh():
    from sttopt_amd import virtual_heat

    return virtual_heat


@pytest.mark.parametrize("base", ["bottom_edge", "left_edge", "right_edge"])
def test_plate_distance_on_a_rectangle_is_the_distance_to_the_plate_edge(base):
    vh = _vh()
    nelx, nely = 7, 4
    mesh = vh.ScalarMesh.build(nelx, nely, base, "cpu")
    d = (vh._plate_distance(mesh) * mesh.unit_length).numpy()
    r, c = np.divmod(mesh.arc.numpy(), nelx + 1)
    want = {"bottom_edge": nely - r, "left_edge": c, "right_edge": nelx - c}[base]
    np.testing.assert_allclose(d, want, atol=1e-12)
    w = vh.distance_wall(mesh).numpy()
    np.testing.assert_allclose(w, want / (2 * want.max()), atol=1e-12)
    assert w.max() == pytest.approx(0.5)


def test_plate_distance_goes_around_a_reentrant_corner():
    """An L-shaped domain (the top-right block of elements removed), plate at the
    bottom: distances along the walls of the notch, and the node graph has no link
    through a removed element."""
    vh = _vh()
    nelx, nely = 8, 6
    active = np.ones((nely, nelx), bool)
    active[:3, 4:] = False  # element rows 0-2, columns 4-7 removed
    mesh = vh.ScalarMesh.build(nelx, nely, "bottom_edge", "cpu", active=active)
    d = dict(zip(mesh.arc.numpy().tolist(),
                 (vh._plate_distance(mesh) * mesh.unit_length).numpy().tolist()))
    node = lambda r, c: r * (nelx + 1) + c  # noqa: E731
    # the top-left corner of the domain: 6 rows above the plate
    assert d[node(0, 0)] == pytest.approx(6.0)
    # the node on the notch's floor at (3, 8): 3 rows above the plate
    assert d[node(3, 8)] == pytest.approx(3.0)
    # the notch's inner corner (3, 4) and a node on its vertical side (1, 4)
    assert d[node(3, 4)] == pytest.approx(3.0)
    assert d[node(1, 4)] == pytest.approx(5.0)
    # edges only through domain elements: a diagonal across a removed element is absent
    src, dst, step = vh._node_edges(active)
    pairs = set(zip(src.tolist(), dst.tolist()))
    assert (node(2, 4), node(3, 5)) not in pairs and (node(3, 4), node(3, 5)) in pairs
    assert np.allclose(np.sort(step)[[0, -1]], [1.0, np.sqrt(2.0)])
    assert len(pairs) == src.size  # each link once per direction


def _filter_by_extension(b, rmin):
    """The cone filter of b continued past both ends by odd reflection about the
    plate corners -1 and n, as an explicit convolution of the extended signal."""
    n = b.size
    pad = int(np.ceil(rmin)) + n
    pos = np.arange(-pad, n + pad)
    ext = np.zeros(pos.size)
    for j in range(n):
        ext[pos == j] += b[j]
        ext[pos == -2 - j] -= b[j]
        ext[pos == 2 * n - j] -= b[j]
    own = max(rmin, 1.0)
    cone = lambda d: np.where(d == 0, own, np.maximum(0.0, rmin - np.abs(d)))  # noqa: E731
    total = cone(np.arange(-pad, pad + 1)).sum()
    return np.array([(cone(i - pos) * ext).sum() for i in range(n)]) / total


@pytest.mark.parametrize("n, rmin", [(12, 3.0), (9, 2.5), (5, 4.0), (6, 1.0), (7, 0.5)])
def test_wall_filter_is_the_reflected_cone(n, rmin):
    vh = _vh()
    H = vh.wall_filter(n, rmin)
    assert H.shape == (n, n)
    rng = np.random.default_rng(n)
    b = rng.uniform(0, 1, n)
    np.testing.assert_allclose(H @ b, _filter_by_extension(b, rmin), atol=1e-14)
    np.testing.assert_allclose(H, H.T, atol=1e-15)
    assert np.all(H @ b >= -1e-15)  # a nonnegative wall stays nonnegative
    if rmin <= 1:
        np.testing.assert_allclose(H, np.eye(n), atol=1e-15)


def test_wall_filter_passes_a_ramp_from_the_plate():
    """A wall rising linearly from the plate corner at -1 passes unchanged wherever the
    cone does not reach the other corner at n."""
    vh = _vh()
    n, rmin = 20, 3.0
    ramp = (np.arange(n) + 1.0) / n
    out = vh.wall_filter(n, rmin) @ ramp
    near = np.arange(n) < n - 1 - int(np.ceil(rmin))
    np.testing.assert_allclose(out[near], ramp[near], atol=1e-14)
