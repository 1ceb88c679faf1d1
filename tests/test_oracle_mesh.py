"""The oracle's voxel -> hex mesher against hand-worked cases and the
invariants that fix first-touch numbering, coordinates and the keep rule
(f-src/mod_gen_quadmesh.f90:22-679)."""
import numpy as np
import pytest

from helpers import random_mask
from oracle import oracle as O

OFF8 = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
OFF20 = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [2, 1, 0], [2, 2, 0], [1, 2, 0], [0, 2, 0], [0, 1, 0],
                  [0, 0, 1], [2, 0, 1], [2, 2, 1], [0, 2, 1], [0, 0, 2], [1, 0, 2], [2, 0, 2], [2, 1, 2],
                  [2, 2, 2], [1, 2, 2], [0, 2, 2], [0, 1, 2]])


def test_single_voxel():
    nodes, elems = O.mesh(np.ones((1, 1, 1), np.int8), 1, 8, (0.5, 0.5, 0.5), (3, 3, 3))
    assert elems.tolist() == [[1, 2, 3, 4, 5, 6, 7, 8]]
    assert np.array_equal(nodes, (OFF8 + 2) * 0.5)


def test_two_voxels_along_x_and_z():
    _, e = O.mesh(np.ones((1, 1, 2), np.int8), 1)           # (nz, ny, nx) = (1, 1, 2)
    assert e.tolist() == [[1, 2, 3, 4, 5, 6, 7, 8], [2, 9, 10, 3, 6, 11, 12, 7]]
    _, e = O.mesh(np.ones((2, 1, 1), np.int8), 1)
    assert e.tolist() == [[1, 2, 3, 4, 5, 6, 7, 8], [5, 6, 7, 8, 9, 10, 11, 12]]


def test_threshold_and_signed_kinds():
    phi = np.array([[[5000, 4999], [20000, -1000]]], np.int16)   # llimit 5000 (BINARIZE_LO)
    _, e = O.mesh(phi, 5000)
    assert e.shape == (2, 8)
    u8 = np.array([[[200, 10]]], np.uint8)
    _, e = O.mesh(u8.view(np.int8), 100)          # 200 reads as -56 under ik1 (Appendix A12)
    assert e.shape[0] == 0


def test_empty_domain():
    nodes, elems = O.mesh(np.zeros((3, 3, 3), np.int8), 1)
    assert nodes.shape[0] == 0 and elems.shape[0] == 0


def test_component_inside_bbox_is_dropped():
    phi = np.zeros((6, 6, 6), np.int8)
    phi[0, :, :] = 1      # plane z = 0
    phi[5, :, :] = 1      # plane z = 5
    phi[2, 2, 2] = 1      # floating voxel strictly inside the node bounding box
    nodes, elems = O.mesh(phi, 1)
    assert elems.shape[0] == 72 and nodes.shape[0] == 2 * 7 * 7 * 2
    assert elems.min() >= 1 and set(np.unique(elems)) == set(range(1, nodes.shape[0] + 1))
    z = nodes[elems - 1][:, :, 2]
    assert np.all((z <= 1) | (z >= 5))


def test_component_touching_bbox_is_kept():
    phi = np.zeros((6, 6, 6), np.int8)
    phi[0, :, :] = 1
    phi[3, 5, 5] = 1      # touches x/y max of the bounding box, disconnected from the plane
    _, elems = O.mesh(phi, 1)
    assert elems.shape[0] == 37


@pytest.mark.parametrize("seed,shape,p", [(1, (5, 6, 7), 0.5), (2, (8, 8, 8), 0.3), (3, (3, 9, 4), 0.7),
                                          (4, (10, 10, 10), 0.15)])
def test_first_touch_and_coordinates(seed, shape, p):
    phi = random_mask(shape, p, seed)
    delta, xa = (0.01, 0.02, 0.015), (3, 5, 7)
    nodes, elems = O.mesh(phi, 1, 8, delta, xa)
    # first appearance order of node ids while scanning elements is 1, 2, 3, ...
    seen, nxt = set(), 1
    for e in elems:
        for v in e:
            if v not in seen:
                assert v == nxt
                seen.add(v)
                nxt += 1
    # each element is a unit voxel: node l sits at its corner offset
    base = nodes[elems[:, 0] - 1]
    for l in range(8):
        assert np.allclose(nodes[elems[:, l] - 1] - base, OFF8[l] * np.array(delta), atol=1e-12)
    # kept elements are the solid voxels of kept components, in scan order
    vox = np.argwhere(phi[:, :, :] >= 1)[:, ::-1]           # (x, y, z) in scan order
    corner = np.rint(base / np.array(delta) - (np.array(xa) - 1)).astype(int)
    assert set(map(tuple, corner)).issubset(set(map(tuple, vox)))
    lin = corner[:, 0] + shape[2] * (corner[:, 1] + shape[1] * corner[:, 2])
    assert np.all(np.diff(lin) > 0)


def test_hex20_single_voxel_and_coordinates():
    nodes, elems = O.mesh(np.ones((1, 1, 1), np.int32), 1, 20, (0.2, 0.2, 0.2), (2, 2, 2))
    assert elems.tolist() == [list(range(1, 21))]
    # first-touch order over el_nn(1:20) on the doubled lattice; coordinate (k/2 + xa - 1) * delta
    assert np.allclose(nodes, (OFF20 / 2.0 + 1) * 0.2)


def test_hex20_two_voxels_share_face_nodes():
    nodes, elems = O.mesh(np.ones((1, 1, 2), np.int8), 1, 20)
    assert nodes.shape[0] == 20 + 20 - 8
    shared = set(elems[0]) & set(elems[1])
    assert len(shared) == 8
