"""classify_contour_type and mesh_is_closed of the oracle (intersection_resolver2.cpp:
311-375, :2823-2841).

Known answers of UnitTest/test_prp_selection.cpp:518-573: EF pairs built as the
test's make_pair (a non-loop pair i has edge vertices (2i, 2i + 1), a loop-vertex
pair has both edge vertices i), no boundary pairs, all on mesh 0:
  three non-loop pairs in a triangle -> Closed; in a chain (two degree-one
  endpoints) -> Undefined; one loop pair (its own degree one) -> Eight; two loop
  pairs -> LL.
Cases with boundary pairs follow the same function's branches: two boundary pairs
on one mesh -> BBII, on both meshes -> BIBI, one loop pair with one boundary pair
-> BLI, with two -> Cross, three boundary pairs -> Undefined; and a role is closed
when it owns a pair and none of its non-loop pairs has degree one."""
import numpy as np

import oracle

UNDEF, CLOSED, EIGHT, LL, BLI, CROSS, BIBI, BBII = range(8)


def make_pairs(loops):
    P = np.zeros((len(loops), 12), np.uint32)
    for i, lp in enumerate(loops):
        P[i, :2] = (i, i) if lp else (2 * i, 2 * i + 1)
        P[i, 2:5] = (10 + i, 20 + i, 30 + i)
    return P


def classify(loops, adjacency, boundary=None, mesh=None):
    n = len(loops)
    adjn = [len(a) for a in adjacency]
    b = np.zeros(n, np.uint8) if boundary is None else np.asarray(boundary, np.uint8)
    mi = np.zeros(n, np.uint32) if mesh is None else np.asarray(mesh, np.uint32)
    t, closed = oracle.classify_contours(make_pairs(loops), b, adjn, np.zeros(n, np.uint32), mi, 1)
    return int(t[0]), closed[0].tolist()


def test_reference_known_answers():
    assert classify([0, 0, 0], [[1, 2], [0, 2], [0, 1]])[0] == CLOSED
    assert classify([0, 0, 0], [[1], [0, 2], [1]])[0] == UNDEF
    assert classify([1], [[0]])[0] == EIGHT
    assert classify([1, 1], [[1], [0]])[0] == LL


def test_open_contours():
    chain = [[1], [0, 2], [1]]
    assert classify([0, 0, 0], chain, boundary=[1, 0, 1], mesh=[0, 1, 0])[0] == BBII
    assert classify([0, 0, 0], chain, boundary=[1, 0, 1], mesh=[0, 1, 1])[0] == BIBI
    assert classify([0, 0, 0], chain, boundary=[1, 1, 1])[0] == UNDEF
    assert classify([0, 1, 0], chain, boundary=[1, 0, 0])[0] == BLI
    assert classify([0, 1, 0], chain, boundary=[1, 0, 1])[0] == CROSS
    # a closed cycle with a loop pair inside but a boundary pair: not Eight
    assert classify([0, 1, 0], [[1, 2], [0, 2], [0, 1]], boundary=[1, 0, 0])[0] == BLI


def test_mesh_is_closed_roles():
    # triangle on mesh 0: role 0 closed, role 1 owns nothing
    assert classify([0, 0, 0], [[1, 2], [0, 2], [0, 1]])[1] == [1, 0]
    # chain across both meshes: the degree-one endpoints open their roles
    assert classify([0, 0, 0], [[1], [0, 2], [1]], mesh=[0, 1, 1])[1] == [0, 0]
    # a degree-one loop pair does not open its role
    assert classify([1, 0, 0], [[1], [0, 2], [1, 2]], mesh=[1, 0, 0])[1] == [1, 1]
