"""Small affine-body scenes for the parity tests: a few bodies placed by hand so
that proximity pairs (gap below d_hat) and intersections (overlap) occur."""
import numpy as np

import oracle
from untangling26_amd import initializer, scenes


def hand_scene(specs, seed=0, rotate=True):
    """specs: list of (kind, radius, centre) with kind 0 icosphere (level 2), 1 box (4x4x4);
    random orientations from the seed (identity with rotate=False)."""
    rng = np.random.default_rng(seed)
    sv, sf = scenes.icosphere(2, 1.0)
    bv, bf = scenes.subdivided_box(4, 1.0)
    xs, fs, body, xbar, q0, off = [], [], [], [], [], [0]
    for b, (kind, r, c) in enumerate(specs):
        v, f = (sv, sf) if kind == 0 else (bv, bf)
        R = scenes._random_rotation(rng) if rotate else np.eye(3)
        model = v * r
        xs.append(np.asarray(c) + model @ R.T)
        fs.append(f + off[-1])
        body.append(np.full(len(v), b, np.uint32))
        xbar.append(model)
        q0.append(np.stack([np.asarray(c, np.float64), R[:, 0], R[:, 1], R[:, 2]]))
        off.append(off[-1] + len(v))
    mesh = scenes.Mesh(np.concatenate(xs), np.concatenate(fs), name="hand_bodies")
    return scenes.RigidScene(mesh, np.concatenate(body), np.concatenate(xbar).astype(np.float32), np.asarray(q0),
                             np.array([s[1] for s in specs]), np.array([s[0] for s in specs], np.uint8),
                             np.asarray(off, np.uint32))


def near_pair():
    """A box and a sphere whose surfaces come within d_hat of each other without
    intersecting, plus a third body far away: proximity records between bodies 0
    and 1 only.  The sphere slides along -x towards the box: bisection for the first
    shift with an edge-face intersection, then back in 0.05 mm steps until no
    intersection remains and proximity pairs do."""
    sc = hand_scene([(1, 0.1, (0.0, 0.0, 0.0)), (0, 0.08, (0.5, 0.0, 0.0)), (0, 0.05, (2.0, 2.0, 2.0))], seed=3)
    x0 = sc.mesh.x.astype(np.float64)
    sph = sc.body == 1
    st0 = setup(sc)

    def at(shift):
        x = x0.copy()
        x[sph] -= np.array([shift, 0.0, 0.0])
        return x.astype(np.float32)

    def ef(shift):
        return len(oracle.ef_detect(at(shift), st0.edges, st0.faces, st0.d_hat, st0.offset, vprop=st0.vprop)[0])

    lo, hi = 0.0, 0.5
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        lo, hi = (lo, mid) if ef(mid) else (mid, hi)
    shift = hi
    while ef(shift) or len(records(st0, at(shift))) == 0:
        shift -= 5e-5
        if shift < hi - 2e-3:
            raise RuntimeError("near_pair: no separating shift with proximity pairs")
    sc.q0[1, 0] -= np.array([shift, 0.0, 0.0])
    sc.mesh.x[:] = at(shift)
    return sc


def stacked_boxes(gap=5e-4):
    """Two axis-aligned boxes (half-extent 0.1) with parallel faces `gap` apart and
    partly overlapping in the plane, plus a sphere resting on the upper one at the
    same gap along y: many VF and EE proximity records between three bodies."""
    top = 0.2 + gap
    return hand_scene([(1, 0.1, (0.0, 0.0, 0.0)), (1, 0.1, (0.035, top, 0.02)),
                       (0, 0.06, (0.03, top + 0.1 + 0.06 + gap, 0.0))], rotate=False)


def records(st, x=None, energy=1, kappa=1e8):
    """Oracle VF + EE proximity records of the body mesh (pairs inside one body skipped)."""
    x = st.x if x is None else np.asarray(x, np.float32)
    qv = oracle.query_boxes(0, x, x, None, st.d_hat, st.offset)
    vf, _ = oracle.dcd_list(0, oracle.overlap_pairs(qv, oracle.leaf_boxes(1, x, x, st.faces, st.offset)), x,
                            st.edges, st.faces, st.d_hat, st.offset, st.rest_vert_area, st.rest_edge_area,
                            st.rest_face_area, kappa, energy, vprop=st.vprop)
    qe = oracle.query_boxes(1, x, x, st.edges, st.d_hat, st.offset)
    ee, _ = oracle.dcd_list(1, oracle.overlap_pairs(qe, oracle.leaf_boxes(2, x, x, st.edges, st.offset), cull=True),
                            x, st.edges, st.faces, st.d_hat, st.offset, st.rest_vert_area, st.rest_edge_area,
                            st.rest_face_area, kappa, energy, vprop=st.vprop)
    return np.concatenate([vf, ee])


def setup(sc, **kw):
    return initializer.setup_bodies(sc, initializer.RigidMaterial(**kw))
