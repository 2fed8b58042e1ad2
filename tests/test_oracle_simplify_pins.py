"""Pins the oracle's vertex-valence simplifications with hand-built quad meshes whose
results were derived by hand from the reference kernels:

  simplify3X3Y (colour)   markElementsP3X3Y, mergeVerticesP3X3YColor, handleMerge
                          (HalfedgeMesh.cu:317-339, :368-449), compaction :500-538
  simplify3X3YOld         countVerticesP3X3Y, mergeVerticesP3X3Y (:341-366, :455-498)
  simplify3333            markElementsP3333, removeVerticesP3333, removeQuadsP3333
                          (:547-683)
  valence                 computeVertexValence (:296-311): one per outgoing halfedge, plus
                          one at the end vertex of every halfedge without a twin

Every expected array below is a literal worked out on paper (the valences, the pattern
marks, the merge target, the midpoint in float32, the stable compaction's new ids).  Each
planted mistake (oracle/mdmc_oracle.cpp, orc_set_plant) must make at least one of these
cases fail, so a misreading of that kind cannot pass unnoticed.

Mesh "P" (P3X3Y), vertex ids and positions (z = 0 unless given):
   0 x (-1,0)      1 w (0,-1,0.5)   2 z (1.5,0)     3 y (0.5,1.5,-0.25)
   4 a (-2,0.5)    5 p (-1,1.5)     6 q (-1,-1.5)   7 c (2.5,0)
   8 p' (1.5,1.5)  9 q' (1.5,-1.5) 10 t (0.5,2.5)  11 u1 (-1,2.5)
  12 u2 (1.5,2.5) 13 s (0,-2.5)    14 v1 (-1,-2.5) 15 v2 (1.5,-2.5)
  16 h (-2,2)     17 g1 (-2.5,-1)  18 g2 (-2,-2)
 quads q0 (x,w,z,y)  A (x,y,p,a)  B (w,x,a,q)  C (y,z,c,p')  D (z,w,q',c)
       E1 (p,y,t,u1) E2 (y,p',u2,t) F1 (w,q,v1,s) F2 (q',w,s,v2) [G1 (q,a,g1,g2)] H (a,p,u1,h)
 Without G1 the valences are x 3, w 5, z 3, y 5, a 4, p 3, q 3, ...: only q0 matches
 3-(>=5)-3-(>=5).  With G1, a has 4 outgoing halfedges plus the twinless h->a: valence 5,
 so A = (x 3, y 5, p 3, a 5) is a second pattern quad, sharing the edge x-y with q0.
"""
import numpy as np
import pytest

import oracle_api as oa

COLOR_BITS, NM = 0x1F, 0x20

# ---------------------------------------------------------------------------------------
# mesh P
P_POS = np.array([
    (-1, 0, 0), (0, -1, 0.5), (1.5, 0, 0), (0.5, 1.5, -0.25),
    (-2, 0.5, 0), (-1, 1.5, 0), (-1, -1.5, 0), (2.5, 0, 0),
    (1.5, 1.5, 0), (1.5, -1.5, 0), (0.5, 2.5, 0), (-1, 2.5, 0),
    (1.5, 2.5, 0), (0, -2.5, 0), (-1, -2.5, 0), (1.5, -2.5, 0),
    (-2, 2, 0), (-2.5, -1, 0), (-2, -2, 0)], np.float32)
P_NRM = np.tile(np.array([0, 0, 1], np.float32), (19, 1))
P_NRM[1] = (1, 0, 0)   # w
P_NRM[3] = (0, 1, 0)   # y
P_NRM[4] = (0, 0, -1)  # a
X, W, Z, Y, A_, PP, Q_, C_, P2, Q2, T, U1, U2, S, V1, V2, H_, G1, G2 = range(19)
P_Q0, P_A, P_B = (X, W, Z, Y), (X, Y, PP, A_), (W, X, A_, Q_)
P_C, P_D = (Y, Z, C_, P2), (Z, W, Q2, C_)
P_E1, P_E2 = (PP, Y, T, U1), (Y, P2, U2, T)
P_F1, P_F2 = (W, Q_, V1, S), (Q2, W, S, V2)
P_G1, P_H = (Q_, A_, G1, G2), (A_, PP, U1, H_)

SQRT_HALF = np.float32(np.sqrt(0.5))  # normalize((0.5, 0.5, 0)) = (sqrt 1/2, sqrt 1/2, 0)


def mesh_isolated():
    """Mesh P without G1 (vertices 0..16): q0 is the only pattern quad."""
    quads = [P_Q0, P_A, P_B, P_C, P_D, P_E1, P_E2, P_F1, P_F2, P_H]
    return P_POS[:17].copy(), P_NRM[:17].copy(), quads


def mesh_pair():
    """Mesh P with G1: q0 and A are adjacent pattern quads."""
    quads = [P_Q0, P_A, P_B, P_C, P_D, P_E1, P_E2, P_F1, P_F2, P_G1, P_H]
    return P_POS.copy(), P_NRM.copy(), quads


def expect_q0_merged(n_vertices):
    """q0 merges: x moves to w + 0.5 (y - w) = (0.25, 0.25, 0.125), its normal to
    normalize((1,0,0) + 0.5 ((0,1,0) - (1,0,0))); z (id 2) is dropped and maps onto x."""
    keep = [v for v in range(n_vertices) if v != Z]
    pos = P_POS[keep].copy()
    nrm = P_NRM[keep].copy()
    pos[0] = (0.25, 0.25, 0.125)
    nrm[0] = (SQRT_HALF, SQRT_HALF, 0)
    # new ids: x 0, w 1, y 2, a 3, p 4, q 5, c 6, p' 7, q' 8, t 9, u1 10, u2 11, s 12, v1 13,
    # v2 14, h 15, g1 16, g2 17; z -> x
    faces = [(0, 2, 4, 3), (1, 0, 3, 5), (2, 0, 6, 7), (0, 1, 8, 6), (4, 2, 9, 10), (2, 7, 11, 9),
             (1, 5, 13, 12), (8, 1, 12, 14)]
    if n_vertices == 19:
        faces.append((5, 3, 16, 17))
    faces.append((3, 4, 10, 15))
    return pos, nrm, faces


def expect_a_merged():
    """A (x,y,p,a) merges as pattern 3-(>=5)-3-(>=5): x moves to y + 0.5 (a - y) =
    (-0.75, 1, -0.125), normal normalize((0,1,0) + 0.5 ((0,0,-1) - (0,1,0))) =
    (0, sqrt 1/2, -sqrt 1/2); p (id 5) is dropped onto x."""
    keep = [v for v in range(19) if v != PP]
    pos = P_POS[keep].copy()
    nrm = P_NRM[keep].copy()
    pos[0] = (-0.75, 1.0, -0.125)
    nrm[0] = (0, SQRT_HALF, -SQRT_HALF)
    # new ids: x 0, w 1, z 2, y 3, a 4, q 5, c 6, p' 7, q' 8, t 9, u1 10, u2 11, s 12, v1 13,
    # v2 14, h 15, g1 16, g2 17; p -> x
    faces = [(0, 1, 2, 3), (1, 0, 4, 5), (3, 2, 6, 7), (2, 1, 8, 6), (0, 3, 9, 10), (3, 7, 11, 9),
             (1, 5, 13, 12), (8, 1, 12, 14), (5, 4, 16, 17), (4, 0, 10, 15)]
    return pos, nrm, faces


def expect_unchanged(pos, nrm, quads):
    return pos.copy(), nrm.copy(), [tuple(q) for q in quads]


# ---------------------------------------------------------------------------------------
# mesh "R" (P3333): a centre quad c0..c3 inside o0..o3, four trapezoids N_i =
# (c_{i+1}, c_i, o_i, o_{i+1}) and optional outer quads M_i = (o_{i+1}, o_i, m_i, m'_i).
# c_i has valence 3 (centre, N_{i-1}, N_i).  The far vertex across the centre's edge i
# (twin of c_i -> c_{i+1} is c_{i+1} -> c_i in N_i; next, next starts at o_i) is o_i.
# o_i's valence: one per face at o_i plus one twinless halfedge ending there, so 3 with
# no M at o_i, 4 with one, 5 with two.
R_C = [(-1, -1), (1, -1), (1, 1), (-1, 1)]
R_O = [(-3, -3), (3, -3), (3, 3), (-3, 3)]
R_OUT = [(0, -2), (2, 0), (0, 2), (-2, 0)]  # outward offsets of edges 0..3


def mesh_ring(ms, extra_nm=False):
    """Centre + trapezoids + the outer quads M_i for i in ms.  Vertex ids: c_i = i, o_i = 4 + i,
    then m_i, m'_i for each i in ms in order.  extra_nm adds two quads on the edge o0-o1
    (X1 = (o0, o1, e1, e2), X2 = (o1, o0, e3, e4)), so that edge has four halfedges (N0, M0,
    X1, X2) and N0 is non-manifold."""
    pos = [(x, y, 0) for x, y in R_C] + [(x, y, 0) for x, y in R_O]
    quads = [(0, 1, 2, 3)] + [((i + 1) % 4, i, 4 + i, 4 + (i + 1) % 4) for i in range(4)]
    for i in ms:
        m = len(pos)
        ox, oy = R_O[i]
        px, py = R_O[(i + 1) % 4]
        dx, dy = R_OUT[i]
        pos += [(ox + dx, oy + dy, 0), (px + dx, py + dy, 0)]
        quads.append((4 + (i + 1) % 4, 4 + i, m, m + 1))
    if extra_nm:
        e = len(pos)
        pos += [(-3, -5, 1), (3, -5, 1), (3, -5, 2), (-3, -5, 2)]
        quads += [(4, 5, e, e + 1), (5, 4, e + 2, e + 3)]
    pos = np.array(pos, np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (len(pos), 1))
    return pos, nrm, quads


def expect_ring_collapsed(ms):
    """The centre's corners are dropped (ids 0..3), each mapped onto o_i; the trapezoids go;
    the centre becomes (o0, o1, o2, o3) = (0, 1, 2, 3) and M_i keeps its corners, renumbered.
    No position or normal changes."""
    pos, nrm, _ = mesh_ring(ms)
    faces = [(0, 1, 2, 3)]
    for k, i in enumerate(ms):
        faces.append(((i + 1) % 4, i, 4 + 2 * k, 5 + 2 * k))
    return pos[4:].copy(), nrm[4:].copy(), faces


# ---------------------------------------------------------------------------------------
def run(pos, nrm, quads, flags, colors=None):
    attr = np.full(len(quads), COLOR_BITS, np.uint16)
    if colors:
        for f, c in colors.items():
            attr[f] = c
    return oa.simplify_quads(pos, nrm, np.array(quads, np.int32), attr, flags)


def faces_of(m):
    f = []
    for h in m.face_he:
        cyc = []
        for _ in range(4):
            cyc.append(int(m.he[h, 0]))
            h = m.he[h, 2]
        f.append(tuple(cyc))
    return f


def mismatches(m, expected, nm_faces=()):
    pos, nrm, faces = expected
    out = []
    if len(m.pos) != len(pos):
        return ["vertex count %d != %d" % (len(m.pos), len(pos))]
    if faces_of(m) != [tuple(f) for f in faces]:
        out.append("faces %s != %s" % (faces_of(m), faces))
    if not np.array_equal(m.pos.view(np.uint32), pos.view(np.uint32)):
        out.append("positions differ")
    if not np.array_equal(m.nrm.view(np.uint32), nrm.view(np.uint32)):
        out.append("normals differ")
    want_attr = [COLOR_BITS | (NM if f in nm_faces else 0) for f in range(len(faces))]
    if [int(a) for a in m.face_attr] != want_attr:
        out.append("face attributes %s" % list(m.face_attr))
    # halfedges of the rebuilt mesh: 4f+i starts at corner i of face f, twins pair up
    he = m.he
    for h in range(len(he)):
        t = he[h, 3]
        if t >= 0 and he[t, 3] != h and not nm_faces:
            out.append("twin of %d" % h)
            break
    return out


# each case: (flags, builder) -> (mesh result, expected, nm face ids of the result)
def case_isolated_colour():
    pos, nrm, quads = mesh_isolated()
    return run(pos, nrm, quads, oa.P3X3Y_COLOR), expect_q0_merged(17), ()


def case_isolated_old():
    pos, nrm, quads = mesh_isolated()
    return run(pos, nrm, quads, oa.P3X3Y_OLD), expect_q0_merged(17), ()


def case_pair_colour(cq0, ca):
    pos, nrm, quads = mesh_pair()
    m = run(pos, nrm, quads, oa.P3X3Y_COLOR, colors={0: cq0, 1: ca})
    if cq0 < ca:
        exp = expect_q0_merged(19)   # A's colour is higher: only q0 merges
    elif ca < cq0:
        exp = expect_a_merged()      # q0 sees a pattern neighbour of lower colour: only A merges
    else:
        exp = expect_unchanged(pos, nrm, quads)  # equal colours stop both (<=)
    return m, exp, ()


def case_pair_old():
    # x is a valence-3 corner of both pattern quads (elementCount 2): neither merges
    pos, nrm, quads = mesh_pair()
    return run(pos, nrm, quads, oa.P3X3Y_OLD), expect_unchanged(pos, nrm, quads), ()


def case_isolated_nm_neighbour():
    # three more quads on D's boundary edge q'-c make it a four-halfedge edge: D is
    # non-manifold, so q0 (a neighbour of D) must not merge.  They also raise c's valence
    # to 6, which makes C = (y 5, z 3, c 6, p' 3) a pattern quad; q0 gets colour 1 (the
    # others keep 31) so that C's colour does not stop it: only D's NM flag does.
    pos, nrm, quads = mesh_isolated()
    e = len(pos)
    pos = np.vstack([pos, np.array([(2.5, -2.5, 1), (3, -2, 1), (2.5, -2.5, 2), (3, -2, 2), (2.5, -2.5, 3), (3, -2, 3)],
                                   np.float32)])
    nrm = np.vstack([nrm, np.tile(np.array([0, 0, 1], np.float32), (6, 1))])
    quads = quads + [(C_, Q2, e, e + 1), (Q2, C_, e + 2, e + 3), (C_, Q2, e + 4, e + 5)]
    nm_faces = (4, 10, 11, 12)  # D and the three added quads
    return run(pos, nrm, quads, oa.P3X3Y_COLOR, colors={0: 1}), expect_unchanged(pos, nrm, quads), nm_faces


def case_ring(ms, extra_nm=False):
    pos, nrm, quads = mesh_ring(ms, extra_nm)
    m = run(pos, nrm, quads, oa.P3333)
    if extra_nm:
        # N0 (face 1), M0 (face 5) and the two added quads are non-manifold: the centre keeps
        return m, expect_unchanged(pos, nrm, quads), (1, 5, 9, 10)
    val = [2 + sum(1 for i in ms if i in (k, (k - 1) % 4)) + 1 for k in range(4)]  # o_k's valence
    if 3 in val or all(v == 4 for v in val):
        return m, expect_unchanged(pos, nrm, quads), ()
    return m, expect_ring_collapsed(ms), ()


CASES = {
    "p3x3y_isolated_colour": case_isolated_colour,
    "p3x3y_isolated_old": case_isolated_old,
    "p3x3y_pair_q0_lower": lambda: case_pair_colour(1, 2),
    "p3x3y_pair_a_lower": lambda: case_pair_colour(2, 1),
    "p3x3y_pair_equal": lambda: case_pair_colour(3, 3),
    "p3x3y_pair_old": case_pair_old,
    "p3x3y_nm_neighbour": case_isolated_nm_neighbour,
    "p3333_ring_all": lambda: case_ring([0, 1, 2, 3]),          # o valences 5 5 5 5: collapse
    "p3333_ring_4554": lambda: case_ring([0, 1, 2]),            # 4 5 5 4: collapse
    "p3333_ring_all_four": lambda: case_ring([0, 2]),           # 4 4 4 4: rejected
    "p3333_ring_valence3": lambda: case_ring([]),               # 3 3 3 3: rejected
    "p3333_ring_nm_neighbour": lambda: case_ring([0, 1, 2, 3], extra_nm=True),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_hand_derived_case(name):
    m, exp, nm_faces = CASES[name]()
    assert mismatches(m, exp, nm_faces) == []
    assert m.stats["anomalies"] == 0


def test_expected_counts():
    # the V / F deltas of a P3X3Y merge (-1 / -1) and a P3333 collapse (-4 / -4)
    m, _, _ = case_isolated_colour()
    assert (len(m.pos), len(m.face_he)) == (17 - 1, 10 - 1)
    m, _, _ = case_ring([0, 1, 2, 3])
    assert (len(m.pos), len(m.face_he)) == (16 - 4, 9 - 4)


# each planted mistake and the cases that must catch it
PLANT_CATCHES = {
    oa.PLANT_SWAP_SURVIVOR: ["p3x3y_isolated_colour", "p3x3y_isolated_old", "p3x3y_pair_q0_lower"],
    oa.PLANT_MID_XZ: ["p3x3y_isolated_colour", "p3x3y_pair_a_lower"],
    oa.PLANT_NO_RENORM: ["p3x3y_isolated_colour", "p3x3y_pair_a_lower"],
    oa.PLANT_COLOR_LT: ["p3x3y_pair_equal"],
    oa.PLANT_OLD_NO_COUNT: ["p3x3y_pair_old"],
    oa.PLANT_NO_NEIGHBOR_NM: ["p3x3y_nm_neighbour", "p3333_ring_nm_neighbour"],
    oa.PLANT_FAR_CORNER: ["p3333_ring_all", "p3333_ring_4554"],
    oa.PLANT_NO_ALL_FOUR: ["p3333_ring_all_four"],
}


@pytest.mark.parametrize("plant", sorted(PLANT_CATCHES))
def test_planted_mistake_is_caught(plant):
    with oa.planted(plant):
        for name in PLANT_CATCHES[plant]:
            m, exp, nm_faces = CASES[name]()
            assert mismatches(m, exp, nm_faces) != [], (plant, name)
    # and the unplanted oracle passes them again
    for name in PLANT_CATCHES[plant]:
        m, exp, nm_faces = CASES[name]()
        assert mismatches(m, exp, nm_faces) == []
