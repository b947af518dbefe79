"""UG multigrid files (ug_amd.mgio over include/ugmg_mgio.h), host only.

Files are written here as SaveMultiGrid writes them (tests/ug_mgfile.py,
a restatement of gm/mgio.cc's Write_* and low/bio.cc's encodings) for
hierarchies whose geometry and transfer are known, then read back:
uniform red refinement of the unit square (triangles) and of the unit cube
(hexahedra), a locally refined triangle mesh with copies and a green
bisection, and a moved corner.  Checks: node counts and positions level by
level, node IDs ascending, elements tiling the domain, the interpolation
reproducing linear functions, the hexahedral transfer equal to the Q1
generator's (problems.jump3d's P, the trilinear GNs weights) after mapping
nodes by position, the three encodings giving the same multigrid, and the
refusals (parallel files, debug extensions, damaged files).  No file
written by UG itself exists in the reference, so the format is pinned
against mgio.cc's writers as restated, not against a UG-written file.
"""
import itertools

import numpy as np
import pytest

import ug_mgfile as W

TRI, HEX = 3, 7
TRI_GE = (3, [(0, 1), (1, 2), (2, 0)], [(0, 1), (1, 2), (2, 0)])
# rules for triangles: copy, red, green bisection of edge 0
TRI_RULES = [
    (0, [0, 0, 0, 0, 0], [(TRI, [0, 1, 2])]),
    (1, [1, 1, 1, 0, 0], [(TRI, [0, 3, 5]), (TRI, [3, 1, 4]), (TRI, [5, 4, 2]), (TRI, [4, 5, 3])]),
    (2, [1, 0, 0, 0, 0], [(TRI, [0, 3, 2]), (TRI, [3, 1, 2])]),
]
# the unit hexahedron's corners, edges and sides (local coordinates)
HEX_C = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
HEX_E = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5), (5, 6), (6, 7), (7, 4)]
HEX_S = [(0, 3, 2, 1), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7), (4, 5, 6, 7)]


def hex_red_rule():
    """8 sons over NodeList = corners 0-7, edge midnodes 8-19, side nodes
    20-25, centre 26 (= 8 + centre index 18)."""
    loc = {}
    for i, c in enumerate(HEX_C):
        loc[tuple(2 * np.array(c))] = i
    for e, (a, b) in enumerate(HEX_E):
        loc[tuple(np.array(HEX_C[a]) + np.array(HEX_C[b]))] = 8 + e
    for s, sd in enumerate(HEX_S):
        loc[tuple((sum(np.array(HEX_C[k]) for k in sd) // 2))] = 20 + s
    loc[(1, 1, 1)] = 26
    sons = []
    for o in itertools.product((0, 1), repeat=3):
        corners = [loc[tuple(np.array(o) + np.array(c))] for c in HEX_C]
        sons.append((HEX, corners))
    return (1, [1] * 19, sons)


class Builder:
    """Refines level by level with the given rules; node IDs level by level
    in creation order (level 0: vertex IDs), positions by the rule's
    local points.  Records are then written depth first."""

    def __init__(self, dim, ge, rules, points, elements):
        self.dim, self.ge, self.rules = dim, ge, rules
        self.pos = [dict(enumerate(map(tuple, points)))]            # level -> {id: position}
        self.elems = [[(t, list(c), None) for t, c in elements]]   # (tag, corner ids, record)
        self.next_id = len(points)

    def refine(self, choose):
        """choose(level elements index, tag, corner positions) -> rule index or None."""
        l = len(self.elems) - 1
        by_pos = {}
        newpos, newel = {}, []

        def node(p):
            key = tuple(np.round(p, 12))
            if key not in by_pos:
                by_pos[key] = self.next_id
                newpos[self.next_id] = tuple(p)
                self.next_id += 1
            return by_pos[key]

        for k, (tag, cs, _) in enumerate(self.elems[l]):
            P = np.array([self.pos[l][c] for c in cs])
            r = choose(k, tag, P)
            if r is None:
                continue
            rclass, pattern, sons = self.rules[tag][r]
            nc, edges, sides = self.ge[tag]
            nl = {i: P[i] for i in range(nc)}
            ids = [node(P[i]) for i in range(nc)]
            for e, (a, b) in enumerate(edges):
                if pattern[e] == 1:
                    nl[nc + e] = (P[a] + P[b]) / 2
                    ids.append(node(nl[nc + e]))
            if self.dim == 3:
                for s, sd in enumerate(sides):
                    if pattern[len(edges) + s] == 1:
                        nl[nc + len(edges) + s] = P[list(sd)].mean(axis=0)
                        ids.append(node(nl[nc + len(edges) + s]))
            ci = 4 if self.dim == 2 else 18
            if len(pattern) > ci and pattern[ci] == 1:
                nl[nc + ci] = P.mean(axis=0)
                ids.append(node(nl[nc + ci]))
            rec = {"rule": self.rules_offset(tag) + r, "ids": ids, "sons": []}
            for st, sc in sons:
                corner_ids = [node(nl[j]) for j in sc]
                newel.append((st, corner_ids, None))
                rec["sons"].append(len(newel) - 1)
            self.elems[l][k] = (tag, cs, rec)
        self.pos.append(newpos)
        self.elems.append(newel)

    def rules_offset(self, tag):
        return sum(len(self.rules[t]) for t in sorted(self.rules) if t < tag)

    def write(self, path, mode, **kw):
        levels = len(self.elems)

        def tree(l, k):
            rec = self.elems[l][k][2]
            if rec is None:
                return None
            return {"rule": rec["rule"], "ids": rec["ids"], "moved": rec.get("moved", []),
                    "sons": [tree(l + 1, s) for s in rec["sons"]]}

        flat_rules = [r for t in sorted(self.rules) for r in self.rules[t]]
        pts = [self.pos[0][i] for i in range(len(self.pos[0]))]
        W.write(path, mode, self.dim, self.ge, flat_rules, pts, len(pts),
                [(t, c) for t, c, _ in self.elems[0]], [tree(0, k) for k in range(len(self.elems[0]))],
                levels, **kw)


def _square():
    return Builder(2, {TRI: TRI_GE}, {TRI: TRI_RULES}, [(0, 0), (1, 0), (1, 1), (0, 1)],
                   [(TRI, [0, 1, 2]), (TRI, [0, 2, 3])])


def _areas(lv):
    x = lv["xyz"]
    out = []
    for k in range(len(lv["elem_tag"])):
        c = lv["elem_corners"][lv["elem_ptr"][k]:lv["elem_ptr"][k + 1]]
        a, b, d = x[c[0]], x[c[1]], x[c[2]]
        out.append(0.5 * ((b[0] - a[0]) * (d[1] - a[1]) - (b[1] - a[1]) * (d[0] - a[0])))
    return np.array(out)


def _p_apply(lv, u):
    rows = np.repeat(np.arange(len(lv["ids"])), np.diff(lv["p_ptr"]))
    out = np.zeros(len(lv["ids"]))
    np.add.at(out, rows, lv["p_w"] * u[lv["p_idx"]])
    return out


@pytest.mark.parametrize("mode", [W.XDR, W.ASCII, W.BIN])
def test_uniform_triangles(tmp_path, mode):
    from ug_amd import mgio
    B = _square()
    for _ in range(3):
        B.refine(lambda k, t, P: 1)
    path = tmp_path / "sq.ug.mg"
    B.write(path, mode)
    mg = mgio.read(path)
    _matches_builder(mg, B)
    info = mg["info"]
    assert info["dim"] == 2 and info["nlevels"] == 4 and info["ncoarse_points"] == 4
    f = lambda X: 1.0 + 2.0 * X[:, 0] - 3.0 * X[:, 1]
    for l, lv in enumerate(mg["levels"]):
        N = 2 ** l + 1
        assert len(lv["ids"]) == N * N and np.all(np.diff(lv["ids"]) > 0)
        grid = sorted(map(tuple, np.round(lv["xyz"] * (N - 1)).astype(int).tolist()))
        assert grid == sorted(itertools.product(range(N), repeat=2))
        ar = _areas(lv)
        assert len(ar) == 2 * 4 ** l and np.all(ar > 0) and abs(ar.sum() - 1.0) < 1e-14
        if l > 0:
            np.testing.assert_array_equal(_p_apply(lv, f(mg["levels"][l - 1]["xyz"])), f(lv["xyz"]))
            # corner sons: one entry of weight 1; midnodes: 1/2 + 1/2
            cnt = np.diff(lv["p_ptr"])
            assert set(cnt.tolist()) == {1, 2}
            assert np.all(lv["p_w"][lv["p_ptr"][:-1][cnt == 1]] == 1.0)
    assert mg["levels"][0]["ids"].tolist() == [0, 1, 2, 3]


def test_encodings_agree(tmp_path):
    from ug_amd import mgio
    B = _square()
    B.refine(lambda k, t, P: 1)
    B.refine(lambda k, t, P: 1 if P[:, 0].max() <= 0.5 else 0)
    out = []
    for mode in (W.XDR, W.ASCII, W.BIN):
        p = tmp_path / f"m{mode}.ug.mg"
        B.write(p, mode)
        out.append(mgio.read(p))
    for a in out[1:]:
        for la, lb in zip(out[0]["levels"], a["levels"]):
            for k in la:
                np.testing.assert_array_equal(la[k], lb[k])


def _matches_builder(mg, B):
    """Per level: the positions and the elements (as sets of corner
    positions) the builder made."""
    for l, lv in enumerate(mg["levels"]):
        want = sorted(tuple(np.round(p, 12)) for p in B.pos[l].values())
        got = sorted(tuple(np.round(p, 12)) for p in lv["xyz"])
        assert got == want, l
        ids = sorted(B.pos[l])
        assert lv["ids"].tolist() == ids
        we = sorted(tuple(sorted(tuple(np.round(B.pos[l][c], 12)) for c in cs)) for _, cs, _ in B.elems[l])
        ge = sorted(tuple(sorted(tuple(np.round(lv["xyz"][c], 12))
                                 for c in lv["elem_corners"][lv["elem_ptr"][k]:lv["elem_ptr"][k + 1]]))
                    for k in range(len(lv["elem_tag"])))
        assert ge == we, l


def test_local_refinement_copies_and_green(tmp_path):
    """Level 1 red; level 2: red inside [0, 1/2]^2, a green bisection on one
    element, copies elsewhere (the copies carry their nodes up)."""
    from ug_amd import mgio
    B = _square()
    B.refine(lambda k, t, P: 1)
    state = {"green": False}

    def choose(k, t, P):
        if P[:, 0].max() <= 0.5 and P[:, 1].max() <= 0.5:
            return 1
        if not state["green"] and P[:, 0].min() >= 0.5:
            state["green"] = True
            return 2
        return 0

    B.refine(choose)
    path = tmp_path / "loc.ug.mg"
    B.write(path, W.XDR)
    mg = mgio.read(path)
    _matches_builder(mg, B)
    lv1, lv2 = mg["levels"][1], mg["levels"][2]
    ar = _areas(lv2)
    assert np.all(ar > 0) and abs(ar.sum() - 1.0) < 1e-14
    assert len(lv2["ids"]) > len(lv1["ids"]) and len(lv2["elem_tag"]) > len(lv1["elem_tag"])
    f = lambda X: 0.5 - X[:, 0] + 4.0 * X[:, 1]
    np.testing.assert_array_equal(_p_apply(lv2, f(lv1["xyz"])), f(lv2["xyz"]))


def test_moved_corner_weights_are_barycentric(tmp_path):
    from ug_amd import mgio
    B = _square()
    B.refine(lambda k, t, P: 1)
    # move the midnode of the first element's first edge off the edge
    rec = B.elems[0][0][2]
    mid_id = rec["ids"][3]
    rec["moved"] = [(mid_id, (0.45, 0.1))]
    path = tmp_path / "mv.ug.mg"
    B.write(path, W.BIN)
    mg = mgio.read(path)
    lv = mg["levels"][1]
    q = int(np.flatnonzero(lv["ids"] == mid_id)[0])
    np.testing.assert_allclose(lv["xyz"][q], [0.45, 0.1])
    row = slice(lv["p_ptr"][q], lv["p_ptr"][q + 1])
    assert lv["p_ptr"][q + 1] - lv["p_ptr"][q] == 3
    assert abs(lv["p_w"][row].sum() - 1.0) < 1e-14
    f = lambda X: 2.0 + X[:, 0] + 5.0 * X[:, 1]
    assert abs(_p_apply(lv, f(mg["levels"][0]["xyz"]))[q] - f(lv["xyz"][q:q + 1])[0]) < 1e-14


@pytest.mark.parametrize("mode", [W.XDR, W.ASCII])
def test_hexahedra_match_the_q1_generator(tmp_path, mode):
    from ug_amd import mgio, problems
    rule = hex_red_rule()
    B = Builder(3, {HEX: (8, HEX_E, HEX_S)}, {HEX: [rule]}, HEX_C, [(HEX, list(range(8)))])
    for _ in range(3):
        B.refine(lambda k, t, P: 0)
    path = tmp_path / "cube.ug.mg"
    B.write(path, mode)
    mg = mgio.read(path)
    Q = problems.jump3d(3)
    for l, lv in enumerate(mg["levels"]):
        N = 2 ** l + 1
        ijk = np.round(lv["xyz"] * (N - 1)).astype(np.int64)
        lex = ijk[:, 0] + N * (ijk[:, 1] + N * ijk[:, 2])
        assert np.array_equal(np.sort(lex), np.arange(N ** 3))
        assert len(lv["elem_tag"]) == 8 ** l and np.all(lv["elem_tag"] == HEX)
        if l == 0:
            continue
        Nc = 2 ** (l - 1) + 1
        ijc = np.round(mg["levels"][l - 1]["xyz"] * (Nc - 1)).astype(np.int64)
        lexc = ijc[:, 0] + Nc * (ijc[:, 1] + Nc * ijc[:, 2])
        Pm = np.zeros((N ** 3, Nc ** 3))
        rows = np.repeat(np.arange(len(lv["ids"])), np.diff(lv["p_ptr"]))
        np.add.at(Pm, (lex[rows], lexc[lv["p_idx"]]), lv["p_w"])
        G = Q.levels[l]
        Pg = G.dense_p(Q.levels[l - 1].n)
        np.testing.assert_array_equal(Pm, Pg)
        # entries in the father element's corner order: corner sons weight
        # 1, edge midnodes 1/2, side nodes 1/4, centres 1/8
        assert set(np.unique(lv["p_w"]).tolist()) == {1.0, 0.5, 0.25, 0.125}


@pytest.mark.parametrize("kw,msg", [({"nparfiles": 2}, "parallel"), ({"debug": 1}, "debug")])
def test_refusals(tmp_path, kw, msg):
    from ug_amd import mgio
    B = _square()
    B.refine(lambda k, t, P: 1)
    p = tmp_path / "bad.ug.mg"
    B.write(p, W.XDR, **kw)
    with pytest.raises(mgio.MgError, match=msg):
        mgio.read(p)


def test_damaged_files(tmp_path):
    from ug_amd import mgio
    B = _square()
    B.refine(lambda k, t, P: 1)
    p = tmp_path / "ok.ug.mg"
    B.write(p, W.BIN)
    data = p.read_bytes()
    (tmp_path / "cut.ug.mg").write_bytes(data[: len(data) - 7])
    with pytest.raises(mgio.MgError):
        mgio.read(tmp_path / "cut.ug.mg")
    (tmp_path / "title.ug.mg").write_bytes(data.replace(b"sparse.mg", b"sparse.xx", 1))
    with pytest.raises(mgio.MgError, match="title"):
        mgio.read(tmp_path / "title.ug.mg")
    with pytest.raises(mgio.MgError):
        mgio.read(tmp_path / "missing.ug.mg")
