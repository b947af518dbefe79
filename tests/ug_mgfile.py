"""Test helper: writes UG multigrid files the way SaveMultiGrid does
(gm/ugio.cc:1283-1690 over gm/mgio.cc's Write_* functions and low/bio.cc's
three encodings), for hierarchies built here by refinement rules, so the
reader (ug_amd.mgio) can be checked against known geometry and transfer.

A hierarchy is given as element types (corners, edges, sides), rules
(pattern over the new-corner slots, sons as NodeList indices: corners,
then edge midnodes, then side nodes, then the centre at nCorner +
centre index), coarse points and elements, and per coarse element a tree
of refinement records {rule, ids: new corner node IDs, moved: [(id, pos)],
sons: [record or None per son]}.
"""
from __future__ import annotations

import struct

TITLE = "####.sparse.mg.storage.format.####"
XDR, ASCII, BIN = 0, 1, 2
TAGS = 8
MAX_CORNERS, MAX_EDGES, MAX_SIDES = 8, 12, 6


class Bio:
    """low/bio.cc's writers; n_byte counts as Bio_Jump_From/_To do."""

    def __init__(self, f):
        self.f, self.mode, self.n_byte = f, ASCII, 0

    def _put(self, b: bytes):
        self.f.write(b)
        self.n_byte += len(b)

    def ints(self, v):
        for x in v:
            if self.mode == ASCII:
                self._put(b"%d\n" % x)
            elif self.mode == BIN:
                self._put(struct.pack("=i", x))
            else:
                self._put(struct.pack(">I", x & 0xFFFFFFFF))

    def doubles(self, v):
        for x in v:
            if self.mode == ASCII:
                self._put(("%g\n" % x).encode())
            elif self.mode == BIN:
                self._put(struct.pack("=d", x))
            else:
                self._put(struct.pack(">d", x))

    def string(self, s: str):
        b = s.encode()
        self._put((b"%d\n" if self.mode == ASCII else b"%d ") % len(b))
        self._put(b + (b"\n" if self.mode == XDR else b" "))

    def jump_from(self):
        self.n_byte = 0
        self._jpos = self.f.tell()
        self.f.write(b" %20d " % 0)

    def jump_to(self):
        here = self.f.tell()
        self.f.seek(self._jpos)
        self.f.write(b" %20d " % self.n_byte)
        self.f.seek(here)


def ectrl(rf, nf, nm, rc, on=0):
    return (nf & 31) | ((nm & 31) << 5) | ((rf & ((1 << 18) - 1)) << 10) | ((rc & 7) << 28) | ((on & 1) << 31)


def write(path, mode, dim, ge, rules, points, nbnd, elements, records, nlevels, counts=(0, 0, 0),
          version="UG_IO_2.5", nparfiles=1, debug=0, names=("---", "dom", "mg", "fmt")):
    """ge: {tag: (ncorners, edges [(a, b)], sides [(a, b, c, d)])};
    rules: [(rclass, pattern, [(son_tag, corners)])]; points: coarse
    positions by vertex ID (boundary vertices first, nbnd of them);
    elements: [(tag, corner IDs)]; records: one tree per coarse element
    (None: not refined)."""
    max_new = 5 if dim == 2 else 19
    with open(path, "wb") as f:
        io = Bio(f)
        io.string(TITLE)                                  # Write_MG_General, head in ASCII
        io.ints([mode])
        io.mode = mode
        io.string(version)
        for s in names[:1] + names[1:]:
            io.string(s)
        io.ints([dim, 4711, 1024, nlevels, counts[0], counts[1], counts[2], 0, 0, nparfiles, debug])
        io.ints([TAGS])                                   # Write_GE_General
        for t in range(TAGS):                             # Write_GE_Elements
            nc, edges, sides = ge.get(t, (0, [], []))
            io.ints([t, nc, len(edges), len(sides)])
            if edges or sides:
                v = []
                for a, b in edges:
                    v += [a, b]
                for sd in sides:
                    v += list(sd) + [-1] * (4 - len(sd))
                io.ints(v)
        io.ints([len(rules)] + [0] * TAGS)                # Write_RR_General (offsets unused here)
        for rclass, pattern, sons in rules:               # Write_RR_Rules
            io.ints([rclass, len(sons)])
            v = list(pattern) + [0] * (max_new - len(pattern))
            v += [0, 0] * max_new                         # sonandnode
            for tag, corners in sons:
                v += [tag] + list(corners) + [-1] * (MAX_CORNERS - len(corners))
                v += [-1] * MAX_SIDES + [0]               # nb, path
            io.ints(v)
        ne = len(elements)
        io.ints([len(points), nbnd, len(points) - nbnd, ne, 0, ne])   # Write_CG_General
        for p in points:                                  # Write_CG_Points
            io.doubles(list(p)[:dim])

        def count(rec):
            return 0 if rec is None else 1 + sum(count(s) for s in rec["sons"])

        for (tag, corners), rec in zip(elements, records):   # Write_CG_Elements
            io.ints([tag])
            nsides = len(ge[tag][2])
            io.ints([count(rec)] + list(corners) + [-1] * nsides + [0, 0])
        io.jump_from()                                    # boundary section
        io.ints([nbnd])                                   # Write_BD_General
        for i in range(nbnd):                             # std_domain BNDP_SaveBndP
            io.ints([0, 1])
            io.doubles([0.25 * i] * (dim - 1))
        io.jump_to()

        def rec_out(rec):                                 # Write_Refinement, depth first
            moved = rec.get("moved", [])
            sonref = sum(1 << s for s, r in enumerate(rec["sons"]) if r is not None)
            io.ints([ectrl(rec["rule"] + 1, len(rec["ids"]), len(moved), rec.get("rclass", 1)), sonref]
                    + list(rec["ids"]) + [m[0] for m in moved])
            if moved:
                io.doubles([x for m in moved for x in list(m[1])[:dim]])
            for r in rec["sons"]:
                if r is not None:
                    rec_out(r)

        for rec in records:
            if rec is not None:
                rec_out(rec)
