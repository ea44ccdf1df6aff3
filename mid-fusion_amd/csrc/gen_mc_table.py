"""Generates mf_mc_table.h: the marching-cubes triangle table of this
library, built from the cube's geometry rather than typed in.

Corner and edge numbering follow meshing::compute_index / interp_vertexes
(lib/algorithms/meshing.hpp:21-112): corners c0..c7 at offsets
(0,0,0) (1,0,0) (1,0,1) (0,0,1) (0,1,0) (1,1,0) (1,1,1) (0,1,1); edges
e0..e11 = (c0,c1) (c1,c2) (c2,c3) (c0,c3) (c4,c5) (c5,c6) (c6,c7) (c4,c7)
(c0,c4) (c1,c5) (c2,c6) (c3,c7).

For every inside/outside pattern the iso-lines on the six faces are joined
into closed loops on the cube's surface: a face with two sign changes gets
one segment; a face with four (diagonal corners inside) gets two segments,
each cutting off one inside corner (the "inside corners separated" rule, so
neighbouring cubes agree on the shared face).  Every segment is oriented
with its inside corner on the left seen from outside the cube, which makes
the loops consistently wound: triangle normals (v1 - v0) x (v2 - v0) point
from inside (tsdf < 0) to outside.  Each loop is triangulated as a fan from
its lowest-numbered edge.

Run: python gen_mc_table.py > mf_mc_table.h
"""
import itertools

import numpy as np

CORNERS = [(0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)]
EDGES = [(0, 1), (1, 2), (2, 3), (0, 3), (4, 5), (5, 6), (6, 7), (4, 7), (0, 4), (1, 5), (2, 6), (3, 7)]


def faces():
    """Each face as its 4 corners in cyclic order, counter-clockwise seen from
    outside the cube."""
    out = []
    for axis in range(3):
        for side in (0, 1):
            cs = [i for i, c in enumerate(CORNERS) if c[axis] == side]
            centre = np.array([0.5, 0.5, 0.5])
            centre[axis] = side
            normal = np.zeros(3)
            normal[axis] = 1 if side else -1
            u = np.zeros(3)
            u[(axis + 1) % 3] = 1
            v = np.cross(normal, u)
            ang = [np.arctan2(np.dot(np.array(CORNERS[c]) - centre, v), np.dot(np.array(CORNERS[c]) - centre, u))
                   for c in cs]
            out.append([c for _, c in sorted(zip(ang, cs))])
    return out


def edge_of(a, b):
    for i, (p, q) in enumerate(EDGES):
        if {p, q} == {a, b}:
            return i
    raise ValueError((a, b))


def case(mask):
    inside = [(mask >> c) & 1 for c in range(8)]
    nxt = {}
    for f in faces():
        # segments around each inside corner of the face: from the edge before
        # it to the edge after it (counter-clockwise from outside) keeps the
        # inside corner on the left
        for k in range(4):
            c, prv, nx = f[k], f[k - 1], f[(k + 1) % 4]
            if not inside[c]:
                continue
            if inside[prv] and inside[nx]:
                continue
            if inside[prv]:
                # a run of inside corners: the segment spans from the run's
                # first crossing to its last; handled at the run's last corner
                continue
            # walk forward over the run of inside corners starting at c
            j = k
            while inside[f[(j + 1) % 4]]:
                j = (j + 1) % 4
            a = edge_of(f[(k - 1) % 4], c)
            b = edge_of(f[j], f[(j + 1) % 4])
            assert a not in nxt
            nxt[a] = b
    loops, seen = [], set()
    for s in sorted(nxt):
        if s in seen:
            continue
        loop, e = [], s
        while e not in seen:
            seen.add(e)
            loop.append(e)
            e = nxt[e]
        assert e == s, (mask, loop)
        loops.append(loop)
    tris = []
    for loop in loops:
        r = loop.index(min(loop))
        loop = loop[r:] + loop[:r]
        for i in range(1, len(loop) - 1):
            tris.append((loop[0], loop[i], loop[i + 1]))
    crossing = {i for i, (p, q) in enumerate(EDGES) if inside[p] != inside[q]}
    assert {e for t in tris for e in t} == crossing and sum(len(l) for l in loops) == len(crossing)
    return tris


def table():
    return [case(m) for m in range(256)]


if __name__ == '__main__':
    t = table()
    assert max(len(x) for x in t) <= 5
    print('// Generated by gen_mc_table.py (marching-cubes triangles from the cube')
    print('// geometry; see that script).  Do not edit.')
    print('#pragma once')
    print('#include <cstdint>')
    print('namespace mf {')
    print('// kMcTri[case]: up to 5 triangles as edge triples, -1 terminated.')
    print('static const int8_t kMcTri[256][16] = {')
    for m, tris in enumerate(t):
        flat = [e for tri in tris for e in tri] + [-1] * (16 - 3 * len(tris))
        print('  {%s},  // %d' % (', '.join('%d' % v for v in flat), m))
    print('};')
    print('}  // namespace mf')
