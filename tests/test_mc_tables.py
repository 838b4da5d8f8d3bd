"""Marching-cubes case table (vhm_mc_tables.h) pins.

The triangle table must be the reference's (marching_cubes_tables.h:44-301)
entry for entry; the edge-mask table is derived from it at compile time and
must equal the reference's (marching_cubes_tables.h:9-42).  Independent
invariants: complementary cases cut the same edges, every triangle's edges
are cut, case 0/255 are empty, triangle counts stay in [0, 5], and every
triangle uses three distinct edges.
"""
import numpy as np
import pytest

import oracle_lib as O


def _product_tables():
    import re
    from pathlib import Path
    src = (Path(__file__).resolve().parents[1] / "voxelhashmt_amd" / "csrc" / "vhm_mc_tables.h").read_text()
    body = src[src.index("kMcCaseRows[256] = {"):src.index("};", src.index("kMcCaseRows[256] = {"))]
    rows = re.findall(r'"([0-9a-f]*)"', body)
    assert len(rows) == 256
    tri = np.full((256, 16), -1, np.int32)
    edge = np.zeros(256, np.int32)
    for i, r in enumerate(rows):
        for j, ch in enumerate(r):
            tri[i, j] = int(ch, 16)
            edge[i] |= 1 << int(ch, 16)
    return tri, edge


def test_invariants():
    tri, edge = _product_tables()
    assert edge[0] == 0 and edge[255] == 0
    for i in range(256):
        assert edge[i] == edge[255 - i]
        n = int((tri[i] >= 0).sum())
        assert n % 3 == 0 and 0 <= n // 3 <= 5
        assert (tri[i][n:] == -1).all()
        for t in range(n // 3):
            e = tri[i][3 * t: 3 * t + 3]
            assert len(set(e.tolist())) == 3
        # a cut edge joins corners on opposite sides of the iso level
        corners = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
        for e in range(12):
            a, b = corners[e]
            cut = ((i >> a) & 1) != ((i >> b) & 1)
            assert bool(edge[i] & (1 << e)) == cut, (i, e)
    # total triangle count over all cases (classic table: 2460 edges -> 820 triangles? check stable sum)
    assert int((tri >= 0).sum()) % 3 == 0


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU build (oracle/_ref) not present")
def test_equals_reference_tables():
    import ctypes as C
    e = np.zeros(256, np.int32)
    t = np.zeros(4096, np.int32)
    O.ref().vref_mc_tables(e.ctypes.data, t.ctypes.data)
    tri, edge = _product_tables()
    assert np.array_equal(edge, e)
    assert np.array_equal(tri, t.reshape(256, 16))
