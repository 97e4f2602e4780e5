"""Union-mesh layout of gospl_amd.ensemble (CPU): members at workgroup-
aligned offsets, neighbours shifted into their member's range, padding
isolated, ordering ids a permutation and tie-break ids the members' own."""

import numpy as np

from gospl_amd import _lib, workloads
from gospl_amd.ensemble import BLOCK, union_mesh


def test_union_mesh_layout():
    lib = _lib.load()
    ms = [workloads.sphere_case(lib, level=lv, seed=s, lmax=8)[0] for lv, s in ((2, 1), (3, 2), (2, 3))]
    u, off, seg, tgid = union_mesh(ms)
    assert off[0] == 0 and off[-1] == u.N
    assert np.all(off % BLOCK == 0)
    assert sorted(u.gid.tolist()) == list(range(u.N))
    for k, m in enumerate(ms):
        o, n = int(off[k]), m.N
        assert int(off[k + 1]) - o == -(-n // BLOCK) * BLOCK
        assert np.all(seg[o:o + n] == k) and np.all(seg[o + n:off[k + 1]] == -1)
        nb = u.nbr[:m.width, o:o + n]
        assert np.array_equal(np.where(nb >= 0, nb - o, -1), m.nbr)
        assert np.all(u.nbr[m.width:, o:o + n] == -1)
        assert np.array_equal(tgid[o:o + n], m.gid)
        assert np.array_equal(u.gid[o:o + n] - o, m.gid)
        # padding vertices have no neighbours and nobody points at them
        assert np.all(u.deg[o + n:off[k + 1]] == 0)
    pad = seg < 0
    assert not np.isin(np.nonzero(pad)[0], u.nbr[u.nbr >= 0]).any()
