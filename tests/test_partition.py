"""One rank's block of a partitioned patch (igap4standard_amd/partition.py)
against the reference's own partitioner: every rank's Nurbs fields for the
meshes and rank counts of tests/golden/partition.json (written by running
mesh/nurbs/nurbs_init.c under mpiexec), and the index maps of
src/snesiga_init.c:80-339 checked for what they must satisfy ("fake ranks"):
owned rows tile the PETSc numbering, the local ghosted window is owned plus
received, and every ghost receives the PETSc row of the same natural dof from
its owner through the send/recv blocks."""
import json
import os

import numpy as np
import pytest

from igap4standard_amd import partition as PT

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "partition.json")
CASES = json.load(open(GOLD))


@pytest.mark.parametrize("case", range(len(CASES)), ids=lambda k: "%s_p%d" % (CASES[k]["nelem"], CASES[k]["nproc"]))
def test_partition_matches_reference(case):
    c = CASES[case]
    bl = PT.blocks(c["nelem"], c["nproc"])
    assert len(bl) == len(c["ranks"])
    for b, r in zip(bl, c["ranks"]):
        assert r["rank"] == b.rank
        assert list(b.nelem) == r["nelem"] and list(b.nbasis) == r["nbasis"]
        assert [len(k) for k in b.knots] == r["nknot"]
        for d in range(3):
            assert np.array_equal(b.knots[d], np.array(r["knot%d" % d]))
        assert list(b.boun) == r["boun"]
        assert b.sendPart == r["sendPart"] and b.sendBlock == r["sendBlock"]
        assert b.recvPart == r["recvPart"] and b.recvBlock == r["recvBlock"]
        assert np.array_equal(b.universalIdx, np.array(r["universalIdx"]))


@pytest.mark.parametrize("case", range(len(CASES)), ids=lambda k: "%s_p%d" % (CASES[k]["nelem"], CASES[k]["nproc"]))
def test_index_maps(case):
    c = CASES[case]
    ng = c["nelem"]
    bl = PT.blocks(ng, c["nproc"])
    ndof = 3 * int(np.prod([n + 2 for n in ng]))
    p2n = PT.petsc_to_natural(bl, ng)
    assert len(p2n) == ndof and np.array_equal(np.sort(p2n), np.arange(ndof))   # owned rows tile the mesh
    n2p = np.empty(ndof, np.int64)
    n2p[p2n] = np.arange(ndof)
    for b in bl:
        assert b.nlocal == b.nself + len(b.recvIdx)
        loc = np.concatenate([b.selfIdx, b.recvIdx])
        assert np.array_equal(np.sort(loc), np.arange(b.nlocal))                # owned + received = the window
        nat = 3 * np.repeat(PT.local_node_natural(b, ng), 3) + np.tile(np.arange(3), b.nlocal // 3)
        assert np.array_equal(b.globalIdx, n2p[nat])                            # ghosts carry their owners' rows
        # the blocks a rank sends are the ones its partners receive, in the same order
        for j, src in enumerate(b.recvPart):
            s = bl[src]
            k = s.sendPart.index(b.rank)
            assert s.sendBlock[k] == b.recvBlock[j]


def test_block_layout_checked_at_create():
    """igap4h_create accepts a rank's block only with the iga_setmpicomm
    layout (checked before any device work, so this runs without a GPU)."""
    from igap4standard_amd import _lib as L
    from igap4standard_amd import problem as P
    from igap4standard_amd.snesiga import App
    pb = P.stretch_problem((6, 5, 7), a=0.01)
    b = PT.blocks(pb.nelem, 8)[0]
    b.selfIdx = b.selfIdx.copy()
    b.selfIdx[[3, 4]] = b.selfIdx[[4, 3]]
    with pytest.raises(L.IgapError) as e:
        App(pb, rank=b)
    assert e.value.code == L.IGAP4H_EINVAL and "selfIdx" in str(e.value)
    b = PT.blocks(pb.nelem, 8)[0]
    b.recvIdx = b.recvIdx.copy()
    b.recvIdx[0] = b.selfIdx[0]
    with pytest.raises(L.IgapError) as e:
        App(pb, rank=b)
    assert e.value.code == L.IGAP4H_EINVAL and "recvIdx" in str(e.value)
