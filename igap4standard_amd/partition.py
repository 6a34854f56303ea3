"""One rank's block of a partitioned patch, as the reference builds it for
`mpiexec -n P` (non-periodic meshes, p = 2, ndof = 3):

    set_mpi_part / nurbs_init   mesh/nurbs/nurbs_init.c:9-199
        parts per direction, each rank's elements, local knot window,
        owned basis functions (nbasis), boundary faces (boun), the send/recv
        neighbour lists and blocks, universal (natural) node numbers
    iga_setmpicomm              src/snesiga_init.c:134-339
        selfIdx (owned -> local ghosted numbering), sendIdx/sendPtr,
        recvIdx/recvPtr (the p-thick ghost layers from the higher neighbours)
    globalIdx                   src/snesiga_init.c:80-107
        local ghosted -> PETSc row: the owned rows of rank r are contiguous
        from the sum of the lower ranks' sizes; the ghosts' rows arrive from
        their owners through the send/recv blocks

`blocks(nelem_global, nproc)` returns one RankBlock per rank; the exchange of
globalIdx is simulated from the send/recv lists exactly as the reference's
MPI messages carry it.  A context made from a RankBlock (App(pb, rank=...))
assembles that rank's contributions in its local ghosted numbering.
"""
from dataclasses import dataclass, field

import numpy as np

PORDER = 2
NDOF = 3


@dataclass
class RankBlock:
    rank: int
    nproc: int
    npart: tuple
    ipart: tuple
    nelem: tuple            # local elements per direction
    displ: tuple            # first global element (= first global basis function) per direction
    nbasis: tuple           # owned basis functions per direction
    knots: list             # local knot windows (nelem + 5 each)
    boun: tuple             # global boundary faces this rank touches
    sendPart: list
    sendBlock: list
    recvPart: list
    recvBlock: list
    universalIdx: np.ndarray  # natural node number of each owned basis function (owned order)
    selfIdx: np.ndarray = field(default=None)
    sendIdx: np.ndarray = field(default=None)
    sendPtr: np.ndarray = field(default=None)
    recvIdx: np.ndarray = field(default=None)
    recvPtr: np.ndarray = field(default=None)
    globalIdx: np.ndarray = field(default=None)
    dof_displ: int = 0      # first PETSc row of this rank

    @property
    def nactive(self):
        """Basis functions per direction of the local (ghosted) window."""
        return tuple(n + PORDER for n in self.nelem)

    @property
    def nself(self):
        return NDOF * int(np.prod(self.nbasis))

    @property
    def nlocal(self):
        return NDOF * int(np.prod(self.nactive))


def _round(x):
    # C's round(): halves away from zero
    return int(np.floor(x + 0.5)) if x >= 0 else -int(np.floor(-x + 0.5))


def parts(nelem_global, nproc):
    """npart of set_mpi_part (nurbs_init.c:92-100)."""
    ng = nelem_global
    nepp = _round((float(ng[0] * ng[1] * ng[2]) / float(nproc)) ** (1.0 / 3.0))
    n0 = 1
    for i in range(ng[0] // nepp + 1, 0, -1):
        if nproc % i == 0:
            n0 = i
            break
    nepp = _round((float(ng[1] * ng[2]) / float(nproc // n0)) ** 0.5)
    n1 = 1
    for i in range(ng[1] // nepp + 1, 0, -1):
        if (nproc // n0) % i == 0:
            n1 = i
            break
    return (n0, n1, nproc // n0 // n1)


def _neighbours(npart, ipart, sign):
    """send (sign -1) / recv (sign +1) partner lists and block ids in the
    reference's order (nurbs_init.c:152-190): blocks 0..6 = z, y, x, yz,
    xz, xy, xyz."""
    n0, n1, n2 = npart
    i0, i1, i2 = ipart
    if sign < 0:
        c0, c1, c2 = i0 != 0, i1 != 0, i2 != 0
    else:
        c0, c1, c2 = i0 != n0 - 1, i1 != n1 - 1, i2 != n2 - 1
    combos = [(0, 0, 1), (0, 1, 0), (1, 0, 0), (0, 1, 1), (1, 0, 1), (1, 1, 0), (1, 1, 1)]
    conds = [c2, c1, c0, c1 and c2, c0 and c2, c0 and c1, c0 and c1 and c2]
    part, block = [], []
    for b, (dx, dy, dz) in enumerate(combos):
        if not conds[b]:
            continue
        j0, j1, j2 = (i0 + sign * dx) % n0, (i1 + sign * dy) % n1, (i2 + sign * dz) % n2
        part.append(n1 * n2 * j0 + n2 * j1 + j2)
        block.append(b)
    return part, block


def _block_ranges(blk, nbasis, ghost):
    """Basis ranges per direction of send block blk (ghost=False: the first
    p layers of the owned range) or recv block blk (ghost=True: the p layers
    past it), src/snesiga_init.c:180-334."""
    ext = {0: (0, 0, 1), 1: (0, 1, 0), 2: (1, 0, 0), 3: (0, 1, 1), 4: (1, 0, 1), 5: (1, 1, 0), 6: (1, 1, 1)}[blk]
    rng = []
    for d in range(3):
        if not ext[d]:
            rng.append(range(0, nbasis[d]))
        elif ghost:
            rng.append(range(nbasis[d], nbasis[d] + PORDER))
        else:
            rng.append(range(0, PORDER))
    return rng


def _local_dofs(ranges, nactive):
    out = []
    for x in ranges[0]:
        for y in ranges[1]:
            for z in ranges[2]:
                ib = (nactive[1] * nactive[2]) * x + nactive[2] * y + z
                out.extend(NDOF * ib + i for i in range(NDOF))
    return np.array(out, dtype=np.int64)


def blocks(nelem_global, nproc, knots_global=None):
    """Every rank's block for an nelem_global mesh on nproc ranks."""
    ng = tuple(int(n) for n in nelem_global)
    if knots_global is None:
        from .problem import uniform_knots
        knots_global = [uniform_knots(n) for n in ng]
    npart = parts(ng, nproc)
    n1p, n2p = npart[1], npart[2]
    # elements per part and displacements per direction (nurbs_init.c:109-128)
    nel, dsp = [], []
    for d in range(3):
        w = [ng[d] // npart[d]] * npart[d]
        for i in range(ng[d] % npart[d]):
            w[i] += 1
        nel.append(w)
        dsp.append([int(sum(w[:i])) for i in range(npart[d])])
    nbg = [n + PORDER for n in ng]
    out = []
    for r in range(nproc):
        ip = (r // (n1p * n2p), (r % (n1p * n2p)) // n2p, r % n2p)
        ne = tuple(nel[d][ip[d]] for d in range(3))
        di = tuple(dsp[d][ip[d]] for d in range(3))
        nb = tuple(ne[d] + PORDER if ip[d] == npart[d] - 1 else ne[d] for d in range(3))
        kn = [np.asarray(knots_global[d][di[d]:di[d] + ne[d] + 5], dtype=np.float64) for d in range(3)]
        boun = []
        for d in range(3):
            if ip[d] == 0:
                boun.append(2 * d)
            if ip[d] == npart[d] - 1:
                boun.append(2 * d + 1)
        sp, sb = _neighbours(npart, ip, -1)
        rp, rb = _neighbours(npart, ip, +1)
        uni = np.array([nbg[1] * nbg[2] * (di[0] + x) + nbg[2] * (di[1] + y) + (di[2] + z)
                        for x in range(nb[0]) for y in range(nb[1]) for z in range(nb[2])], dtype=np.int64)
        out.append(RankBlock(rank=r, nproc=nproc, npart=npart, ipart=ip, nelem=ne, displ=di, nbasis=nb, knots=kn,
                             boun=tuple(boun), sendPart=sp, sendBlock=sb, recvPart=rp, recvBlock=rb, universalIdx=uni))
    # iga_setmpicomm (src/snesiga_init.c:134-339)
    for b in out:
        na = b.nactive
        b.selfIdx = _local_dofs([range(0, n) for n in b.nbasis], na)
        idx, ptr = [], [0]
        for blk in b.sendBlock:
            idx.append(_local_dofs(_block_ranges(blk, b.nbasis, False), na))
            ptr.append(ptr[-1] + len(idx[-1]))
        b.sendIdx = np.concatenate(idx) if idx else np.zeros(0, np.int64)
        b.sendPtr = np.array(ptr, dtype=np.int64)
        idx, ptr = [], [0]
        for blk in b.recvBlock:
            idx.append(_local_dofs(_block_ranges(blk, b.nbasis, True), na))
            ptr.append(ptr[-1] + len(idx[-1]))
        b.recvIdx = np.concatenate(idx) if idx else np.zeros(0, np.int64)
        b.recvPtr = np.array(ptr, dtype=np.int64)
    # globalIdx (src/snesiga_init.c:80-107): owned rows, then the exchange
    disp = 0
    for b in out:
        b.dof_displ = disp
        disp += b.nself
        g = np.full(b.nlocal, -1, dtype=np.int64)
        g[b.selfIdx] = b.dof_displ + np.arange(b.nself)
        b.globalIdx = g
    for b in out:                                        # receiver
        for j, src in enumerate(b.recvPart):
            s = out[src]
            k = s.sendPart.index(b.rank)                 # the sender's message to this rank
            buf = s.globalIdx[s.sendIdx[s.sendPtr[k]:s.sendPtr[k + 1]]]
            dst = b.recvIdx[b.recvPtr[j]:b.recvPtr[j + 1]]
            if len(buf) != len(dst):
                raise RuntimeError("rank %d: block from %d has %d entries, %d expected" % (b.rank, src, len(buf),
                                                                                              len(dst)))
            b.globalIdx[dst] = buf
    return out


def local_node_natural(b, nelem_global):
    """Natural (universal) node number of every local ghosted basis function
    of block b (x slowest)."""
    nbg = [n + PORDER for n in nelem_global]
    na = b.nactive
    x, y, z = np.meshgrid(*[np.arange(n) for n in na], indexing="ij")
    return ((nbg[1] * nbg[2]) * (b.displ[0] + x) + nbg[2] * (b.displ[1] + y) + (b.displ[2] + z)).reshape(-1)


def petsc_to_natural(bl, nelem_global):
    """PETSc row -> natural dof for the whole partition."""
    n = sum(b.nself for b in bl)
    m = np.empty(n, dtype=np.int64)
    for b in bl:
        nat = NDOF * np.repeat(b.universalIdx, NDOF) + np.tile(np.arange(NDOF), len(b.universalIdx))
        m[b.dof_displ:b.dof_displ + b.nself] = nat
    return m
