"""Meshers: a lattice of blocks, each a box of nodes, and how the blocks connect.

Mirrors ``peregrine/mesher/base.py:6-80`` and ``peregrine/mesher/cube.py:6-53``:
blocks are numbered i fastest over the lattice ``mbdims``; each block has
``blkdims`` nodes per axis; interfaces are numbered nface = 1..6 (i low, i high,
j low, j high, k low, k high); an interface either faces a neighbour block
(a periodic one carries the translation by the period) or is a boundary named
by ``bcnames`` or ``defaultbc``.
"""

import numpy as np


class Interface:
    """One of a block's six interfaces. Facing a neighbour it carries the orientation: digit m
    (of three) names the neighbour's axis our axis m runs along, 1 2 3 for +i +j +k and 4 5 6 for
    -i -j -k (``multiblock/interface/base.py:95-136``)."""

    def __init__(self, nface):
        self.nface = nface
        self.nbr = None
        self.orient = None
        self.periodic = False
        self.xlate = (0.0, 0.0, 0.0)
        self.bcname = None

    @property
    def axis(self):
        return (self.nface - 1) // 2

    @property
    def is_low(self):
        return self.nface % 2 == 1

    @staticmethod
    def signedaxis(code):
        n = int(code)
        return (n - 1) % 3, n > 3

    @staticmethod
    def orientcode(axis, flip):
        return str(axis + 1 + (3 if flip else 0))

    def connect(self, nbr, orient="123"):
        self.nbr, self.orient, self.bcname = nbr, orient, None

    @property
    def nbrnface(self):
        """The neighbour's interface this one meets: along the neighbour's axis our normal runs
        along, its high end for our low one unless that axis runs backwards."""
        axis, flip = self.signedaxis(self.orient[self.axis])
        return 2 * axis + (1 if self.is_low == flip else 2)

    @property
    def nbrorient(self):
        """The neighbour's orientation digits back to us (the inverse signed permutation)."""
        out = [None] * 3
        for ours, code in enumerate(self.orient):
            theirs, flip = self.signedaxis(code)
            out[theirs] = self.orientcode(ours, flip)
        if None in out:
            raise ValueError(f"orientation {self.orient!r} does not name three distinct axes")
        return "".join(out)


class BaseMesher:
    name = None

    def __init__(self, mbdims=(1, 1, 1), blkdims=(10, 10, 10), bcnames=None, defaultbc=None):
        self.mbdims = [int(n) for n in mbdims]
        self.blkdims = [int(n) for n in blkdims]
        self.bcnames = dict(bcnames or {})
        self.defaultbc = defaultbc

    @property
    def nblks(self):
        return int(np.prod(self.mbdims))

    @property
    def periodicaxes(self):
        return (False, False, False)

    def fill(self, mb):
        if mb.blocks:
            raise ValueError(f"a mesher fills an empty multiblock; this one has {len(mb.blocks)} blocks")
        for _ in range(self.nblks):
            mb.add_blk()
        mbd = self.mbdims
        for k in range(mbd[2]):
            for j in range(mbd[1]):
                for i in range(mbd[0]):
                    nblki = k * mbd[1] * mbd[0] + j * mbd[0] + i
                    blk = mb.blocks[nblki]
                    self.shape_blk(blk, i, j, k)
                    self.connect(blk, nblki, (i, j, k))

    def connect(self, blk, nblki, ijk):
        for axis, (n, index, wraps) in enumerate(zip(self.mbdims, ijk, self.periodicaxes)):
            stride = int(np.prod(self.mbdims[:axis]))
            for nface, end, step in ((2 * axis + 1, 0, -1), (2 * axis + 2, n - 1, 1)):
                face = blk.intface(nface)
                if index != end:
                    face.connect(nblki + step * stride)
                elif wraps:
                    face.periodic = True
                    face.connect(nblki - step * stride * (n - 1))
                    span = -self.lengths[axis] if face.is_low else self.lengths[axis]
                    face.xlate = tuple(span if m == axis else 0.0 for m in range(3))
                else:
                    face.bcname = self.bcnames.get(nface, self.defaultbc)


class CubeMesher(BaseMesher):
    name = "cube"

    def __init__(self, origin=(0, 0, 0), lengths=(1, 1, 1), periodic=(False, False, False), **kwargs):
        super().__init__(**kwargs)
        self.origin = [float(o) for o in origin]
        self.lengths = [float(L) for L in lengths]
        self.periodic = [bool(p) for p in periodic]
        # the whole lattice's nodes along each axis, sliced per block: a node on an interface has the
        # same coordinate in both blocks and in a single block spanning the lattice
        self.lines = [np.linspace(o, o + L, m * (n - 1) + 1)
                      for o, L, m, n in zip(self.origin, self.lengths, self.mbdims, self.blkdims)]

    @property
    def periodicaxes(self):
        return tuple(self.periodic)

    def shape_blk(self, blk, i, j, k):
        x, y, z = (line[m * (n - 1):(m + 1) * (n - 1) + 1]
                   for line, m, n in zip(self.lines, (i, j, k), self.blkdims))
        blk.set_extents(*self.blkdims)
        blk.interior_nodes = np.stack(np.meshgrid(x, y, z, indexing="ij"), axis=-1)
        # the whole lattice's period, for a periodic interface's halo nodes
        blk.period = list(self.lengths)
