"""Meshers: a lattice of blocks, each a box of nodes, and how the blocks connect.

Mirrors ``peregrine/mesher/base.py:6-80``, ``peregrine/mesher/cube.py:6-53`` and
``peregrine/mesher/annulus.py:6-107``: blocks are numbered i fastest over the lattice
``mbdims``; each block has ``blkdims`` nodes per axis; interfaces are numbered
nface = 1..6 (i low, i high, j low, j high, k low, k high); an interface either faces a
neighbour block (a periodic one carries the translation by the period, or the rotation
of a swept sector) or is a boundary named by ``bcnames`` or ``defaultbc``.
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
        # a rotational periodic's turn of what arrives through it (vectors, and the halo nodes)
        self.rot = None
        self.bcname = None

    def set_periodic(self, rot=None, xlate=(0.0, 0.0, 0.0)):
        self.periodic = True
        self.xlate = tuple(float(x) for x in xlate)
        self.rot = None if rot is None else np.array(rot, np.float64).reshape(3, 3)

    @property
    def rotperiodic(self):
        return self.rot is not None and not np.allclose(self.rot, np.eye(3))

    @staticmethod
    def rotabout(axis, deg):
        """The rotation by deg degrees about axis (Rodrigues)."""
        u = np.array(axis, np.float64)
        u = u / np.linalg.norm(u)
        th = np.radians(deg)
        c, s = np.cos(th), np.sin(th)
        x, y, z = u
        K = np.array([[0.0, -z, y], [z, 0.0, -x], [-y, x, 0.0]])
        return c * np.eye(3) + s * K + (1.0 - c) * np.outer(u, u)

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
                    face.connect(nblki - step * stride * (n - 1))
                    self.wrap(face, axis)
                else:
                    face.bcname = self.bcnames.get(nface, self.defaultbc)

    def wrap(self, face, axis):
        """What a periodic interface along axis does to what crosses it: a translation by the
        lattice's length."""
        span = -self.lengths[axis] if face.is_low else self.lengths[axis]
        face.set_periodic(xlate=tuple(span if m == axis else 0.0 for m in range(3)))


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


class AnnulusMesher(BaseMesher):
    """A sector of an annulus: i along the axis from p1 toward p2, j outward from the inner radius
    |p3 - p1| by thickness, k around the axis through sweep degrees. Periodic in k: a full turn
    joins its ends as an interior interface; a sector's ends are rotational periodics, each
    turning what arrives by the sweep back (``mesher/annulus.py:6-107``)."""

    name = "annulus"

    def __init__(self, p1=(0, 0, 0), p2=(1, 0, 0), p3=(0, 1, 0), sweep=45, thickness=0.1, periodic=False, **kwargs):
        super().__init__(**kwargs)
        self.p1, self.p2, self.p3 = (np.array(p, np.float64) for p in (p1, p2, p3))
        self.sweep, self.thickness, self.periodic = float(sweep), float(thickness), bool(periodic)
        length, rin = np.linalg.norm(self.p2 - self.p1), np.linalg.norm(self.p3 - self.p1)
        if length == 0.0 or rin == 0.0:
            raise ValueError("the annulus needs p2 off p1 (the axis) and p3 off the axis (the inner radius)")
        self.axis, self.radial = (self.p2 - self.p1) / length, (self.p3 - self.p1) / rin
        if abs(np.dot(self.axis, self.radial)) > 1e-7:
            raise ValueError("the line p1-p3 is not orthogonal to the axis p1-p2")
        if not 0.0 < self.sweep <= 360.0:
            raise ValueError(f"sweep must be in (0, 360] degrees, not {sweep}")
        if self.periodic and self.sweep != 360.0:
            # a rotational periodic is a rotation about the origin
            if np.linalg.norm(self.p1 - np.dot(self.p1, self.axis) * self.axis) > 1e-7:
                raise ValueError("a rotationally periodic sector must be swept about an axis through the origin")
        self.edges = [np.linspace(0.0, length, self.mbdims[0] + 1),
                      np.linspace(rin, rin + self.thickness, self.mbdims[1] + 1),
                      np.linspace(0.0, self.sweep, self.mbdims[2] + 1)]

    @property
    def periodicaxes(self):
        return (False, False, self.periodic)

    def _turned(self, theta):
        n, p = self.axis, self.radial
        c, s = np.cos(theta)[:, None], np.sin(theta)[:, None]
        return c * p + s * np.cross(n, p) + (1.0 - c) * np.dot(n, p) * n

    def shape_blk(self, blk, i, j, k):
        x, r, th = (np.linspace(e[m], e[m + 1], n) for e, m, n in zip(self.edges, (i, j, k), self.blkdims))
        radial = self._turned(np.radians(th))
        blk.set_extents(*self.blkdims)
        blk.interior_nodes = (self.p1 + x[:, None, None, None] * self.axis
                              + r[None, :, None, None] * radial[None, None, :, :])
        blk.period = [0.0, 0.0, 0.0]

    def wrap(self, face, axis):
        if self.sweep == 360.0:
            face.periodic = False
            return
        # the low end takes its halo from the far end of the sweep, turned back by it
        face.set_periodic(rot=face.rotabout(self.axis, -self.sweep if face.is_low else self.sweep))


class BatchMesher(BaseMesher):
    """Independent cases side by side: block n is its own periodic box of dims[n] cells on
    lengths[n] (or one length for all), joined to nothing else, so one solver steps them together
    (the uneven multiblock batch: blocks of different sizes in one step)."""

    name = "batch"

    def __init__(self, dims, lengths=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), **kwargs):
        self.dims = [tuple(int(n) for n in d) for d in dims]
        super().__init__(mbdims=(len(self.dims), 1, 1), **kwargs)
        L = np.asarray(lengths, float)
        self.lengths_of = [tuple(L) if L.ndim == 1 else tuple(L[n]) for n in range(len(self.dims))]
        self.origin = np.asarray(origin, float)

    def fill(self, mb):
        if mb.blocks:
            raise ValueError(f"a mesher fills an empty multiblock; this one has {len(mb.blocks)} blocks")
        for n, (dims, lengths) in enumerate(zip(self.dims, self.lengths_of)):
            blk = mb.add_blk()
            lines = [np.linspace(o, o + L, m + 1) for o, L, m in zip(self.origin, lengths, dims)]
            blk.set_extents(*(m + 1 for m in dims))
            blk.interior_nodes = np.stack(np.meshgrid(*lines, indexing="ij"), axis=-1)
            blk.period = list(lengths)
            for axis in range(3):
                for nface in (2 * axis + 1, 2 * axis + 2):
                    face = blk.intface(nface)
                    face.connect(n)
                    span = -lengths[axis] if face.is_low else lengths[axis]
                    face.set_periodic(xlate=tuple(span if m == axis else 0.0 for m in range(3)))
