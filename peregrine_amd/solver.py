"""The multiblock solver: blocks, their arrays, and the slots of an RK stage.

Mirrors ``peregrine/multiblock/solver.py:23-391``. A stage is
  prerhs   halo exchange of Q, the EOS over the halos, the chemistry's source into dQ (coupled or
           sub-stepped)
  rhs      the advective flux through every face, its divergence into dQ
  advance  the RK combine into the stage's output register
  postrhs  the EOS over the interior of the output register
(``simulator/advec/simulator.py:57-87``, ``integrators/rungekutta.py:33-39``),
each slot a sequence of launches of the HIP library on one stream.
"""

import ctypes
import gc

import numpy as np

from peregrine_amd import grid, native
from peregrine_amd.backend import HostScalar
from peregrine_amd.integrators import get_integrator
from peregrine_amd.mesher import Interface
from peregrine_amd.simulator import NativeCase, get_simulator



class Block:
    def __init__(self, nblki, mb):
        self.nblki = nblki
        self.mb = mb
        self.faces = {n: Interface(n) for n in range(1, 7)}
        self.interior_nodes = None
        self.dijk = None
        self.substeps = None
        self.sweeps = None

    def intface(self, nface):
        return self.faces[nface]

    def set_extents(self, ni, nj, nk):
        self.ni, self.nj, self.nk = int(ni), int(nj), int(nk)

    @property
    def extents(self):
        return (self.ni, self.nj, self.nk)

    @property
    def ncell(self):
        return (self.ni - 1, self.nj - 1, self.nk - 1)

    @property
    def ncells(self):
        return int(np.prod(self.ncell))

    @property
    def ng(self):
        return self.mb.ng

    @property
    def cellshape(self):
        return tuple(n + 2 * self.ng for n in self.ncell)

    def faceshape(self, axis):
        s = list(self.cellshape)
        s[axis] += 1
        return tuple(s)

    @property
    def interior(self):
        ng = self.ng
        return np.s_[ng:-ng, ng:-ng, ng:-ng]

    def interior_box(self):
        return native.box((self.ng,) * 3, self.ncell)

    def face_box(self, axis):
        ext = list(self.ncell)
        ext[axis] += 1
        return native.box((self.ng,) * 3, ext)

    def halo_boxes(self, depth=None):
        """Each interface's halo slab, depth (ng) deep next to the interior, over the interior of
        the other axes."""
        depth = self.ng if depth is None else depth
        boxes = []
        for nface, f in self.faces.items():
            if f.nbr is None:
                continue
            start, ext = [self.ng] * 3, list(self.ncell)
            start[f.axis] = self.ng - depth if f.is_low else self.ng + self.ncell[f.axis]
            ext[f.axis] = depth
            boxes.append(native.box(start, ext))
        return boxes

    def alloc(self, backend, sim, nregs):
        ng, cs = self.ng, self.cellshape
        ne = sim.ne
        # every array's rows start a cache line at the first interior index, i = ng
        def alloc(shape, name):
            return backend.alloc(shape, name=name, lead=ng)

        self.regs = [alloc(cs + (ne,), f"Q{r}") for r in range(nregs)]
        self.state = alloc(cs + (sim.nstate,), "state")
        self.iF, self.jF, self.kF = (alloc(self.faceshape(a) + (ne,), f"{'ijk'[a]}F") for a in range(3))
        self.iS, self.jS, self.kS = (alloc(self.faceshape(a) + (3,), f"{'ijk'[a]}S") for a in range(3))
        self.detJinv = alloc(cs + (1,), "detJinv")
        self.xcc = alloc(cs + (3,), "xcc")
        self.q = alloc(cs + (ne,), "q")
        self.viscous = sim.viscous
        if sim.viscous:
            ns = sim.ns
            # navstokes/simulator.py:30-36, advecdiff/simulator.py:12-15: the transport state, the
            # species enthalpies, the gradients; the metrics the gradients and viscous flux read
            self.trstate = alloc(cs + (2 + ns,), "trstate")
            self.hsp = alloc(cs + (ns,), "hsp")
            self.grads = alloc(cs + (3 * sim.ngrad,), "grads")
            self.Jinv = alloc(cs + (9,), "Jinv")
            self.xf = [alloc(self.faceshape(a) + (3,), f"{'ijk'[a]}xc") for a in range(3)]

    @property
    def Q(self):
        return self.regs[0]

    def make_metrics(self, nodes):
        """The metrics of the block's nodes with their halo layers."""
        ng = self.ng
        self.nodes = nodes
        self.xcc.set(grid.cellcenters(nodes))
        for a, S in enumerate((self.iS, self.jS, self.kS)):
            S.set(grid.areavecs(nodes, a))
        self.detJinv.set(grid.detJinv(nodes, ng)[..., None])
        if self.viscous:
            with np.errstate(divide="ignore", invalid="ignore"):
                Jinv = grid.Jinv(nodes)
            self.Jinv.set(Jinv.reshape(self.cellshape + (9,)))
            for a, xf in enumerate(self.xf):
                xf.set(grid.facecenters(nodes, a))


def block_nodes(blocks, ng):
    """Every block's nodes with ng halo layers (``multiblock/grid.py``, ``gen_halo`` then the node
    exchange): linear extrapolation first, a translated periodic with itself moved by its period;
    then across every interface the neighbour's node layers placed through the orientation, turned
    by a rotational periodic and moved by a translation. Corner and edge halos stay extrapolated."""
    nodes = {}
    for blk in blocks:
        lows = [blk.faces[2 * a + 1] for a in range(3)]
        moved = [f.periodic and f.nbr == blk.nblki and f.rot is None for f in lows]
        nodes[blk.nblki] = grid.halo_nodes(blk.interior_nodes, ng, moved, blk.period)
    for blk in blocks:
        for f in blk.faces.values():
            if f.nbr is not None:
                place_halo_nodes(nodes[blk.nblki], blk, f, blocks[f.nbr], ng)
    return nodes


def place_halo_nodes(nodes, blk, f, nbr, ng):
    """The halo node layers across interface f from nbr's interior nodes: our node x (interior-
    relative; the layers n_a + 1 .. n_a + ng, or -1 .. -ng, over the plane's nodes 0 .. n_m) is the
    neighbour's node y, y_A = x_m (or N_A - x_m backwards) in the plane, and across it the
    neighbour's nodes counted inward from the face it meets."""
    a, n = f.axis, blk.ncell
    N = nbr.ncell
    ranges = [np.arange(n[m] + 1) for m in range(3)]
    ranges[a] = np.arange(-ng, 0) if f.is_low else np.arange(n[a] + 1, n[a] + ng + 1)
    x = np.meshgrid(*ranges, indexing="ij")
    y = [None] * 3
    for m in range(3):
        A, back = f.signedaxis(f.orient[m])
        if m != a:
            y[A] = N[A] - x[m] if back else x[m]
        elif f.is_low:
            y[A] = -x[m] if back else N[A] + x[m]
        else:
            y[A] = N[A] - (x[m] - n[a]) if back else x[m] - n[a]
    values = nbr.interior_nodes[y[0], y[1], y[2]]
    if f.rot is not None:
        values = values @ f.rot.T
    values = values + np.array(f.xlate)
    nodes[ng + x[0], ng + x[1], ng + x[2]] = values


def placement(blk, f, nbr, depth):
    """(dst_start, src_start, extent, orientation digits) of the halo slab `depth` deep across
    interface f of blk from nbr's interior (``exchange/collection.py:55-96``): the slab's first
    cell x (interior-relative: -depth or n along the normal, 0 along the plane) is the neighbour's
    cell y, along each of its axes A_m = digit m's axis: y = x_m, or n_A - 1 - x_m when it runs
    backwards, in the plane; and across the interface the neighbour's cells next to the face it
    meets, counted outward from it as ours are. The device steps from there along the digits."""
    ng = blk.ng
    a, n, N = f.axis, blk.ncell, nbr.ncell
    ext = list(n)
    ext[a] = depth
    dst = [ng] * 3
    dst[a] = ng - depth if f.is_low else ng + n[a]
    x = [0, 0, 0]
    x[a] = -depth if f.is_low else n[a]
    src = [0, 0, 0]
    for m in range(3):
        A, back = f.signedaxis(f.orient[m])
        if m != a:
            y = N[A] - 1 - x[m] if back else x[m]
        elif f.is_low:
            y = -1 - x[m] if back else x[m] + N[A]
        else:
            y = N[A] - 1 - (x[m] - n[a]) if back else x[m] - n[a]
        src[A] = ng + y
    return dst, src, ext, [int(c) for c in f.orient]


def turn_of(f):
    """The row-major turn a rotational periodic gives what crosses it, or None."""
    return native.dbls(f.rot.ravel()) if f.rotperiodic else None


def stage_slabs(nk, n):
    """The k-slabs (k0, k1, f0, f1) a fused stage runs a block of nk cells along k in, n of them
    (at most nk): the stage of cells k0 <= k < k1 follows the fluxes it reads and writes Q and the
    state in place, so a slab also takes every k face whose stencil (cells k - 2 to k + 1) reaches
    one of its cells before they change: f1 = k1 + 1 (nk at the last), and the next slab's k faces
    start at f1 + 1. Every face is computed once."""
    n = max(1, min(n, nk))
    slabs, f0 = [], 0
    for c in range(n):
        k0, k1 = nk * c // n, nk * (c + 1) // n
        f1 = min(k1 + 1, nk)
        slabs.append((k0, k1, f0, f1))
        f0 = f1 + 1
    return slabs


class Solver:
    def __init__(self, cfg, mesh, backend, sim, integ, state=None):
        self.cfg, self.mesh, self.backend, self.sim, self.integ = cfg, mesh, backend, sim, integ
        self.nrt = 0
        # register -> the derivative register its prerhs began (None: halos ready, no derivative)
        self.readied = {}
        self._cflwork = None
        self.tme = HostScalar(0.0)
        # the step scalar is nothing until a step sets it (multiblock/solver.py:36, allocation.py:80-85):
        # a right-hand side before any step judges no equilibrium exit over no time
        self.dt = backend.scalar("dt", 0.0)
        self.ng = sim.ng
        self.blocks = []
        mesh.fill(self)
        self._check_bcs()
        for blk in self.blocks:
            blk.alloc(backend, sim, integ.nregs)
        nodes = block_nodes(self.blocks, self.ng)
        for blk in self.blocks:
            blk.make_metrics(nodes[blk.nblki])
        self.case = NativeCase(sim, backend)
        self.fused = self._fused_supported()
        if self.fused:
            for blk in self.blocks:
                rshape = blk.cellshape + (max(sim.ns - 1, 1),)
                blk.rates = backend.alloc(rshape, name="rates", lead=self.ng) if sim.chemistry else None
        self.batched = self._batch_supported()
        self.batch = None
        self.tiles = None
        self.plugins = self._plugins()
        self._set_state(state)
        for p in self.plugins:
            p.start(self)

    # -- construction -------------------------------------------------------
    def add_blk(self):
        blk = Block(len(self.blocks), self)
        self.blocks.append(blk)
        return blk

    @property
    def nblocks(self):
        return len(self.blocks)

    # the boundary kinds this build has (simulator/euler/bcs.py, simulator/navstokes/bcs.py)
    # the wall kinds each simulator has (simulator/euler/bcs.py:15-27, navstokes/bcs.py:27-53) and
    # the values each reads from its bcvals section
    BC_KINDS = {"euler": ("slp-adia-wall", "slp-isot-wall"), "navier-stokes": native.WALL_KINDS}
    BC_VALUES = {"slp-isot-wall": ("T",), "no-slp-isot-wall": ("T",), "mov-adia-wall": ("u", "v", "w"),
                 "mov-isot-wall": ("u", "v", "w", "T")}

    def _wall(self, where, bcname, sect):
        kind = sect.get("kind")
        kinds = self.BC_KINDS[self.sim.name]
        if kind not in kinds:
            raise ValueError(f"{where}: bcvals.{bcname} is of kind {kind!r}; this build's {self.sim.name} has {kinds}")
        need = self.BC_VALUES.get(kind, ())
        missing = [v for v in need if v not in sect]
        if missing:
            raise ValueError(f"{where}: bcvals.{bcname} ({kind}) needs {missing}")
        w = native.Wall()
        w.kind = native.WALL_KINDS.index(kind)
        for d, name in enumerate("uvw"):
            w.u[d] = float(sect.get(name, 0.0)) if name in need else 0.0
        w.T = float(sect["T"]) if "T" in need else 0.0
        return w

    def _check_bcs(self):
        """Every boundary named and its name in bcvals with a kind this build has
        (multiblock/solver.py:158-173)."""
        for blk in self.blocks:
            blk.walls = []
            for f in blk.faces.values():
                if f.nbr is not None:
                    continue
                where = f"block {blk.nblki} interface {f.nface}"
                if f.bcname is None:
                    raise ValueError(f"{where} is a boundary the grid did not name; name it (a mesher's bcnames or "
                                     "defaultbc) and say what it is in bcvals")
                if f.bcname not in self.cfg["bcvals"]:
                    raise ValueError(f"{where} is named {f.bcname!r}, which bcvals does not have; it has "
                                     f"{sorted(self.cfg['bcvals'])}")
                wall = self._wall(where, f.bcname, dict(self.cfg["bcvals"][f.bcname]))
                blk.walls.append((f.axis, 0 if f.is_low else 1, wall))
            # a neighbour's interface meets this one back, with the inverse orientation, over a plane
            # of the same cells
            for f in blk.faces.values():
                if f.nbr is None:
                    continue
                where = f"block {blk.nblki} interface {f.nface}"
                if not 0 <= f.nbr < len(self.blocks):
                    raise ValueError(f"{where} faces block {f.nbr}, which the grid does not have")
                nbr = self.blocks[f.nbr]
                back = nbr.faces[f.nbrnface]
                if back.nbr != blk.nblki or back.orient != f.nbrorient:
                    raise ValueError(f"{where} (orientation {f.orient}) meets block {nbr.nblki} interface "
                                     f"{f.nbrnface}, which faces block {back.nbr} with orientation {back.orient}, "
                                     f"not this one with {f.nbrorient}")
                for m in range(3):
                    A = f.signedaxis(f.orient[m])[0]
                    if m != f.axis and blk.ncell[m] != nbr.ncell[A]:
                        raise ValueError(f"{where}: blocks {blk.nblki} and {nbr.nblki} do not match across it "
                                         f"({blk.ncell} and {nbr.ncell} cells, orientation {f.orient})")
            # an exchanged interface copies ng interior planes; a wall reflects however thin the block
            for f in blk.faces.values():
                if f.nbr is not None and blk.ncell[f.axis] < self.ng:
                    raise ValueError(f"block {blk.nblki}: {blk.ncell} cells, thinner along axis {f.axis} than the "
                                     f"{self.ng}-deep halo its interface {f.nface} exchanges")

    def _plugins(self):
        from peregrine_amd.plugins import claim

        plugins, self.sponges = claim(self.cfg["plugins"], self)
        return plugins

    def _batch_supported(self):
        """The fused stage's phases run as one launch over every block (pga_batch_*) when the case has
        more than one block, each an inviscid block with no walls and no rotational periodic
        (``backend-mi355x.batched``: auto, true or false)."""
        want = self.cfg["backend-mi355x"]["batched"]
        if want is False or want == "false" or not self.fused or len(self.blocks) < 2 or self.integ.name == "rk4":
            return False
        if self.sim.viscous or int(self.cfg["backend-mi355x"]["stage-slabs"]) != 1:
            return False
        for blk in self.blocks:
            if blk.walls or any(f.nbr is not None and f.rotperiodic for f in blk.faces.values()):
                return False
        return True

    def _fused_supported(self):
        """The fused stage runs when the config asks for it, the integrator takes one combine per
        stage, and the library has the case's variant compiled."""
        sect = self.cfg["backend-mi355x"]
        if not sect["fused-stage"] or self.integ.name not in ("rk1", "rk2", "tvd-rk3", "rk4"):
            return False
        return bool(self.case.L.pga_stage_supported(self.case.handle))

    @property
    def regs(self):
        return tuple(range(self.integ.nregs))

    @property
    def ncells(self):
        return sum(b.ncells for b in self.blocks)

    # -- state ---------------------------------------------------------------
    def _set_state(self, state):
        """The case's initial values everywhere, then a restart's primitives over the interior
        with its step and time (``run.py``, ``readers/restart.py:30-54``)."""
        values = self.sim.ics()
        for blk in self.blocks:
            q = np.empty(blk.cellshape + (self.sim.ne,))
            q[...] = values
            blk.q.set(q)
        if state is not None:
            state.fill(self)
            self.nrt = state.nrt
            self.tme.set(state.tme)
        self.from_prims()

    def set_prims(self, prims):
        for blk, values in zip(self.blocks, prims):
            blk.q.set(values)
        self.from_prims()

    def from_prims(self):
        for blk in self.blocks:
            full = native.box((0, 0, 0), blk.cellshape)
            self._call("pga_fromprims", self.case.handle, blk.q.view(), blk.Q.view(), blk.state.view(), full)
        self.readied.clear()
        self.postrhs(0)
        # the next step skips its first halo pass: it is not the steady step a graph holds
        self._fresh = True

    # -- launches --------------------------------------------------------------
    timers = None

    def time_kernels(self, on=True):
        """Bracket every launch with device events on the launch stream, by entry point."""
        self.timers = {} if on else None

    def _call(self, name, *args):
        """Launch one entry point of the library on the backend's stream; raise on failure."""
        L = self.case.L
        fn = getattr(L, name)
        if self.timers is None:
            native.check(fn(*args, self.backend.stream), name, L)
            return
        torch = self.backend.torch
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        native.check(fn(*args, self.backend.stream), name, L)
        e1.record()
        self.timers.setdefault(name, []).append((e0, e1))

    def kernel_times(self):
        """{entry point: (launches, mean ms)} of the launches timed so far."""
        self.synchronize()
        return {k: (len(v), sum(a.elapsed_time(b) for a, b in v) / len(v)) for k, v in (self.timers or {}).items()}

    # -- the slots -----------------------------------------------------------
    def exchange(self, reg):
        """Each interface's halo slab of Q from its partner's interior planes (place, per interface),
        the momentum turned across a rotational periodic."""
        self._exchange(lambda b: b.regs[reg], self.sim.ne, self.ng, 1, 0)

    def _exchange(self, array, ncomp, depth, vec, rows):
        """The halo slabs of `array` across every interface, a block's slabs in one launch."""
        eye = np.eye(3).ravel()
        for blk in self.blocks:
            slabs = []
            for f in blk.faces.values():
                if f.nbr is None:
                    continue
                nbr = self.blocks[f.nbr]
                dst, src, ext, orient = placement(blk, f, nbr, depth)
                rot = f.rot.ravel() if f.rotperiodic else eye
                slabs.append((array(blk).view(), array(nbr).view(), dst, src, ext, orient, rot))
            if not slabs:
                continue
            n = len(slabs)
            dv = (native.View * n)(*[t[0] for t in slabs])
            sv = (native.View * n)(*[t[1] for t in slabs])
            dst, src, ext, orient = (native.ints([v for t in slabs for v in t[m]]) for m in (2, 3, 4, 5))
            rots = native.dbls([v for t in slabs for v in t[6]])
            self._call("pga_copy_slabs", n, dv, sv, dst, src, ext, ncomp, orient, rots, vec, rows)

    def transport(self, Q):
        """The transport state and species enthalpies over the interior and the first halo layer
        ("shallow", launch/ranges.py:52-53), after the halo EOS (advec/simulator.py:68-75)."""
        for blk in self.blocks:
            boxes = [blk.interior_box()] + blk.halo_boxes(1)
            arr = (native.Box * len(boxes))(*boxes)
            self._call("pga_transport", self.case.handle, blk.regs[Q].view(), blk.state.view(), blk.trstate.view(),
                       blk.hsp.view(), arr, len(boxes))
            for sponge in self.sponges:
                self._call("pga_viscsponge", blk.xcc.view(), blk.trstate.view(), sponge.line, arr, len(boxes))

    def gradients(self, Q):
        """The cell gradients over the interior and their one-plane exchange
        (advecdiff/simulator.py:20-27)."""
        for blk in self.blocks:
            self._call("pga_grads", self.case.handle, blk.regs[Q].view(), blk.state.view(), blk.Jinv.view(),
                       blk.grads.view(), blk.interior_box())
        # the rows u, v, w, T, Y: each a gradient, the first three a velocity gradient
        self._exchange(lambda b: b.grads, 3 * self.sim.ngrad, 1, 0, 1)

    def transport_gradients(self, Q):
        """transport() and gradients() for the fused step: a small block's two kernels in one launch
        (pga_transport_grads, the same values), since it pays more for a launch than for its work; a
        block of more than 2^16 cells in two (one launch holds the registers of the heavier kernel
        for both: 0.8% slower per step on the 256 x 128 x 64 layer)."""
        for blk in self.blocks:
            boxes = [blk.interior_box()] + blk.halo_boxes(1)
            arr = (native.Box * len(boxes))(*boxes)
            Qv, st = blk.regs[Q].view(), blk.state.view()
            if blk.ncells <= 1 << 16:
                self._call("pga_transport_grads", self.case.handle, Qv, st, blk.trstate.view(), blk.hsp.view(), arr,
                           len(boxes), blk.Jinv.view(), blk.grads.view(), blk.interior_box())
            else:
                self._call("pga_transport", self.case.handle, Qv, st, blk.trstate.view(), blk.hsp.view(), arr,
                           len(boxes))
                self._call("pga_grads", self.case.handle, Qv, st, blk.Jinv.view(), blk.grads.view(),
                           blk.interior_box())
            for sponge in self.sponges:
                self._call("pga_viscsponge", blk.xcc.view(), blk.trstate.view(), sponge.line, arr, len(boxes))
        self._exchange(lambda b: b.grads, 3 * self.sim.ngrad, 1, 0, 1)

    def fluxes(self, blk, Qv, st, slab=None, axes=(0, 1, 2), one_launch=False):
        """The advective flux of each axis, with the viscous flux added for a viscous case, then the
        walls' own flux through their faces (the postflux hook). A slab (k0, k1, f0, f1) limits the
        i and j faces to the cells k0 <= k < k1 and the k faces to f0 <= k <= f1. one_launch: every
        axis's faces, advective and viscous, in one launch (pga_faces, the same bits), else a launch
        per axis and kind as the reference sequences them."""
        sides = ((blk.iS, blk.iF), (blk.jS, blk.jF), (blk.kS, blk.kF))
        boxes = []
        for a in range(3):
            box = blk.face_box(a)
            if slab is not None:
                k0, k1, f0, f1 = slab
                lo, hi = (f0, f1 + 1) if a == 2 else (k0, k1)
                start, ext = list(box.start), list(box.extent)
                start[2], ext[2] = self.ng + lo, max(0, hi - lo)
                box = native.box(start, ext)
            boxes.append(box)
        if one_launch and tuple(axes) == (0, 1, 2):
            S = (native.View * 3)(*[sd[0].view() for sd in sides])
            F = (native.View * 3)(*[sd[1].view() for sd in sides])
            B = (native.Box * 3)(*boxes)
            if self.sim.viscous:
                xf = (native.View * 3)(*[blk.xf[a].view() for a in range(3)])
                self._call("pga_faces", self.case.handle, Qv, st, S, F, B, 1, self.sim.alpha, blk.trstate.view(),
                           blk.hsp.view(), blk.grads.view(), blk.xcc.view(), xf)
            else:
                self._call("pga_faces", self.case.handle, Qv, st, S, F, B, 0, 0.0, native.View(), native.View(),
                           native.View(), native.View(), None)
        else:
            for a, (S, F) in enumerate(sides):
                if a not in axes:
                    continue
                self._call("pga_flux", self.case.handle, a, Qv, st, S.view(), F.view(), boxes[a])
                if self.sim.viscous:
                    self._call("pga_diffflux", self.case.handle, a, self.sim.alpha, Qv, st, blk.trstate.view(),
                               blk.hsp.view(), blk.grads.view(), blk.xcc.view(), blk.xf[a].view(), S.view(),
                               F.view(), boxes[a])
        for axis, side, wall in blk.walls:
            S, F = (blk.iS, blk.jS, blk.kS)[axis], (blk.iF, blk.jF, blk.kF)[axis]
            visc = self.sim.viscous
            self._call("pga_wall_flux", self.case.handle, axis, side, ctypes.byref(wall), native.ints(blk.ncell),
                       self.ng, Qv, st, blk.trstate.view() if visc else native.View(), blk.xcc.view(),
                       blk.xf[axis].view() if visc else native.View(), S.view(), F.view())

    def wall_halos(self, Q):
        """Each wall's halo as the interior's mirror image (the preadvflux hook)."""
        for blk in self.blocks:
            for axis, side, wall in blk.walls:
                S = (blk.iS, blk.jS, blk.kS)[axis]
                self._call("pga_wall_mirror", self.case.handle, axis, side, ctypes.byref(wall), native.ints(blk.ncell),
                           self.ng, blk.regs[Q].view(), blk.state.view(), S.view())

    def prerhs(self, Q, dQ=None):
        self.exchange(Q)
        for blk in self.blocks:
            boxes = blk.halo_boxes()
            arr = (native.Box * len(boxes))(*boxes)
            self._call("pga_fromQ", self.case.handle, blk.regs[Q].view(), blk.state.view(), arr, len(boxes))
        self.wall_halos(Q)
        if dQ is not None:
            self.begin_dQ(Q, dQ)
        if self.sim.viscous:
            self.transport(Q)
        self.readied[Q] = dQ

    def begin_dQ(self, Q, dQ):
        """The derivative's first term, the chemistry's source, before the fluxes append to it."""
        if self.sim.chemistry:
            for blk in self.blocks:
                self.chemistry_source(blk, blk.regs[Q].view(), blk.regs[dQ].view(), True)
        self.readied[Q] = dQ

    def chemistry_source(self, blk, Qv, out, begin_dQ):
        """The chemistry's term of the derivative over the interior: the coupled source (the
        production rates at the stage's state), the sub-stepped mean rate over the step
        (``chemint/coupled.py:10-13``, ``chemint/substepped.py:52-68``) or the implicit form's mean
        rate (linearly implicit Euler, ``pga_implicit``, in the sub-stepped form's slot); into dQ,
        rows 0..4 zero, when begin_dQ, else into the rates array the fused stage adds."""
        box = blk.interior_box()
        if self.sim.chemkind == "coupled":
            if begin_dQ:
                self._call("pga_source", self.case.handle, Qv, blk.state.view(), out, box)
            else:
                self._call("pga_rates", self.case.handle, Qv, blk.state.view(), out, box)
            return
        torch = self.backend.torch
        if self.sim.chemkind == "implicit":
            opts = self.sim.implicit
            if blk.sweeps is None:
                blk.imopts = native.ImplicitOpts(opts["parts"], opts["sweeps"], opts["tol"])
                blk.sweeps = torch.zeros(blk.ncells, dtype=torch.int32, device=self.backend.device)
                n = int(self.case.L.pga_implicit_worksize(self.case.handle, box))
                blk.imwork = torch.zeros(max(n, 1), dtype=torch.float64, device=self.backend.device)
            self._call("pga_implicit", self.case.handle, ctypes.byref(blk.imopts), Qv, blk.state.view(), out,
                       int(begin_dQ), blk.sweeps.data_ptr(), blk.ncells, blk.imwork.data_ptr(), blk.imwork.numel(),
                       self.dt.ptr, box)
            return
        opts = self.sim.substep
        if blk.substeps is None:
            o = native.SubstepOpts(opts["cap"], opts["bulk"], opts["bins"], 0, opts["eqtol"])
            blk.subopts = o
            blk.substeps = torch.zeros(blk.ncells, dtype=torch.int32, device=self.backend.device)
            n = int(self.case.L.pga_substep_listsize(ctypes.byref(o), box)) if opts["bulk"] < opts["cap"] else 0
            blk.stifflist = torch.zeros(max(n, 1), dtype=torch.int32, device=self.backend.device)
            blk.nlist = n
        self._call("pga_substep", self.case.handle, ctypes.byref(blk.subopts), Qv, blk.state.view(), out,
                   int(begin_dQ), blk.substeps.data_ptr(), blk.ncells, blk.stifflist.data_ptr(), blk.nlist,
                   self.dt.ptr, box)

    def _ready(self, Q, dQ):
        """Halos of Q exchanged and their EOS taken once per solution: a register whose halos are
        ready (after from_prims or a postrhs with its prerhs) only begins its derivative
        (peregrine/multiblock/solver.py:251-268). A second halo EOS from the first's T would move
        T's last bits, and a block's halo would no longer hold its partner cells' state."""
        standing = self.readied.get(Q, False)
        if standing == dQ:
            return
        if standing is None:
            self.begin_dQ(Q, dQ)
        else:
            self.prerhs(Q, dQ)

    def _spent(self, dQ):
        self.readied = {q: d for q, d in self.readied.items() if d != dQ}

    def rhs(self, *, Q, dQ):
        self._ready(Q, dQ)
        if self.sim.viscous:
            self.gradients(Q)
        for blk in self.blocks:
            Qv, st = blk.regs[Q].view(), blk.state.view()
            self.fluxes(blk, Qv, st)
            self._call("pga_applyflux", self.sim.ne, blk.regs[dQ].view(), blk.iF.view(), blk.jF.view(), blk.kF.view(),
                       blk.detJinv.view(), int(self.sim.chemistry), blk.interior_box())
        self._spent(dQ)

    def advance(self, n):
        Qout = self.integ.stageregs(self.regs)[n].Qout
        for c in self.integ.stagekerns(n, self.regs):
            for blk in self.blocks:
                X = (native.View * max(len(c.X), 1))(*[blk.regs[r].view() for r in c.X])
                w = (ctypes.c_double * max(len(c.w), 1))(*c.w)
                self._call("pga_combine", self.sim.ne, blk.regs[c.out].view(), int(c.own), c.wout, len(c.X), X, w,
                           self.dt.ptr, int(c.step), blk.interior_box())
        self.readied.pop(Qout, None)
        self.postrhs(Qout, prerhs=False)

    def postrhs(self, Q=0, prerhs=True):
        for blk in self.blocks:
            b = blk.interior_box()
            self._call("pga_fromQ", self.case.handle, blk.regs[Q].view(), blk.state.view(), (native.Box * 1)(b), 1)
        if prerhs:
            self.prerhs(Q)

    def stage(self, n, dt=None):
        regs = self.integ.stageregs(self.regs)[n]
        self.rhs(Q=regs.Q, dQ=regs.dQ)
        self.advance(n)

    # -- the fused stage ------------------------------------------------------
    def _halo_pass(self, qin):
        """Each block's halo slabs from its partners' interior planes, and the EOS over them."""
        for blk in self.blocks:
            dsts, srcs, ds, ss, es, os, rots = [], [], [], [], [], [], []
            for f in blk.faces.values():
                if f.nbr is None:
                    continue
                nbr = self.blocks[f.nbr]
                dst, src, ext, orient = placement(blk, f, nbr, self.ng)
                dsts.append(blk.regs[qin].view())
                srcs.append(nbr.regs[qin].view())
                ds += dst
                ss += src
                es += ext
                os += orient
                rots += list(f.rot.ravel()) if f.rotperiodic else [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
            n = len(dsts)
            turns = native.dbls(rots) if any(f.rotperiodic for f in blk.faces.values() if f.nbr is not None) else None
            self._call("pga_halo", self.case.handle, n, (native.View * n)(*dsts), (native.View * n)(*srcs),
                       native.ints(ds), native.ints(ss), native.ints(es), native.ints(os), turns, blk.state.view(),
                       blk.state.view())

    def _combine_terms(self, n):
        """The stage's combine as (source, weight, times dt) terms in the reference's order: its
        registers read as the step's start (Q0), the stage's input (Qin) or its derivative. An rk4
        stage has two combines, [first, second]: the first's own register reads as 3, and the second
        reads the first's output (the running total) as the value the first just took, 4."""
        kerns = self.integ.stagekerns(n, self.regs)
        dQ = self.regs[-1]
        first = kerns[0].out if len(kerns) == 2 else None

        def src(r, own=False):
            if own:
                return 3
            if r == first:
                return 4
            return 2 if r == dQ else (1 if r == 0 else 0)

        out = []
        for m, c in enumerate(kerns):
            terms = [(src(c.out, own=m == 0 and len(kerns) == 2), c.wout, 0)] if c.own else []
            for k, (r, w) in enumerate(zip(c.X, c.w)):
                terms.append((src(r), w, int(c.step and k == len(c.X) - 1)))
            out.append(terms)
        return out

    _streams = None

    def _block_streams(self):
        """Each block with the stream its stage work goes on. Launched eagerly (no graph replay),
        several blocks take side streams (default priority) in turn, forked from the current
        stream after what it has queued (the halo pass, which reads other blocks) and joined back
        once every block is queued, so small blocks' sweeps and stages overlap on the device; a
        block's stage reads and writes only its own arrays. Under graph replay the blocks stay on
        the one stream: forking side streams inside the captured graph crashed its replay."""
        import contextlib

        torch = self.backend.torch
        nside = int(self.cfg["backend-mi355x"]["block-streams"])
        if len(self.blocks) < 2 or nside < 2 or self.cfg["backend-mi355x"]["graphs"]:
            for blk in self.blocks:
                yield blk, contextlib.nullcontext()
            return
        if self._streams is None:
            self._streams = [torch.cuda.Stream(device=self.backend.device) for _ in range(nside)]
        main = torch.cuda.current_stream(self.backend.device)
        used = self._streams[:min(nside, len(self.blocks))]
        for s in used:
            s.wait_stream(main)
        for n, blk in enumerate(self.blocks):
            yield blk, torch.cuda.stream(used[n % len(used)])
        for s in used:
            main.wait_stream(s)

    class _Batch:
        """A table of blocks on the device (pga_batch_create) and the buffer it lives in."""

        def __init__(self, L, handle, table, keep):
            self.L, self.handle, self.table, self.keep = L, handle, table, keep
            self.imwork = None

        def close(self):
            if self.handle:
                self.L.pga_batch_destroy(self.handle)
                self.handle = None

        def __del__(self):
            self.close()

    def _make_batch(self, tiles, regs=None, states=None):
        """The table of every block: its registers (those named by regs, in order, else all), state
        (or the given stand-ins), faces, 1/J, rates, interior and halo slabs, and the tile shapes of its
        size class (tiles: {class: {"stage": (wi, kc), "fluxk": ki}})."""
        from peregrine_amd import autotune

        torch = self.backend.torch
        regs = tuple(range(self.integ.nregs)) if regs is None else tuple(regs)
        recs = (native.BatchBlock * len(self.blocks))()
        nslab = 0
        for n, blk in enumerate(self.blocks):
            r = recs[n]
            for m, q in enumerate(regs):
                r.Q[m] = blk.regs[q].view()
            r.state = (states[n] if states is not None else blk.state).view()
            r.iS, r.jS, r.kS = blk.iS.view(), blk.jS.view(), blk.kS.view()
            r.iF, r.jF, r.kF = blk.iF.view(), blk.jF.view(), blk.kF.view()
            r.detJinv = blk.detJinv.view()
            r.rates = blk.rates.view() if blk.rates is not None else native.View()
            r.interior = blk.interior_box()
            t = 0
            for f in blk.faces.values():
                if f.nbr is None:
                    continue
                dst, src, ext, orient = placement(blk, f, self.blocks[f.nbr], self.ng)
                r.slab_src[t] = f.nbr
                for d in range(3):
                    r.slab_dst[t][d], r.slab_from[t][d] = dst[d], src[d]
                    r.slab_extent[t][d], r.slab_orient[t][d] = ext[d], orient[d]
                t += 1
            r.nslab = t
            nslab += t
            choice = (tiles or {}).get(autotune.size_class(blk.ncell[0]), autotune.DEFAULT)
            r.stage_wi, r.stage_kc = choice["stage"]
            r.flux_ki = choice["fluxk"]
        L = self.case.L
        nbytes = int(L.pga_batch_bytes(len(self.blocks), nslab))
        table = torch.zeros(nbytes, dtype=torch.uint8, device=self.backend.device)
        handle = ctypes.c_void_p()
        native.check(L.pga_batch_create(self.case.handle, len(self.blocks), recs, len(regs), table.data_ptr(), nbytes,
                                        ctypes.byref(handle)), "pga_batch_create", L)
        return self._Batch(L, handle, table, recs)

    def _batched_step(self):
        """The fused step with each phase one launch over every block: the halo pass (unless from_prims
        left the halos ready), the chemistry rates, the three flux sweeps, and the stage with its EOS."""
        if self.batch is None:
            if self.tiles is None:
                from peregrine_amd import autotune

                self.tiles = autotune.tune(self) if self.cfg["backend-mi355x"]["autotune"] else {}
            self.batch = self._make_batch(self.tiles)
        h = self.batch.handle
        for n, regs in enumerate(self.integ.stageregs(self.regs)):
            if regs.Q not in self.readied:
                self._call("pga_batch_halo", h, regs.Q)
            self.readied.clear()
            (terms,) = self._combine_terms(n)
            src = native.ints([t[0] for t in terms])
            w = (ctypes.c_double * len(terms))(*[t[1] for t in terms])
            withdt = native.ints([t[2] for t in terms])
            if self.sim.chemistry:
                if self.sim.chemkind == "coupled":
                    self._call("pga_batch_rates", h, regs.Q)
                elif self.sim.chemkind == "implicit":
                    self._batch_implicit(regs.Q)
                else:
                    for blk in self.blocks:
                        self.chemistry_source(blk, blk.regs[regs.Q].view(), blk.rates.view(), False)
            for a in range(3):
                self._call("pga_batch_flux", h, regs.Q, a)
            self._call("pga_batch_stage", h, regs.Q, regs.Qout, len(terms), src, w, withdt, self.dt.ptr)

    def _batch_implicit(self, q):
        """Implicit chemistry over every block in one launch per phase, into each block's rates."""
        torch = self.backend.torch
        L = self.case.L
        if getattr(self.batch, "imwork", None) is None:
            opts = self.sim.implicit
            self.batch.imopts = native.ImplicitOpts(opts["parts"], opts["sweeps"], opts["tol"])
            n = int(L.pga_batch_implicit_worksize(self.batch.handle))
            self.batch.imwork = torch.zeros(max(n, 1), dtype=torch.float64, device=self.backend.device)
            cells = int(sum(-(-blk.ncells // 256) for blk in self.blocks)) * 256
            self.batch.sweeps = torch.zeros(cells, dtype=torch.int32, device=self.backend.device)
        b = self.batch
        self._call("pga_batch_implicit", b.handle, ctypes.byref(b.imopts), q, b.sweeps.data_ptr(), b.sweeps.numel(),
                   b.imwork.data_ptr(), b.imwork.numel(), self.dt.ptr)

    def _march(self, blk, regs, first):
        """The k-planes a thread of the march walks for this block and stage, or 0 for the split
        k sweep: the march takes an inviscid block without walls (whose hooks change the k faces'
        fluxes), a stage with one combine whose output is not its input, and one slab."""
        n = int(self.cfg["backend-mi355x"]["k-march"])
        if n < 1 or first is not None or regs.Qout == regs.Q or self.sim.viscous or blk.walls:
            return 0
        if int(self.cfg["backend-mi355x"]["stage-slabs"]) != 1:
            return 0
        return n

    def _slabs(self, blk):
        return stage_slabs(blk.ncell[2], int(self.cfg["backend-mi355x"]["stage-slabs"]))

    def _fused_step(self):
        """A step by stages of five launches per block: the halo pass (exchange and halo EOS, unless
        from_prims left the halos ready), the chemistry rates, the three flux sweeps, and the fused
        divergence + source + combine + EOS, in place on the reference's stage registers."""
        kc = int(self.cfg["backend-mi355x"]["k-chunk"])
        for n, regs in enumerate(self.integ.stageregs(self.regs)):
            if regs.Q not in self.readied:
                self._halo_pass(regs.Q)
                self.wall_halos(regs.Q)
                if self.sim.viscous:
                    self.transport_gradients(regs.Q)
            elif self.sim.viscous:
                self.gradients(regs.Q)
            self.readied.clear()
            combines = self._combine_terms(n)
            terms = combines[-1]
            src = native.ints([t[0] for t in terms])
            w = (ctypes.c_double * len(terms))(*[t[1] for t in terms])
            withdt = native.ints([t[2] for t in terms])
            if len(combines) == 2:
                tA = combines[0]
                first = (len(tA), native.ints([t[0] for t in tA]), (ctypes.c_double * len(tA))(*[t[1] for t in tA]),
                         native.ints([t[2] for t in tA]), self.integ.stagekerns(n, self.regs)[0].out)
            else:
                first = None
            for blk, stream in self._block_streams():
                with stream:
                    box = blk.interior_box()
                    Qin, st = blk.regs[regs.Q].view(), blk.state.view()
                    if self.sim.chemistry:
                        self.chemistry_source(blk, Qin, blk.rates.view(), False)
                    rates = blk.rates.view() if self.sim.chemistry else native.View()
                    march = self._march(blk, regs, first)
                    if march:
                        # the i and j sweeps, then the k sweep inside the stage's launch
                        self.fluxes(blk, Qin, st, axes=(0, 1))
                        self._call("pga_stage_march", self.case.handle, Qin, st, blk.regs[0].view(),
                                   blk.regs[regs.Qout].view(), st, blk.iF.view(), blk.jF.view(), blk.kS.view(),
                                   blk.detJinv.view(), rates, len(terms), src, w, withdt, self.dt.ptr, box, march)
                        continue
                    for slab in self._slabs(blk):
                        self.fluxes(blk, Qin, st, slab, one_launch=True)
                        k0, k1 = slab[:2]
                        sbox = native.box((self.ng, self.ng, self.ng + k0), (blk.ncell[0], blk.ncell[1], k1 - k0))
                        if first is None:
                            self._call("pga_stage", self.case.handle, Qin, st, blk.regs[0].view(),
                                       blk.regs[regs.Qout].view(), st, blk.iF.view(), blk.jF.view(), blk.kF.view(),
                                       blk.detJinv.view(), rates, len(terms), src, w, withdt, self.dt.ptr, sbox, kc)
                        else:
                            nA, srcA, wA, dtA, outA = first
                            self._call("pga_stage2", self.case.handle, Qin, st, blk.regs[0].view(),
                                       blk.regs[outA].view(), blk.regs[regs.Qout].view(), st, blk.iF.view(),
                                       blk.jF.view(), blk.kF.view(), blk.detJinv.view(), rates, nA, srcA, wA, dtA,
                                       len(terms), src, w, withdt, self.dt.ptr, sbox, kc)

    _graph = None
    _fresh = True
    replays = 0

    def _standing(self):
        return tuple(sorted(self.readied.items(), key=str))

    def _graphed(self):
        """A step replays one captured HIP graph once its launches are the steady ones: not the
        first step after from_prims (which skips its first halo pass, and makes the arrays a step
        creates on first use), with the registers' halos standing as they stood when the graph was
        captured and no launch timing. Everything a step reads that changes between steps is on the
        device (Q, the state, dt), so the same graph serves every such step, a restarted case's
        included."""
        if not self.cfg["backend-mi355x"]["graphs"] or self._fresh or self.timers is not None:
            return False
        return self._graph is None or self._graph[1] == self._standing()

    def _replay(self, launches):
        if self._graph is None:
            torch = self.backend.torch
            graph = torch.cuda.CUDAGraph()
            before = self._standing()
            # captured on torch's capture stream, which the backend's launches follow; nothing runs,
            # and the host's bookkeeping moves on as the step's would. The cyclic collector is held off
            # meanwhile: a collection inside the capture could free another solver's graph, and
            # destroying a graph is refused while a stream captures.
            collecting = gc.isenabled()
            gc.disable()
            try:
                with torch.cuda.graph(graph):
                    launches()
            finally:
                if collecting:
                    gc.enable()
            self._graph = (graph, before, dict(self.readied))
        self._graph[0].replay()
        self.readied = dict(self._graph[2])
        self.replays += 1

    def release(self):
        """Drop the captured graph and the table of blocks now, on the host's schedule, rather than
        whenever the collector frees the solver (which may fall inside another solver's capture)."""
        self._graph = None
        if self.batch is not None:
            self.batch.close()
            self.batch = None

    def step(self, dt):
        start = self.tme.get()
        self.dt.set(dt)
        times = self.integ.stagetimes(start, dt)
        if self.fused:
            launches = self._batched_step if self.batched else self._fused_step
        else:
            def launches():
                for n in range(self.integ.nstages):
                    self.stage(n)
                    self.tme.set(times[n])
        if self._graphed():
            self._replay(launches)
        else:
            launches()
        self._fresh = False
        self.nrt += 1
        self.tme.set(start + dt)

    def run(self):
        try:
            for _ in range(self.cfg["simulation"]["niter"]):
                self.step(self.integ.controller.stepsize(self))
                for p in self.plugins:
                    if p.due(self):
                        p(self)
        finally:
            for p in self.plugins:
                p.finalize(self)

    def cflmax(self):
        """{acoustic, convective, combined}: the largest inverse time scales over every block's
        interior, reduced on the device per block and across blocks here
        (``multiblock/solver.py:292-296``)."""
        best = np.full(3, -np.inf)
        torch = self.backend.torch
        for blk in self.blocks:
            if blk.dijk is None:
                blk.dijk = self.backend.alloc(blk.cellshape + (3,), name="dijk", lead=self.ng)
                blk.dijk.set(grid.dijk(blk.nodes, blk.ncell))
            box = blk.interior_box()
            n = int(self.case.L.pga_cflmax_worksize(box))
            if self._cflwork is None or self._cflwork.numel() < n + 3:
                self._cflwork = torch.empty(n + 3, dtype=torch.float64, device=self.backend.device)
            work = self._cflwork
            self._call("pga_cflmax", self.case.handle, blk.Q.view(), blk.state.view(), blk.dijk.view(), blk.iS.view(),
                       blk.jS.view(), blk.kS.view(), box, work.data_ptr() + 24, n, work.data_ptr())
            best = np.fmax(best, work[:3].cpu().numpy())
        return dict(zip(self.sim.cflspeeds, best.tolist()))

    def synchronize(self):
        self.backend.synchronize()

    def __repr__(self):
        return f"  Blocks: {self.nblocks}\n" + self.sim.report() + f"  Time Integrator: {self.integ.name}\n"


def get_solver(cfg, mesh, backend, state=None, cls=Solver):
    return cls(cfg, mesh, backend, get_simulator(cfg), get_integrator(cfg), state)
