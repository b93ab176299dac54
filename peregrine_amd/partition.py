"""Relabelling a block's axes before its arrays are made: the reorientation step of the
reference's partitioner (``peregrine/partition/orient.py:14-70``), which the interface
orientations exist for. The block's nodes are laid out anew, its interfaces renumbered along
their new axes, and every orientation that names its axes (its own, and its neighbours' digits
toward it) rewritten, so the halo exchange still places each physical cell where it was.
"""

import numpy as np

from peregrine_amd.mesher import Interface


class Conditioner:
    def reorient(self, mb, blk, perm, flips):
        """New axis m of blk holds its old axis perm[m], running backwards where flips[m]."""
        perm = [int(p) for p in perm]
        flips = [bool(f) for f in flips]
        if sorted(perm) != [0, 1, 2]:
            raise ValueError(f"reorient: {perm} is not a permutation of the three axes")
        if perm == [0, 1, 2] and not any(flips):
            return
        if any(f.nbr == blk.nblki for f in blk.faces.values()):
            # its periodic halo nodes are moved by the period along its own index axes
            raise ValueError(f"reorient: block {blk.nblki} is joined to itself; this build keeps such blocks as laid out")
        new_of = {old: new for new, old in enumerate(perm)}
        signed, code = Interface.signedaxis, Interface.orientcode

        # a neighbour's digits toward this block name this block's axes: each takes its new label,
        # reversed where that new axis runs backwards
        for other in mb.blocks:
            if other is blk:
                continue
            for f in other.faces.values():
                if f.nbr != blk.nblki or f.orient is None:
                    continue
                digits = []
                for c in f.orient:
                    old, flip = signed(c)
                    m = new_of[old]
                    digits.append(code(m, flip != flips[m]))
                f.orient = "".join(digits)

        # this block's own digits are indexed by its axes: new axis m takes old axis perm[m]'s
        for f in blk.faces.values():
            if f.orient is None:
                continue
            digits = []
            for m in range(3):
                theirs, flip = signed(f.orient[perm[m]])
                digits.append(code(theirs, flip != flips[m]))
            f.orient = "".join(digits)

        # each interface keeps its partner and name; its number follows its axis, a flip swaps its end
        renumbered = {}
        for f in list(blk.faces.values()):
            m = new_of[f.axis]
            low = f.is_low != flips[m]
            f.nface = 2 * m + (1 if low else 2)
            renumbered[f.nface] = f
        blk.faces = dict(sorted(renumbered.items()))

        nodes = np.moveaxis(blk.interior_nodes, perm, (0, 1, 2))
        for m in range(3):
            if flips[m]:
                nodes = np.flip(nodes, axis=m)
        blk.interior_nodes = np.ascontiguousarray(nodes)
        blk.set_extents(*nodes.shape[:3])
        blk.period = [blk.period[p] for p in perm]
