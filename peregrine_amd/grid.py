"""Block geometry: halo nodes and the metrics the kernels read, computed once on the host.

Restates ``peregrine/multiblock/block/grid.py:84-202``:
- face-area vectors as half the cross product of the face diagonals, taken over
  the cyclic axis pair so they point along the axis (``_areavecs``);
- cell volume as the body diagonal dotted with the high faces' area sum over 3
  (``_volumes``), stored as its inverse (``_detJinv``);
- cell centres as the mean of the 8 nodes, face centres as the mean of 4;
- the inverse Jacobian as the adjugate of the averaged edge vectors over the
  volume (``_Jinv``).
Halo nodes (``gen_halo``, ``move_periodics``) are linear extrapolation across
a boundary, and the partner's nodes moved by the period across a periodic
interface (identical on the uniform grids of a cube mesher).
"""

import numpy as np


def halo_nodes(interior, ng, periodic, lengths):
    """Nodes with ng halo layers: interior (ni, nj, nk, 3) -> (ni+2ng, nj+2ng, nk+2ng, 3)."""
    x = np.pad(interior, [(ng, ng)] * 3 + [(0, 0)])
    for axis in range(3):
        n = interior.shape[axis]
        a = np.moveaxis(x, axis, 0)
        for layer in range(ng):
            lo, hi = ng - 1 - layer, ng + n + layer
            if periodic[axis] and n - 1 >= layer + 1:
                shift = np.zeros(3)
                shift[axis] = lengths[axis]
                # the node layer+1 in from the far end, moved back by one period (the end planes coincide)
                a[lo] = a[ng + n - 2 - layer] - shift
                a[hi] = a[ng + 1 + layer] + shift
            else:
                a[lo] = 2.0 * a[lo + 1] - a[lo + 2]
                a[hi] = 2.0 * a[hi - 1] - a[hi - 2]
    return x


def _corner(nodes, axes, *highs):
    s = [slice(None)] * 3
    for n, high in zip(axes, highs):
        s[n] = np.s_[1:] if high else np.s_[:-1]
    return nodes[tuple(s)]


def cellcenters(nodes):
    return 0.125 * (
        nodes[:-1, :-1, :-1] + nodes[:-1, :-1, 1:] + nodes[:-1, 1:, :-1] + nodes[:-1, 1:, 1:]
        + nodes[1:, :-1, :-1] + nodes[1:, :-1, 1:] + nodes[1:, 1:, :-1] + nodes[1:, 1:, 1:]
    )


def facecenters(nodes, a):
    inplane = [n for n in range(3) if n != a]
    c = [_corner(nodes, inplane, *h) for h in ((0, 0), (0, 1), (1, 0), (1, 1))]
    return 0.25 * (c[0] + c[1] + c[2] + c[3])


def dijk(nodes, ncell):
    """Per cell and axis, the distance between the centres of its two faces along that axis; an
    axis one cell thick is not marched in, and its length is infinite so that no step limit binds
    on it (``multiblock/block/grid.py:164-176``)."""
    lengths = []
    for a in range(3):
        centers = facecenters(nodes, a)
        far, near = [slice(None)] * 3, [slice(None)] * 3
        far[a], near[a] = np.s_[1:], np.s_[:-1]
        span = centers[tuple(far)] - centers[tuple(near)]
        length = np.sqrt((span**2).sum(axis=-1))
        lengths.append(np.full_like(length, np.inf) if ncell[a] == 1 else length)
    return np.stack(lengths, axis=-1)


def areavecs(nodes, a):
    diagonal = [(a + 1) % 3, (a + 2) % 3]
    c00, c01, c10, c11 = (_corner(nodes, diagonal, *h) for h in ((0, 0), (0, 1), (1, 0), (1, 1)))
    return 0.5 * np.cross(c10 - c01, c11 - c00)


def volumes(nodes):
    bodydiag = nodes[1:, 1:, 1:] - nodes[:-1, :-1, :-1]
    S = [areavecs(nodes, a) for a in range(3)]
    high = S[0][1:] + S[1][:, 1:] + S[2][:, :, 1:]
    return sum(bodydiag[..., n] * high[..., n] for n in range(3)) / 3.0


def detJinv(nodes, ng):
    J = volumes(nodes)
    inner = J[ng:-ng, ng:-ng, ng:-ng] if ng else J
    bad = int((inner <= 0.0).sum())
    if bad:
        raise ValueError(f"{bad} cells with no volume or a negative one")
    with np.errstate(divide="ignore"):
        return 1.0 / J


def Jinv(nodes):
    corners = [
        _corner(nodes, (0, 1, 2), *h)
        for h in ((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1))
    ]
    c1, c2, c3, c4, c5, c6, c7, c8 = corners
    dxi = 0.25 * ((c4 - c1) + (c8 - c5) + (c3 - c2) + (c7 - c6))
    deta = 0.25 * ((c2 - c1) + (c3 - c4) + (c7 - c8) + (c6 - c5))
    dzeta = 0.25 * ((c5 - c1) + (c8 - c4) + (c6 - c2) + (c7 - c3))
    rows = np.cross(deta, dzeta), np.cross(dzeta, dxi), np.cross(dxi, deta)
    vol = volumes(nodes)[..., None, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.stack(rows, axis=-2) / vol
