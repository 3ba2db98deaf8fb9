"""Test meshes shared by the oracle pinning tests and the GPU parity tests."""
import numpy as np


def quadtree_mesh(nc=4, h=16.0, refine=((1, 1), (2, 1), (1, 2), (2, 2)), x0=20.0, y0=30.0,
                  all_midside=False):
    """A two-level quadtree on an nc x nc grid of coarse cells of size 2h: the cells in
    `refine` are split into four; a coarse cell next to a split one gets the midside
    node of that edge (nodes 5..8 = right, top, left, bottom edge midpoints).  With
    all_midside every coarse cell gets all four midside nodes (no element lacks one)."""
    nodes = {}

    def node(x, y):
        k = (round(x * 4), round(y * 4))
        if k not in nodes:
            nodes[k] = len(nodes)
        return nodes[k]

    ref = set(refine)
    elems = []
    for cy in range(nc):
        for cx in range(nc):
            X, Y, H = x0 + 2 * h * cx, y0 + 2 * h * cy, 2 * h
            if (cx, cy) in ref:
                for sy in range(2):
                    for sx in range(2):
                        xa, ya = X + h * sx, Y + h * sy
                        elems.append([node(xa, ya), node(xa + h, ya), node(xa + h, ya + h),
                                      node(xa, ya + h), -1, -1, -1, -1])
            else:
                e = [node(X, Y), node(X + H, Y), node(X + H, Y + H), node(X, Y + H)]
                mids = [((cx + 1, cy), (X + H, Y + h)), ((cx, cy + 1), (X + h, Y + H)),
                        ((cx - 1, cy), (X, Y + h)), ((cx, cy - 1), (X + h, Y))]
                e += [node(*p) if (nb in ref or all_midside) else -1 for nb, p in mids]
                elems.append(e)
    xy = np.zeros((len(nodes), 2))
    for (kx, ky), i in nodes.items():
        xy[i] = (kx / 4.0, ky / 4.0)
    return xy, np.array(elems, np.int64)
