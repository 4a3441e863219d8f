"""Closed triangle meshes for the seeding tests (counter-clockwise seen from outside), and an
OBJ writer.  Test helpers, not collected."""
import numpy as np


def cube(lo=(-1.0, -1.0, -1.0), hi=(1.0, 1.0, 1.0)):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    v = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    # vertex index = 4 x + 2 y + z with x, y, z in {0, 1}
    f = np.array([[0, 1, 3], [0, 3, 2],      # x = lo (normal -x)
                  [4, 6, 7], [4, 7, 5],      # x = hi
                  [0, 4, 5], [0, 5, 1],      # y = lo
                  [2, 3, 7], [2, 7, 6],      # y = hi
                  [0, 2, 6], [0, 6, 4],      # z = lo
                  [1, 5, 7], [1, 7, 3]])     # z = hi
    return v, f


def icosphere(subdiv=3, radius=1.0, centre=(0.0, 0.0, 0.0), bump=0.0):
    """Subdivided icosahedron; bump > 0 moves every vertex radially by bump * radius *
    sin(3 theta) cos(2 phi) (a star-shaped blob, still closed and outward-wound)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    v = [np.asarray(x, float) / np.linalg.norm(x) for x in v]
    for _ in range(subdiv):
        mid = {}

        def m(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                x = v[a] + v[b]
                v.append(x / np.linalg.norm(x))
                mid[k] = len(v) - 1
            return mid[k]

        f = [g for a, b, c in f for g in ([a, m(a, b), m(c, a)], [b, m(b, c), m(a, b)], [c, m(c, a), m(b, c)],
                                          [m(a, b), m(b, c), m(c, a)])]
    v = np.asarray(v)
    if bump:
        th = np.arccos(np.clip(v[:, 2], -1, 1))
        ph = np.arctan2(v[:, 1], v[:, 0])
        v = v * (1.0 + bump * np.sin(3 * th) * np.cos(2 * ph))[:, None]
    return v * radius + np.asarray(centre, float), np.asarray(f)


def flip(f):
    return np.asarray(f)[:, [0, 2, 1]]


def triangles(v, f):
    return np.ascontiguousarray(np.asarray(v, float)[np.asarray(f)].reshape(-1, 9))


def write_obj(path, v, f, y_up_from_z=False):
    """OBJ text; with y_up_from_z the vertices are written Y-up, (x, y, z) -> (x, z, -y), the
    inverse of the seeding's Y-up to Z-up turn."""
    v = np.asarray(v, float)
    if y_up_from_z:
        v = np.stack([v[:, 0], v[:, 2], -v[:, 1]], 1)
    with open(path, "w") as fh:
        for x in v:
            fh.write(f"v {x[0]:.17g} {x[1]:.17g} {x[2]:.17g}\n")
        for a, b, c in np.asarray(f):
            fh.write(f"f {a + 1} {b + 1} {c + 1}\n")
