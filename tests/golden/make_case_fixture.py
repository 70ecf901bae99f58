"""Writes tests/golden/case_rect3x2/input/{nlist00,elist00,blist00}: a 3 x 2-cell rectangle
(12 nodes, 12 triangles, '/' diagonals, two interior nodes moved off the lattice) in the
layout the reference's `biofem sweep` writes (bin/biofem:551-575: tab-separated node and
element lines, coordinates with 6 decimals, 1-based numbers, four open boundary sides in
counter-clockwise order, the number of distinct boundary nodes on line two of blist).
Run from the repository root; the files are committed."""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
NX, NY, LX, LY = 3, 2, 1.5, 1.0
MOVE = {(1, 1): (0.0123456, -0.0065432), (2, 1): (-0.0201, 0.0102)}


def mesh():
    pts = []
    for j in range(NY + 1):
        for i in range(NX + 1):
            dx, dy = MOVE.get((i, j), (0.0, 0.0))
            pts.append((i * LX / NX + dx, j * LY / NY + dy))
    nid = lambda i, j: j * (NX + 1) + i
    tri = []
    for j in range(NY):
        for i in range(NX):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            tri += [(a, b, c), (a, c, d)]
    sides = [[nid(i, 0) for i in range(NX + 1)],                 # bottom, x ascending
             [nid(NX, j) for j in range(NY + 1)],                # right, y ascending
             [nid(i, NY) for i in range(NX, -1, -1)],            # top, x descending
             [nid(0, j) for j in range(NY, -1, -1)]]             # left, y descending
    return pts, tri, sides


def main():
    pts, tri, sides = mesh()
    d = os.path.join(HERE, "case_rect3x2", "input")
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "nlist00"), "w") as f:
        f.write(f"{len(pts)}\n")
        for i, (x, y) in enumerate(pts, start=1):
            f.write(f"{i}\t0\t0\t{x:.6f} {y:.6f} {0.0:.6f}\n")
    with open(os.path.join(d, "elist00"), "w") as f:
        f.write(f"{len(tri)}\n")
        for e, (a, b, c) in enumerate(tri, start=1):
            f.write(f"1\t1\t0\t0\t0\t0\t0\t0\t4\t0\t{e}\t{a+1}\t{b+1}\t{c+1}\t{c+1}\n")
    with open(os.path.join(d, "blist00"), "w") as f:
        f.write("4\n")
        f.write(f"{len(set(v for s in sides for v in s))}\n")
        for s in sides:
            f.write(f"{len(s)}\n")
        for s in sides:
            f.write(" ".join(str(v + 1) for v in s) + "\n")


if __name__ == "__main__":
    main()
