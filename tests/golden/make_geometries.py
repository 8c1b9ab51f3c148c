"""Write the seeded synthetic geometries used by the tests and the bench as .xyz files (angstrom).

Workloads (seeds are numpy PCG64 ``default_rng`` seeds):
  h2o.xyz          water, O (0,0,0), H (+-0.7572, 0, 0.5865)
  h2o_8.xyz        (H2O)8: O on a 2x2x2 cubic grid of 2.9 A, +-0.2 A jitter, random rotations, seed 1
  c12h26.xyz       all-trans C12H26, C-C 1.54 A, C-H 1.09 A, tetrahedral angles, +-0.02 A jitter, seed 2
  h2o_16.xyz       (H2O)16 on a 4x2x2 grid, seed 3
  batch_h2o.xyz    512 clusters (multi-frame), n ~ U{1..10} (seed 5), cluster i uses seed 1000+i
Monomer: r(OH) = 0.9572 A, angle HOH = 104.52 deg.
"""
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROH = 0.9572
HOH = np.deg2rad(104.52)


def monomer():
    s, c = np.sin(HOH / 2), np.cos(HOH / 2)
    return np.array([[0.0, 0.0, 0.0], [ROH * s, 0.0, ROH * c], [-ROH * s, 0.0, ROH * c]])


def quat_rot(u):
    """Uniform random rotation from three uniforms (Shoemake 1992)."""
    u1, u2, u3 = u
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    q = np.array([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)])
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(n, dims, seed, spacing=2.9, jitter=0.2):
    rng = np.random.default_rng(seed)
    grid = [(i, j, k) for i in range(dims[0]) for j in range(dims[1]) for k in range(dims[2])][:n]
    jit = rng.uniform(-jitter, jitter, size=(n, 3))
    rots = rng.random(size=(n, 3))
    atoms = []
    m = monomer()
    for w in range(n):
        o = np.array(grid[w], dtype=float) * spacing + jit[w]
        r = quat_rot(rots[w])
        for a, sym in enumerate("OHH"):
            atoms.append((sym, o + r @ m[a]))
    return atoms


def alkane(nc=12, seed=2, jitter=0.02):
    cc, ch = 1.54, 1.09
    half = np.deg2rad(109.4712) / 2
    carbons = [np.array([i * cc * np.sin(half), (i % 2) * cc * np.cos(half), 0.0]) for i in range(nc)]
    atoms = [("C", c) for c in carbons]
    z = np.array([0.0, 0.0, 1.0])
    for i, c in enumerate(carbons):
        if 0 < i < nc - 1:
            u1 = carbons[i - 1] - c
            u2 = carbons[i + 1] - c
            b = -(u1 / np.linalg.norm(u1) + u2 / np.linalg.norm(u2))
            b /= np.linalg.norm(b)
            for sgn in (1, -1):
                d = np.cos(half) * b + sgn * np.sin(half) * z
                atoms.append(("H", c + ch * d))
        else:
            nb = carbons[1] if i == 0 else carbons[nc - 2]
            u = (nb - c) / np.linalg.norm(nb - c)
            # anti (in-plane) hydrogen continues the zig-zag, two out of plane
            perp = np.cross(z, u)
            perp /= np.linalg.norm(perp)
            other = carbons[2] if i == 0 else carbons[nc - 3]
            if np.dot(perp, other - c) > 0:
                perp = -perp
            ct = np.cos(np.deg2rad(109.4712))
            st = np.sin(np.deg2rad(109.4712))
            atoms.append(("H", c + ch * (ct * u + st * perp)))
            for sgn in (1, -1):
                d = ct * u + st * (-0.5 * perp + sgn * np.sqrt(3) / 2 * z)
                atoms.append(("H", c + ch * d))
    rng = np.random.default_rng(seed)
    jit = rng.uniform(-jitter, jitter, size=(len(atoms), 3))
    return [(s, p + jit[k]) for k, (s, p) in enumerate(atoms)]


def write_xyz(path, atoms, comment):
    with open(path, "w") as f:
        f.write(f"{len(atoms)}\n{comment}\n")
        for s, p in atoms:
            f.write(f"{s:2s} {p[0]:18.12f} {p[1]:18.12f} {p[2]:18.12f}\n")


def main():
    write_xyz(os.path.join(HERE, "h2o.xyz"),
              [("O", np.zeros(3)), ("H", np.array([0.7572, 0, 0.5865])), ("H", np.array([-0.7572, 0, 0.5865]))],
              "water, angstrom")
    write_xyz(os.path.join(HERE, "h2o_8.xyz"), water_cluster(8, (2, 2, 2), 1), "(H2O)8 2x2x2 seed 1, angstrom")
    write_xyz(os.path.join(HERE, "c12h26.xyz"), alkane(), "all-trans C12H26 seed 2, angstrom")
    write_xyz(os.path.join(HERE, "h2o_16.xyz"), water_cluster(16, (4, 2, 2), 3), "(H2O)16 4x2x2 seed 3, angstrom")
    ns = np.random.default_rng(5).integers(1, 11, size=512)
    with open(os.path.join(HERE, "batch_h2o.xyz"), "w") as f:
        for i, n in enumerate(ns):
            k = 1
            while k ** 3 < n:
                k += 1
            atoms = water_cluster(int(n), (k, k, k), 1000 + i)
            f.write(f"{len(atoms)}\n(H2O){n} seed {1000 + i}, angstrom\n")
            for sym, pos in atoms:
                f.write(f"{sym:2s} {pos[0]:18.12f} {pos[1]:18.12f} {pos[2]:18.12f}\n")

if __name__ == "__main__":
    main()
