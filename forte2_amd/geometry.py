"""Geometries: XYZ / Z-matrix parsing and the seeded workload generators.

Units follow forte2 (forte2/system/geom_utils.py:78-140, data/atom_data.py:1-2):
input in angstrom by default, stored in bohr with BOHR_TO_ANGSTROM =
0.529177210903.

Generators (the benchmarked systems):
  water()              O at the origin, O-H 0.9572 A, H-O-H 104.52 deg
  water_cluster(n, L)  n rigid waters with Haar-random orientations at uniform
                       random positions in a cube of side L (A), rejecting any
                       O-O distance below 2.7 A; seeded
  alkane(n)            all-trans C_nH_{2n+2}: C-C 1.54 A, C-H 1.09 A,
                       tetrahedral angles
"""

import math
import re

import numpy as np

BOHR_TO_ANGSTROM = 0.529177210903
ANGSTROM_TO_BOHR = 1.0 / BOHR_TO_ANGSTROM

SYMBOLS = ["X", "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P",
           "S", "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn",
           "Ga", "Ge", "As", "Se", "Br", "Kr"]
SYMBOL_TO_Z = {s.upper(): z for z, s in enumerate(SYMBOLS)}


def _z_of(sym):
    s = re.sub(r"\d+$", "", sym).upper()
    if s not in SYMBOL_TO_Z:
        raise ValueError(f"unknown element {sym!r}")
    return SYMBOL_TO_Z[s]


def parse_geometry(geom, unit="angstrom"):
    """[(Z, xyz_bohr)] from an XYZ or Z-matrix string (';' also ends a line)."""
    if unit not in ("angstrom", "bohr"):
        raise ValueError("unit must be 'angstrom' or 'bohr'")
    scale = ANGSTROM_TO_BOHR if unit == "angstrom" else 1.0
    lines = [l.strip() for l in geom.replace(";", "\n").splitlines() if l.strip()]
    if not lines:
        raise ValueError("Empty geometry string provided.")
    if re.match(r"^([a-zA-Z][a-zA-Z]?)$", lines[0]):
        return _parse_zmatrix(lines, scale)
    atoms = []
    for line in lines:
        tok = line.split()
        if len(tok) != 4:
            raise ValueError(f"Invalid line in XYZ input: {line}")
        atoms.append((_z_of(tok[0]), np.array([float(x) for x in tok[1:]]) * scale))
    return atoms


def _parse_zmatrix(lines, scale):
    xyz, Zs = [], []
    for k, line in enumerate(lines):
        tok = line.split()
        Zs.append(_z_of(tok[0]))
        if k == 0:
            xyz.append(np.zeros(3))
            continue
        a, r = int(tok[1]) - 1, float(tok[2]) * scale
        if k == 1:
            xyz.append(xyz[a] + np.array([0.0, 0.0, r]))
            continue
        b, theta = int(tok[3]) - 1, math.radians(float(tok[4]))
        if k == 2:
            phi, c = 0.0, None
        else:
            c, phi = int(tok[5]) - 1, math.radians(float(tok[6]))
        A, Bp = xyz[a], xyz[b]
        Cp = xyz[c] if c is not None else Bp + np.array([1.0, 0.0, 0.0]) + (np.array([0, 1.0, 0]) if abs((Bp - A)[0]) > 0.99 * np.linalg.norm(Bp - A) else 0)
        bc = (A - Bp) / np.linalg.norm(A - Bp)
        n = np.cross(Bp - Cp, bc)
        if np.linalg.norm(n) < 1e-12:
            n = np.cross(bc, np.array([0.0, 1.0, 0.0]))
            if np.linalg.norm(n) < 1e-12:
                n = np.cross(bc, np.array([1.0, 0.0, 0.0]))
        n /= np.linalg.norm(n)
        m = np.cross(n, bc)
        d = np.array([-r * math.cos(theta), r * math.sin(theta) * math.cos(phi), r * math.sin(theta) * math.sin(phi)])
        xyz.append(A + d[0] * bc + d[1] * m + d[2] * n)
    return [(z, p) for z, p in zip(Zs, xyz)]


def format_xyz(atoms):
    """XYZ string (angstrom) of [(Z, xyz_bohr)]."""
    return "\n".join(f"{SYMBOLS[z]:2s} {p[0]*BOHR_TO_ANGSTROM: .10f} {p[1]*BOHR_TO_ANGSTROM: .10f} "
                     f"{p[2]*BOHR_TO_ANGSTROM: .10f}" for z, p in atoms)


# ------------------------------------------------------------------ generators (angstrom)
WATER_OH = 0.9572
WATER_HOH = 104.52


def water_frame():
    """Rigid water in its own frame (angstrom): O at origin, molecule in the xz plane."""
    t = math.radians(WATER_HOH / 2)
    return [(8, np.zeros(3)),
            (1, np.array([WATER_OH * math.sin(t), 0.0, WATER_OH * math.cos(t)])),
            (1, np.array([-WATER_OH * math.sin(t), 0.0, WATER_OH * math.cos(t)]))]


def water():
    """XYZ string of the single water molecule."""
    return "\n".join(f"{SYMBOLS[z]} {p[0]:.12f} {p[1]:.12f} {p[2]:.12f}" for z, p in water_frame())


def haar_rotation(rng):
    """Haar-distributed rotation: QR of a Gaussian matrix with the sign fix."""
    A = rng.standard_normal((3, 3))
    Q, R = np.linalg.qr(A)
    Q = Q * np.sign(np.diag(R))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


def water_cluster(n, box=None, seed=0, min_oo=2.7, max_tries=200000):
    """XYZ string (angstrom) of n rigid waters in a cube of side `box` (default 3.6 n^(1/3))."""
    if box is None:
        box = 3.6 * n ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    frame = water_frame()
    O = []
    tries = 0
    while len(O) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"could not place {n} waters in a {box} A cube")
        p = rng.uniform(0.0, box, size=3)
        if all(np.linalg.norm(p - q) >= min_oo for q in O):
            O.append(p)
    lines = []
    for p in O:
        Rm = haar_rotation(rng)
        for z, r in frame:
            x = p + Rm @ r
            lines.append(f"{SYMBOLS[z]} {x[0]:.12f} {x[1]:.12f} {x[2]:.12f}")
    return "\n".join(lines)


def alkane(n):
    """XYZ string (angstrom) of all-trans C_nH_{2n+2}, n >= 2.

    Carbons zig-zag in the xy plane (C-C-C = tetrahedral angle).  Every carbon
    has four tetrahedral bonds: for bonds b1, b2 to its (real or, at a chain
    end, continuing-zig-zag virtual) carbon neighbours, the two out-of-plane
    hydrogens point along -u cos(t/2) +- z sin(t/2) with u the bisector of b1,
    b2; a terminal carbon also carries a hydrogen where the next carbon of the
    zig-zag would be.
    """
    if n < 2:
        raise ValueError("alkane needs n >= 2")
    cc, ch = 1.54, 1.09
    tet = math.acos(-1.0 / 3.0)
    dx, dy = cc * math.sin(tet / 2), cc * math.cos(tet / 2)

    def cpos(i):  # zig-zag, valid for any integer i (virtual carbons at -1 and n)
        return np.array([i * dx, dy if i % 2 else 0.0, 0.0])

    C = [cpos(i) for i in range(n)]
    H = []
    zhat = np.array([0.0, 0.0, 1.0])
    for i in range(n):
        b1 = cpos(i - 1) - C[i]
        b2 = cpos(i + 1) - C[i]
        b1 /= np.linalg.norm(b1)
        b2 /= np.linalg.norm(b2)
        u = (b1 + b2) / np.linalg.norm(b1 + b2)
        for sgn in (+1.0, -1.0):
            H.append(C[i] + ch * (-u * math.cos(tet / 2) + sgn * zhat * math.sin(tet / 2)))
        if i == 0:
            H.append(C[i] + ch * b1)
        if i == n - 1:
            H.append(C[i] + ch * b2)
    lines = [f"C {c[0]:.12f} {c[1]:.12f} {c[2]:.12f}" for c in C]
    lines += [f"H {h[0]:.12f} {h[1]:.12f} {h[2]:.12f}" for h in H]
    return "\n".join(lines)
