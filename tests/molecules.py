"""Small molecules the reference's own tests use (geometries in Angstrom)."""
import numpy as np

ANGSTROM_TO_BOHR = 1.0 / 0.529167
_Z = {"H": 1, "C": 6, "N": 7, "O": 8, "F": 9, "Cl": 17, "S": 16, "P": 15}

# tests/auxiliary_integral_impact.rs:22-50 (recorded heats of formation, kcal/mol)
RECORDED_HOF = {
    "HCN": ("H 0.0000 0.0 0.0\nC 1.0655 0.0 0.0\nN 2.2261 0.0 0.0", 31.0125),
    "CH3NH2": ("C -0.7000 0.0000 0.0000\nN 0.7000 0.0000 0.0000\nH -1.0800 1.0180 0.0000\n"
               "H -1.0800 -0.5090 0.8815\nH -1.0800 -0.5090 -0.8815\nH 1.0400 -0.4680 0.8300\n"
               "H 1.0400 -0.4680 -0.8300", -2.4802),
    "CH3OH": ("C -0.7000 0.0000 0.0000\nO 0.7000 0.0000 0.0000\nH -1.0800 1.0180 0.0000\n"
              "H -1.0800 -0.5090 0.8815\nH -1.0800 -0.5090 -0.8815\nH 0.9800 -0.8900 0.2000", -54.7464),
    "CH3Cl": ("C 0.0000 0.0000 0.0000\nCl 1.7810 0.0000 0.0000\nH -0.3600 1.0270 0.0000\n"
              "H -0.3600 -0.5135 0.8894\nH -0.3600 -0.5135 -0.8894", -17.5923),
    "H2O": ("O 0.0000 0.0000 0.0\nH 0.9614 0.0000 0.0\nH -0.2246 0.9348 0.0", -59.2408),
}

# scf.rs:1297-1300 / :1347-1348 / :1375 ; tests/mopac_reference.rs:98-104
WATER = "O 0.0000 0.0000 0.0000\nH 0.9584 0.0000 0.0000\nH -0.2400 0.9278 0.0000"
FORMALDEHYDE = "C 0.0 0.0 0.0\nO 0.0 0.0 1.21\nH 0.94 0.0 -0.54\nH -0.94 0.0 -0.54"
METHYL = "C 0.0 0.0 0.0\nH 1.079 0.0 0.0\nH -0.5395 0.9344 0.0\nH -0.5395 -0.9344 0.0"
METHANE = ("C 0.0000 0.0000 0.0000\nH 0.6276 0.6276 0.6276\nH -0.6276 -0.6276 0.6276\n"
           "H -0.6276 0.6276 -0.6276\nH 0.6276 -0.6276 -0.6276")


def co2(bond=1.16):
    return f"C 0.0 0.0 0.0\nO {bond} 0.0 0.0\nO -{bond} 0.0 0.0"


def parse(text):
    """XYZ body -> (z int32 array, coordinates in Bohr)."""
    z, x = [], []
    for line in text.strip().splitlines():
        p = line.split()
        z.append(_Z[p[0]])
        x.append([float(v) * ANGSTROM_TO_BOHR for v in p[1:4]])
    return np.array(z, dtype=np.int32), np.array(x)


def water_chain(n):
    """tests/divide_conquer.rs:22-43: waters along x at 2.9 A with alternating tilt (Bohr out)."""
    z, x = [], []
    for i in range(n):
        base = np.array([i * 2.9, 0.0 if i % 2 == 0 else 0.35, 0.0])
        for zz, d in ((8, (0.0, 0.0, 0.0)), (1, (0.9584, 0.0, 0.0)), (1, (-0.2400, 0.9279, 0.0))):
            z.append(zz)
            x.append((base + np.array(d)) * ANGSTROM_TO_BOHR)
    return np.array(z, dtype=np.int32), np.array(x)


def benzene(cc=1.39, ch=1.09):
    """D6h benzene (doubly degenerate e1g / e2u orbitals)."""
    import math
    lines = []
    for k in range(6):
        a = math.pi / 3.0 * k
        lines.append(f"C {cc * math.cos(a):.10f} {cc * math.sin(a):.10f} 0.0")
        lines.append(f"H {(cc + ch) * math.cos(a):.10f} {(cc + ch) * math.sin(a):.10f} 0.0")
    return "\n".join(lines)


def cubane(cc=1.57, ch=1.09):
    """Oh cubane C8H8 (triply degenerate t orbitals)."""
    import math
    h = cc / 2.0
    s = ch / math.sqrt(3.0)
    lines = []
    for x in (-1, 1):
        for y in (-1, 1):
            for z in (-1, 1):
                lines.append(f"C {x * h:.10f} {y * h:.10f} {z * h:.10f}")
                lines.append(f"H {x * (h + s):.10f} {y * (h + s):.10f} {z * (h + s):.10f}")
    return "\n".join(lines)
