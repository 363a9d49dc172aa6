"""Synthetic systems of the benchmark workloads (host-side definitions).

Slater-Jastrow systems are plain data: nuclei, a basis (exponent, centre and
kind per function: s-type STO exp(-zeta r), Gaussian exp(-zeta r^2), or a p-type
STO x exp(-zeta r), y exp(-zeta r), z exp(-zeta r)), orbital coefficients, spin
occupations and the five Jastrow parameters (b_anti, b_para, c, a, d) of
    J = sum_{i<j} b_ij r_ij/(1 + c r_ij) - sum_{i,I} a Z_I r_iI/(1 + d r_iI).
The same arrays feed the device (through the C ABI) and the CPU oracle.
"""
import math

import numpy as np

DEFAULT_JASTROW = (0.5, 0.25, 1.0, 0.3, 1.0)
# basis kinds (the C ABI's MOLE_BASIS_*)
STO, GTO, STO_PX, STO_PY, STO_PZ = 0, 1, 2, 3, 4
ELEMENTS = {1: "H", 2: "He", 3: "Li", 4: "Be", 5: "B", 6: "C", 7: "N", 8: "O", 9: "F", 10: "Ne"}


class SJSystem:
    """A Slater-Jastrow system: determinants of n_up and n_dn electrons over
    orbitals of STO (s and p) or Gaussian basis functions, times the Pade e-e
    and e-n Jastrow factor."""

    def __init__(self, nuclei, charges, n_up, n_dn, zeta, center, coef, jastrow=DEFAULT_JASTROW,
                 name="", kind=None):
        self.nuclei = np.ascontiguousarray(nuclei, dtype=np.float64).reshape(-1, 3)
        self.charges = np.ascontiguousarray(charges, dtype=np.int32)
        self.n_up = int(n_up)
        self.n_dn = int(n_dn)
        self.zeta = np.ascontiguousarray(zeta, dtype=np.float64)
        self.center = np.ascontiguousarray(center, dtype=np.int32)
        self.coef = np.ascontiguousarray(coef, dtype=np.float64).reshape(-1, self.zeta.size)
        self.jastrow = np.ascontiguousarray(jastrow, dtype=np.float64)
        self.name = name
        # basis kinds (STO, GTO, STO_PX..STO_PZ); None: all s-type STOs
        self.kind = None if kind is None else np.ascontiguousarray(kind, dtype=np.int32)
        assert self.coef.shape[0] >= max(self.n_up, self.n_dn)
        assert self.center.size == self.zeta.size
        assert np.all((self.center >= 0) & (self.center < self.nuclei.shape[0]))

    @property
    def n_electrons(self):
        return self.n_up + self.n_dn

    @property
    def n_basis(self):
        return self.zeta.size


def atom(Z, seed=1, jastrow=DEFAULT_JASTROW):
    """Neutral atom Z at the origin: n_up = ceil(Z/2), n_dn = floor(Z/2), in a
    minimal STO basis of five functions -- two s functions exp(-zeta r) and
    one p shell (x, y, z) exp(-zeta r), the 1s, 2s and 2p of the first row --
    with the three exponents uniform in [0.5, Z] and N(0,1) coefficients for
    n_up orbitals, from numpy's default_rng(seed).  (The p shell gives the
    spin determinants of B..Ne their angular rank: over s functions of one
    centre alone they are numerically singular from four orbitals on.)"""
    rng = np.random.default_rng(seed)
    n_up = (Z + 1) // 2
    n_dn = Z // 2
    zs = rng.uniform(0.5, float(Z), 3)
    zeta = [zs[0], zs[1], zs[2], zs[2], zs[2]]
    kind = [STO, STO, STO_PX, STO_PY, STO_PZ]
    coef = rng.normal(size=(n_up, 5))
    return SJSystem([(0.0, 0.0, 0.0)], [Z], n_up, n_dn, zeta, np.zeros(5, np.int32), coef,
                    jastrow, name=ELEMENTS.get(Z, f"Z{Z}"), kind=kind)


def hydrogenic_2p(Z=1, axis=2):
    """One electron in the 2p orbital (x, y or z) exp(-Z r/2) of a charge-Z
    nucleus: an exact eigenfunction, E_L = -Z^2/8."""
    return SJSystem([(0.0, 0.0, 0.0)], [Z], 1, 0, [0.5 * Z], [0], [[1.0]], (0, 0, 0, 0, 0),
                    name=f"hydrogenic 2p Z={Z}", kind=[STO_PX + axis])


def hydrogenic(Z=1):
    """One electron in exp(-Z r): an exact eigenfunction, E_L = -Z^2/2."""
    return SJSystem([(0.0, 0.0, 0.0)], [Z], 1, 0, [float(Z)], [0], [[1.0]], (0, 0, 0, 0, 0),
                    name=f"hydrogenic Z={Z}")


def harmonic_oscillator(omega=1.0, a=None):
    """One electron in exp(-(r/a)^2) about a charge-0 centre: the Gaussian
    trial function and oscillator of examples/custom_operator.rs:64-100,
    a = sqrt(2/omega) by default (exact: E_L = 3 omega/2 under
    T + 1/2 omega^2 r^2)."""
    a = math.sqrt(2.0 / omega) if a is None else a
    return SJSystem([(0.0, 0.0, 0.0)], [0], 1, 0, [1.0 / (a * a)], [0], [[1.0]], (0, 0, 0, 0, 0),
                    name=f"oscillator w={omega}", kind=[1])


def helium_product(zeta=2.0, b=0.0, c=0.3):
    """He: exp(-zeta r) for both electrons, optional e-e Pade factor."""
    return SJSystem([(0.0, 0.0, 0.0)], [2], 1, 1, [zeta], [0], [[1.0]], (b, 0.0, c, 0.0, 1.0),
                    name="He product")


def h2_lcao(R=1.4, zeta=1.0, b=0.5, c=0.5):
    """H2 on the z axis (nuclei at +-R/2): both electrons in the sigma_g
    orbital exp(-zeta |r - A|) + exp(-zeta |r - B|), e-e Pade factor."""
    return SJSystem([(0.0, 0.0, -0.5 * R), (0.0, 0.0, 0.5 * R)], [1, 1], 1, 1, [zeta, zeta], [0, 1], [[1.0, 1.0]],
                    (b, 0.0, c, 0.0, 1.0), name="H2 LCAO")


def h2_ion_product(R=2.5, alpha=1.0):
    """H2+ as tests/hydrogen_molecular_ion_lcao.rs:60-77 writes it: one
    electron in exp(-alpha |r - A|) exp(-alpha |r - B|), nuclei at (+-R/2, 0,
    0).  As a Slater-Jastrow function: a one-by-one determinant of a constant
    orbital (one STO with zeta 0) times the e-n Pade factor with a = alpha,
    d = 0, exp(-alpha r_A - alpha r_B)."""
    return SJSystem([(-0.5 * R, 0.0, 0.0), (0.5 * R, 0.0, 0.0)], [1, 1], 1, 0, [0.0], [0], [[1.0]],
                    (0.0, 0.0, 1.0, alpha, 0.0), name="H2+ product")


def poisson_disk(n, min_sep, rng, box=None, max_tries=10000):
    """n points with pairwise distance >= min_sep by dart throwing in a cube
    whose side grows with n (seeded)."""
    side = box if box is not None else max(2.0, min_sep * (2.0 * n) ** (1.0 / 3.0) * 1.2)
    pts = []
    tries = 0
    while len(pts) < n:
        p = rng.uniform(0.0, side, 3)
        if all(np.linalg.norm(p - q) >= min_sep for q in pts):
            pts.append(p)
        tries += 1
        if tries > max_tries:
            side *= 1.2
            tries = 0
    pts = np.array(pts)
    return pts - pts.mean(axis=0)


def molecule(charges, positions, basis_per_nucleus=2, seed=0, jastrow=DEFAULT_JASTROW, name="", p_shells=True):
    """Neutral molecule: electrons = sum Z, closed shell where the count is
    even (n_up = ceil(N/2)); basis_per_nucleus s-type STOs per nucleus and,
    with p_shells, one p shell (x, y, z) on every nucleus with Z >= 3, all
    with exponents uniform in [0.5, max(1, Z)]; N(0,1) coefficients.  (A
    nucleus then carries at least ceil(Z/2) functions for Z <= 10, so the
    orbitals span n_up dimensions.)"""
    rng = np.random.default_rng(seed)
    charges = np.asarray(charges, dtype=np.int32)
    N = int(charges.sum())
    n_up, n_dn = (N + 1) // 2, N // 2
    zeta, center, kind = [], [], []
    for I, Z in enumerate(charges):
        zeta.extend(rng.uniform(0.5, max(1.0, float(Z)), basis_per_nucleus))
        center.extend([I] * basis_per_nucleus)
        kind.extend([STO] * basis_per_nucleus)
        if p_shells and Z >= 3:
            zp = rng.uniform(0.5, float(Z))
            zeta.extend([zp] * 3)
            center.extend([I] * 3)
            kind.extend([STO_PX, STO_PY, STO_PZ])
    nb = len(zeta)
    coef = rng.normal(size=(n_up, nb)) / math.sqrt(nb)
    return SJSystem(positions, charges, n_up, n_dn, zeta, center, coef, jastrow, name=name, kind=kind)


def hydrogen_chain(n, spacing=1.4, basis_per_nucleus=2, seed=0):
    pos = np.zeros((n, 3))
    pos[:, 2] = (np.arange(n) - (n - 1) / 2.0) * spacing
    return molecule(np.ones(n, np.int32), pos, basis_per_nucleus, seed, name=f"H{n} chain")


def random_molecule(n_electrons, seed, min_sep=1.2, basis_per_nucleus=2, max_z=8):
    """Nuclei with Z in 1..max_z summing to n_electrons (neutral), placed by
    seeded Poisson-disk sampling with minimum separation min_sep."""
    rng = np.random.default_rng(seed)
    charges = []
    left = n_electrons
    while left > 0:
        z = int(rng.integers(1, min(max_z, left) + 1))
        charges.append(z)
        left -= z
    pos = poisson_disk(len(charges), min_sep, rng)
    return molecule(charges, pos, basis_per_nucleus, seed + 1000, name=f"mol{n_electrons}-{seed}")
