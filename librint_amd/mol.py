"""Molecules and basis sets in libcint's ``atm | bas | env`` layout.

The layout is the one the reference reads and passes through its C ABI
(reference src/utils.rs:17-81, src/p2c.rs:28-46):

* ``atm[6*a:6*a+6]`` = {Z, ptr_coord, nuc_mod, ptr_zeta, 0, 0}
* ``bas[8*s:8*s+8]`` = {atom, l, nprim, nctr, kappa, ptr_exp, ptr_coeff, 0}
* ``env``: float64 pool; the first 20 slots are reserved (env[0] = expcutoff
  override, env[8] = range-separation omega, both 0 here), then coordinates
  (bohr) and basis parameters.  Coefficients are stored already normalised in
  pyscf's convention (radial normalisation per primitive times contraction
  renormalisation), ``[nctr][nprim]``.

Builders cover the workloads used by tests and the benchmark:
H2O/STO-3G, n-alkanes CnH2n+2/6-31G, and seeded water clusters with an
even-tempered s,p,d basis.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

BOHR = 0.52917721092  # angstrom per bohr (pyscf's constant)
PTR_ENV_START = 20

# --- basis data -------------------------------------------------------------
# Exponents and *normalised* contraction coefficients, as stored in the env
# pools of the reference's molecule files (reference molecules/h2o/sto3g.txt,
# molecules/c6h6/631g.txt).  Entries: (l, exponents, coefficients).
STO3G = {
    1: [(0, [3.42525091, 0.62391373, 0.1688554],
         [0.981706749037157, 0.9494640084217566, 0.295906454323468])],
    8: [(0, [130.70932, 23.808861, 6.4436083],
         [15.072746491370875, 14.57770167457583, 4.543233586562156]),
        (0, [5.0331513, 1.1695961, 0.380389],
         [-0.8486969982288295, 1.13520078524025, 0.856753041179671]),
        (1, [5.0331513, 1.1695961, 0.380389],
         [3.4290657073966995, 2.156288562438211, 0.34159238646916873])],
}
B631G = {
    1: [(0, [18.731137, 2.8253937, 0.6401217],
         [0.7619262198165349, 1.292370996702112, 1.4713190025449525]),
        (0, [0.1612778], [0.6429777834241421])],
    6: [(0, [3047.5249, 457.36951, 103.94869, 29.210155, 9.286663, 3.163927],
         [1.9012558827120676, 3.507511079352715, 5.662218081995158,
          7.370526227802586, 6.289283974418693, 2.1715389431907424]),
        (0, [7.8682724, 1.8812885, 0.5442493],
         [-1.4163903881029998, -0.6528133260732446, 1.8305562330069238]),
        (0, [0.1687144], [0.665088039862392]),
        (1, [7.8682724, 1.8812885, 0.5442493],
         [2.6526321210673007, 2.0338692577523014, 1.015043315744541]),
        (1, [0.1687144], [0.3154456786322328])],
}
# d polarisation shell of carbon in molecules/c6h6/631g.txt (6-31G*)
C_D_POL = (2, [0.8], [1.7657810968610126])


def gaussian_int(n, alpha):
    """int_0^inf r^n exp(-alpha r^2) dr."""
    n1 = (n + 1) * 0.5
    return math.gamma(n1) / (2.0 * alpha ** n1)


def gto_norm(l, alpha):
    """Radial normalisation of r^l exp(-alpha r^2) (pyscf gto_norm)."""
    return 1.0 / math.sqrt(gaussian_int(2 * l + 2, 2 * alpha))


def normalize_contraction(l, exps, coefs_raw):
    """pyscf convention: multiply by gto_norm, then renormalise the contraction.

    ``coefs_raw``: array [nprim] or [nprim, nctr] of raw coefficients.
    Returns normalised coefficients with the same shape.
    """
    es = np.asarray(exps, dtype=np.float64)
    cs = np.asarray(coefs_raw, dtype=np.float64)
    one = cs.ndim == 1
    if one:
        cs = cs[:, None]
    cs = cs * np.array([gto_norm(l, a) for a in es])[:, None]
    ee = es[:, None] + es[None, :]
    gi = np.vectorize(lambda a: gaussian_int(2 * l + 2, a))(ee)
    s = np.einsum("pi,pq,qi->i", cs, gi, cs)
    cs = cs / np.sqrt(s)[None, :]
    return cs[:, 0] if one else cs


@dataclass
class Mole:
    """Minimal stand-in for the pyscf ``Mole`` the reference API consumes."""

    _atm: np.ndarray
    _bas: np.ndarray
    _env: np.ndarray
    nelectron: int
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def nelec(self):
        return (self.nelectron // 2, self.nelectron - self.nelectron // 2)

    @property
    def natm(self):
        return len(self._atm)

    @property
    def nbas(self):
        return len(self._bas)

    def nao(self, coord=0):
        n = 0
        for b in self._bas:
            l, nctr = int(b[1]), int(b[3])
            n += ((l + 1) * (l + 2) // 2 if coord == 0 else 2 * l + 1) * nctr
        return n

    def coords(self):
        return np.array([self._env[a[1]:a[1] + 3] for a in self._atm])

    def charges(self):
        return np.array([a[0] for a in self._atm])

    def arrays(self):
        """C-contiguous int32/int32/float64 flat arrays for the C ABI."""
        return (np.ascontiguousarray(self._atm, dtype=np.int32).ravel(),
                np.ascontiguousarray(self._bas, dtype=np.int32).ravel(),
                np.ascontiguousarray(self._env, dtype=np.float64).ravel())

    def max_l(self):
        return int(self._bas[:, 1].max())


def build(atoms, basis, name="", charge=0, normalized=True):
    """Assemble atm/bas/env.

    ``atoms``: list of (Z, (x, y, z) in bohr).
    ``basis``: {Z: [(l, exps, coefs), ...]}.  With ``normalized`` the
    coefficients are used as given; otherwise they are raw and get pyscf
    normalisation.  Shells of one element share one env block.
    """
    env = [0.0] * PTR_ENV_START
    atm = []
    for Z, xyz in atoms:
        p = len(env)
        env.extend(float(c) for c in xyz)
        env.append(0.0)  # zeta slot (point nucleus)
        atm.append([int(Z), p, 1, p + 3, 0, 0])
    shell_ptr = {}
    for Z in sorted({a[0] for a in atoms}):
        blocks = []
        for (l, exps, coefs) in basis[Z]:
            c = np.asarray(coefs, dtype=np.float64)
            if not normalized:
                c = normalize_contraction(l, exps, c)
            pe = len(env)
            env.extend(float(e) for e in exps)
            pc = len(env)
            env.extend(float(v) for v in (c.T.ravel() if c.ndim == 2 else c))
            nctr = 1 if c.ndim == 1 else c.shape[1]
            blocks.append((l, len(exps), nctr, pe, pc))
        shell_ptr[Z] = blocks
    bas = []
    for ia, (Z, _) in enumerate(atoms):
        for (l, nprim, nctr, pe, pc) in shell_ptr[Z]:
            bas.append([ia, l, nprim, nctr, 0, pe, pc, 0])
    nel = sum(int(Z) for Z, _ in atoms) - charge
    return Mole(np.array(atm, dtype=np.int32), np.array(bas, dtype=np.int32),
                np.array(env, dtype=np.float64), nel, name)


def read_reference_txt(path, nelectron=None, name=""):
    """Read the reference's whitespace ``atm | bas | env`` molecule file."""
    with open(path) as f:
        toks = f.read().split()
    parts, cur = [], []
    for t in toks:
        if t == "|" and len(parts) < 2:
            parts.append(cur)
            cur = []
        elif t != "|":
            cur.append(t)
    parts.append(cur)
    atm = np.array([int(v) for v in parts[0]], dtype=np.int32).reshape(-1, 6)
    bas = np.array([int(v) for v in parts[1]], dtype=np.int32).reshape(-1, 8)
    env = np.array([float(v) for v in parts[2]], dtype=np.float64)
    if nelectron is None:
        nelectron = int(atm[:, 0].sum())
    return Mole(atm, bas, env, nelectron, name or os.path.basename(path))


# --- geometries ---------------------------------------------------------------

def _ang(xyz):
    return tuple(float(v) / BOHR for v in xyz)


WATER_INTERNAL = np.array([[0.0, 0.0, 0.0], [0.757, 0.587, 0.0], [-0.757, 0.587, 0.0]])


def h2o_sto3g():
    """H2O/STO-3G, O(0,0,0), H(+-0.757, 0.587, 0) angstrom: 7 cartesian functions."""
    atoms = [(8, _ang(WATER_INTERNAL[0])), (1, _ang(WATER_INTERNAL[1])), (1, _ang(WATER_INTERNAL[2]))]
    return build(atoms, STO3G, "h2o_sto3g")


def alkane_geometry(n_carbon, r_cc=1.54, r_ch=1.09):
    """All-trans zig-zag CnH2n+2 (angstrom), ideal tetrahedral angles.

    Carbons lie in the xy plane; each CH2 carries its hydrogens out of plane
    along +-z, each terminal CH3 has one in-plane H continuing the zig-zag.
    """
    theta = math.acos(-1.0 / 3.0)  # 109.4712 deg
    half = theta / 2.0
    dx = r_cc * math.sin(half)
    dy = r_cc * math.cos(half)
    C = np.array([[i * dx, (i % 2) * dy, 0.0] for i in range(n_carbon)])
    # virtual carbons beyond each end continue the zig-zag
    ext = np.array([[(i) * dx, (i % 2) * dy, 0.0] for i in (-1, n_carbon)])
    full = np.vstack([ext[0], C, ext[1]])
    H = []
    zhat = np.array([0.0, 0.0, 1.0])
    for i in range(n_carbon):
        c = full[i + 1]
        u1 = full[i] - c
        u1 /= np.linalg.norm(u1)
        u2 = full[i + 2] - c
        u2 /= np.linalg.norm(u2)
        b = -(u1 + u2)
        b /= np.linalg.norm(b)
        for s in (1.0, -1.0):
            d = math.cos(half) * b + s * math.sin(half) * zhat
            H.append(c + r_ch * d)
        if i == 0:
            H.append(c + r_ch * u1)
        if i == n_carbon - 1:
            H.append(c + r_ch * u2)
    if n_carbon == 1:
        raise ValueError("n_carbon >= 2")
    return C, np.array(H)


def alkane_631g(n_carbon, perturb=None):
    C, H = alkane_geometry(n_carbon)
    if perturb is not None:
        C = C + perturb[: len(C)]
        H = H + perturb[len(C):]
    atoms = [(6, _ang(c)) for c in C] + [(1, _ang(h)) for h in H]
    return build(atoms, B631G, "C%dH%d_631g" % (n_carbon, 2 * n_carbon + 2))


def c16h34_631g():
    """n-hexadecane, all-trans, 6-31G (s,p): 148 shells, 212 functions."""
    return alkane_631g(16)


def _even_tempered_basis():
    """Even-tempered s,p,d basis for water clusters.

    O: s exponents 0.15*3^k (k=0..6) -> 4s: the four tightest primitives form
       one contracted function with equal raw weights, the three most diffuse
       are uncontracted; p exponents 0.2*3^k (k=0..3) -> 3p: the two tightest
       contracted with equal raw weights, two free; one d with zeta 0.8.
    H: s exponents 0.1*4^k (k=0..2), uncontracted; one p with zeta 0.75.
    Coefficients get pyscf normalisation.
    """
    os_ = [0.15 * 3 ** k for k in range(7)][::-1]
    op = [0.2 * 3 ** k for k in range(4)][::-1]
    hs = [0.1 * 4 ** k for k in range(3)][::-1]
    O = [(0, os_[:4], [1.0] * 4)] + [(0, [e], [1.0]) for e in os_[4:]]
    O += [(1, op[:2], [1.0, 1.0])] + [(1, [e], [1.0]) for e in op[2:]]
    O += [(2, [0.8], [1.0])]
    H = [(0, [e], [1.0]) for e in hs] + [(1, [0.75], [1.0])]
    return {8: O, 1: H}


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
        [2 * (b * c + a * d), a * a - b * b + c * c - d * d, 2 * (c * d - a * b)],
        [2 * (b * d - a * c), 2 * (c * d + a * b), a * a - b * b - c * c + d * d]])


def water_cluster_geometry(n_water, seed, box=10.0, dmin=2.6):
    """Seeded water cluster (angstrom): O uniform in a box, rejection on O-O < dmin,
    orientation uniform on SO(3) (random unit quaternion), rigid internal geometry."""
    rng = np.random.default_rng(seed)
    Os = []
    while len(Os) < n_water:
        p = rng.uniform(0.0, box, size=3)
        if all(np.linalg.norm(p - q) >= dmin for q in Os):
            Os.append(p)
    atoms = []
    for p in Os:
        R = random_rotation(rng)
        for k, Z in enumerate((8, 1, 1)):
            atoms.append((Z, p + R @ WATER_INTERNAL[k]))
    return atoms


WATER_SEED = 20240611


def water_cluster(n_water, seed=WATER_SEED):
    atoms = [(Z, _ang(x)) for Z, x in water_cluster_geometry(n_water, seed)]
    return build(atoms, _even_tempered_basis(), "w%d_et" % n_water, normalized=False)


def h2o_fg():
    """H2O with f and g shells: STO-3G plus, on O, a d (zeta 0.8), a
    two-primitive f (1.3, 0.45) and a g (1.1); on H a p (0.75) and a d
    (1.0).  Parity molecule for the l = 3, 4 paths (not a published basis)."""
    atoms = [(8, _ang(WATER_INTERNAL[0])), (1, _ang(WATER_INTERNAL[1])), (1, _ang(WATER_INTERNAL[2]))]
    basis = {8: STO3G[8] + [(2, [0.8], [1.0]), (3, [1.3, 0.45], [0.6, 0.5]), (4, [1.1], [1.0])],
             1: STO3G[1] + [(1, [0.75], [1.0]), (2, [1.0], [1.0])]}
    # STO-3G coefficients are already normalised; normalise the added shells
    norm = {}
    for Z, shells in basis.items():
        norm[Z] = [sh if k < len(STO3G[Z]) else (sh[0], sh[1], normalize_contraction(sh[0], sh[1], sh[2]))
                   for k, sh in enumerate(shells)]
    return build(atoms, norm, "h2o_fg")


def h2o_gc():
    """H2O with generally contracted shells (nctr = 2 on shared primitives):
    on O an s pair over the six STO-3G s exponents, a p pair over the three
    sp exponents and a d (zeta 0.8); on H an s pair over the STO-3G
    exponents and a p (zeta 0.75).  24 cartesian functions.  Parity molecule
    for the general-contraction paths (not a published basis)."""
    atoms = [(8, _ang(WATER_INTERNAL[0])), (1, _ang(WATER_INTERNAL[1])), (1, _ang(WATER_INTERNAL[2]))]
    os_ = [130.70932, 23.808861, 6.4436083, 5.0331513, 1.1695961, 0.380389]
    os_c = np.array([[0.15432897, 0.53532814, 0.44463454, 0.02, 0.01, 0.005],
                     [-0.005, -0.02, -0.09996723, 0.39951283, 0.70011547, 0.3]]).T
    op = [5.0331513, 1.1695961, 0.380389]
    op_c = np.array([[0.15591627, 0.60768372, 0.39195739], [0.1, -0.4, 1.0]]).T
    hs = [3.42525091, 0.62391373, 0.1688554]
    hs_c = np.array([[0.15432897, 0.53532814, 0.44463454], [0.3, -0.5, 1.0]]).T
    basis = {8: [(0, os_, os_c), (1, op, op_c), (2, [0.8], [1.0])],
             1: [(0, hs, hs_c), (1, [0.75], [1.0])]}
    return build(atoms, basis, "h2o_gc", normalized=False)


def h2o_631gstar_like():
    """Small d-shell test molecule: H2O with the carbon 6-31G* d shell on O
    (parameters only; used for d-shell parity, not a published basis)."""
    atoms = [(8, _ang(WATER_INTERNAL[0])), (1, _ang(WATER_INTERNAL[1])), (1, _ang(WATER_INTERNAL[2]))]
    basis = {8: STO3G[8] + [C_D_POL], 1: STO3G[1]}
    return build(atoms, basis, "h2o_sto3g_d")
