"""Seeded synthetic molecules written as real FCHK text, for the benchmark
workloads (the configurations of the evaluator's benchmark plan) and the tests.

Shells follow the FCHK conventions the reader expects (src/basis.cpp:21-120):
shell type -1 is SP, <= -2 pure, >= 0 Cartesian. Exponents are log-uniform
in [0.1, 5000]; contraction coefficients uniform in (0.1, 1) and scaled so
that each contracted AO has unit norm; MO coefficients are N(0, 1) and then
orthonormalised in the AO overlap metric (C^T S C = 1); P_total and P_spin are
built from the occupied MOs, so the orbital representation is eligible for
the AO-MO contraction. Reals are written with 17 significant digits (the
reader accepts any float token) so that C C^T reproduces P to rounding.
"""
import math
import os

import numpy as np

# Cartesian component order of FCHK files (src/basis_evaluator.cpp:135-152).
_CART = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
         2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
         3: [(3, 0, 0), (0, 3, 0), (0, 0, 3), (1, 2, 0), (2, 1, 0), (2, 0, 1), (1, 0, 2), (0, 1, 2), (0, 2, 1),
             (1, 1, 1)]}
# g shells and above: lexicographic in (lx, ly) (zzzz, yzzz, yyzz, ...).
for _l in (4, 5):
    _CART[_l] = [(lx, ly, _l - lx - ly) for lx in range(_l + 1) for ly in range(_l - lx + 1)]
MAX_L = 5

# Shell recipes: (FCHK shell type, primitives). 6-31G*-shaped: Cartesian d.
POPLE_HEAVY = [(0, 6), (-1, 3), (-1, 1), (2, 1)]
POPLE_H = [(0, 3), (0, 1)]
# def2-TZVP-shaped: pure d and f; heavy 5s3p2d1f, H 3s1p.
TZVP_HEAVY = [(0, 6), (0, 2), (0, 1), (0, 1), (0, 1), (1, 4), (1, 1), (1, 1), (-2, 1), (-2, 1), (-3, 1)]
TZVP_H = [(0, 3), (0, 1), (0, 1), (1, 1)]


_BINOM = np.array([[math.comb(n, k) for k in range(MAX_L + 1)] for n in range(MAX_L + 1)], dtype=float)


def _dfact(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2, dtype=np.float64)))


def _norm(alpha, p):
    l = sum(p)
    return (2 * alpha / math.pi) ** 0.75 * math.sqrt((4 * alpha) ** l / (_dfact(2 * p[0] - 1) * _dfact(2 * p[1] - 1) *
                                                                       _dfact(2 * p[2] - 1)))


def _pure_transform(l):
    """Schlegel-Frisch Cartesian->pure matrix, rows m = 0, +1, -1, ..."""
    f = math.factorial
    cart = _CART[l]

    def coef(m, lx, ly):
        lz = l - lx - ly
        ma = abs(m)
        if ma > l or lz < 0:
            return 0.0
        a = math.sqrt(f(2 * lx) * f(2 * ly) * f(2 * lz) * f(l) / (f(2 * l) * f(lx) * f(ly) * f(lz)))
        b = math.sqrt(f(l - ma) / f(l + ma)) / (2.0 ** l * f(l))
        j2 = l - ma - lz
        if j2 % 2:
            return 0.0
        j = j2 // 2
        g = 0.0
        for i in range((l - ma) // 2 + 1):
            if j < 0 or j > i:
                continue
            c = f(l) / (f(i) * f(l - i)) * f(2 * l - 2 * i) * (-1) ** i / f(l - ma - 2 * i) * f(i) / (f(j) * f(i - j))
            for k in range(j + 1):
                px = lx - 2 * k
                if px < 0 or px > ma:
                    continue
                d = f(j) / (f(k) * f(j - k)) * f(ma) / (f(px) * f(ma + 2 * k - lx))
                e = 0.0
                kt = (2 * k + ma - lx) // 2
                if m == 0 and lx % 2 == 0:
                    e = (-1.0) ** (k - lx // 2)
                if m > 0 and abs(ma - lx) % 2 == 0:
                    e = math.sqrt(2.0) * (-1.0) ** kt
                if m < 0 and abs(ma - lx) % 2 == 1:
                    e = math.sqrt(2.0) * (-1.0) ** kt
                g += c * d * e
        return a * b * g
    rows = [0] + [x for m in range(1, l + 1) for x in (m, -m)]
    return np.array([[coef(m, p[0], p[1]) for p in cart] for m in rows])


class Basis:
    """Normalized AO blocks: per block centre, l, exponents, weights
    (ncart x nprim) and pure transform (or None)."""

    def __init__(self, shell_types, nprim, shell_xyz, exps, coefs, pcoefs):
        self.blocks = []
        off = 0
        for s, t in enumerate(shell_types):
            sl = slice(off, off + nprim[s])
            off += nprim[s]
            parts = [(0, coefs[sl], False), (1, pcoefs[sl], False)] if t == -1 else [(abs(t), coefs[sl], t <= -2)]
            for l, c, pure in parts:
                cart = _CART[l]
                e = np.asarray(exps[sl])
                w = np.array([[c[p] * _norm(e[p], q) for p in range(len(e))] for q in cart])
                self.blocks.append((np.asarray(shell_xyz[s]), l, e, w, _pure_transform(l) if pure and l >= 2 else None))
        self.nao = sum(b[4].shape[0] if b[4] is not None else len(_CART[b[1]]) for b in self.blocks)

    def overlap(self):
        """Analytic AO overlap (the expansion of src/basis_evaluator.cpp:210-278),
        vectorised over primitive pairs."""
        # Expand to Cartesian primitive functions.
        cen, ex, pw, wt, ao = [], [], [], [], []
        cart_index = 0
        cart_of_block = []
        for (o, l, e, w, T) in self.blocks:
            cart_of_block.append(cart_index)
            for k, q in enumerate(_CART[l]):
                for p in range(len(e)):
                    cen.append(o)
                    ex.append(e[p])
                    pw.append(q)
                    wt.append(w[k, p])
                    ao.append(cart_index + k)
            cart_index += len(_CART[l])
        cen, ex, pw, wt, ao = map(np.asarray, (cen, ex, pw, wt, ao))
        # Primitive pairs a <= b whose Gaussian product prefactor can matter
        # (exp(-mu R^2) >= 1e-24; dropped pairs change S by far less than its
        # last bit), evaluated as compressed arrays.
        A, B = np.triu_indices(len(ex))
        a, b = ex[A], ex[B]
        p = a + b
        r2 = ((cen[A] - cen[B]) ** 2).sum(1)
        keep = a * b / p * r2 <= 55.3
        A, B, a, b, p = A[keep], B[keep], a[keep], b[keep], p[keep]
        s = np.ones(A.shape)
        for axis in range(3):
            xa, xb = cen[A, axis], cen[B, axis]
            P = (a * xa + b * xb) / p
            pa, pb = P - xa, P - xb
            na, nb = pw[A, axis], pw[B, axis]
            pref = np.exp(-a * b / p * (xa - xb) ** 2) * np.sqrt(math.pi / p)
            pap = [pa ** q for q in range(MAX_L + 1)]
            pbp = [pb ** q for q in range(MAX_L + 1)]
            tot = np.zeros(A.shape)
            for i in range(MAX_L + 1):
                for j in range(MAX_L + 1):
                    k = i + j
                    if k % 2:
                        continue
                    sel = (na >= i) & (nb >= j)
                    if not sel.any():
                        continue
                    mk = _dfact(k - 1) / (2 * p[sel]) ** (k // 2)
                    ci = _BINOM[na[sel], i]
                    cj = _BINOM[nb[sel], j]
                    ea, eb = na[sel] - i, nb[sel] - j
                    fa = np.choose(ea, [pp[sel] for pp in pap])
                    fb = np.choose(eb, [pp[sel] for pp in pbp])
                    tot[sel] += ci * cj * fa * fb * mk
            s *= pref * tot
        S = np.zeros((len(ex), len(ex)))
        S[A, B] = s
        S[B, A] = s
        s = S
        ncart = cart_index
        W = np.zeros((ncart, len(ex)))
        W[ao, np.arange(len(ex))] = wt
        Sc = W @ (s @ W.T)
        # Cartesian -> AO (pure) transform
        rows = []
        for (o, l, e, w, T), c0 in zip(self.blocks, cart_of_block):
            nc = len(_CART[l])
            if T is None:
                r = np.zeros((nc, ncart))
                r[:, c0:c0 + nc] = np.eye(nc)
            else:
                r = np.zeros((T.shape[0], ncart))
                r[:, c0:c0 + nc] = T
            rows.append(r)
        M = np.vstack(rows)
        return M @ Sc @ M.T


class SyntheticMolecule:
    """Arrays of an FCHK (names as tests/support/fchk.py FchkData)."""

    def __init__(self, rng, z, xyz, recipe, nalpha, nbeta, unrestricted):
        shell_types, nprim, shell_atom, sxyz, exps, coefs, pcoefs = [], [], [], [], [], [], []
        for a, (za, pos) in enumerate(zip(z, xyz)):
            for t, n in recipe(za):
                e = np.sort(np.exp(rng.uniform(math.log(0.1), math.log(5000.0), n)))[::-1]
                c = rng.uniform(0.1, 1.0, n)
                pc = rng.uniform(0.1, 1.0, n) if t == -1 else np.zeros(n)
                shell_types.append(t)
                nprim.append(n)
                shell_atom.append(a + 1)
                sxyz.append(pos)
                exps.extend(e)
                coefs.extend(c)
                pcoefs.extend(pc)
        self.shell_types = np.array(shell_types, np.int32)
        self.nprim = np.array(nprim, np.int32)
        self.shell_atom = np.array(shell_atom, np.int32)
        self.shell_xyz = np.array(sxyz, dtype=float)
        self.exponents = np.array(exps)
        self.coefficients = np.array(coefs)
        self.p_coefficients = np.array(pcoefs) if (self.shell_types == -1).any() else None
        # Normalise each contracted shell: scale its coefficients so that the
        # diagonal of S is one.
        basis = Basis(self.shell_types, self.nprim, self.shell_xyz, self.exponents, self.coefficients,
                      self.p_coefficients if self.p_coefficients is not None else np.zeros_like(self.exponents))
        S = basis.overlap()
        off = 0
        ao = 0
        scale = np.ones(S.shape[0])  # per AO: the factor its shell's coefficients are scaled by
        for s, t in enumerate(self.shell_types):
            sl = slice(off, off + self.nprim[s])
            off += self.nprim[s]
            if t == -1:
                fs, fp = 1.0 / math.sqrt(S[ao, ao]), 1.0 / math.sqrt(S[ao + 1, ao + 1])
                self.coefficients[sl] *= fs
                self.p_coefficients[sl] *= fp
                scale[ao] = fs
                scale[ao + 1:ao + 4] = fp
                ao += 4
            else:
                l = abs(t)
                cnt = 2 * l + 1 if t <= -2 else len(_CART[l])
                # Cartesian d/f components are not all of unit norm; scale by the first (x^l) one.
                f = 1.0 / math.sqrt(S[ao, ao])
                self.coefficients[sl] *= f
                scale[ao:ao + cnt] = f
                ao += cnt
        # the overlap of the rescaled functions (every AO of a shell scales alike)
        S = S * scale[:, None] * scale[None, :]
        self.overlap = S
        self.nbasis = basis.nao
        self.atomic_numbers = np.asarray(z, np.int32)
        self.charges = self.atomic_numbers.astype(float)
        self.xyz = np.asarray(xyz, dtype=float)
        self.nalpha, self.nbeta = nalpha, nbeta
        n = self.nbasis
        # S^-1/2-type orthonormalisation of random coefficients: C (C^T S C)^-1/2.
        # Random exponents on atoms 1.8 bohr apart can make the basis nearly
        # linearly dependent; an SCF program removes such combinations (Gaussian
        # drops overlap eigenvalues below 1e-6), so no FCHK carries one, and the
        # caller redraws (ValueError) when the smallest eigenvalue is below 1e-6.
        w, V = np.linalg.eigh(S)
        if w.min() < 1e-6:
            raise ValueError("synthetic basis is numerically linearly dependent")

        def orbitals():
            C = rng.normal(size=(n, n))
            M = C.T @ S @ C
            mw, mv = np.linalg.eigh(M)
            return C @ (mv @ np.diag(mw ** -0.5) @ mv.T)
        Ca = orbitals()
        Cb = orbitals() if unrestricted else Ca
        Pa = Ca[:, :nalpha] @ Ca[:, :nalpha].T
        Pb = Cb[:, :nbeta] @ Cb[:, :nbeta].T
        self.p_total = Pa + Pb
        self.p_spin = Pa - Pb
        self.c_alpha = Ca.T.ravel().copy()  # FCHK: MO-major, AO fastest
        self.c_beta = Cb.T.ravel().copy() if unrestricted else None
        self.e_alpha = np.sort(rng.uniform(-20.0, 2.0, n))
        self.e_beta = np.sort(rng.uniform(-20.0, 2.0, n)) if unrestricted else None

    def write_fchk(self, path, title="synthetic molecule"):
        def ints(label, v):
            v = np.asarray(v).ravel()
            out = [f"{label:<43}I   N={len(v):>12d}"]
            for i in range(0, len(v), 6):
                out.append("".join(f"{int(x):12d}" for x in v[i:i + 6]))
            return out

        def reals(label, v):
            v = np.asarray(v, dtype=float).ravel()
            out = [f"{label:<43}R   N={len(v):>12d}"]
            for i in range(0, len(v), 5):
                out.append("".join(f"{x:24.16E}" for x in v[i:i + 5]))
            return out

        def scalar(label, v):
            return [f"{label:<43}I     {int(v):>12d}"]
        tri = np.tril_indices(self.nbasis)
        lines = [title, "SP        RB3LYP                                                      synthetic"]
        lines += scalar("Number of atoms", len(self.atomic_numbers))
        lines += scalar("Number of electrons", self.nalpha + self.nbeta)
        lines += scalar("Number of alpha electrons", self.nalpha)
        lines += scalar("Number of beta electrons", self.nbeta)
        lines += scalar("Number of basis functions", self.nbasis)
        lines += ints("Atomic numbers", self.atomic_numbers)
        lines += reals("Nuclear charges", self.charges)
        lines += reals("Current cartesian coordinates", self.xyz)
        lines += ints("Shell types", self.shell_types)
        lines += ints("Number of primitives per shell", self.nprim)
        lines += ints("Shell to atom map", self.shell_atom)
        lines += reals("Primitive exponents", self.exponents)
        lines += reals("Contraction coefficients", self.coefficients)
        if self.p_coefficients is not None:
            lines += reals("P(S=P) Contraction coefficients", self.p_coefficients)
        lines += reals("Coordinates of each shell", self.shell_xyz)
        lines += reals("Alpha Orbital Energies", self.e_alpha)
        if self.e_beta is not None:
            lines += reals("Beta Orbital Energies", self.e_beta)
        lines += reals("Alpha MO coefficients", self.c_alpha)
        if self.c_beta is not None:
            lines += reals("Beta MO coefficients", self.c_beta)
        lines += reals("Total SCF Density", self.p_total[tri])
        lines += reals("Spin SCF Density", self.p_spin[tri])
        with open(path, "w") as fh:
            fh.write("\n".join(lines) + "\n")
        return path


def _positions(rng, natoms, side, min_dist, first=None):
    pos = [] if first is None else [np.asarray(first, float)]
    tries = 0
    while len(pos) < natoms:
        p = rng.uniform(-side / 2, side / 2, 3)
        tries += 1
        if tries > 200000:
            raise RuntimeError("cannot place atoms")
        if all(np.linalg.norm(p - q) >= min_dist for q in pos):
            pos.append(p)
    return np.array(pos)


def grid_around(xyz, margin, spacing):
    """Grid::around (src/grid.cpp:15-38)."""
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    counts = [max(2, int(math.ceil((hi[k] - lo[k] + 2 * margin) / spacing) + 1)) for k in range(3)]
    return {"origin": lo - margin, "steps": np.eye(3) * spacing, "counts": np.array(counts, np.int32)}


def cube_grid(center, edge, spacing):
    half = (edge - 1) * spacing / 2.0
    return {"origin": np.asarray(center, float) - half, "steps": np.eye(3) * spacing,
            "counts": np.array([edge, edge, edge], np.int32)}


def water_like(seed=0):
    """Configuration 0: O at the origin, two H at 1.7-1.9 bohr in random
    directions; 6-31G*-shaped Cartesian basis (19 AOs); 5 doubly occupied MOs;
    grid around the atoms with a 3 bohr margin at 0.1 bohr."""
    rng = np.random.default_rng(seed)
    xyz = [np.zeros(3)]
    for _ in range(2):
        d = rng.normal(size=3)
        xyz.append(d / np.linalg.norm(d) * rng.uniform(1.7, 1.9))
    xyz = np.array(xyz)
    mol = SyntheticMolecule(rng, [8, 1, 1], xyz, lambda z: POPLE_HEAVY if z > 1 else POPLE_H, 5, 5, False)
    return mol, grid_around(xyz, 3.0, 0.1)


def _element_mix(rng, nheavy, nh):
    return list(rng.choice([6, 7, 8], size=nheavy)) + [1] * nh


def tzvp_cube(seed=1, natoms=40, nheavy=16, nalpha=60, nbeta=59, edge=160, spacing=0.15):
    """Configuration 1: 40 atoms (16 from C, N, O + 24 H) at least 1.8 bohr
    apart in a 16 bohr cube; def2-TZVP-shaped pure shells (640 AOs); UHF
    60 alpha / 59 beta; a 160^3 grid at 0.15 bohr centred on the atoms."""
    rng = np.random.default_rng(seed)
    z = _element_mix(rng, nheavy, natoms - nheavy)
    xyz = _positions(rng, natoms, 16.0, 1.8)
    mol = SyntheticMolecule(rng, z, xyz, lambda zz: TZVP_HEAVY if zz > 1 else TZVP_H, nalpha, nbeta, True)
    return mol, cube_grid(xyz.mean(axis=0) * 0 + (xyz.min(0) + xyz.max(0)) / 2, edge, spacing)


def batch_molecule(seed, natoms, family, edge, spacing=0.2):
    """One molecule of configuration 2: `natoms` atoms (40 % heavy), 6-31G*-
    or TZVP-shaped basis, neutral occupations (ceil/floor of N_e / 2), a cube
    grid of `edge` points per axis at 0.2 bohr."""
    rng = np.random.default_rng(seed)
    nheavy = max(1, round(0.4 * natoms))
    z = _element_mix(rng, nheavy, natoms - nheavy)
    side = (102.4 * natoms) ** (1 / 3)
    xyz = _positions(rng, natoms, max(side, 3.0), 1.8)
    ne = int(sum(z))
    na, nb = (ne + 1) // 2, ne // 2
    recipe = (lambda zz: TZVP_HEAVY if zz > 1 else TZVP_H) if family == "tzvp" else \
        (lambda zz: POPLE_HEAVY if zz > 1 else POPLE_H)
    mol = SyntheticMolecule(rng, z, xyz, recipe, na, nb, na != nb)
    return mol, cube_grid((xyz.min(0) + xyz.max(0)) / 2, edge, spacing)


# Shells up to h: pure and Cartesian g and h beside s, p, SP and d shells.
HIGH_L_A = [(0, 4), (1, 2), (-2, 1), (-4, 2), (5, 1)]      # pure g, Cartesian h
HIGH_L_B = [(0, 3), (-1, 2), (4, 1), (-5, 1)]               # Cartesian g, pure h
HIGH_L_H = [(0, 2), (1, 1)]


def high_l_molecule(seed=7):
    """Three atoms (O, C, H) 2.2-2.6 bohr apart with g and h shells, pure and
    Cartesian (74 AOs), unrestricted 8 alpha / 7 beta, and a grid around them
    at 0.3 bohr with a 2.5 bohr margin."""
    while True:
        rng = np.random.default_rng(seed)
        xyz = np.array([[0.0, 0.0, 0.0], [2.4, 0.3, -0.2], [-0.9, 2.1, 0.6]]) + rng.normal(size=(3, 3)) * 0.1
        recipes = {8: HIGH_L_A, 6: HIGH_L_B, 1: HIGH_L_H}
        try:
            mol = SyntheticMolecule(rng, [8, 6, 1], xyz, lambda zz: recipes[zz], 8, 7, True)
            return mol, grid_around(xyz, 2.5, 0.3)
        except ValueError:
            seed += 1000


def no_beta_molecule(seed=5):
    """Two H atoms 1.4 bohr apart, def2-TZVP-shaped (3s1p each), unrestricted
    with two alpha and no beta electrons (a triplet): the beta channel is
    empty."""
    rng = np.random.default_rng(seed)
    xyz = np.array([[0.0, 0.0, 0.0], [0.3, 0.2, 1.4]])
    mol = SyntheticMolecule(rng, [1, 1], xyz, lambda zz: TZVP_H, 2, 0, True)
    return mol, grid_around(xyz, 3.0, 0.25)


def cached_fchk(mol_fn, name, directory=None):
    """Writes the molecule's FCHK once under `directory` and returns its path."""
    directory = directory or os.path.join(os.environ.get("TMPDIR", "/tmp"), "kumo_hip_synthetic")
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, name + ".fchk")
    mol, grid = mol_fn()
    if not os.path.exists(path):
        mol.write_fchk(path + ".tmp")
        os.replace(path + ".tmp", path)
    return path, mol, grid


# Bondi van der Waals radii (angstrom), J. Phys. Chem. 68, 441 (1964).
BONDI = {1: 1.20, 6: 1.70, 7: 1.55, 8: 1.52}
BOHR_PER_ANGSTROM = 1.8897261246257702  # src/cube.cpp:32


def pople_molecule(seed=3, natoms=20, nheavy=12):
    """Configuration 3 molecule: `natoms` atoms (12 C/N/O + 8 H) at least
    1.8 bohr apart, 6-31G*-shaped Cartesian shells (~200 AOs), closed shell."""
    rng = np.random.default_rng(seed)
    z = _element_mix(rng, nheavy, natoms - nheavy)
    xyz = _positions(rng, natoms, (102.4 * natoms) ** (1 / 3), 1.8)
    ne = int(sum(z))
    if ne % 2:
        z[-1] = 1 if z[-1] != 1 else 1
        ne = int(sum(z))
    na = ne // 2
    mol = SyntheticMolecule(rng, z, xyz, lambda zz: POPLE_HEAVY if zz > 1 else POPLE_H, na, ne - na, ne % 2 == 1)
    return mol


def vdw_shell_points(z, xyz, npoints, scale=1.4, seed=11):
    """Points uniformly on the union of spheres of radius scale x r_vdW around
    the atoms, rejecting points inside any other sphere (a molecular surface
    for the electrostatic potential)."""
    rng = np.random.default_rng(seed)
    radii = np.array([BONDI.get(int(a), 1.7) * BOHR_PER_ANGSTROM * scale for a in z])
    area = 4 * np.pi * radii ** 2
    out = []
    while sum(len(o) for o in out) < npoints:
        k = rng.choice(len(z), size=4 * npoints, p=area / area.sum())
        d = rng.normal(size=(4 * npoints, 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        p = xyz[k] + d * radii[k][:, None]
        dist = np.linalg.norm(p[:, None, :] - xyz[None, :, :], axis=2)
        inside = (dist < radii[None, :] * (1 - 1e-12)).any(axis=1)
        out.append(p[~inside])
    return np.concatenate(out)[:npoints]
