"""Seeded synthetic molecules, primitive bases and HF-like DM2 streams.

The generators write the reference's own input formats (.fchk subset,
.bas, .dm2p), and can also hand records straight to device memory for the
benchmark.  Definitions:

  * Molecule: atoms grown one at a time, each bonded to a random earlier
    atom at a bond length drawn from U[bond_lo, bond_hi] bohr and kept at
    least bond_lo from every other atom.
  * Basis: Cartesian primitives per atom (s, p = x,y,z, d = xx,yy,zz,xy,xz,yz,
    the component order of the reference's .fchk reader,
    src/mod_read_files.f90:556-571), exponents log-uniform in [lo, hi].
  * Orbitals: nocc seeded random vectors in the normalised-primitive basis,
    orthonormalised in the overlap metric S; P = 2 C C^T.
  * DM2 in the file convention of the reference (electron 1 on the (i,k)
    pair, electron 2 on (j,l); src/mod_intracule.f90:345-358):
        D(i,j,k,l) = P_ik P_jl - 1/2 P_il P_jk,
    so that sum_q D_q S_ik S_jl = N(N-1) with N = 2 nocc.
  * "unique" streams keep one record per class of the 8 index images that
    give the same f_q (i<->k, j<->l, (ik)<->(jl)) and store the sum of D
    over the distinct images, so the sum rule is unchanged.
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial.hermite import hermgauss

ZNUM = {"H": 1, "C": 6, "N": 7, "O": 8}
SHELLS = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
}


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def normprim(alpha, t):
    L = sum(t)
    df = dfact(2 * t[0] - 1) * dfact(2 * t[1] - 1) * dfact(2 * t[2] - 1)
    return (2 * alpha / np.pi) ** 0.75 * np.sqrt((4 * alpha) ** L / df)


def grow_molecule(rng, elements, bond_lo=1.8, bond_hi=2.9):
    pos = [np.zeros(3)]
    for _ in range(1, len(elements)):
        for _try in range(10000):
            anchor = pos[rng.integers(len(pos))]
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            p = anchor + v * rng.uniform(bond_lo, bond_hi)
            if min(np.linalg.norm(p - q) for q in pos) >= bond_lo - 1e-12:
                pos.append(p)
                break
        else:
            raise RuntimeError("could not place atom")
    return np.array(pos)


def make_basis(rng, elements, coords, shells_heavy, shells_h, exp_lo, exp_hi):
    alpha, tmn, cen, ra = [], [], [], []
    for a, (el, c) in enumerate(zip(elements, coords)):
        for L, cnt in (shells_h if el == "H" else shells_heavy).items():
            for _ in range(cnt):
                e = float(np.exp(rng.uniform(np.log(exp_lo), np.log(exp_hi))))
                for t in SHELLS[L]:
                    alpha.append(e)
                    tmn.append(t)
                    cen.append(c)
                    ra.append(a + 1)
    return (np.array(alpha), np.array(tmn, dtype=np.int32), np.array(cen),
            np.array(ra, dtype=np.int32))


def overlap_matrix(alpha, tmn, cen):
    """Overlap of normalised Cartesian primitives (Gauss-Hermite per axis)."""
    n = len(alpha)
    ui, wi = hermgauss(8)  # exact for total degree <= 15
    a = alpha[:, None]
    b = alpha[None, :]
    p = a + b
    S = np.ones((n, n))
    for ax in range(3):
        A = cen[:, ax][:, None]
        B = cen[:, ax][None, :]
        P = (a * A + b * B) / p
        pref = np.exp(-a * b / p * (A - B) ** 2) / np.sqrt(p)
        ta = tmn[:, ax][:, None]
        tb = tmn[:, ax][None, :]
        acc = np.zeros((n, n))
        for u, w in zip(ui, wi):
            x = P + u / np.sqrt(p)
            acc += w * (x - A) ** ta * (x - B) ** tb
        S *= pref * acc
    N = np.array([normprim(al, t) for al, t in zip(alpha, tmn)])
    return S * N[:, None] * N[None, :]


def orbitals(rng, S, nocc):
    n = S.shape[0]
    C = rng.normal(size=(n, nocc))
    M = C.T @ S @ C
    w, V = np.linalg.eigh(M)
    return C @ (V @ np.diag(w ** -0.5) @ V.T)


class System:
    """A synthetic molecule with its basis, orbitals and density matrix."""

    def __init__(self, elements, coords, alpha, tmn, cen, ra, C):
        self.elements = list(elements)
        self.an = np.array([ZNUM[e] for e in elements], dtype=np.int32)
        self.coords = np.asarray(coords, dtype=np.float64)
        self.alpha, self.tmn, self.cen, self.ra = alpha, tmn, cen, ra
        self.C = C
        self.P = 2.0 * C @ C.T
        self.nelec = 2 * C.shape[1]
        self.S = overlap_matrix(alpha, tmn, cen)

    @property
    def nprim(self):
        return len(self.alpha)

    def prims(self):
        return dict(alpha=self.alpha, tmn=self.tmn, x=self.cen[:, 0].copy(),
                    y=self.cen[:, 1].copy(), z=self.cen[:, 2].copy(), an=self.an,
                    cartes=self.coords, nelec=self.nelec)

    # ---- DM2 records -------------------------------------------------------
    def dm2_full(self):
        """All nprim^4 records, i fastest last: (i,j,k,l) with l innermost."""
        n = self.nprim
        P = self.P
        idx = np.indices((n, n, n, n)).reshape(4, -1).T
        i, j, k, l = idx.T
        d = P[i, k] * P[j, l] - 0.5 * P[i, l] * P[j, k]
        return np.ascontiguousarray(idx + 1, dtype=np.int32), d

    def unique_pairs(self):
        n = self.nprim
        i, k = np.triu_indices(n)
        return i, k

    def dm2_unique_torch(self, device="cuda", chunk=1 << 24):
        """Symmetry-unique records as a (Q, 6) int32 torch tensor laid out as
        24-byte .dm2p records (i, j, k, l, D as two words), on `device`."""
        import torch
        n = self.nprim
        pi_, pk_ = self.unique_pairs()
        npair = len(pi_)
        P = torch.tensor(self.P, dtype=torch.float64, device=device)
        pi_t = torch.tensor(pi_, dtype=torch.int64, device=device)
        pk_t = torch.tensor(pk_, dtype=torch.int64, device=device)
        # pair index p <= q, p outer
        counts = npair - torch.arange(npair, device=device)
        Q = int(counts.sum().item())
        out = torch.empty((Q, 6), dtype=torch.int32, device=device)
        starts = torch.cumsum(counts, 0) - counts
        pos = 0
        p0 = 0
        while p0 < npair:
            # take a block of p rows with about `chunk` records
            p1 = p0
            tot = 0
            while p1 < npair and tot + (npair - p1) <= max(chunk, npair - p0):
                tot += npair - p1
                p1 += 1
            p = torch.repeat_interleave(torch.arange(p0, p1, device=device),
                                        counts[p0:p1])
            q = torch.arange(tot, device=device) - (starts[p] - starts[p0]) + p
            i, k = pi_t[p], pk_t[p]
            j, l = pi_t[q], pk_t[q]
            d = _folded(P, i, j, k, l)
            blk = out[pos:pos + tot]
            blk[:, 0] = (i + 1).to(torch.int32)
            blk[:, 1] = (j + 1).to(torch.int32)
            blk[:, 2] = (k + 1).to(torch.int32)
            blk[:, 3] = (l + 1).to(torch.int32)
            blk[:, 4:6] = d.contiguous().view(torch.int32).view(-1, 2)
            pos += tot
            p0 = p1
        assert pos == Q
        return out

    def dm2_unique(self):
        """Symmetry-unique records on the host: (ijkl int32 (Q,4), D (Q,))."""
        import torch
        t = self.dm2_unique_torch(device="cpu")
        ijkl = t[:, :4].numpy().copy()
        d = t[:, 4:6].contiguous().view(torch.float64).reshape(-1).numpy().copy()
        return ijkl, d

    def sum_rule(self, ijkl, d):
        i, j, k, l = (ijkl[:, c] - 1 for c in range(4))
        return float(np.sum(d * self.S[i, k] * self.S[j, l]))

    # ---- files ----------------------------------------------------------------
    def write_bas(self, path):
        with open(path, "w") as f:
            f.write("# Z              X_A             Y_A             Z_A"
                    "             Exp             Quant(nx,ny,nz)\n")
            for a, t, c, r in zip(self.alpha, self.tmn, self.cen, self.ra):
                f.write("%3d %22.15E %22.15E %22.15E %22.15E %5d %5d %5d\n" % (
                    self.an[r - 1], c[0], c[1], c[2], a, t[0], t[1], t[2]))

    def write_fchk(self, path, title="synthetic"):
        """Minimal .fchk with the sections the reference's reader uses
        (src/mod_read_files.f90:449-539): one shell per primitive set."""
        shells, pps, s2a, pexp = [], [], [], []
        seen = 0
        n = self.nprim
        while seen < n:
            t = tuple(self.tmn[seen])
            L = sum(t)
            shells.append(L)
            pps.append(1)
            s2a.append(int(self.ra[seen]))
            pexp.append(self.alpha[seen])
            seen += len(SHELLS[L])
        na = self.nelec // 2

        def ivec(name, v):
            s = "%-40s   I   N=%12d\n" % (name, len(v))
            for k in range(0, len(v), 6):
                s += "".join("%12d" % x for x in v[k:k + 6]) + "\n"
            return s

        def rvec(name, v):
            s = "%-40s   R   N=%12d\n" % (name, len(v))
            for k in range(0, len(v), 5):
                s += "".join("%16.8E" % x for x in v[k:k + 5]) + "\n"
            return s

        with open(path, "w") as f:
            f.write("%s\n" % title)
            f.write("SP        RHF                                                         STO-3G\n")
            f.write("%-40s   I     %12d\n" % ("Number of atoms", len(self.an)))
            f.write("%-40s   I     %12d\n" % ("Number of electrons", self.nelec))
            f.write("%-40s   I     %12d\n" % ("Number of alpha electrons", na))
            f.write("%-40s   I     %12d\n" % ("Number of beta electrons", na))
            f.write(ivec("Atomic numbers", list(self.an)))
            f.write(rvec("Nuclear charges", list(self.an.astype(float))))
            f.write(rvec("Current cartesian coordinates", list(self.coords.reshape(-1))))
            f.write("%-40s   I     %12d\n" % ("Number of contracted shells", len(shells)))
            f.write("%-40s   I     %12d\n" % ("Number of primitive shells", len(shells)))
            f.write(ivec("Shell types", shells))
            f.write(ivec("Number of primitives per shell", pps))
            f.write(ivec("Shell to atom map", s2a))
            f.write(rvec("Primitive exponents", pexp))
            f.write("%-40s   R     %22.15E\n" % ("Total Energy", -1.0))


WFX_TYPES = {(0, 0, 0): 1, (1, 0, 0): 2, (0, 1, 0): 3, (0, 0, 1): 4, (2, 0, 0): 5,
             (0, 2, 0): 6, (0, 0, 2): 7, (1, 1, 0): 8, (1, 0, 1): 9, (0, 1, 1): 10}


def write_wfx(system, path, title="synthetic"):
    """AIMAll .wfx with the sections the reference's reader uses
    (src/mod_read_files.f90:12-287): primitives on their atoms, closed-shell
    occupations, MO coefficients of the unnormalised primitives."""
    s = system
    n = s.nprim
    nocc = s.nelec // 2
    norm = np.array([normprim(a, t) for a, t in zip(s.alpha, s.tmn)])

    def block(tag, lines):
        return "<%s>\n%s\n</%s>\n" % (tag, "\n".join(lines), tag)

    def rows(vals, fmt, per):
        return [" ".join(fmt % v for v in vals[k:k + per]) for k in range(0, len(vals), per)]

    out = [block("Title", [" " + title]), block("Keywords", [" GTO"]),
           block("Number of Nuclei", [" %d" % len(s.an)]),
           block("Number of Occupied Molecular Orbitals", [" %d" % nocc]),
           block("Number of Perturbations", [" 0"]), block("Net Charge", [" 0"]),
           block("Number of Electrons", [" %d" % s.nelec]),
           block("Number of Alpha Electrons", [" %d" % nocc]),
           block("Number of Beta Electrons", [" %d" % nocc]),
           block("Electronic Spin Multiplicity", [" 1"]),
           block("Nuclear Names", [" %s%d" % (e, i + 1) for i, e in enumerate(s.elements)]),
           block("Atomic Numbers", [" %d" % z for z in s.an]),
           block("Nuclear Charges", [" %.15E" % float(z) for z in s.an]),
           block("Nuclear Cartesian Coordinates",
                 [" %.15E %.15E %.15E" % tuple(c) for c in s.coords]),
           block("Number of Primitives", [" %d" % n]),
           block("Primitive Centers", rows(list(s.ra), " %d", 10)),
           block("Primitive Types", rows([WFX_TYPES[tuple(t)] for t in s.tmn], " %d", 10)),
           block("Primitive Exponents", rows(list(s.alpha), " %.15E", 4)),
           block("Molecular Orbital Occupation Numbers", [" %.15E" % 2.0] * nocc),
           block("Molecular Orbital Energies", [" %.15E" % (-1.0)] * nocc)]
    mo = ["<Molecular Orbital Primitive Coefficients>"]
    for k in range(nocc):
        mo += ["<MO Number>", " %d" % (k + 1), "</MO Number>"]
        mo += rows(list(s.C[:, k] * norm), " %.15E", 4)
    mo.append("</Molecular Orbital Primitive Coefficients>")
    out.append("\n".join(mo) + "\n")
    with open(path, "w") as f:
        f.write("".join(out))


def write_dm2p(path, ijkl, d):
    """Flat .dm2p stream: 24-byte records (int32 i, j, k, l; float64 D)."""
    rec = np.empty(len(d), dtype=[("ijkl", "<i4", 4), ("d", "<f8")])
    rec["ijkl"], rec["d"] = ijkl, d
    rec.tofile(path)


def _folded(P, i, j, k, l):
    """Sum of D over the distinct images of (i,j,k,l) under i<->k, j<->l and
    the electron swap (ik)<->(jl)."""
    import torch

    def D(a, b, c, d):
        return P[a, c] * P[b, d] - 0.5 * P[a, d] * P[b, c]

    imgs = [(i, j, k, l), (k, j, i, l), (i, l, k, j), (k, l, i, j),
            (j, i, l, k), (l, i, j, k), (j, k, l, i), (l, k, j, i)]
    total = torch.zeros_like(P[i, k])
    keys = []
    for a, b, c, d in imgs:
        key = ((a * 4096 + b) * 4096 + c) * 4096 + d
        dup = torch.zeros_like(key, dtype=torch.bool)
        for kk in keys:
            dup |= (kk == key)
        total += torch.where(dup, torch.zeros_like(total), D(a, b, c, d))
        keys.append(key)
    return total


def config_ch4(seed=1):
    """CH4 geometry with 27 seeded primitives (C: 6 s + 3 p-sets, H: 3 s)."""
    rng = np.random.default_rng(seed)
    h = 1.1967636
    coords = np.array([[0, 0, 0], [h, h, h], [-h, -h, h], [-h, h, -h], [h, -h, -h]], float)
    elements = ["C", "H", "H", "H", "H"]
    alpha, tmn, cen, ra = make_basis(rng, elements, coords, {0: 6, 1: 3}, {0: 3}, 0.1, 100.0)
    C = orbitals(rng, overlap_matrix(alpha, tmn, cen), 5)
    return System(elements, coords, alpha, tmn, cen, ra, C)


def config_cluster(seed=2, natoms=10, nheavy=7, nocc=20, exp_lo=0.05, exp_hi=500.0,
                   shells_heavy=None, shells_h=None):
    """Seeded C/N/O/H cluster with s/p/d primitives (about 150 for the
    defaults: heavy atoms 4 s + 3 p + 1 d sets, H 3 s + 1 p)."""
    rng = np.random.default_rng(seed)
    heavy = list(rng.choice(["C", "N", "O"], size=nheavy))
    elements = heavy + ["H"] * (natoms - nheavy)
    coords = grow_molecule(rng, elements)
    shells_heavy = shells_heavy or {0: 4, 1: 3, 2: 1}
    shells_h = shells_h or {0: 3, 1: 1}
    alpha, tmn, cen, ra = make_basis(rng, elements, coords, shells_heavy, shells_h,
                                     exp_lo, exp_hi)
    S = overlap_matrix(alpha, tmn, cen)
    C = orbitals(rng, S, nocc)
    return System(elements, coords, alpha, tmn, cen, ra, C)


CONFIG4_SEEDS = range(1000, 1256)
CONFIG4_NRAD = (50, 75, 100, 150)
CONFIG4_NANG = (110, 194, 302, 590)


def _config4_layout(seed):
    """Atoms, geometry, basis and grid orders of a batch molecule (the rng
    draws before the orbitals')."""
    rng = np.random.default_rng(seed)
    nat = int(rng.integers(4, 21))
    nheavy = max(1, min(nat, int(round(nat * rng.uniform(0.4, 0.8)))))
    elements = list(rng.choice(["C", "N", "O"], size=nheavy)) + ["H"] * (nat - nheavy)
    nelec = sum(ZNUM[e] for e in elements)
    nocc = nelec // 2
    target = int(rng.integers(max(30, nocc + 4, nat), 301))
    nrad = int(rng.choice(CONFIG4_NRAD))
    nang = int(rng.choice(CONFIG4_NANG))
    coords = grow_molecule(rng, elements)
    # shells: one s per atom, then random s/p/d (heavy) or s/p (H) shells
    shells = [[0] for _ in elements]
    count = nat
    while count < target:
        a = int(rng.integers(nat))
        L = int(rng.choice([0, 1, 2])) if elements[a] != "H" else int(rng.choice([0, 1]))
        if count + len(SHELLS[L]) > 300:
            L = 0
        shells[a].append(L)
        count += len(SHELLS[L])
    alpha, tmn, cen, ra = [], [], [], []
    for a, (el, c) in enumerate(zip(elements, coords)):
        for L in shells[a]:
            e = float(np.exp(rng.uniform(np.log(0.05), np.log(500.0))))
            for t in SHELLS[L]:
                alpha.append(e)
                tmn.append(t)
                cen.append(c)
                ra.append(a + 1)
    return (rng, elements, coords, np.array(alpha), np.array(tmn, dtype=np.int32),
            np.array(cen), np.array(ra, dtype=np.int32), nocc, nrad, nang)


def config4_molecule(seed):
    """One molecule of the batched workload: 4-20 atoms (C/N/O/H, 40-80%
    heavy), bonds 1.8-2.9 bohr, 30-300 Cartesian s/p/d primitives with
    exponents log-uniform in [0.05, 500] (shells added at random until the
    seeded target count is reached; every atom keeps one s shell), an HF-like
    DM2 from N/2 seeded occupied orbitals (N the electron count of the
    neutral molecule, rounded down to even), and the seeded autogrid orders nrad in
    {50, 75, 100, 150} and nang in {110, 194, 302, 590}.  Returns
    (System, nrad, nang)."""
    rng, elements, coords, alpha, tmn, cen, ra, nocc, nrad, nang = _config4_layout(seed)
    C = orbitals(rng, overlap_matrix(alpha, tmn, cen), nocc)
    return System(elements, coords, alpha, tmn, cen, ra, C), nrad, nang


def config4_sizes(seed):
    """(atoms, primitives, electrons, unique records, nrad, nang) of a batch
    molecule, without its orbitals."""
    _, elements, _, alpha, _, _, _, nocc, nrad, nang = _config4_layout(seed)
    nprim = len(alpha)
    npair = nprim * (nprim + 1) // 2
    return len(elements), nprim, 2 * nocc, npair * (npair + 1) // 2, nrad, nang
