"""Python readers for the INCA input formats, used by the tests to feed the
oracle and the C-ABI with the same primitives the Fortran host reads.

Test infrastructure: these mirror the reference readers
(src/mod_read_files.f90:432-726 for .fchk/.bas, src/mod_intracule.f90:164-262
for .dm2p) closely enough for the fixtures; the product's readers are the
Fortran ones under inca_parallel_amd/fortran.
"""
from __future__ import annotations

import re

import numpy as np

# Cartesian component order per shell type (src/mod_read_files.f90:556-647)
_SHELL_TMN = {
    0: [(0, 0, 0)],
    1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    2: [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
    3: [(3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (2, 0, 1), (0, 2, 1),
        (1, 2, 0), (1, 0, 2), (0, 1, 2), (1, 1, 1)],
}


def _fchk_block(lines, key):
    for n, ln in enumerate(lines):
        if ln.startswith(key):
            m = re.search(r"N=\s*(\d+)", ln)
            cnt = int(m.group(1))
            vals = []
            k = n + 1
            while len(vals) < cnt:
                vals += lines[k].split()
                k += 1
            return vals
    raise KeyError(key)


def _fchk_scalar(lines, key):
    for ln in lines:
        if ln.startswith(key):
            return ln.split()[-1]
    raise KeyError(key)


def read_fchk(path):
    lines = open(path).read().splitlines()
    natoms = int(_fchk_scalar(lines, "Number of atoms"))
    na = int(_fchk_scalar(lines, "Number of alpha electrons"))
    nb = int(_fchk_scalar(lines, "Number of beta electrons"))
    an = np.array([int(v) for v in _fchk_block(lines, "Atomic numbers")])
    cart = np.array([float(v) for v in
                     _fchk_block(lines, "Current cartesian coordinates")]
                    ).reshape(natoms, 3)
    st = [int(v) for v in _fchk_block(lines, "Shell types")]
    pps = [int(v) for v in _fchk_block(lines, "Number of primitives per shell")]
    s2a = [int(v) for v in _fchk_block(lines, "Shell to atom map")]
    pexp = [float(v) for v in _fchk_block(lines, "Primitive exponents")]
    alpha, tmn, ra = [], [], []
    l = 0
    for t, n, a in zip(st, pps, s2a):
        if t == -1:
            for comp in [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]:
                for k in range(n):
                    alpha.append(pexp[l + k]); tmn.append(comp); ra.append(a)
            l += n
            continue
        comps = _SHELL_TMN[abs(t)]
        for k in range(n):
            for comp in comps:
                alpha.append(pexp[l]); tmn.append(comp); ra.append(a)
            l += 1
    return dict(natoms=natoms, nelec=na + nb, an=an, cartes=cart,
                alpha=np.array(alpha), tmn=np.array(tmn, dtype=np.int32),
                ra=np.array(ra))


def read_bas(path, mol):
    """Override centres, exponents and t,m,n (src/mod_read_files.f90:695-726)."""
    lines = open(path).read().splitlines()
    start = next(i for i, ln in enumerate(lines) if "#" in ln) + 1
    nprim = len(mol["alpha"])
    x, y, z, alpha, tmn = [], [], [], [], []
    for ln in lines[start:start + nprim]:
        f = ln.split()
        x.append(float(f[1])); y.append(float(f[2])); z.append(float(f[3]))
        alpha.append(float(f[4]))
        tmn.append((int(f[5]), int(f[6]), int(f[7])))
    out = dict(mol)
    out.update(x=np.array(x), y=np.array(y), z=np.array(z),
               alpha=np.array(alpha), tmn=np.array(tmn, dtype=np.int32))
    return out


DM2P_DTYPE = np.dtype([("i", "<i4"), ("j", "<i4"), ("k", "<i4"),
                       ("l", "<i4"), ("d", "<f8")])


def read_dm2p(path):
    """Records (ijkl[n,4] int32, d[n]) with the reference's format detection."""
    raw = np.fromfile(path, dtype=np.uint8)
    head = raw[:8].view("<i4") if raw.size >= 8 else np.zeros(2, np.int32)
    if head[0] == 24 and head[1] == 0:
        raise ValueError("8-byte record markers are unsupported")
    if head[0] == 0:
        raise ValueError("zero-marker format is unsupported")
    if head[0] == 24:
        n = raw.size // 32
        rec = raw[:n * 32].reshape(n, 32)
        good = (rec[:, :4].copy().view("<i4")[:, 0] == 24) & \
               (rec[:, 28:].copy().view("<i4")[:, 0] == 24)
        bad = np.nonzero(~good)[0]
        if bad.size:
            n = int(bad[0])
        body = np.ascontiguousarray(rec[:n, 4:28]).view(DM2P_DTYPE).reshape(-1)
    else:
        n = raw.size // 24
        body = raw[:n * 24].view(DM2P_DTYPE)
    ijkl = np.stack([body["i"], body["j"], body["k"], body["l"]], axis=1)
    return np.ascontiguousarray(ijkl, dtype=np.int32), np.array(body["d"])
