"""Pins of the CPU FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a wrong sign or index, a transposed operand) fails it:
  * the reference's own CH4 goldens (tests/run_tests.py:19-28, ch4_ref.rad,
    ch4_ref.cube) through the whole path;
  * Gauss-Hermite / Gauss-Legendre tables against numpy;
  * unit self-overlap of normalised primitives (normprim);
  * per-quartet f_q(u): closed form for s quartets and the sum rule
    int f_q = S_ik S_jl for s/p/d quartets;
  * TFVC partition of unity and inversion symmetry;
  * hand-computed automatic centres.
"""
import numpy as np
import pytest
from numpy.polynomial.hermite import hermgauss

import ch4_golden as GD
from inca_parallel_amd import synth
from oracle import oracle as O

THRESH = 1e-12


def test_ch4_integral_zero_two_points(ch4):
    cent, Ps, sf = O.centercalc(ch4["an"], ch4["cartes"], True, False)
    assert len(Ps) == 1 and Ps[0] == 15.0 and abs(sf[0] - 0.70 * 1.8897) < 1e-15
    g = O.gridpoints(cent, [50], [110], sf, Ps)
    assert (g["maxgrid"], g["rgrid"], g["rrgrid"]) == (5500, 3050, 3050)
    iv, cnt = O.intracule(ch4, ch4["ijkl"], ch4["d"], g["r"], THRESH)
    assert cnt[0] == 142884
    tot = np.sum(g["w"] * iv)
    assert abs(tot - GD.INTEGRAL) / GD.INTEGRAL < 1e-12
    iv0, _ = O.intracule(ch4, ch4["ijkl"], ch4["d"], np.zeros((1, 3)), THRESH)
    assert abs(iv0[0] / 2 - GD.ZERO_HALF) / GD.ZERO_HALF < 1e-12
    iv2, _ = O.intracule(ch4, ch4["ijkl"], ch4["d"], np.array([[0, 0, 0], [1.0, 0, 0]]), THRESH)
    assert np.allclose(iv2, GD.TWO_POINTS, rtol=1e-12, atol=0)


def test_ch4_radial_scan_and_cube(ch4):
    gp = O.gridpoints2([[0.0, 5.0]], [0.05], [110])
    iv, _ = O.intracule(ch4, ch4["ijkl"], ch4["d"], gp["r"], THRESH)
    rint, s = [], 0
    for n in gp["smn"]:
        rint.append(np.sum(gp["w_ang"][s:s + n] * iv[s:s + n]))
        s += n
    ref = GD.rad_table()
    assert np.abs(gp["radi"] - ref[:, 0]).max() < 1e-12
    assert (np.abs(np.array(rint) - ref[:, 1]) / np.abs(ref[:, 1])).max() < 1e-8
    r = O.gridpoints3([0, 0, 0], [0.1, 0.1, 0.1], [20, 20, 20])
    ivc, _ = O.intracule(ch4, ch4["ijkl"], ch4["d"], r, THRESH)
    vr = GD.cube_values()
    assert (np.abs(ivc - vr) / np.abs(vr)).max() < 1e-12


@pytest.mark.parametrize("L", range(0, 22))
def test_gauherm_against_numpy(L):
    rh = np.zeros(11)
    wh = np.zeros(11)
    n = O.lib().orc_gauherm(L, rh, wh)
    assert n == L // 2 + 1
    x, w = hermgauss(n)
    assert np.abs(np.sort(rh[:n]) - np.sort(x)).max() < 1e-14
    assert np.abs(wh[:n][np.argsort(rh[:n])] - w[np.argsort(x)]).max() < 1e-13 * w.max()


def test_gauleg_against_numpy():
    for n in (1, 2, 7, 50, 150):
        t, w = O.gauleg(n)
        t2, w2 = np.polynomial.legendre.leggauss(n)
        assert np.abs(t - t2).max() < 1e-14 and np.abs(w - w2).max() < 1e-13


def test_normprim_unit_self_overlap():
    rng = np.random.default_rng(0)
    for _ in range(20):
        a = float(np.exp(rng.uniform(np.log(0.05), np.log(500))))
        t = tuple(int(v) for v in rng.integers(0, 3, size=3))
        N = O.normprim([a], [t])[0]
        # int x^2t exp(-2 a x^2) dx = (2t-1)!! / (4a)^t sqrt(pi/(2a))
        s = 1.0
        for ti in t:
            s *= synth.dfact(2 * ti - 1) / (4 * a) ** ti * np.sqrt(np.pi / (2 * a))
        assert abs(N * N * s - 1.0) < 1e-13


def _one_quartet(prims, q, r):
    """Ivec of a single record (i,j,k,l,D=1) without normalisation."""
    ijkl = np.array([q], dtype=np.int32)
    iv, cnt = O.intracule(prims, ijkl, np.array([1.0]), r, 0.0, normalize=False)
    return iv, cnt


def _prims(cen, alpha, tmn):
    cen = np.asarray(cen, float)
    return dict(alpha=np.asarray(alpha, float), tmn=np.asarray(tmn, np.int32),
                x=cen[:, 0].copy(), y=cen[:, 1].copy(), z=cen[:, 2].copy())


def test_s_quartet_closed_form():
    rng = np.random.default_rng(1)
    cen = rng.uniform(-1.5, 1.5, size=(4, 3))
    al = np.exp(rng.uniform(np.log(0.1), np.log(20), size=4))
    pr = _prims(cen, al, [(0, 0, 0)] * 4)
    r = rng.uniform(-2, 2, size=(50, 3))
    iv, _ = _one_quartet(pr, (1, 2, 3, 4), r)
    ai, aj, ak, al_ = al
    aik, ajl = ai + ak, aj + al_
    eik, ejl = ai * ak / aik, aj * al_ / ajl
    Pik = (ai * cen[0] + ak * cen[2]) / aik
    Pjl = (aj * cen[1] + al_ * cen[3]) / ajl
    zeta = aik + ajl
    e = aik * ajl / zeta
    R2ik = np.sum((cen[0] - cen[2]) ** 2)
    R2jl = np.sum((cen[1] - cen[3]) ** 2)
    ref = (np.pi / zeta) ** 1.5 * np.exp(-eik * R2ik - ejl * R2jl) * \
        np.exp(-e * np.sum((r + Pik - Pjl) ** 2, axis=1))
    assert np.abs(iv - ref).max() < 1e-13 * np.abs(ref).max()


def _overlap_1d(a, A, ta, b, B, tb):
    p = a + b
    P = (a * A + b * B) / p
    u, w = hermgauss(10)
    x = P + u / np.sqrt(p)
    return np.exp(-a * b / p * (A - B) ** 2) / np.sqrt(p) * np.sum(w * (x - A) ** ta * (x - B) ** tb)


@pytest.mark.parametrize("seed", [2, 3, 4, 5, 6])
def test_quartet_sum_rule(seed):
    """int f_q(u) d3u = S_ik S_jl for s/p/d quartets (unnormalised
    primitives), integrated with a tensor Gauss-Hermite rule that is exact for
    this integrand.  The residual is the conditioning of the Vandermonde
    solve inside polycoef; the bar is relative to the L1 mass of f_q because
    signed p/d overlaps can cancel."""
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-1.0, 1.0, size=(4, 3))
    al = np.exp(rng.uniform(np.log(0.3), np.log(4), size=4))
    tmn = [tuple(int(v) for v in rng.integers(0, 2, size=3)) for _ in range(4)]
    pr = _prims(cen, al, tmn)
    # 3D tensor Gauss-Hermite grid centred on the quartet centre
    ai, aj, ak, al_ = al
    Pik = (ai * cen[0] + ak * cen[2]) / (ai + ak)
    Pjl = (aj * cen[1] + al_ * cen[3]) / (aj + al_)
    e = (ai + ak) * (aj + al_) / (ai + ak + aj + al_)
    u, w = hermgauss(24)
    d = -(Pik - Pjl)
    X, Y, Z = np.meshgrid(u, u, u, indexing="ij")
    W = (w[:, None, None] * w[None, :, None] * w[None, None, :]).reshape(-1)
    sc = 1.0 / np.sqrt(e)
    r = np.stack([X.reshape(-1) * sc + d[0], Y.reshape(-1) * sc + d[1],
                  Z.reshape(-1) * sc + d[2]], axis=1)
    iv, cnt = _one_quartet(pr, (1, 2, 3, 4), r)
    assert cnt[3] == 0
    # int f d3u = sum_w f(u) exp(+|y|^2) / e^{3/2}
    y2 = np.sum(((r - d) * np.sqrt(e)) ** 2, axis=1)
    integral = np.sum(W * iv * np.exp(y2)) * sc ** 3
    mass = np.sum(W * np.abs(iv) * np.exp(y2)) * sc ** 3  # L1 scale of f_q
    Sik = np.prod([_overlap_1d(ai, cen[0][x], tmn[0][x], ak, cen[2][x], tmn[2][x]) for x in range(3)])
    Sjl = np.prod([_overlap_1d(aj, cen[1][x], tmn[1][x], al_, cen[3][x], tmn[3][x]) for x in range(3)])
    assert abs(integral - Sik * Sjl) < 1e-11 * mass


def _exact_fq(cen, al, tmn, r):
    """f_q(u) from the exact axis polynomials: p(u) = sum_n w_n prod over the
    four primitives of (t_n/sqrt(zeta) + (alpha' -+ 1/2) u + P - R)^t, expanded
    with numpy polynomial arithmetic (no Vandermonde solve)."""
    P = np.polynomial.polynomial
    ai, aj, ak, al_ = al
    aik, ajl = ai + ak, aj + al_
    zeta = aik + ajl
    Pik = (ai * cen[0] + ak * cen[2]) / aik
    Pjl = (aj * cen[1] + al_ * cen[3]) / ajl
    Pq = (aik * Pik + ajl * Pjl) / zeta
    alf = 0.5 * (aik - ajl) / zeta
    e = aik * ajl / zeta
    A = zeta ** -1.5 * np.exp(-ai * ak / aik * np.sum((cen[0] - cen[2]) ** 2)
                              - aj * al_ / ajl * np.sum((cen[1] - cen[3]) ** 2))
    f = A * np.exp(-e * np.sum((r + Pik - Pjl) ** 2, axis=1))
    for ax in range(3):
        L = sum(t[ax] for t in tmn)
        tn, wn = hermgauss(L // 2 + 1)
        poly = np.zeros(1)
        for t_, w_ in zip(tn, wn):
            term = np.array([w_])
            for a, sgn in ((0, -1), (1, 1), (2, -1), (3, 1)):
                lin = np.array([t_ / np.sqrt(zeta) + Pq[ax] - cen[a][ax], alf + 0.5 * sgn])
                term = P.polymul(term, P.polypow(lin, tmn[a][ax]))
            poly = P.polyadd(poly, term)
        f = f * P.polyval(r[:, ax], poly)
    return f


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_quartet_against_exact_polynomials(seed):
    """s/p/d quartets pointwise against the exact expansion.  The reference's
    Vandermonde fit loses up to ~1e-7 of max|f| for Lx ~ 8, so the bar is
    1e-6 of max|f|; a wrong sign, index or operand gives O(1) errors."""
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-1.0, 1.0, size=(4, 3))
    al = np.exp(rng.uniform(np.log(0.3), np.log(4), size=4))
    tmn = [tuple(int(v) for v in rng.integers(0, 3, size=3)) for _ in range(4)]
    pr = _prims(cen, al, tmn)
    r = rng.uniform(-2.5, 2.5, size=(200, 3))
    iv, cnt = _one_quartet(pr, (1, 2, 3, 4), r)
    assert cnt[3] == 0
    ref = _exact_fq(cen, al, tmn, r)
    assert np.abs(iv - ref).max() < 1e-6 * np.abs(ref).max()


def test_tfvc_partition_of_unity_and_inversion():
    rng = np.random.default_rng(5)
    cent = np.array([[0, 0, 0], [1.5, 0.3, 0.0], [-1.5, -0.3, 0.0], [0.2, 2.0, 0.7],
                     [-0.2, -2.0, -0.7]])
    Ps = np.array([15.0, 18.0, 18.0, 24.0, 24.0])
    pts = rng.uniform(-4, 4, size=(400, 3))
    tot = np.zeros(len(pts))
    for c in range(len(Ps)):
        w = O.tfvc(pts, np.full(len(pts), c, np.int32), cent, Ps)
        tot += w
        # inversion: the weight of centre c at p equals that of -c at -p
        mirror = {0: 0, 1: 2, 2: 1, 3: 4, 4: 3}[c]
        wm = O.tfvc(-pts, np.full(len(pts), mirror, np.int32), cent, Ps)
        assert np.abs(w - wm).max() < 1e-12
    assert np.abs(tot - 1.0).max() < 1e-12


def test_multicentre_grid_integrates_gaussians():
    """Multicentre grid with TFVC integrates normalised Gaussians placed
    anywhere to quadrature accuracy."""
    cent = np.array([[0, 0, 0], [2.0, 0.5, 0.0], [-2.0, -0.5, 0.0]])
    Ps = np.array([15.0, 18.0, 18.0])
    sf = np.array([1.3, 1.2, 1.2])
    g = O.gridpoints(cent, [80] * 3, [302] * 3, sf, Ps, nosym=True)
    for c0, a in (((0.3, -0.2, 0.1), 1.0), ((2.1, 0.4, -0.2), 2.0), ((-1.0, 1.0, 0.5), 0.7)):
        f = (a / np.pi) ** 1.5 * np.exp(-a * np.sum((g["r"] - np.array(c0)) ** 2, axis=1))
        # radial weights carry 2 pi (not 4 pi): the grid integrates to half
        assert abs(2 * np.sum(g["w"] * f) - 1.0) < 1e-6


def test_centercalc_two_heavy_atoms():
    an = np.array([6, 8, 1])
    cart = np.array([[0.0, 0.0, 0.0], [1.2, 0.4, 2.0], [-1.0, 0.0, 0.3]])
    cent, Ps, sf = O.centercalc(an, cart, True, False)
    rC, rO = 0.70 * 1.8897, 0.60 * 1.8897
    assert len(Ps) == 3
    assert Ps[0] == 15.0 + 28.0
    assert abs(sf[0] - (15 * rC + 28 * rO) / 43) < 1e-14
    assert np.allclose(cent[1], [-1.2, -0.4, -2.0]) and np.allclose(cent[2], [1.2, 0.4, 2.0])
    assert Ps[1] == Ps[2] == 24.0 and abs(sf[1] - (rC + rO) / 2) < 1e-14
    # z-snap below 0.5 bohr
    cart2 = cart.copy()
    cart2[1, 2] = 0.3
    cent2, _, _ = O.centercalc(an, cart2, True, False)
    assert cent2[1][2] == 0.0 and cent2[2][2] == 0.0


def test_smearing_conserves_the_integral():
    rng = np.random.default_rng(7)
    pts = rng.uniform(-3, 3, size=(2000, 3))
    w = rng.uniform(0.1, 1.0, size=2000)
    dens = rng.uniform(0, 1, size=2000)
    u, P, Ia = O.smearing(pts, w, dens, u_max=12.0, n_bins=1000, sigma=0.15)
    bw = u[1] - u[0]
    assert abs(np.sum(P) * bw - np.sum(w * dens)) < 2e-3 * np.sum(w * dens)


def _smear_closed_form(u, pts, w, dens, sigma):
    """P(u_j) of radial_smearing (mod_intracule.f90:576-649) written from its
    definition: every point p spreads w I onto the bins as a normal density
    of width sigma centred at |p|, over the bins from 4 sigma above |p| down
    to 4 sigma plus one bin below it (the window's integer truncation), and
    nowhere else."""
    bw = u[1] - u[0]
    P = np.zeros_like(u)
    for p, wi, di in zip(np.atleast_2d(pts), np.atleast_1d(w), np.atleast_1d(dens)):
        u0 = float(np.linalg.norm(p))
        off = u - u0
        inside = (off <= 4 * sigma) & (off > -4 * sigma - bw)
        g = np.exp(-off ** 2 / (2 * sigma ** 2)) / (sigma * np.sqrt(2 * np.pi))
        P += np.where(inside, wi * di * g, 0.0)
    return P


@pytest.mark.parametrize("u0", [2.3456, 0.1, 9.87, 4.0])
def test_smearing_single_point_closed_form(u0):
    """One point of weight w and density I at |p| = u0: P is w I N(u; u0,
    0.15) on its window, zero elsewhere; I_avg = P / (4 pi u^2), 0 at u = 0;
    the axis is u_j = j u_max / (n_bins - 1)."""
    d = np.array([0.3, -0.48, 0.824])
    p = d / np.linalg.norm(d) * u0
    u, P, Ia = O.smearing(p[None, :], [0.7], [1.9], u_max=10.0, n_bins=1000, sigma=0.15)
    assert np.array_equal(u, np.arange(1000) * (10.0 / 999))
    ref = _smear_closed_form(u, p[None, :], [0.7], [1.9], 0.15)
    assert np.count_nonzero(P) == np.count_nonzero(ref) > 50
    assert np.allclose(P, ref, rtol=1e-13, atol=0.0)
    Iref = np.where(u > 1e-6, ref / (4 * np.pi * np.where(u > 0, u, 1.0) ** 2), 0.0)
    assert Ia[0] == 0.0
    assert np.allclose(Ia, Iref, rtol=1e-13, atol=0.0)
    # the peak sits on the bin nearest u0, at w I / (sigma sqrt(2 pi)) there
    j = int(np.argmax(P))
    assert abs(u[j] - u0) <= 0.5 * (u[1] - u[0]) + 1e-12


def test_smearing_is_linear_in_the_points():
    rng = np.random.default_rng(3)
    pts = rng.uniform(-4, 4, size=(40, 3))
    w = rng.uniform(0.1, 1.0, size=40)
    dens = rng.uniform(-1, 1, size=40)
    u, P, _ = O.smearing(pts, w, dens, u_max=9.0, n_bins=700, sigma=0.15)
    ref = _smear_closed_form(u, pts, w, dens, 0.15)
    assert np.allclose(P, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


# ---- one-electron functions (mod_functions.f90) -----------------------------

def _onee_system():
    from inca_parallel_amd import synth
    rng = np.random.default_rng(41)
    elements = ["O", "H", "H"]
    coords = np.array([[0.0, 0.0, 0.0], [1.43, 1.1, 0.0], [-1.43, 1.1, 0.3]])
    alpha, tmn, cen, ra = synth.make_basis(rng, elements, coords, {0: 2, 1: 1, 2: 1}, {0: 2},
                                           0.2, 3.0)
    C = synth.orbitals(rng, synth.overlap_matrix(alpha, tmn, cen), 3)
    s = synth.System(elements, coords, alpha, tmn, cen, ra, C)
    norm = np.array([synth.normprim(a, t) for a, t in zip(alpha, tmn)])
    T = (s.C * norm[:, None]).T  # MO coefficients of the unnormalised primitives
    return s, T


def test_onee_primitive_laplacian_by_finite_differences():
    """sd (the Laplacian of a primitive, mod_functions.f90:113-152) against a
    fourth-order central difference of Prim, for s, p and d primitives: the
    Laplacian of one primitive p is Lapl with T = e_p and w = 1/2 divided by
    Prim_p."""
    s, _ = _onee_system()
    p = s.prims()
    rng = np.random.default_rng(2)
    pts = rng.uniform(-1.5, 1.5, size=(6, 3))
    h = 1e-3
    for q in range(s.nprim):
        T = np.zeros((1, s.nprim))
        T[0, q] = 1.0
        v = O.onee(1, q + 1, p, pts)
        lap_v = O.onee(6, 0, p, pts, T, [0.5])
        fd = np.zeros(len(pts))
        for ax in range(3):
            e = np.zeros(3)
            e[ax] = h
            f = [O.onee(1, q + 1, p, pts + k * e) for k in (-2, -1, 1, 2)]
            fd += (-f[0] + 16 * f[1] - 30 * v + 16 * f[2] - f[3]) / (12 * h * h)
        ok = np.abs(v) > 1e-3 * np.abs(v).max()
        assert np.allclose(lap_v[ok] / v[ok], fd[ok], rtol=1e-5, atol=1e-6 * np.abs(fd).max())


def test_onee_orbitals_and_density_integrate_to_their_counts():
    """On a fine cube the S-orthonormal orbitals (MoOr) integrate to 1 in
    square, and the closed-shell density (Density, w = 2) to N = 6; the
    primitive (Prim) of an s function to (pi/alpha)^(3/2)."""
    s, T = _onee_system()
    p = s.prims()
    h = 0.12
    ax = np.arange(-7.0, 7.0 + h / 2, h)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    dv = h ** 3
    rho = O.onee(5, 0, p, pts, T, [2.0] * 3)
    assert abs(rho.sum() * dv - 6.0) < 1e-5
    for i in range(3):
        phi = O.onee(3, i + 1, p, pts, T)
        assert abs((phi ** 2).sum() * dv - 1.0) < 1e-5
    q = int(np.nonzero(np.asarray(p["tmn"]).sum(1) == 0)[0][0])
    g = O.onee(1, q + 1, p, pts)
    assert abs(g.sum() * dv - (np.pi / p["alpha"][q]) ** 1.5) < 1e-6 * (np.pi / p["alpha"][q]) ** 1.5
    # the reference's Laplacian integrates to -4 sum_i int |grad phi_i|^2;
    # check it pointwise against 2 w phi lap(phi) by central differences
    lap = O.onee(6, 0, p, pts[::997], T, [2.0] * 3)
    hd = 1e-3
    ref = np.zeros(len(lap))
    for i in range(3):
        phi0 = O.onee(3, i + 1, p, pts[::997], T)
        l2 = -6 * phi0
        for a in range(3):
            e = np.zeros(3)
            e[a] = hd
            l2 = l2 + O.onee(3, i + 1, p, pts[::997] + e, T) + O.onee(3, i + 1, p, pts[::997] - e, T)
        ref += 4.0 * phi0 * l2 / hd ** 2
    assert np.allclose(lap, ref, rtol=1e-4, atol=1e-5 * np.abs(ref).max())
