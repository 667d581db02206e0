"""Pins of the CPU oracle (oracle/oracle.c) and of the integral engine it is fed by.

Golden values are the reference's own test expectations (water.xyz, STO-3G, TEST_TOLERANCE = 1e-8):
test/hartree_fock.zig:24,56; test/moller_plesset.zig:27,62,96,131,165,200;
test/configuration_interaction.zig:27-31,66-70,105-109.  Everything else pins the oracle through
mathematics the reference's algorithm must satisfy (closed forms, invariants, identities).
"""
import numpy as np
import pytest

from conftest import system_ints
from oracle import oracle as O
from zinq_amd.system import MolecularSystem, spin_block_1e, spin_block_2e

TOL = 1e-8  # TEST_TOLERANCE of the reference's tests


@pytest.fixture(scope="module")
def rhf():
    s, H, S, g, VN = system_ints("sto3g")
    r = O.scf(H, S, g, VN, s.nel // 2)
    return s, H, S, g, VN, r


@pytest.fixture(scope="module")
def ghf():
    s, H, S, g, VN = system_ints("sto3g")
    Hs, Ss, gs = spin_block_1e(H), spin_block_1e(S), spin_block_2e(g)
    return s, Hs, Ss, gs, VN, O.scf(Hs, Ss, gs, VN, s.nel, generalized=True)


def test_rhf_energy_matches_reference(rhf):
    assert rhf[-1]["E"] == pytest.approx(-74.9659012172971900, abs=TOL)  # test/hartree_fock.zig:24


def test_ghf_energy_matches_reference(ghf):
    assert ghf[-1]["E"] == pytest.approx(-74.9659012172972300, abs=TOL)  # test/hartree_fock.zig:56


@pytest.mark.parametrize("order,ref", [(2, -75.0048550569061500), (3, -75.0154555119455900), (4, -75.0187039890576300)])
def test_rmp_energies_match_reference(rhf, order, ref):
    s, H, S, g, VN, r = rhf
    e = O.mp(order, g, r["C"], r["e"], s.nel // 2)
    assert r["E"] + e[2:].sum() == pytest.approx(ref, abs=TOL)  # test/moller_plesset.zig:27,96,165


@pytest.mark.parametrize("order,ref", [(2, -75.0048550569062000), (3, -75.0154555119455300), (4, -75.0187039890575700)])
def test_gmp_energies_match_reference(ghf, order, ref):
    s, H, S, g, VN, r = ghf
    e = O.mp(order, g, r["C"], r["e"], s.nel, generalized=True)
    assert r["E"] + e[2:].sum() == pytest.approx(ref, abs=TOL)  # test/moller_plesset.zig:62,131,200


def test_rcis_energies_match_reference(rhf):
    s, H, S, g, VN, r = rhf
    e = O.ci(g, H, r["C"], s.nel // 2, [1], VN)
    ref = [-74.9659012172972900, -74.5822026772564000, -74.5822026772563200, -74.5822026772562900, -74.5070515329961900]
    np.testing.assert_allclose(e, ref, atol=TOL)  # test/configuration_interaction.zig:27-31


def test_gcis_energies_match_reference(ghf):
    s, H, S, g, VN, r = ghf
    e = O.ci(g, H, r["C"], s.nel, [1], VN, generalized=True)
    ref = [-74.9659012172973100, -74.5822026772563600, -74.5822026772563500, -74.5822026772562800, -74.5070515329962000]
    np.testing.assert_allclose(e, ref, atol=TOL)  # test/configuration_interaction.zig:66-70


def test_rcisd_energies_match_reference():
    s, H, S, g, VN = system_ints("sto3g")
    r = O.scf(H, S, g, VN, s.nel // 2, diis=None)  # .diis = null in the reference test
    e = O.ci(g, H, r["C"], s.nel // 2, [1, 2], VN)
    ref = [-75.0195290856878400, -74.6224197038007400, -74.6224197038007200, -74.6224197038007200, -74.5642383868465500]
    np.testing.assert_allclose(e, ref, atol=TOL)  # test/configuration_interaction.zig:105-109


# ---------------------------------------------------------------- mathematics the oracle must satisfy

def test_closed_shell_mp2_equals_spin_orbital_rspt(rhf):
    """The determinant-space order-2 term equals the closed-shell sum over (ia|jb): E2 = sum (ia|jb)[2(ia|jb) - (ib|ja)] / (e_i + e_j - e_a - e_b)."""
    s, H, S, g, VN, r = rhf
    o = s.nel // 2
    e2_det = O.mp(2, g, r["C"], r["e"], o)[2]
    e2_cs = O.mp2_closed(O.ao2mo_oovv(g, r["C"], o), r["e"])
    assert e2_cs == pytest.approx(e2_det, abs=1e-13)


def test_closed_shell_mp2_equals_spin_orbital_rspt_ccpvdz():
    s, H, S, g, VN = system_ints("ccpvdz")
    o = s.nel // 2
    r = O.scf(H, S, g, VN, o, threshold=1e-10)
    e2_det = O.mp(2, g, r["C"], r["e"], o)[2]
    e2_cs = O.mp2_closed(O.ao2mo_oovv(g, r["C"], o), r["e"])
    assert e2_cs == pytest.approx(e2_det, abs=1e-12)


def test_fock_is_symmetric_linear_and_coulomb_positive(rhf):
    s, H, S, g, VN, r = rhf
    rng = np.random.default_rng(1)
    A, B = rng.standard_normal((2, s.nbf, s.nbf))
    P1, P2 = A + A.T, B + B.T
    Z = np.zeros_like(H)
    G = lambda P, x=False: O.get_fock(g, Z, P, x)
    np.testing.assert_allclose(G(P1 + P2), G(P1) + G(P2), atol=1e-12)
    F = O.get_fock(g, H, P1)
    assert np.array_equal(F, F.T)
    # Coulomb part alone (J = G_RHF + K/2 with K = G_GHF(x=1) - J ... use x=0 via difference)
    J = 2 * G(P1) - G(P1, True)  # G_RHF = J - K/2, G_GHF = J - K  =>  J = 2 G_RHF - G_GHF
    assert np.sum(P1 * J) >= 0.0  # (ij|kl) is a positive semidefinite kernel


def test_transform_identity_orthogonal_and_index_order(rhf):
    s, H, S, g, VN, r = rhf
    n = s.nbf
    np.testing.assert_array_equal(O.ao2mo_pppp(g, np.eye(n)), g)
    Q, _ = np.linalg.qr(np.random.default_rng(2).standard_normal((n, n)))
    gq = O.ao2mo_pppp(g, Q)
    assert np.linalg.norm(gq) == pytest.approx(np.linalg.norm(g), rel=1e-13)
    # canonical RHF identity ties the transform to the SCF: eps_p = h_pp + sum_i [2 (pp|ii) - (pi|ip)]
    C, e, o = r["C"], r["e"], s.nel // 2
    mo = O.ao2mo_pppp(g, C)  # <pq|rs>
    h = O.ao2mo_pp(H, C)
    eps = [h[p, p] + sum(2 * mo[p, i, p, i] - mo[p, i, i, p] for i in range(o)) for p in range(n)]
    np.testing.assert_allclose(eps, e, atol=1e-7)  # SCF converged to 1e-8
    # ao2mo_oovv is the occupied/virtual block of ao2mo_pppp
    np.testing.assert_allclose(O.ao2mo_oovv(g, C, o), mo[:o, :o, o:, o:], atol=1e-13)


def test_geigh_solves_generalized_problem(rhf):
    s, H, S, g, VN, r = rhf
    e, C = O.geigh(H, S)
    np.testing.assert_allclose(H @ C, S @ C * e, atol=1e-12)
    np.testing.assert_allclose(C.T @ S @ C, np.eye(s.nbf), atol=1e-12)
    assert np.all(np.diff(e) >= 0)


def test_diis_reset_on_singular_history(rhf):
    """Two identical history entries make the DIIS matrix singular: the LU must report failure."""
    s, H, S, g, VN, r = rhf
    F = O.get_fock(g, H, r["P"])
    err = O.get_error(F, r["P"], S) + 1e-3
    _, ok = O.diis([F, F], [err, err])
    assert not ok


def test_scf_reports_non_convergence(rhf):
    s, H, S, g, VN, r = rhf
    with pytest.raises(O.ScfDidNotConverge):
        O.scf(H, S, g, VN, s.nel // 2, iterations=1)


# ---------------------------------------------------------------- integral engine (libint stand-in)

def test_h2_sto3g_matches_szabo_ostlund():
    """H2, R = 1.4 bohr, STO-3G (zeta = 1.24): Szabo & Ostlund, Modern Quantum Chemistry, sec. 3.5."""
    s = MolecularSystem.from_atoms([1, 1], [0, 0, 0, 0, 0, 1.4], "builtin:sto-3g")
    S, T, V, g = s.overlap(), s.kinetic(), s.nuclear(), s.coulomb()
    assert S[0, 1] == pytest.approx(0.6593, abs=1e-4)
    assert T[0, 0] == pytest.approx(0.7600, abs=1e-4)
    assert T[0, 1] == pytest.approx(0.2365, abs=1e-4)
    assert (T + V)[0, 0] == pytest.approx(-1.1204, abs=1e-4)
    assert (T + V)[0, 1] == pytest.approx(-0.9584, abs=1e-4)
    assert g[0, 0, 0, 0] == pytest.approx(0.7746, abs=1e-4)   # (11|11)
    assert g[0, 1, 0, 1] == pytest.approx(0.5697, abs=1e-4)   # (11|22) = <12|12>
    assert g[1, 0, 0, 0] == pytest.approx(0.4441, abs=1e-4)   # (21|11) = <20|10>... = g[1,0,0,0]
    assert g[1, 1, 0, 0] == pytest.approx(0.2970, abs=1e-4)   # (21|21) = <22|11>
    r = O.scf(T + V, S, g, s.nrep(), 1)
    assert r["E"] == pytest.approx(-1.1167, abs=1e-4)


def test_single_primitive_ssss_closed_form(tmp_path):
    """(aa|bb) of normalised s Gaussians: (pi/2 ... ) closed form with F0(T) = sqrt(pi/T) erf(sqrt T) / 2."""
    from math import erf, pi, sqrt
    alpha, R = 0.8, 1.7
    (tmp_path / "b.g94").write_text("H     0\nS    1   1.00\n      %.10f       1.0000000\n****\n" % alpha)
    s = MolecularSystem.from_atoms([1, 1], [0, 0, 0, 0, 0, R], str(tmp_path / "b.g94"))
    g = s.coulomb()
    # normalised s: N = (2a/pi)^(3/4); (AA|BB): p = q = 2a, P = A, Q = B, T = p q/(p+q) R^2 = a R^2
    N4 = (2 * alpha / pi) ** 3
    T = alpha * R * R
    F0 = 0.5 * sqrt(pi / T) * erf(sqrt(T))
    ref = N4 * 2 * pi ** 2.5 / ((2 * alpha) ** 2 * sqrt(4 * alpha)) * F0
    assert g[0, 1, 0, 1] == pytest.approx(ref, rel=1e-13)
    assert g[0, 0, 0, 0] == pytest.approx(N4 * 2 * pi ** 2.5 / ((2 * alpha) ** 2 * sqrt(4 * alpha)), rel=1e-13)


def _ssss(mp, a, b, c, d):
    """(ss|ss) of unnormalised primitive s Gaussians exp(-x|r - X|^2) as an mpmath function of the 12
    centre coordinates (Szabo & Ostlund, Modern Quantum Chemistry, eq. A.41):
    2 pi^(5/2) / (p q sqrt(p + q)) exp(-ab/p |AB|^2 - cd/q |CD|^2) F0(pq/(p+q) |PQ|^2)."""
    p, q = a + b, c + d

    def G(*X):
        A, B, C, D = X[0:3], X[3:6], X[6:9], X[9:12]
        P = [(a * A[i] + b * B[i]) / p for i in range(3)]
        Q = [(c * C[i] + d * D[i]) / q for i in range(3)]
        ab2 = sum((A[i] - B[i]) ** 2 for i in range(3))
        cd2 = sum((C[i] - D[i]) ** 2 for i in range(3))
        T = p * q / (p + q) * sum((P[i] - Q[i]) ** 2 for i in range(3))
        F0 = mp.mpf(1) if T == 0 else mp.sqrt(mp.pi / T) / 2 * mp.erf(mp.sqrt(T))
        return 2 * mp.pi ** 2.5 / (p * q * mp.sqrt(p + q)) * mp.exp(-a * b / p * ab2 - c * d / q * cd2) * F0
    return G


CART_D = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]  # xx yy zz xy xz yz


def _mixed(mp, f, X, orders, h):
    """The mixed derivative prod_i d^{o_i} / dX[i]^{o_i} f (orders = {i: o_i}, o_i in {1, 2}) by the
    product of central-difference stencils (first: (f(+h) - f(-h)) / 2h, second: (f(+h) - 2f + f(-h)) / h^2;
    error O(h^2)); at 80 digits h = 1e-8 leaves round-off near 1e-80 / h^(sum o_i)."""
    import itertools
    taps = {1: ((1, 1), (-1, -1)), 2: ((1, 1), (0, -2), (-1, 1))}  # (offset / h, weight)
    idx = list(orders)
    tot = mp.mpf(0)
    for combo in itertools.product(*(taps[orders[i]] for i in idx)):
        Y = list(X)
        w = 1
        for i, (off, wt) in zip(idx, combo):
            Y[i] += off * h
            w *= wt
        tot += w * f(*Y)
    den = 1
    for i in idx:
        den *= (2 * h) if orders[i] == 1 else h * h
    return tot / den


def _d_basis(tmp_path, shells):
    """A Gaussian94 file with one single-primitive shell per element: shells = [(symbol, L, exponent)]."""
    text = "".join("%s     0\n%s    1   1.00\n      %.12f       1.0000000\n****\n" % (sym, L, e) for sym, L, e in shells)
    (tmp_path / "d.g94").write_text(text)
    return str(tmp_path / "d.g94")


def test_d_shell_integrals_from_derivatives_of_the_ss_closed_form(tmp_path):
    """d-shell ERIs of the integral engine against derivatives of the (ss|ss) closed form, at 30 digits
    (mpmath): a Cartesian d primitive (x_i - A_i)(x_j - A_j) exp(-a|r - A|^2) is
    (d^2/dA_i dA_j + 2a delta_ij) / (4a^2) applied to the s primitive on A.
    1. (d s|s s) at 8 seeded geometries: the engine's five pure d functions are fitted as linear
       combinations of the six Cartesian ones; the fit must be exact (a wrong Hermite recursion is not a
       fixed linear map of the true integrals), each combination traceless (a solid harmonic), the five
       mutually orthogonal with equal norms (tr Q_m Q_n, the pure-d inner product), and d_xy carries the
       analytic normalisation 4a (2a/pi)^(3/4).
    2. (d d|d d), L = 8, against 8th mixed derivatives (the stencils of _mixed): the xy functions on
       four centres, and the x^2 - y^2 functions (Hermite indices up to 8 in x)."""
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    rng = np.random.default_rng(2024)
    ex = dict(He=0.9, H=0.6, Li=0.75, Be=0.45)
    Ns = {k: (2 * v / np.pi) ** 0.75 for k, v in ex.items()}
    # 1. d on He; s on H, Li, Be
    path = _d_basis(tmp_path, [("He", "D", ex["He"]), ("H", "S", ex["H"]), ("Li", "S", ex["Li"]), ("Be", "S", ex["Be"])])
    G = _ssss(mpmath, *(mpmath.mpf(ex[k]) for k in ("He", "H", "Li", "Be")))
    a = ex["He"]
    eng, cart = [], []
    for _ in range(8):
        xyz = rng.uniform(-1.2, 1.2, (4, 3))
        s = MolecularSystem.from_atoms([2, 1, 3, 4], xyz, path)
        assert s.nbf == 8
        g = s.coulomb()
        eng.append([g[m, 6, 5, 7] for m in range(5)])  # (d_m s_H | s_Li s_Be) = <d_m s_Li | s_H s_Be>
        X = [mpmath.mpf(float(v)) for v in xyz.reshape(-1)]
        row = []
        for i, j in CART_D:
            order = [0] * 12
            order[i] += 1
            order[j] += 1
            v = mpmath.diff(G, X, order)
            if i == j:
                v += 2 * a * G(*X)
            row.append(float(v / (4 * a * a)) * Ns["H"] * Ns["Li"] * Ns["Be"])
        cart.append(row)
    eng, cart = np.array(eng), np.array(cart)
    T, *_ = np.linalg.lstsq(cart, eng, rcond=None)  # eng = cart @ T, T: 6 x 5
    resid = np.abs(cart @ T - eng).max()
    print(f"(d s|s s): max |fit residual| {resid:.1e} of max |integral| {np.abs(eng).max():.1e}")
    assert resid < 1e-10 * np.abs(eng).max()
    Q = np.zeros((5, 3, 3))
    for c, (i, j) in enumerate(CART_D):
        for m in range(5):
            if i == j:
                Q[m, i, i] = T[c, m]
            else:
                Q[m, i, j] = Q[m, j, i] = T[c, m] / 2
    norm = np.einsum("mij,nij->mn", Q, Q)
    assert np.abs(np.trace(Q, axis1=1, axis2=2)).max() < 1e-9 * np.sqrt(norm[0, 0])
    np.testing.assert_allclose(norm, norm[0, 0] * np.eye(5), atol=1e-9 * norm[0, 0])
    m_xy = int(np.argmax(np.abs(T[3])))
    assert np.abs(np.delete(T[:, m_xy], 3)).max() < 1e-9 * abs(T[3, m_xy])  # d_xy is the pure xy function
    assert abs(T[3, m_xy]) == pytest.approx(4 * a * (2 * a / np.pi) ** 0.75, rel=1e-10)
    # 2. (d d|d d) of the xy functions: d on He, Li, Be, B
    ex4 = dict(He=0.9, Li=0.7, Be=0.55, B=0.8)
    path4 = _d_basis(tmp_path, [(k, "D", v) for k, v in ex4.items()])
    al = [ex4[k] for k in ("He", "Li", "Be", "B")]
    mpmath.mp.dps = 80
    G4 = _ssss(mpmath, *(mpmath.mpf(v) for v in al))
    mpmath.mp.dps = 30
    Nxy = np.prod([4 * x * (2 * x / np.pi) ** 0.75 for x in al])
    # the x^2 - y^2 function: c (xx - yy), c scaling with the primitive normalisation (alpha / a)^(7/4)
    m_d = [m for m in range(5) if abs(T[0, m]) > 1e-3 and abs(T[0, m] + T[1, m]) < 1e-9 * abs(T[0, m])
           and np.abs(T[2:, m]).max() < 1e-9 * abs(T[0, m])]
    assert len(m_d) == 1
    m_d = m_d[0]
    cd = np.prod([T[0, m_d] * (x / a) ** 1.75 for x in al])
    den = np.prod([4 * x * x for x in al])
    for _ in range(2):
        xyz = rng.uniform(-1.0, 1.0, (4, 3))
        s = MolecularSystem.from_atoms([2, 3, 4, 5], xyz, path4)
        assert s.nbf == 20
        g = s.coulomb()
        mpmath.mp.dps = 80
        X = [mpmath.mpf(float(v)) for v in xyz.reshape(-1)]
        h = mpmath.mpf("1e-8")
        # (d_A d_B | d_C d_D) = <d_A d_C | d_B d_D>, functions 5k + m
        val = g[m_xy, 10 + m_xy, 5 + m_xy, 15 + m_xy]
        ref = float(_mixed(mpmath, G4, X, {i: 1 for i in (0, 1, 3, 4, 6, 7, 9, 10)}, h)) / den * Nxy
        print(f"(dxy dxy|dxy dxy): engine {val:.15e} reference {ref:.15e}")
        assert val == pytest.approx(ref, rel=1e-10, abs=1e-14)
        # (xx - yy) on every centre: 16 terms of second derivatives, x or y on each centre (the 2a terms cancel)
        val = g[m_d, 10 + m_d, 5 + m_d, 15 + m_d]
        tot = mpmath.mpf(0)
        for axes in np.ndindex(2, 2, 2, 2):
            sign = (-1) ** sum(axes)
            tot += sign * _mixed(mpmath, G4, X, {3 * k + ax: 2 for k, ax in enumerate(axes)}, h)
        ref = float(tot) / den * cd
        mpmath.mp.dps = 30
        print(f"(dx2-y2 ...|...): engine {val:.15e} reference {ref:.15e}")
        assert val == pytest.approx(ref, rel=1e-10, abs=1e-14)


def test_pure_d_shells_orthonormal_and_eri_symmetric():
    s, H, S, g, VN = system_ints("ccpvdz")
    assert s.nbf == 24  # pure d: 14 on O + 2 x 5 on H
    np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-14)
    # the five d functions on O are orthonormal (one-centre)
    d = slice(9, 14)
    np.testing.assert_allclose(S[d, d], np.eye(5), atol=1e-14)
    for perm in [(1, 0, 3, 2), (2, 3, 0, 1), (2, 1, 0, 3), (0, 3, 2, 1)]:
        np.testing.assert_array_equal(g, g.transpose(perm))
    w = np.linalg.eigvalsh(g.transpose(0, 2, 1, 3).reshape(s.nbf ** 2, -1))
    assert w.min() > -1e-12  # ERI matrix is positive semidefinite


def test_packed_matches_dense():
    s, H, S, g, VN = system_ints("ccpvdz")
    n = s.nbf
    pk = s.coulomb_packed()
    idx = np.array([(p, q) for p in range(n) for q in range(p + 1)])
    chem = g.transpose(0, 2, 1, 3)
    M = chem[idx[:, 0], idx[:, 1]][:, idx[:, 0], idx[:, 1]]
    np.testing.assert_array_equal(M[np.tril_indices(len(idx))], pk)


def test_many_core_oracle_is_bitwise_the_faithful_one():
    """The CPU baseline's many-core mode (OpenMP over independent outputs) reproduces the faithful
    single-core arithmetic bit for bit."""
    s, H, S, g, VN = system_ints("ccpvdz")
    n, o = s.nbf, s.nel // 2
    rng = np.random.default_rng(11)
    P = rng.standard_normal((n, n))
    P = P + P.T
    np.testing.assert_array_equal(O.get_fock_par(g, H, P), O.get_fock(g, H, P))
    C = np.linalg.qr(rng.standard_normal((n, n)))[0]
    np.testing.assert_array_equal(O.ao2mo_oovv_par(g, C, o), O.ao2mo_oovv(g, C, o))
    assert O.max_threads() >= 1


@pytest.mark.parametrize("kind", ["sto3g", "ccpvdz", "w2"])
def test_packed_oracle_is_bitwise_the_dense_one(kind):
    """The packed-ERI restatements used at n >= 144 (where the dense tensor does not fit) give bit for bit
    the rows of orc_get_fock (any exchange factor, also for a non-symmetric P) and the (ia|jb) blocks of
    orc_ao2mo_oovv (same values fetched, same summation order)."""
    s, H, S, g, VN = system_ints(kind)
    n, o = s.nbf, s.nel // 2
    pk = s.coulomb_packed()
    rng = np.random.default_rng(3)
    P = rng.standard_normal((n, n))
    rows = [0, n // 3, n - 1]
    for Pm, gen in ((P + P.T, False), (P, True)):
        F = O.get_fock(g, H, Pm, gen)
        np.testing.assert_array_equal(O.fock_rows_packed(pk, H, Pm, rows, 1.0 if gen else 0.5), F[rows])
    C = np.linalg.qr(rng.standard_normal((n, n)))[0]
    T = O.ao2mo_oovv(g, C, o)
    pairs = [(0, 0), (0, o - 1), (o - 1, 1)]
    np.testing.assert_array_equal(O.ovov_pairs_packed(pk, C, o, pairs), np.stack([T[i, j] for i, j in pairs]))


def test_dft_branch_fock():
    """getFock's DFT branch: exx = 1 with no Vxc is the Hartree-Fock Fock matrix (loops in another order,
    so to rounding); exx = 0 leaves the Coulomb part alone, linear in P with Tr(P J(P)) > 0; a symmetric
    Vxc is added unchanged."""
    s, H, S, g, VN = system_ints("ccpvdz")
    n = s.nbf
    rng = np.random.default_rng(4)
    P = rng.standard_normal((n, n))
    P = P + P.T
    F_hf = O.get_fock(g, H, P)
    np.testing.assert_allclose(O.get_fock_dft(g, H, P, 1.0), F_hf, rtol=0, atol=1e-12)
    Z = np.zeros_like(H)
    J = O.get_fock_dft(g, Z, P, 0.0)
    np.testing.assert_allclose(O.get_fock_dft(g, Z, 2 * P, 0.0), 2 * J, rtol=0, atol=1e-11)
    assert np.sum(P * J) > 0
    K = (J - O.get_fock(g, Z, P)) / 0.5  # F_hf - H = J - K/2
    np.testing.assert_allclose(O.get_fock_dft(g, H, P, 0.25), H + J - 0.125 * K, rtol=0, atol=1e-11)
    V = rng.standard_normal((n, n))
    V = V + V.T
    np.testing.assert_allclose(O.get_fock_dft(g, H, P, 0.25, V), H + J - 0.125 * K + V, rtol=0, atol=1e-11)


def test_packed_scf_is_bitwise_the_dense_one():
    """orc_scf_packed (the Fock builds read from packed ERIs, used for the committed cluster energies
    at n >= 144) runs the same SCF bit for bit as orc_scf."""
    s, H, S, g, VN = system_ints("w2")
    o = s.nel // 2
    a = O.scf(H, S, g, VN, o, threshold=1e-10)
    b = O.scf(H, S, None, VN, o, threshold=1e-10, packed=s.coulomb_packed())
    assert a["E"] == b["E"] and a["iterations"] == b["iterations"]
    np.testing.assert_array_equal(a["C"], b["C"])
    c = O.scf(H, S, g, VN, o, threshold=1e-10, par=True)
    assert c["E"] == a["E"] and np.array_equal(c["C"], a["C"])
    np.testing.assert_array_equal(O.get_fock_packed(s.coulomb_packed(), H, a["P"]), O.get_fock(g, H, a["P"]))


def test_gradient_eri_contraction(rhf):
    """The gradient's two-electron contraction (src/hartree_fock.zig:263-273) taken on g itself is the
    two-electron energy 1/2 sum P (F - H) of getFock/getEnergy (a transposed index in either term would
    give an exchange-like value instead); it is linear in the tensor and scales exchange by exx."""
    s, H, S, g, VN, r = rhf
    P = r["P"]
    F = O.get_fock(g, H, P)
    e2 = 0.5 * np.sum(P * (F - H))
    assert O.grad_eri(g[None], P)[0] == pytest.approx(e2, abs=1e-12)
    dg = np.stack([g, 2.0 * g - 0.5 * g.transpose(2, 3, 0, 1)])
    out = O.grad_eri(dg, P)
    assert out[1] == pytest.approx(1.5 * e2, abs=1e-11)  # <rs|pq> = <pq|rs>
    Fj = O.get_fock_dft(g, np.zeros_like(H), P, 0.0)  # H + J only
    ej = 0.5 * np.sum(P * Fj)
    assert O.grad_eri(g[None], P, exx=0.0)[0] == pytest.approx(ej, abs=1e-12)


def test_dual_transform(rhf):
    """ao2mo_pppp over dual numbers: the value part is ao2mo_pppp bit for bit; the derivative part is the
    directional derivative of the transform, checked against a central difference (the transform is
    quartic in C and linear in g, so the difference error is O(h^2)); zero tangents give zero."""
    s, H, S, g, VN, r = rhf
    rng = np.random.default_rng(23)
    dg = rng.standard_normal(g.shape)
    dg = dg + dg.transpose(1, 0, 3, 2)
    dg = dg + dg.transpose(2, 3, 0, 1)
    dg = dg + dg.transpose(2, 1, 0, 3)  # <pq|rs> symmetries
    dC = rng.standard_normal(r["C"].shape)
    out, dout = O.ao2mo_pppp_dual(g, dg, r["C"], dC)
    np.testing.assert_array_equal(out, O.ao2mo_pppp(g, r["C"]))
    h = 1e-4
    fd = (O.ao2mo_pppp(g + h * dg, r["C"] + h * dC) - O.ao2mo_pppp(g - h * dg, r["C"] - h * dC)) / (2 * h)
    assert np.abs(dout - fd).max() < 1e-6 * max(1.0, np.abs(dout).max())
    z = O.ao2mo_pppp_dual(g, np.zeros_like(g), r["C"], np.zeros_like(dC))[1]
    assert not z.any()


@pytest.mark.parametrize("k", [1, 2])
def test_energies_invariant_under_rotation_and_translation(k):
    """cc-pVDZ (H2O)_k rigidly rotated (seeded uniform rotations) and translated: the RHF and MP2 energies
    and the spectra of S and of the ERI supermatrix do not change.  A wrong solid-harmonic combination of
    the pure d functions, or a wrong Hermite term of their (dd|dd) recursions, makes the basis (or its
    integrals) depend on the orientation, which these catch."""
    from zinq_amd.molecules import water_cluster, BOHR_TO_ANGSTROM, _rotation
    Z, xyz = water_cluster(k)
    xyz = xyz / BOHR_TO_ANGSTROM

    def energies(coords):
        s = MolecularSystem.from_atoms(Z, coords, "builtin:cc-pvdz")
        H, S, g, VN, o = s.hcore(), s.overlap(), s.coulomb(), s.nrep(), s.nel // 2
        r = O.scf(H, S, g, VN, o, threshold=1e-11)
        e2 = O.mp2_closed(O.ao2mo_oovv(g, r["C"], o), r["e"])
        n = s.nbf
        w = np.linalg.eigvalsh(g.transpose(0, 2, 1, 3).reshape(n * n, n * n))
        return r["E"], e2, np.linalg.eigvalsh(S), w

    E0, c0, s0, w0 = energies(xyz)
    rng = np.random.default_rng(100 + k)
    centre = xyz.mean(axis=0)
    for t in range(2):
        R = _rotation(*rng.uniform(size=3))
        moved = (xyz - centre) @ R.T + centre + rng.uniform(-2.0, 2.0, size=3)
        E, c, s_, w = energies(moved)
        assert E == pytest.approx(E0, abs=1e-10)
        assert c == pytest.approx(c0, abs=1e-10)
        np.testing.assert_allclose(s_, s0, rtol=0, atol=1e-12)
        np.testing.assert_allclose(w, w0, rtol=0, atol=1e-11)
