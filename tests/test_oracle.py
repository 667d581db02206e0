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
