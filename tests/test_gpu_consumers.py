"""GPU consumers of the FULL transform and the other dense-ERI contractions, through the C ABI, against the
reference's published energies (TEST_TOLERANCE 1e-8) and the CPU oracle (1e-10 Eh energies, 1e-12 on
one-electron transforms, 1e-11 on Fock matrices).

- CIS / CISD (computeCiStates, src/configuration_interaction.zig:597-683) restricted and generalized:
  test/configuration_interaction.zig:27-31,66-70,105-109.
- MP2 / MP3 / MP4 by the determinant recursion (mp, src/moller_plesset.zig:340-481), restricted and
  generalized: test/moller_plesset.zig:27,62,96,131,165,200; cc-pVDZ water MP2 over 32016
  determinants against the oracle and the closed-shell sum.
- ao2mo_pp (src/integral_transform.zig:12-34).
- getFock's DFT branch (src/hartree_fock.zig:701-719): Coulomb + exx-scaled exchange + Vxc.
"""
import numpy as np
import pytest

from conftest import system_ints
from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-8

RCIS = [-74.9659012172972900, -74.5822026772564000, -74.5822026772563200, -74.5822026772562900, -74.5070515329961900]
GCIS = [-74.9659012172973100, -74.5822026772563600, -74.5822026772563500, -74.5822026772562800, -74.5070515329962000]
RCISD = [-75.0195290856878400, -74.6224197038007400, -74.6224197038007200, -74.6224197038007200, -74.5642383868465500]
RMP = {2: -75.0048550569061500, 3: -75.0154555119455900, 4: -75.0187039890576300}
GMP = {2: -75.0048550569062000, 3: -75.0154555119455300, 4: -75.0187039890575700}


@pytest.fixture(scope="module")
def ctx():
    from zinq_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def _rhf(ctx, kind="sto3g", diis=8):
    s, H, S, g, VN = system_ints(kind)
    ctx.eri_load_packed(s.coulomb_packed(), s.nbf)
    o = s.nel // 2
    r = ctx.scf(H, S, VN, o, diis=diis, threshold=1e-10)
    return s, H, g, VN, o, r


def _ghf(ctx):
    from zinq_amd.system import spin_block_1e, spin_block_2e
    s, H, S, g, VN = system_ints("sto3g")
    Hs, Ss, gs = spin_block_1e(H), spin_block_1e(S), spin_block_2e(g)
    ctx.eri_load(gs)
    r = ctx.scf(Hs, Ss, VN, s.nel, threshold=1e-10, generalized=True)
    return s, Hs, gs, VN, s.nel, r


def test_rcis_and_rcisd(ctx):
    for exc, ref, diis in (([1], RCIS, 8), ([1, 2], RCISD, None)):
        s, H, g, VN, o, r = _rhf(ctx, diis=diis)
        E, nd = ctx.ci(H, r["C"], o, exc, VN)
        Eo = O.ci(g, H, r["C"], o, exc, VN)
        print(f"R excitations {exc}: {nd} determinants, max|dE| vs oracle {np.abs(E - Eo).max():.1e}", flush=True)
        assert nd == {1: 41, 2: 311}[max(exc)]
        np.testing.assert_allclose(E, ref, rtol=0, atol=TOL)
        np.testing.assert_allclose(E, Eo, rtol=0, atol=1e-10)


def test_gcis(ctx):
    s, H, g, VN, nel, r = _ghf(ctx)
    E, nd = ctx.ci(H, r["C"], nel, [1], VN, generalized=True)
    Eo = O.ci(g, H, r["C"], nel, [1], VN, generalized=True)
    np.testing.assert_allclose(E, GCIS, rtol=0, atol=TOL)
    np.testing.assert_allclose(E, Eo, rtol=0, atol=1e-10)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_rmp_and_gmp_orders(ctx, order):
    for gen in (False, True):
        s, H, g, VN, o, r = _ghf(ctx) if gen else _rhf(ctx)
        E, nd = ctx.mp(r["C"], r["e"], o, order, generalized=gen)
        Eo = O.mp(order, g, r["C"], r["e"], o, generalized=gen)
        print(f"order {order} generalized={gen}: {nd} determinants, E={r['E'] + E[2:].sum():.13f}, "
              f"max|dE_k| vs oracle {np.abs(E - Eo).max():.1e}", flush=True)
        np.testing.assert_allclose(E, Eo, rtol=0, atol=1e-10)
        assert r["E"] + E[2:].sum() == pytest.approx((GMP if gen else RMP)[order], abs=TOL)


def test_mp_mirror_runs_the_reference_tests(water_xyz):
    """The reference's own MP3 / GMP4 tests through zinq_amd's Options / run mirror."""
    from zinq_amd import hartree_fock as hf, moller_plesset as mp
    m = mp.run(mp.Options(hartree_fock=hf.Options(system=water_xyz, basis="builtin:sto-3g"), order=3))
    assert m.energy[0] == pytest.approx(RMP[3], abs=TOL)  # test/moller_plesset.zig:96
    m = mp.run(mp.Options(hartree_fock=hf.Options(system=water_xyz, basis="builtin:sto-3g", generalized=True), order=4))
    assert m.energy[0] == pytest.approx(GMP[4], abs=TOL)  # test/moller_plesset.zig:200


def test_ci_mirror_runs_the_reference_test(water_xyz):
    from zinq_amd import configuration_interaction as ci, hartree_fock as hf
    r = ci.run(ci.Options(hartree_fock=hf.Options(system=water_xyz, basis="builtin:sto-3g", diis=None)))
    np.testing.assert_allclose(r.energy, RCISD, rtol=0, atol=TOL)  # test/configuration_interaction.zig:105-109


def test_ccpvdz_mp2_by_determinants(ctx):
    """cc-pVDZ water, 48 spin orbitals: order 2 over the 32016 singles and doubles (row 0 of H_CI) against
    the oracle's determinant path and the closed-shell sum of zq_mp2."""
    s, H, g, VN, o, r = _rhf(ctx, "ccpvdz")
    E, nd = ctx.mp(r["C"], r["e"], o, 2)
    assert nd == 32016
    Eo = O.mp(2, g, r["C"], r["e"], o)
    print(f"cc-pVDZ MP2 by determinants: dE2 vs oracle {E[2] - Eo[2]:+.1e}, vs zq_mp2 "
          f"{E[2] - ctx.mp2(r['C'], r['e'], o):+.1e}", flush=True)
    np.testing.assert_allclose(E, Eo, rtol=0, atol=1e-10)
    assert E[2] == pytest.approx(ctx.mp2(r["C"], r["e"], o), abs=1e-10)


def test_ao2mo_pp(ctx):
    s, H, S, g, VN = system_ints("ccpvdz")
    C = np.linalg.qr(np.random.default_rng(13).standard_normal((s.nbf, s.nbf)))[0]
    assert np.abs(ctx.ao2mo_pp(H, C) - O.ao2mo_pp(H, C)).max() < 1e-12


def test_dft_branch_coulomb_and_scaled_exchange(ctx):
    """getFock's DFT branch F = H + J - x exx K + Vxc (symmetric Vxc), symmetrised: zq_fock_jk with
    x_exch = exx / 2 on H + Vxc."""
    s, H, S, g, VN = system_ints("ccpvdz")
    ctx.eri_load_packed(s.coulomb_packed(), s.nbf)
    rng = np.random.default_rng(17)
    P = rng.standard_normal((s.nbf, s.nbf))
    P = P + P.T
    V = rng.standard_normal((s.nbf, s.nbf))
    V = V + V.T
    for exx in (0.0, 0.2, 0.25, 1.0):
        F = ctx.fock_jk(H + V, P, 0.5 * exx)
        Fo = O.get_fock_dft(g, H, P, exx, V)
        assert np.abs(F - Fo).max() < 1e-11, exx


def _derivative_eris(s, basis, h=1e-4):
    """Central differences (step h bohr) of the ERIs with respect to every nuclear coordinate: dense
    (for the oracle) and canonical packed (for the GPU) stacks.  Both contractions read the same numbers,
    so the comparison is of the contraction alone."""
    from zinq_amd.system import MolecularSystem
    dense, packed = [], []
    for x in range(3 * len(s.atoms)):
        xyz = s.coors.copy()
        xyz[x] += h
        sp = MolecularSystem.from_atoms(s.atoms, xyz, basis)
        xyz[x] -= 2 * h
        sm = MolecularSystem.from_atoms(s.atoms, xyz, basis)
        dense.append((sp.coulomb() - sm.coulomb()) / (2 * h))
        packed.append((sp.coulomb_packed() - sm.coulomb_packed()) / (2 * h))
    return np.stack(dense), np.stack(packed)


@pytest.mark.parametrize("basis,kind", [("builtin:sto-3g", "sto3g"), ("builtin:cc-pvdz", "ccpvdz")])
def test_gradient_eri_contraction(ctx, basis, kind):
    """The two-electron part of the HF nuclear gradient (src/hartree_fock.zig:263-273) for the 9 water
    coordinates, RHF exchange (exx 1) and a hybrid-like exx = 0.2: against the oracle to 1e-10."""
    s, H, g, VN, o, r = _rhf(ctx, kind)
    dgd, dgp = _derivative_eris(s, basis)
    for exx in (1.0, 0.2):
        out = ctx.grad_eri(dgp, s.nbf, r["P"], 0.5 * exx)
        ref = O.grad_eri(dgd, r["P"], exx=exx)
        print(f"{kind} exx={exx}: max|d g_G| = {np.abs(out - ref).max():.2e} (|g_G| <= {np.abs(ref).max():.2e})", flush=True)
        np.testing.assert_allclose(out, ref, rtol=0, atol=1e-10)
    # the context's own ERIs stay loaded
    assert np.abs(ctx.fock_jk(H, r["P"]) - O.get_fock(g, H, r["P"])).max() < 1e-11


@pytest.mark.parametrize("kind,basis", [("sto3g", "builtin:sto-3g"), ("ccpvdz", "builtin:cc-pvdz")])
def test_dual_number_transform(ctx, kind, basis):
    """The transform over dual numbers of mp's gradient path (src/moller_plesset.zig:304-333): ERIs + eps
    (their derivative with respect to one nuclear coordinate), C + eps dC.  FULL and OVOV blocks against
    the oracle's dual ao2mo_pppp (value 1e-11; derivative 1e-10 relative to its largest element)."""
    from zinq_amd import hip
    s, H, g, VN, o, r = _rhf(ctx, kind)
    dgd, dgp = _derivative_eris(s, basis)
    x = 2  # the z coordinate of the oxygen
    dC = np.random.default_rng(31).standard_normal(r["C"].shape) * 0.1
    ref, dref = O.ao2mo_pppp_dual(g, dgd[x], r["C"], dC)
    dset = hip.EriSet(dgp[x], s.nbf, 0, False)
    try:
        out, dout = ctx.ao2mo_dual(dset, r["C"], dC)
        scale = max(1.0, np.abs(dref).max())
        print(f"{kind} FULL: max|d out| {np.abs(out - ref).max():.1e}, max|d dout| {np.abs(dout - dref).max():.1e} "
              f"(|dout| <= {np.abs(dref).max():.1e})", flush=True)
        assert np.abs(out - ref).max() < 1e-11
        assert np.abs(dout - dref).max() < 1e-10 * scale
        out2, dout2 = ctx.ao2mo_dual(dset, r["C"], dC, o, "ovov")
        assert np.abs(out2 - ref[:o, :o, o:, o:]).max() < 1e-11
        assert np.abs(dout2 - dref[:o, :o, o:, o:]).max() < 1e-10 * scale
    finally:
        dset.close()
