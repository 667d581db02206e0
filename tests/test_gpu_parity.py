"""GPU parity of the HIP hot path (through the C ABI of libzinq_hip.so) against the CPU oracle.

Bar: FP64 J/K and transforms agree with the oracle's restatement of getFock / ao2mo_oovv / ao2mo_pppp
to 1e-11 (absolute, different summation order); energies to 1e-10 Eh; the reference's published STO-3G
water energies to its own TEST_TOLERANCE 1e-8.  Larger sizes are checked through size-independent
properties (identity and orthogonal transforms, symmetry, bitwise determinism).
"""
import numpy as np
import pytest

from conftest import system_ints
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from zinq_amd import hip
    c = hip.Context(0)
    yield c
    c.close()


def _load(ctx, kind, packed=True):
    s, H, S, g, VN = system_ints(kind)
    if packed:
        ctx.eri_load_packed(s.coulomb_packed(), s.nbf)
    else:
        ctx.eri_load(g)
    return s, H, S, g, VN


def _rand_sym(n, seed):
    A = np.random.default_rng(seed).standard_normal((n, n))
    return A + A.T


@pytest.mark.parametrize("kind", ["sto3g", "ccpvdz", "w2"])
@pytest.mark.parametrize("x,gen", [(0.5, False), (1.0, True)])
def test_fock_jk_matches_oracle(ctx, kind, x, gen):
    s, H, S, g, VN = _load(ctx, kind)
    P = _rand_sym(s.nbf, 5)
    F = ctx.fock_jk(H, P, x)
    Fo = O.get_fock(g, H, P, gen)
    assert np.abs(F - Fo).max() < 1e-11
    assert np.array_equal(F, F.T)


def test_fock_jk_dense_and_packed_loads_agree(ctx):
    s, H, S, g, VN = system_ints("ccpvdz")
    P = _rand_sym(s.nbf, 6)
    ctx.eri_load(g)
    F1 = ctx.fock_jk(H, P)
    ctx.eri_load_packed(s.coulomb_packed(), s.nbf)
    F2 = ctx.fock_jk(H, P)
    assert np.array_equal(F1, F2)


def test_fock_jk_bitwise_reproducible(ctx):
    s, H, S, g, VN = _load(ctx, "w2")
    P = _rand_sym(s.nbf, 7)
    assert np.array_equal(ctx.fock_jk(H, P), ctx.fock_jk(H, P))


@pytest.mark.parametrize("kind", ["sto3g", "ccpvdz", "w2"])
def test_ao2mo_ovov_matches_oracle(ctx, kind):
    s, H, S, g, VN = _load(ctx, kind)
    o = s.nel // 2
    r = O.scf(H, S, g, VN, o, threshold=1e-10)
    T = ctx.ao2mo(r["C"], o, "ovov")
    To = O.ao2mo_oovv(g, r["C"], o)
    assert np.abs(T - To).max() < 1e-11


@pytest.mark.parametrize("kind", ["sto3g", "ccpvdz"])
def test_ao2mo_full_matches_oracle(ctx, kind):
    s, H, S, g, VN = _load(ctx, kind)
    C = np.linalg.qr(np.random.default_rng(8).standard_normal((s.nbf, s.nbf)))[0]
    assert np.abs(ctx.ao2mo(C, 0, "full") - O.ao2mo_pppp(g, C)).max() < 1e-11


def test_ao2mo_identity_returns_the_eris(ctx):
    s, H, S, g, VN = _load(ctx, "w2")
    out = ctx.ao2mo(np.eye(s.nbf), 0, "full")
    assert np.abs(out - g).max() < 1e-14


@pytest.mark.parametrize("kind", ["sto3g", "ccpvdz", "w2"])
def test_rhf_mp2_match_oracle(ctx, kind):
    s, H, S, g, VN = _load(ctx, kind)
    o = s.nel // 2
    r = ctx.scf(H, S, VN, o, threshold=1e-10)
    ro = O.scf(H, S, g, VN, o, threshold=1e-10)
    assert r["E"] == pytest.approx(ro["E"], abs=1e-10)
    e2 = ctx.mp2(r["C"], r["e"], o)
    e2o = O.mp2_closed(O.ao2mo_oovv(g, ro["C"], o), ro["e"])
    assert e2 == pytest.approx(e2o, abs=1e-10)


def test_reference_anchors_through_gpu_path(water_xyz):
    """The reference's own tests, run through zinq_amd's HF/MP mirrors (test/hartree_fock.zig:24,56;
    test/moller_plesset.zig:27)."""
    from zinq_amd import hartree_fock as hf, moller_plesset as mp
    r = hf.run(hf.Options(system=water_xyz, basis="builtin:sto-3g"))
    assert r.energy[0] == pytest.approx(-74.9659012172971900, abs=1e-8)
    g = hf.run(hf.Options(system=water_xyz, basis="builtin:sto-3g", generalized=True))
    assert g.energy[0] == pytest.approx(-74.9659012172972300, abs=1e-8)
    m = mp.run(mp.Options(hartree_fock=hf.Options(system=water_xyz, basis="builtin:sto-3g"), order=2))
    assert m.energy[0] == pytest.approx(-75.0048550569061500, abs=1e-8)


def test_scf_not_converged_maps_to_error(ctx):
    from zinq_amd import hip
    s, H, S, g, VN = _load(ctx, "sto3g")
    with pytest.raises(hip.ScfDidNotConverge):
        ctx.scf(H, S, VN, s.nel // 2, iterations=1)


def test_invalid_inputs_raise(ctx):
    from zinq_amd import hip
    s, H, S, g, VN = _load(ctx, "sto3g")
    with pytest.raises(hip.InvalidInput):
        ctx.ao2mo(np.eye(s.nbf), s.nbf, "ovov")
    fresh = hip.Context(0)
    with pytest.raises(hip.BadState):
        fresh.fock_jk(H, H)
    fresh.close()


# ---------------------------------------------------------------- reduced-precision quarter transforms
# Bars (absolute, on (ia|jb) and on E_MP2).  AO integrals reach ~5 Eh (oxygen core), so each quarter
# carries absolute rounding of ~5 x unit roundoff x sqrt(terms): FP32 (2^-24) 5e-5 / 1e-6 Eh; BF16X3
# (hi*hi + hi*lo + lo*hi of a two-term bf16 split, ~2^-16 relative per product) 5e-3 / 5e-5 Eh.
# FP32C keeps FP32 operands but sums whole K blocks in FP64, so only one block's FP32 rounding (and the
# FP32 rounding of the operands and intermediates, ~5 x 2^-24 per element) remains: 1e-5 / 5e-7 Eh,
# a factor 2 inside the mixed path's 1e-6 Eh budget.  The MP2 energy counts each i > j block of the
# symmetric transform twice (its mirror is the same block transposed), so the rounding of the two blocks
# no longer averages out: measured 2.1e-7 Eh at n = 96.
PREC_TOL = {"fp32": (5e-5, 1e-6), "bf16x3": (5e-3, 5e-5), "fp32c": (1e-5, 5e-7)}


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp32c"])
@pytest.mark.parametrize("kind", ["sto3g", "ccpvdz", "w2"])
def test_reduced_precision_ovov_and_mp2(ctx, kind, prec):
    s, H, S, g, VN = _load(ctx, kind)
    o = s.nel // 2
    r = O.scf(H, S, g, VN, o, threshold=1e-10)
    To = O.ao2mo_oovv(g, r["C"], o)
    T = ctx.ao2mo(r["C"], o, "ovov", prec)
    e2 = ctx.mp2(r["C"], r["e"], o, prec)
    e2o = O.mp2_closed(To, r["e"])
    tol_t, tol_e = PREC_TOL[prec]
    print(f"{kind} {prec}: max|d(ia|jb)|={np.abs(T - To).max():.3e} |dE_MP2|={abs(e2 - e2o):.3e}")
    assert np.abs(T - To).max() < tol_t
    assert abs(e2 - e2o) < tol_e


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "fp32c"])
def test_reduced_precision_full_transform(ctx, prec):
    s, H, S, g, VN = _load(ctx, "ccpvdz")
    C = np.linalg.qr(np.random.default_rng(9).standard_normal((s.nbf, s.nbf)))[0]
    err = np.abs(ctx.ao2mo(C, 0, "full", prec) - O.ao2mo_pppp(g, C)).max()
    print(f"full {prec}: {err:.3e}")
    assert err < PREC_TOL[prec][0] * 5


# ---------------------------------------------------------------- device-resident SCF
@pytest.mark.parametrize("kind", ["sto3g", "ccpvdz", "w2"])
def test_device_scf_matches_oracle(ctx, kind):
    s, H, S, g, VN = _load(ctx, kind)
    o = s.nel // 2
    r = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
    assert ctx.last_scf_driver() == "device"  # decided on the device, no host-synchronised restart
    # the later densities fold with certified HOMO/LUMO bounds, and still give the aufbau density
    sp = ctx.last_scf_sp2()
    print(kind, r["iterations"], sp)
    assert sp["densities"] >= r["iterations"] and 1 <= sp["folded"] < sp["densities"]
    ro = O.scf(H, S, g, VN, o, threshold=1e-10)
    assert r["E"] == pytest.approx(ro["E"], abs=1e-10)
    # geigh contract: F C = S C e, C^T S C = I, e ascending; MP2 from these orbitals agrees
    C, e = r["C"], r["e"]
    np.testing.assert_allclose(C.T @ S @ C, np.eye(s.nbf), atol=1e-10)
    np.testing.assert_allclose(r["F"] @ C, S @ C * e, atol=1e-9)
    assert np.all(np.diff(e) >= -1e-12)
    e2 = ctx.mp2(C, e, o)
    e2o = O.mp2_closed(O.ao2mo_oovv(g, ro["C"], o), ro["e"])
    assert e2 == pytest.approx(e2o, abs=1e-10)


def test_full_size_water8_device_vs_host_driver():
    """(H2O)_8 cc-pVDZ (n = 192, the largest molecule of the three-molecule bench step), where the dense
    oracle does not fit: the device-decided SCF (SP2 densities, folds, device DIIS) against the
    host-steps driver (dsyevd every iteration, DIIS solved on the host), both over the same J/K; the
    geigh contract; MP2 from both; and the mixed path within its 1e-6 Eh bar."""
    from zinq_amd import hip
    from zinq_amd.molecules import water_system
    s = water_system(8)
    n, o = s.nbf, s.nel // 2
    H, S, VN = s.hcore(), s.overlap(), s.nrep()
    with hip.Context(0) as c:
        c.eri_load_packed(s.coulomb_packed(), n)
        rh = c.scf(H, S, VN, o, threshold=1e-10)
        rd = c.scf(H, S, VN, o, threshold=1e-10, device=True)
        assert c.last_scf_driver() == "device"
        sp = c.last_scf_sp2()
        assert sp["folded"] >= 1
        print(f"n={n}: E host {rh['E']:.12f} device {rd['E']:.12f} it {rh['iterations']}/{rd['iterations']} {sp}")
        assert rd["E"] == pytest.approx(rh["E"], abs=1e-9)
        C, e = rd["C"], rd["e"]
        np.testing.assert_allclose(C.T @ S @ C, np.eye(n), atol=1e-9)
        np.testing.assert_allclose(rd["F"] @ C, S @ C * e, atol=1e-8)
        np.testing.assert_allclose(e, rh["e"], atol=1e-8)
        e2d, e2h = c.mp2(C, e, o), c.mp2(rh["C"], rh["e"], o)
        assert e2d == pytest.approx(e2h, abs=1e-9)
        rm = c.scf(H, S, VN, o, threshold=1e-6, device=True, fock_precision="fp32")
        e2m = c.mp2(rm["C"], rm["e"], o, "fp32c")
        print(f"mixed dE_MP2 = {rm['E'] + e2m - rd['E'] - e2d:+.2e}")
        assert abs(rm["E"] + e2m - rd["E"] - e2d) < 1e-6


def test_device_scf_ghf_and_reference_anchor(ctx):
    from zinq_amd.system import spin_block_1e, spin_block_2e
    s, H, S, g, VN = system_ints("sto3g")
    Hs, Ss, gs = spin_block_1e(H), spin_block_1e(S), spin_block_2e(g)
    ctx.eri_load(gs)
    r = ctx.scf(Hs, Ss, VN, s.nel, generalized=True, device=True)
    assert r["E"] == pytest.approx(-74.9659012172972300, abs=1e-8)  # test/hartree_fock.zig:56
    ctx.eri_load_packed(s.coulomb_packed(), s.nbf)
    r = ctx.scf(H, S, VN, s.nel // 2, device=True)
    assert r["E"] == pytest.approx(-74.9659012172971900, abs=1e-8)  # test/hartree_fock.zig:24


def test_device_scf_no_diis_and_not_converged(ctx):
    from zinq_amd import hip
    s, H, S, g, VN = _load(ctx, "sto3g")
    o = s.nel // 2
    r = ctx.scf(H, S, VN, o, diis=None, device=True)
    ro = O.scf(H, S, g, VN, o, diis=None)
    assert r["E"] == pytest.approx(ro["E"], abs=1e-9)
    with pytest.raises(hip.ScfDidNotConverge):
        ctx.scf(H, S, VN, o, iterations=1, device=True)


def test_device_scf_graph_replays_match_eager(ctx):
    """The captured SCF iteration (one hipGraph per context, replayed every iteration and reused by later
    calls) gives bitwise the same SCF as the eagerly launched kernels (profiling on disables capture),
    call after call, with eager work in between."""
    s, H, S, g, VN = _load(ctx, "w2")
    o = s.nel // 2
    ctx.profile(True)
    eager = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
    ctx.profile(False)
    assert ctx.last_scf_driver() == "device"
    for _ in range(3):
        r = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
        assert ctx.last_scf_driver() == "device"
        assert r["E"] == eager["E"] and r["iterations"] == eager["iterations"]
        np.testing.assert_array_equal(r["P"], eager["P"])
        np.testing.assert_array_equal(r["F"], eager["F"])
        ctx.fock_jk(H, r["P"])  # eager kernels on the same stream between replays


def test_device_scf_sp2_budget_fallback(ctx):
    """A too small SP2 step budget ends the asynchronous run (state 'SP2 did not settle'); the driver
    then reruns the host-synchronised SCF and the result is unchanged."""
    from zinq_amd import hip
    s, H, S, g, VN = _load(ctx, "ccpvdz")
    o = s.nel // 2
    ref = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
    assert ctx.last_scf_driver() == "device"
    ctx.set_sp2_budget(4)
    r = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
    assert ctx.last_scf_driver() == "synced"
    ctx.set_sp2_budget(100)
    assert r["E"] == pytest.approx(ref["E"], abs=1e-11)
    assert r["iterations"] == ref["iterations"]
    with pytest.raises(hip.InvalidInput):
        ctx.set_sp2_budget(0)
    with pytest.raises(hip.InvalidInput):
        ctx.set_sp2_budget(129)


def test_device_scf_unsettled_guess_stops_the_replays(ctx):
    """A guess density whose SP2 does not settle (a 2-step budget) marks the state before the first
    iteration; that iteration's convergence kernel hands the mark to the host, which stops replaying
    after at most two iterations and reruns the host-synchronised driver.  The result equals the oracle's
    SCF; a plain run enqueues one iteration past the converged one."""
    from zinq_amd import hip
    s, H, S, g, VN = _load(ctx, "ccpvdz")
    o = s.nel // 2
    ro = O.scf(H, S, g, VN, o, threshold=1e-10)
    ref = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
    assert ctx.last_scf_driver() == "device"
    assert ref["iterations"] + 1 <= ctx.scf_replays() <= ref["iterations"] + 2
    ctx.set_sp2_guess_budget(2)
    try:
        r = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
        replays = ctx.scf_replays()
    finally:
        ctx.set_sp2_guess_budget(128)
        ctx.set_sp2_budget(100)
    print(f"unsettled guess: {replays} replays, driver {ctx.last_scf_driver()}, dE vs oracle {r['E'] - ro['E']:+.1e}",
          flush=True)
    assert ctx.last_scf_driver() == "synced"
    assert 1 <= replays <= 2
    assert r["E"] == pytest.approx(ro["E"], abs=1e-10)
    assert r["iterations"] == ref["iterations"]
    for bad in (0, 129):
        with pytest.raises(hip.InvalidInput):
            ctx.set_sp2_guess_budget(bad)


def test_device_scf_singular_diis_resets(ctx):
    """The device DIIS reset on a singular system (src/hartree_fock.zig:821-845), in both drivers.  With
    the stored error vectors scaled by 0, every system [B -1; -1 0] from the second iteration on has a
    zero B block and is singular: each iteration cuts the history to its own entry and keeps F
    unextrapolated, so the SCF is the oracle's SCF without DIIS (48 iterations for cc-pVDZ water,
    against 16 with DIIS), with one reset per iteration after the first.  Scale 1 gives the DIIS SCF
    back with no reset."""
    from zinq_amd import hip
    s, H, S, g, VN = _load(ctx, "ccpvdz")
    o = s.nel // 2
    plain = O.scf(H, S, g, VN, o, diis=None, threshold=1e-10)
    with_diis = O.scf(H, S, g, VN, o, threshold=1e-10)
    ctx.set_diis_error_scale(0.0)
    try:
        r = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
        runs = [("device", r, ctx.last_scf_driver(), ctx.last_diis_resets())]
        ctx.set_sp2_guess_budget(2)  # an unsettled guess: the host-synchronised driver finishes the run
        try:
            r = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
            runs.append(("synced", r, ctx.last_scf_driver(), ctx.last_diis_resets()))
        finally:
            ctx.set_sp2_guess_budget(128)
            ctx.set_sp2_budget(100)
    finally:
        ctx.set_diis_error_scale(1.0)
    for want, r, driver, resets in runs:
        print(f"zeroed DIIS errors, {driver} driver: {r['iterations']} iterations (oracle without DIIS "
              f"{plain['iterations']}), {resets} resets, dE {r['E'] - plain['E']:+.1e}", flush=True)
        assert driver == want
        assert abs(r["iterations"] - plain["iterations"]) <= 1
        assert resets == r["iterations"] - 1
        assert r["E"] == pytest.approx(plain["E"], abs=1e-10)
    r = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
    assert ctx.last_scf_driver() == "device" and ctx.last_diis_resets() == 0
    assert abs(r["iterations"] - with_diis["iterations"]) <= 1
    assert r["E"] == pytest.approx(with_diis["E"], abs=1e-10)
    for bad in (-1.0, 2.0, float("nan")):
        with pytest.raises(hip.InvalidInput):
            ctx.set_diis_error_scale(bad)


def test_mp2_resident_matches_mp2(ctx):
    """zq_mp2_resident: the MP2 energy from the orbitals and energies the device SCF leaves on the device
    is bitwise zq_mp2's on the C and e that SCF returned, in every precision, after the plain and the
    mixed SCF; fresh ERIs and a failed SCF leave no result (ZQ_ERR_STATE)."""
    from zinq_amd import hip
    for kind in ("ccpvdz", "w2"):
        s, H, S, g, VN = _load(ctx, kind)
        o = s.nel // 2
        with pytest.raises(hip.BadState):
            ctx.mp2_resident(o)
        for fock, thr in (("fp64", 1e-10), ("fp32", 1e-6)):
            r = ctx.scf(H, S, VN, o, threshold=thr, device=True, fock_precision=fock)
            for prec in ("fp64", "fp32", "fp32c", "bf16x3"):
                a, b = ctx.mp2(r["C"], r["e"], o, prec), ctx.mp2_resident(o, prec)
                assert a == b, (kind, fock, prec, a, b)
    with pytest.raises(hip.ScfDidNotConverge):
        ctx.scf(H, S, VN, o, iterations=1, device=True)
    with pytest.raises(hip.BadState):
        ctx.mp2_resident(o)


@pytest.mark.parametrize("kind", ["ccpvdz", "w2"])
def test_mixed_path_within_bar(ctx, kind):
    """Mixed path: incremental Fock builds from FP32 ERIs to 1e-6, one full FP64 build, FP32C transform.
    Bar: |dE_MP2| <= 1e-6 Eh against the FP64 oracle (the total MP2 energy and its correlation part)."""
    s, H, S, g, VN = _load(ctx, kind)
    o = s.nel // 2
    ro = O.scf(H, S, g, VN, o, threshold=1e-10)
    e2o = O.mp2_closed(O.ao2mo_oovv(g, ro["C"], o), ro["e"])
    r = ctx.scf(H, S, VN, o, threshold=1e-6, device=True, fock_precision="fp32")
    assert ctx.last_scf_driver() == "device"  # the FP32 incremental builds ran (no FP64 restart)
    e2 = ctx.mp2(r["C"], r["e"], o, "fp32c")
    print(f"{kind} mixed: dE_rhf={r['E'] - ro['E']:+.2e} dE_corr={e2 - e2o:+.2e} it={r['iterations']}")
    assert abs(r["E"] - ro["E"]) < 1e-7
    assert abs(e2 - e2o) < 1e-6
    assert abs(r["E"] + e2 - ro["E"] - e2o) < 1e-6
    # the FP32 ERI copy stays resident; a plain FP64 SCF afterwards is unaffected
    r64 = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
    assert r64["E"] == pytest.approx(ro["E"], abs=1e-10)


# ---------------------------------------------------------------- grouped many-molecule path
def _batch_systems(count, seed=77):
    from zinq_amd.molecules import batch_specs, batch_system
    return [batch_system(s) for s in batch_specs(count, seed=seed)]


def test_batch_matches_single_molecule_path():
    """Grouped J/K + per-molecule workgroup SCF + batched dsyevd + MP2 equals the single-molecule
    device path (same FP64 arithmetic; reductions differ in order only)."""
    from zinq_amd import hip
    from zinq_amd.batch import Batch
    systems = _batch_systems(10)
    b = Batch(0)
    for s in systems:
        b.add_system(s)
    r = b.run(threshold=1e-10)
    assert (r["status"] == 0).all(), r["status"]
    with hip.Context(0) as ctx:
        for m, s in enumerate(systems):
            ctx.eri_load_packed(s.coulomb_packed(), s.nbf)
            o = s.nel // 2
            ref = ctx.scf(s.hcore(), s.overlap(), s.nrep(), o, threshold=1e-10, device=True)
            e2 = ctx.mp2(ref["C"], ref["e"], o)
            print(f"mol {m} n={s.nbf} dE={r['E_rhf'][m] - ref['E']:+.1e} dEc={r['E_corr'][m] - e2:+.1e} "
                  f"it={r['iterations'][m]}/{ref['iterations']}")
            assert r["E_rhf"][m] == pytest.approx(ref["E"], abs=1e-10)
            assert r["E_corr"][m] == pytest.approx(e2, abs=1e-10)
            e, C = b.orbitals(m)
            np.testing.assert_allclose(e, ref["e"], atol=1e-8)
    b.close()


def test_batch_result_same_alone_and_in_the_mix():
    """Each molecule's numbers do not depend on which other molecules share the batch: bitwise."""
    from zinq_amd.batch import Batch
    systems = _batch_systems(6, seed=91)
    mix = Batch(0)
    for s in systems:
        mix.add_system(s)
    rm = mix.run(threshold=1e-10)
    for m in (0, 3, 5):
        alone = Batch(0)
        alone.add_system(systems[m])
        ra = alone.run(threshold=1e-10)
        assert ra["E_rhf"][0] == rm["E_rhf"][m] and ra["E_corr"][0] == rm["E_corr"][m]
        assert ra["iterations"][0] == rm["iterations"][m]
        alone.close()
    mix.close()


def test_batch_autotune_and_rerun():
    from zinq_amd.batch import Batch
    systems = _batch_systems(8, seed=5)
    b = Batch(0)
    for s in systems:
        b.add_system(s)
    r0 = b.run(threshold=1e-10)
    chosen = b.autotune()
    assert chosen["item_quartets"] in (1, 2, 4, 8, 16) and chosen["workgroup"] in (256, 512)
    r1 = b.run(threshold=1e-10)
    np.testing.assert_allclose(r1["E_rhf"], r0["E_rhf"], atol=1e-11)
    np.testing.assert_allclose(r1["E_corr"], r0["E_corr"], atol=1e-11)
    b.close()


def test_batch_grouped_mp2_is_zq_mp2():
    """The batch's MP2 stage (one grouped launch per transform stage over all molecules, C and e on the
    device) gives, for every molecule, bitwise the energy zq_mp2 gives for the same orbitals: FP64, FP32
    and FP32C grouped, BF16X3 on per-molecule streams.  (H2O)_4 cc-pVDZ (n = 96, o = 20) puts two 16-row
    tiles of occupied orbitals in the grouped Q1 where the single-molecule launch takes one workgroup of
    two tiles."""
    from zinq_amd import hip
    from zinq_amd.batch import Batch
    from zinq_amd.molecules import water_system
    systems = _batch_systems(7, seed=3) + [water_system(4)]
    b = Batch(0)
    for s in systems:
        b.add_system(s)
    with hip.Context(0) as ctx:
        for prec in ("fp64", "fp32", "fp32c", "bf16x3"):
            r = b.run(threshold=1e-10, mp2_precision=prec)
            assert (r["status"] == 0).all(), r["status"]
            t = b.timings()
            print(f"{prec}: stage ms {t}", flush=True)
            assert t["total"] >= t["scf"] + t["eigen"] + t["mp2"] - 1e-3
            for m, s in enumerate(systems):
                ctx.eri_load_packed(s.coulomb_packed(), s.nbf)
                e, C = b.orbitals(m)
                e2 = ctx.mp2(C, e, s.nel // 2, prec)
                assert r["E_corr"][m] == e2, (prec, m, s.nbf, r["E_corr"][m] - e2)
    b.close()


# ---------------------------------------------------------------- serving
def _serve_molecules():
    from zinq_amd.molecules import water_system
    from zinq_amd.serve import Molecule
    return [Molecule(f"w{k}", water_system(k)) for k in (1, 2, 3)] + \
           [Molecule("s2", water_system(2, basis="builtin:sto-3g"))]


def test_serve_same_alone_and_in_the_mix():
    """Requests overlapping on 4 slots (continuous batching) give bitwise the result each molecule gets
    when served alone; also under a budget so tight that idle resident sets are evicted and rebuilt."""
    from zinq_amd.serve import Server, poisson_arrivals, eri_set_bytes, scratch_bytes
    mols = _serve_molecules()
    keys = [m.key for m in mols]
    alone = {}
    for m in mols:
        srv = Server(mols, slots=1)
        rec, _ = srv.serve([(0.0, m.key)])
        srv.close()
        assert "error" not in rec[0], rec[0]
        alone[m.key] = rec[0]
    reqs = poisson_arrivals(500.0, 16, keys, seed=3)
    srv = Server(mols, slots=4)
    rec, wall = srv.serve(reqs)
    srv.close()
    assert len({r["slot"] for r in rec}) > 1
    for r in rec:
        assert "error" not in r, r
        a = alone[r["key"]]
        assert r["E_mp2"] == a["E_mp2"] and r["iterations"] == a["iterations"], (r, a)
    big = max(eri_set_bytes(m.n, True) for m in mols)
    tight = 2 * max(scratch_bytes(m.n, m.o, 4) for m in mols) + int(1.5 * big)
    srv = Server(mols, slots=2, hbm_budget=tight)
    rec, _ = srv.serve(poisson_arrivals(500.0, 10, keys, seed=4))
    srv.close()
    for r in rec:
        assert "error" not in r, r
        assert r["E_mp2"] == alone[r["key"]]["E_mp2"]


def test_serve_burst_reattach_eager():
    """A burst of 64 requests on 8 slots whose contexts switch molecules between requests (each attach
    points the context at the set's J/K plan and zeroes its P), launched eagerly (profiling on: no
    captured graph), so that a context's first kernels run while other slots keep the device busy.
    Every result equals bitwise the molecule's result served alone.  Regression: the attach's P fill
    once went to the null stream, which is not ordered before the non-blocking context stream; landing
    after the first density, it made the incremental mixed-path Fock builds add G(P_0) twice, and the
    SCF converged to a wrong state."""
    from zinq_amd.serve import Server
    mols = _serve_molecules()
    keys = [m.key for m in mols]
    alone = {}
    srv = Server(mols, slots=1)
    for m in mols:
        rec, _ = srv.serve([(0.0, m.key)])
        assert "error" not in rec[0], rec[0]
        alone[m.key] = rec[0]["E_mp2"]
    srv.close()
    pick = np.random.default_rng(11).integers(0, len(keys), size=64)
    srv = Server(mols, slots=8)
    for c in srv.ctx:
        c.profile(True)
    rec, _ = srv.serve([(0.0, keys[int(j)]) for j in pick])
    srv.close()
    bad = [(r["key"], r["slot"], r.get("iterations"), r.get("E_mp2", r.get("error"))) for r in rec
           if "error" in r or r["E_mp2"] != alone[r["key"]]]
    assert not bad, bad


def test_attach_switch_back_replays_cached_iteration():
    """One context attached to set A, then B, then A again (plain and mixed Fock), A the larger molecule
    so that B's run grows no buffer: the SCF iteration captured for A is kept across the switch and
    replayed (the second A run captures nothing), A's results are bitwise those of its first run, and
    B's equal those of a fresh context."""
    from zinq_amd import hip
    mols = _serve_molecules()
    a, b = mols[2], mols[1]  # (H2O)_3 then (H2O)_2: the switch back to A finds every buffer large enough
    assert a.n > b.n
    sa, sb = hip.EriSet(a.packed, a.n, 0, True), hip.EriSet(b.packed, b.n, 0, True)
    try:
        for fock in ("fp64", "fp32"):
            def run(c, m, s):
                c.attach(s)
                r = c.scf(m.H, m.S, m.VN, m.o, diis=8, threshold=1e-8, device=True, fock_precision=fock)
                assert c.last_scf_driver() == "device"
                return r["E"], r["iterations"], r["C"]
            with hip.Context(0) as c, hip.Context(0) as fresh:
                first = run(c, a, sa)
                assert c.scf_captures() == 1
                mid = run(c, b, sb)
                assert c.scf_captures() == 2
                again = run(c, a, sa)
                assert c.scf_captures() == 2, "A's cached iteration was captured again"
                ref_b = run(fresh, b, sb)
            assert first[:2] == again[:2] and np.array_equal(first[2], again[2]), (fock, first[:2], again[:2])
            assert mid[:2] == ref_b[:2], (fock, mid[:2], ref_b[:2])
    finally:
        sa.close()
        sb.close()


def test_serve_mixed_results_within_bar():
    """Served mixed-path energies against the plain FP64 device path: |dE_MP2| <= 1e-6 Eh."""
    from zinq_amd import hip
    from zinq_amd.serve import Server
    mols = _serve_molecules()[:3]
    srv = Server(mols, slots=3)
    rec, _ = srv.serve([(0.0, m.key) for m in mols])
    srv.close()
    with hip.Context(0) as ctx:
        for m, r in zip(mols, rec):
            ctx.eri_load_packed(m.packed, m.n)
            ref = ctx.scf(m.H, m.S, m.VN, m.o, threshold=1e-10, device=True)
            e2 = ctx.mp2(ref["C"], ref["e"], m.o)
            print(f"{m.key}: served dE_MP2 = {r['E_mp2'] - ref['E'] - e2:+.2e}")
            assert abs(r["E_mp2"] - ref["E"] - e2) < 1e-6


def test_fock_jk_multi_matches_single_builds(ctx):
    """Several densities in one ERI pass equal the single-density builds bitwise; H = None gives the
    bare two-electron part G(P) used by the CPHF response contraction."""
    s, H, S, g, VN = _load(ctx, "w2")
    rng = np.random.default_rng(21)
    Ps = rng.standard_normal((3, s.nbf, s.nbf))
    Ps = Ps + Ps.transpose(0, 2, 1)
    F = ctx.fock_jk_multi(H, Ps)
    G = ctx.fock_jk_multi(None, Ps, x_exch=1.0)
    for d in range(3):
        np.testing.assert_array_equal(F[d], ctx.fock_jk(H, Ps[d]))
        np.testing.assert_allclose(G[d], O.get_fock(g, np.zeros_like(H), Ps[d], generalized=True), atol=1e-11)


def _check_eigh(A, W, U):
    n = A.shape[0]
    nrm = max(1.0, np.abs(A).sum(1).max())
    tol = 1e-13 * max(n, 8) * nrm
    assert np.all(np.diff(W) >= 0)
    assert np.allclose(W, np.linalg.eigvalsh(A), rtol=0, atol=tol)
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-13 * max(n, 8)
    assert np.abs((U * W) @ U.T - A).max() < tol


@pytest.mark.parametrize("n", [1, 2, 3, 17, 33, 64, 96, 150, 192, 200])
def test_eigh_matches_lapack(ctx, n):
    """zq_eigh (the eigh of src/linear_algebra.zig:45-71) against numpy's LAPACK: eigenvalues, U^T U = I
    and U diag(W) U^T = A to 1e-13 max(n, 8) ||A||_inf.  Sizes cover every register tiling of the
    one-workgroup reduction (n <= 192); n = 200 takes the dsyevd path."""
    A = _rand_sym(n, 11 + n)
    W, U = ctx.eigh(A)
    _check_eigh(A, W, U)


def test_eigh_degenerate_and_fock_like_spectrum(ctx):
    """Exactly repeated eigenvalues and 1e-10 splittings in a Fock-like spectrum (deep core, valence,
    virtual bands): the reduction must leave the invariant subspaces intact."""
    n = 150
    Q = np.linalg.qr(np.random.default_rng(5).standard_normal((n, n)))[0]
    lam = np.concatenate([np.full(8, -20.55), -20.5 + 1e-10 * np.arange(4), np.linspace(-1.4, -0.4, 30),
                          np.repeat([0.2, 0.5, 0.9], 6), np.linspace(1.0, 6.0, n - 60)])
    A = (Q * lam) @ Q.T
    A = (A + A.T) / 2
    W, U = ctx.eigh(A)
    _check_eigh(A, W, U)
