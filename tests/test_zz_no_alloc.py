"""zq_mp2_noalloc (the MP2 energy from the library's static device scratch, no allocation) against zq_mp2
through the C ABI: two calls in a row on one GPU, the second with a smaller molecule, in FP64, BF16X3
and FP32C; the context's and the device's memory unchanged by the calls; the size and state errors.
Both entry points run the same kernels on the same shapes.  Bars: FP64 1e-12 Eh; the reduced-precision
forms bitwise."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _prepared(ctx, k):
    from zinq_amd.molecules import water_system
    s = water_system(k)
    o = s.nel // 2
    ctx.eri_load_packed(s.coulomb_packed(), s.nbf)
    r = ctx.scf(s.hcore(), s.overlap(), s.nrep(), o, threshold=1e-10, device=True)
    return r["C"], r["e"], o


def test_mp2_noalloc_matches_mp2():
    from zinq_amd import hip
    with hip.Context(0) as big, hip.Context(0) as small:
        runs = [(big, *_prepared(big, 3)), (small, *_prepared(small, 1))]  # n = 72, then n = 24
        precs = ("fp64", "bf16x3", "fp32c")
        ref = {(i, p): c.mp2(C, e, o, p) for i, (c, C, e, o) in enumerate(runs) for p in precs}
        for p in precs:  # first launches of this entry point's own kernels, before memory is compared
            runs[1][0].mp2_noalloc(*runs[1][1:], p)
        for p in precs:
            before = [c.device_bytes() for c, *_ in runs]
            free0 = hip.mem_info(0)[0]
            got = [c.mp2_noalloc(C, e, o, p) for c, C, e, o in runs]  # two calls in a row, large then small
            assert hip.mem_info(0)[0] == free0
            assert [c.device_bytes() for c, *_ in runs] == before
            for i, g in enumerate(got):
                print(f"{p} run {i}: noalloc {g:.15f} mp2 {ref[(i, p)]:.15f}", flush=True)
                if p == "fp64":
                    assert abs(g - ref[(i, p)]) < 1e-12
                else:
                    assert g == ref[(i, p)]
        # again in the other order: nothing left over from the larger call changes the smaller one
        c, C, e, o = runs[1]
        assert abs(c.mp2_noalloc(C, e, o) - ref[(1, "fp64")]) < 1e-12


def test_mp2_noalloc_errors():
    from zinq_amd import hip
    with hip.Context(0) as ctx:
        C, e, o = _prepared(ctx, 1)
        with pytest.raises(hip.BadState):  # no FP32 ERI copy resident yet
            ctx.mp2_noalloc(C, e, o, "fp32")
        C, e, o = _prepared(ctx, 4)  # n = 96: above the static scratch
        with pytest.raises(hip.InvalidInput, match="64 MiB"):
            ctx.mp2_noalloc(C, e, o)
