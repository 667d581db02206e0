"""Benchmark of the MI355X hot path on the water-cluster workload.

Workload (default): RHF + MP2 of (H2O)_k, k = 4, 6, 8, cc-pVDZ (n = 96, 144, 192), seeded geometries
(zinq_amd.molecules), dense FP64 ERIs generated on the host by the integral engine (untimed, reported
separately) and resident on the GPU 8-fold packed.  One step = for every molecule: the DIIS SCF to
1e-10 (GPU J/K each iteration, FP64 host DIIS/diagonalisation) followed by the MP2 energy (GPU AO->MO
transform + FP64 sum).  value = molecules / second over the timed steps.

Per-kernel durations come from HIP events recorded on the stream the kernels run on (zq_profile_*),
in one extra profiled step after the timed region; the roofline object is computed from them for the
kernel with the largest total time, and a second one for the transform's Q1 stage.  cpu_baseline
times the CPU oracle (test infrastructure) end to end on the step's smallest molecule; oracle_parity
compares the step's FP64 energies with the oracle's.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per zq_jk_records launch measured with FETCH_SIZE / WRITE_SIZE counters, scaled by a 1 GiB
# copy (profiles/zq_jk_records_water8_hbm_traffic.md), by basis size n
JK_TRAFFIC = {192: 1561121731}
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0 / 3}
METRIC = "MO-transform TFLOP/s + molecules/s, mixed vs FP64 at 1µEh; p50/p99 latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def jk_bytes(n):
    p = n * (n + 1) // 2
    return 8 * (p * (p + 1) // 2) + 24 * n * n


def ovov_flops(n, o):
    v = n - o
    return 2 * n * (o * n ** 3 + o * o * n * n + o * o * v * n + o * o * v * v)


def prepare(ks, basis):
    from zinq_amd import hip
    from zinq_amd.molecules import water_system
    mols = []
    for k in ks:
        t = time.time()
        s = water_system(k, basis=basis)
        pk = s.coulomb_packed()
        t_int = time.time() - t
        ctx = hip.Context(0)
        ctx.eri_load_packed(pk, s.nbf)
        mols.append(dict(k=k, n=s.nbf, o=s.nel // 2, H=s.hcore(), S=s.overlap(), VN=s.nrep(), ctx=ctx, t_int=t_int,
                         sys=s, packed=pk))
        log(f"(H2O)_{k}: n={s.nbf} integrals {t_int:.1f}s, device {ctx.device_bytes() / 1e9:.2f} GB")
    return mols


# The plain FP64 form (the reference's arithmetic, SCF to 1e-10) and the mixed path (incremental Fock
# builds from FP32 ERIs to 1e-6, then one full FP64 build; FP32 products with FP64 block sums in the
# transform).  MP2 denominators and sums are FP64 in both.
PLAIN = dict(fock="fp64", transform="fp64", threshold=1e-10)
MIXED = dict(fock="fp32", transform="fp32c", threshold=1e-6)


def run_one(m, mode, device_scf):
    r = m["ctx"].scf(m["H"], m["S"], m["VN"], m["o"], diis=8, iterations=100, threshold=mode["threshold"],
                     device=device_scf, fock_precision=mode["fock"] if device_scf else "fp64")
    driver = m["ctx"].last_scf_driver() if device_scf else "host"
    # after the device SCF its orbitals are still on the device: the MP2 takes them there
    e2 = (m["ctx"].mp2_resident(m["o"], mode["transform"]) if device_scf
          else m["ctx"].mp2(r["C"], r["e"], m["o"], mode["transform"]))
    return dict(k=m["k"], E_rhf=r["E"], E_mp2=r["E"] + e2, iterations=r["iterations"], scf_driver=driver)


def step(mols, mode, results=None, device_scf=True, concurrent=True):
    """All molecules of the step; with concurrent=True each runs in its own host thread on its own
    context/stream (the C calls release the GIL), so their GPU work overlaps."""
    if concurrent:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(len(mols)) as ex:
            out = list(ex.map(lambda m: run_one(m, mode, device_scf), mols))
    else:
        out = [run_one(m, mode, device_scf) for m in mols]
    if results is not None:
        results.extend(out)


def timed(mols, mode, args, dev, conc):
    for _ in range(args.warmup):
        step(mols, mode, device_scf=dev, concurrent=conc)
    results = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(mols, mode, results if i == 0 else None, device_scf=dev, concurrent=conc)
    return time.perf_counter() - t0, results


def cpu_baseline(mols):
    """The CPU oracle (oracle.c: the reference's loops restated; test infrastructure) timed on the host
    cores: the smallest molecule of the step, RHF (DIIS-8 to 1e-10; the Fock builds on every core,
    bitwise the faithful loops; the Jacobi eigensolver serial) + MP2 (ao2mo_oovv on every core + the
    closed-shell sum), end to end.  value = that molecule per second.  Also returns the oracle's
    energies, against which the GPU results of the same molecule are compared."""
    from oracle import oracle as O
    m = min(mols, key=lambda x: x["n"])
    s = m["sys"]
    g = s.coulomb()
    t = time.time()
    r = O.scf(m["H"], m["S"], g, m["VN"], m["o"], threshold=1e-10, par=True)
    t_scf = time.time() - t
    e2 = O.mp2_closed(O.ao2mo_oovv_par(g, r["C"], m["o"]), r["e"])
    total = time.time() - t
    del g
    threads = O.max_threads()
    base = dict(value=1.0 / total, unit="molecules/s", cores=threads, kind="port",
                sample=f"(H2O)_{m['k']} cc-pVDZ (n={m['n']}) RHF + MP2 end to end with the oracle on {threads} threads: "
                       f"{total:.2f} s (SCF {t_scf:.2f} s, {r['iterations']} iterations; MP2 {total - t_scf:.2f} s); "
                       f"value = that molecule per second (the step's other molecules take longer)")
    return base, dict(k=m["k"], E_rhf=r["E"], E_mp2=r["E"] + e2)


def oracle_parity(results, live):
    """|dE| of the step's FP64 GPU results against the CPU oracle: the molecule the cpu_baseline leg ran
    live, the others from the committed oracle energies (tests/golden/water_clusters_oracle.json, written
    by tests/golden/make_clusters.py with the same oracle)."""
    path = os.path.join(ROOT, "tests", "golden", "water_clusters_oracle.json")
    fixed = json.load(open(path)) if os.path.exists(path) else {}
    out = {}
    for r in results:
        if live is not None and r["k"] == live["k"]:
            ref, src = live, "oracle run in this bench"
        elif str(r["k"]) in fixed:
            ref, src = fixed[str(r["k"])], "committed oracle energies"
        else:
            continue
        out[f"k{r['k']}"] = dict(abs_dE_rhf=abs(r["E_rhf"] - ref["E_rhf"]), abs_dE_mp2=abs(r["E_mp2"] - ref["E_mp2"]),
                                 source=src)
    return out


def sweep(k, basis):
    """Precision sweep on one (H2O)_k: converged FP64 RHF orbitals, then the OVOV transform and the MP2
    energy in every transform precision, against the FP64 GPU form (itself matched to the oracle at
    smaller sizes).  Also the whole mixed path (mixed Fock + FP32C transform) against the plain form."""
    from zinq_amd import hip
    from zinq_amd.molecules import water_system
    s = water_system(k, basis=basis)
    n, o = s.nbf, s.nel // 2
    ctx = hip.Context(0)
    ctx.eri_load_packed(s.coulomb_packed(), n)
    H, S, VN = s.hcore(), s.overlap(), s.nrep()
    r = ctx.scf(H, S, VN, o, threshold=1e-10, device=True)
    rows, ref_T, ref_e2 = [], None, None
    for prec in ("fp64", "fp32c", "fp32", "bf16x3"):
        ctx.mp2(r["C"], r["e"], o, prec)  # warm-up (scratch allocation)
        t = time.perf_counter()
        e2 = ctx.mp2(r["C"], r["e"], o, prec)
        dt = time.perf_counter() - t
        T = ctx.ao2mo(r["C"], o, "ovov", prec)
        if prec == "fp64":
            ref_T, ref_e2 = T, e2
        rows.append(dict(precision=prec, E_corr=e2, abs_dE_mp2=abs(e2 - ref_e2),
                         max_abs_d_iajb=float(np.abs(T - ref_T).max()), mp2_ms=round(dt * 1e3, 2),
                         transform_TFLOPs=round(ovov_flops(n, o) / dt / 1e12, 2)))
        del T
        log(rows[-1])
    rm = ctx.scf(H, S, VN, o, threshold=1e-6, device=True, fock_precision="fp32")
    e2m = ctx.mp2(rm["C"], rm["e"], o, "fp32c")
    mixed = dict(E_rhf=rm["E"], dE_rhf=rm["E"] - r["E"], E_mp2=rm["E"] + e2m,
                 abs_dE_mp2=abs(rm["E"] + e2m - r["E"] - ref_e2), iterations=rm["iterations"])
    ctx.close()
    return dict(sweep=f"(H2O)_{k} {basis.split(':')[-1]}", n=n, o=o, E_rhf=r["E"], E_mp2=r["E"] + ref_e2,
                rows=rows, mixed_path=mixed,
                reference="FP64 GPU form (MP2 transform + FP64 sums); tests/test_gpu_sizes.py matches it to the CPU "
                          "oracle on (H2O)_4 whole and on sampled rows, blocks and energies at n = 144..240")


def batch(count, steps, warmup):
    """Batched workload: `count` seeded molecules (k ~ U{1,2,3} waters, STO-3G or cc-pVDZ, bond +-0.05 A
    and angle +-5 deg distortions; n in {7,14,21,24,48,72}) through the grouped RHF + MP2 (zinq_batch.h):
    grouped J/K, one workgroup per molecule for the rest of each iteration, the iteration captured as a
    hipGraph, item size and workgroup size from the autotuner; MP2 as one grouped launch per transform
    stage over all molecules.  FP64, SCF to 1e-10.  stage_ms: the batch's own stage times (SCF,
    eigenpairs, MP2, total)."""
    from zinq_amd.batch import Batch
    from zinq_amd.molecules import batch_specs, batch_system
    t = time.time()
    specs = batch_specs(count)
    systems = [batch_system(s) for s in specs]
    b = Batch(0)
    for s in systems:
        b.add_system(s)
    t_int = time.time() - t
    t = time.time()
    tuned = b.autotune()
    t_tune = time.time() - t
    for _ in range(warmup):
        b.run(threshold=1e-10)
    t0 = time.perf_counter()
    stages = []
    for _ in range(steps):
        r = b.run(threshold=1e-10)
        stages.append(b.timings())
    el = time.perf_counter() - t0
    sizes = np.array(b.sizes)
    per_n = {int(n): dict(count=int((sizes == n).sum()), iterations_mean=round(float(r["iterations"][sizes == n].mean()), 2))
             for n in sorted(set(b.sizes))}
    b.close()
    return dict(metric=METRIC, value=round(count * steps / el, 1), unit="molecules/s", n_gpus=1, steps=steps,
                warmup=warmup, ms_per_step=round(1e3 * el / steps, 2), higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype="f64", data="synthetic",
                config=dict(workload=f"batched RHF+MP2, {count} molecules, (H2O)_k k~U{{1,2,3}}, STO-3G/cc-pVDZ, "
                                     "distorted; grouped J/K + per-molecule workgroup SCF, graph-captured iteration, "
                                     "grouped MP2 transform",
                            autotuned=tuned, autotune_s=round(t_tune, 2)),
                status_ok=int((r["status"] == 0).sum()), per_n=per_n, integrals_s=round(t_int, 2),
                stage_ms={k: round(float(np.mean([t[k] for t in stages])), 2) for k in stages[0]},
                E_rhf_sample=[float(x) for x in r["E_rhf"][:4]], E_corr_sample=[float(x) for x in r["E_corr"][:4]])


def serve(lams, slots, kmax):
    """Serving workload: Poisson request streams (seeded exponential gaps, rate lam per second) of RHF+MP2
    on (H2O)_k, k ~ U{1..kmax}, cc-pVDZ; the mixed path (incremental FP32-ERI Fock builds to 1e-6, one
    FP64 build, FP32C transform); ERI sets resident and admitted by footprint against the HBM budget;
    `slots` requests in flight on their own streams.  Latency = completion - arrival.  |dE_MP2| of every
    served result against the plain FP64 device path of the same molecule."""
    from zinq_amd import hip
    from zinq_amd.molecules import water_system
    from zinq_amd.serve import Molecule, Server, poisson_arrivals, summarize
    t = time.time()
    mols = [Molecule(f"w{k}", water_system(k)) for k in range(1, kmax + 1)]
    t_int = time.time() - t
    ref = {}
    with hip.Context(0) as ctx:
        for m in mols:
            ctx.eri_load_packed(m.packed, m.n)
            r = ctx.scf(m.H, m.S, m.VN, m.o, threshold=1e-10, device=True)
            ref[m.key] = r["E"] + ctx.mp2(r["C"], r["e"], m.o)
    keys = [m.key for m in mols]
    srv = Server(mols, slots=slots)
    srv.serve([(0.0, k) for k in keys])  # warm-up: every set resident, every slot's scratch grown
    rows, worst = [], 0.0
    for i, lam in enumerate(lams):
        count = int(min(64, max(24, 2 * lam)))
        rec, wall = srv.serve(poisson_arrivals(lam, count, keys, seed=100 + i))
        row = summarize(rec, wall)
        row["lambda"] = lam
        row["max_abs_dE_mp2_vs_fp64"] = max(abs(r["E_mp2"] - ref[r["key"]]) for r in rec if "error" not in r)
        worst = max(worst, row["max_abs_dE_mp2_vs_fp64"])
        rows.append(row)
        log(row)
    resident_gb = srv.reserved / 1e9
    srv.close()
    return dict(metric=METRIC, serve=rows, slots=slots, unit="molecules/s", workload=f"Poisson RHF+MP2 requests, "
                f"(H2O)_k k~U{{1..{kmax}}} cc-pVDZ (n=24..{24 * kmax}), mixed path, continuous batching",
                max_abs_dE_mp2_vs_fp64=worst, resident_sets_gb=round(resident_gb, 2), integrals_s=round(t_int, 1))


def kernel_only(which, k, reps, basis, precision="fp64"):
    """Single-kernel runs for hardware-counter passes: `jk` repeats the FP64 Fock build of (H2O)_k on a
    fixed density; `mp2` repeats the MP2 energy (OVOV transform + sum) of (H2O)_k from converged orbitals
    in the given transform precision; `copy` copies a device buffer of 1 GiB (reads 1 GiB, writes 1 GiB), the calibration
    for FETCH_SIZE / WRITE_SIZE."""
    if which == "copy":
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(reps):
            y.copy_(x)
        torch.cuda.synchronize()
        return dict(kernel="copy", bytes_read=x.numel() * 8, bytes_written=y.numel() * 8, reps=reps)
    from zinq_amd import hip
    from zinq_amd.molecules import water_system
    s = water_system(k, basis=basis)
    ctx = hip.Context(0)
    ctx.eri_load_packed(s.coulomb_packed(), s.nbf)
    H = s.hcore()
    if which == "mp2":
        o = s.nel // 2
        r = ctx.scf(H, s.overlap(), s.nrep(), o, threshold=1e-10, device=True)
        for _ in range(reps):
            e2 = ctx.mp2(r["C"], r["e"], o, precision)
        ctx.close()
        return dict(kernel="zq_q1/zq_q2/zq_gemm", n=s.nbf, o=o, precision=precision,
                    algorithmic_flops=ovov_flops(s.nbf, o), reps=reps, E_corr=e2)
    P = np.eye(s.nbf) * (s.nel / s.nbf)
    for _ in range(reps):
        ctx.fock_jk(H, P)
    ctx.close()
    return dict(kernel="zq_jk_records", n=s.nbf, algorithmic_bytes=jk_bytes(s.nbf), reps=reps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ks", default="4,6,8")
    ap.add_argument("--basis", default="builtin:cc-pvdz")
    ap.add_argument("--precision", default="fp64", help="transform precision of the plain line")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-path pass")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--scf", default="device", choices=["device", "host"])
    ap.add_argument("--serial", action="store_true", help="run the molecules of a step one after another")
    ap.add_argument("--sweep", type=int, default=0, metavar="K", help="precision sweep on (H2O)_K instead")
    ap.add_argument("--batch", type=int, default=0, metavar="N", help="batched workload of N molecules instead")
    ap.add_argument("--serve", default="", metavar="L1,L2,..", help="serving workload at these arrival rates instead")
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--kmax", type=int, default=10)
    ap.add_argument("--kernel", default="", choices=["", "jk", "mp2", "copy"], help="single-kernel run for counter passes")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the short batched (512 molecules) and serving (k = 1..10, 100 req/s) runs of the default line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: only the one-GPU path is built; a multi-GPU run is not supported.", file=sys.stderr)
        sys.exit(2)
    if args.sweep:
        print(json.dumps(sweep(args.sweep, args.basis)), flush=True)
        return
    if args.batch:
        print(json.dumps(batch(args.batch, args.steps, args.warmup)), flush=True)
        return
    if args.kernel:
        print(json.dumps(kernel_only(args.kernel, int(args.ks.split(",")[-1]), args.steps, args.basis, args.precision)),
              flush=True)
        return
    if args.serve:
        print(json.dumps(serve([float(x) for x in args.serve.split(",")], args.slots, args.kmax)), flush=True)
        return
    ks = [int(k) for k in args.ks.split(",")]
    mols = prepare(ks, args.basis)

    dev = args.scf == "device"
    conc = not args.serial
    plain = dict(PLAIN, transform=args.precision)
    elapsed, results = timed(mols, plain, args, dev, conc)
    ms_per_step = 1e3 * elapsed / args.steps
    mixed = None
    if not args.no_mixed:
        el_m, res_m = timed(mols, MIXED, args, dev, conc)
        de = [abs(a["E_mp2"] - b["E_mp2"]) for a, b in zip(res_m, results)]
        mixed = dict(value=round(len(mols) * args.steps / el_m, 3), unit="molecules/s",
                     speedup_vs_fp64=round(elapsed / el_m, 3), ms_per_step=round(1e3 * el_m / args.steps, 2),
                     max_abs_dE_mp2_vs_fp64=max(de), bar=1e-6, within_bar=bool(max(de) <= 1e-6),
                     config=dict(MIXED, note="Fock: incremental builds from FP32 ERIs (FP64 products/sums) to "
                                             "the threshold, then one full FP64 build; transform: FP32 MFMA "
                                             "products, FP64 sums per K block"),
                     results=res_m)
    if mixed is not None:
        # the same mixed arithmetic converged to the plain form's threshold: precision and threshold apart
        mixed_tight = dict(MIXED, threshold=PLAIN["threshold"])
        el_t, res_t = timed(mols, mixed_tight, args, dev, conc)
        de_t = [abs(a["E_mp2"] - b["E_mp2"]) for a, b in zip(res_t, results)]
        mixed["at_plain_threshold"] = dict(threshold=PLAIN["threshold"], value=round(len(mols) * args.steps / el_t, 3),
                                           speedup_vs_fp64=round(elapsed / el_t, 3),
                                           ms_per_step=round(1e3 * el_t / args.steps, 2),
                                           max_abs_dE_mp2_vs_fp64=max(de_t),
                                           iterations=[x["iterations"] for x in res_t])
        # one profiled mixed pass on the largest molecule: the FP32-ERI incremental Fock build
        big_m = max(mols, key=lambda x: x["n"])
        c = big_m["ctx"]
        c.profile(True)
        c.profile_reset()
        run_one(big_m, MIXED, dev)
        ms32, cnt32 = c.profile_get("jk_records32")
        c.profile(False)
        n_b = big_m["n"]
        p_b = n_b * (n_b + 1) // 2
        b32 = 4 * (p_b * (p_b + 1) // 2) + 24 * n_b * n_b
        if cnt32:
            mixed["jk_fp32"] = dict(n=n_b, ms_avg=round(ms32 / cnt32, 4), launches=cnt32,
                                    GBs=round(b32 / (ms32 / cnt32 * 1e-3) / 1e9, 1),
                                    per_launch=f"4*N_u + 24 n^2 = {b32} bytes")
    for r in results:
        for m in mols:
            if m["k"] == r["k"]:
                m["iters"] = r["iterations"]

    # profiled step: kernel durations on the context streams
    kern = {}
    for m in mols:
        m["ctx"].profile(True)
        m["ctx"].profile_reset()
    wall = {}
    for m in mols:  # one molecule at a time, so that the wall times are the molecule's own
        t = time.perf_counter()
        r = m["ctx"].scf(m["H"], m["S"], m["VN"], m["o"], diis=8, iterations=100, threshold=plain["threshold"], device=dev)
        t_scf = time.perf_counter() - t
        m["ctx"].mp2(r["C"], r["e"], m["o"], args.precision)
        wall[m["k"]] = (t_scf * 1e3, (time.perf_counter() - t - t_scf) * 1e3)
    per_mol = []
    for m in mols:
        c = m["ctx"]
        d = {name: c.profile_get(name) for name in ("jk_records", "jk_reduce", "ao2mo_q1", "ao2mo_q2", "ao2mo_q3",
                                                     "ao2mo_q4", "mp2_energy", "scf_eig", "scf_gemm", "scf_sp2")}
        c.profile(False)
        t_tr = sum(d[q][0] for q in ("ao2mo_q1", "ao2mo_q2", "ao2mo_q3", "ao2mo_q4")) * 1e-3
        jk_ms, jk_cnt = d["jk_records"]
        per_mol.append(dict(k=m["k"], n=m["n"], o=m["o"], scf_iterations=m.get("iters"),
                            jk_ms_avg=jk_ms / max(jk_cnt, 1), jk_reduce_ms_avg=d["jk_reduce"][0] / max(d["jk_reduce"][1], 1),
                            jk_GBs=jk_bytes(m["n"]) / (jk_ms / max(jk_cnt, 1) * 1e-3) / 1e9,
                            transform_ms=t_tr * 1e3, transform_TFLOPs=ovov_flops(m["n"], m["o"]) / t_tr / 1e12,
                            q_ms={q: d[q][0] for q in ("ao2mo_q1", "ao2mo_q2", "ao2mo_q3", "ao2mo_q4")},
                            q_launches={q: d[q][1] for q in ("ao2mo_q1", "ao2mo_q2", "ao2mo_q3", "ao2mo_q4")},
                            eig_ms=d["scf_eig"][0], scf_gemm_ms=d["scf_gemm"][0], sp2_ms=d["scf_sp2"][0],
                            sp2_calls=d["scf_sp2"][1], scf_wall_ms=wall[m["k"]][0], mp2_wall_ms=wall[m["k"]][1]))
        for name, (ms, cnt) in d.items():
            kern.setdefault(name, [0.0, 0, 0.0])
            kern[name][0] += ms
            kern[name][1] += cnt

    # roofline of the dominant kernel, for the largest molecule
    big = max(per_mol, key=lambda x: x["n"])
    dom = max(("jk_records", "ao2mo_q1"), key=lambda k: kern[k][0])  # hot-path kernels only
    if dom.startswith("jk"):
        achieved = big["jk_GBs"]
        roof = dict(kernel="zq_jk_records", bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=JK_TRAFFIC.get(big["n"]),
                    per_launch=f"8*N_u + 24 n^2 = {jk_bytes(big['n'])} bytes at n={big['n']}")
    else:
        q1_ms = big["q_ms"]["ao2mo_q1"]
        n, o = big["n"], big["o"]
        fl = 2 * o * n ** 4
        achieved = fl / (q1_ms * 1e-3) / 1e12
        peak = PEAK_TFLOPS[args.precision]
        roof = dict(kernel="zq_q1", bound="mfma", achieved=round(achieved, 2), peak=peak, unit="TFLOP/s",
                    frac=round(achieved / peak, 4), traffic=None, per_launch=f"2 o n^4 = {fl} flops at n={n}, o={o}")

    # the transform's dominant stage, Q1 (matrix-core bound), for the largest molecule: algorithmic
    # dense-chain flops 2 o n^4 per launch over its average launch time (ket-pair symmetry halves the
    # flops it executes; its executed MFMA utilisation is in profiles/)
    q1_ms, q1_cnt = big["q_ms"]["ao2mo_q1"], max(1, big["q_launches"]["ao2mo_q1"])
    n_b, o_b = big["n"], big["o"]
    fl_q1 = 2 * o_b * n_b ** 4
    ach_q1 = fl_q1 / (q1_ms / q1_cnt * 1e-3) / 1e12
    peak_t = PEAK_TFLOPS[args.precision]
    roof_t = dict(kernel="zq_q1", bound="mfma", achieved=round(ach_q1, 2), peak=peak_t, unit="TFLOP/s",
                  frac=round(ach_q1 / peak_t, 4), traffic=None,
                  per_launch=f"2 o n^4 = {fl_q1} flops at n={n_b}, o={o_b}",
                  transform_TFLOPs=round(big["transform_TFLOPs"], 2),
                  transform_per_launch=f"2 n (o n^3 + o^2 n^2 + o^2 v n + o^2 v^2) = {ovov_flops(n_b, o_b)} flops")

    value = len(mols) * args.steps / elapsed
    out = dict(metric=METRIC, value=round(value, 3), unit="molecules/s", n_gpus=1, steps=args.steps, warmup=args.warmup,
               ms_per_step=round(ms_per_step, 2), higher_is_better=True, scaling="weak", vs_baseline=None,
               dtype="f64" if args.precision == "fp64" else args.precision, data="synthetic",
               config=dict(workload=f"RHF+MP2 (H2O)_k k={ks} {args.basis.split(':')[-1]}, DIIS SCF to 1e-10, "
                                    f"ERIs resident 8-fold packed", precision=args.precision, scf=args.scf,
                           molecules_concurrent=conc),
               transform_TFLOPs={f"k{p['k']}": round(p["transform_TFLOPs"], 2) for p in per_mol},
               jk_GBs={f"k{p['k']}": round(p["jk_GBs"], 1) for p in per_mol},
               roofline=roof, roofline_transform=roof_t, results=results, mixed=mixed, per_molecule=per_mol,
               integrals_s={f"k{m['k']}": round(m["t_int"], 2) for m in mols})
    live = None
    if not args.no_cpu_baseline:
        out["cpu_baseline"], live = cpu_baseline(mols)
    out["oracle_parity"] = oracle_parity(results, live)
    if not args.no_extras:
        # the batched and serving workloads, shortened, so that every default run reports them too
        b = batch(512, 2, 1)
        out["batch"] = {k: b[k] for k in ("value", "unit", "steps", "ms_per_step", "status_ok", "config", "integrals_s",
                                          "stage_ms")}
        sv = serve([100.0], 8, 10)
        out["serve"] = {k: sv[k] for k in ("serve", "slots", "unit", "workload", "max_abs_dE_mp2_vs_fp64", "resident_sets_gb")}
    print(json.dumps(out), flush=True)
    for m in mols:
        m["ctx"].close()


if __name__ == "__main__":
    main()
