"""CPU tests of the host side: the C ABI libraries load and export what include/*.h declares, the host
FP64 eigensolver used by the SCF driver, the J/K block-quartet algorithm (numpy emulation of the
kernel's arithmetic against the oracle's getFock), and the synthetic-molecule generator."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, system_ints
from oracle import oracle as O


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//.*", "", text)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", text)
    return sorted(set(n for n in names if not n.startswith("__")))


@pytest.mark.parametrize("header,lib", [("zinq_hip.h", "libzinq_hip.so"), ("zinq_scf.h", "libzinq_hip.so"),
                                        ("zinq_batch.h", "libzinq_hip.so"), ("zinq_ints.h", "libzinq_ints.so")])
def test_library_exports_every_declared_symbol(header, lib):
    names = _declared(header)
    assert len(names) >= 2
    L = ctypes.CDLL(os.path.join(ROOT, "zinq_amd", lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_sygv_matches_numpy():
    from zinq_amd import hip
    for kind in ("sto3g", "ccpvdz"):
        s, H, S, g, VN = system_ints(kind)
        w, C = hip.host_sygv(H, S)
        np.testing.assert_allclose(H @ C, S @ C * w, atol=1e-11)
        np.testing.assert_allclose(C.T @ S @ C, np.eye(s.nbf), atol=1e-12)
        np.testing.assert_allclose(w, O.geigh(H, S)[0], atol=1e-11)
    # clustered / degenerate spectrum
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    d = np.repeat(np.arange(10.0), 4)
    A = Q @ np.diag(d) @ Q.T
    w, V = hip.host_sygv(A, np.eye(40))
    np.testing.assert_allclose(w, d, atol=1e-12)
    np.testing.assert_allclose(A @ V, V * w, atol=1e-12)


# ---------------------------------------------------------------- J/K algorithm emulation
B = 8


def _bpair(a, b):
    return a * (a + 1) // 2 + b


def _kets(bp, bq):
    return [(br, bs) for bs in range(bp + 1) for br in range(bs, (bp if bs <= bq else bp - 1) + 1)]


def _pack(chem, n):
    nb = (n + B - 1) // B
    g = np.zeros((nb * B,) * 4)
    g[:n, :n, :n, :n] = chem
    Nb = nb * (nb + 1) // 2
    blocks = np.zeros((Nb * (Nb + 1) // 2, 64, 64))
    r, s, p, q = np.meshgrid(*[np.arange(8)] * 4, indexing="ij")
    for bp in range(nb):
        for bq in range(bp + 1):
            Ib = _bpair(bp, bq)
            for k, (br, bs) in enumerate(_kets(bp, bq)):
                Jb = _bpair(br, bs)
                keep = ~(((bp == bq) & (p < q)) | ((br == bs) & (r < s)) | ((Ib == Jb) & (p * 8 + q < r * 8 + s)))
                vals = g[bp * 8 + p, bq * 8 + q, br * 8 + r, bs * 8 + s] * keep
                blocks[Ib * (Ib + 1) // 2 + k] = vals.reshape(64, 64)
    return blocks, nb


def _jk_emulated(blocks, nb, P, H, n, x, CH):
    Pp = np.zeros((nb * B, nb * B))
    Pp[:n, :n] = P
    G = np.zeros_like(Pp)
    lane = np.arange(64)
    lp, lq = lane // 8, lane % 8
    for bp in range(nb):
        for bq in range(bp + 1):
            Ib, kl = _bpair(bp, bq), _kets(bp, bq)
            Pg, Qg = bp * 8 + lp, bq * 8 + lq
            fI = np.where((bp == bq) & (lp == lq), 0.5, 1.0)
            for k0 in range(0, len(kl), CH):
                for k in range(k0, min(len(kl), k0 + CH)):
                    br, bs = kl[k]
                    Jb = _bpair(br, bs)
                    blk = blocks[Ib * (Ib + 1) // 2 + k]
                    for r in range(8):
                        for s in range(8):
                            w = fI * (0.5 if (br == bs and r == s) else 1.0) * np.where((Ib == Jb) & (lane == r * 8 + s), .5, 1)
                            a = blk[r * 8 + s] * w
                            R, S = br * 8 + r, bs * 8 + s
                            np.add.at(G, (Pg, Qg), 2 * Pp[R, S] * a)
                            G[R, S] += 2 * np.sum(Pp[Pg, Qg] * a)
                            np.add.at(G, (Qg, S), -x * Pp[Pg, R] * a)
                            np.add.at(G, (Pg, S), -x * Pp[Qg, R] * a)
                            np.add.at(G, (Qg, R), -x * Pp[Pg, S] * a)
                            np.add.at(G, (Pg, R), -x * Pp[Qg, S] * a)
    F = H + (G + G.T)[:n, :n]
    return F


@pytest.mark.parametrize("kind", ["sto3g", "ccpvdz"])
def test_block_quartet_jk_algorithm_matches_getfock(kind):
    s, H, S, g, VN = system_ints(kind)
    n = s.nbf
    blocks, nb = _pack(g.transpose(0, 2, 1, 3), n)
    rng = np.random.default_rng(4)
    A = rng.standard_normal((n, n))
    P = A + A.T
    for x, gen in [(0.5, False), (1.0, True)]:
        F = _jk_emulated(blocks, nb, P, H, n, x, CH=2)
        np.testing.assert_allclose(F, O.get_fock(g, H, P, gen), atol=1e-12)
    # every unique integral stored exactly once
    npair = n * (n + 1) // 2
    assert np.count_nonzero(blocks) == np.count_nonzero(s.coulomb_packed())


def test_water_cluster_generator_is_deterministic():
    from zinq_amd.molecules import water_cluster
    Z1, x1 = water_cluster(4)
    Z2, x2 = water_cluster(4)
    np.testing.assert_array_equal(x1, x2)
    assert list(Z1) == [8, 1, 1] * 4
    # rigid monomers: O-H 0.9571 A, HOH 104.5 deg
    for m in range(4):
        o, h1, h2 = x1[3 * m:3 * m + 3]
        assert np.linalg.norm(h1 - o) == pytest.approx(np.hypot(0.757, 0.587), abs=1e-12)
        cosang = np.dot(h1 - o, h2 - o) / np.linalg.norm(h1 - o) / np.linalg.norm(h2 - o)
        assert cosang == pytest.approx((0.587 ** 2 - 0.757 ** 2) / (0.757 ** 2 + 0.587 ** 2), abs=1e-12)
    # the committed fixtures of the benchmark clusters (tests/golden/make_clusters.py) match the generator
    for k in (4, 6, 8, 10):
        Zk, xk = water_cluster(k)
        lines = open(os.path.join(ROOT, "tests", "golden", f"water_k{k}.xyz")).read().splitlines()
        assert int(lines[0]) == 3 * k and [ln.split()[0] for ln in lines[2:]] == ["O", "H", "H"] * k
        ref = np.array([[float(v) for v in ln.split()[1:]] for ln in lines[2:]])
        np.testing.assert_allclose(xk, ref, atol=1e-9)


def test_batch_specs_cover_the_workload():
    """512 seeded molecules, k in 1..3, STO-3G and cc-pVDZ: sizes n in {7, 14, 21, 24, 48, 72}."""
    from zinq_amd.molecules import batch_specs, batch_system
    specs = batch_specs()
    assert len(specs) == 512 and specs == batch_specs()
    assert {s["k"] for s in specs} == {1, 2, 3}
    sizes = {(s["k"], s["basis"]): None for s in specs}
    assert len(sizes) == 6
    n = sorted({batch_system(s).nbf for s in specs[:40]})
    assert set(n) <= {7, 14, 21, 24, 48, 72}


def test_matrix_text_format_round_trip(tmp_path):
    """The reference's text format (read_write.zig:24-80): header, {d:20.14} elements, n^2 x n^2 ERIs."""
    from zinq_amd import matio
    from conftest import system_ints
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3, 5)) * 10.0 ** rng.integers(-3, 3, (3, 5))
    p = tmp_path / "a.mat"
    matio.write_matrix(p, A)
    lines = p.read_text().splitlines()
    assert lines[0] == "3 5" and len(lines) == 4
    assert len(lines[1]) == 5 * 20 + 4
    np.testing.assert_allclose(matio.read_matrix(p), A, rtol=0, atol=5e-15)
    s, H, S, g, VN = system_ints("sto3g")
    q = tmp_path / "coulomb.mat"
    matio.write_tensor(q, g)
    assert q.read_text().splitlines()[0] == f"{s.nbf ** 2} {s.nbf ** 2}"
    np.testing.assert_allclose(matio.read_coulomb(q), g, rtol=0, atol=5e-15)
    # readMatrix ignores where the line breaks fall
    r = tmp_path / "b.mat"
    r.write_text("2 2\n1.5 -2.25\n3.0\n4.0\n")
    np.testing.assert_array_equal(matio.read_matrix(r), [[1.5, -2.25], [3.0, 4.0]])


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; a request for several exits with a message and a non-zero status,
    before touching the GPU."""
    import subprocess
    import sys
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr and p.stdout == ""
