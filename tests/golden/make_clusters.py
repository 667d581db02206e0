"""Writes the seeded water-cluster fixtures: water_k{4..10}.xyz (the geometries of the (H2O)_k cc-pVDZ
workloads, from zinq_amd.molecules.water_cluster) and water_clusters_oracle.json, the CPU oracle's RHF
(DIIS-8, threshold 1e-10) and closed-shell MP2 energies of each.  The oracle reads the canonical packed
ERIs (orc_scf_packed / orc_ovov_pairs_packed, bitwise the dense loops), so n = 240 fits in memory.

    python tests/golden/make_clusters.py [k ...]      (default 4..10; minutes of CPU time for k = 10)
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from zinq_amd.molecules import water_cluster, to_xyz, water_system  # noqa: E402

OUT = os.path.join(HERE, "water_clusters_oracle.json")


def main(ks):
    res = json.load(open(OUT)) if os.path.exists(OUT) else {}
    for k in ks:
        Z, xyz = water_cluster(k)
        with open(os.path.join(HERE, f"water_k{k}.xyz"), "w") as f:
            f.write(to_xyz(Z, xyz, f"(H2O)_{k} seed {1000 + k}, Angstrom"))
        t0 = time.time()
        s = water_system(k)
        n, o = s.nbf, s.nel // 2
        pk = s.coulomb_packed()
        H, S, VN = s.hcore(), s.overlap(), s.nrep()
        r = O.scf(H, S, None, VN, o, threshold=1e-10, packed=pk)
        t1 = time.time()
        print(f"k={k} n={n}: E_RHF={r['E']:.12f} it={r['iterations']} ({t1 - t0:.0f} s)", flush=True)
        e2 = 0.0
        for i in range(o):
            T = O.ovov_pairs_packed(pk, r["C"], o, [(i, j) for j in range(o)])
            e = r["e"]
            d = e[i] + e[:o, None, None] - e[None, o:, None] - e[None, None, o:]
            e2 += float(np.sum(T * (2 * T - T.transpose(0, 2, 1)) / d))
        print(f"k={k}: E_MP2={r['E'] + e2:.12f} ({time.time() - t1:.0f} s)", flush=True)
        res[str(k)] = dict(n=n, nocc=o, E_rhf=r["E"], E_corr=e2, E_mp2=r["E"] + e2, iterations=r["iterations"],
                           threshold=1e-10, basis="cc-pvdz")
        json.dump(res, open(OUT, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main([int(a) for a in sys.argv[1:]] or list(range(4, 11)))
