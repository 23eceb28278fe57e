#!/usr/bin/env python3
"""Oracle fixtures for the batched 10-300-atom PM6 parity test (tests/golden/mix_pm6_oracle.json).

The FP64 CPU oracle (oracle/, a restatement of the reference algorithm) solves each molecule of a
fixed sample of the seeded 10-300-atom mix tightly (|dE| < 1e-11 Eh, RMSD(dP) < 1e-10, DIIS 8,
standard guess) and the fixture keeps, per molecule, its index in the mix, the total energy, the
iteration count and the atom populations q_A = sum over A's orbitals of P_mumu.  These solves take
minutes on a CPU, which is why they are precomputed; the GPU test recomputes the molecules from
the same generator and seed.  Run by hand: python tests/golden/make_oracle_fixtures.py
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
from sparrow_amd import generators as gen  # noqa: E402
from sparrow_amd.structure import BOHR_PER_ANGSTROM  # noqa: E402

OUT = os.path.join(HERE, "mix_pm6_oracle.json")
MIX = dict(count=256, lo=10, hi=300, seed=1)
# eight molecules of the mix whose DIIS path is sensitive to projection noise in its first
# iterations (they reached other SCF solutions when low precision started from the guess), and
# the first eight others
SAMPLE = [7, 51, 52, 66, 82, 154, 169, 201, 0, 1, 2, 3, 4, 5, 6, 8]


def populations(Z, P, orc):
    q, o = [], 0
    for z in Z:
        k = orc.nao([z])
        q.append(float(np.trace(P[o:o + k, o:o + k])))
        o += k
    return q


def main():
    O.set_threads(os.cpu_count() or 1)
    orc = O.Oracle("PM6")
    mols = gen.size_mix(MIX["count"], MIX["lo"], MIX["hi"], seed=MIX["seed"])
    out = {"method": "PM6", "mix": MIX, "convergence": {"dE": 1e-11, "rmsd": 1e-10, "maxit": 800, "mixer": "diis",
                                                          "diis_size": 8},
           "source": "oracle/nddo_oracle.cpp orc_scf (FP64 CPU restatement of the reference)", "molecules": []}
    for i in SAMPLE:
        Z, X = mols[i]
        t = time.time()
        r = orc.scf(Z, X * BOHR_PER_ANGSTROM, 0, dE=1e-11, rmsd=1e-10, maxit=800)
        out["molecules"].append({"index": i, "atoms": len(Z), "nao": int(r["P"].shape[0]), "energy": r["energy"],
                                 "iterations": r["iterations"], "converged": bool(r["converged"]),
                                 "populations": populations(Z, r["P"], orc)})
        print(f"{i}: {len(Z)} atoms, E = {r['energy']:.10f}, {r['iterations']} iterations, "
              f"{time.time() - t:.1f} s", flush=True)
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
