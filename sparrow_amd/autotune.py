"""Autotuner for the SP2 tile tables (the tile edge a molecule's density steps use, by size).

For each form (one molecule alone: latency; a batch: throughput) and each probe size it times
three SCF iterations of synthetic PM6 molecules of that size with every tile edge (32: k split
over 16 waves; 64, 128: whole k per wave) and keeps the fastest by SP2 time per SP2 flop.  The
winners become the tables {nmax: tile} that sparrow_amd._native.Context loads at start-up from
sp2_tiles.json next to this file (the C ABI default tables apply without it).

Usage (on the GPU):  python -m sparrow_amd.autotune [--out PATH]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "sp2_tiles.json")
TILES = (32, 64, 128)
ATOMS = (60, 120, 200, 300, 450, 700)  # about 150 to 1750 basis functions


def _measure(ctx, mols, st, tile, batched):
    table = [(1 << 30, tile)]
    ctx.set_sp2_tiles(table, batched=batched)
    run = (lambda: ctx.scf_batch("PM6", mols, settings=st, want_density=False)) if batched else \
        (lambda: ctx.scf("PM6", mols[0][0], mols[0][1], 0, settings=st, want_eps=False))
    run()  # warm-up
    ctx.set_profiling(True)
    ctx.reset_profile()
    t0 = time.perf_counter()
    run()
    wall = time.perf_counter() - t0
    ctx.set_profiling(False)
    ms, _ = ctx.kernel_time("gemm_sp2")
    flops = ctx.solve_counters()["sp2_flops"]
    return ms / max(flops, 1), wall


def _table(winners):
    """[(nao, tile)] ascending -> [(nmax, tile)]: each size's winner up to the midpoint to the next."""
    out = []
    for k, (n, t) in enumerate(winners):
        nmax = (n + winners[k + 1][0]) // 2 if k + 1 < len(winners) else 1 << 30
        if out and out[-1][1] == t:
            out[-1] = (nmax, t)
        else:
            out.append((nmax, t))
    return out


def tune(ctx, log=print):
    from sparrow_amd import _native
    from sparrow_amd import generators as gen
    from sparrow_amd.structure import BOHR_PER_ANGSTROM
    result = {}
    for batched, precision in ((False, "fp64"), (True, "mixed")):
        st = _native.settings_struct(1e-8, 1e-5, 3, "diis", 8, precision=precision)
        winners = []
        for atoms in ATOMS:
            Z, X = gen.random_molecule(atoms, seed=atoms)
            mol = (Z, X * BOHR_PER_ANGSTROM, 0)
            n = ctx.nao("PM6", Z)
            count = max(4, int(64 * (500.0 / n) ** 2)) if batched else 1
            times = {t: _measure(ctx, [mol] * count, st, t, batched) for t in TILES}
            best = min(times, key=lambda t: times[t][0])
            log(f"[autotune] {'batched' if batched else 'single'} n={n} x{count}: " +
                ", ".join(f"{t}: {times[t][0] * 1e9:.3f} ms/TFLOP" for t in TILES) + f" -> {best}")
            winners.append((n, best))
        result["batched" if batched else "single"] = _table(winners)
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    args = ap.parse_args()
    from sparrow_amd import _native
    ctx = _native.Context(0, apply_tuning=False)
    res = tune(ctx, log=lambda m: print(m, file=sys.stderr, flush=True))
    ctx.close()
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
