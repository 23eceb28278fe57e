"""Autotuner for the SP2 tile tables (the tile edge a molecule's density steps use, by size).

One molecule alone (latency): for each probe size it times three FP64 SCF iterations with every
tile edge (32: k split over 16 waves; 64, 128: whole k per wave) and keeps the fastest by SP2
time per SP2 flop; the per-size winners become the table.  Batches (throughput): a batch spends
its last iterations on the few molecules that converge slowly, where large tiles leave most of
the GPU idle, so candidate tables are timed on whole solves of a seeded 64-molecule 10-300-atom
mix and of 16 molecules of 350-500 atoms (above 900 basis functions, the upper end of serving
requests), each in FP64 and mixed precision (iteration cap 200; not the benchmark's seeds), and
the fastest is kept.  The tables {nmax: tile} go to sp2_tiles.json next to this file, which
sparrow_amd._native.Context loads at start-up (the C ABI default tables apply without it).

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
ATOMS = (60, 120, 200, 300, 450, 700, 2000)  # about 150 to 4900 basis functions
BATCH_TABLES = {  # candidate batched tables: {nmax: tile}
    "32": [(1 << 30, 32)],
    "64": [(1 << 30, 64)],
    "32|64@256": [(255, 32), (1 << 30, 64)],
    "64|128@936": [(936, 64), (1 << 30, 128)],
    "32|128@256": [(255, 32), (1 << 30, 128)],
    "32|64@256|128@512": [(255, 32), (511, 64), (1 << 30, 128)],
}


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
    st = _native.settings_struct(1e-8, 1e-5, 3, "diis", 8, precision="fp64")
    winners = []
    for atoms in ATOMS:
        Z, X = gen.random_molecule(atoms, seed=atoms)
        mol = (Z, X * BOHR_PER_ANGSTROM, 0)
        n = ctx.nao("PM6", Z)
        times = {t: _measure(ctx, [mol], st, t, False)[0] for t in TILES}
        best = min(times, key=lambda t: times[t])
        log(f"[autotune] single n={n}: " + ", ".join(f"{t}: {times[t] * 1e9:.3f} ms/TFLOP" for t in TILES) +
            f" -> {best}")
        winners.append((n, best))
    result["single"] = _table(winners)
    mix = [(z, x * BOHR_PER_ANGSTROM, 0) for z, x in gen.size_mix(64, 10, 300, seed=12345)]
    big = []
    for k, atoms in enumerate(range(350, 500, 10)):
        Z, X = gen.random_molecule(atoms, seed=20000 + k)
        big.append((Z, X * BOHR_PER_ANGSTROM, 0))
    walls = {}
    for name, table in BATCH_TABLES.items():
        ctx.set_sp2_tiles(table, batched=True)
        walls[name] = 0.0
        for batch in (mix, big):
            for precision in ("fp64", "mixed"):
                stb = _native.settings_struct(1e-8, 1e-5, 200, "diis", 8, precision=precision)
                ctx.scf_batch("PM6", batch[:2], settings=stb, want_density=False)  # warm-up
                t0 = time.perf_counter()
                ctx.scf_batch("PM6", batch, settings=stb, want_density=False)
                walls[name] += time.perf_counter() - t0
        log(f"[autotune] batched table {name}: {walls[name]:.3f} s (FP64 + mixed solves of both batches)")
    best = min(walls, key=lambda k: walls[k])
    log(f"[autotune] batched -> {best}")
    result["batched"] = BATCH_TABLES[best]
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
