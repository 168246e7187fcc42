"""Writes batched_slabs_analytic.npz: deterministic T and R for the 4096
slabs of the batched sweep (decks.batched_slabs(4096)), from the oracle's
slab solver.  Run: python tests/golden/make_batched_analytic.py"""
import importlib
import os
import sys
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def solve(args):
    import slab_analytic as sa
    L, c = args
    return sa.slab_leakage(L, c)


def main():
    decks = importlib.import_module("mc-old_amd.decks")
    L, c, N = decks.batched_slabs(4096)
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        TR = list(ex.map(solve, zip(L.tolist(), c.tolist()), chunksize=16))
    T = np.array([t for t, _ in TR])
    R = np.array([r for _, r in TR])
    np.savez(os.path.join(HERE, "batched_slabs_analytic.npz"), L=L, c=c, N=N, T=T, R=R)
    print("wrote", len(T))


if __name__ == "__main__":
    main()
