"""Writes systematic_fcc32.json: exhaustive search of the 32-site FCC A16B16 cell, 2 shells
(weights 1, 1/2), default pair weights, all C(32,16) permutations, solved by the oracle
(oracle/sqsgen_oracle.c driver, multithreaded).  Records the optimal objective, the number
of optimal configurations with a SHA-256 digest of their sorted reported ranks (little-endian
uint64), and the next objective level.

Run from the repository root: python tests/golden/make_systematic_golden.py
"""
import hashlib
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402

A = 4.05
BASIS = [[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]]


def problem():
    lat, frac, z = O.supercell(np.diag([A, A, A]), BASIS, [13] * 4, 2, 2, 2)
    # composition {Al: 16, Ti: 16}: ascending Z over ascending sites (structure.h:390-399)
    z = np.array([13] * 16 + [22] * 16, dtype=np.uint8)
    dm = O.distance_matrix(lat, frac)
    radii = O.distances_histogram(dm, 0.05, 0.25)
    shells = [1, 2]
    sm = O.shell_matrix(dm, radii)
    pairs = O.pairs(sm, shells)
    pref = O.compute_prefactors(sm, shells, z)
    pw = O.scaled_pair_weights(np.repeat((np.ones((2, 2)) - np.eye(2))[None], 2, axis=0),
                               [1.0, 0.5])
    packed = (z == 22).astype(np.uint8)
    return O.Problem(pairs, packed, [0, 32], pref, pw, np.zeros((2, 2, 2))), radii


def main():
    prob, radii = problem()
    total = O.num_permutations(prob.species)
    t0 = time.time()
    st, res = O.run([prob], total, chunk_size=1 << 20, systematic=True, keep=1,
                    threads=min(8, os.cpu_count() or 1))
    dt = time.time() - t0
    levels = sorted({r[0] for r in res})
    best = levels[0]
    ranks = sorted(O.rank_permutation(r[1]) for r in res if r[0] == best)
    second = [r for r in res if r[0] == levels[1]] if len(levels) > 1 else []
    r = np.array(ranks, dtype=np.uint64)
    out = dict(description="exhaustive search of the 32-site FCC A16B16 cell (a = 4.05), 2 shells "
                           "(weights 1, 1/2), default pair weights, solved by the oracle",
               total=total, radii=list(radii), best_objective=best,
               best_objective_hex=float(best).hex(), num_best=len(ranks),
               best_ranks_sha256=hashlib.sha256(r.astype("<u8").tobytes()).hexdigest(),
               best_ranks_head=r[:64].tolist(), best_ranks_tail=r[-64:].tolist(),
               second_objective=levels[1] if len(levels) > 1 else None,
               second_objective_hex=float(levels[1]).hex() if len(levels) > 1 else None,
               num_second=len(second), oracle_seconds=dt)
    path = Path(__file__).with_name("systematic_fcc32.json")
    path.write_text(json.dumps(out, indent=1))
    print(f"{total} permutations in {dt:.1f}s; best {best} x{len(ranks)}; "
          f"second {out['second_objective']} x{len(second)}")


if __name__ == "__main__":
    main()
