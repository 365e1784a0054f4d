"""Synthetic workloads of the benchmark and parity suites (dict configs for parse_config)."""
from __future__ import annotations

import numpy as np

FCC_BASIS = [[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]]
BCC_BASIS = [[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]


def fcc(a, cell, composition, shells, **kw):
    d = dict(structure=dict(lattice=np.diag([a, a, a]).tolist(), coords=FCC_BASIS,
                            species=["Al"] * 4, supercell=list(cell)),
             composition=composition,
             shell_weights={str(s): 1.0 / s for s in range(1, shells + 1)})
    d.update(kw)
    return d


def bcc(a, cell, composition, shells, **kw):
    d = dict(structure=dict(lattice=np.diag([a, a, a]).tolist(), coords=BCC_BASIS,
                            species=["Nb"] * 2, supercell=list(cell)),
             composition=composition,
             shell_weights={str(s): 1.0 / s for s in range(1, shells + 1)})
    d.update(kw)
    return d


def hcp(a, c_over_a, cell, composition, shells, **kw):
    c = c_over_a * a
    lat = [[a, 0.0, 0.0], [-0.5 * a, 0.5 * np.sqrt(3.0) * a, 0.0], [0.0, 0.0, c]]
    d = dict(structure=dict(lattice=lat, coords=[[1 / 3, 2 / 3, 0.25], [2 / 3, 1 / 3, 0.75]],
                            species=["Ti"] * 2, supercell=list(cell)),
             composition=composition,
             shell_weights={str(s): 1.0 / s for s in range(1, shells + 1)})
    d.update(kw)
    return d


def plain_mc_32(iterations=1_000_000, seed=42):
    """FCC 2x2x2 (a = 4.05), A16B16, 3 shells (1, 1/2, 1/3), all pair weights 1, target 0"""
    return fcc(4.05, (2, 2, 2), {"Al": 16, "Ti": 16}, 3, pair_weights=1.0,
               iterations=iterations, seed=seed, thread_config=1, chunk_size=100000)


def systematic_32():
    """same cell, A16B16, 2 shells (1, 1/2), all C(32,16) permutations"""
    d = fcc(4.05, (2, 2, 2), {"Al": 16, "Ti": 16}, 2, iteration_mode="systematic")
    return d


def quinary_256(iterations=1 << 22, seed=7):
    """FCC 4x4x4 (256 sites), Fe/Co/Ni/Cr 51 + Mn 52, 7 shells with weights 1/s"""
    return fcc(4.05, (4, 4, 4), {"Fe": 51, "Co": 51, "Ni": 51, "Cr": 51, "Mn": 52}, 7,
               iterations=iterations, seed=seed, thread_config=1)


def quinary_864(iterations=1 << 20, seed=7):
    return fcc(4.05, (6, 6, 6), {"Fe": 173, "Co": 173, "Ni": 173, "Cr": 173, "Mn": 172}, 7,
               iterations=iterations, seed=seed, thread_config=1)


def ternary_64(iterations=1_000_000, seed=1005):
    """FCC 2x2x4 (64 sites), Al 22 / Ti 21 / Mo 21, 4 shells (a serving-mix cell size)"""
    return fcc(4.05, (2, 2, 4), {"Al": 22, "Ti": 21, "Mo": 21}, 4, iterations=iterations,
               seed=seed, thread_config=1)


def quaternary_500(iterations=1_000_000, seed=1006):
    """FCC 5x5x5 (500 sites), Fe/Co/Ni/Cr 125 each, 6 shells (a serving-mix cell size)"""
    return fcc(4.05, (5, 5, 5), {"Fe": 125, "Co": 125, "Ni": 125, "Cr": 125}, 6,
               iterations=iterations, seed=seed, thread_config=1)


def ternary_108(iterations=10_000_000, seed=1001):
    return fcc(4.05, (3, 3, 3), {"Al": 36, "Ti": 36, "Mo": 36}, 3, iterations=iterations,
               seed=seed, thread_config=1)


def quaternary_bcc_128(iterations=10_000_000, seed=1002):
    return bcc(3.165, (4, 4, 4), {"Nb": 32, "Mo": 32, "Ta": 32, "W": 32}, 4,
               iterations=iterations, seed=seed, thread_config=1)


def hcp_128(iterations=10_000_000, seed=1003):
    return hcp(2.95, 1.633, (4, 4, 4), {"Ti": 64, "Zr": 64}, 5, iterations=iterations,
               seed=seed, thread_config=1)


def rocksalt_split_216(iterations=10_000_000, seed=1004):
    """templates/ti-al-b-n.json: rocksalt 3x3x3 (a = 4.253534), cation Ti/Al and anion N/B
    sublattices optimised independently (split mode), targets [-1, 1] as shipped, 3 shells"""
    a = 4.253534
    coords = [[0.5, 0.0, 0.0], [0.5, 0.5, 0.5], [0.0, 0.0, 0.5], [0.0, 0.5, 0.0],
              [0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]
    return dict(structure=dict(lattice=np.diag([a, a, a]).tolist(), coords=coords,
                               species=["Ti"] * 4 + ["N"] * 4, supercell=[3, 3, 3]),
                sublattice_mode="split",
                composition=[{"sites": "Ti", "Ti": 54, "Al": 54},
                             {"sites": "N", "N": 54, "B": 54}],
                target_objective=[-1, 1],
                shell_weights=[{"1": 1.0, "2": 0.5, "3": 1.0 / 3.0}] * 2,
                iterations=iterations, seed=[seed, seed + 1], thread_config=1)


PEROVSKITE_BASIS = [[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5],
                    [0.0, 0.5, 0.5]]


def perovskite_bx(seed=21, mixed=False, iterations=100_000):
    """cubic ABX3 perovskite 3x3x3 (a = 3.9; A Sr, B Ti, X O; 135 sites), the A sites left
    out.  Sublattices B (Ti/Zr) and X (O/F), whose sites have different neighbour counts
    (6 and 2 at a/2); mixed: one sublattice over the B and X sites together (Ti/Zr/O), where
    the counts differ inside the shuffle window"""
    a = 3.9
    d = dict(structure=dict(lattice=np.diag([a, a, a]).tolist(), coords=PEROVSKITE_BASIS,
                            species=["Sr", "Ti", "O", "O", "O"], supercell=[3, 3, 3]),
             shell_weights={"1": 1.0, "2": 0.5, "3": 1.0 / 3.0},
             iterations=iterations, seed=seed, thread_config=1)
    if mixed:
        from sqsgenerator_amd import parse_config
        st = parse_config(dict(d, composition={"sites": "Ti", "Ti": 27})).structure()
        bx = [i for i, z in enumerate(st.species) if z in (8, 22)]
        d["composition"] = {"sites": bx, "Ti": 36, "Zr": 36, "O": 36}
    else:
        d["composition"] = [{"sites": "Ti", "Ti": 14, "Zr": 13},
                            {"sites": "O", "O": 41, "F": 40}]
    return d


def reference_test_16(**kw):
    """python/tests/test_core.py:27-42: 16 sites, Ti8Al8"""
    d = dict(iteration_mode="systematic",
             structure=dict(lattice=np.diag([1, 2, 3]).tolist(),
                            coords=[[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.0, 0.5, 0.5],
                                    [0.5, 0.0, 0.5]],
                            species=[1] * 4, supercell=[2, 2, 1]),
             pair_weights=0, composition={"sites": "H", "Ti": 8, "Al": 8}, thread_config=1)
    d.update(kw)
    return d
