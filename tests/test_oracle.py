"""Pins of the C oracle against the reference's golden values and known answers, plus checks
fixed by mathematics (closed forms, brute force, invariants).  CPU only."""
import itertools
import json
import math
from collections import Counter
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = Path(__file__).parent / "golden"
M64 = (1 << 64) - 1


# ---- RNG: rapidhash.h vectors (built from the reference header into oracle/_ref) ----------

def test_rapid_mix_golden_vectors():
    d = json.loads((GOLDEN / "rapid_mix.json").read_text())
    for a, b, out in d["mix"]:
        assert O.rapid_mix(a, b) == out


def test_rapid_mix_against_reference_build():
    ref = O.ref_lib()
    if ref is None:
        pytest.skip("reference tree not mounted: oracle/_ref not built")
    rng = np.random.default_rng(5)
    for a, b in rng.integers(0, 2 ** 63, size=(500, 2), dtype=np.uint64).tolist():
        assert O.rapid_mix(a << 1 | 1, b) == ref.ref_rapid_mix(a << 1 | 1, b)


def test_rapidrand_streams():
    """shuffle.h:23-26 driven by the reference's own mixer and secrets"""
    d = json.loads((GOLDEN / "rapid_mix.json").read_text())
    for seed, stream in d["rapidrand_streams"].items():
        assert O.rapidrand_stream(int(seed), len(stream)).tolist() == stream


def test_rapidrand_is_a_counter():
    """draw k of a stream only depends on seed + k * secret0 (jump-ahead used on device)"""
    d = json.loads((GOLDEN / "rapid_mix.json").read_text())
    s0, s1 = d["secret0"], d["secret1"]
    seed = 42
    stream = O.rapidrand_stream(seed, 50)
    for k in (0, 7, 49):
        st = (seed + (k + 1) * s0) & M64
        assert int(stream[k]) == O.rapid_mix(st, st ^ s1)


def test_random_bounded_range_and_formula():
    st = 12345
    for rng in (1, 2, 3, 17, 255, 864):
        v, st2 = O.random_bounded(rng, st)
        x = int(O.rapidrand_stream(st, 1)[0]) & 0xFFFFFFFF
        assert v == (x * rng) >> 32 and 0 <= v < rng
        st = st2


def test_fisher_yates_uniform():
    """all 4! orders appear with equal frequency (Sattolo's variant would give only 3!)"""
    counts = Counter()
    st = 99
    base = np.arange(4, dtype=np.uint8)
    n = 24000
    for _ in range(n):
        out, st = O.shuffle_random(base, [0, 4], st)
        counts[tuple(out.tolist())] += 1
    assert len(counts) == 24
    chi2 = sum((c - n / 24) ** 2 / (n / 24) for c in counts.values())
    assert chi2 < 60  # 23 dof: p ~ 4e-5


def test_fisher_yates_windows_are_independent():
    conf = np.array([0, 0, 1, 1, 2, 2, 3, 3], dtype=np.uint8)
    out, _ = O.shuffle_random(conf, [0, 4, 4, 8], 7)
    assert sorted(out[:4]) == [0, 0, 1, 1] and sorted(out[4:]) == [2, 2, 3, 3]


# ---- permutations: tests/test_permutation.cpp:15-102 ------------------------------------

first_16 = [0] * 8 + [1] * 8
last_16 = [1] * 8 + [0] * 8
first_24 = [0] * 8 + [1] * 8 + [2] * 8
last_24 = [2] * 8 + [1] * 8 + [0] * 8
diff_20 = list(range(20))
same_20 = [0] * 20


def test_num_permutations_known_answers():
    assert O.num_permutations([]) == 1
    assert O.num_permutations([0, 1]) == 2 and O.num_permutations([1, 0]) == 2
    assert O.num_permutations(first_16) == 12870 == O.num_permutations(last_16)
    assert O.num_permutations(first_24) == 9465511770 == O.num_permutations(last_24)
    assert O.num_permutations(diff_20) == 2432902008176640000
    assert O.num_permutations(same_20) == 1


@pytest.mark.parametrize("conf", [[], [0, 1], first_16, first_24, same_20, diff_20])
def test_rank_first_is_one(conf):
    assert O.rank_permutation(conf) == 1
    assert O.unrank_permutation(conf, 1).tolist() == list(conf)


@pytest.mark.parametrize("conf", [[], [1, 0], last_16, last_24, same_20])
def test_rank_last_is_total(conf):
    assert O.rank_permutation(conf) == O.num_permutations(conf)
    assert O.unrank_permutation(conf, O.num_permutations(conf)).tolist() == list(conf)


def test_unrank_rank_round_trip_all_12870():
    out = np.array(first_16, dtype=np.uint8)
    for r in range(1, 12871):
        assert O.rank_permutation(O.unrank_permutation(out, r)) == r


def test_next_permutation_matches_unrank():
    out = np.array(first_16, dtype=np.uint8)
    for r in range(1, 12870):
        assert out.tolist() == O.unrank_permutation(first_16, r).tolist()
        out, ok = O.next_permutation(out)
        assert ok and O.rank_permutation(out) == r + 1
    assert not O.next_permutation(out)[1]


@pytest.mark.parametrize("ms", [[0, 0, 1, 1, 2], [0, 1, 1, 2, 2, 2], [0, 0, 0, 1, 2, 3]])
def test_rank_is_lexicographic_position(ms):
    """brute force: rank = 1 + position in the sorted list of distinct permutations"""
    perms = sorted(set(itertools.permutations(ms)))
    assert O.num_permutations(ms) == len(perms)
    for pos, p in enumerate(perms):
        assert O.rank_permutation(p) == pos + 1
        assert tuple(O.unrank_permutation(ms, pos + 1).tolist()) == p


def test_unrank_out_of_range():
    with pytest.raises(IndexError):
        O.unrank_permutation(first_16, 12871)


# ---- geometry: tests/test_structure.cpp, tests/assets ------------------------------------

def _case1():
    return json.loads((GOLDEN / "case1.structure.json").read_text())


def test_distance_matrix_case1_fixture():
    """test_structure.cpp:113-118 (pymatgen-generated fixture, tolerance 1e-7)"""
    d = _case1()
    dm = O.distance_matrix(np.array(d["lattice"]), np.array(d["frac_coords"]))
    assert np.allclose(dm, np.array(d["distance_matrix"]), atol=1e-7, rtol=0)
    assert np.all(dm == dm.T) and np.all(np.diag(dm) == 0)


def test_shell_matrix_case1_fixture():
    """test_structure.cpp:120-136: naive radii (eps, 1e-9) reproduce the fixture exactly"""
    d = _case1()
    dm = O.distance_matrix(np.array(d["lattice"]), np.array(d["frac_coords"]))
    sm = O.shell_matrix(dm, O.distances_naive(dm))
    assert np.array_equal(sm, np.array(d["shell_matrix"]))


def test_supercell_case1_fixture():
    """test_structure.cpp:146-159: every computed site matches a fixture site"""
    d = json.loads((GOLDEN / "case1.supercell.json").read_text())
    st, sc = d["structure"], d["supercell"]
    L, f, z = O.supercell(np.array(st["lattice"]), np.array(st["frac_coords"]),
                          np.array(st["species"]), *d["shape"])
    exp_f, exp_z = np.array(sc["frac_coords"]), np.array(sc["species"])
    assert len(f) == len(exp_f)
    for k in range(len(f)):
        close = np.all(np.abs(exp_f - f[k]) <= 1e-5 + 1e-3 * np.abs(exp_f), axis=1)
        assert np.any(close & (exp_z == z[k]))


def _fcc(a, cell):
    return O.supercell(np.eye(3) * a, [[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]],
                       [1] * 4, *cell)


def test_fcc_naive_distances_and_coordination():
    """test_structure.cpp:209-238"""
    L, f, z = _fcc(1.0, (3, 3, 3))
    dm = O.distance_matrix(L, f)
    d = O.distances_naive(dm)
    expect = [0, 1 / math.sqrt(2), 1, math.sqrt(1.5), math.sqrt(2), math.sqrt(2.5), math.sqrt(3),
              math.sqrt(3.5), math.sqrt(4.5), math.sqrt(5.5)]
    assert np.allclose(d, expect, atol=1e-5, rtol=1e-3)
    pref = O.compute_prefactors(O.shell_matrix(dm, d), [1, 2, 3, 4, 5, 6], z)
    cn = [int((1.0 / p) / len(z)) for p in pref.ravel()]
    assert cn == [12, 6, 24, 12, 12, 8]


@pytest.mark.parametrize("cell", [(2, 2, 2), (3, 3, 3)])
def test_peak_equals_naive_on_perfect_lattice(cell):
    """test_parser.cpp:281-309: with bin_width 0.001 both detections give the same radii
    (bit-identical on the reference's 2x2x2 cell; the running average of "naive" and the bin
    maximum of "peak" may differ in the last ulp on larger cells)"""
    L, f, _ = O.supercell(np.diag([1.0, 2.0, 3.0]), [[0, 0, 0], [0, .5, .5], [.5, 0, .5],
                                                      [.5, .5, 0]], [1] * 4, *cell)
    dm = O.distance_matrix(L, f)
    peak, naive = O.distances_histogram(dm, 0.001, 0.25), O.distances_naive(dm, 1e-3, 1e-5)
    if cell == (2, 2, 2):
        assert peak.tolist() == naive.tolist()
    else:
        assert len(peak) == len(naive) and np.allclose(peak, naive, rtol=1e-14, atol=0)


def test_prefactor_closed_form():
    """pref = 1 / (M x_a x_b N) with M the coordination number of the shell"""
    L, f, _ = _fcc(4.05, (2, 2, 2))
    z = np.array([13] * 16 + [22] * 16, dtype=np.uint8)
    dm = O.distance_matrix(L, f)
    sm = O.shell_matrix(dm, O.distances_histogram(dm))
    pref = O.compute_prefactors(sm, [1, 2, 3], z)
    M = [12, 3, 12]  # a 2x2x2 FCC cell folds shell 2 (6 -> 3) and shell 3 (24 -> 12) images
    for s in range(3):
        assert pref[s, 0, 1] == 1.0 / (M[s] * 0.5 * 0.5 * 32)
        assert np.all(pref[s] == pref[s, 0, 1])


def test_scaled_pair_weights_closed_form():
    pw = np.arange(18, dtype=float).reshape(2, 3, 3)
    w = [0.5, 0.25]
    out = O.scaled_pair_weights(pw, w)
    assert np.array_equal(out, pw * np.array(w)[:, None, None])


# ---- hot path: count_bonds / compute_objective ---------------------------------------------

def _random_problem(seed=0):
    rng = np.random.default_rng(seed)
    L, f, _ = _fcc(4.05, (3, 3, 3))
    dm = O.distance_matrix(L, f)
    radii = O.distances_histogram(dm)
    sm = O.shell_matrix(dm, radii)
    pairs = O.pairs(sm, [1, 2, 3, 4])
    conf = rng.permutation(np.repeat(np.arange(3, dtype=np.uint8), 36))
    return sm, pairs, conf


def test_count_bonds_equals_dense_contraction():
    """bonds(s) = C^T A_s^+ C with C the one-hot species matrix and A_s^+ the strict upper
    triangle of the shell-s adjacency (a different formula than the pair loop)"""
    sm, pairs, conf = _random_problem()
    bonds = O.count_bonds(pairs, conf, 4, 3)
    C = np.eye(3, dtype=np.int64)[conf]
    for s, shell in enumerate([1, 2, 3, 4]):
        A = np.triu((sm == shell).astype(np.int64), k=1)
        assert np.array_equal(bonds[s].astype(np.int64), C.T @ A @ C)


def test_count_bonds_totals():
    """python/tests/test_core.py:245-248: the bonds of a shell sum to its pair count"""
    sm, pairs, conf = _random_problem(1)
    bonds = O.count_bonds(pairs, conf, 4, 3)
    assert np.array_equal(bonds.sum(axis=(1, 2)), np.bincount(pairs[:, 2].astype(int)))


def test_objective_zero_weights():
    """python/tests/test_core.py:39,59: pair_weights 0 -> objective 0"""
    sm, pairs, conf = _random_problem(2)
    bonds = O.count_bonds(pairs, conf, 4, 3)
    o, _ = O.compute_objective(bonds, np.full((4, 3, 3), 0.01), np.zeros((4, 3, 3)),
                               np.zeros((4, 3, 3)))
    assert o == 0.0


def test_sro_with_unit_prefactors():
    """python/tests/test_core.py:249: with prefactors 1, sro = 1 - symmetrised bonds"""
    sm, pairs, conf = _random_problem(3)
    bonds = O.count_bonds(pairs, conf, 4, 3).astype(np.int64)
    o, sro = O.compute_objective(bonds, np.ones((4, 3, 3)), np.ones((4, 3, 3)),
                                 np.zeros((4, 3, 3)))
    sym = bonds + np.transpose(bonds, (0, 2, 1))
    sym[:, np.arange(3), np.arange(3)] //= 2
    assert np.array_equal(sro, 1.0 - sym)
    iu = np.triu_indices(3)
    assert o == float(np.sum(np.abs(1.0 - sym[:, iu[0], iu[1]])))


def test_objective_closed_form():
    """sum over s, a <= b of w |1 - n p - t| for a hand-made binary case"""
    bonds = np.array([[[3, 1], [2, 4]]], dtype=np.uint64)
    pref = np.full((1, 2, 2), 0.125)
    pw = np.array([[[2.0, 0.5], [0.5, -1.0]]])
    tgt = np.array([[[0.25, -0.5], [-0.5, 0.0]]])
    o, sro = O.compute_objective(bonds, pref, pw, tgt)
    # diagonal (0,0): 1 - 3/8 = 0.625; off (0,1): 1 - 3/8 = 0.625; diagonal (1,1): 0.5
    assert o == 2.0 * abs(0.625 - 0.25) + 0.5 * abs(0.625 + 0.5) + -1.0 * abs(0.5 - 0.0)
    assert sro[0].tolist() == [[0.625, 0.625], [0.625, 0.5]]


def test_objective_summation_order():
    """optimization.h:85-100 adds (0,0), (0,1), (1,1): with terms 1e16, 1, -1e16 that order
    loses the 1 (a diagonal-first order would return 1)"""
    bonds = np.zeros((1, 2, 2), dtype=np.uint64)
    pw = np.array([[[1e16, 1.0], [1.0, -1e16]]])
    o, _ = O.compute_objective(bonds, np.zeros((1, 2, 2)), pw, np.zeros((1, 2, 2)))
    assert o == 0.0


def test_term_zero_sign_never_changes_the_objective():
    """the dense kernels take |v - tgt| as fabs where the reference's absolute (numeric.h:15)
    gives -0.0 for +0.0: sums from +0.0 of the same terms with either zero sign are equal bit
    for bit, including all-zero sums, exact cancellations and negative pair weights"""
    rng = np.random.default_rng(5)

    def ref_abs(a):
        return a if a > 0 else -a

    def total(terms_v, pw, absf):
        obj = 0.0
        for v, w in zip(terms_v, pw):
            obj += w * absf(v)
        return obj

    cases = [([0.0, 0.0, 0.0], [1.0, -2.0, 0.5]),  # every term a zero
             ([0.25, 0.0, -0.25], [4.0, 3.0, -4.0]),  # cancels exactly to zero, then a zero
             ([0.0, 0.125, 0.0], [-1.0, -8.0, 2.0])]
    for _ in range(2000):
        n = int(rng.integers(1, 30))
        v = rng.choice([0.0, -0.0, 0.5, -0.5, 1.0 / 3.0], size=n) * rng.choice([1.0, 1e-300, 3.0],
                                                                             size=n)
        w = rng.choice([1.0, -1.0, 0.5, 2.0, 0.0], size=n)
        cases.append((list(v), list(w)))
    for v, w in cases:
        a = np.float64(total(v, w, ref_abs)).view(np.uint64)
        b = np.float64(total(v, w, math.fabs)).view(np.uint64)
        assert a == b, (v, w)


# ---- driver: result collection semantics (sqs.h:348-455, results.h:42-99) --------------

def _reference_test_16_problem(pair_weights=None):
    L, f, _ = O.supercell(np.diag([1.0, 2.0, 3.0]), [[0, 0, 0], [.5, .5, 0], [0, .5, .5],
                                                      [.5, 0, .5]], [1] * 4, 2, 2, 1)
    dm = O.distance_matrix(L, f)
    radii = O.distances_histogram(dm)
    shells = list(range(1, len(radii)))
    sm = O.shell_matrix(dm, radii)
    z = np.array([13] * 8 + [22] * 8, dtype=np.uint8)
    pref = O.compute_prefactors(sm, shells, z)
    S = len(shells)
    pw = np.zeros((S, 2, 2)) if pair_weights == 0 else np.repeat((1 - np.eye(2))[None], S, 0)
    pw = O.scaled_pair_weights(pw, [1.0 / s for s in shells])
    return O.Problem(O.pairs(sm, shells), (z == 22).astype(np.uint8), [0, 16], pref, pw,
                     np.zeros((S, 2, 2)))


def test_systematic_all_results_oracle():
    """python/tests/test_core.py:45-67 on the oracle driver"""
    prob = _reference_test_16_problem(pair_weights=0)
    st, res = O.run([prob], 12870, systematic=True, threads=4)
    assert st.finished == 12870 and len(res) == 12870
    assert {O.rank_permutation(sp) for _, sp, _ in res} == set(range(1, 12871))
    assert all(o == 0.0 for o, _, _ in res)


def test_keep_levels_contain_all_ties():
    """every candidate with objective <= the keep-th best distinct level is inserted,
    whatever the order (results.h:78-82 threshold)"""
    prob = _reference_test_16_problem()
    objs = {}
    conf = np.array(prob.species)
    for r in range(1, 12871):
        c = O.unrank_permutation(conf, r)
        o, _ = O.compute_objective(O.count_bonds(prob.pairs, c, prob.S, prob.K), prob.pref,
                                   prob.pw, prob.tgt)
        objs.setdefault(o, set()).add(r)
    levels = sorted(objs)
    for keep in (1, 2):
        st, res = O.run([prob], 12870, chunk_size=1000, systematic=True, keep=keep, threads=4)
        got = {}
        for o, sp, _ in res:
            got.setdefault(o, set()).add(O.rank_permutation(sp))
        for lv in levels[:keep + 1]:
            assert got[lv] == objs[lv]


def test_max_results_per_objective():
    """sqs.h:423-426: a level stops growing once it holds more than max entries"""
    prob = _reference_test_16_problem(pair_weights=0)
    st, res = O.run([prob], 12870, systematic=True, max_results_per_objective=5, threads=1)
    assert len(res) == 6


# ---- the collection rule on an explicit candidate stream ---------------------------------
# A plain-Python statement of the worker loop on one thread (sqs.h:384-441, 461-466) that the
# oracle's driver must reproduce, in both its candidate modes.

SECRET0 = 0x2d358dccaa6c78a5  # rapidhash.h:129-131


def _block_of(iterations, chunk):
    """BS::thread_pool's detach_blocks(0, iterations, f, iterations / chunk): near-equal
    blocks, the first `iterations % blocks` one larger"""
    nb = max(1, iterations // chunk)
    size, rem = divmod(iterations, nb)
    out, t = [], 0
    for b in range(nb):
        n = size + (1 if b < rem else 0)
        out += [b] * n
        t += n
    assert t == iterations
    return out


def _collect(stream, blocks, keep, max_results, interact=True):
    """stream: (configuration bytes, objective) per candidate in index order"""
    levels, search, stalled = {}, np.finfo(np.float64).max, None
    for (conf, o), b in zip(stream, blocks):
        if b == stalled or not (o <= search):
            continue
        if max_results is not None and o in levels and len(levels[o]) > max_results:
            if interact:  # the skipped configuration is scored again until the block ends
                stalled = b
            continue
        lv = levels.setdefault(o, [])
        if conf not in lv:
            lv.append(conf)
        distinct = sorted(levels)
        if len(distinct) > keep:
            search = min(search, distinct[keep])
    return {o: set(v) for o, v in levels.items()}


def _score(prob, c):
    return O.compute_objective(O.count_bonds(prob.pairs, c, prob.S, prob.K), prob.pref, prob.pw,
                               prob.tgt)[0]


def _counter_stream(prob, iterations, chunk, seed):
    """the device path's Monte Carlo candidates: candidate t is the sorted configuration
    shuffled with the draws [t D, (t + 1) D) of the seed's stream, the first candidate of an
    interact block unshuffled"""
    D = sum(int(hi) - int(lo) - 1 for lo, hi in np.asarray(prob.bounds).reshape(-1, 2))
    blocks = _block_of(iterations, chunk)
    out = []
    for t in range(iterations):
        c = prob.species.copy()
        if t and blocks[t] == blocks[t - 1]:
            c, _ = O.shuffle_random(c, prob.bounds, (seed + t * D * SECRET0) % (1 << 64))
        out.append((c.tobytes(), _score(prob, c)))
    return out, blocks


def _pack(res):
    out = {}
    for o, sp, _ in res:
        out.setdefault(o, set()).add(sp.tobytes())
    return out


def test_block_split_is_near_equal():
    assert [_block_of(5001, 700).count(b) for b in range(7)] == [715] * 3 + [714] * 4
    assert _block_of(10, 10) == [0] * 10
    assert [_block_of(12870, 1000).count(b) for b in range(12)] == [1073] * 6 + [1072] * 6


def test_counter_stream_holds_every_candidate():
    """keep above the candidate count inserts every distinct configuration: the oracle's
    counter-stream pack is exactly the set of candidates built from the golden-pinned
    shuffle and draw stream (block starts unshuffled)"""
    prob = _reference_test_16_problem()
    prob.seed = 77
    stream, _ = _counter_stream(prob, 2001, 300, 77)
    _, res = O.run([prob], 2001, chunk_size=300, keep=10 ** 9, counter_stream=True)
    got = _pack(res)
    want = {}
    for conf, o in stream:
        want.setdefault(o, set()).add(conf)
    assert got == want


@pytest.mark.parametrize("max_results", [None, 0, 1])
@pytest.mark.parametrize("keep", [1, 2])
def test_counter_stream_collection_rule(keep, max_results):
    prob = _reference_test_16_problem()
    prob.seed = 5
    stream, blocks = _counter_stream(prob, 1501, 400, 5)
    _, res = O.run([prob], 1501, chunk_size=400, keep=keep, max_results_per_objective=max_results,
                   counter_stream=True)
    assert _pack(res) == _collect(stream, blocks, keep, max_results)


@pytest.mark.parametrize("max_results", [None, 0, 2])
def test_systematic_collection_rule(max_results):
    """systematic blocks: rank order, the stall after a max_results skip included"""
    prob = _reference_test_16_problem()
    conf = np.array(prob.species)
    stream = []
    for r in range(1, 12871):
        c = O.unrank_permutation(conf, r)
        stream.append((c.tobytes(), _score(prob, c)))
    blocks = _block_of(12870, 1000)
    _, res = O.run([prob], 12870, chunk_size=1000, systematic=True, keep=3,
                   max_results_per_objective=max_results)
    assert _pack(res) == _collect(stream, blocks, 3, max_results)
