"""Pins of the oracle's paired-end primitives (bwa/src/aux.cu:411-500,
703-714, 1000-1070) by hand-worked cases and closed forms.

The reference never launches these ports, so it holds no vectors for them:
each case below is worked out from the definitions (orientation and distance
of two hits, the unique-hit rule, the quartile/bound/mean/deviation formulas
on size sets whose statistics are known in closed form, the pair score of
two hits under a normal insert-size model).  A wrong orientation code, a
swapped quartile index, a bound constant off, a population/sample deviation
mix-up or a wrong pair-score sign fails one of them."""
import math

import numpy as np
import pytest

import oracle_lib

L = 1000  # l_pac of the synthetic texts below


@pytest.fixture(scope="module")
def orc():
    return oracle_lib.Oracle()


def check_pairs(orc):
    """Every pin of this file against one oracle build (for the mutation checks)."""
    for t in (test_infer_dir_orientations, test_cal_sub_cases, test_pestat_uniform_sizes_closed_form,
              test_pestat_outliers_and_stddev_bound, test_pestat_filters_and_direction_counts,
              test_mem_pair_single_candidate, test_mem_pair_picks_the_best_and_counts_near_ones):
        t(orc)


def rev(fwd_end):
    """Text coordinate of a reverse-strand hit whose forward-strand end base is fwd_end."""
    return 2 * L - 1 - fwd_end


def test_infer_dir_orientations(orc):
    # FF: both forward, mate further right
    assert orc.infer_dir(L, 100, 300) == (0, 200)
    # FR: read 1 forward, mate reverse ending at 400 (the usual innie pair)
    assert orc.infer_dir(L, 100, rev(400)) == (1, 300)
    # RF: read 1 forward at 500, mate reverse to its left
    assert orc.infer_dir(L, 500, rev(200)) == (2, 300)
    # RR: both forward, mate to the left
    assert orc.infer_dir(L, 300, 100) == (3, 200)
    # both on the reverse strand: the distance is measured there
    assert orc.infer_dir(L, rev(400), rev(100)) == (0, 300)
    # the distance is symmetric in the two hits
    for b1, b2 in [(100, rev(400)), (700, 20), (rev(5), rev(900))]:
        assert orc.infer_dir(L, b1, b2)[1] == orc.infer_dir(L, b2, b1)[1]


def test_cal_sub_cases(orc):
    floor = orc.opt.min_seed_len * orc.opt.a  # 19 at the defaults
    assert floor == 19
    assert orc.cal_sub([0], [150], [150]) == floor                       # one hit
    assert orc.cal_sub([0, 0], [150, 150], [150, 120]) == 120            # full overlap
    assert orc.cal_sub([0, 100], [60, 150], [60, 50]) == floor           # disjoint on the read
    # overlap of 20 against a shorter hit of 50: 20 < 0.5 * 50 -> not a sub-hit
    assert orc.cal_sub([0, 40], [60, 90], [60, 50]) == floor
    # overlap of 30 against 50: 30 >= 25 -> sub-hit
    assert orc.cal_sub([0, 30], [60, 80], [60, 50]) == 50
    # overlap of exactly half the shorter hit counts (>=)
    assert orc.cal_sub([0, 35], [60, 85], [60, 50]) == 50
    # the first qualifying hit is taken, not the best
    assert orc.cal_sub([0, 0, 0], [150, 150, 150], [150, 90, 140]) == 90


def _pairs_to_arrays(pairs):
    """pairs: list of (hits1, hits2), each a list of (rid, rb, qb, qe, score)."""
    off, rid, rb, qb, qe, sc = [0], [], [], [], [], []
    for h1, h2 in pairs:
        for hits in (h1, h2):
            for r, b, q0, q1, s in hits:
                rid.append(r), rb.append(b), qb.append(q0), qe.append(q1), sc.append(s)
            off.append(len(rid))
    return off, rid, rb, qb, qe, sc


def _fr(sizes, read_len=100):
    """FR pairs of the given insert sizes: read 1 forward at 0, mate reverse ending at size."""
    return [([(0, 10, 0, read_len, read_len)], [(0, rev(10 + s), 0, read_len, read_len)]) for s in sizes]


def test_pestat_uniform_sizes_closed_form(orc):
    sizes = list(range(1, 101))  # quartile indices 25, 50, 75 -> 26, 51, 76
    pes = orc.pestat(L, *_pairs_to_arrays(_fr(sizes)))
    fr = pes[1]
    assert fr["failed"] == 0
    assert fr["avg"] == 50.5
    assert fr["std"] == pytest.approx(math.sqrt((100 ** 2 - 1) / 12), rel=1e-14)  # population deviation
    # mapping bounds: 26 - 3*50 -> clamped to 1; 76 + 3*50 = 226 (wider than avg +- 4 sd)
    assert (fr["low"], fr["high"]) == (1, 226)
    assert [p["failed"] for p in pes] == [1, 0, 1, 1]
    assert pes[0] == {"low": 0, "high": 0, "failed": 1, "avg": 0.0, "std": 0.0}


def test_pestat_outliers_and_stddev_bound(orc):
    # 40 sizes at 300 and 2 far outliers: quartiles 300, outlier bounds [300, 300]
    sizes = [300] * 40 + [900, 950]
    fr = orc.pestat(L, *_pairs_to_arrays(_fr(sizes)))[1]
    assert (fr["avg"], fr["std"]) == (300.0, 0.0)
    assert (fr["low"], fr["high"]) == (300, 300)
    # sorted: 200, 290 x10, 300 x20, 310 x10, 400 -> quartiles at indices 10, 21, 31
    # = 290, 300, 310; outlier band [250, 350] drops 200 and 400
    sizes = [290] * 10 + [300] * 20 + [310] * 10 + [200, 400]
    fr = orc.pestat(L, *_pairs_to_arrays(_fr(sizes)))[1]
    assert fr["avg"] == 300.0
    assert fr["std"] == pytest.approx(math.sqrt(50.0), rel=1e-14)
    assert (fr["low"], fr["high"]) == (230, 370)  # 290 - 60, 310 + 60
    sizes = [280] * 10 + [300] * 10 + [320] * 10 + [340] * 10
    fr = orc.pestat(L, *_pairs_to_arrays(_fr(sizes)))[1]
    # p25 = 300, p75 = 340, outlier band [220, 420]: all 40 kept, avg 310, sd sqrt(500)
    assert fr["avg"] == 310.0
    assert fr["std"] == pytest.approx(math.sqrt(500.0), rel=1e-14)
    # mapping band [300 - 120, 340 + 120] = [180, 460]; avg +- 4 sd = [220.6, 399.4] lies inside
    assert (fr["low"], fr["high"]) == (180, 460)
    # n = 41, sorted 300 x10, 310 x10, 320 x11, 330 x9, 365: quartile indices
    # (int)(10.25 + .499) = 10 and (int)(30.75 + .499) = 31 -> 310, 330; the
    # outlier band [270, 370] keeps 365 (a band of 1.5 IQR, or p75 read at
    # index 30, would drop it)
    sizes = [300] * 10 + [310] * 10 + [320] * 11 + [330] * 9 + [365]
    fr = orc.pestat(L, *_pairs_to_arrays(_fr(sizes)))[1]
    assert fr["avg"] == 12955 / 41
    assert fr["std"] == pytest.approx(float(np.std(np.array(sizes, dtype=np.float64))), rel=1e-12)
    assert (fr["low"], fr["high"]) == (250, 390)


def test_pestat_filters_and_direction_counts(orc):
    base = _fr([300 + (i % 7) for i in range(300)])
    # a pair whose read 1 has a strong second hit overlapping the first: not counted
    amb = ([(0, 10, 0, 100, 100), (0, 500, 0, 100, 95)], [(0, rev(410), 0, 100, 100)])
    # a pair on two contigs: not counted
    two = ([(0, 10, 0, 100, 100)], [(1, rev(310), 0, 100, 100)])
    # pairs beyond max_ins: not counted
    far = [([(0, 10, 0, 100, 100)], [(0, rev(10 + 700), 0, 100, 100)])] * 200
    # 9 RF pairs: below the 10-pair minimum
    rf = [([(0, 500, 0, 100, 100)], [(0, rev(200), 0, 100, 100)])] * 9
    pes_all = orc.pestat(L, *_pairs_to_arrays(base + [amb, two] + far + rf), max_ins=600)
    pes_base = orc.pestat(L, *_pairs_to_arrays(base), max_ins=600)
    assert pes_all[1] == pes_base[1]
    assert pes_all[2]["failed"] == 1
    # with the far pairs inside max_ins the upper quartile moves to 700
    assert orc.pestat(L, *_pairs_to_arrays(base + far), max_ins=800)[1]["high"] > 1000
    # 11 FF pairs against 300 FR: enough pairs, but under 5% of the top orientation
    ff = [([(0, 100, 0, 100, 100)], [(0, 400, 0, 100, 100)])] * 11
    pes = orc.pestat(L, *_pairs_to_arrays(base + ff))
    assert pes[0]["failed"] == 1 and pes[0]["avg"] == 300.0  # its statistics are still computed
    ff = [([(0, 100, 0, 100, 100)], [(0, 400, 0, 100, 100)])] * 10
    pes = orc.pestat(L, *_pairs_to_arrays(_fr([300] * 150) + ff))
    assert pes[0]["failed"] == 0  # 10 >= 150 * 0.05
    # max_ins cuts sizes off
    pes = orc.pestat(L, *_pairs_to_arrays(_fr([100] * 20 + [500] * 20)), max_ins=400)
    assert pes[1]["avg"] == 100.0


def test_hash_64_is_a_bijection_on_samples(orc):
    rng = np.random.default_rng(5)
    keys = [int(k) for k in rng.integers(0, 2 ** 63, size=2000, dtype=np.uint64)] + list(range(64))
    hs = [orc.hash_64(k) for k in keys]
    assert len(set(hs)) == len(set(keys))
    # Wang's mix: low-bit changes reach the high half
    assert (orc.hash_64(0) ^ orc.hash_64(1)) >> 32 != 0


def _q(s1, s2, dist, avg, sd, a=1):
    ns = (dist - avg) / sd
    return max(0, int(s1 + s2 + .721 * math.log(2. * math.erfc(abs(ns) / math.sqrt(2))) * a + .499))


PES_FR = [{"low": 0, "high": 0, "failed": 1, "avg": 0.0, "std": 0.0},
          {"low": 100, "high": 700, "failed": 0, "avg": 400.0, "std": 50.0},
          {"low": 0, "high": 0, "failed": 1, "avg": 0.0, "std": 0.0},
          {"low": 0, "high": 0, "failed": 1, "avg": 0.0, "std": 0.0}]


def test_mem_pair_single_candidate(orc):
    ctg = np.array([0], dtype=np.int64)
    # read 1 forward at 100, read 2 reverse with forward-strand end 549: distance 449
    o, sub, n_sub, z0, z1 = orc.mem_pair(L, ctg, PES_FR, 7, [0, 0], [100, rev(549)], [150, 140], 1)
    assert (o, sub, n_sub, z0, z1) == (_q(150, 140, 449, 400, 50), 0, 0, 0, 0)
    # a normal deviate of 0 costs nothing: q = s1 + s2
    o, *_ = orc.mem_pair(L, ctg, PES_FR, 7, [0, 0], [100, rev(500)], [150, 140], 1)
    assert o == 290
    # outside [low, high]: no pair
    assert orc.mem_pair(L, ctg, PES_FR, 7, [0, 0], [100, rev(900)], [150, 140], 1) == (0, 0, 0, -1, -1)
    # orientation FF has failed: two forward hits never pair
    assert orc.mem_pair(L, ctg, PES_FR, 7, [0, 0], [100, 500], [150, 140], 1) == (0, 0, 0, -1, -1)


def test_mem_pair_picks_the_best_and_counts_near_ones(orc):
    ctg = np.array([0], dtype=np.int64)
    # read 2 has three hits; the one at distance 400 scores best as a pair
    rid = [0, 0, 0, 0]
    rb = [100, rev(600), rev(500), rev(350)]
    sc = [150, 148, 140, 145]
    o, sub, n_sub, z0, z1 = orc.mem_pair(L, ctg, PES_FR, 3, rid, rb, sc, 1)
    qs = {1: _q(150, 148, 500, 400, 50), 2: _q(150, 140, 400, 400, 50), 3: _q(150, 145, 250, 400, 50)}
    best = max(qs, key=lambda k: qs[k])
    second = sorted(qs.values())[-2]
    assert (o, z0, z1) == (qs[best], 0, best - 1)
    assert sub == second
    tmp = max(orc.opt.a + orc.opt.b, orc.opt.o_del + orc.opt.e_del, orc.opt.o_ins + orc.opt.e_ins)
    assert n_sub == sum(1 for k, q in qs.items() if k != best and second - q <= tmp)
    # read 1 and read 2 exchanged: the same pair is found from the other end
    o2, sub2, _, z0b, z1b = orc.mem_pair(L, ctg, PES_FR, 3, rid[1:] + rid[:1], rb[1:] + rb[:1], sc[1:] + sc[:1], 3)
    assert (o2, sub2, z0b, z1b) == (o, sub, best - 1, 0)
