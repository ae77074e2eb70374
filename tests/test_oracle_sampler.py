"""Oracle sampler: Philox4x32-10, Lemire's bounded integers, Floyd triples.

Pinned by: the Random123 known-answer vectors of Philox4x32-10 (Salmon et
al., SC'11), 64 vectors from ROCm's independent rocRAND implementation
(tests/golden/philox_rocrand.json), hand-computed Lemire cases, and a
chi-square test of the triple distribution against the uniform law over
3-subsets on tiny N.
"""
import itertools
import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "philox_rocrand.json")

# Random123 kat_vectors for philox4x32 with 10 rounds.
RANDOM123_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def test_philox_random123_kat(orc):
    for ctr, key, out in RANDOM123_KAT:
        assert orc.philox(ctr, key) == out


def test_philox_matches_rocrand(orc):
    vec = json.load(open(GOLDEN))["vectors"]
    assert len(vec) == 64
    for v in vec:
        assert orc.philox(v["ctr"], v["key"]) == v["out"]


def test_lemire_hand_cases(orc):
    # bound 3: word 0 has low product 0 < (2^32 mod 3) = 1 and is rejected;
    # 0x80000000 * 3 = 0x1_8000_0000 gives 1 with low word 0x80000000.
    assert orc.bounded_from_words([0, 0x80000000], 3) == (1, 2)
    # bound 2^31: 2^32 mod 2^31 = 0, nothing is rejected; x * 2^31 >> 32 = x >> 1.
    assert orc.bounded_from_words([0xFFFFFFFF], 1 << 31) == (0x7FFFFFFF, 1)
    assert orc.bounded_from_words([0], 1 << 31) == (0, 1)
    # bound 5: 2^32 mod 5 = 1, so word 0 (low 0) is rejected, word 1 gives 0.
    assert orc.bounded_from_words([0, 1], 5) == (0, 2)
    # bound 7: 2^32 mod 7 = 4; x = 1 -> low 7 >= 4, accepted, value 0.
    assert orc.bounded_from_words([1], 7) == (0, 1)
    # high word: x = 0xFFFFFFFF, bound 10 -> floor(0xFFFFFFFF * 10 / 2^32) = 9.
    assert orc.bounded_from_words([0xFFFFFFFF], 10) == (9, 1)
    # the list running out reports -1
    assert orc.bounded_from_words([0, 0], 3)[0] == -1


def test_lemire_threshold_is_2_32_mod_bound(orc):
    rng = np.random.default_rng(0)
    for bound in [3, 5, 6, 7, 1000, 65537, 123456789, (1 << 31) + 11]:
        thr = (1 << 32) % bound
        if thr == 0:
            continue
        # the largest rejected low word is thr - 1: x with x*bound mod 2^32 = thr - 1
        inv = pow(bound, -1, 1 << 32) if bound % 2 else None
        if inv is None:
            continue
        x_rej = ((thr - 1) * inv) % (1 << 32)
        x_acc = (thr * inv) % (1 << 32)
        follow = int(rng.integers(0, 1 << 32))
        v, used = orc.bounded_from_words([x_rej, follow], bound)
        assert used == 2 or ((follow * bound) % (1 << 32) < thr)
        v, used = orc.bounded_from_words([x_acc], bound)
        assert used == 1 and v == (x_acc * bound) >> 32


def test_triples_distinct_in_range(orc):
    for n in [3, 4, 5, 17, 1000, 2**20 + 7]:
        t = orc.sample_triples(seed=99, n=n, first=0, count=4000)
        assert t.min() >= 0 and t.max() < n
        assert np.all(t[:, 0] != t[:, 1]) and np.all(t[:, 0] != t[:, 2]) and np.all(t[:, 1] != t[:, 2])


def test_triples_n3_is_a_permutation(orc):
    t = orc.sample_triples(seed=5, n=3, first=0, count=200)
    assert np.all(np.sort(t, axis=1) == np.array([0, 1, 2]))


def test_triples_uniform_over_subsets(orc):
    """Floyd's algorithm gives every 3-subset the same probability."""
    for n in [5, 6, 8]:
        count = 60000
        t = np.sort(orc.sample_triples(seed=1234 + n, n=n, first=0, count=count), axis=1)
        subsets = list(itertools.combinations(range(n), 3))
        idx = {s: i for i, s in enumerate(subsets)}
        hist = np.zeros(len(subsets))
        for row in map(tuple, t):
            hist[idx[row]] += 1
        exp = count / len(subsets)
        chi2 = float(((hist - exp) ** 2 / exp).sum())
        dof = len(subsets) - 1
        # mean dof, sd sqrt(2 dof): six sigma is far beyond a fair generator
        assert chi2 < dof + 6 * math.sqrt(2 * dof), (n, chi2)


def test_triples_stream_is_counter_based(orc):
    """Sample g depends only on (seed, g): any window reproduces the stream."""
    a = orc.sample_triples(seed=42, n=5000, first=0, count=300)
    b = orc.sample_triples(seed=42, n=5000, first=100, count=100)
    assert np.array_equal(a[100:200], b)
    c = orc.sample_triples(seed=43, n=5000, first=0, count=300)
    assert not np.array_equal(a, c)


def test_triples_follow_floyd_on_philox_words(orc):
    """Recomputes samples from raw Philox words with Python integers."""
    def words(seed, g):
        blk = 0
        while True:
            for w in orc.philox([g & 0xFFFFFFFF, g >> 32, blk, 0], [seed & 0xFFFFFFFF, seed >> 32]):
                yield w
            blk += 1

    def bounded(it, b):
        x = next(it)
        m = x * b
        thr = (1 << 32) % b
        while (m & 0xFFFFFFFF) < thr:
            x = next(it)
            m = x * b
        return m >> 32

    seed, n = 0xDEADBEEF12345678, 777
    got = orc.sample_triples(seed, n, 1000, 50)
    for s in range(50):
        it = words(seed, 1000 + s)
        i0 = bounded(it, n - 2)
        i1 = bounded(it, n - 1)
        i1 = n - 2 if i1 == i0 else i1
        i2 = bounded(it, n)
        i2 = n - 1 if i2 in (i0, i1) else i2
        assert list(got[s]) == [i0, i1, i2]
