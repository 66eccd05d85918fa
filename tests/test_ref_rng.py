"""Pins the oracle's random-number restatement (oracle/jms_oracle.cpp OrcRng:
MT19937, rand = next() >> 1, unirand, Marsaglia-polar normrand) to the
reference's own CRng (rng.cpp:30-89), compiled from its source by
oracle/Makefile (target `ref`) into oracle/_ref/.  The library's variant
sequences and frozen-phonon sites are checked bit for bit against the oracle
elsewhere, so this closes the chain to the reference.  Skipped where the
reference tree (and so oracle/_ref) is absent."""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref_rng.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (reference sources absent)")


def _ref():
    L = ctypes.CDLL(REF)
    L.ref_rng_create.restype = ctypes.c_void_p
    L.ref_rng_create.argtypes = [ctypes.c_int]
    L.ref_rng_destroy.argtypes = [ctypes.c_void_p]
    L.ref_rng_seed.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.ref_rng_rand.argtypes = [ctypes.c_void_p]
    L.ref_rng_rand.restype = ctypes.c_int
    for f in ("ref_rng_unirand", "ref_rng_unirand_oub", "ref_rng_normrand"):
        getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        getattr(L, f).restype = ctypes.c_double
    return L


@pytest.mark.parametrize("seed", [1, 42, 1000, 1013, 123456789])
def test_rand_and_unirand_bit_exact(seed):
    L = _ref()
    r = L.ref_rng_create(seed)
    o = O.Rng(seed)
    for i in range(20000):
        if i % 3 == 0:
            assert L.ref_rng_rand(r) == o.rand()
        else:
            a, b = (0.0, float(i % 17 + 1))
            assert L.ref_rng_unirand(r, a, b) == o.unirand(a, b)
    L.ref_rng_destroy(r)


@pytest.mark.parametrize("seed", [7, 1000, 1008, 1015])
def test_normrand_bit_exact_interleaved(seed):
    """Normal deviates come in pairs (the second is cached); interleaving them
    with uniform draws and re-seeding must keep both generators in step."""
    L = _ref()
    r = L.ref_rng_create(seed)
    o = O.Rng(seed)
    for i in range(20000):
        if i % 5 == 4:
            assert L.ref_rng_unirand(r, 0.0, 8.0) == o.unirand(0.0, 8.0)
        else:
            assert L.ref_rng_normrand(r, 0.0, 0.08) == o.normrand(0.0, 0.08)
    L.ref_rng_destroy(r)


def test_frozen_phonon_draws_match_reference_rng():
    """The library's displacement draws of one variant (x then y per site,
    rms 0.08 A, seed 1000 + 8*type + v) are the reference generator's normal
    deviates, bit for bit."""
    from jmultislicelib_amd import configs
    L = _ref()
    xy0, _ = configs.perovskite_sites(2, 0, 1, 0, 0.08, 1000)
    xy, _ = configs.perovskite_sites(2, 0, 8, 3, 0.08, 1000)
    r = L.ref_rng_create(1000 + 8 * 0 + 3)
    d = np.array([[L.ref_rng_normrand(r, 0.0, 1.0) for _ in range(2)] for _ in range(len(xy))])
    L.ref_rng_destroy(r)
    expect = xy0 + 0.08 * d  # the library adds rms * normrand(0, 1) in double
    assert np.array_equal(xy, expect)
