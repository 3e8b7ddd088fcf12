"""Pin the oracle's random-number pieces against library routines and
closed-form recurrences (CPU only)."""
import ctypes as C
import ctypes.util

import numpy as np
import pytest
from scipy import stats

import oracle as O
from gfx import PRIMES, mwc_state


def test_mt19937_against_numpy_legacy_seeding():
    # gsl_rng_set(mt19937, s) is init_genrand(s); numpy's legacy seeding is too
    for seed in (1, 42, 5489, 123456789):
        out = np.zeros(2000, dtype=np.uint32)
        O.lib().gfo_mt19937(seed, 2000, O.ptr(out))
        bg = np.random.MT19937()
        bg._legacy_seeding(seed)
        ref = bg.random_raw(2000).astype(np.uint32)
        assert np.array_equal(out, ref)


def test_mt19937_cpp_standard_known_answer():
    # [rand.predef]: the 10000th output of a default-seeded mt19937 is 4123659995
    out = np.zeros(10000, dtype=np.uint32)
    O.lib().gfo_mt19937(5489, 10000, O.ptr(out))
    assert int(out[-1]) == 4123659995


def test_rand48_against_glibc():
    libc = C.CDLL(ctypes.util.find_library("c"))
    libc.drand48.restype = C.c_double
    libc.lrand48.restype = C.c_long
    libc.srand48.argtypes = [C.c_long]
    for seed in (0, 1, 2271, 0xFFFFFFFF, 16777215):
        d = np.zeros(300, dtype=np.float64)
        l = np.zeros(300, dtype=np.uint32)
        O.lib().gfo_rand48(seed, 300, O.ptr(d), None)
        libc.srand48(seed)
        assert all(libc.drand48() == d[i] for i in range(300))
        O.lib().gfo_rand48(seed, 300, None, O.ptr(l))
        libc.srand48(seed)
        assert all(libc.lrand48() == int(l[i]) for i in range(300))


def test_mwc_seeding_interleaves_drand_lrand_and_overflows_like_x86():
    seed = 987654
    st = mwc_state(seed, 65536)
    ns = 65536
    a, c, x = st[:ns], st[ns:2 * ns], st[2 * ns:]
    assert np.array_equal(a, PRIMES[:ns])
    # recompute with an independent rand48 in python integers
    X = (seed << 16) | 0x330E
    A_, C_ = 0x5DEECE66D, 0xB
    M48 = (1 << 48) - 1
    for i in list(range(0, 64)) + list(range(54730, 54800)) + [65535]:
        pass
    cs, xs = [], []
    for i in range(ns):
        X = (A_ * X + C_) & M48
        v = (X / float(1 << 48)) * float(PRIMES[i])
        cs.append(0x80000000 if v >= 2**31 else int(v))
        X = (A_ * X + C_) & M48
        xs.append(X >> 17)
    assert np.array_equal(c, np.array(cs, dtype=np.uint32))
    assert np.array_equal(x, np.array(xs, dtype=np.uint32))
    big = PRIMES[:ns] > 2**31
    assert big.sum() == 10801
    # the overflow value appears for the large multipliers
    assert (c[big] == 0x80000000).sum() > 5000


def test_seed_from_pipeline_seed():
    # rng_gsl_t(seed).uniform() * 2^24 through float (common/gpu.h:377, cux.cpp:410)
    for s in (1, 42, 22, 0):
        bg = np.random.MT19937()
        bg._legacy_seeding(4357 if s == 0 else s)
        u = np.float32(bg.random_raw(1)[0] / 4294967296.0)
        assert O.lib().gfo_seed_from_pipeline_seed(s) == int(np.float32(u * np.float32(1 << 24)))


def test_mwc_uniform_is_the_64bit_recurrence():
    st = mwc_state(7, 256)
    ns = 256
    n = 200
    out = np.zeros(n * ns, dtype=np.float32)
    a = st[:ns].astype(np.uint64)
    c = st[ns:2 * ns].astype(np.uint64)
    x = st[2 * ns:].astype(np.uint64)
    O.lib().gfo_mwc_sample(O.ptr(st), ns, 0, C.c_float(0), n, O.ptr(out))
    out = out.reshape(n, ns)
    for i in range(n):
        t = a * x + c
        c = t >> np.uint64(32)
        x = t & np.uint64(0xFFFFFFFF)
        ref = np.float32(2.32830643708e-10) * x.astype(np.float32)
        assert np.array_equal(out[i], ref)


@pytest.mark.parametrize("kind,param", [(1, 1.0), (2, 0.3), (2, 4.5), (2, 37.0), (2, 1500.0)])
def test_samplers_match_their_distributions(kind, param):
    ns, n = 1024, 200
    st = mwc_state(11, ns)
    dt = np.float32 if kind == 1 else np.uint32
    out = np.zeros(n * ns, dtype=dt)
    O.lib().gfo_mwc_sample(O.ptr(st), ns, kind, C.c_float(param), n, O.ptr(out))
    if kind == 1:
        assert stats.kstest(out, "norm").pvalue > 1e-3
    else:
        mu = param
        assert abs(out.mean() - mu) < 6 * np.sqrt(mu / out.size)
        assert abs(out.var() / mu - 1) < 0.05
        if mu < 50:
            k = np.arange(0, int(mu + 8 * np.sqrt(mu) + 5))
            obs = np.bincount(out.astype(np.int64), minlength=len(k))[: len(k)]
            exp = stats.poisson.pmf(k, mu) * out.size
            sel = exp > 20
            chi2 = ((obs[sel] - exp[sel]) ** 2 / exp[sel]).sum()
            assert stats.chi2.sf(chi2, sel.sum() - 1) > 1e-4


def _step_numpy(st, ns, n):
    """n steps of every stream by the 64-bit recurrence, vectorised."""
    a = st[:ns].astype(np.uint64)
    c = st[ns:2 * ns].astype(np.uint64)
    x = st[2 * ns:].astype(np.uint64)
    for _ in range(n):
        t = a * x + c
        c = t >> np.uint64(32)
        x = t & np.uint64(0xFFFFFFFF)
    return np.concatenate([a, c, x]).astype(np.uint32)


def test_mwc_jump_equals_stepping_for_every_multiplier():
    # all 65,536 seeded streams (multipliers below and above 2^31, carries
    # from the x86 overflow rule); jumps of n steps against n recurrence steps
    O.lib().gfo_mwc_jump.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64]
    ns = 65536
    st0 = mwc_state(2024, ns)
    cur, done = st0.copy(), 0
    for n in (1, 2, 3, 63, 64, 65, 700):
        ref = _step_numpy(cur, ns, n - done)
        got = st0.copy()
        O.lib().gfo_mwc_jump(O.ptr(got), ns, n)
        assert np.array_equal(got, ref), n
        cur, done = ref, n
    # a wrong modulus or exponent fails: one step fewer differs everywhere
    got = st0.copy()
    O.lib().gfo_mwc_jump(O.ptr(got), ns, 699)
    assert (got[ns:] != cur[ns:]).reshape(2, ns).any(axis=0).all()


def test_mwc_jump_long_distances_and_composition():
    O.lib().gfo_mwc_jump.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64]
    O.lib().gfo_mwc_step.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64]
    ns = 64
    st0 = mwc_state(99, 65536)
    sub = np.concatenate([st0[j * 65536 + 54720: j * 65536 + 54784] for j in range(3)])   # a > 2^31 among them
    assert (sub[:ns] > 2 ** 31).any() and (sub[:ns] < 2 ** 31).any()
    n = (1 << 20) + 12345
    a = sub.copy()
    O.lib().gfo_mwc_step(O.ptr(a), ns, n)
    b = sub.copy()
    O.lib().gfo_mwc_jump(O.ptr(b), ns, n)
    assert np.array_equal(a, b)
    # jumps compose: 2^40 + 2^33 + 5 in three jumps and in one
    c1 = sub.copy()
    for k in (1 << 40, 1 << 33, 5):
        O.lib().gfo_mwc_jump(O.ptr(c1), ns, k)
    c2 = sub.copy()
    O.lib().gfo_mwc_jump(O.ptr(c2), ns, (1 << 40) + (1 << 33) + 5)
    assert np.array_equal(c1, c2)


def test_product_host_jump_equals_oracle():
    # the library's Montgomery products (the device code compiled for the
    # host) against the oracle's 128-bit square-and-multiply
    import galfast_amd as GA
    L = GA.lib()
    L.gf_mwc_jump_host.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64]
    O.lib().gfo_mwc_jump.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64]
    ns = 65536
    st0 = mwc_state(5, ns)
    for n in (1, 64, 8 * 1000 + 3, (1 << 31) + 8 * 17, (7 << 32) + (1 << 31), (1 << 63) + 12345, 2 ** 64 - 1):
        a = st0.copy()
        b = st0.copy()
        assert L.gf_mwc_jump_host(O.ptr(a), ns, n) == 0
        O.lib().gfo_mwc_jump(O.ptr(b), ns, n)
        assert np.array_equal(a, b), n
