"""Multiply-with-carry streams on the GPU, bit-exact against the oracle."""
import ctypes as C

import numpy as np
import pytest

import galfast_amd as G
import oracle as O
from devutil import DevRng, stream_ptr, torch, sync
from gfx import mwc_state

pytestmark = pytest.mark.gpu


def _run(kind, param, n, seed, ns=65536, chunks=1):
    T = torch()
    host = mwc_state(seed, ns)
    dev = DevRng(host)
    dt = np.float32 if kind <= 1 else np.uint32
    for c in range(chunks):
        out = T.zeros(n * ns, dtype=T.int32, device="cuda")
        G.check(G.lib().gf_mwc_sample(C.byref(dev.rng), kind, C.c_float(param), n, C.c_void_p(out.data_ptr()),
                                      stream_ptr()), "gf_mwc_sample")
        ref = np.zeros(n * ns, dtype=dt)
        O.lib().gfo_mwc_sample(O.ptr(host), ns, kind, C.c_float(param), n, O.ptr(ref))
        got = out.cpu().numpy().view(dt)
        assert np.array_equal(got, ref), "chunk %d: %d mismatches" % (c, (got != ref).sum())
    assert np.array_equal(dev.state(), host)


def test_uniform_all_streams_1e4_draws():
    _run(0, 0.0, 1000, seed=42, chunks=10)


def test_uniform_seed_with_large_multipliers():
    _run(0, 0.0, 64, seed=3, ns=65536)


@pytest.mark.parametrize("sigma", [1.0, 0.3])
def test_gaussian(sigma):
    _run(1, sigma, 100, seed=5)


@pytest.mark.parametrize("mu", [0.0, 0.02, 0.7, 6.5, 10.5, 37.0, 2500.0])
def test_poisson(mu):
    _run(2, mu, 20, seed=9, ns=16384)


def test_poisson_sweep_of_means():
    # stream s draws with mean 0.01*(1+s): covers the product loop, the
    # gamma_int/gamma_large split and the binomial tail
    _run(3, 0.01, 4, seed=13, ns=65536)


def test_device_jump_equals_oracle():
    # gf_mwc_jump (device Montgomery products) against the oracle's 128-bit
    # square-and-multiply, for every one of 65,536 seeded streams
    O.lib().gfo_mwc_jump.argtypes = [C.c_void_p, C.c_uint32, C.c_uint64]
    L = G.lib()
    L.gf_mwc_jump.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
    st = mwc_state(31, 65536)
    d = DevRng(st)
    ref = st.copy()
    for n in (1, 1000, (1 << 33) + 17, (1 << 62) + 3):
        G.check(L.gf_mwc_jump(C.byref(d.rng), n, stream_ptr()), "gf_mwc_jump")
        O.lib().gfo_mwc_jump(O.ptr(ref), 65536, n)
        assert np.array_equal(d.state(), ref), n


@pytest.mark.parametrize("seed,ns", [(0, 64), (1234, 4096), (16777215, 65536), (2 ** 32 - 1, 70720)])
def test_device_seeding_equals_host_seeding(seed, ns):
    """gf_mwc_seed_device (every stream's drand48/lrand48 pair reached by
    jump-ahead, a lane per stream) writes gf_mwc_seed_host's tables and the
    oracle's, including the multipliers above 2^31 (their carries overflow
    to INT32_MIN) at the end of the prime list."""
    from gfx import PRIMES
    T = torch()
    host = np.zeros(3 * ns, dtype=np.uint32)
    assert G.lib().gf_mwc_seed_host(seed, PRIMES.ctypes.data_as(C.POINTER(C.c_uint32)), len(PRIMES), ns,
                                    host.ctypes.data_as(C.POINTER(C.c_uint32))) == ns
    assert np.array_equal(host, mwc_state(seed, ns))
    primes = T.from_numpy(PRIMES.view(np.int32).copy()).cuda()
    out = T.zeros(3 * ns, dtype=T.int32, device="cuda")
    G.check(G.lib().gf_mwc_seed_device(seed, primes.data_ptr(), len(PRIMES), ns, out.data_ptr(), stream_ptr()),
            "gf_mwc_seed_device")
    sync()
    assert np.array_equal(out.cpu().numpy().view(np.uint32), host)


def test_device_seeding_rejects_bad_counts():
    from gfx import PRIMES
    T = torch()
    primes = T.from_numpy(PRIMES.view(np.int32).copy()).cuda()
    out = T.zeros(3 * 128, dtype=T.int32, device="cuda")
    L = G.lib()
    assert L.gf_mwc_seed_device(1, primes.data_ptr(), len(PRIMES), 100, out.data_ptr(), stream_ptr()) != 0
    assert L.gf_mwc_seed_device(1, primes.data_ptr(), 64, 128, out.data_ptr(), stream_ptr()) != 0
    assert L.gf_mwc_seed_device(1, None, len(PRIMES), 64, out.data_ptr(), stream_ptr()) != 0
