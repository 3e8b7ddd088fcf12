"""gf_skygen_count_noalloc (the count with its scratch in a fixed device
array of the library, no allocation) against gf_skygen_count on the same
inputs: totals, per-pixel sums and (pixel, DM-bin) expectations identical,
on two calls in a row, the second with fewer pixels, in every accumulation
arithmetic; sizes beyond the fixed scratch are refused on the host."""
import ctypes as C

import numpy as np
import pytest

import galfast_amd as G
from galfast_amd import _abi as A
from devutil import dev_array, stream_ptr, torch
import gfx

pytestmark = pytest.mark.gpu


def sky(dx, b1):
    """Config-0 models over l 85-95, b 55-b1 in dx-degree pixels."""
    T = torch()
    pix = gfx.pixelize_rect(85, 95, 55, b1, dx)
    sp = gfx.skygen_params(15.0, 21.0, 0.05, 3.0, 15.0, len(pix), 7680)
    M, phi = gfx.seeded_lf()
    lf, tc = gfx.lf_texture(M, phi)
    lf_dev = dev_array(lf)
    models = []
    for m in gfx.c0_models(lf, tc):
        md = A.gf_model.from_buffer_copy(m)
        md.lf.data = lf_dev.data_ptr()
        models.append(md)
    pix_dev = T.from_numpy(np.frombuffer(bytes(pix), dtype=np.uint8).copy()).cuda()
    return sp, models, pix_dev, lf_dev


def run(fn, sp, model, pix_dev, mode, noalloc):
    T = torch()
    K = sp.nm + sp.nM - 1
    per = T.full((sp.npixels,), -1.0, dtype=T.float64, device="cuda")
    w = T.full((sp.npixels * K,), -1.0, dtype=T.float64, device="cuda")
    L = G.lib()
    if noalloc:
        L.gf_skygen_count_noalloc.argtypes = [C.c_void_p] * 4 + [C.c_int] + [C.c_void_p] * 4
        res_dev = T.full((4,), -1.0, dtype=T.float64, device="cuda")
        G.check(L.gf_skygen_count_noalloc(C.byref(sp), C.byref(model), C.c_void_p(pix_dev.data_ptr()), None, mode,
                                          C.c_void_p(res_dev.data_ptr()), C.c_void_p(per.data_ptr()),
                                          C.c_void_p(w.data_ptr()), stream_ptr()), "gf_skygen_count_noalloc")
        T.cuda.synchronize()
        res = res_dev.cpu().numpy()
    else:
        r = A.gf_count_result()
        G.check(L.gf_skygen_count(C.byref(sp), C.byref(model), C.c_void_p(pix_dev.data_ptr()), None, mode,
                                  C.byref(r), C.c_void_p(per.data_ptr()), C.c_void_p(w.data_ptr()), stream_ptr()),
                "gf_skygen_count")
        T.cuda.synchronize()
        res = np.array([r.nstarsExpectedToGenerate, r.nstarsExpected, r.maxCount, r.ncells])
    return res, per.cpu().numpy(), w.cpu().numpy()


@pytest.mark.parametrize("mode", [A.GF_ACC_FP64, A.GF_ACC_KAHAN32, A.GF_ACC_FP32])
def test_noalloc_count_equals_count_twice_smaller_second(mode):
    G.check(G.lib().gf_init(0), "gf_init")
    big = sky(0.5, 65)
    small = sky(1.0, 58)
    assert small[0].npixels < big[0].npixels
    for sp, models, pix_dev, _ in (big, small):
        for m in models:
            a = run(None, sp, m, pix_dev, mode, noalloc=False)
            b = run(None, sp, m, pix_dev, mode, noalloc=True)
            assert a[0][0] > 0
            assert np.array_equal(a[0], b[0])
            assert np.array_equal(a[1], b[1])
            assert np.array_equal(a[2], b[2])


def test_noalloc_count_refuses_sizes_beyond_its_scratch():
    sp, models, pix_dev, _ = sky(1.0, 58)
    L = G.lib()
    L.gf_skygen_count_noalloc.argtypes = [C.c_void_p] * 4 + [C.c_int] + [C.c_void_p] * 4
    res = torch().zeros(4, dtype=torch().float64, device="cuda")
    too_many = A.gf_skygen_params.from_buffer_copy(sp)
    too_many.npixels = 262145
    assert L.gf_skygen_count_noalloc(C.byref(too_many), C.byref(models[0]), C.c_void_p(pix_dev.data_ptr()), None,
                                     0, C.c_void_p(res.data_ptr()), None, None, stream_ptr()) == -9
    too_fine = A.gf_skygen_params.from_buffer_copy(sp)
    too_fine.nM = 4097
    assert L.gf_skygen_count_noalloc(C.byref(too_fine), C.byref(models[0]), C.c_void_p(pix_dev.data_ptr()), None,
                                     0, C.c_void_p(res.data_ptr()), None, None, stream_ptr()) == -9
