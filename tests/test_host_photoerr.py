"""gf_photoerr_records: the interval records the photometric-error kernel
reads instead of dividing out each interval's slope per star."""
import numpy as np

import galfast_amd as G
import gfx


def _records(koff, kx, ky):
    nb = len(koff) - 1
    koffa = np.array(koff, dtype=np.int32)
    kq = np.full(4 * max(1, len(kx)), np.nan)
    gx0, gs = np.full(18, np.nan), np.full(18, np.nan)
    rc = G.lib().gf_photoerr_records(nb, koffa.ctypes.data, kx.ctypes.data, ky.ctypes.data, kq.ctypes.data,
                                     gx0.ctypes.data, gs.ctypes.data)
    return rc, kq[:4 * len(kx)].reshape(-1, 4), gx0[:nb], gs[:nb]


def test_records_hold_the_quotients_the_evaluation_forms():
    rng = np.random.default_rng(3)
    kx0, ky0, _ = gfx.photoerr_knots(1)
    m1 = np.sort(rng.uniform(10, 30, 500))
    parts = [(kx0, ky0), (m1, 0.02 + np.exp(0.8 * (m1 - 24))), (np.array([21.5]), np.array([0.3])),
             (np.zeros(0), np.zeros(0))]
    kx = np.concatenate([p[0] for p in parts])
    ky = np.concatenate([p[1] for p in parts])
    koff = [0]
    for p in parts:
        koff.append(koff[-1] + len(p[0]))
    rc, q, gx0, gs = _records(koff, kx, ky)
    assert rc == 0
    assert np.array_equal(q[:, 0], kx) and np.array_equal(q[:, 1], ky)
    for b in range(len(parts)):
        k0, k1 = koff[b], koff[b + 1]
        if k1 == k0:
            assert gx0[b] == 0 and gs[b] == 0
            continue
        # every interval: IEEE double (y1 - y0) / (x1 - x0), bit for bit, and its right end
        slope = (ky[k0 + 1:k1] - ky[k0:k1 - 1]) / (kx[k0 + 1:k1] - kx[k0:k1 - 1])
        assert np.array_equal(q[k0:k1 - 1, 2], slope)
        assert np.array_equal(q[k0:k1 - 1, 3], kx[k0 + 1:k1])
        # the band's last knot, and the interval guess from its first knot and scale
        n = k1 - k0
        assert q[k1 - 1, 2] == 0 and q[k1 - 1, 3] == kx[k1 - 1]
        scale = (n - 1) / (kx[k1 - 1] - kx[k0]) if n > 1 else 0.0
        assert gs[b] == scale and gx0[b] == kx[k0]


def test_records_reject_bad_arguments():
    kx, ky, koff = gfx.photoerr_knots(2)
    assert _records(koff, kx, ky)[0] == 0
    assert _records([0, 5, 3], kx, ky)[0] != 0            # decreasing offsets
    assert _records([1, 5, 9], kx, ky)[0] != 0            # not starting at 0
    koffa = np.zeros(19, dtype=np.int32)
    g = np.zeros(18)
    assert G.lib().gf_photoerr_records(18, koffa.ctypes.data, None, None, None, g.ctypes.data,
                                       g.ctypes.data) != 0   # more than 17 bands
