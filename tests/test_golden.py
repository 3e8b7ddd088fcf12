"""The reference's own golden vectors: gal2other radec and vel2pm proper
motions of tests/os_vel2pm/result/sky.obs.txt, and the known answer
commented in modules/vel2pm_gpu.cu.h:223-231."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle as O
import gfx

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "os_vel2pm_sky.obs.txt")


def gold():
    return np.loadtxt(GOLD)


def test_oracle_gal2other_reproduces_golden_radec():
    d = gold()
    assert d.shape == (371, 16)
    ra, de = C.c_double(), C.c_double()
    err = []
    for r in d:
        O.lib().gfo_galequ(math.radians(r[0]), math.radians(r[1]), C.byref(ra), C.byref(de))
        err.append([math.degrees(ra.value) - r[8], math.degrees(de.value) - r[9]])
    # the golden file prints 8 decimals
    assert np.abs(err).max() < 1e-8


def test_oracle_vel2pm_reproduces_commented_known_answer():
    # X=8100, Y=100, Z=3000 (galactocentric), v=(10, 50, -30) km/s
    # => (mu_l, mu_b, v_r) = (12.72 mas/yr, -14.75 mas/yr, -27.34 km/s)
    p = gfx.vel2pm_params(0)
    X, Y, Z = 8100.0, 100.0, 3000.0
    l = math.atan2(-Y, 8000 - X)
    b = math.asin(Z / math.sqrt((8000 - X) ** 2 + Y * Y + Z * Z))
    pm = np.zeros(3, dtype=np.float32)
    O.lib().gfo_vcyl2pm(C.byref(p), l, b, 10.0, 50.0, -30.0, X, Y, Z, O.ptr(pm))
    assert pm[0] == pytest.approx(12.72, abs=0.005)
    assert pm[1] == pytest.approx(-14.75, abs=0.005)
    assert pm[2] == pytest.approx(-27.34, abs=0.005)


def test_oracle_vel2pm_reproduces_golden_pmlb():
    d = gold()
    p = gfx.vel2pm_params(0)
    pm = np.zeros(3, dtype=np.float32)
    err = []
    for r in d:
        lr = np.float32(0.017453292519943295769) * np.float32(r[0])
        br = np.float32(0.017453292519943295769) * np.float32(r[1])
        O.lib().gfo_vcyl2pm(C.byref(p), lr, br, r[10], r[11], r[12], r[5], r[6], r[7], O.ptr(pm))
        err.append(pm - r[13:16])
    err = np.abs(np.array(err))
    # pmlb is printed to 0.1; vcyl (the input) to 0.1 km/s and XYZ to 0.01 pc
    assert err[:, :2].max() < 0.06
    assert err[:, 2].max() < 0.15
    assert err.mean() < 0.04
