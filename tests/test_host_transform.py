"""The library's coordinate transforms (load_transform,
common/transform.cpp:345-388, with get_gal2xxx_matrix :70-148,
xxx2galxyz :150-224 and get_rotation_matrix :226-343), host code that builds
every model's and module's (T, M), against constructions made here from the
geometry: the Galactic-plane frame from the Sun's position, the equatorial
frame from the oracle's Galactic -> equatorial conversion (itself pinned
against the reference's golden (ra, dec) columns), axis-angle and Euler
rotations from scipy."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import galfast_amd as GA
import oracle as O
from test_oracle_pins import galplane_geometric, unit

RG, Z0 = 8000.0, 25.0


def transform(center, orientation, Rg=RG, z0=Z0):
    L = GA.lib()
    L.gf_transform_host.argtypes = [C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    T = np.zeros(3, np.float32)
    M = np.zeros(9, np.float32)
    rc = L.gf_transform_host(center.encode(), orientation.encode(), Rg, z0, T.ctypes.data, M.ctypes.data)
    return rc, T.astype(np.float64), M.reshape(3, 3).astype(np.float64)


def gal2equ_matrix():
    """Galactic -> equatorial cartesian, column by column from the oracle's
    (l, b) -> (ra, dec)."""
    lib = O.lib()
    cols = []
    for l, b in [(0.0, 0.0), (math.pi / 2, 0.0), (0.0, math.pi / 2)]:
        ra, dec = C.c_double(), C.c_double()
        lib.gfo_galequ(l, b, C.byref(ra), C.byref(dec))
        cols.append(unit(ra.value, dec.value))
    return np.column_stack(cols)


def test_galplane_from_the_geometry():
    rc, T, M = transform("galplane", "galplane")
    assert rc == 0
    Tg, Mg = galplane_geometric(RG, Z0)
    assert np.abs(T - Tg).max() < 1e-3 and np.abs(M - Mg).max() < 1e-7
    # the Sun sits z0 above the plane at distance Rg from the centre
    sun = M @ (np.zeros(3) + T)
    assert abs(np.linalg.norm(sun) - RG) < 1e-3 and abs(sun[2] - Z0) < 1e-3 and abs(sun[1]) < 1e-6
    assert np.abs(M @ M.T - np.eye(3)).max() < 1e-6 and np.linalg.det(M) > 0
    # planted mistake: the Sun below the plane
    Tb, Mb = galplane_geometric(RG, -Z0)
    assert np.abs(M - Mb).max() > 1e-3


def test_centres():
    # an empty type is the system's origin
    assert np.allclose(transform("gal", "")[1], 0.0)
    assert np.allclose(transform("galcentric", "")[1], [-RG, 0, 0])
    # gal sph: (l, b, D) in degrees and parsecs
    rc, T, M = transform("gal sph 30 60 2500", "")
    assert rc == 0 and np.allclose(-T, 2500 * unit(math.radians(30), math.radians(60)), atol=1e-3)
    assert np.allclose(M, np.eye(3))
    # galcentric xyz: x from the centre away from the Sun, y against the Galactic y
    _, T, _ = transform("galcentric xyz 100 200 300", "")
    assert np.allclose(-T, [RG - 100, -200, 300], atol=1e-3)
    # galplane xyz: the plane frame's point, back through the plane frame's own matrix
    v = np.array([1500.0, -700.0, 420.0])
    _, T, M = transform("galplane xyz 1500 -700 420", "galplane")
    gc = np.array([RG, 0.0, 0.0])
    assert np.allclose(M @ (-T - gc), v, atol=2e-3)
    # equ sph: the direction of (ra, dec) through the equatorial frame
    _, T, _ = transform("equ sph 150 -20 1000", "")
    E = gal2equ_matrix()
    assert np.allclose(E @ (-T), 1000 * unit(math.radians(150), math.radians(-20)), atol=2e-3)


def test_equatorial_orientation():
    rc, T, M = transform("gal", "equ")
    assert rc == 0
    E = gal2equ_matrix()
    assert np.abs(M - E).max() < 1e-6
    assert np.abs(M @ M.T - np.eye(3)).max() < 1e-6 and np.linalg.det(M) > 0
    # the north Galactic pole at (ra, dec) = (192.8595, 27.1283) deg, the
    # Galactic centre at (266.405, -28.936) deg (J2000)
    ngp = M @ np.array([0.0, 0.0, 1.0])
    assert np.allclose(ngp, unit(math.radians(192.85948), math.radians(27.12825)), atol=2e-5)
    gc = M @ np.array([1.0, 0.0, 0.0])
    assert np.allclose(gc, unit(math.radians(266.40500), math.radians(-28.93617)), atol=5e-5)


@pytest.mark.parametrize("axis,angle", [((0, 0, 1), 30.0), ((1, 2, 3), -75.0), ((0.3, -1, 0.2), 160.0)])
def test_axisangle_is_the_frame_rotated_about_the_axis(axis, angle):
    rc, _, M = transform("gal", "gal axisangle %g %g %g %g" % (*axis, angle))
    assert rc == 0
    ax = np.array(axis, dtype=np.float64)
    ax /= np.linalg.norm(ax)
    # the target frame is the origin frame turned by +angle about the axis:
    # point coordinates turn by -angle
    R = Rotation.from_rotvec(math.radians(angle) * ax).as_matrix()
    assert np.abs(M - R.T).max() < 1e-6
    assert np.abs(M - R).max() > 1e-2            # planted: active for passive
    # composed after the starting system's rotation
    rc, _, Mp = transform("gal", "galplane axisangle %g %g %g %g" % (*axis, angle))
    assert np.abs(Mp - R.T @ galplane_geometric(RG, Z0)[1]).max() < 1e-6


@pytest.mark.parametrize("phi,theta,psi", [(30.0, 0.0, 0.0), (10.0, 20.0, 30.0), (-120.0, 75.0, 200.0)])
def test_euler_zxz(phi, theta, psi):
    rc, _, M = transform("gal", "equ euler %g %g %g" % (phi, theta, psi))
    assert rc == 0
    # ccw turns of the frame about z by phi, the new x by theta, the new z by psi
    R = Rotation.from_euler("ZXZ", [phi, theta, psi], degrees=True).as_matrix()
    assert np.abs(M - R.T @ gal2equ_matrix()).max() < 1e-6
    if theta != 0.0:
        Rb = Rotation.from_euler("ZYZ", [phi, theta, psi], degrees=True).as_matrix()
        assert np.abs(M - Rb.T @ gal2equ_matrix()).max() > 1e-2     # planted: y for x


def test_errors():
    for c, o in [("nowhere", ""), ("gal", "nowhere"), ("gal", "gal spin 1 2 3"), ("gal", "gal axisangle 1 2"),
                 ("gal pol 1 2 3", "")]:
        rc = transform(c, o)[0]
        assert rc != 0, (c, o)
