"""The plane-wave basis (apw/basis.py) against the reference's own PW_Basis / PW_Basis_K
unit tests, with their inputs: a 10 Angstrom cubic cell (lat0 = 1.8897261254578281 bohr,
latvec = 10 I), gridecut 10 Ry, and the k-points (0,0,0), (.1,.2,.3), (.4,.5,.6).

Each expected value is the one the reference test asserts:
  source/module_basis/module_pw/test_serial/pw_basis_k_test.cpp:55-80   (Initgrids1: nx = 20)
  source/module_basis/module_pw/test_serial/pw_basis_k_test.cpp:82-105  (Initgrids2: explicit box)
  source/module_basis/module_pw/test_serial/pw_basis_k_test.cpp:169-188 (SetupTransform: npw 3695)
  source/module_basis/module_pw/test_serial/pw_basis_k_test.cpp:190-212 (CollectLocalPW: npwk_max 2721)
  source/module_basis/module_pw/test_serial/pw_basis_test.cpp:189-210   (gamma-only half sphere:
                                                                         npwtot 1994, nstot 156)
"""
import math

import numpy as np
import pytest

from apw.basis import Lattice, PWBasisK, grid_from_ecut

LAT0 = 1.8897261254578281
KVEC_D = np.array([[0.0, 0.0, 0.0], [0.1, 0.2, 0.3], [0.4, 0.5, 0.6]])


@pytest.fixture(scope="module")
def lat():
    return Lattice(LAT0, 10.0 * np.eye(3))


def test_initgrids_from_cutoff(lat):
    assert lat.tpiba == pytest.approx(2.0 * math.pi / LAT0, rel=0, abs=0)
    assert lat.tpiba2 == lat.tpiba * lat.tpiba
    assert grid_from_ecut(lat, 10.0) == (20, 20, 20)
    b = PWBasisK(lat, 10.0, kvec_d=KVEC_D, gridecut=10.0)
    assert (b.nx, b.ny, b.nz) == (20, 20, 20)
    assert b.gridecut_lat == 10.0 / lat.tpiba2
    assert abs(b.gridecut_lat - 0.904561) < 1e-4
    assert np.allclose(lat.GT, np.linalg.inv(lat.latvec)) and np.allclose(lat.G, lat.GT.T)


def test_initgrids_explicit_box(lat):
    b = PWBasisK(lat, 10.0, 20, 20, 20)
    assert (b.nx, b.ny, b.nz) == (20, 20, 20)
    assert abs(b.gridecut_lat - 0.999999) < 1e-4
    assert abs(b.gridecut_lat * lat.tpiba2 - 11.0551) < 1e-4


@pytest.mark.parametrize("gk_ecut", [10.0, 11.0])
def test_big_sphere_and_k_spheres(lat, gk_ecut):
    b = PWBasisK(lat, gk_ecut, kvec_d=KVEC_D, gridecut=10.0)
    assert b.npw_big == 3695
    assert b.npwk_max == 2721
    # the per-k spheres sit inside the big one, in its (stick) order
    assert (b.npwk <= b.npw_big).all()
    for ik in range(3):
        n = b.npwk[ik]
        assert (b.gk2[ik, :n] <= b.gk_ecut).all()


def test_gamma_only_half_sphere_counts(lat):
    """pw_basis_test.cpp:189-210: with gamma_only and x' layout (fftnx = nx/2 + 1) the
    distributed sphere holds 1994 plane waves on 156 sticks: the x >= 0 half of the
    Gamma-point sphere of this basis."""
    b = PWBasisK(lat, 11.0, kvec_d=np.zeros((1, 3)), gridecut=10.0)
    half = b.f_big[b.f_big[:, 0] >= 0]
    assert half.shape[0] == 1994
    assert len({(int(x), int(y)) for x, y in half[:, :2]}) == 156
