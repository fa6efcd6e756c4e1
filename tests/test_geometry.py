"""Geometry of TLTrapSim.__init__ (tl_trap_lab.py:33-39) and Box.seed (engine.py:124-129),
pinned by closed forms and the values worked out from the reference's configs."""
import math

import numpy as np
import pytest

from mcluminescence_amd import _abi
from mcluminescence_amd import workloads as W
from mcluminescence_amd._abi import RunParams
from oracle import orc


def _tl12(holes=2000.0, N_e=2000.0):
    return RunParams.make(alpha=1.6e9, b=1e12, s=1e12, E_cb=1000.0, E_loc_1=1.2, E_loc_2=1.2,
                          D0=400.35, Retrap=0.5, k_b=8.617343e-5, rho_prime=3e-4, holes=holes,
                          boundary_factor=1.2, N_e=N_e, e_ratio_start=1.0, T_start=0.0,
                          T_rate=1.0, dt=1.0, n_steps=800, sims=2)


def test_tl12_geometry_closed_form():
    g = orc.geometry(_tl12())
    rho = 3e-4 * 3 / (4 * math.pi) * 1.6e9 ** 3
    assert g.rho == pytest.approx(2.934e23, rel=1e-3) and g.rho == pytest.approx(rho, rel=1e-15)
    assert g.side_m == pytest.approx((2000 / rho) ** (1 / 3), rel=1e-15)
    assert g.side_m == pytest.approx(1.896e-7, rel=1e-3)
    assert (g.n_e0, g.n_h) == (2000, 3455)       # int(2000 * 1.2**3) = 3455


@pytest.mark.parametrize("n,expected", [(500, 863), (2000, 3455), (2048, 3538), (400, 691)])
def test_boundary_hole_truncation(n, expected):
    r = W.single_trap(1.3, 1e12, 4e9, 1e23, n, T_start_K=300, T_rate=1, dt=0.1, n_steps=10,
                      sims=1)
    assert orc.geometry(r).n_h == expected == int(n * 1.2 ** 3)


def test_glow_workload_side_and_density():
    r = W.tl_glow_16x500()[0]
    g = orc.geometry(r)
    assert r.rho_prime == pytest.approx(6.545e-6, rel=1e-3)
    assert g.rho == pytest.approx(1e23, rel=1e-12)
    assert g.side_m == pytest.approx(1.710e-7, rel=1e-3)


def test_library_geometry_equals_oracle():
    for r in (_tl12(), W.tl_glow_16x500()[0], W.iso_decay_4096x400()[0]):
        a, b = _abi.geometry(r), orc.geometry(r)
        assert (a.rho, a.side_m, a.n_e0, a.n_h) == (b.rho, b.side_m, b.n_e0, b.n_h)


def test_coordinates_inside_boxes_and_uniform():
    r = W.tl_glow_16x500()[0]
    g = orc.geometry(r)
    st = orc.init_state(r, 0, 0)
    side_nm = g.side_m * 1e9
    e, h = st["e_xyz"], st["h_xyz"]
    assert e.shape == (500, 3) and h.shape == (863, 3)
    assert (e >= 0).all() and (e < np.float32(side_nm) * (1 + 1e-6)).all()
    assert (h >= 0).all() and (h < np.float32(side_nm * 1.2) * (1 + 1e-6)).all()
    assert abs(h.mean() / (side_nm * 1.2) - 0.5) < 0.03


def test_initial_nearest_hole_is_bruteforce_argmin():
    r = W.tl_glow_16x500()[0]
    st = orc.init_state(r, 2, 5)
    # independent float64 check, tolerant only to float32 near-ties
    d = np.linalg.norm(st["e_xyz"][:, None, :].astype(np.float64) -
                       st["h_xyz"][None, :, :].astype(np.float64), axis=2)
    nn64 = d.argmin(axis=1)
    dmin = d.min(axis=1)
    same = nn64 == st["nn"]
    assert same.mean() > 0.99
    # where they differ the two holes are equidistant to float32 precision
    for e in np.nonzero(~same)[0]:
        assert abs(d[e, st["nn"][e]] - dmin[e]) <= 1e-6 * dmin[e]
    alpha_nm = 4e9 * 1e-9
    ar = alpha_nm * d[np.arange(500), st["nn"]]
    flushed = ar > 87.0                  # exp below FLT_MIN is flushed to 0 by design
    assert (st["w"][flushed] == 0).all()
    np.testing.assert_allclose(st["w"][~flushed], np.exp(-ar[~flushed]), rtol=2e-5)
