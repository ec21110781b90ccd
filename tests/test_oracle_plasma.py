"""Pins of the oracle's plasma optics and PlasmaSample restatement.

Reference-held numbers: delta_free = 1.077221e-5 at n_e = 1e24 cm^-3, 8 keV
(tests/plasma_sample/TEST_REPORT.md:76) and the Kramers beta_ff of Al at
n_e = 1e23 cm^-3, T_e = 100 eV, 8 keV: 7.61e-12 fully ionised, 4.68e-12 at
Z* = 8 (TEST_REPORT.md:100-106).  The rest is fixed by the physics: the
scaling laws of each term, the clipping of q, vacuum, linearity in the
Chantler factors, the modulo-2pi phases of a uniform slab by the reference's
gold-standard ratio method (test_plasma_phase_absorption.py:60-101), and
equality with the material-id Sample path on the same per-voxel values.
"""

import importlib

import numpy as np
import pytest

import oracle as O

ORACLE_MOD = importlib.import_module("oracle.ravesim_oracle")


def db(ne, ni, te, zs, optics=(0.0, 0.0, 13.0, 169.0), energy=8000.0):
    d, b = O.plasma_delta_beta_grid_multi(np.atleast_1d(ne), np.atleast_1d(ni), np.atleast_1d(te),
                                          np.atleast_1d(zs), *optics, energy)
    return d[0], b[0]


def test_delta_free_golden():
    d, _ = db(1e24, 1e24, 100.0, 1.0, optics=(0.0, 0.0, 1.0, 1.0))
    assert d == pytest.approx(1.077221e-5, rel=5e-7)


@pytest.mark.parametrize("zstar,expected", [(13.0, 7.61e-12), (8.0, 4.68e-12)])
def test_kramers_beta_ff_golden(zstar, expected):
    # f2eff_bar = 0: only the free-free term is left in beta
    _, b = db(1e23, 1e23 / zstar, 100.0, zstar, optics=(0.0, 0.0, 13.0, 169.0))
    assert b == pytest.approx(expected, rel=2e-3)


def test_kramers_scaling_laws():
    base = db(1e23, 5e21, 100.0, 8.0)[1]
    assert db(2e23, 5e21, 100.0, 8.0)[1] == pytest.approx(2 * base, rel=1e-14)   # n_e
    assert db(1e23, 1e22, 100.0, 8.0)[1] == pytest.approx(2 * base, rel=1e-14)   # n_i
    assert db(1e23, 5e21, 400.0, 8.0)[1] == pytest.approx(base / 2, rel=1e-14)   # T_e^-1/2
    assert db(1e23, 5e21, 100.0, 4.0)[1] == pytest.approx(base / 4, rel=1e-14)   # Z*^2
    # nu^-3 lambda = E^-4 for the free-free term; lambda^2 for delta_free
    b2 = db(1e23, 5e21, 100.0, 8.0, energy=16000.0)[1]
    assert b2 == pytest.approx(base / 16, rel=1e-12)
    d1 = db(1e23, 0.0, 100.0, 8.0)[0]
    d2 = db(1e23, 0.0, 100.0, 8.0, energy=16000.0)[0]
    assert d2 == pytest.approx(d1 / 4, rel=1e-12)


def test_ionisation_clipping_vacuum_and_bound_linearity():
    opt = (13.2, 0.37, 13.0, 169.0)
    free = db(1e23, 1e22, 100.0, 13.0, optics=(0.0, 0.0, 13.0, 169.0))
    # Z* >= zbar: no bound electrons left, the Chantler factors drop out
    assert db(1e23, 1e22, 100.0, 20.0, optics=opt)[0] == free[0]
    assert db(1e23, 1e22, 100.0, 13.0, optics=opt) == free
    # Z* <= 0: q = 0, no free-free absorption, all electrons bound
    d0, b0 = db(1e23, 1e22, 100.0, -1.0, optics=opt)
    lamb = 1e-8 * O.PLANCK_HC / 8000.0
    P = O.R_ELECTRON_CM * lamb * lamb / (2 * np.pi)
    assert d0 == pytest.approx(1e23 * P + 1e22 * P * 13.2, rel=1e-14)
    assert b0 == pytest.approx(1e22 * P * 0.37, rel=1e-14)
    # vacuum
    assert db(0.0, 0.0, 0.0, 0.0, optics=opt) == (0.0, 0.0)
    # bound terms linear in f1bar / f2eff_bar with ratio f2eff_bar / f1bar
    a = db(1e23, 1e22, 100.0, 5.0, optics=(0.0, 0.0, 13.0, 169.0))
    b = db(1e23, 1e22, 100.0, 5.0, optics=(1.0, 1.0, 13.0, 169.0))
    c = db(1e23, 1e22, 100.0, 5.0, optics=(2.0, 2.0, 13.0, 169.0))
    assert c[0] - a[0] == pytest.approx(2 * (b[0] - a[0]), rel=1e-12)
    assert (b[1] - a[1]) == pytest.approx(b[0] - a[0], rel=1e-12)


def _uniform_spec(nz, ny_g, nx_g, ne, zs, pz, px, optics=(0.0, 0.0, 1.0, 1.0)):
    shape = (nz, ny_g, nx_g)
    return O.PlasmaSpec(ne=np.full(shape, ne), ni=np.full(shape, ne / zs), te=np.full(shape, 100.0),
                        zstar=np.full(shape, zs), optics=optics, pixel_size_x=px,
                        pixel_size_z=pz)


@pytest.mark.parametrize("target", [np.pi / 2, np.pi, 2 * np.pi])
def test_uniform_slab_phase_by_ratio_method(target):
    """Fully ionised H slab covering the field: T = u_plasma / u_empty is the
    material factor everywhere (propagation is linear), with |phase| the
    target modulo 2 pi and |T| = exp(-2 pi beta t / lambda)."""
    n, dx, wl = 64, 1e-6, O.convert_energy_wavelength(8000.0)
    ne_cm = 1e24
    t = target / (ne_cm * 1e6 * 2.8179403262e-15 * wl)   # phase = n_e r_e lambda t
    spec = _uniform_spec(1, n + 8, n + 8, ne_cm, 1.0, t, dx)
    rng = np.random.default_rng(3)
    u0 = np.exp(1j * rng.uniform(0, 1, (n, n)))
    cut = 0.3 / dx
    up, _ = O.plasma_sample_apply_2d(u0, spec, dx, dx, wl, cut)
    ue, _ = O.propagate_2d(u0, dx, dx, wl, t, cut)
    T = up / ue
    delta, beta = O.plasma_delta_beta_grid_multi(ne_cm, ne_cm, 100.0, 1.0, *spec.optics,
                                                 O.convert_energy_wavelength(wl))
    # the slab factor exp(-2 pi i t/lambda (delta - i beta)) at every pixel
    # (the grid covers the field with margin)
    expect = np.exp(-2j * np.pi * t / wl * (delta - 1j * beta))
    assert np.allclose(T, expect, rtol=1e-12, atol=0)
    # |phase| is the target modulo 2 pi (to the 3e-8 offset between
    # big-wave's wavelength constants and nist_lookup's PLANCK_HC)
    assert np.allclose(np.abs(np.angle(T)), abs(np.angle(np.exp(1j * target))), atol=1e-6)
    assert np.allclose(np.abs(T), np.exp(-2 * np.pi * beta * t / wl), rtol=1e-12)


def test_plasma_apply_equals_material_sample_on_same_values():
    """Voxels in three plasma states == a material-id volume whose table
    holds those states' (delta, beta): identical arithmetic after the lookup."""
    rng = np.random.default_rng(8)
    nz, ny_g, nx_g, n = 3, 40, 44, 32
    states = np.array([[0.0, 0.0, 0.0, 0.0], [1e23, 2e22, 150.0, 5.0], [4e22, 4e22, 30.0, 1.0]])
    ids = rng.integers(0, 3, (nz, ny_g, nx_g))
    g = states[ids]
    optics = (13.3, 0.21, 13.0, 169.0)
    wl = O.convert_energy_wavelength(9000.0)
    energy = O.convert_energy_wavelength(wl)
    spec = O.PlasmaSpec(ne=g[..., 0], ni=g[..., 1], te=g[..., 2], zstar=g[..., 3], optics=optics,
                        pixel_size_x=0.9e-6, pixel_size_z=0.4e-6, pixel_size_y=1.1e-6,
                        x_pos=0.3e-6, y_pos=-0.2e-6)
    d, b = O.plasma_delta_beta_grid_multi(states[:, 0], states[:, 1], states[:, 2], states[:, 3],
                                          *optics, energy)
    sample = O.SampleSpec(grid=ids.astype(np.uint32), db_list=d + 1j * b, pixel_size_x=0.9e-6,
                          pixel_size_y=1.1e-6, pixel_size_z=0.4e-6, x_pos=0.3e-6, y_pos=-0.2e-6)
    u0 = np.exp(1j * rng.uniform(0, 1, (n, n)))
    up, Up = O.plasma_sample_apply_2d(u0, spec, 1e-6, 1e-6, wl, 0.3e6)
    us, Us = O.sample_apply_2d(u0, sample, 1e-6, 1e-6, wl, 0.3e6)
    assert O.relative_l2(up, us) < 1e-15
    assert O.relative_l2(Up, Us) < 1e-15
