"""The ADVECTION case service (casTest/ADVECTION/ADVECTIONService.cc) in the oracle.

The reference holds no dataset or output for this service, so these checks pin it by what
the mathematics fixes rather than by fixtures:
* the initial square (ADVECTIONService.cc:11-65, 67-130, 146-231) against its closed form
  on the cell centres, including the linear-density option and the rotation's centre;
* the prescribed velocity fields (:131-144, 233-246): translations are constants, the
  rotation is omega (-(y - 0.5), x - 0.5) with omega = 4 Pi;
* the reverse option scales the field by cos(Pi t parametre) (MahycoModule.cc:702-720);
* order-1 remap at Courant number 1 along x, y or both is a shift by one cell per step: the
  swept volume of each face is the whole upwind cell.  (At exactly 1 the last interior node
  row can land an ulp beyond the fixed boundary nodes, a negative volume that stops the
  reference too (MahycoModule.cc:1164), so the tests run at 1 - 2^-30.)"""
import numpy as np
import pytest

from mahyco_amd import cases
from mahyco_amd import types as T

PI = 3.14159265359  # TypesMahyco.h:42


def _oracle(case):
    from oracle.pyoracle import OracleCase
    return OracleCase(case)


def _centres(n):
    x = (np.arange(n) / n + (np.arange(n) + 1) / n) * .5
    return np.meshgrid(x, x, indexing="xy")


@pytest.mark.parametrize("test", [7, 10, 18, 21])
def test_initial_square(test):
    n = 40
    o = _oracle(cases.advection("adv", test, 1., n=n))
    rho = o.get(T.MH_F_DENSITY).reshape(n, n)
    xc, yc = _centres(n)
    yb = .5 if test in (7, 18) else .75
    inside = (np.abs(xc - .5) < .25) & (np.abs(yc - yb) < .25)
    assert inside.sum() > 0 and (~inside).sum() > 0
    np.testing.assert_array_equal(rho, np.where(inside, 1., 0.))
    mask = o.get(T.MH_F_ENV_MASK).reshape(n, n)
    np.testing.assert_array_equal(mask, np.where(inside, 2, 1) if test >= 18 else np.ones((n, n)))
    np.testing.assert_array_equal(o.get(T.MH_F_INTERNAL_ENERGY), 0.)


def test_linear_density():
    n = 40
    c = cases.advection("adv", 18, 1., n=n, lineaire=True)
    c.envs[1].densite_initiale = 2.
    o = _oracle(c)
    rho = o.get(T.MH_F_DENSITY).reshape(n, n)
    xc, yc = _centres(n)
    inside = (np.abs(xc - .5) < .25) & (np.abs(yc - .5) < .25)
    r = np.hypot(xc - .5, yc - .5)
    np.testing.assert_allclose(rho[inside], (2. / .25 * r)[inside], rtol=1e-14, atol=0)
    assert np.all(rho[~inside] == 0.)


@pytest.mark.parametrize("test,ux,uy", [(7, 1., 0.), (8, 0., 1.), (9, 1., 1.), (18, 1., 0.), (20, 1., 1.)])
def test_translation_fields(test, ux, uy):
    o = _oracle(cases.advection("adv", test, 1., n=20))
    np.testing.assert_array_equal(o.get(T.MH_N_VELOCITY, -1, 0), ux)
    np.testing.assert_array_equal(o.get(T.MH_N_VELOCITY, -1, 1), uy)


@pytest.mark.parametrize("test", [10, 21])
def test_rotation_field(test):
    o = _oracle(cases.advection("adv", test, 1., n=30))
    x = o.get(T.MH_N_COORD, -1, 0)
    y = o.get(T.MH_N_COORD, -1, 1)
    w = 4. * PI
    np.testing.assert_allclose(o.get(T.MH_N_VELOCITY, -1, 0), -w * (y - .5), rtol=0, atol=1e-14)
    np.testing.assert_allclose(o.get(T.MH_N_VELOCITY, -1, 1), w * (x - .5), rtol=0, atol=1e-14)


def test_reverse_option_scales_the_field():
    dt, steps, par = .01, 7, .8
    o = _oracle(cases.advection("adv", 9, 1., n=20, dt=dt, reverse=par))
    o.run(steps)
    f = np.cos(PI * o.time * par)
    np.testing.assert_allclose(o.get(T.MH_N_VELOCITY, -1, 0), f, rtol=1e-14)
    np.testing.assert_allclose(o.get(T.MH_N_VELOCITY, -1, 1), f, rtol=1e-14)


@pytest.mark.parametrize("test,sx,sy", [(7, 1, 0), (8, 0, 1), (9, 1, 1)])
def test_courant_one_is_a_shift(test, sx, sy):
    n, k = 50, 5
    o = _oracle(cases.advection("adv", test, 1., n=n, dt=(1. - 2. ** -30) / n, ordre=1))
    r0 = o.get(T.MH_F_DENSITY).reshape(n, n).copy()
    o.run(k)
    assert o.iteration == k
    r = o.get(T.MH_F_DENSITY).reshape(n, n)
    shifted = np.zeros_like(r0)
    shifted[k * sy:, k * sx:] = r0[:n - k * sy, :n - k * sx]
    np.testing.assert_allclose(r, shifted, rtol=0, atol=1e-8)


@pytest.mark.parametrize("test,sx,sy", [(18, 1, 0), (19, 0, 1)])
def test_two_environments_courant_one(test, sx, sy):
    """The square in its own environment moves as in one environment, but the empty
    environment (density 0) never re-enters the cells it left: an environment is added only
    with a nonzero mass (RemapADIFinal.cc:47-58), so those cells keep the square's
    environment alone, at its density.  (Along the diagonal the x sweep also carries the
    square into cells that the y sweep then leaves, which keep it the same way.)"""
    n, k = 50, 5
    o = _oracle(cases.advection("adv", test, 1., n=n, dt=(1. - 2. ** -30) / n, ordre=1))
    r0 = o.get(T.MH_F_DENSITY).reshape(n, n).copy()
    o.run(k)
    assert o.iteration == k
    r = o.get(T.MH_F_DENSITY).reshape(n, n)
    mask = o.get(T.MH_F_ENV_MASK).reshape(n, n)
    shifted = np.zeros_like(r0)
    shifted[k * sy:, k * sx:] = r0[:n - k * sy, :n - k * sx]
    left = (r0 == 1.) & (shifted == 0.)
    ahead = (r0 == 0.) & (shifted == 0.)
    assert left.sum() > 0
    np.testing.assert_allclose(r[shifted == 1.], 1., rtol=0, atol=1e-8)
    np.testing.assert_allclose(r[ahead], 0., rtol=0, atol=1e-8)
    assert np.all(mask[left] == 2)
    np.testing.assert_allclose(r[left], 1., rtol=0, atol=1e-8)


def test_bubble_cases_need_two_environments():
    """The RIDER and ADVECTION services give the bubble environment 1's density (also in
    their Mono tests), so a single environment is refused."""
    for c in (cases.advection("adv", 7, 1.), cases.rider("r", 23, 1.)):
        c.envs = c.envs[:1]
        with pytest.raises(Exception):
            _oracle(c)
