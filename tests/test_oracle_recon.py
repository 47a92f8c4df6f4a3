"""Pins the oracle's reconstruction (recon.cu:216-378, PPM.cu:7-113) by
mathematics the reference's algorithm must satisfy, not by retyping it:

* get_4th_parameters fits the cubic whose cell means (under the geometric
  weight 1, x, x^2) equal the data; for a monotone cubic with inactive
  limiter it returns exactly <f> - f(x_L) and f(x_R) - <f>;
* the limiter zeroes both slopes at an extremum and clips each slope to its
  one-sided difference;
* adjust_PPM_parameters equals the Colella & Woodward (1984) monotonicity
  rule written in edge-value form;
* get_PPM_aveL/aveR are the means of the cell parabola over the first /
  last fraction of the cell (checked by quadrature);
* flatten vanishes without compression or for linear pressure and is
  bounded by 1/2.
"""
import numpy as np
import pytest

import oracle


def cell_mean(f, xl, xr, geom, n=400):
    """Mean of f over [xl, xr] with weight 1 (geom 0/3), x (1), x^2 (2)."""
    x, w = np.polynomial.legendre.leggauss(24)
    xs = 0.5 * (xr - xl) * x + 0.5 * (xr + xl)
    ws = 0.5 * (xr - xl) * w
    wt = {0: 1.0, 3: 1.0, 1: xs, 2: xs * xs}[geom]
    return float(np.sum(ws * wt * f(xs)) / np.sum(ws * wt))


@pytest.mark.parametrize("geom", [0, 1, 2, 3])
@pytest.mark.parametrize("grid", ["uniform", "log"])
def test_fourth_order_exact_for_cubics(geom, grid):
    rng = np.random.default_rng(3)
    for trial in range(20):
        x0 = rng.uniform(0.5, 2.0)
        if grid == "uniform":
            dx = rng.uniform(0.005, 0.05)
            x6 = x0 + dx * np.arange(6)
        else:
            g = 1.0 + rng.uniform(0.005, 0.05)
            x6 = x0 * g ** np.arange(6)
        # monotone increasing cubic with moderate curvature
        c = [rng.uniform(-1, 1), rng.uniform(1.0, 2.0), rng.uniform(-0.2, 0.2), rng.uniform(-0.05, 0.05)]
        f = lambda x: c[0] + c[1] * x + c[2] * x**2 + c[3] * x**3
        a5 = [cell_mean(f, x6[k], x6[k + 1], geom) for k in range(5)]
        par = oracle.fourth_parameters(geom, x6, a5)
        sL_exact = a5[2] - f(x6[2])
        sR_exact = f(x6[3]) - a5[2]
        assert par[1] == a5[2]
        assert par[0] == pytest.approx(sL_exact, rel=2e-6, abs=1e-12)
        assert par[2] == pytest.approx(sR_exact, rel=2e-6, abs=1e-12)


def test_fourth_order_limiter():
    x6 = np.linspace(1.0, 1.06, 6)
    # local maximum in the middle cell: both slopes zero
    par = oracle.fourth_parameters(0, x6, [1.0, 2.0, 3.0, 2.0, 1.0])
    assert list(par) == [0.0, 3.0, 0.0]
    # flat neighbour on one side: zero
    par = oracle.fourth_parameters(0, x6, [1.0, 2.0, 2.0, 3.0, 4.0])
    assert par[0] == 0.0 and par[2] == 0.0
    # steep jump after a flat-ish region: slope clipped to the one-sided difference
    par = oracle.fourth_parameters(0, x6, [0.0, 0.01, 0.02, 10.0, 20.0])
    assert 0.0 < par[0] <= 0.01 + 1e-15
    assert 0.0 < par[2] <= 9.98 + 1e-12


def cw84(aL, a, aR):
    """Colella & Woodward (1984) eq. 1.10 monotonicity correction."""
    if (aR - a) * (a - aL) <= 0.0:
        return a, a
    d = aR - aL
    a6 = 6.0 * (a - 0.5 * (aL + aR))
    if d * a6 > d * d:
        aL = 3.0 * a - 2.0 * aR
    elif -d * d > d * a6:
        aR = 3.0 * a - 2.0 * aL
    return aL, aR


def test_ppm_adjust_is_cw84():
    rng = np.random.default_rng(7)
    for _ in range(2000):
        a = rng.uniform(-1, 1)
        sL, sR = rng.uniform(0.01, 1.0, size=2) * rng.choice([-1, 1])
        if sL * sR <= 0:
            continue
        par = oracle.ppm_adjust(0.0, [sL, a, sR])
        aL, aR = cw84(a - sL, a, a + sR)
        assert a - par[0] == pytest.approx(aL, abs=1e-14)
        assert a + par[2] == pytest.approx(aR, abs=1e-14)
    # flattening scales both slopes before the correction
    par = oracle.ppm_adjust(0.25, [0.4, 1.0, 0.5])
    assert par[0] == pytest.approx(0.3) and par[2] == pytest.approx(0.375)


def test_ppm_averages_are_parabola_means():
    rng = np.random.default_rng(11)
    xg, wg = np.polynomial.legendre.leggauss(12)
    for _ in range(200):
        sL, a, sR = rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1)
        aL, aR = a - sL, a + sR
        a6 = 6.0 * (a - 0.5 * (aL + aR))
        prof = lambda z: aL + z * ((aR - aL) + a6 * (1.0 - z))
        # the parabola's mean over the cell is the cell value
        assert np.sum(0.5 * wg * prof(0.5 * xg + 0.5)) == pytest.approx(a, abs=1e-14)
        q = rng.uniform(0.0, 1.0)
        zl = 0.5 * q * xg + 0.5 * q
        meanL = np.sum(0.5 * q * wg * prof(zl)) / q
        assert oracle.ppm_aveL(q, [sL, a, sR]) == pytest.approx(meanL, abs=1e-13)
        zr = 0.5 * (1 - q) * xg + 0.5 * (1 + q)
        meanR = np.sum(0.5 * (1 - q) * wg * prof(zr)) / (1 - q)
        assert oracle.ppm_aveR(q, [sL, a, sR]) == pytest.approx(meanR, abs=1e-13)
    # limits: aveL(0) is the left edge value, aveR(1) the right edge value
    assert oracle.ppm_aveL(0.0, [0.2, 1.0, 0.3]) == pytest.approx(0.8)
    assert oracle.ppm_aveR(1.0, [0.2, 1.0, 0.3]) == pytest.approx(1.3)


def test_flatten_properties():
    r = np.ones(5)
    # expansion everywhere: no flattening
    assert oracle.flatten(1.0, r, [1, 2, 3, 4, 5], [0, 1, 2, 3, 4]) == 0.0
    # compression but linear pressure: dp1/dp2 = 1/2, no flattening
    assert oracle.flatten(1.0, r, [1, 2, 3, 4, 5], [4, 3, 2, 1, 0]) == 0.0
    # compressive pressure jump: flattening switches on, capped at 1/2
    f = oracle.flatten(1.0, r, [1.0, 1.0, 1.0, 10.0, 10.0], [1, 1, 1, 0, 0])
    assert f == 0.5
    # weak jump: positive, proportional to the jump strength
    f1 = oracle.flatten(1.0, r, [1.0, 1.0, 1.0, 1.001, 1.001], [1, 1, 1, 0, 0])
    f2 = oracle.flatten(1.0, r, [1.0, 1.0, 1.0, 1.002, 1.002], [1, 1, 1, 0, 0])
    assert 0.0 < f1 < f2 < 0.5
    assert f2 / f1 == pytest.approx(2.0, rel=1e-3)
