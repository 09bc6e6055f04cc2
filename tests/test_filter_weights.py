"""Filter weights pinned by their mathematics (CPU), for the oracle's
restatement and the host mirror alike, so that neither side can carry a
mistake the other shares: normalisation, symmetry, the Gaussian's discrete
second moment sigma^2 = Delta^2 / 12, Julia's ties-to-even radius rounding,
the trapezoid as a box convolved with a two-point mean, and the moment
additivity of the discrete convolution (composekernel, build_doublekernel).

References: tophat / gaussian / composekernel (src/N.jl:592-625),
tophat_weights / gaussian_weights (src/Burgers.jl:146-162),
build_doublekernel (src/Burgers.jl:217-233), the LES kernels of
compute_fractions (src/Burgers.jl:1030-1048)."""
import math

import numpy as np
import pytest


def _moments(w):
    w = np.asarray(w, dtype=np.float64)
    r = np.arange(len(w)) - (len(w) - 1) / 2
    m0 = w.sum()
    m1 = (w * r).sum() / m0
    return m0, m1, (w * (r - m1) ** 2).sum() / m0


def _gaussians(ec, O):
    """(name, f(h, delta, nsigma) -> weights) for every Gaussian builder."""
    NS, B = ec.navierstokes, ec.burgers
    return [("oracle gaussian", lambda h, d, s: np.asarray(O.gaussian(h, d, s))),
            ("oracle gaussian_weights", lambda h, d, s: np.asarray(O.gaussian_weights(h, d, s)[1])),
            ("mirror gaussian", lambda h, d, s: np.asarray(NS.gaussian(NS.Grid(1.0, int(round(1 / h))), d, s))),
            ("mirror gaussian_weights", lambda h, d, s: np.asarray(B.gaussian_weights(h, d, s)[1]))]


@pytest.mark.parametrize("c", [1, 3, 5, 9, 15])
def test_tophat_weights(ec, O, c):
    """c weights of 1/c: sum 1, symmetric, variance (c^2 - 1) / 12 cells^2 (the
    discrete uniform distribution)."""
    for w in (O.tophat(c), O.tophat_weights(c)[1], ec.navierstokes.tophat(ec.navierstokes.Grid(1.0, 8), c),
              ec.burgers.tophat_weights(c)[1]):
        w = np.asarray(w)
        assert len(w) == c
        m0, m1, var = _moments(w)
        assert abs(m0 - 1) < 1e-15 and abs(m1) < 1e-14 and abs(var - (c * c - 1) / 12) < 1e-12
    assert O.tophat_weights(c)[0] == ec.burgers.tophat_weights(c)[0] == c // 2


@pytest.mark.parametrize("c", [2, 4, 8, 16])
def test_trapezoid_is_box_times_two_point_mean(ec, O, c):
    """The even-compression face kernel: the box of c cells convolved with the
    two-point mean [1/2, 1/2] (c + 1 taps, ends 1/(2c)); variance (c^2 - 1)/12 + 1/4."""
    want = np.convolve(np.full(c, 1.0 / c), [0.5, 0.5])
    for w in (O.trapezoid_weights(c)[1], ec.burgers.tophat_weights(c)[1], ec.navierstokes.trapezoid_face(c),
              O._face(c)[0]):
        assert np.abs(np.asarray(w) - want).max() < 1e-16
    assert abs(_moments(want)[2] - ((c * c - 1) / 12 + 0.25)) < 1e-12


@pytest.mark.parametrize("sigma_cells", [1.5, 3.0, 7.25])
def test_gaussian_second_moment(ec, O, sigma_cells):
    """With a wide support (nsigma = 8) the normalised, symmetric discrete
    Gaussian has second moment sigma^2 = Delta^2 / 12 to 1e-8 for sigma >= 1.5
    cells (sampling error ~ exp(-2 pi^2 sigma^2), truncation ~ erfc(8/sqrt 2));
    a wrong sigma (Delta/12, Delta/sqrt 6, ...) or an unnormalised kernel fails."""
    h = 1.0 / 64
    delta = sigma_cells * math.sqrt(12) * h
    for name, f in _gaussians(ec, O):
        w = f(h, delta, 8)
        m0, m1, var = _moments(w)
        assert abs(m0 - 1) < 1e-14, name
        assert np.abs(w - w[::-1]).max() < 1e-17, name
        assert abs(m1) < 1e-12, name
        assert abs(var / sigma_cells ** 2 - 1) < 1e-8, (name, var, sigma_cells ** 2)


def test_gaussian_radius_rounds_half_to_even(ec, O):
    """r = round(nsigma sigma / h) with Julia's default rounding (ties to
    even): for nsigma sigma / h exactly 2.5, 3.5, 4.5 the radius is 2, 4, 4
    (len 5, 9, 9), where rounding half up would give 3, 4, 5."""
    h = 1.0
    for x, r in ((2.5, 2), (3.5, 4), (4.5, 4), (5.5, 6)):
        # delta such that nsigma * (delta / sqrt 12) / h is exactly x, nsigma = 1
        delta = x * math.sqrt(12)
        assert (delta / math.sqrt(12)) / h == x and math.sqrt((delta / h) ** 2 / 12) == x
        for name, f in _gaussians(ec, O):
            if name.startswith("mirror gaussian") and "weights" not in name:
                continue  # the NS mirror takes h from a grid of l = 1
            assert len(f(h, delta, 1)) == 2 * r + 1, (name, x)
    # the NS mirror with a grid: h = 1/8, sigma = 2.5 h
    NS = ec.navierstokes
    g = NS.Grid(1.0, 8)
    d = 2.5 * math.sqrt(12) / 8
    if (d / math.sqrt(12)) / (1.0 / 8) == 2.5:
        assert len(NS.gaussian(g, d, 1)) == 5


def test_gaussian_zero_width_is_identity(ec, O):
    """Delta = 0 is the identity kernel [1] (the reference's gaussian would
    divide 0 by 0 there; src/Burgers.jl:149-150 special-cases it)."""
    for name, f in _gaussians(ec, O):
        assert list(f(0.1, 0.0, 3)) == [1.0], name


def test_convolution_moments_add(ec, O):
    """For kernels f, g (asymmetric here) the convolution has mass m0(f) m0(g),
    mean m1(f) + m1(g) and variance var(f) + var(g); a correlation (f
    reversed) or an index shift fails the mean."""
    rng = np.random.default_rng(3)
    f = rng.uniform(0.1, 1.0, 7)
    g = rng.uniform(0.1, 1.0, 5)
    mf, mg = _moments(f), _moments(g)
    outs = [np.asarray(O.composekernel(f, g)), np.asarray(ec.navierstokes.composekernel(f, g)),
            np.asarray(O.build_doublekernel((3, list(f)), (2, list(g)))[1]),
            np.asarray(ec.burgers.build_doublekernel((3, f), (2, g))[1])]
    for w in outs:
        assert len(w) == 11
        m0, m1, var = _moments(w)
        assert abs(m0 - mf[0] * mg[0]) < 1e-12
        assert abs(m1 - (mf[1] + mg[1])) < 1e-12
        assert abs(var - (mf[2] + mg[2])) < 1e-12
    assert O.build_doublekernel((3, list(f)), (2, list(g)))[0] == ec.burgers.build_doublekernel((3, f), (2, g))[0] == 5


@pytest.mark.parametrize("nh,nH,dr,kind", [(13500, 300, 2.0, "gaussian"), (13500, 900, 0.0, "gaussian"),
                                           (225, 15, 1.0, "tophat"), (135, 45, 3.0, "tophat"),
                                           (600, 100, 1.5, "tophat"), (600, 100, 2.5, "tophat"),
                                           (64, 16, 1.0, "tophat"), (600, 100, 0.75, "tophat"),
                                           (600, 100, 1.25, "tophat")])
def test_burgers_les_kernels(ec, O, nh, nH, dr, kind):
    """The kernels of compute_fractions: the LES top-hat spans 2 (round(Delta/h)
    div 2) + 1 fine cells (Delta/h = 9 -> 9, 6 -> 7 (the even width made odd),
    4.5 and 7.5 round to 4 and 8 -> 5 and 9), the Gaussian has variance
    (Delta/h)^2 / 12 up to its 3-sigma truncation, and the double kernel has the
    variance of the finite-volume kernel plus that of the LES kernel; the
    oracle's and the mirror's kernels agree."""
    L = 2 * math.pi
    (Ro, wo), (Ko, do) = O.burgers_les_kernels(nh, nH, L, dr, kind)
    (Rm, wm), (Km, dm) = ec.burgers.les_kernels(nh, nH, L, dr, kind)
    assert (Ro, Ko) == (Rm, Km) and len(wo) == 2 * Ro + 1 and len(do) == 2 * Ko + 1
    assert np.abs(np.asarray(wo) - np.asarray(wm)).max() < 1e-15
    assert np.abs(np.asarray(do) - np.asarray(dm)).max() < 1e-15
    comp = nh // nH
    dh = dr * nh / nH  # Delta / h
    if kind == "tophat":
        want = {9.0: 9, 6.0: 7, 4.5: 5, 7.5: 9, 4.0: 5, 3.0: 3}
        if dh in want:
            assert len(wo) == want[dh], (dh, len(wo))
        assert len(wo) == 2 * (int(round(dh)) // 2) + 1
        var_les = (len(wo) ** 2 - 1) / 12
    else:
        var_les = _moments(wo)[2]
        if dr > 0:
            assert abs(var_les / (dh * dh / 12) - 1) < 0.03  # 3-sigma truncation lowers it by ~2.6%
    var_fvm = (comp * comp - 1) / 12 + (0 if comp % 2 else 0.25)
    m0, m1, var = _moments(do)
    assert abs(m0 - 1) < 1e-13 and abs(m1) < 1e-9 and abs(var - (var_fvm + var_les)) < 1e-8 * max(1, var)
