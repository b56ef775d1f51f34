"""Pins of the 3-D oracle (the reference has no 3-D route; these extend its 2-D
definitions axis by axis): the direct adjoint against a brute-force numpy sum, the
KB gridding adjoint against the exact direct adjoint on a ladder of kernel widths
(its approximation error must fall as the width grows, to the levels the 2-D
oracle reaches), the binning rule against floor(w G / 2 pi) with wrap, and the
golden-means kooshball against its defining recurrence."""
import numpy as np
import pytest

import oracle as O
from tests.helpers import rel_l2


def ball_trajectory(samples, seed):
    """Random frequencies in the ball |w| < pi (radians per pixel), float32-exact."""
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((samples, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    r = np.pi * rng.uniform(0, 1, samples) ** (1 / 3) * 0.999
    return (v * r[:, None]).astype(np.float32).astype(np.float64)


def test_direct_adjoint3_is_the_brute_force_sum():
    rng = np.random.default_rng(1)
    shape = (3, 4, 5)
    t = rng.uniform(-np.pi, np.pi, (23, 3))
    y = rng.standard_normal(23) + 1j * rng.standard_normal(23)
    origin = (1.0, 1.5, 2.0)
    z, yy, x = np.meshgrid(*(np.arange(n) - o for n, o in zip(shape, origin)), indexing="ij")
    ref = np.einsum("s,szyx->zyx", y, np.exp(1j * (t[:, 0, None, None, None] * z + t[:, 1, None, None, None] * yy
                                                   + t[:, 2, None, None, None] * x)))
    assert rel_l2(O.nudft3_adjoint(y, t, shape, origin), ref) <= 1e-13


@pytest.mark.parametrize("width,bound", [(4, 2e-3), (6, 5e-5), (8, 1e-6)])
def test_kb_adjoint3_converges_to_the_direct_adjoint(width, bound):
    shape = (10, 10, 10)
    t = ball_trajectory(1500, 2)
    rng = np.random.default_rng(3)
    y = rng.standard_normal(t.shape[0]) + 1j * rng.standard_normal(t.shape[0])
    dcf = O.ramp_sq_dcf3(t)
    exact = O.nudft3_adjoint(y * dcf, t, shape)
    assert rel_l2(O.kb_adjoint3(y, t, dcf, shape, width=width), exact) <= bound


def test_kb_adjoint3_error_falls_with_the_width():
    shape = (8, 12, 10)  # unequal axes: every axis has its own grid and deapodization
    t = ball_trajectory(900, 4)
    y = np.exp(1j * np.arange(t.shape[0]))
    exact = O.nudft3_adjoint(y, t, shape)
    errs = [rel_l2(O.kb_adjoint3(y, t, None, shape, width=w), exact) for w in (2, 4, 6, 8)]
    assert all(b < a / 10 for a, b in zip(errs, errs[1:])), errs


def test_bin3_follows_the_wrapped_floor_rule_and_sorts_stably():
    rng = np.random.default_rng(5)
    t = rng.uniform(-np.pi, np.pi, (4000, 3)).astype(np.float32)
    t[:7] = [[0, 0, 0], [np.float32(np.pi), 0, 0], [-np.float32(np.pi), 0, 0], [0, 1e-30, -1e-30], [3.0, -3.0, 0.5],
             [0, 0, 0], [1, 1, 1]]
    grid = (12, 10, 14)
    bins, start, perm = O.bin3(t, grid)
    u = t.astype(np.float64) * np.array(grid) / (2 * np.pi)
    u = np.where(u < 0, u + np.array(grid), u)
    u = np.where(u >= np.array(grid), u - np.array(grid), u)
    c = np.floor(u).astype(np.int64)
    assert np.array_equal(bins, (c[:, 0] * grid[1] + c[:, 1]) * grid[2] + c[:, 2])
    assert np.array_equal(perm, np.argsort(bins, kind="stable"))
    assert np.array_equal(start, np.searchsorted(np.sort(bins), np.arange(np.prod(grid) + 1)))


def test_golden_means3_directions():
    g1, g2 = 0.465571231876768, 0.682327803828019
    lam = 1.0 + g1
    assert abs(lam ** 3 - lam ** 2 - 1.0) < 1e-12 and abs(g2 - 1.0 / lam) < 1e-12
    spokes, readout = 50, 16
    t = O.golden_means3(spokes, readout).reshape(spokes, readout, 3)
    k = 2 * np.pi * (np.arange(readout) - readout // 2) / readout
    d = t[:, 0, :] / k[0]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-14)
    assert np.allclose(t, k[None, :, None] * d[:, None, :], atol=1e-14)
    p = np.arange(spokes)
    assert np.allclose(d[:, 0], np.mod(p * g1, 1.0), atol=1e-14)
    assert np.allclose(np.mod(np.arctan2(d[:, 1], d[:, 2]), 2 * np.pi)[1:], 2 * np.pi * np.mod(p * g2, 1.0)[1:],
                       atol=1e-12)


def test_ramp_sq_dcf3_is_the_squared_norm():
    """|k|^2 in (rad/px)^2: a KAT, numpy's sum of squares, and not |k| (a |k|-for-|k|^2 slip
    would pass every test that applies the same weights on both sides)."""
    kat = O.ramp_sq_dcf3(np.array([[1.0, 2.0, 2.0], [0.0, 0.0, 0.0], [-np.pi, 0.0, 0.0], [0.5, -0.5, 0.5]]))
    assert np.array_equal(kat, [9.0, 0.0, np.pi ** 2, 0.75])
    t = ball_trajectory(5000, 8)
    d = O.ramp_sq_dcf3(t)
    assert np.allclose(d, (t ** 2).sum(axis=1), rtol=1e-15, atol=0)
    assert not np.allclose(d, np.linalg.norm(t, axis=1))


@pytest.mark.parametrize("shape,width", [((8, 8, 8), 4), ((6, 10, 12), 6), ((9, 7, 5), 2)])
def test_kb_adjoint3_voxels_equal_the_full_volume(shape, width):
    """The lattice evaluation (used at sizes where the full FP64 volume takes minutes) is
    the same sum as kb_adjoint3 at those voxels, also with the samples split over threads."""
    t = ball_trajectory(700, 9)
    rng = np.random.default_rng(10)
    y = rng.standard_normal(t.shape[0]) + 1j * rng.standard_normal(t.shape[0])
    dcf = O.ramp_sq_dcf3(t)
    full = O.kb_adjoint3(y, t, dcf, shape, width=width)
    zs, ys, xs = [0, shape[0] // 2, shape[0] - 1], [1, shape[1] - 2], list(range(shape[2]))
    ref = full[np.ix_(zs, ys, xs)]
    for threads in (1, 3):
        got = O.kb_adjoint3_voxels(y, t, dcf, shape, zs, ys, xs, width=width, threads=threads)
        assert rel_l2(got, ref) <= 1e-12
    # a wrong crop index (the voxel next to it) is caught
    other = O.kb_adjoint3_voxels(y, t, dcf, shape, zs, ys, [x + 1 for x in xs[:-1]] + [0], width=width)
    assert rel_l2(other, ref) > 1e-3


def test_nudft3_forward_is_the_brute_force_sum_and_the_adjoint_transpose():
    """ora_nudft3_forward against a numpy sum, and <A x, y> = <x, A^H y> with the adjoint."""
    rng = np.random.default_rng(31)
    shape, S = (3, 4, 5), 17
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    t = rng.uniform(-np.pi, np.pi, (S, 3))
    o = ((shape[0] - 1) / 2, (shape[1] - 1) / 2, (shape[2] - 1) / 2)
    zz, yy, xx = np.meshgrid(*[np.arange(n) - oo for n, oo in zip(shape, o)], indexing="ij")
    want = np.array([np.sum(x * np.exp(-1j * (w[0] * zz + w[1] * yy + w[2] * xx))) for w in t])
    got = O.nudft3_forward(x, t)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    y = rng.standard_normal(S) + 1j * rng.standard_normal(S)
    lhs = np.vdot(y, got)
    rhs = np.vdot(O.nudft3_adjoint(y, t, shape), x)
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    shifted = O.nudft3_forward(x, t, origin=(o[0] + 1, o[1], o[2]))
    assert np.max(np.abs(shifted - want)) > 1e-3  # the origin is honoured
