"""Pins of the oracle's light_sample (template.wgsl:2178-2229) at a surface point, by closed
forms of each light kind:

- point: unit direction to the light, colour * brightness / max(d^2, 0.01), distance d;
- global: the normalised direction, colour * brightness, distance 2 max_dist;
- ray (a Gaussian beam from its position along ray_direction): on the axis at depth t,
  colour * brightness / w^2 with w = waist + t tan(spread); off the axis by r at the same
  depth, exp(-r^2 / w^2) of that; nothing behind the source; the spread clamped to 1.5 rad;
  the axis normalised (its length does not matter); the direction and distance those of a
  point light at the beam's source.
The inputs pass through the light record's float32 fields, so the closed forms use the float32
values of spread and waist.
"""
import math

import numpy as np

import oracle


def f32(x):
    return float(np.float32(x))


def test_point_and_global():
    d, c, dist = oracle.light_sample(0, (1.0, 2.0, 3.0), (0.5, 1.0, 0.25), 4.0, (1.0, 2.0, 1.0))
    assert np.allclose(d, [0, 0, 1]) and abs(dist - 2.0) < 1e-12
    assert np.allclose(c, np.array([0.5, 1.0, 0.25]) * 4.0 / 4.0)
    d, c, dist = oracle.light_sample(0, (0.0, 0.05, 0.0), (1, 1, 1), 1.0, (0.0, 0.0, 0.0))
    assert np.allclose(c, [100, 100, 100])  # d^2 = 0.0025 floors at 0.01
    d, c, dist = oracle.light_sample(1, (0.0, 3.0, 4.0), (1, 0.5, 0.25), 2.0, (9.0, 9.0, 9.0), max_dist=7.0)
    assert np.allclose(d, [0, 0.6, 0.8]) and np.allclose(c, [2, 1, 0.5]) and dist == 14.0


def test_ray_on_axis_and_off_axis():
    src, axis, spread, waist = (0.0, 3.0, 0.0), (0.0, -1.0, 0.0), 0.2, 0.05
    s, w0 = f32(spread), f32(waist)
    for t in (0.5, 1.0, 2.5):
        pos = (0.0, 3.0 - t, 0.0)
        d, c, dist = oracle.light_sample(2, src, (1.0, 0.5, 0.25), 3.0, pos, ray_direction=axis, spread=spread,
                                         waist=waist)
        w = w0 + t * math.tan(s)
        assert np.allclose(c, np.array([1.0, 0.5, 0.25]) * 3.0 / w ** 2, rtol=1e-12)
        assert np.allclose(d, [0, 1, 0]) and abs(dist - t) < 1e-12
        for r in (0.01, 0.1, 0.3):
            _, c2, dist2 = oracle.light_sample(2, src, (1.0, 0.5, 0.25), 3.0, (r, 3.0 - t, 0.0), ray_direction=axis,
                                               spread=spread, waist=waist)
            assert np.allclose(c2, c * math.exp(-r * r / (w * w)), rtol=1e-12)
            assert abs(dist2 - math.hypot(r, t)) < 1e-12


def test_ray_behind_source_clamp_and_axis_length():
    src = (1.0, 1.0, 1.0)
    _, c, _ = oracle.light_sample(2, src, (1, 1, 1), 1.0, (1.0, 2.0, 1.0), ray_direction=(0, -1, 0), spread=0.3)
    assert np.all(c == 0)  # behind the source
    _, c, _ = oracle.light_sample(2, src, (1, 1, 1), 1.0, (1.0, 1.0, 1.0), ray_direction=(0, -1, 0), spread=0.3)
    assert np.all(c == 0)  # at the source plane: t0 = 0 is not in front
    pos = (1.2, -1.0, 0.7)
    _, a, _ = oracle.light_sample(2, src, (1, 1, 1), 1.0, pos, ray_direction=(0.3, -1.0, 0.2), spread=0.3)
    _, b, _ = oracle.light_sample(2, src, (1, 1, 1), 1.0, pos, ray_direction=(2.1, -7.0, 1.4), spread=0.3)
    assert np.allclose(a, b, rtol=1e-6)
    _, wide, _ = oracle.light_sample(2, src, (1, 1, 1), 1.0, pos, ray_direction=(0, -1, 0), spread=3.0)
    _, cap, _ = oracle.light_sample(2, src, (1, 1, 1), 1.0, pos, ray_direction=(0, -1, 0), spread=1.5)
    assert np.array_equal(wide, cap)
    t, r2 = 2.0, 0.2 ** 2 + 0.3 ** 2
    w = f32(0.05) + t * math.tan(1.5)
    assert np.allclose(cap, [math.exp(-r2 / w ** 2) / w ** 2] * 3, rtol=1e-6)
