"""The closed-form room model the surrogate learns a correction to (surrogate.py): the free-field
Gaussian pulse and its images in a box with rigid and absorbing walls.

Pinned by what the mathematics fixes, not by restating the formula: the free-field solution
satisfies the spherical wave equation (r p)_tt = c^2 (r p)_rr and the initial data (the Gaussian,
at rest); images make the normal pressure gradient vanish on rigid walls; a receiver far from
every wall sees the free field until the first reflection arrives.
"""
import numpy as np

import surrogate

C0, SIG = 343.0, 0.109


def test_free_field_solves_the_wave_equation_and_starts_at_rest():
    r = np.linspace(0.3, 3.0, 60)[:, None]
    t = np.linspace(1e-4, 6e-3, 50)[None, :]
    hr, ht = 1e-4, 1e-7
    f = lambda rr, tt: rr * surrogate.gaussian_pulse_free_field(rr, tt, SIG, C0)
    tt = (f(r, t + ht) - 2 * f(r, t) + f(r, t - ht)) / ht ** 2
    rr = (f(r + hr, t) - 2 * f(r, t) + f(r - hr, t)) / hr ** 2
    scale = np.abs(tt).max()
    assert np.abs(tt - C0 ** 2 * rr).max() <= 1e-4 * scale
    r1 = np.linspace(0.01, 1.0, 200)
    assert np.allclose(surrogate.gaussian_pulse_free_field(r1, 0.0, SIG, C0), np.exp(-r1 ** 2 / (2 * SIG ** 2)))
    dt = 1e-8
    d = (surrogate.gaussian_pulse_free_field(r1, dt, SIG, C0) - surrogate.gaussian_pulse_free_field(r1, 0.0, SIG, C0))
    assert np.abs(d / dt).max() <= 1e-3 * C0 / SIG  # p_t(r, 0) = 0 up to O(dt)


def test_images_make_rigid_walls_reflect_and_absorbing_walls_transparent():
    L = (4.0, 3.0, 2.5)
    t = np.linspace(1e-4, 2e-2, 400)
    src = np.array([[1.3, 0.9, 0.7]])
    eps = 1e-4
    # rigid x = 0 wall: dp/dx = 0 there (central difference across the wall plane)
    for rcv in ([eps, 1.7, 1.1], [L[0] - eps, 2.2, 0.4], [2.0, eps, 1.0], [2.1, 1.0, eps]):
        rp = np.array([rcv])
        k = int(np.argmin([min(rcv[j], L[j] - rcv[j]) for j in range(3)]))
        rm = rp.copy()
        rm[0, k] = -rcv[k] if rcv[k] < 1 else 2 * L[k] - rcv[k]
        a = surrogate.image_source_series(src, rp, L, SIG, C0, t)
        b = surrogate.image_source_series(src, rm, L, SIG, C0, t)
        assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()
    # an absorbing ceiling has no image: the field just below it equals the free field plus the
    # images of the other walls, and differs from the field with a rigid ceiling
    rp = np.array([[2.0, 1.5, L[2] - 0.05]])
    absorbing = surrogate.image_source_series(src, rp, L, SIG, C0, t)
    rigid = surrogate.image_source_series(src, rp, L, SIG, C0, t, rigid=(True,) * 6)
    assert np.abs(rigid - absorbing).max() > 0.1 * np.abs(absorbing).max()


def test_free_field_until_the_first_reflection():
    L = (20.0, 15.0, 8.0)
    src, rcv = np.array([[10.0, 7.5, 1.0]]), np.array([[11.0, 7.5, 1.0]])
    t = np.arange(1, 2049) * 1.04e-5
    y = surrogate.image_source_series(src, rcv, L, SIG, C0, t)[0]
    free = surrogate.gaussian_pulse_free_field(1.0, t, SIG, C0)
    first = np.sqrt(1.0 + 2.0 ** 2) / C0  # the floor image, 2.24 m away
    early = t < first - 9 * SIG / C0
    assert np.abs(y[early] - free[early]).max() <= 1e-12 * np.abs(free).max()
    assert np.abs(y[~early] - free[~early]).max() > 1e-3 * np.abs(free).max()
