"""Airfoil input formats and the application entry (aero2d/geometry.py:244-448,
aero2d/core.py:220-260), on the CPU: normalisation of scaled/rotated and
reversed contours, Selig and Lednicer files, re-distribution along the
arc-length spline (points on the input curve, cosine spacing, kept corners,
blunt trailing edges), smoothing of noisy coordinates, and the M < 0.2 guard
of run_euler. The reference's own tests of these functions compare panel
method results (out of scope here); the properties below pin the same
behaviour directly."""
import numpy as np
import pytest

from aero2d_amd.geometry import Airfoil, load_airfoil, naca4, normalize, repanel, smooth_airfoil


def test_normalize_undoes_scale_rotation_and_shift():
    a = naca4("2412", 60)
    t = np.radians(10)
    x = 2 * (a.x * np.cos(t) - a.y * np.sin(t)) + 5
    y = 2 * (a.x * np.sin(t) + a.y * np.cos(t)) - 1
    n = normalize(Airfoil(x, y, "moved"))
    assert np.allclose(n.x, a.x, atol=1e-12) and np.allclose(n.y, a.y, atol=1e-12)


def test_normalize_makes_the_contour_counter_clockwise():
    a = naca4("2412", 40)
    n = normalize(Airfoil(a.x[::-1].copy(), a.y[::-1].copy()))
    assert np.array_equal(n.x, a.x) and np.array_equal(n.y, a.y)


def _write(path, name, xs, ys, header=None):
    lines = ([name] if name else []) + ([header] if header else [])
    path.write_text("\n".join(lines + [f"{x:.8f} {y:.8f}" for x, y in zip(xs, ys)]) + "\n")


def test_selig_reversed_and_lednicer_files_give_one_contour(tmp_path):
    a = naca4("2412", 35)
    n = 35
    _write(tmp_path / "selig.dat", "NACA 2412", a.x, a.y)
    _write(tmp_path / "reversed.dat", None, a.x[::-1], a.y[::-1])
    up_x, up_y = a.x[:n][::-1], a.y[:n][::-1]                 # leading edge -> trailing edge
    lo_x, lo_y = a.x[n - 1:], a.y[n - 1:]
    (tmp_path / "lednicer.dat").write_text(
        "NACA 2412\n35. 35.\n\n" + "\n".join(f"{x:.8f} {y:.8f}" for x, y in zip(up_x, up_y)) + "\n\n"
        + "\n".join(f"{x:.8f}, {y:.8f}" for x, y in zip(lo_x, lo_y)) + "\n")
    got = [load_airfoil(str(tmp_path / f), 65) for f in ("selig.dat", "reversed.dat", "lednicer.dat")]
    assert got[0].name == "NACA 2412" and got[2].name == "NACA 2412"
    for g in got:
        assert len(g.x) == 2 * 65 - 1
        assert np.allclose(g.x, got[0].x, atol=1e-12) and np.allclose(g.y, got[0].y, atol=1e-12)
    # the files hold the points to 8 decimals: the same as re-distributing them in memory
    ref = repanel(a, 65)
    assert np.allclose(got[0].x, ref.x, atol=1e-6) and np.allclose(got[0].y, ref.y, atol=1e-6)


def test_too_few_points_are_refused(tmp_path):
    _write(tmp_path / "short.dat", "few", np.linspace(1, 0, 5), np.zeros(5))
    with pytest.raises(ValueError):
        load_airfoil(str(tmp_path / "short.dat"))


def test_repanel_places_points_on_the_contour_with_cosine_spacing():
    r = repanel(naca4("0012", 120), 65)
    assert len(r.x) == 129 and r.x[0] == r.x[-1] == 1.0 and r.y[0] == r.y[-1]
    # on the airfoil: |y| is the closed-trailing-edge NACA 00xx half thickness at x
    xc = np.clip(r.x, 0, 1)
    yt = 0.6 * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc**2 + 0.2843 * xc**3 - 0.1036 * xc**4)
    assert np.max(np.abs(np.abs(r.y) - yt)[r.x > 0.01]) < 2e-6
    assert np.max(np.abs(np.abs(r.y) - yt)) < 2e-4                 # interpolation near the nose
    i = int(np.argmin(r.x))
    assert i == 64 and abs(r.x[i]) < 1e-6 and abs(r.y[i]) < 1e-5  # leading edge in the middle
    # arc length of the upper side follows (1 - cos) / 2 (cosine spacing in s)
    s = np.concatenate([[0], np.cumsum(np.hypot(np.diff(r.x[:i + 1]), np.diff(r.y[:i + 1])))])
    target = 0.5 * (1 - np.cos(np.linspace(0, np.pi, 65))) * s[-1]
    assert np.max(np.abs(s - target)) < 2e-3 * s[-1]
    # the symmetric section stays symmetric
    assert np.allclose(r.y[:i + 1][::-1], -r.y[i:], atol=2e-6)


def test_repanel_keeps_corners():
    """A contour with the two corners of a flap-cove step keeps a node exactly
    at each corner and no overshoot next to them."""
    t = np.linspace(0, 1, 31)
    x = np.r_[1 - t, 0.5 * t[1:], np.full(10, 0.5), 0.5 + 0.5 * t[1:]]
    y = np.r_[0.06 * np.sin(np.pi * (1 - t)), -0.04 * np.sin(0.5 * np.pi * t[1:]),
              np.linspace(-0.04, -0.01, 11)[1:], -0.01 * (1 - t[1:])]
    r = repanel(Airfoil(x, y, "cove"), 60, keep_frame=True)
    for cx, cy in ((0.5, -0.04), (0.5, -0.01)):
        assert np.min(np.hypot(r.x - cx, r.y - cy)) < 1e-12
    aft = r.x > 0.5 + 1e-6
    assert r.y[aft & (r.y < 0)].min() >= -0.01 - 1e-9


def test_blunt_trailing_edge_is_closed_by_a_segment():
    a = naca4("0012", 80, closed_te=False)                      # a thin base, 0.25 % of the chord
    gap = np.hypot(a.x[0] - a.x[-1], a.y[0] - a.y[-1])
    assert gap > 2e-3
    r = repanel(a, 50)
    assert len(r.x) == 2 * 50 and r.x[0] == r.x[-1] and r.y[0] == r.y[-1]
    assert abs(abs(r.y[0] - r.y[-2]) - gap) < 1e-3            # the base kept as its own segment
    m = repanel(a, 50, te_merge=5e-3)                          # within te_merge: ends merged
    assert len(m.x) == 2 * 50 - 1 and m.y[0] == m.y[-1]


def test_smoothing_removes_noise_and_keeps_the_shape():
    af = naca4("2412", 81)
    rng = np.random.default_rng(0)
    noisy = Airfoil(af.x, af.y + rng.normal(0, 3e-4, len(af.y)) * (af.x > 0.002))

    def upper(a):
        i = int(np.argmin(a.x))
        return np.interp(np.linspace(0.01, 0.95, 200), a.x[:i + 1][::-1], a.y[:i + 1][::-1])
    true = upper(af)
    assert np.max(np.abs(upper(smooth_airfoil(af)) - true)) < 3e-4
    assert np.std(upper(smooth_airfoil(noisy)) - true) < 0.6 * 3e-4


def test_run_euler_refuses_low_mach_before_any_work():
    from aero2d_amd.core import run_euler
    with pytest.raises(ValueError):
        run_euler(naca4("0012", 60), [0.0], 0.1)


def test_grid_presets():
    from aero2d_amd.core import GRIDS, grid_key
    assert GRIDS[grid_key("coarse")] == (65, 49, 2e-3)
    assert grid_key("средняя") == "средняя"
    with pytest.raises(KeyError):
        grid_key("huge")
