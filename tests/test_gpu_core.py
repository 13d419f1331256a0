"""The application entry on the GPU: run_euler (aero2d/core.py:220-260) builds
the preset grid from a re-distributed contour and solves every angle in one
mixed-precision batch."""
import numpy as np
import pytest

from aero2d_amd import naca4
from aero2d_amd.core import run_euler

pytestmark = pytest.mark.gpu


def test_run_euler_polar_and_callback():
    """NACA 0012 at M 0.5, alpha 0 and 2 on the coarse preset (128 x 48): both
    converge, Cl(0) = 0 by symmetry, Cl(2) within 3e-3 of the reference's
    published 0.2816 (its grid comes from naca4 directly, this one from the
    re-distributed contour, as the reference's run_euler does), and the
    callback receives progress text and fractions."""
    msgs = []
    p = run_euler(naca4("0012", 100), [0.0, 2.0], 0.5, callback=lambda t, f: msgs.append((t, f)) and False)
    print(p.label, p.cl, p.cd, p.cm, p.ok, p.notes)
    assert p.method == "euler" and p.ok.all()
    assert abs(p.cl[0]) < 1e-4 and abs(p.cl[1] - 0.28162520444340794) < 3e-3
    assert len(p.cp) == 2 and len(p.cp[0]) == len(p.cp_x[0]) == 128
    assert any(t.startswith("Iteration") for t, _ in msgs)
    assert 0 < max(f for _, f in msgs) <= 1


def test_run_euler_stops_when_the_callback_asks():
    calls = []

    def cb(text, frac):
        calls.append(text)
        return text.startswith("Iteration 3/")
    p = run_euler(naca4("0012", 100), [1.0], 0.6, callback=cb)
    assert calls[-1].startswith("Iteration 3/")
    assert not p.ok[0] and "Not converged" in p.notes[0]
