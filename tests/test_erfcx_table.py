"""The erfcx pieces the Ewald correction evaluates erfc and erf with (ryoanji_amd/csrc/erfcx_table.hpp):
evaluated by Horner in double as on the device, within 2.5e-16 of erfcx(x) = exp(x^2) erfc(x) computed in
50-digit arithmetic, on every interval and at its ends (continuity), and reproducible from the generator."""

import os
import re
import subprocess
import sys

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "ryoanji_amd", "csrc")


def _table():
    text = open(os.path.join(CSRC, "erfcx_table.hpp")).read()
    body = text[text.index("{", text.index("kErfcx[")) + 1:text.rindex("};")]
    rows = re.findall(r"\{([^}]*)\}", body)
    return np.array([[float(v) for v in r.split(",") if v.strip()] for r in rows])


def _eval(tab, x):
    j = min(int(x * 4.0), tab.shape[0] - 1)
    t = (x - (j + 0.5) * 0.25) * 8.0
    v = 0.0
    for c in tab[j][::-1]:
        v = v * t + c
    return v


def test_pieces_match_erfcx():
    mp.mp.dps = 40
    tab = _table()
    assert tab.shape == (32, 12)
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(0, 8, 3000), np.arange(33) * 0.25 - 1e-12, np.arange(32) * 0.25])
    xs = xs[(xs >= 0) & (xs < 8)]
    worst = 0.0
    for x in xs:
        ref = mp.exp(mp.mpf(x) ** 2) * mp.erfc(mp.mpf(x))
        worst = max(worst, abs(float((mp.mpf(_eval(tab, x)) - ref) / ref)))
    assert worst < 2.5e-16, worst


def test_generator_reproduces_the_table():
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_erfcx_table.py")], capture_output=True, text=True,
                         check=True).stdout
    assert out == open(os.path.join(CSRC, "erfcx_table.hpp")).read()
