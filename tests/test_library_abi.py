"""The C-ABI library loads and exports every symbol include/harmstep.h declares
(CPU only: no compute call needs a GPU here)."""
import os
import re

import harm_harm_amd as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "harmstep.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(harm_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_all_declared_symbols():
    L = H.lib()
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing


def test_default_config_is_harm_torus():
    c = H.default_config(256, 128, 64)
    assert c.n1 == 256 and c.ng == 4
    assert abs(c.a - 0.9375) < 1e-15 and abs(c.gam - 4.0 / 3.0) < 1e-15
    assert abs(c.Rout - 40.0) < 1e-15 and c.Rin < 1.0 + (1 - c.a ** 2) ** 0.5  # inside the horizon
    assert c.gammafail == 100.0 * c.gammamax
