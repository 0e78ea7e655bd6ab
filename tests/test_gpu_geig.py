"""GPU side of the GEIG v1 tooling (geig_tool.py make / check): the reference's
fixtures re-solved by the library against the LAPACK outputs they store, and a
`make` file that reproduces the reference's inputs for its seed."""
import importlib.util
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

PKG = os.path.join(ROOT, "sunny-batched-geneig_amd")


@pytest.fixture(scope="module")
def tool():
    spec = importlib.util.spec_from_file_location("geig_tool_under_test", os.path.join(PKG, "geig_tool.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.mark.parametrize("name,form", [("ref_n16_random.bin", "plain"), ("ref_n16_nearid.bin", "plain"),
                                       ("ref_n54_random.bin", "plain"), ("ref_n96_nearid.bin", "plain"),
                                       ("ref_n16_random.bin", "precision")])
def test_check_reference_fixture(tool, device, name, form):
    G = tool._geig()
    d = G.read(os.path.join(GOLDEN, name))
    W, U, info = tool.solve_gpu(d["H"], d["S"], form)
    eig, ph = tool.compare(W, U, d["W"], d["U"], d["S"])
    res = tool.rel_residual(d["H"], d["S"], W, U)
    print("%s %s: eig %.2e phase %.2e residual %.2e" % (name, form, eig, ph, res), flush=True)
    assert (info == 0).all()
    assert eig <= 1e-12 and ph <= 1e-10 and res <= 1e-13
    assert tool.main(["check", os.path.join(GOLDEN, name), "--form", form]) == 0


def test_make_reproduces_reference_inputs(tool, device, tmp_path, capsys):
    out = str(tmp_path / "n16.bin")
    assert tool.main(["make", "--n", "16", "--batch", "2", "--seed", "42", "--out", out]) == 0
    line = capsys.readouterr().out
    assert "N=16, B=2, seed=42, s_mode=random" in line and "max_rel_residual" in line
    G = tool._geig()
    mine, ref = G.read(out), G.read(os.path.join(GOLDEN, "ref_n16_random.bin"))
    assert mine["H"].tobytes() == ref["H"].tobytes() and mine["S"].tobytes() == ref["S"].tobytes()
    scale = np.maximum(np.abs(ref["W"]).max(axis=1, keepdims=True), 1.0)
    assert (np.abs(mine["W"] - ref["W"]) / scale).max() <= 1e-12
