"""The results.out comparator (gpu_determ_amd.compare), on the CPU: the
reference's RMS figures (test.py:1-17), the thresholds, and the failure
cases.  The reference file is rebuilt from the golden rows in
tests/golden/reference_rows.tsv, whose slot / cell pairs give the integer
columns (16^3 box, 256 cells: index = slot * 256 + cell)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from gpu_determ_amd import compare as CMP

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def golden_rows():
    return np.genfromtxt(os.path.join(GOLD, "reference_rows.tsv"), names=True, delimiter="\t")


def write_results(path, g, vals=None):
    """results.out rows (determ.cu:146-160 format) from the golden rows, with
    optional replacement value columns [n, 5]."""
    idx = (g["slot"] * 256 + g["cell"]).astype(int)
    v = np.stack([g[c] for c in CMP.VALUE_COLS], axis=1) if vals is None else vals
    with open(path, "w") as fh:
        for k, ix in enumerate(idx):
            fh.write("%d %d %d %d %g %g %g %g %g\n" % ((ix, ix // 256, (ix % 256) // 16, ix % 16) + tuple(v[k])))
    return v


def test_rebuilt_reference_rows(tmp_path):
    """The first rows of the reference's true_results.out, as data."""
    write_results(tmp_path / "t.out", golden_rows())
    lines = (tmp_path / "t.out").read_text().splitlines()
    assert len(lines) == 512
    assert lines[0] == "0 0 0 0 7.44729e-39 844909 2.87827e-35 -6.26349e-33 -217.613"
    assert lines[1] == "1 0 0 1 7.44729e-39 659.371 2.87827e-35 2.38721e-35 0.829393"


def test_rms_figures_are_test_py(tmp_path):
    """test.py accumulates (x - y)^2 over the rows per value column and prints
    sqrt(sum / rows): checked on a perturbed copy against that loop."""
    g = golden_rows()
    ref = write_results(tmp_path / "a.out", g)
    rng = np.random.default_rng(3)
    got = ref * (1 + 1e-3 * rng.standard_normal(ref.shape))
    write_results(tmp_path / "b.out", g, got)
    _, rv = CMP.read_results(tmp_path / "a.out")
    _, gv = CMP.read_results(tmp_path / "b.out")
    acc = [0.0] * 5
    for i in range(len(rv)):
        for k in range(5):
            acc[k] += (rv[i, k] - gv[i, k]) ** 2
    want = [(a / len(rv)) ** 0.5 for a in acc]
    assert np.allclose(CMP.rms_line(rv, gv), want, rtol=1e-12, atol=0)


def test_identical_and_thresholds(tmp_path):
    g = golden_rows()
    ref = write_results(tmp_path / "a.out", g)
    ok, rep = CMP.compare(tmp_path / "a.out", tmp_path / "a.out", out=open(os.devnull, "w"))
    assert ok and all(c["failed_rows"] == 0 for c in rep["columns"].values())
    # one IC row off by 1.5e-4 relative: inside the default 2e-4
    got = ref.copy()
    got[100, 2] *= 1 + 1.5e-4
    write_results(tmp_path / "b.out", g, got)
    ok, rep = CMP.compare(tmp_path / "a.out", tmp_path / "b.out", out=open(os.devnull, "w"))
    assert ok and rep["columns"]["IC"]["worst_row"] == 100
    # ... and outside a tighter one
    ok, rep = CMP.compare(tmp_path / "a.out", tmp_path / "b.out", rtol={"IC": 1e-4}, out=open(os.devnull, "w"))
    assert not ok and rep["columns"]["IC"]["failed_rows"] == 1
    # F off by 3e-5 on two rows: fails the default 1e-5
    got = ref.copy()
    got[[5, 7], 0] *= 1 + 3e-5
    write_results(tmp_path / "c.out", g, got)
    ok, rep = CMP.compare(tmp_path / "a.out", tmp_path / "c.out", out=open(os.devnull, "w"))
    assert not ok and rep["columns"]["F"]["failed_rows"] == 2
    # an unchecked column may differ
    got = ref.copy()
    got[:, 4] += 1.0
    write_results(tmp_path / "d.out", g, got)
    ok, _ = CMP.compare(tmp_path / "a.out", tmp_path / "d.out", out=open(os.devnull, "w"))
    assert ok
    ok, _ = CMP.compare(tmp_path / "a.out", tmp_path / "d.out", check=("ACCURACY",), out=open(os.devnull, "w"))
    assert not ok


def test_structure_failures(tmp_path):
    g = golden_rows()
    write_results(tmp_path / "a.out", g)
    lines = (tmp_path / "a.out").read_text().splitlines()
    (tmp_path / "short.out").write_text("\n".join(lines[:-1]) + "\n")
    ok, rep = CMP.compare(tmp_path / "a.out", tmp_path / "short.out", out=open(os.devnull, "w"))
    assert not ok and rep["error"] == "row count"
    t = lines[3].split()
    t[3] = str(int(t[3]) + 1)
    (tmp_path / "swap.out").write_text("\n".join(lines[:3] + [" ".join(t)] + lines[4:]) + "\n")
    ok, rep = CMP.compare(tmp_path / "a.out", tmp_path / "swap.out", out=open(os.devnull, "w"))
    assert not ok and rep["error"] == "integer columns"
    (tmp_path / "bad.out").write_text("0 0 0 0 1.0 2.0\n")
    with pytest.raises(ValueError):
        CMP.read_results(tmp_path / "bad.out")


def test_command_line(tmp_path):
    g = golden_rows()
    ref = write_results(tmp_path / "a.out", g)
    got = ref.copy()
    got[10, 1] *= 1 + 5e-5
    write_results(tmp_path / "b.out", g, got)
    cmd = [sys.executable, "-m", "gpu_determ_amd.compare", str(tmp_path / "a.out")]
    r = subprocess.run(cmd + [str(tmp_path / "a.out")], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 0 and r.stdout.splitlines()[-1] == "PASS"
    r = subprocess.run(cmd + [str(tmp_path / "b.out")], capture_output=True, text=True, cwd=ROOT, timeout=120)
    assert r.returncode == 1 and "DC" in r.stdout and r.stdout.splitlines()[-1] == "FAIL"
    r = subprocess.run(cmd + [str(tmp_path / "b.out"), "--rtol", "DC=1e-4"], capture_output=True, text=True,
                       cwd=ROOT, timeout=120)
    assert r.returncode == 0
