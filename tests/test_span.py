"""Exact row spans (csrc/pbr_span.h) against per-pixel int64/int128 edge
tests on the host: random snapped triangles from sub-pixel size to the
guard-band extreme, horizontal and vertical edges, every fill-rule bias, and
the float estimate perturbed by up to +-0.24.  A planted mistake (the
downward fix-up step removed) must be caught."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "native" / "span_check.cpp"
HDR = ROOT / "pybatchrender_amd" / "csrc" / "pbr_span.h"


def _build(tmp: Path, header: Path) -> Path:
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    src = SRC.read_text().replace("../../pybatchrender_amd/csrc/pbr_span.h", str(header))
    cpp = tmp / "span_check.cpp"
    cpp.write_text(src)
    exe = tmp / "span_check"
    subprocess.run([cxx, "-O2", "-std=c++17", "-o", str(exe), str(cpp)], check=True)
    return exe


def test_row_span_matches_pixel_tests(tmp_path):
    exe = _build(tmp_path, HDR)
    res = subprocess.run([str(exe), "20000"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout
    assert res.stdout.startswith("ok ")
    assert int(res.stdout.split()[1]) > 1_000_000


def test_planted_fixup_error_is_caught(tmp_path):
    bad = tmp_path / "pbr_span_bad.h"
    text = HDR.read_text()
    assert "g -= (r > 0.0) ? 1 : 0;" in text
    bad.write_text(text.replace("g -= (r > 0.0) ? 1 : 0;", ""))
    exe = _build(tmp_path, bad)
    res = subprocess.run([str(exe), "20000"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 1 and res.stdout.startswith("mismatch")
