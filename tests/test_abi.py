"""The C-ABI library loads and exports every function include/fr_amd.h
declares (no device needed: nothing is launched)."""
import os
import re
import subprocess

import pytest

import first_repo_amd as fr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = []
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"\b(fr_[a-z0-9_]+)\s*\(", txt)
    return sorted(set(names))


def test_header_declares_entry_points():
    n = declared()
    assert "fr_sim_step" in n and "fr_box_poisson_solve" in n and len(n) >= 30


def test_library_exports_all_declared():
    lib = fr.lib()
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", fr.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r" T (fr_[a-z0-9_]+)", out))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in fr.lib().fr_version()


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(fr, "_lib", None)
    monkeypatch.setattr(fr, "LIB_PATH", "/nonexistent/libfr_amd.so")
    with pytest.raises(fr.FrError):
        fr.lib()
