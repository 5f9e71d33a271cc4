"""The C-ABI library loads and exports every entry point include/pggc.h
declares (no GPU calls)."""
import ctypes
import os
import re
import subprocess

from pennsylvaniagpugraphclustering_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = "".join(open(os.path.join(ROOT, "include", h)).read() for h in ("pggc.h", "pggc_io.h"))
    return sorted(set(re.findall(r"\b(pggc_\w+)\s*\(", src)))


def test_header_declares_expected_api():
    names = declared()
    for n in ["pggc_louvain", "pggc_leiden", "pggc_contract", "pggc_evaluate", "pggc_orderings", "pggc_sweep",
              "pggc_local_move", "pggc_leiden_refine", "pggc_cluster"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for n in declared():
        assert hasattr(lib, n), n
    out = subprocess.check_output(["nm", "-D", "--defined-only", _lib.LIB_PATH]).decode()
    exported = set(re.findall(r" T (pggc_\w+)", out))
    assert set(declared()) <= exported
    # the python binding covers the whole header
    assert set(declared()) == set(_lib.SIGNATURES)


def test_version_and_error_string_without_gpu():
    lib = _lib.load()
    assert lib.pggc_version() >= 100
    assert isinstance(lib.pggc_last_error(), bytes)


def test_null_context_is_rejected():
    lib = _lib.load()
    rc = lib.pggc_ctx_set_seed(None, 1, 0)
    assert rc == _lib.EINVAL
    assert b"null context" in lib.pggc_last_error()


def test_knob_ids_match_the_header():
    """The Python mirror's knob names map to the header's PGGC_KNOB_* ids,
    and every id below PGGC_KNOB_COUNT has a name."""
    from pennsylvaniagpugraphclustering_amd.clustering import Context
    src = open(os.path.join(ROOT, "include", "pggc.h")).read()
    ids = {m.group(1).lower(): int(m.group(2)) for m in re.finditer(r"#define PGGC_KNOB_(\w+) (\d+)", src)}
    count = ids.pop("count")
    assert sorted(ids.values()) == list(range(count))
    assert Context.KNOBS == ids
