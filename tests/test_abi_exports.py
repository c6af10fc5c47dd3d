"""The C-ABI library loads and exports every entry point include/*.h declares."""
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    inc = os.path.join(ROOT, "include", "quarry")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"QR_API\s+[\w\s\*]+?\b(qr_\w+)\s*\(", text))
    return names


def test_library_exports_header_symbols(qabi):
    names = declared()
    assert len(names) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", qabi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = sorted(names - exported)
    assert not missing, missing
    # nothing but the C ABI is exported from the product's namespace
    assert not any(s.startswith("_ZN2qr") for s in exported)


def test_ctypes_binding_covers_header(qabi):
    lib = qabi.lib()
    for name in declared():
        assert hasattr(lib, name), name


def test_no_device_reports_cleanly(qabi):
    # Without a GPU (this container) describe returns None rather than crashing.
    d = qabi.describe()
    assert d is None or "gfx" in d


def test_process_wide_knobs_check_their_range(qabi):
    # Host-only setters: in-range values are taken, out-of-range ones are refused
    # with the invalid-argument code; the defaults are restored afterwards.
    lib = qabi.lib()
    try:
        for v in (0, 1, 2):
            assert lib.qr_set_build_phase_timing(v) == 0
        for v in (-1, 3):
            assert lib.qr_set_build_phase_timing(v) != 0
        for v in (0, 5, 32):
            assert lib.qr_set_frontier_cap(v) == 0
        for v in (-1, 33):
            assert lib.qr_set_frontier_cap(v) != 0
    finally:
        assert lib.qr_set_build_phase_timing(1) == 0
        assert lib.qr_set_frontier_cap(32) == 0
