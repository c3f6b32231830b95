"""The C ABI library loads on the CPU and exports every function include/*.h declares
(no compute calls: there is no GPU here)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(mpmc_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "mpmc_flash_batch" in names and "mpmc_mixture_create" in names, names


def test_library_exports_every_declared_symbol():
    from mpmc_hnu_amd import native

    lib = native.library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(native.EXPORTED) <= set(declared_functions())


def test_options_struct_layout_matches_header():
    from mpmc_hnu_amd import native

    lib = native.library()
    o = native.FlashOptions()
    lib.mpmc_flash_default_options(ctypes.byref(o))
    assert o.fugacity_tolerance == 1e-11 and o.max_phases == 3 and o.precision == 0
    assert ctypes.sizeof(native.FlashOptions) == 8 + 4 + 4 + 4 + 12
    assert o.schedule == 0 and o.group_mixtures == 0
    assert lib.mpmc_flash_workspace_bytes(6, 1000) > 1000 * 8 * 6
    assert lib.mpmc_flash_workspace_bytes(13, 1000) == 0
    assert ctypes.sizeof(native.FlashOutputs) == 8 * 8


def test_invalid_arguments_are_rejected_without_device_work():
    from mpmc_hnu_amd import native

    lib = native.library()
    P = ctypes.c_void_p
    h = P()
    tc = (ctypes.c_double * 2)(300.0, 400.0)
    rc = lib.mpmc_mixture_create(13, tc, tc, tc, tc, ctypes.byref(h))
    assert rc == 1 and not h.value
    assert b"nc" in lib.mpmc_last_error()
    rc = lib.mpmc_flash_batch(None, 0, None, None, None, None, None, None)
    assert rc == 1
    # the phase-transition scan: null handles and outputs are argument errors, and the
    # summary is left untouched
    summary = native.ScanSummary(7, 5, 3)
    scan = native.ScanOutputs(None, None, None)
    rc = lib.mpmc_transition_scan(None, 4, None, None, None, None, None, 0, None, ctypes.byref(scan),
                                  ctypes.byref(summary), None)
    assert rc == 1 and b"transition_scan" in lib.mpmc_last_error()
    assert (summary.status, summary.first_stop, summary.proposals) == (7, 5, 3)
    rc = lib.mpmc_transition_scan_2p(None, None, None, None, None, None, None, None, 0, None, ctypes.byref(scan),
                                     ctypes.byref(summary), None)
    assert rc == 1
    n = ctypes.c_int64(-1)
    assert lib.mpmc_tpfa_grid_cells(None, ctypes.byref(n)) == 1 and n.value == -1
    assert ctypes.sizeof(native.ScanSummary) == 24 and ctypes.sizeof(native.ScanOutputs) == 24
