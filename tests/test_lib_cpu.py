"""CPU-side checks of the C-ABI library: it loads and exports every function
that include/fireball_hip.h declares (no GPU calls)."""
from fireball_hip import lib


def test_library_exports_every_header_symbol():
    L = lib.lib()
    names = lib.header_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_null_context_is_rejected():
    L = lib.lib()
    assert L.fb_destroy(None) == -1
    assert L.fb_assemble_S(None) == -1
    assert L.fb_build_lists(None, None, None) == -1
