"""The C-ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(ocfd_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    for n in ("ocfd_compute_dsdt", "ocfd_rk_stage", "ocfd_fill_boundary", "ocfd_estdt", "ocfd_advance",
              "ocfd_level_create"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import ocfd
    lib = ctypes.CDLL(str(ocfd.LIB_PATH))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_parm_initialize_host_only():
    import ocfd
    p = ocfd.default_parm()
    assert p.eos_gamma == 1.4 and abs(p.Rg - 8.31446261815324 / 28.97e-3) < 1e-12
    assert abs(p.kOverMu - p.cp / p.Pr) < 1e-12
