"""The C ABI library loads and exports every entry point include/*.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

from mcluminescence_amd import _abi

HEADERS = sorted((Path(__file__).resolve().parent.parent / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(mcl_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(str(_abi.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 24
    for n in names:
        assert hasattr(lib, n), n


def test_binding_covers_the_header():
    assert set(declared_functions()) == set(_abi.EXPORTED)


def test_abi_version_and_host_geometry():
    lib = _abi.library()
    assert lib.mcl_abi_version() == _abi.ABI_VERSION
    from mcluminescence_amd import workloads as W
    g = _abi.geometry(W.iso_decay_4096x400()[0])
    assert (g.n_e0, g.n_h) == (400, 691)
