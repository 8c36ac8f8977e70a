"""CPU checks of the C-ABI library: it is built, loads, exports every symbol
include/tmvp.h declares, and its host-side logic (parameter table, error
strings, argument checks, device probing) behaves without a GPU."""
import ctypes
import os
import re

import pytest

import oracle as O
import tmvp_amd as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tmvp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?(\w+)\s*\(", text, flags=re.M)))


def test_library_is_built_and_loads():
    assert os.path.exists(T.LIB_PATH), "libtmvp.so missing: run __graft_entry__.build()"
    lib = T.load_library()
    assert lib.tmvp_version() >= 1


def test_every_declared_symbol_is_exported():
    lib = T.load_library()
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) == set(T.EXPORTS)


@pytest.mark.parametrize("ps", range(6))
def test_param_table_matches_oracle(ps):
    mine = T.get_params(ps)
    ref = O.params(ps)
    for key in ("l", "eq", "ep", "et", "mu", "sable", "h1", "h2", "pk_bytes", "sk_bytes", "ct_bytes"):
        assert mine[key] == ref[key], key


def test_param_lookup_rejects_bad_ids():
    with pytest.raises(T.TmvpError):
        T.get_params(6)
    with pytest.raises(T.TmvpError):
        T.get_params(-1)


def test_strerror():
    lib = T.load_library()
    for code, text in T.ERRORS.items():
        assert lib.tmvp_strerror(code).decode() == text
    assert lib.tmvp_strerror(-99).decode() == "unknown error"


def test_null_context_rejected():
    lib = T.load_library()
    assert lib.tmvp_polymul(None, 0, None, None, None, 1, 13, 0, None) == -1
    assert lib.tmvp_matvec(None, 1, 0, 0, None, None, None, 1, 0, 0, None) == -1
    assert lib.tmvp_innerprod(None, 1, 0, 0, None, None, None, None, 1, 0, None) == -1
    assert lib.tmvp_destroy(None) == -1
    assert lib.tmvp_ctx_status(None, None) == -1
    assert lib.tmvp_init(None, 0) == -1


def test_init_without_gfx950_reports_nodev():
    # This container has no GPU: the library must report it, not crash.
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = T.load_library()
    h = ctypes.c_void_p()
    assert lib.tmvp_init(ctypes.byref(h), 0) == -6
    assert not h.value


def test_reference_named_entry_points_without_gpu():
    """saber_enc & co. keep the reference's void signatures; without a GPU they
    record the failure in tmvp_host_status() instead of crashing."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    lib = T.load_library()
    assert lib.tmvp_host_configure(0, 4, T.SABER, 8, T.SABLE2) == 0
    assert lib.tmvp_host_configure(0, 0, T.SABER, 8, T.SABLE2) == -1
    assert lib.tmvp_host_configure(0, 4, T.SABLE1, 8, T.SABLE2) == -1
    assert lib.tmvp_host_configure(0, 4, T.SABER, 8, T.FIRESABER) == -1
    pk, c, k = (ctypes.create_string_buffer(n) for n in (4 * 992, 4 * 1088, 4 * 32))
    lib.saber_enc(0, k, pk, c)
    assert lib.tmvp_host_status() == -6
    lib.saber_enc(9, k, pk, c)
    assert lib.tmvp_host_status() == -1
    assert lib.tmvp_host_release() == 0


def test_missing_library_fails_loudly(tmp_path):
    """With no built library there is no fallback: loading it raises, in a
    fresh process, before any operator can run."""
    import subprocess
    import sys

    code = (
        "import sys\n"
        "sys.path.insert(0, %r)\n"
        "import tmvp_amd as T\n"
        "try:\n"
        "    T.load_library(%r)\n"
        "except FileNotFoundError as e:\n"
        "    print('RAISED', e)\n"
        "    sys.exit(3)\n"
        "print('LOADED')\n"
    ) % (os.path.join(ROOT, "asfand-tmvp_amd"), str(tmp_path / "libtmvp.so"))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 3, (r.stdout, r.stderr)
    assert "RAISED" in r.stdout and "not built" in r.stdout
