"""C ABI checks that need no GPU: the library loads, exports every symbol
include/bfe.h declares, and its error paths report through status codes."""
import ctypes as C
import re

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import PKG, ROOT


def _declared():
    hdr = (ROOT / "include" / "bfe.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:bfe_status|const char\*)\s+(bfe_\w+)\s*\(", hdr, re.M)))


def test_header_declares_expected_entry_points():
    names = _declared()
    assert {"bfe_create", "bfe_run", "bfe_run_batched", "bfe_destroy", "bfe_last_error"} <= set(names)


def test_library_exports_every_declared_symbol():
    from bristolfe_v2_amd import _lib
    lib = C.CDLL(str(_lib.LIB_PATH))
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_declared()) == set(_lib.EXPORTS)


def test_structs_match_header_sizes():
    from bristolfe_v2_amd import _lib
    hdr = (ROOT / "include" / "bfe.h").read_text()
    body = re.search(r"typedef struct bfe_run_stats \{(.*?)\} bfe_run_stats;", hdr, re.S).group(1)
    fields = re.findall(r"^\s*(?:int64_t|double|int32_t)\s+(\w+);", body, re.M)
    assert fields == [f for f, _ in _lib.BfeRunStats._fields_]
    body = re.search(r"typedef struct bfe_info \{(.*?)\} bfe_info;", hdr, re.S).group(1)
    fields = re.findall(r"^\s*(?:int64_t|double|int32_t)\s+(\w+);", body, re.M)
    assert fields == [f for f, _ in _lib.BfeInfo._fields_]


def test_null_handle_and_bad_arguments_return_status():
    from bristolfe_v2_amd import _lib
    lib = _lib.load()
    st = _lib.BfeRunStats()
    assert lib.bfe_last_run_stats(None, C.byref(st)) == _lib.BFE_ERR_ARG
    assert b"NULL" in lib.bfe_last_error()
    assert lib.bfe_destroy(None) == _lib.BFE_OK
    h = C.c_void_p()
    # n = 0 is rejected before any device call
    assert lib.bfe_create(0, None, None, None, None, None, None, None, 1e-3, 0, 0, 0, C.byref(h)) \
        == _lib.BFE_ERR_ARG


def test_create_fails_loudly_without_a_device():
    """On a host without a HIP device the product path raises; it never
    falls back to the CPU."""
    from bristolfe_v2_amd import _lib
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import __graft_entry__ as g; g.load_package();"
            "from bristolfe_v2_amd import _lib; import numpy as np, scipy.sparse as sp;"
            "n = 0\n"
            "import ctypes\n"
            "try:\n"
            "    _lib.Plan.from_scipy(sp.eye(3, format='csr'), np.ones(3), None, 1e-3, 0, 0)\n"
            "    print('DEVICE')\n"
            "except _lib.BfeError as e:\n"
            "    print('STATUS', e.status)\n") % str(ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    if "DEVICE" in out.stdout:
        pytest.skip("a HIP device is present")
    assert "STATUS 2" in out.stdout, out.stdout + out.stderr


def test_solver_rejects_prescribed_displacements_and_bad_modes():
    K = sp.eye(2, format="csr")
    t = np.arange(5) * 0.1
    with pytest.raises(NotImplementedError):
        PKG.fn_explicit_dynamic_solver_v6(K, K * 0, K, t, [1], np.ones(5), [1], np.ones(5), [1], np.inf)
    with pytest.raises(ValueError):
        PKG.make_plan(K, K * 0, K, 0.1, "vel at next step")
