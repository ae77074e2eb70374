"""The C ABI without a GPU: the library loads, exports every entry point
include/turboransac.h declares, agrees with the ctypes struct layouts, and
fails with status codes (never exits) when no device is present."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from turboransac_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "turboransac.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(tr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = _abi.load_library()
    decl = declared_functions()
    assert len(decl) >= 25
    assert sorted(_abi.EXPORTED_SYMBOLS) == decl
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (tr_\w+)", out))
    for name in decl:
        assert name in exported, name
        assert getattr(lib, name) is not None


def test_engine_does_not_link_the_oracle():
    out = subprocess.run(["ldd", _abi.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    syms = subprocess.run(["nm", "-D", _abi.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in syms


def test_struct_layouts_match_the_header():
    src = r"""
    #include <stdio.h>
    #include <stddef.h>
    #include "turboransac.h"
    int main(void) {
      printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(tr_camera), sizeof(tr_ransac_options),
             sizeof(tr_bundle_options), sizeof(tr_pose_result), sizeof(tr_frame),
             sizeof(tr_server_options), sizeof(tr_debug_frame));
      printf("%zu %zu %zu\n", offsetof(tr_pose_result, status), offsetof(tr_frame, bundle),
             offsetof(tr_ransac_options, adaptive_stopping));
      printf("%zu %zu\n", sizeof(tr_bundle_stats), offsetof(tr_pose_result, lm_grad_norm));
      return 0;
    }
    """
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(v) for v in out]
    want = [C.sizeof(t) for t in (_abi.tr_camera, _abi.tr_ransac_options, _abi.tr_bundle_options,
                                  _abi.tr_pose_result, _abi.tr_frame, _abi.tr_server_options,
                                  _abi.tr_debug_frame)]
    assert sizes[:7] == want
    assert sizes[7] == _abi.tr_pose_result.status.offset
    assert sizes[8] == _abi.tr_frame.bundle.offset
    assert sizes[9] == _abi.tr_ransac_options.adaptive_stopping.offset
    assert sizes[10] == C.sizeof(_abi.tr_bundle_stats)
    assert sizes[11] == _abi.tr_pose_result.lm_grad_norm.offset


def test_default_options_match_reference_defaults():
    lib = _abi.load_library()
    r = _abi.tr_ransac_options()
    lib.tr_ransac_options_default(C.byref(r))
    ref = _abi.default_ransac_options()
    for f, _ in r._fields_:
        assert getattr(r, f) == getattr(ref, f), f
    b = _abi.tr_bundle_options()
    lib.tr_bundle_options_default(C.byref(b))
    refb = _abi.default_bundle_options()
    for f, _ in b._fields_:
        assert getattr(b, f) == getattr(refb, f), f
    # utils.h:14-60 values
    assert (r.max_iterations, r.min_iterations, r.max_reproj_error) == (100000, 1000, 12.0)
    assert (b.max_iterations, b.loss_scale, b.initial_lambda) == (100, 1.0, 1e-3)


def test_status_strings_and_abi_version():
    lib = _abi.load_library()
    assert lib.tr_abi_version() == 5 == _abi.ABI_VERSION
    for code, name in _abi.STATUS_NAMES.items():
        assert lib.tr_status_string(code).decode() == name


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_is_a_status_not_an_exit():
    lib = _abi.load_library()
    h = C.c_void_p()
    st = lib.tr_context_create(0, C.byref(h))
    assert st == _abi.TR_ERR_NO_DEVICE
    assert not h.value


def test_null_arguments_are_rejected():
    lib = _abi.load_library()
    assert lib.tr_context_create(0, None) == _abi.TR_ERR_INVALID_ARGUMENT
    assert lib.tr_estimate_absolute_pose(None, None, None, 0, None, None, None, None, None) \
        == _abi.TR_ERR_INVALID_ARGUMENT
    assert lib.tr_batch_launch(None) == _abi.TR_ERR_INVALID_ARGUMENT


def test_missing_library_raises():
    with pytest.raises(RuntimeError, match="not built"):
        _abi.load_library("/nonexistent/libturboransac_amd.so")


def test_cpp_wrapper_example_links():
    exe = os.path.join(ROOT, "turboransac_amd", "lib", "tr_example")
    out = subprocess.run([exe, "--abi"], capture_output=True, text=True, check=True).stdout
    assert out.strip() == f"turboransac abi {_abi.ABI_VERSION}"
