"""CPU-side checks of the C ABI library: it loads without a GPU and exports
every function include/vortex_hip.h declares (no compute calls here)."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "vortex_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b((?:cuda|vx)_[A-Za-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_reference_symbols():
    names = declared_functions()
    for must in ["cuda_circle_ntt_evaluate", "cuda_circle_ntt_interpolate", "cuda_merkle_tiled_soa4",
                 "cuda_fold_line_soa", "cuda_fold_circle_into_line_soa", "cuda_fibonacci_quotient",
                 "cuda_batch_inverse_m31", "vx_prover_create", "vx_prover_prove_device",
                 "cuda_stwo_ntt_evaluate", "cuda_stwo_ntt_interpolate", "cuda_build_leaves_lifted",
                 "cuda_pack_leaves", "cuda_bit_reverse_qm31", "vx_wide_create", "vx_wide_commit_device"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    import vortexstark_amd

    lib = ctypes.CDLL(vortexstark_amd._ffi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared symbol also has a ctypes signature in the package
    assert set(declared_functions()) <= set(vortexstark_amd._ffi.SIGNATURES)


def test_no_device_needed_to_load():
    import vortexstark_amd

    assert vortexstark_amd.lib.vx_device_count() >= 0
