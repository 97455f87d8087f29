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
                 "cuda_pack_leaves", "cuda_bit_reverse_qm31", "vx_wide_create", "vx_wide_commit_device",
                 "cuda_grind_pow", "cuda_grind_pow_m31_output", "vx_grind_blake2s",
                 "cuda_accumulate_numerators", "cuda_compute_quotients_combine", "cuda_accumulate_lift",
                 "vx_deep_create", "vx_deep_quotient", "vx_circle_domain_points"]:
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


def test_prover_estimate_bytes_host_only():
    """vx_prover_estimate_bytes is host arithmetic: unsupported parameters give
    0, the reference defaults apply to zero arguments, and a batch costs one
    arena plus one 32-byte call slot more per proof."""
    import vortexstark_amd

    est = vortexstark_amd.lib.vx_prover_estimate_bytes
    assert est(1, 80, 3, 1) == 0 and est(30, 80, 3, 1) == 0 and est(12, 80, 3, 0) == 0
    # 2^29 rows evaluate on the whole circle group (2^31 points): the largest prover
    assert 80e9 < est(29, 80, 3, 1) < 100e9
    assert est(12, 80, 11, 1) == 0 and est(12, 5000, 3, 1) == 0
    assert est(12, 0, 0, 1) == est(12, 80, 3, 1)
    one, two, nine = est(12, 80, 3, 1), est(12, 80, 3, 2), est(12, 80, 3, 9)
    arena = two - one
    assert arena > 0 and nine - one == 8 * arena
    for log_n in range(4, 29):
        assert est(log_n + 1, 80, 3, 1) > est(log_n, 80, 3, 1)
    # at least the trace, its LDE and FRI layer 0 of every proof
    for log_n in (10, 16, 20):
        E = 4 << log_n
        assert est(log_n, 80, 3, 2) - est(log_n, 80, 3, 1) >= 4 * (E // 4) + 4 * E + 16 * (E // 2)


def test_argument_errors_are_reported_without_a_device():
    """Entry points reject bad arguments on the host, before any device call,
    and report them through vx_get_last_error (never by aborting)."""
    import ctypes

    import vortexstark_amd as vx

    lib = vx.lib
    lib.vx_get_last_error()  # clear
    fake = ctypes.c_void_p(0x1000)
    # a PoW target above 128 bits can never be met
    assert lib.vx_grind_blake2s(fake, 129, 0, 0, 1 << 20, fake, 0, None) == -1
    assert lib.vx_get_last_error() != 0
    assert lib.vx_grind_blake2s(None, 10, 0, 0, 1, fake, 0, None) == -1
    assert lib.vx_get_last_error() != 0
    # DEEP objects: 0 columns, 5 points
    assert not lib.vx_deep_create(0, 2) and lib.vx_get_last_error() != 0
    assert not lib.vx_deep_create(8, 5) and lib.vx_get_last_error() != 0
    assert lib.vx_deep_quotient(None, None, None, 0, 0, None, None, 0, None, None, None, None, None, None,
                                None) == -1
    assert lib.vx_get_last_error() != 0
    # domain points outside 1..30
    assert lib.vx_circle_domain_points(31, fake, fake, None) == -1 and lib.vx_get_last_error() != 0
    # host-trace proving without a prover
    assert lib.vx_prover_prove_batch_host(None, None, None, 1) == -1 and lib.vx_get_last_error() != 0
    # servers whose protocol no prover supports are refused at creation
    # (more than 4096 queries, a last layer above 2^10)
    assert not lib.vx_server_create(0, 0, 5000, 3) and lib.vx_get_last_error() == -2
    assert not lib.vx_server_create(0, 0, 80, 11) and lib.vx_get_last_error() == -2
    assert lib.vx_server_inject_fault(None, 0, 0) == -1 and lib.vx_get_last_error() == -2
    assert lib.vx_get_last_error() == 0  # reading clears it


def test_krak_proof_size_counts_the_layout():
    """vx_krak_proof_size (host arithmetic over the KRAK v1 layout,
    stark_cli.rs:948-975) equals the length of the oracle's proof bytes."""
    import vortexstark_amd as vx
    import oracle_lib as O

    for log_n, nq, last in [(2, 80, 3), (4, 80, 3), (5, 1, 1), (8, 64, 4), (10, 100, 4), (12, 4096, 10), (13, 80, 3)]:
        assert vx.lib.vx_krak_proof_size(log_n, nq, last) == len(O.prove(5, 7, log_n, nq, last)), (log_n, nq, last)
    assert vx.lib.vx_krak_proof_size(12, 0, 0) == vx.lib.vx_krak_proof_size(12, 80, 3)
    assert vx.lib.vx_krak_proof_size(30, 80, 3) == 0 and vx.lib.vx_krak_proof_size(29, 80, 3) > 0
