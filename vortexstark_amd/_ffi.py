"""ctypes binding of libvortex_hip.so (declarations in include/vortex_hip.h).

The library is the product: this module only loads it and declares argument
types. There is no fallback; if the shared library is missing the import of
the package fails with an error that says how to build it.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvortex_hip.so")

u32 = C.c_uint32
i32 = C.c_int32
u64 = C.c_uint64
sz = C.c_size_t
vp = C.c_void_p
P32 = C.POINTER(C.c_uint32)

# name -> (restype, argtypes); mirrors include/vortex_hip.h
SIGNATURES = {
    # Part 1: reference kernel ABI (src/cuda/ffi.rs)
    "cuda_m31_add": (None, [vp, vp, vp, u32]),
    "cuda_m31_mul": (None, [vp, vp, vp, u32]),
    "cuda_device_sync": (None, []),
    "cuda_circle_ntt_evaluate": (None, [vp, vp, vp, P32, P32, u32, u32]),
    "cuda_circle_ntt_interpolate": (None, [vp, vp, vp, P32, P32, u32, u32]),
    "cuda_circle_ntt_evaluate_batch": (None, [vp, vp, vp, P32, P32, u32, u32, u32]),
    "cuda_circle_ntt_interpolate_batch": (None, [vp, vp, vp, P32, P32, u32, u32, u32]),
    "cuda_circle_ntt_layer": (None, [vp, vp, u32, u32, i32]),
    "cuda_bit_reverse_m31": (None, [vp, u32]),
    "cuda_fold_line_soa": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp, P32, u32]),
    "cuda_fold_circle_into_line_soa": (None, [vp, vp, vp, vp, vp, vp, vp, vp, vp, P32, P32, u32]),
    "cuda_compute_fold_twiddle_sources": (None, [u32, u32, u32, u32, vp, u32, u32, i32]),
    "cuda_batch_inverse_m31": (None, [vp, vp, u32]),
    "cuda_compute_fold_twiddle_sources_stream": (None, [u32, u32, u32, u32, vp, u32, u32, i32, vp]),
    "cuda_batch_inverse_m31_stream": (None, [vp, vp, u32, vp]),
    "cuda_compute_coset_points": (None, [u32, u32, u32, u32, vp, vp, u32]),
    "cuda_squash_x": (None, [vp, vp, u32]),
    "cuda_extract_and_squash": (None, [vp, vp, vp, u32]),
    "cuda_fibonacci_quotient": (None, [vp, vp, vp, vp, vp, P32, u32]),
    "cuda_zero_pad": (None, [vp, vp, u32, u32]),
    "cuda_fibonacci_quotient_chunk": (None, [vp, vp, vp, vp, vp, P32, u32, u32, u32]),
    "cuda_fibonacci_quotient_chunk_stream": (None, [vp, vp, vp, vp, vp, P32, u32, u32, u32, vp]),
    "cuda_merkle_hash_leaves": (None, [vp, vp, u32, u32]),
    "cuda_merkle_hash_nodes": (None, [vp, vp, u32]),
    "cuda_merkle_reduce_to_root": (None, [vp, vp, u32]),
    "cuda_merkle_commit_small_soa4": (None, [vp, vp, vp, vp, vp, u32]),
    "cuda_merkle_hash_leaves_merge_soa4": (None, [vp, vp, vp, vp, vp, u32]),
    "cuda_merkle_tiled_soa4": (None, [vp, vp, vp, vp, vp, u32]),
    "cuda_merkle_tiled_soa4_stream": (None, [vp, vp, vp, vp, vp, u32, vp]),
    "cuda_merkle_tiled_generic": (None, [vp, vp, u32, u32]),
    "cuda_accumulate_numerators": (None, [vp, vp, vp, vp, u32, u32, vp, vp, vp, vp]),
    "cuda_compute_quotients_combine": (None, [vp, vp, vp, vp, vp, vp, vp, vp, u32, vp, vp, u32, u32,
                                              vp, vp, vp, vp]),
    "cuda_accumulate_lift": (None, [vp, vp, u32, u32]),
    "vx_deep_create": (vp, [u32, u32]),
    "vx_deep_quotient": (i32, [vp, vp, vp, u32, u32, vp, vp, u32, vp, vp, vp, vp, vp, vp, vp]),
    "vx_deep_quotient_noalloc": (i32, [vp, vp, u32, u32, vp, vp, u32, vp, vp, vp, vp, vp, vp, vp]),
    "vx_deep_destroy": (None, [vp]),
    "vx_circle_domain_points": (i32, [u32, vp, vp, vp]),
    "cuda_grind_pow": (None, [vp, vp, u32, u64, u32]),
    "cuda_grind_pow_m31_output": (None, [vp, vp, u32, u64, u32]),
    "vx_grind_blake2s": (i32, [vp, u32, i32, u64, u64, vp, i32, vp]),
    "cuda_stwo_ntt_evaluate": (None, [vp, vp, u32]),
    "cuda_stwo_ntt_interpolate": (None, [vp, vp, u32]),
    "cuda_build_leaves_lifted": (None, [vp, vp, u32, u32, vp, u32]),
    "cuda_bit_reverse_qm31": (None, [vp, vp, u32, u32]),
    "cuda_pack_leaves": (None, [vp, vp, u32]),
    # Part 2: runtime
    "vx_get_last_error": (i32, []),
    "vx_get_last_error_message": (C.c_char_p, []),
    "vx_device_count": (i32, []),
    "vx_set_device": (i32, [i32]),
    "vx_malloc": (i32, [C.POINTER(vp), sz]),
    "vx_free": (i32, [vp]),
    "vx_malloc_host": (i32, [C.POINTER(vp), sz]),
    "vx_free_host": (i32, [vp]),
    "vx_memcpy_h2d": (i32, [vp, vp, sz]),
    "vx_memcpy_d2h": (i32, [vp, vp, sz]),
    "vx_memcpy_d2d": (i32, [vp, vp, sz]),
    "vx_memset": (i32, [vp, i32, sz]),
    "vx_device_synchronize": (i32, []),
    "vx_mem_get_info": (i32, [C.POINTER(sz), C.POINTER(sz)]),
    "vx_stream_create": (i32, [C.POINTER(vp)]),
    "vx_stream_synchronize": (i32, [vp]),
    "vx_stream_destroy": (i32, [vp]),
    "vx_set_stream": (None, [vp]),
    "vx_event_create": (i32, [C.POINTER(vp)]),
    "vx_event_record": (i32, [vp, vp]),
    "vx_event_elapsed_ms": (i32, [vp, vp, C.POINTER(C.c_float)]),
    "vx_event_destroy": (i32, [vp]),
    "vx_ntt_autotune": (i32, [u32, u32, vp]),
    "vx_measure_valu_peak": (i32, [vp, vp]),
    "vx_measure_hash_latency": (i32, [vp, vp]),
    "vx_ntt_plan_set": (i32, [u32, u32, vp, vp]),
    "vx_ntt_plan_get": (u32, [u32, vp, vp]),
    "vx_ntt_plan_reset": (None, []),
    "vx_ntt_lde_plan_set": (i32, [u32, u32, vp, vp, vp]),
    "vx_ntt_lde_plan_get": (u32, [u32, vp, vp, vp]),
    "vx_ntt_set_lde_bridge": (None, [i32]),
    "vx_get_stream": (vp, []),
    # Part 3: fused prover
    "vx_prover_create": (vp, [u32, u32, u32]),
    "vx_prover_create_batch": (vp, [u32, u32, u32, u32]),
    "vx_prover_capacity": (u32, [vp]),
    "vx_prover_device_bytes": (u64, [vp]),
    "vx_prover_estimate_bytes": (u64, [u32, u32, u32, u32]),
    "vx_merkle_autotune": (i32, [u32, u32, vp]),
    "vx_merkle_stage_plan": (i32, [u32]),
    "vx_merkle_stage_plan_set": (i32, [u32, i32]),
    "vx_merkle_stage_plan_reset": (None, []),
    "vx_prover_prove_batch_device": (i32, [vp, vp, vp, u32]),
    "vx_prover_prove_batch_host": (i32, [vp, vp, vp, u32]),
    "vx_prove_fibonacci_batch": (i32, [vp, vp, u32]),
    "vx_prover_set_graph": (i32, [vp, i32]),
    "vx_prover_proof_device_at": (vp, [vp, u32]),
    "vx_prover_proof_copy_at": (u64, [vp, u32, vp, u64]),
    "vx_prover_enqueue_copy": (i32, [vp, u32, vp, vp, vp]),
    "vx_prover_set_copy_kernel": (i32, [vp, i32]),
    "vx_krak_proof_size": (u64, [u32, u32, u32]),
    "vx_prover_destroy": (None, [vp]),
    "vx_prover_stream": (vp, [vp]),
    "vx_prover_proof_size": (u64, [vp]),
    "vx_prover_prove_device": (i32, [vp, vp, u32, u32]),
    "vx_prover_proof_device": (vp, [vp]),
    "vx_prover_proof_copy": (u64, [vp, vp, u64]),
    "vx_prove_fibonacci": (u64, [vp, u32, u32, vp, u64]),
    "vx_fibonacci_trace": (i32, [u32, u32, u32, vp, vp]),
    "vx_prover_eval_device": (vp, [vp]),
    "vx_prover_fri_layer_device": (vp, [vp, u32, u32]),
    "vx_prover_set_timing": (i32, [vp, i32]),
    "vx_prover_timeline_count": (u32, [vp]),
    "vx_prover_timeline_label": (C.c_char_p, [vp, u32]),
    "vx_prover_timeline_ms": (C.c_float, [vp, u32]),
    # Part 4: batched wide commitment
    "vx_stwo_twiddles": (i32, [u32, vp, vp, vp]),
    "vx_wide_create": (vp, [u32, P32, u32, u32]),
    "vx_wide_destroy": (None, [vp]),
    "vx_wide_stream": (vp, [vp]),
    "vx_wide_n_fri_layers": (u32, [vp]),
    "vx_wide_lifting_log_size": (u32, [vp]),
    "vx_wide_commit_device": (i32, [vp, vp]),
    "vx_wide_results": (i32, [vp, vp, vp, vp, vp]),
    "vx_wide_lde_device": (vp, [vp, u32]),
    "vx_wide_fri_layer_device": (vp, [vp, u32, u32]),
    "vx_wide_set_timing": (i32, [vp, i32]),
    "vx_wide_timeline_count": (u32, [vp]),
    "vx_wide_timeline_label": (C.c_char_p, [vp, u32]),
    "vx_wide_timeline_ms": (C.c_float, [vp, u32]),
    # Part 5: proof server
    "vx_server_create": (vp, [u64, u64, u32, u32]),
    "vx_server_destroy": (None, [vp]),
    "vx_server_submit": (C.c_int64, [vp, u32, u32, vp]),
    "vx_server_poll": (i32, [vp]),
    "vx_server_prewarm": (i32, [vp, u32, u32]),
    "vx_server_set_graphs": (i32, [vp, i32]),
    "vx_server_set_order": (i32, [vp, i32]),
    "vx_server_set_big_batches": (i32, [vp, u32, u32]),
    "vx_server_set_stagger": (i32, [vp, i32]),
    "vx_server_set_copy_kernel": (i32, [vp, i32]),
    "vx_server_wait_all": (i32, [vp]),
    "vx_server_inject_fault": (i32, [vp, C.c_int64, u32]),
    "vx_server_jobs": (u64, [vp]),
    "vx_server_completed": (u64, [vp]),
    "vx_server_job_status": (i32, [vp, C.c_int64]),
    "vx_server_job_latency_ms": (C.c_double, [vp, C.c_int64]),
    "vx_server_proof_copy": (u64, [vp, C.c_int64, u32, vp, u64]),
    "vx_server_release_job": (None, [vp, C.c_int64]),
    "vx_server_run_trace": (i32, [vp, u32, vp, vp, vp, vp, vp]),
    "vx_server_stats": (None, [vp, vp]),
}


def load(path: str = LIB_PATH) -> C.CDLL:
    if not os.path.exists(path):
        raise ImportError(
            f"libvortex_hip.so not found at {path}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C vortexstark_amd/csrc`"
        )
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
