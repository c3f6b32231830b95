"""ctypes binding of the C ABI in include/mpmc_hnu_amd.h.

The library is the in-tree gfx950 build ``mpmc_hnu_amd/libmpmc_hnu_amd.so``. There is
no fallback: when it is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

# MPMC_HNU_AMD_LIBRARY may name another build of the same library (A/B timing of
# compile variants); by default the in-tree build is loaded.
LIB_PATH = Path(os.environ.get("MPMC_HNU_AMD_LIBRARY", Path(__file__).resolve().parent / "libmpmc_hnu_amd.so"))

MAX_COMPONENTS = 12

# include/mpmc_hnu_amd.h symbols, each checked by tests/test_abi.py
EXPORTED = (
    "mpmc_mixture_create",
    "mpmc_mixture_destroy",
    "mpmc_mixture_components",
    "mpmc_flash_default_options",
    "mpmc_flash_batch",
    "mpmc_flash_batch_ws",
    "mpmc_flash_batch_static",
    "mpmc_flash_workspace_bytes",
    "mpmc_flash_multi_max_mixtures",
    "mpmc_flash_multi_workspace_bytes",
    "mpmc_flash_multi",
    "mpmc_flash_batch_hinted",
    "mpmc_flash_batch_warm",
    "mpmc_continuation_default_options",
    "mpmc_pt_continuation_workspace_bytes",
    "mpmc_pt_continuation",
    "mpmc_flash_plan_create",
    "mpmc_flash_plan_launch",
    "mpmc_flash_plan_info",
    "mpmc_flash_plan_destroy",
    "mpmc_phase_batch",
    "mpmc_rr_batch",
    "mpmc_closure_default_options",
    "mpmc_closure_batch",
    "mpmc_tpfa_grid_create",
    "mpmc_tpfa_grid_destroy",
    "mpmc_tpfa_grid_adjacency",
    "mpmc_tpfa_grid_cells",
    "mpmc_tpfa_grid_max_degree",
    "mpmc_assembly_default_model",
    "mpmc_assembly_2p",
    "mpmc_inventory_2p",
    "mpmc_assembly_reduce_2p",
    "mpmc_assembly_plan_create",
    "mpmc_assembly_plan_launch",
    "mpmc_assembly_plan_domain_error",
    "mpmc_assembly_plan_destroy",
    "mpmc_transition_scan",
    "mpmc_transition_scan_2p",
    "mpmc_server_default_options",
    "mpmc_server_create",
    "mpmc_server_destroy",
    "mpmc_server_slots",
    "mpmc_server_device_bytes",
    "mpmc_server_submit",
    "mpmc_server_submit_hinted",
    "mpmc_server_wait",
    "mpmc_server_poll",
    "mpmc_server_last_error",
    "mpmc_last_error",
)


class ScanOutputs(ctypes.Structure):
    _fields_ = [("code", ctypes.c_void_p), ("target_count", ctypes.c_void_p), ("target_density", ctypes.c_void_p)]


class ScanSummary(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("first_stop", ctypes.c_int64), ("proposals", ctypes.c_int64)]


class FlashOptions(ctypes.Structure):
    _fields_ = [
        ("fugacity_tolerance", ctypes.c_double),
        ("max_phases", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("schedule", ctypes.c_int32),
        ("group_mixtures", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 2),
    ]


class FlashOutputs(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_void_p),
        ("phase_count", ctypes.c_void_p),
        ("beta", ctypes.c_void_p),
        ("zfactor", ctypes.c_void_p),
        ("branch", ctypes.c_void_p),
        ("x", ctypes.c_void_p),
        ("ln_phi", ctypes.c_void_p),
        ("counters", ctypes.c_void_p),
    ]


class ContinuationOptions(ctypes.Structure):
    _fields_ = [
        ("carry_compositions", ctypes.c_int32),
        ("carry_three_phase", ctypes.c_int32),
        ("nfallback_initial", ctypes.c_int32),
        ("nfallback_final", ctypes.c_int32),
        ("three_phase_start", ctypes.c_int32),
        ("fallback_initial", (ctypes.c_double * 12) * 3),
        ("fallback_final", (ctypes.c_double * 12) * 3),
        ("start", (ctypes.c_double * 12) * 3),
        ("start_seed", ctypes.c_double * 2),
    ]


class FlashSegment(ctypes.Structure):
    _fields_ = [
        ("mixture", ctypes.c_void_p),
        ("ncell", ctypes.c_int64),
        ("p", ctypes.c_void_p),
        ("t", ctypes.c_void_p),
        ("z", ctypes.c_void_p),
        ("out", FlashOutputs),
    ]


class ClosureOptions(ctypes.Structure):
    _fields_ = [
        ("minimum_derivative_phase_fraction", ctypes.c_double),
        ("minimum_reciprocal_condition", ctypes.c_double),
        ("fugacity_tolerance", ctypes.c_double),
        ("reserved", ctypes.c_int32 * 2),
    ]


class ClosureOutputs(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_void_p),
        ("primal", ctypes.c_void_p),
        ("jac", ctypes.c_void_p),
        ("log_k_jac", ctypes.c_void_p),
        ("diag", ctypes.c_void_p),
    ]


class ServerOptions(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("max_phases", ctypes.c_int32),
        ("reserved0", ctypes.c_int32),
        ("fugacity_tolerance", ctypes.c_double),
        ("chunk_cells", ctypes.c_int64),
        ("small_job_cells", ctypes.c_int64),
        ("hbm_cap_bytes", ctypes.c_size_t),
        ("stream", ctypes.c_void_p),
    ]


class HostOutputs(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("status", "phase_count", "beta", "zfactor", "branch", "x", "ln_phi")]


class AssemblyModel(ctypes.Structure):
    _fields_ = [("mu_a0", ctypes.c_double), ("mu_a1", ctypes.c_double), ("cp", ctypes.c_double),
                ("t0", ctypes.c_double), ("cv_rock", ctypes.c_double), ("molar_mass", ctypes.c_double * 12)]


class AssemblyOutputs(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("cell_status", "residual", "diag", "offdiag")]


class ReducedOutputs(ctypes.Structure):
    _fields_ = [(k, ctypes.c_void_p) for k in ("status", "residual", "diag", "offdiag", "back")]


class MpmcError(RuntimeError):
    pass


_lib = None


def library() -> ctypes.CDLL:
    """Load the gfx950 library; raises MpmcError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise MpmcError(f"{LIB_PATH} is not built: run `python -m mpmc_hnu_amd.build`")
    # torch carries its own HIP runtime under the same soname as the system one. Whichever
    # is loaded first serves the whole process, and torch cannot initialise its devices on
    # the other; importing torch first keeps one runtime shared by torch's tensors and
    # this library's kernels.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.c_void_p
    lib.mpmc_mixture_create.argtypes = [ctypes.c_int, P, P, P, P, ctypes.POINTER(P)]
    lib.mpmc_mixture_create.restype = ctypes.c_int
    lib.mpmc_mixture_destroy.argtypes = [P]
    lib.mpmc_mixture_components.argtypes = [P]
    lib.mpmc_flash_default_options.argtypes = [ctypes.POINTER(FlashOptions)]
    lib.mpmc_flash_default_options.restype = None
    lib.mpmc_flash_batch.argtypes = [P, ctypes.c_int64, P, P, P, ctypes.POINTER(FlashOptions),
                                     ctypes.POINTER(FlashOutputs), P]
    lib.mpmc_flash_batch_static.argtypes = [P, ctypes.c_int64, P, P, P, ctypes.POINTER(FlashOptions),
                                            ctypes.POINTER(FlashOutputs), P]
    lib.mpmc_flash_batch_ws.argtypes = [P, ctypes.c_int64, P, P, P, ctypes.POINTER(FlashOptions),
                                        ctypes.POINTER(FlashOutputs), P, ctypes.c_size_t, P]
    lib.mpmc_flash_batch_hinted.argtypes = [P, ctypes.c_int64, P, P, P, ctypes.c_int, P, ctypes.POINTER(FlashOptions),
                                            ctypes.POINTER(FlashOutputs), P, ctypes.c_size_t, P]
    lib.mpmc_flash_batch_warm.argtypes = [P, ctypes.c_int64, P, P, P, ctypes.c_int, P, ctypes.c_int, P,
                                          ctypes.POINTER(FlashOptions), ctypes.POINTER(FlashOutputs), P,
                                          ctypes.c_size_t, P]
    lib.mpmc_continuation_default_options.argtypes = [ctypes.POINTER(ContinuationOptions)]
    lib.mpmc_continuation_default_options.restype = None
    lib.mpmc_pt_continuation_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_int64]
    lib.mpmc_pt_continuation_workspace_bytes.restype = ctypes.c_size_t
    lib.mpmc_pt_continuation.argtypes = [P, ctypes.c_int64, ctypes.c_int, P, P, P, ctypes.POINTER(ContinuationOptions),
                                         ctypes.POINTER(FlashOptions), ctypes.POINTER(FlashOutputs), P, P,
                                         ctypes.c_size_t, P]
    lib.mpmc_flash_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_int64]
    lib.mpmc_flash_workspace_bytes.restype = ctypes.c_size_t
    lib.mpmc_flash_multi_max_mixtures.argtypes = [ctypes.c_int]
    lib.mpmc_flash_multi_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64]
    lib.mpmc_flash_multi_workspace_bytes.restype = ctypes.c_size_t
    lib.mpmc_flash_multi.argtypes = [ctypes.c_int, P, P, P, P, P, ctypes.POINTER(FlashOptions),
                                     ctypes.POINTER(FlashOutputs), P, ctypes.c_size_t, P]
    lib.mpmc_flash_plan_create.argtypes = [ctypes.c_int, ctypes.POINTER(FlashSegment), ctypes.POINTER(FlashOptions),
                                           ctypes.POINTER(P)]
    lib.mpmc_flash_plan_launch.argtypes = [P, P]
    lib.mpmc_flash_plan_info.argtypes = [P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_size_t)]
    lib.mpmc_flash_plan_destroy.argtypes = [P]
    lib.mpmc_phase_batch.argtypes = [P, ctypes.c_int64, P, P, P, P, P, P, P, P, P]
    lib.mpmc_rr_batch.argtypes = [ctypes.c_int, ctypes.c_int64] + [P] * 9
    lib.mpmc_closure_default_options.argtypes = [ctypes.POINTER(ClosureOptions)]
    lib.mpmc_closure_default_options.restype = None
    lib.mpmc_closure_batch.argtypes = [P, ctypes.c_int64, P, P, P, ctypes.POINTER(FlashOutputs),
                                       ctypes.POINTER(ClosureOptions), ctypes.POINTER(ClosureOutputs), P]
    lib.mpmc_tpfa_grid_create.argtypes = [ctypes.c_int64, ctypes.c_int64] + [P] * 7 + [ctypes.POINTER(P)]
    lib.mpmc_tpfa_grid_destroy.argtypes = [P]
    lib.mpmc_tpfa_grid_adjacency.argtypes = [P, ctypes.POINTER(P), ctypes.POINTER(P)]
    lib.mpmc_tpfa_grid_max_degree.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
    lib.mpmc_tpfa_grid_cells.argtypes = [P, ctypes.POINTER(ctypes.c_int64)]
    lib.mpmc_assembly_default_model.argtypes = [ctypes.c_int, ctypes.POINTER(AssemblyModel)]
    lib.mpmc_assembly_default_model.restype = None
    lib.mpmc_assembly_2p.argtypes = [P, P, ctypes.POINTER(AssemblyModel), P, P, P, ctypes.c_double,
                                     ctypes.POINTER(AssemblyOutputs), ctypes.POINTER(ctypes.c_int32), P]
    lib.mpmc_assembly_reduce_2p.argtypes = [P, ctypes.c_int, P, ctypes.POINTER(AssemblyOutputs),
                                            ctypes.POINTER(ReducedOutputs), P]
    lib.mpmc_inventory_2p.argtypes = [P, P, ctypes.POINTER(AssemblyModel), P, P, P, P,
                                      ctypes.POINTER(ctypes.c_int32), P]
    lib.mpmc_assembly_plan_create.argtypes = [P, P, ctypes.POINTER(AssemblyModel), P, P, P, ctypes.c_double,
                                              ctypes.POINTER(AssemblyOutputs), ctypes.POINTER(P)]
    lib.mpmc_assembly_plan_launch.argtypes = [P, P]
    lib.mpmc_assembly_plan_domain_error.argtypes = [P, P, ctypes.POINTER(ctypes.c_int32)]
    lib.mpmc_assembly_plan_destroy.argtypes = [P]
    lib.mpmc_transition_scan.argtypes = [P, ctypes.c_int64, P, P, P, P, ctypes.POINTER(FlashOptions),
                                         ctypes.c_int, ctypes.POINTER(FlashOutputs), ctypes.POINTER(ScanOutputs),
                                         ctypes.POINTER(ScanSummary), P]
    lib.mpmc_transition_scan_2p.argtypes = [P, P, ctypes.POINTER(AssemblyModel), P, P, P, P,
                                            ctypes.POINTER(FlashOptions), ctypes.c_int, ctypes.POINTER(FlashOutputs),
                                            ctypes.POINTER(ScanOutputs), ctypes.POINTER(ScanSummary), P]
    lib.mpmc_server_default_options.argtypes = [ctypes.POINTER(ServerOptions)]
    lib.mpmc_server_default_options.restype = None
    lib.mpmc_server_create.argtypes = [ctypes.POINTER(ServerOptions), ctypes.POINTER(P)]
    lib.mpmc_server_destroy.argtypes = [P]
    lib.mpmc_server_slots.argtypes = [P]
    lib.mpmc_server_device_bytes.argtypes = [P]
    lib.mpmc_server_device_bytes.restype = ctypes.c_size_t
    lib.mpmc_server_submit.argtypes = [P, ctypes.c_int, P, P, P, P, ctypes.c_int64, P, P, P,
                                       ctypes.POINTER(HostOutputs), ctypes.POINTER(ctypes.c_int64)]
    lib.mpmc_server_submit_hinted.argtypes = [P, ctypes.c_int, P, P, P, P, ctypes.c_int64, P, P, P, ctypes.c_int, P,
                                              ctypes.POINTER(HostOutputs), ctypes.POINTER(ctypes.c_int64)]
    lib.mpmc_server_wait.argtypes = [P, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]
    lib.mpmc_server_poll.argtypes = [P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]
    lib.mpmc_server_last_error.restype = ctypes.c_char_p
    lib.mpmc_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = library().mpmc_last_error().decode(errors="replace")
        raise MpmcError(f"{what} failed (rc={rc}): {msg}")
