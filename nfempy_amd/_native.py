"""ctypes binding of libnfx (the C ABI declared in include/nfx.h).

The library is built in-tree (``nfempy_amd/lib/libnfx.so``) by
``nfempy_amd.build_lib``. There is no CPU fallback: every compute call goes to the
HIP library, and importing a compute path without the library raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libnfx.so"
# library version (10000 major + 100 minor + patch) whose struct layouts the mirrors below follow
MIN_VERSION = 500

NFX_OK = 0
NFX_E_INVALID = -1
NFX_E_HIP = -2
NFX_E_UNSUPPORTED = -3
NFX_E_NOT_CONVERGED = -4

# Slot / row code flags (include/nfx.h)
SLOT_INDEX_MASK = (1 << 28) - 1
SLOT_ZERO = 1 << 28
SLOT_STAGE = 1 << 29
SLOT_NEG = 1 << 30

_lib = None

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_double = ctypes.c_double
c_size_t = ctypes.c_size_t
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p

# name -> (restype, argtypes)
SIGNATURES = {
    "nfx_version": (c_int, []),
    "nfx_last_error": (c_char_p, []),
    "nfx_device_count": (c_int, [c_void_p]),
    "nfx_assemble_darcy": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int,
                                   c_void_p, c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                   c_void_p, c_void_p]),
    "nfx_assemble_darcy_rows": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_int,
                                        c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "nfx_assemble_darcy_rows_bc": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_int,
                                           c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                           c_void_p]),
    "nfx_element_geometry": (c_int, [c_int, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "nfx_md_coupling": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "nfx_md_scatter": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "nfx_md_update": (c_int, [c_int, c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "nfx_md_rhs": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_md_block_diag": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_assemble_darcy_batched": (c_int, [c_int, c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_void_p]),
    "nfx_darcy_batch_item_size": (c_int, []),
    "nfx_assemble_darcy_bc": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                      c_void_p]),
    "nfx_assemble_elasticity": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                        c_int, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int,
                                        c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_assemble_elasticity_faces": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                              c_int, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int,
                                              c_void_p, c_void_p, c_void_p]),
    "nfx_assemble_elasticity_bc": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "nfx_gather": (c_int, [c_int, c_i64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_stage_reduce": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "nfx_fill": (c_int, [c_int, c_i64, c_double, c_void_p, c_void_p]),
    "nfx_quadrature_points": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_spmv": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_float, c_void_p, c_void_p]),
    "nfx_residual64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_residual64_noalloc": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p]),
    "nfx_csr_diag64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_schur_diag64": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_solver_create": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "nfx_solver_destroy": (c_int, [c_void_p]),
    "nfx_solve_minres_ir": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int, c_int,
                                    c_void_p, c_void_p, c_void_p, c_int]),
    "nfx_field_error": (c_int, [c_int, c_int, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_assemble_projector": (c_int, [c_int] * 8 + [c_void_p] * 6 + [c_int] + [c_void_p] * 5),
    "nfx_field_values": (c_int, [c_int] * 6 + [c_void_p] * 7),
    "nfx_elasticity_mass_schur_pc": (c_int, [c_int] * 3 + [c_void_p] * 7),
    "nfx_solver_set_block3": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "nfx_gnn_num_params": (c_int, []),
    "nfx_schur_ell": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_gnn_features": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p]),
    "nfx_gnn_forward": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_gnn_backward_workspace": (c_int, [c_int]),
    "nfx_gnn_backward": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p]),
    "nfx_ell_spmm": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_float, c_void_p,
                             c_void_p, c_void_p]),
    "nfx_smoother_grad": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_void_p]),
    "nfx_agg_sum": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_coarse_dense": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p]),
    "nfx_twogrid_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_elasticity_tg_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_minres_batch_create": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "nfx_minres_batch_destroy": (c_int, [c_void_p]),
    "nfx_minres_batch_rows": (c_i64, [c_void_p]),
    "nfx_minres_batch_set_pc": (c_int, [c_void_p, c_int, c_void_p]),
    "nfx_batch_tg_schur": (c_int, [c_int, c_int, c_int, c_int] + [c_void_p] * 11 + [c_void_p]),
    "nfx_batch_tg_schur_blk": (c_int, [c_int] + [c_void_p] * 13 + [c_void_p]),
    "nfx_batch_tg_fluxinv": (c_int, [c_int] + [c_void_p] * 5 + [c_void_p]),
    "nfx_batch_tg_smoother": (c_int, [c_int, c_int, c_void_p, c_double, c_void_p, c_void_p, c_void_p]),
    "nfx_batch_tg_coarse": (c_int, [c_int, c_int] + [c_void_p] * 9 + [c_void_p]),
    "nfx_sort_workspace": (c_int, [c_i64, c_void_p]),
    "nfx_stable_sort64": (c_int, [c_i64, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "nfx_first_appearance": (c_int, [c_i64, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t,
                                     c_void_p]),
    "nfx_batch_tg_apply": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_batch_tg_rowsums": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_solve_minres_batch_ir": (c_int, [c_void_p, c_void_p, c_void_p, c_double, c_double, c_int, c_int, c_void_p,
                                          c_void_p, c_void_p]),
    "nfx_elasticity_tg_blocks": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "nfx_elasticity_tg_coarse": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p]),
    "nfx_solve_minres_ir_tg": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_double, c_double, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                       c_int]),
    "nfx_half_rows": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "nfx_sgemm": (c_int, [c_int, c_int, c_int, c_float, c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_float,
                          c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "nfx_spd_inverse_small": (c_int, [c_int, c_void_p, c_int, c_void_p]),
    "nfx_flux_sweep_coef": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_void_p,
                                    c_void_p, c_void_p]),
    "nfx_fgmres_create": (c_int, [c_void_p, c_int, c_int, c_void_p]),
    "nfx_fgmres_destroy": (c_int, [c_void_p]),
    "nfx_solve_fgmres_ir": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_double, c_double, c_int, c_int, c_int, c_int,
                                    c_void_p, c_void_p, c_void_p, c_int]),
}


class TwoGridPC(ctypes.Structure):
    """Mirror of ``nfx_twogrid_pc`` (include/nfx.h)."""

    _fields_ = [("n_u", c_int), ("n_cells", c_int), ("n_agg", c_int), ("nbr", c_void_p), ("Wd", c_void_p),
                ("Wo", c_void_p), ("sd", c_void_p), ("so", c_void_p), ("agg_ptr", c_void_p),
                ("agg_mem", c_void_p), ("agg_of", c_void_p), ("Ainv", c_void_p), ("Ainv16", c_void_p),
                ("Ascale", c_void_p), ("dMinv", c_void_p), ("psign", c_float), ("work", c_void_p),
                ("n_tail", c_int), ("tail_dinv", c_void_p),
                ("slot_tail", c_void_p), ("slot_coef", c_void_p), ("tail_cells", c_void_p),
                ("tail_coef", c_void_p), ("flux_width", c_int), ("flux_col", c_void_p),
                ("flux_coef", c_void_p)]


class MinresBatchItem(ctypes.Structure):
    """Mirror of ``nfx_minres_batch_item`` (include/nfx.h)."""

    _fields_ = [("n_rows", c_int), ("lanes", c_int), ("row_off", c_i64), ("rowptr", c_void_p), ("col", c_void_p),
                ("val32", c_void_p), ("val64", c_void_p), ("sign", c_void_p), ("pinv", c_void_p)]


class ElasticityTG(ctypes.Structure):
    """Mirror of ``nfx_elasticity_tg`` (include/nfx.h)."""

    _fields_ = [("n_s", c_int), ("n_u", c_int), ("n_cells", c_int), ("n_agg", c_int), ("dinv", c_void_p),
                ("nbr", c_void_p), ("sblk", c_void_p), ("winv", c_void_p), ("agg_ptr", c_void_p),
                ("agg_mem", c_void_p), ("agg_of", c_void_p), ("ainv", c_void_p), ("psign", c_float),
                ("work", c_void_p)]


class BatchTG(ctypes.Structure):
    """Mirror of ``nfx_batch_tg`` (include/nfx.h)."""

    _fields_ = [("nb", c_int), ("n_cells", c_int), ("n_agg", c_int)] + \
        [(name, c_void_p) for name in ("cell_inst", "prow", "nbr", "sd", "so", "rs", "wd", "wo", "agg_ptr", "agg_mem",
                                       "agg_of", "arow", "a0", "alen", "ainv", "work")] + \
        [("n_fb", c_int), ("fb_inst", c_void_p), ("fb_rows", c_void_p), ("fb_inv", c_void_p)]


# entry points whose return value is a count, not a status
_RAW = ("nfx_version", "nfx_darcy_batch_item_size", "nfx_gnn_num_params", "nfx_gnn_backward_workspace",
        "nfx_minres_batch_rows")


class NfxError(RuntimeError):
    pass


def library_available() -> bool:
    return LIB_PATH.exists()


def lib():
    """Load libnfx once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise NfxError(
            f"libnfx not built: {LIB_PATH} is missing (run `python -m nfempy_amd.build_lib`)"
        )
    handle = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | getattr(os, "RTLD_GLOBAL", 0))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    version = handle.nfx_version()
    if version < MIN_VERSION:
        # the descriptor structs mirrored here have the layout of MIN_VERSION
        raise NfxError(f"{LIB_PATH} is version {version}, these bindings need {MIN_VERSION} or later "
                       f"(rebuild with `python -m nfempy_amd.build_lib`)")
    _lib = handle
    return _lib


def exported_symbols():
    return list(SIGNATURES)


def check(rc: int, what: str = "") -> int:
    if rc == NFX_OK:
        return rc
    msg = lib().nfx_last_error().decode(errors="replace")
    if rc == NFX_E_INVALID:
        raise ValueError(f"{what}: {msg}")
    if rc == NFX_E_NOT_CONVERGED:
        raise NfxError(f"{what}: not converged: {msg}")
    raise NfxError(f"{what}: {msg} (status {rc})")


def call(name: str, *args) -> int:
    rc = getattr(lib(), name)(*args)
    if name in _RAW:
        return rc
    return check(rc, name)


def ptr(t) -> int | None:
    """Device (or host) address of a torch tensor / numpy array, or None."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    if hasattr(t, "ctypes"):
        return t.ctypes.data
    raise TypeError(type(t))


def stream_handle(stream=None) -> int | None:
    """HIP stream handle of ``stream`` (a torch stream), or of torch's current stream on
    the current device (read without building a stream object: ~0.1 us instead of ~3)."""
    import torch

    if stream is not None:
        return stream.cuda_stream
    raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
    if raw is not None:
        return raw(torch._C._cuda_getDevice())
    return torch.cuda.current_stream().cuda_stream
