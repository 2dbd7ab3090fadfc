"""ctypes binding of libnlccsd.so (declarations in include/nulattice_ccsd.h).

The library is loaded from the package directory; there is no CPU fallback: if it is missing
or a GPU call is attempted without a HIP device, the call raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnlccsd.so")

_i64, _i32, _dbl, _flt = ctypes.c_int64, ctypes.c_int, ctypes.c_double, ctypes.c_float
_vp = ctypes.c_void_p

# name -> (restype, argtypes); mirrors include/nulattice_ccsd.h
SIGNATURES = {
    "nl_gemm_f64": (_i32, [_i32, _i32, _i64, _i64, _i64, _dbl, _vp, _i64, _i64, _vp, _i64, _i64,
                           _dbl, _vp, _i64, _i64, _i64, _i64, _vp, _vp]),
    "nl_contract_ws": (_i64, [_vp]),
    "nl_contract_f64": (_i32, [_vp, _dbl, _vp, _vp, _vp, _vp, _dbl, _vp, _vp, _vp, _i64, _vp,
                               _i32, _vp]),
    "nl_gemm_lp": (_i32, [_i32, _i64, _i64, _i64, _dbl, _vp, _i64, _i64, _vp, _i64, _vp, _i64,
                          _i64, _vp, _i64, _dbl, _vp, _vp, _i64, _i64, _i64, _i64, _vp, _vp]),
    "nl_gemm_lp_split": (_i32, [_i32, _i64, _i64, _i64, _dbl, _vp, _vp, _i64, _i64, _vp, _i64,
                                _vp, _vp, _i64, _i64, _vp, _i64, _dbl, _vp, _vp, _i64, _i64,
                                _i64, _i64, _vp, _vp]),
    "nl_gemm_lp_strided_split": (_i32, [_i32, _i64, _dbl, _vp, _vp, _i64, _vp, _vp, _vp, _i64,
                                        _vp, _dbl, _vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "nl_gemm_lp_rows_split": (_i32, [_i32, _i64, _i64, _i64, _dbl, _vp, _vp, _i64, _i64, _vp,
                                     _i64, _vp, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                                     _vp]),
    "nl_gemm_lp_rows": (_i32, [_i32, _i64, _i64, _i64, _dbl, _vp, _i64, _i64, _vp, _i64, _vp, _i64,
                               _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nl_convert_split": (_i32, [_i32, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _i32,
                                _i32, _vp, _flt, _dbl, _vp]),
    "nl_convert_split_delta": (_i32, [_i32, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _i64,
                                      _i64, _i32, _i32, _vp, _vp, _flt, _dbl, _vp]),
    "nl_gather_convert": (_i32, [_i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _vp]),
    "nl_gather_rows16": (_i32, [_i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _vp]),
    "nl_tau16": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp]),
    "nl_transpose16": (_i32, [_i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nl_transpose16_batched": (_i32, [_i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _vp]),
    "nl_permute16": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp]),
    "nl_transpose16_into": (_i32, [_i64, _i64, _i64, _vp, _i64, _i64, _vp, _i64, _i64, _i64,
                                   _vp]),
    "nl_convert_rows_into": (_i32, [_i64, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _vp]),
    "nl_gemm_lp_strided": (_i32, [_i32, _i64, _dbl, _vp, _i64, _vp, _vp, _i64, _vp, _dbl, _vp,
                                  _vp, _vp, _vp, _i64, _vp, _vp]),
    "nl_permute": (_i32, [_i32, _vp, _vp, _dbl, _vp, _dbl, _vp, _vp]),
    "nl_convert_rows": (_i32, [_i32, _i64, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _flt,
                               _vp]),
    "nl_tau": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp]),
    "nl_segsum": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _i64, _i64, _dbl, _vp, _vp, _i32,
                         _vp]),
    "nl_gather_rows": (_i32, [_i64, _i64, _vp, _vp, _i64, _vp, _vp]),
    "nl_scatter_add_rows": (_i32, [_i64, _i64, _vp, _vp, _dbl, _vp, _i64, _vp]),
    "nl_tau_split_pairs": (_i32, [_i64, _i64, _vp, _vp, _vp, _i64, _i64, _i32, _dbl, _vp]),
    "nl_ring_from_pairs": (_i32, [_i64, _i64, _i64, _vp, _vp, _vp]),
    "nl_split_delta16": (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _i32, _vp]),
    "nl_split_delta8": (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _vp, _vp]),
    "nl_convert_rows8": (_i32, [_i64, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _vp]),
    "nl_hole_pairs": (_i32, [_i64, _i64, _i64, _i64, _vp, _vp, _i32, _vp]),
    "nl_concat_rows": (_i32, [_i64, _i64, _vp, _i64, _i64, _vp, _i64, _vp, _i64, _vp]),
    "nl_update_t1": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _dbl, _vp, _vp]),
    "nl_assemble_t2": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _dbl,
                              _vp, _vp, _vp, _vp]),
    "nl_assemble_t2_16": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _dbl,
                                 _vp, _vp, _vp, _vp, _vp]),
    "nl_assemble_t2_ext": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                  _vp, _dbl, _vp, _vp, _vp, _vp, _i32, _dbl, _i32, _vp]),
    "nl_tau_split": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _i32, _dbl, _vp]),
    "nl_gather_rows16_into": (_i32, [_i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _i64, _vp]),
    "nl_take_diag": (_i32, [_i64, _vp, _vp, _vp]),
    "nl_assemble_t2_adj": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "nl_partials": (_i64, [_i64]),
    "nl_energy_part": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp]),
    "nl_ccsd_energy_noalloc": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nl_finalize": (_i32, [_i64, _vp, _i64, _vp, _vp, _vp, _vp]),
    "nl_dot_part": (_i32, [_i64, _vp, _vp, _vp, _vp]),
    "nl_axpby": (_i32, [_i64, _dbl, _vp, _dbl, _vp, _vp]),
    "nl_sub": (_i32, [_i64, _vp, _vp, _vp, _vp]),
    "nl_copy_sub": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "nl_gram_part": (_i32, [_i64, _i32, _vp, _vp, _vp]),
    "nl_copy_sub_f32": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "nl_gram_part_f32": (_i32, [_i64, _i32, _vp, _vp, _vp]),
    "nl_lincomb_f32": (_i32, [_i64, _i32, _vp, _vp, _vp, _vp]),
    "nl_gram_finalize": (_i32, [_i64, _i32, _vp, _vp, _vp]),
    "nl_lincomb": (_i32, [_i64, _i32, _vp, _vp, _vp, _vp]),
    "nl_gemm_grouped_f64": (_i32, [_i32, _i64, _vp, _vp, _dbl, _vp, _vp, _dbl, _vp, _vp, _vp]),
    "nl_sector_gather": (_i32, [_i64, _vp, _vp, _vp, _vp]),
    "nl_sector_scatter": (_i32, [_i64, _vp, _vp, _vp, _vp]),
    "nl_sector_dress": (_i32, [_i64, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _dbl, _vp, _vp, _vp]),
    "nl_sector_ring": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nl_sector_ring_adj": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "nl_sector_gather_add": (_i32, [_i64, _vp, _vp, _dbl, _vp, _vp]),
    "nl_sector_diag_axpy": (_i32, [_i64, _i64, _vp, _vp, _vp, _dbl, _vp, _vp]),
    "nl_contact_fill": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _dbl, _i64, _vp, _vp]),
    "nl_sector_assemble": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _dbl, _dbl, _vp,
                                  _vp, _vp, _vp]),
    "nl_gather_rows16_split": (_i32, [_i64, _i64, _i64, _vp, _vp, _i64, _vp, _i64, _i64, _vp]),
    "nl_segplan_ws": (_i64, [_i64]),
    "nl_segplan_sort": (_i32, [_i64, _vp, _vp, _vp, _vp, _dbl, _i32, _vp, _i64, _vp, _vp, _vp,
                               _vp]),
    "nl_segplan_finish": (_i32, [_i64, _i64, _vp, _i64, _vp, _vp, _vp, _vp]),
    "nl_pair_outer": (_i32, [_i64, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _vp, _vp]),
    "nl_pw_dense": (_i32, [_i64, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _i32, _dbl, _vp, _vp]),
    "nl_pw_pair_fill": (_i32, [_i64, _i64, _i64, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _i32,
                               _dbl, _vp, _vp]),
    "nl_pw_coo": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _i32,
                         _dbl, _i64, _vp, _vp, _vp]),
}

_LIB = None


class NLError(RuntimeError):
    pass


def load():
    """Load libnlccsd.so once and attach signatures; raises if it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise NLError(f"{LIB_PATH} is not built; run `python -m nulattice_amd.build` "
                      "(there is no CPU fallback for the CCSD kernels)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def call(name, *args):
    """Call an entry point and raise on a nonzero status."""
    rc = getattr(load(), name)(*args)
    if rc != 0:
        raise NLError(f"{name} returned {rc} ({'invalid argument' if rc == -1 else 'HIP error'})")
    return rc
