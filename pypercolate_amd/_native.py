"""ctypes binding of libpercolate_hip.so (the C ABI in include/percolate_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``python -m pypercolate_amd.build``).  There is no CPU fallback: every
compute entry point needs the library and a ROCm GPU, and fails loudly
without them.
"""

import ctypes
import os

import numpy as np

LIB_NAME = 'libpercolate_hip.so'
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

PC_OK = 0
PC_EINVAL = -1
PC_ENOMEM = -2
PC_EHIP = -3
PC_EOVERFLOW = -4

PC_NO_SPAN = 0xFFFFFFFF
PC_STATE_AUTO = 0
PC_STATE_LDS = 1
PC_STATE_HBM = 2
PC_SWEEP_ONE_WAVE = 0x10
PC_SWEEP_TWO_WAVES = 0x20
PC_SWEEP_WAVES_MASK = 0x30
PC_CHUNK = 64
PC_STATE_WORDS = 8
PC_COLUMNS = 7
PC_NOALLOC_SCRATCH_BYTES = 32 << 20

_u32 = ctypes.c_uint32
_i32 = ctypes.c_int32
_u64 = ctypes.c_uint64
_i64 = ctypes.c_int64
_vp = ctypes.c_void_p

# name -> (restype, argtypes); the declarations of include/percolate_hip.h
SIGNATURES = {
    'pc_abi_version': (ctypes.c_int, []),
    'pc_last_error': (ctypes.c_char_p, []),
    'pc_record_bytes': (ctypes.c_int, [_u32]),
    'pc_permute_philox': (ctypes.c_int, [_u64, _u64, _u32, _u32, _vp, _i64, _vp,
                                         _u64, _vp]),
    'pc_permute_philox_workspace_size': (ctypes.c_int, [_u32, _u32,
                                                        ctypes.POINTER(_u64)]),
    'pc_permute_randomstate': (ctypes.c_int, [_vp, _u32, _u32, _vp, _i64, _vp]),
    'pc_sweep_workspace_size': (ctypes.c_int, [_u32, _u32, _u32,
                                               ctypes.POINTER(_u64)]),
    'pc_sweep': (ctypes.c_int, [_u32, _u32, _vp, _vp, _i32, _vp, _i64, _u32,
                                _vp, _i64, _vp, _vp, _u64, _u32, _u32, _vp,
                                _vp]),
    'pc_sweep_occupancy': (ctypes.c_int, [_u32, _u32, ctypes.POINTER(_i32)]),
    'pc_chunk_scan': (ctypes.c_int, [_u32, _u32, _u32, _vp, _i64, _vp, _vp]),
    'pc_microcanonical_sums': (ctypes.c_int, [_u32, _u32, _u32, _vp, _i64, _vp,
                                              _vp, _vp, _vp, _i32, _vp]),
    'pc_expand_traces': (ctypes.c_int, [_u32, _u32, _u32, _vp, _i64, _vp, _vp,
                                        _vp, _vp, _vp, _vp]),
    'pc_binomial_window': (ctypes.c_int, [_u32, ctypes.c_double,
                                          ctypes.POINTER(_i64),
                                          ctypes.POINTER(_i64)]),
    'pc_binomial_weights': (ctypes.c_int, [_u32, _u32, _vp, _vp, _vp, _i64,
                                           _vp, _vp]),
    'pc_contract_workspace_size': (ctypes.c_int, [_u32, _u32, _i32,
                                                  ctypes.POINTER(_u64)]),
    'pc_canonical_contract': (ctypes.c_int, [_u32, _u32, _vp, _i64, _u32, _vp,
                                             _vp, _vp, _i64, _vp, _i32, _vp,
                                             _u64, _vp]),
    'pc_canonical_fused_workspace_size': (ctypes.c_int, [_u32, _u32, _i32,
                                                         ctypes.POINTER(_u64)]),
    'pc_canonical_fused': (ctypes.c_int, [_u32, _u32, _vp, _i64, _u32, _vp, _vp,
                                          _vp, _vp, _i32, _vp, _u64, _vp]),
    'pc_canonical_fused_noalloc': (ctypes.c_int, [_u32, _u32, _vp, _i64, _u32,
                                                  _vp, _vp, _vp, _vp, _i32,
                                                  _vp]),
    'pc_microcanonical_ci': (ctypes.c_int, [_u32, _u32, _vp, _vp,
                                            ctypes.c_double, ctypes.c_double,
                                            _vp, _vp, _vp]),
    'pc_convert_f64_to_f32': (ctypes.c_int, [_vp, _i64, _vp, _vp]),
    'pc_window_traces': (ctypes.c_int, [_u32, _u32, _u32, _vp, _i64, _vp, _vp,
                                        _i64, _i64, _i32, _vp, _vp]),
    'pc_run_moments': (ctypes.c_int, [_vp, _u32, _u32, _u32, _vp, _vp, _vp]),
    'pc_microcanonical_moments': (ctypes.c_int, [_u32, _u32, _vp, _vp, _vp,
                                                 _vp, _vp]),
    'pc_finalize_canonical': (ctypes.c_int, [_u32, _u32, _vp, _vp, _vp, _vp,
                                             _vp, _vp, _vp, _vp]),
    'pc_merge_moments': (ctypes.c_int, [_u32, _u32, _vp, _vp, _vp, _vp, _vp,
                                        _vp, _vp, _vp, _vp]),
    'pc_sweep_multi_workspace_size': (ctypes.c_int, [_u32, _vp,
                                                     ctypes.POINTER(_u64),
                                                     ctypes.POINTER(_u64)]),
    'pc_sweep_multi_prepare': (ctypes.c_int, [_u32, _vp, _vp, _u64]),
    'pc_sweep_multi': (ctypes.c_int, [_u32, _vp, _vp, _u64, _vp, _u64, _vp]),
    'pc_sweep_lattice': (ctypes.c_int, [_u32, _i32, _vp, _i32, _vp, _i64, _u32,
                                        _vp, _i64, _vp, _vp, _u64, _u32, _u32,
                                        _vp, _vp]),
}

class SweepGraph(ctypes.Structure):
    """pc_sweep_graph: one graph of a persistent multi-graph sweep."""
    _fields_ = [('num_nodes', ctypes.c_uint32), ('num_edges', ctypes.c_uint32),
                ('num_runs', ctypes.c_uint32),
                ('span_on_first_merge', ctypes.c_int32),
                ('merge_rounds', ctypes.c_uint32), ('lattice_length', ctypes.c_uint32),
                ('d_edges', ctypes.c_void_p), ('d_node_sides', ctypes.c_void_p),
                ('d_perm', ctypes.c_void_p), ('perm_stride', ctypes.c_int64),
                ('d_records', ctypes.c_void_p), ('rec_stride', ctypes.c_int64),
                ('d_span_step', ctypes.c_void_p),
                ('d_chunk_state', ctypes.c_void_p)]


# precisions of pc_canonical_contract
PC_PREC_FP64 = 0
PC_PREC_FP64_MFMA = 1
PC_PREC_FP32_MFMA = 2
PC_PREC_BF16X3_MFMA = 3
PC_PREC_SPLIT_BF16 = 4
PRECISIONS = {'fp64': PC_PREC_FP64, 'fp64_mfma': PC_PREC_FP64_MFMA,
              'fp32_mfma': PC_PREC_FP32_MFMA, 'bf16x3_mfma': PC_PREC_BF16X3_MFMA,
              'split_bf16': PC_PREC_SPLIT_BF16}

_lib = None


class NativeError(RuntimeError):
    """A HIP runtime failure reported by libpercolate_hip."""


def load_library():
    """Load libpercolate_hip.so and declare its prototypes (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            '{} is not built; run `python -c "import __graft_entry__ as g; '
            'g.build()"` in the repository root'.format(LIB_PATH))
    lib = ctypes.CDLL(LIB_PATH)
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    _lib = lib
    return lib


def check(status, what):
    """Map a negative status of the C ABI to a Python exception."""
    if status == PC_OK:
        return
    msg = load_library().pc_last_error().decode('utf-8', 'replace')
    text = '{}: {}'.format(what, msg or 'status {}'.format(status))
    if status in (PC_EINVAL, PC_EOVERFLOW):
        raise ValueError(text)
    if status == PC_ENOMEM:
        raise MemoryError(text)
    raise NativeError(text)


def torch_cuda():
    """Return torch after checking that a ROCm GPU is present."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError(
            'pypercolate_amd runs its hot path on an AMD Instinct GPU; no '
            'ROCm device is visible (there is no CPU fallback)')
    return torch


def ptr(tensor):
    """Device pointer of a torch tensor (None for None)."""
    return None if tensor is None else ctypes.c_void_p(tensor.data_ptr())


def stream_handle(torch, stream=None):
    """hipStream_t of a torch stream (default: the current stream)."""
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def record_dtype(num_nodes):
    """numpy dtype of one merge record for a graph of num_nodes nodes."""
    return np.uint32 if num_nodes <= 65536 else np.uint64
