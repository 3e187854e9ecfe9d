"""MI355X-native Stockham NTT (32-bit prime field and Goldilocks) for gfx950.

The compute path is libstntt.so (hand-written HIP kernels behind a C ABI,
include/stntt.h); this package is its Python host side.
"""
from .ntt import (GL_P, Batch, Context, DeviceBuffer, NttArgs, auto_ntt, default_context, find_params,  # noqa: F401
                  Server, elem_wise_mul, intt_gs, ntt_ct, ntt_goldilocks, ntt_stockham, ntt_stockham_idx, pwc_convolution, pwc_ntt,
                  restore_pass_split, set_batch_split, set_pass_shape, set_pass_split, version)
