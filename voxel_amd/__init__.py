"""voxel_amd — MI355X-native (gfx950) GPU voxel-editing path.

Hand-written HIP kernels behind a C ABI (include/vx.h): onesweep radix sort,
single-pass scan and stream compaction, slab and cuckoo hash tables, and the
Hash DAG edit, and the pointer-SVO voxel edit. This package is the Python host
mirror of the reference's ``algo::cuda`` primitives and its
``algo::svt::hash_dag_gpu`` and ``algo::svt::cuda`` editing APIs.
"""
from ._lib import (VX_DEVICE_ERROR, VX_INVALID_VALUE, VX_OK, VX_OUT_OF_CAPACITY, VX_UNSUPPORTED, InvalidValue,
                   OutOfCapacity, VxError, lib)

__all__ = ["lib", "VxError", "InvalidValue", "OutOfCapacity", "VX_OK", "VX_INVALID_VALUE", "VX_OUT_OF_CAPACITY",
           "VX_DEVICE_ERROR", "VX_UNSUPPORTED"]
