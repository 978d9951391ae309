"""hivehashtable_amd -- MI355X-native (gfx950) Hive hash table.

The product is libhive.so (HIP kernels + C-ABI, see include/hive.h); this
package builds it (``hivehashtable_amd.build``) and binds it for Python
callers (``HiveTable``).
"""
from ._lib import (HiveError, LAYOUTS, POLICIES, OP_INSERT, OP_LOOKUP, OP_DELETE, OP_ATOMIC_INC, OP_NONE,
                   R_FAILED, R_INSERTED, R_UPDATED, R_STASHED, R_FOUND, R_NOT_FOUND, R_ERASED,
                   R_INVALID_KEY, R_NOOP, lib, declared_symbols, LIB_PATH)
from .table import HiveTable, make_ops, dispatch, dispatch_ycsb

__all__ = [
    "HiveTable", "HiveError", "make_ops", "dispatch", "dispatch_ycsb", "lib", "declared_symbols", "LIB_PATH",
    "POLICIES", "OP_INSERT", "OP_LOOKUP", "OP_DELETE", "OP_ATOMIC_INC", "OP_NONE",
    "R_FAILED", "R_INSERTED", "R_UPDATED", "R_STASHED", "R_FOUND", "R_NOT_FOUND", "R_ERASED", "R_INVALID_KEY",
    "R_NOOP",
]
