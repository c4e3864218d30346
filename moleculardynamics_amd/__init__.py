"""MI355X (gfx950) molecular-dynamics engine behind the reference cpumd step loop.

The compute path is libmdx.so (hand-written HIP kernels behind the C-ABI in
include/mdx.h); this package only loads it and moves arrays across it.
"""
from ._lib import (BOND_REFERENCE, BOND_SURROGATE, BOND_TERSOFF, BOX_PERIODIC, BOX_WALLED, LIST_BOND, LIST_LJ,
                   LJ_REFERENCE, LJ_UNIFORM, PREC_FP64, PREC_MIXED, SG_NFEAT, SG_NPARAM, SG_NTRAIN, Batch,
                   MdxConfig, MdxError, SgTrainer, Sweep, System,
                   header_symbols, load, tuning_lookup)

__all__ = [
    "System", "Batch", "Sweep", "MdxConfig", "MdxError", "load", "header_symbols",
    "BOX_WALLED", "BOX_PERIODIC", "PREC_MIXED", "PREC_FP64", "LJ_REFERENCE", "LJ_UNIFORM",
    "BOND_REFERENCE", "BOND_TERSOFF", "BOND_SURROGATE", "LIST_LJ", "LIST_BOND",
    "SgTrainer", "SG_NFEAT", "SG_NTRAIN", "SG_NPARAM", "tuning_lookup",
]
