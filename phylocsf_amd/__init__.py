"""MI355X-native PhyloCSF scoring path (HIP kernels for gfx950 behind a C ABI).

The public surface mirrors the reference's CamlPaml / PhyloCSF interface; see
api.py.  The HIP library (lib/libphylocsf_amd.so) is required: nothing here
falls back to a CPU implementation.
"""
from ._lib import (PCSF_BF16, PCSF_FP16, PCSF_FP16S, PCSF_FP32, PCSF_FP64,  # noqa
                   PCSF_FP64M, PCSF_MARGINALIZE, PCSF_N_VARIANTS, Failure, InvalidArgument, PcsfError)
from .api import (DB, Batch, Context, PhyloCSFModel, PhyloLik, Tree,  # noqa
                  collect_sufficient_statistics, import_parameters, newick_tree,
                  paramset_prefix, score, score_fixed, score_omega)

__all__ = ["Context", "PhyloCSFModel", "Batch", "PhyloLik", "Tree", "import_parameters",
           "newick_tree", "paramset_prefix", "score", "score_fixed", "score_omega",
           "collect_sufficient_statistics", "DB",
           "PCSF_FP32", "PCSF_FP64", "PCSF_FP64M",
           "PCSF_FP16S", "PCSF_FP16", "PCSF_BF16", "PCSF_N_VARIANTS",
           "PCSF_MARGINALIZE", "PcsfError", "InvalidArgument", "Failure"]
