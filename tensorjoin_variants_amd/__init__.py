"""tensorjoin_variants_amd -- certified exact epsilon-join for MI355X (gfx950).

Public API (the reference's join call as a drop-in):

    join(R, S=None, eps=..., precision='auto')  -> sorted int [P, 2] pairs
    Matrix(T).run('F16' | 'F8' | 'F32', x)       -> (sorted directed u64 ids incl. self, stats)
    join_reference_compat(x, eps)                -> (directed u64 ids, stats)
    Plan(...)                                    -> reusable plan + workspace
    batched_join(sets, eps)                      -> one sorted pair array per set
    BatchPlan(...)                               -> reusable batch plan (+ hipGraph capture)
    join_host(R, S=None, eps=..., out='int32'|'canonical') -> host-to-host pipeline (NumPy in, NumPy out)
    gdsjoin(DATASET, EPSILON, NDIM, DTYPE)       -> GDS-Join's (numNeighbors, neighborTable)
    HostPlan(...)                                -> reusable host-to-host plan
    vecio.read_fvecs / read_idx / validate_canonical -> vector files and the output contract

All computation runs in libtjoin.so (hand-written HIP for gfx950); the
package raises if the library is missing.
"""
from ._lib import TJError, load as load_library
from .batch import BatchPlan, batched_join
from .gds import gdsjoin
from .host import HostPlan, join_host
from .join import JoinResult, Matrix, Plan, canonical_u64, join, join_reference_compat

__all__ = ["join", "Matrix", "Plan", "JoinResult", "canonical_u64", "join_reference_compat",
           "batched_join", "BatchPlan", "gdsjoin", "join_host", "HostPlan", "TJError", "load_library"]
