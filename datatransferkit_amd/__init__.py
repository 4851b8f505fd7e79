"""datatransferkit_amd: MI355X-native meshfree moving-least-squares transfer.

A from-scratch HIP implementation of DataTransferKit's
MovingLeastSquaresOperator path (kNN search, fused MLS setup with the
Jacobi pseudo-inverse, CRS apply and transpose) behind a C ABI
(``include/dtkx/mls.h``), with this Python mirror of the operator surface.
"""
from ._lib import DtkxError  # noqa: F401
from .operators import (  # noqa: F401
    MovingLeastSquaresBatch, MovingLeastSquaresOperator,
    NearestNeighborOperator, SplineOperator, knn, read_graph, release_cached_memory,
    svd_pseudo_inverse)

__all__ = ["MovingLeastSquaresOperator", "MovingLeastSquaresBatch",
           "NearestNeighborOperator", "SplineOperator", "knn", "read_graph", "svd_pseudo_inverse",
           "release_cached_memory", "DtkxError"]
