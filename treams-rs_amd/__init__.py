"""treams-rs_amd: the cluster-coupling hot path of treams-rs on MI355X (gfx950).

The numerical work lives in libtreams_gpu.so (HIP, C ABI in include/treams_gpu.h);
this package is the host-side mirror of the reference's Python interface.
"""
from ._lib import NotConverged, lib  # noqa: F401
from .cluster import (  # noqa: F401
    DEFAULT_ATOL,
    DEFAULT_MAX_ITERATIONS,
    DEFAULT_RESTART,
    DEFAULT_RTOL,
    Convergence,
    DenseFactor,
    FactorContext,
    IterativeContext,
    IterativeSphereCluster,
    SphereClusterFactor,
    mode_index,
    particle_cluster,
    plan_terms,
    plane_wave_field,
    expansion_field,
    rotation_matrix,
    rotate_tmatrix,
    scattering_gradient,
    sphere_cluster_factor,
    translation_blocks,
    translation_blocks_tangent,
    translation_blocks_rect,
    wigner3j,
)
from .geometry import WORKLOADS, random_cluster, workload  # noqa: F401
from . import diff  # noqa: F401,E402
