"""tc -- MI355X-native point-cloud tree construction (host side of libtc.so).

Stage chain of ucsc-redwood/glsl-tree-construction (src/main.rs:60-170):
morton -> sort -> unique -> radix tree -> edge count -> scan -> octree, each a
hand-written HIP kernel for gfx950 behind the C ABI in include/tc.h.
"""
from ._lib import OCT_NODE_DTYPE, TcError, check, lib, LIB_PATH, PKG_DIR  # noqa: F401
from .api import (  # noqa: F401
    Context,
    Tree,
    build_octree,
    build_radix_tree,
    compute_morton,
    default_context,
    edge_count,
    gen_gaussian_mixture,
    gen_uniform,
    histogram,
    morton32_to_xyz,
    morton64_to_xyz,
    octree_node_bound,
    partial_sum,
    radix_sort,
    unique,
)
from .batch import Batch  # noqa: F401,E402
from .serve import Server, plain_run_trace, result_bytes, simulate  # noqa: F401,E402
from . import tune  # noqa: F401,E402
from .treeio import load_tree, validate_tree, write_tree  # noqa: F401,E402
