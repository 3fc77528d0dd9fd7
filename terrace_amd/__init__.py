"""MI355X (gfx950) implementation of Terrace's hot path: batch updates into
the degree-tiered graph container and Ligra-style edgeMap traversal, behind
the C ABI in include/terrace_hip.h."""
from .graph import (BC, CC, PR_S, TC, BFS_with_edge_map, DeviceBuffer, Graph, TerraceError,  # noqa: F401
                    VertexSubset, edgeMap, rmat_device)
