"""lgmi -- LightGraphs.jl's BFS and exact-betweenness hot path on MI355X.

Host-side mirror of the reference's Julia interface for this path, calling
hand-written gfx950 HIP kernels in liblgmi.so through its C ABI
(include/lgmi.h):

* ``gdistances``, ``gdistances_`` (gdistances!), ``bfs_parents``,
  ``bfs_tree``, ``tree``            -- src/traversals/bfs.jl
* ``betweenness_centrality``        -- src/centrality/betweenness.jl
* ``stress_centrality``             -- src/centrality/stress.jl
* ``closeness_centrality``, ``radiality_centrality``, ``eccentricity``,
  ``diameter``, ``radius``, ``center``, ``periphery``
                                    -- src/centrality/{closeness,radiality}.jl,
                                       src/distance.jl
* ``connected_components``, ``is_connected``, ``has_path``, ``period``
                                    -- src/connectivity.jl, src/traversals/bfs.jl
* ``Parallel.*``                    -- src/Parallel/{traversals,centrality}
* ``SimpleGraph``, ``SimpleDiGraph`` -- src/SimpleGraphs (CSR-backed)
* ``loadgraph``, ``loadgraphs``, ``savegraph`` -- src/persistence (LG format;
                                       graphs built on the device)
"""
from . import Parallel, generators
from ._lib import LgmiError, NoDeviceError
from .centrality import betweenness_centrality, betweenness_centrality_batch, sample, stress_centrality
from .connectivity import (components, components_dict, connected_components, connected_components_,
                           has_path, is_connected, is_weakly_connected, neighborhood, neighborhood_dists, period)
from .distances import (center, closeness_centrality, diameter, eccentricity, periphery, radiality_centrality,
                        radius)
from .graphs import SimpleDiGraph, SimpleGraph
from .persistence import loadgraph, loadgraphs, savegraph
from .traversals import bfs_parents, bfs_tree, gdistances, gdistances_, tree, typemax

__all__ = [
    "Parallel", "generators", "LgmiError", "NoDeviceError", "betweenness_centrality", "sample",
    "SimpleDiGraph", "SimpleGraph", "bfs_parents", "bfs_tree", "gdistances", "gdistances_", "stress_centrality",
    "tree", "typemax", "center", "closeness_centrality", "diameter", "eccentricity", "periphery",
    "radiality_centrality", "radius", "components", "components_dict", "connected_components",
    "connected_components_", "has_path", "is_connected", "is_weakly_connected", "period", "neighborhood",
    "neighborhood_dists", "loadgraph", "loadgraphs", "savegraph",
]
