"""nlhip: MI355X-native (gfx950) sort-based neighbour lists.

Host-side mirror of NeighbourLists.jl's sort-based API over the C-ABI HIP
library libnlhip.so (include/nlhip.h).  All neighbour searches run on the
GPU; a missing library or device raises.
"""
from ._lib import (NLHError, NLHInvalidError, NLHMemoryError, NLHOverflowError, Geometry, Handle, geometry,
                   init, lib)
from .api import (Int32, Int64, PairList, SortedCellList, build_cell_list, count_neighbours, cutoff,
                  for_each_neighbour, materialize_pairlist, max_neighbours, max_neigs, maxneigs, neighbour_list,
                  neighbours, neighbours_range, neigs, neigss, nneigs, npairs, nsites, num_neighbours, pairs,
                  sites)
from .batch import BatchBuilder, BatchResult, batch_neighbour_list

__all__ = [
    "NLHError", "NLHInvalidError", "NLHMemoryError", "NLHOverflowError", "Geometry", "Handle", "geometry",
    "init", "lib", "Int32", "Int64", "PairList", "SortedCellList", "build_cell_list", "count_neighbours",
    "cutoff", "for_each_neighbour", "materialize_pairlist", "max_neighbours", "max_neigs", "maxneigs",
    "neighbour_list", "neighbours", "neighbours_range", "neigs", "neigss", "nneigs", "npairs", "nsites", "num_neighbours", "pairs",
    "sites", "BatchBuilder", "BatchResult", "batch_neighbour_list",
]
