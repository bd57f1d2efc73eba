"""
    NeighbourListsHIP

Julia host shim over `libnlhip.so` (the C ABI declared in `include/nlhip.h`).

It offers the sort-based entry points of NeighbourLists.jl with the same
names, arguments and results -- `neighbour_list(X, cutoff, cell, pbc; lazy,
int_type)`, `build_cell_list`, `materialize_pairlist`, `PairList(X, cutoff,
cell, pbc)` -- built on an MI355X instead of through KernelAbstractions
(reference: `src/cell_list.jl:632-679, 848-916`, `src/gpu_kernels.jl:299-364`).
Results are the reference's own types (`NeighbourLists.PairList`,
`NeighbourLists.SortedCellList`) holding ordinary host arrays, so everything
downstream (`neighbours`, `for_each_neighbour`, `npairs`, ...) runs
unchanged.  `HIPCellList` keeps the device-resident build for repeated
materialisation and per-atom queries without copying the cell list back.

Conventions follow the reference: 1-based indices of type `int_type`
(Int32 or Int64), `cell` rows are the lattice vectors, positions are
`Vector{SVector{3,Float64}}`; a cell too fine for `int_type` raises the
reference's error message; a zero-volume cell warns.

Set `ENV["NLHIP_LIB"]` to the library path before loading the module, or
place `libnlhip.so` next to this file's package directory.
"""
module NeighbourListsHIP

using StaticArrays: SVector, SMatrix
import NeighbourLists
using NeighbourLists: PairList, SortedCellList

export hip_neighbour_list, hip_build_cell_list, hip_materialize_pairlist, HIPCellList, hip_neighbours,
       hip_neighbours_range

include("nlhip_abi.jl")

"""Device-resident sort-based cell list (the handle keeps the HIP build)."""
struct HIPCellList{TI}
    handle::Handle
    geom::NLHGeometry
    X::Vector{SVector{3,Float64}}
end

function hip_build_cell_list(X::AbstractVector{SVector{3,Float64}}, cutoff::Real, cell::AbstractMatrix, pbc;
                             int_type::Type = Int32, handle::Handle = Handle())
    g = geometry(cell, cutoff, pbc, int_type)
    Xh = Vector{SVector{3,Float64}}(X)
    GC.@preserve Xh begin
        check(ccall((:nlh_build_cell_list, libnlhip), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Int64, Ref{NLHGeometry}, UInt32),
                    handle.ptr, Ptr{Float64}(pointer(Xh)), length(Xh), g, UInt32(0)), "build_cell_list")
    end
    return HIPCellList{int_type}(handle, g, Xh)
end

function hip_materialize_pairlist(c::HIPCellList{TI}; want_R::Bool = false) where {TI}
    n = Ref{Int64}(0)
    check(ccall((:nlh_materialize, libnlhip), Cint, (Ptr{Cvoid}, UInt32, Ptr{Int64}),
                c.handle.ptr, want_R ? NLH_WANT_R : UInt32(0), n), "materialize_pairlist")
    P, nat = n[], length(c.X)
    i = Vector{TI}(undef, P)
    j = Vector{TI}(undef, P)
    S = Vector{SVector{3,TI}}(undef, P)
    first = Vector{TI}(undef, nat + 1)
    R = want_R ? Vector{SVector{3,Float64}}(undef, P) : nothing
    GC.@preserve i j S first R begin
        check(ccall((:nlh_copy_pairlist, libnlhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
                    c.handle.ptr, pointer(i), pointer(j), pointer(S), pointer(first),
                    want_R ? Ptr{Float64}(pointer(R)) : Ptr{Float64}(C_NULL)), "copy pair list")
    end
    C = SMatrix{3,3,Float64}(c.geom.cell)
    nl = PairList(c.X, C, c.geom.cutoff, i, j, S, first)
    return want_R ? (nl, R) : nl
end

"""Copy a device cell list back as the reference's SortedCellList (types.jl:70-82)."""
function to_sorted_cell_list(c::HIPCellList{TI}) where {TI}
    nat, M = length(c.X), Int(c.geom.ncells_total)
    Xs = Vector{SVector{3,Float64}}(undef, nat)
    perm = Vector{TI}(undef, nat)
    cell_id = Vector{TI}(undef, nat)
    offs = Vector{TI}(undef, M + 1)
    GC.@preserve Xs perm cell_id offs begin
        check(ccall((:nlh_copy_celllist, libnlhip), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                    c.handle.ptr, Ptr{Float64}(pointer(Xs)), pointer(perm), pointer(cell_id), pointer(offs)),
              "copy cell list")
    end
    g = c.geom
    return SortedCellList(Xs, c.X, perm, cell_id, offs, SMatrix{3,3,Float64}(g.cell),
                          SMatrix{3,3,Float64}(g.inv_cell), SVector{3,Bool}(g.pbc .!= 0), g.cutoff,
                          SVector{3,TI}(g.ncells), TI(g.ncells_total))
end

"""Neighbours of atom i (1-based) straight from the device cell list
(cell_list.jl:804-833): returns (j, R, S)."""
function hip_neighbours(c::HIPCellList{TI}, i::Integer; system::Integer = 0) where {TI}
    cnt = Ref{Int64}(0)
    check(ccall((:nlh_neighbours, libnlhip), Cint,
                (Ptr{Cvoid}, Int32, Int64, Int64, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Int64}),
                c.handle.ptr, system, i, 0, C_NULL, C_NULL, C_NULL, cnt), "neighbours")
    n = cnt[]
    j = Vector{TI}(undef, n)
    R = Vector{SVector{3,Float64}}(undef, n)
    S = Vector{SVector{3,TI}}(undef, n)
    GC.@preserve j R S begin
        check(ccall((:nlh_neighbours, libnlhip), Cint,
                    (Ptr{Cvoid}, Int32, Int64, Int64, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Int64}),
                    c.handle.ptr, system, i, n, pointer(j), Ptr{Float64}(pointer(R)), pointer(S), cnt),
              "neighbours")
    end
    return j, R, S
end

"""Neighbours of atoms i_first .. i_first + n - 1 (1-based) in one pass on
the device (nlh_neighbours_range): `(offsets, j, R, S)` where atom
`i_first + a - 1` owns entries `offsets[a] + 1 : offsets[a + 1]` -- what
`neighbours(clist, i)` gives for each of them (cell_list.jl:779-833)."""
function hip_neighbours_range(c::HIPCellList{TI}, i_first::Integer, n::Integer; system::Integer = 0) where {TI}
    offs = zeros(Int64, n + 1)
    check(ccall((:nlh_neighbours_range, libnlhip), Cint,
                (Ptr{Cvoid}, Int32, Int64, Int64, Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}),
                c.handle.ptr, system, i_first, n, 0, offs, C_NULL, C_NULL, C_NULL), "neighbours")
    P = offs[end]
    j = Vector{TI}(undef, P)
    R = Vector{SVector{3,Float64}}(undef, P)
    S = Vector{SVector{3,TI}}(undef, P)
    if P > 0   # the results are still on the device: this call only copies them
        GC.@preserve j R S begin
            check(ccall((:nlh_neighbours_range, libnlhip), Cint,
                        (Ptr{Cvoid}, Int32, Int64, Int64, Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}),
                        c.handle.ptr, system, i_first, n, P, offs, pointer(j), Ptr{Float64}(pointer(R)),
                        pointer(S)), "neighbours")
        end
    end
    return offs, j, R, S
end

"""
    hip_neighbour_list(X, cutoff, cell, pbc; lazy=false, int_type=Int32)

The reference's `neighbour_list` (cell_list.jl:897-916) built on the MI355X.
`lazy=true` returns the reference's `SortedCellList`.
"""
function hip_neighbour_list(X::AbstractVector{SVector{3,Float64}}, cutoff::Real, cell::AbstractMatrix, pbc;
                            lazy::Union{Bool,Val} = false, int_type::Type = Int32)
    c = hip_build_cell_list(X, cutoff, cell, pbc; int_type = int_type)
    L = lazy isa Val ? typeof(lazy).parameters[1] : lazy
    return L ? to_sorted_cell_list(c) : hip_materialize_pairlist(c)
end

"""Many systems in one batched build (nlh_batch_neighbour_list); returns one
PairList per system, each identical to a single-system build."""
function hip_batch_neighbour_list(systems::AbstractVector; int_type::Type = Int32, handle::Handle = Handle())
    nsys = length(systems)
    geoms = [geometry(s.cell, s.cutoff, s.pbc, int_type) for s in systems]
    offs = zeros(Int64, nsys + 1)
    for (k, s) in enumerate(systems)
        offs[k+1] = offs[k] + length(s.X)
    end
    X = reduce(vcat, [Vector{SVector{3,Float64}}(s.X) for s in systems])
    total = Ref{Int64}(0)
    GC.@preserve X offs geoms begin
        check(ccall((:nlh_batch_neighbour_list, libnlhip), Cint,
                    (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Int64}, Ptr{NLHGeometry}, UInt32, Ptr{Int64}),
                    handle.ptr, nsys, Ptr{Float64}(pointer(X)), offs, geoms, UInt32(0), total), "batch build")
    end
    po = zeros(Int64, nsys + 1)
    check(ccall((:nlh_batch_pair_offsets, libnlhip), Cint, (Ptr{Cvoid}, Ptr{Int64}), handle.ptr, po), "offsets")
    P = total[]
    i = Vector{int_type}(undef, P); j = Vector{int_type}(undef, P)
    S = Vector{SVector{3,int_type}}(undef, P); first = Vector{int_type}(undef, offs[end] + nsys)
    GC.@preserve i j S first begin
        check(ccall((:nlh_copy_pairlist, libnlhip), Cint,
                    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
                    handle.ptr, pointer(i), pointer(j), pointer(S), pointer(first), Ptr{Float64}(C_NULL)), "copy")
    end
    return map(1:nsys) do k
        a, b = po[k] + 1, po[k+1]
        xa, xb = offs[k] + 1, offs[k+1]
        PairList(X[xa:xb], SMatrix{3,3,Float64}(geoms[k].cell), geoms[k].cutoff, i[a:b], j[a:b], S[a:b],
                 first[xa+k-1:xb+k])
    end
end

end # module
