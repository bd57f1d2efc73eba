"""
    NeighbourListsAMDGPUExt

Package extension of NeighbourLists.jl (weak dependency: AMDGPU) that puts
libnlhip.so behind the reference's own backend dispatch.  With it loaded, an
unchanged

    nlist = neighbour_list(ROCArray(X), cutoff, cell, pbc)            # PairList
    clist = neighbour_list(ROCArray(X), cutoff, cell, pbc; lazy=true)  # SortedCellList

reaches the MI355X library: `get_array_backend(X)` (src/NeighbourLists.jl:17-18)
returns `ROCBackend()` for a `ROCArray`, and the two methods below are more
specific than the reference's generic ones:

* `_build_sorted_celllist(X::ROCVector{SVec{Float64}}, …, ::ROCBackend)` in
  place of `src/cell_list.jl:647-679` (which would call `_compute_cell_ids`
  (`src/gpu_kernels.jl:378-382`), AcceleratedKernels' `sortperm!`
  (`src/cell_list.jl:711-718`) and `_compute_cell_offsets`
  (`src/gpu_kernels.jl:386-390`));
* `materialize_pairlist(::SortedCellList{Float64,TI,<:ROCVector})` in place
  of `src/gpu_kernels.jl:299-364`.

Results are the reference's types holding device arrays of X's array type
(`similar(X, …)`), as the reference returns them: the SortedCellList holds
`X` as `X_orig`, and the PairList aliases `X` (`src/gpu_kernels.jl:362`).
The geometry comes from the reference's own `analyze_cell`
(`src/cell_list.jl:152-170`), the overflow check and error text of
`src/cell_list.jl:655-659` and the neighbour reach of
`src/gpu_kernels.jl:315-316`, handed to the library as they are, so binning
uses StaticArrays' values.  Lazy per-atom queries on a device cell list from
host code (`neighbours`, `count_neighbours`, `for_each_neighbour`) run on the
device too (the reference's host loop would index device arrays element by
element); inside a user kernel the reference's `for_each_neighbour` on the
adapted cell list (`src/adapt.jl`) is unchanged.

Only Float64 positions take this path (the library computes in FP64); other
element types fall through to the reference's KernelAbstractions methods.

To install it in NeighbourLists.jl, copy this file and `nlhip_abi.jl` into
its `ext/` directory and add to its Project.toml:

    [weakdeps]
    AMDGPU = "21141c5a-9bdb-4563-92ae-f87d6854732e"
    [extensions]
    NeighbourListsAMDGPUExt = "AMDGPU"
"""
module NeighbourListsAMDGPUExt

using NeighbourLists
using NeighbourLists: SortedCellList, PairList, SVec, SMat, analyze_cell
using AMDGPU
using AMDGPU: ROCVector, ROCBackend
using StaticArrays: SVector, SMatrix

include(joinpath(@__DIR__, isfile(joinpath(@__DIR__, "nlhip_abi.jl")) ? "" : "..", "nlhip_abi.jl"))

# One handle (its HIP stream and grow-only device buffers) serves every
# build, so repeated builds allocate nothing once the buffers have grown; it
# remembers, weakly, the perm array of the cell list it built last, so that
# materialising that cell list or querying its atoms needs no rebuild.  One
# lock serialises the extension's calls into the library.
mutable struct _Slot
    handle::Handle
    last::WeakRef
end
const _LOCK = ReentrantLock()
const _SLOT = Ref{Union{_Slot,Nothing}}(nothing)

function _slot()   # call with _LOCK held
    _SLOT[] === nothing && (_SLOT[] = _Slot(Handle(), WeakRef(nothing)))
    return _SLOT[]
end

# analyze_cell (cell_list.jl:152-170, warns on a zero volume), the index-type
# overflow check (cell_list.jl:655-659) and nxyz (gpu_kernels.jl:315-316)
function _geometry(cell::SMat{Float64}, pbc::SVec{Bool}, cutoff::Float64, ::Type{TI}) where {TI}
    inv_cell, ncells, lens = analyze_cell(cell, cutoff, TI)
    if prod(BigInt.(ncells)) > typemax(TI)
        error("""Ratio of simulation cell size to cutoff is very large.
                 Use a larger integer type (e.g. Int64), larger cutoff,
                 or smaller simulation cell.""")
    end
    nxyz = ceil.(TI, cutoff * (ncells ./ abs.(lens)))
    return geometry_from(cell, inv_cell, lens, ncells, nxyz, cutoff, pbc, TI), inv_cell, SVec{TI}(ncells)
end

function _build!(h::Handle, X::ROCVector{SVec{Float64}}, g::NLHGeometry)
    AMDGPU.synchronize()   # X was written by work on AMDGPU's own streams
    check(ccall((:nlh_build_cell_list, libnlhip), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int64, Ref{NLHGeometry}, UInt32),
                h.ptr, Ptr{Float64}(pointer(X)), length(X), g, NLH_X_DEVICE), "build_cell_list")
    return nothing
end

function NeighbourLists._build_sorted_celllist(X::ROCVector{SVec{Float64}}, cell::SMat{Float64},
                                               pbc::SVec{Bool}, cutoff::Float64, ::Type{TI},
                                               backend::ROCBackend) where {TI <: Union{Int32,Int64}}
    g, inv_cell, ncells = _geometry(cell, pbc, cutoff, TI)
    nat = length(X)
    Xs = similar(X, nat)
    perm = similar(X, TI, nat)
    cell_id = similar(X, TI, nat)
    cell_offsets = similar(X, TI, Int(g.ncells_total) + 1)
    lock(_LOCK) do
        s = _slot()
        s.last = WeakRef(nothing)
        _build!(s.handle, X, g)
        # device-to-device copies into the Julia-owned arrays
        check(ccall((:nlh_copy_celllist, libnlhip), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                    s.handle.ptr, Ptr{Float64}(pointer(Xs)), pointer(perm), pointer(cell_id),
                    pointer(cell_offsets)), "copy cell list")
        s.last = WeakRef(perm)
    end
    return SortedCellList{Float64, TI, typeof(Xs), typeof(perm)}(Xs, X, perm, cell_id, cell_offsets, cell,
                                                                inv_cell, pbc, cutoff, ncells, TI(g.ncells_total))
end

# `f(handle)` on a handle holding `clist`'s device build: the shared handle
# when it built this cell list last, else after building it again from
# X_orig with the cell list's own geometry.  Call with _LOCK held.
function _with_built(f, clist::SortedCellList{Float64,TI,<:ROCVector}) where {TI}
    s = _slot()
    if s.last.value !== clist.perm
        lens = NeighbourLists.lengths(clist.cell)
        nxyz = ceil.(TI, clist.cutoff * (clist.ncells ./ abs.(lens)))
        g = geometry_from(clist.cell, clist.inv_cell, lens, clist.ncells, nxyz, clist.cutoff, clist.pbc, TI)
        s.last = WeakRef(nothing)
        _build!(s.handle, clist.X_orig, g)
        s.last = WeakRef(clist.perm)
    end
    return f(s.handle)
end

function NeighbourLists.materialize_pairlist(clist::SortedCellList{Float64,TI,AT,VI};
                                             backend = NeighbourLists.get_array_backend(clist.X_orig)
                                             ) where {TI, AT <: ROCVector, VI}
    nat = length(clist.X_orig)
    return lock(_LOCK) do
        _with_built(clist) do h
            n = Ref{Int64}(0)
            check(ccall((:nlh_materialize, libnlhip), Cint, (Ptr{Cvoid}, UInt32, Ptr{Int64}),
                        h.ptr, UInt32(0), n), "materialize_pairlist")
            P = n[]
            i_arr = similar(clist.X_orig, TI, P)
            j_arr = similar(clist.X_orig, TI, P)
            S_arr = similar(clist.X_orig, SVec{TI}, P)
            first_arr = similar(clist.X_orig, TI, nat + 1)
            check(ccall((:nlh_copy_pairlist, libnlhip), Cint,
                        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
                        h.ptr, P > 0 ? pointer(i_arr) : C_NULL, P > 0 ? pointer(j_arr) : C_NULL,
                        P > 0 ? pointer(S_arr) : C_NULL, pointer(first_arr), Ptr{Float64}(C_NULL)),
                  "copy pair list")
            # the PairList aliases the caller's positions (gpu_kernels.jl:362)
            PairList{Float64, TI, AT, typeof(i_arr), typeof(S_arr)}(clist.X_orig, clist.cell, clist.cutoff,
                                                                    i_arr, j_arr, S_arr, first_arr)
        end
    end
end

# Lazy queries from host code (cell_list.jl:779-833), answered on the device
# by nlh_neighbours_range; results are host vectors.
function _neighbours_of(clist::SortedCellList{Float64,TI,<:ROCVector}, i::Integer) where {TI}
    1 <= i <= length(clist.X_orig) || throw(BoundsError(clist.X_orig, i))
    return lock(_LOCK) do
        _with_built(clist) do h
            offs = zeros(Int64, 2)
            check(ccall((:nlh_neighbours_range, libnlhip), Cint,
                        (Ptr{Cvoid}, Int32, Int64, Int64, Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cvoid}),
                        h.ptr, 0, i, 1, 0, offs, C_NULL, C_NULL, C_NULL), "neighbours")
            P = offs[2]
            js = Vector{TI}(undef, P)
            Rs = Vector{SVec{Float64}}(undef, P)
            Ss = Vector{SVec{TI}}(undef, P)
            if P > 0
                GC.@preserve js Rs Ss begin
                    check(ccall((:nlh_neighbours_range, libnlhip), Cint,
                                (Ptr{Cvoid}, Int32, Int64, Int64, Int64, Ptr{Int64}, Ptr{Cvoid}, Ptr{Float64},
                                 Ptr{Cvoid}),
                                h.ptr, 0, i, 1, P, offs, pointer(js), Ptr{Float64}(pointer(Rs)), pointer(Ss)),
                          "neighbours")
                end
            end
            (js, Rs, Ss)
        end
    end
end

NeighbourLists.neighbours(clist::SortedCellList{Float64,TI,<:ROCVector}, i::Integer) where {TI} =
    _neighbours_of(clist, i)

NeighbourLists.count_neighbours(clist::SortedCellList{Float64,TI,<:ROCVector}, i::Integer) where {TI} =
    length(_neighbours_of(clist, i)[1])

function NeighbourLists.for_each_neighbour(f::F, clist::SortedCellList{Float64,TI,<:ROCVector},
                                           i::Integer) where {F, TI}
    js, Rs, Ss = _neighbours_of(clist, i)
    for q in eachindex(js)
        f(js[q], Rs[q], Ss[q])
    end
    return nothing
end

end # module
