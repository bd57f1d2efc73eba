# Raw ccall bindings of libnlhip.so (the C ABI declared in include/nlhip.h),
# shared by the NeighbourListsHIP module and the NeighbourListsAMDGPUExt
# package extension: library path, status codes, the geometry struct, the
# handle type and the geometry helpers.
#
# Set ENV["NLHIP_LIB"] to the library path before loading, or keep
# libnlhip.so at its in-tree place (neighbourlists.jl_amd/nlhip/).

const libnlhip = get(ENV, "NLHIP_LIB", joinpath(@__DIR__, "..", "nlhip", "libnlhip.so"))

const NLH_OK, NLH_EINVAL, NLH_EOVERFLOW, NLH_ENOMEM, NLH_EDEVICE, NLH_ESTATE = 0, 1, 2, 3, 4, 5
const NLH_X_DEVICE = UInt32(0x1)
const NLH_WANT_R = UInt32(0x2)

# nlh_geometry (include/nlhip.h): identical C layout, 248 bytes
struct NLHGeometry
    cell::NTuple{9,Float64}
    inv_cell::NTuple{9,Float64}
    lens::NTuple{3,Float64}
    cutoff::Float64
    cutoff_sq::Float64
    ncells::NTuple{3,Int64}
    nxyz::NTuple{3,Int64}
    ncells_total::Int64
    pbc::NTuple{3,UInt8}
    zero_volume::UInt8
    ti_bytes::Int32
end

struct NLHCellListView
    nat::Int64
    ncells_total::Int64
    X_orig::Ptr{Float64}
    perm::Ptr{Cvoid}
    cell_id::Ptr{Cvoid}
    cell_offsets::Ptr{Cvoid}
    ti_bytes::Int32
end

last_error() = unsafe_string(ccall((:nlh_last_error, libnlhip), Cstring, ()))

function check(status::Integer, what::AbstractString)
    status == NLH_OK && return nothing
    msg = string(what, ": ", last_error())
    status == NLH_EOVERFLOW && throw(OverflowError(msg))   # cell_list.jl:655-659
    status == NLH_EINVAL && throw(ArgumentError(msg))
    status == NLH_ENOMEM && throw(OutOfMemoryError())
    error(msg)
end

const _initialised = Ref(false)
function init(device::Integer = 0)
    check(ccall((:nlh_init, libnlhip), Cint, (Cint,), device), "nlh_init")
    _initialised[] = true
    nothing
end

"""Owns one `nlh_handle` (a HIP stream and its device buffers); freed by a finalizer."""
mutable struct Handle
    ptr::Ptr{Cvoid}
    function Handle()
        _initialised[] || init(0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:nlh_create, libnlhip), Cint, (Ptr{Ptr{Cvoid}},), r), "nlh_create")
        h = new(r[])
        finalizer(h) do x
            x.ptr == C_NULL || ccall((:nlh_destroy, libnlhip), Cint, (Ptr{Cvoid},), x.ptr)
            x.ptr = C_NULL
        end
        return h
    end
end

ti_bytes(::Type{Int32}) = Int32(4)
ti_bytes(::Type{Int64}) = Int32(8)
ti_bytes(T::Type) = throw(ArgumentError("int_type must be Int32 or Int64, got $T"))

_pbc(pbc::Bool) = (UInt8(pbc), UInt8(pbc), UInt8(pbc))
_pbc(pbc) = (UInt8(pbc[1]), UInt8(pbc[2]), UInt8(pbc[3]))

"""Geometry via the library's host restatement of analyze_cell (cell_list.jl:152-170).
The SMatrix is column-major, which is the ABI's cell layout."""
function geometry(cell::AbstractMatrix, cutoff::Real, pbc, int_type::Type)
    C = SMatrix{3,3,Float64}(cell)
    g = Ref{NLHGeometry}()
    p = collect(_pbc(pbc))
    check(ccall((:nlh_geometry_init, libnlhip), Cint,
                (Ptr{Float64}, Float64, Ptr{UInt8}, Int32, Ptr{NLHGeometry}),
                collect(C), Float64(cutoff), p, ti_bytes(int_type), g), "geometry")
    g[].zero_volume != 0 && @warn("Zero volume cell detected. Results may be unreliable.")
    return g[]
end

"""Geometry from Julia's own cell analysis -- `inv_cell`, `ncells`, `lens`
of `analyze_cell` (cell_list.jl:152-170) and the neighbour reach `nxyz` of
`materialize_pairlist` (gpu_kernels.jl:315-316) -- so that binning on the
device uses StaticArrays' exact values."""
function geometry_from(cell::SMatrix{3,3,Float64}, inv_cell::SMatrix{3,3,Float64}, lens, ncells, nxyz,
                       cutoff::Float64, pbc, ::Type{TI}) where {TI}
    n = (Int64(ncells[1]), Int64(ncells[2]), Int64(ncells[3]))
    return NLHGeometry(Tuple(cell), Tuple(inv_cell), (Float64(abs(lens[1])), Float64(abs(lens[2])), Float64(abs(lens[3]))),
                       cutoff, cutoff^2, n, (Int64(nxyz[1]), Int64(nxyz[2]), Int64(nxyz[3])), n[1] * n[2] * n[3],
                       _pbc(pbc), UInt8(0), ti_bytes(TI))
end
