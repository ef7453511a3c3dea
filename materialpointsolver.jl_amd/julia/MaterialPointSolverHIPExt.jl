#==========================================================================================+
  MaterialPointSolverHIPExt.jl -- binds libmpm_hip.so (include/mpm_hip.h) as the device
  symbol :HIP for MaterialPointSolver.jl.

  Drop this file next to the package's other extensions (ext/) and register it in
  Project.toml as an extension with no trigger package beyond Libdl, or `include` it
  after `using MaterialPointSolver`.  Set ENV["MPM_HIP_LIB"] to the library path.

  What it adds (all by multiple dispatch, nothing in the package is edited):
    host2device(grid, mp, attr, bc, Val(:HIP))      -> HIP handles (mpm_create + uploads)
    procedure!(args, ::HIPGrid, ::HIPParticle, ::HIPProperty, ::HIPVBoundary,
               ΔT, Ti, Val(:OS), Val(scheme), Val(va)) -> one mpm_step
               (scheme :MUSL | :USL | :USF, va :a | :v; one method per pair, so each
               is strictly more specific than the package's own methods)
    device2host!(args, mp, ::HIPParticle, Val(:HIP)) -> mpm_download_particles
    clean_device!(::HIPGrid, ..., Val(:HIP))          -> mpm_destroy
    getBackend(Val(:HIP))                             -> KernelAbstractions.CPU()
                                                         (KAsync is then a no-op; the
                                                         download synchronises)
    Tpeak(Val(:HIP))                                  -> mpm_tpeak
  Args2D/Args3D validate `device` against a fixed set (src/types/modelargs.jl:72,163):
  add :HIP there, or build Args with the default (unvalidated) constructor.
  Julia is not installed where this library was built; the same call sequence is
  exercised from Python through ctypes (materialpointsolver.jl_amd/solver.py).
+==========================================================================================#
module MaterialPointSolverHIPExt

using MaterialPointSolver
using KernelAbstractions: CPU
import MaterialPointSolver: host2device, device2host!, clean_device!, getBackend, Tpeak,
    procedure!
using MaterialPointSolver: DeviceArgs, DeviceGrid, DeviceParticle, DeviceProperty,
    DeviceVBoundary

const LIB = get(ENV, "MPM_HIP_LIB", joinpath(@__DIR__, "..", "libmpm_hip.so"))

# ---- C structs (field order and types of include/mpm_hip.h) -----------------------------
struct MPMConfig
    dim::Int32; precision::Int32; basis::Int32; sort_interval::Int32
    np::Int64; nmat::Int64; nnx::Int64; nny::Int64; nnz::Int64
    x1::Float64; y1::Float64; z1::Float64
    dx::Float64; dy::Float64; dz::Float64
    lpx::Float64; lpy::Float64; lpz::Float64
end

struct MPMArgs
    scheme::Int32; va::Int32; constitutive::Int32; mvl::Int32
    flip::Float64; pic::Float64; zeta_s::Float64; gravity::Float64; Te::Float64
    gravity_x::Float64; gravity_y::Float64   # extension; 0 = the package's gravity
end

# body force on the other axes, read from the user-extension slot when present
extra_gravity(ext, name) = hasproperty(ext, name) ? Float64(getproperty(ext, name)) : 0.0

const CONSTITUTIVE = Dict(:linearelastic => Int32(0), :druckerprager => Int32(1),
                          :mohrcoulomb => Int32(2))
const SCHEME = Dict(:MUSL => Int32(0), :USL => Int32(1), :USF => Int32(2))

# ---- device-side handles: the abstract supertypes the package dispatches on -------------
mutable struct HIPContext
    ptr::Ptr{Cvoid}
end

struct HIPGrid{T1, T2} <: DeviceGrid{T1, T2}
    ctx::HIPContext
end
struct HIPParticle{T1, T2} <: DeviceParticle{T1, T2}
    ctx::HIPContext
    np ::T1
    dim::Int
end
struct HIPProperty{T1, T2} <: DeviceProperty{T1, T2}
    ctx::HIPContext
end
struct HIPVBoundary{T1, T2} <: DeviceVBoundary{T1, T2}
    ctx::HIPContext
end

function check(ctx::Ptr{Cvoid}, rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:mpm_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx))
    error("libmpm_hip ($(rc)): $(msg)")
end

f64(a) = convert(Array{Float64}, a)          # column-major (np, k) kept as is
i64(a) = convert(Array{Int64}, a)

function host2device(
    grid::DeviceGrid{T1, T2}, mp::DeviceParticle{T1, T2}, attr::DeviceProperty{T1, T2},
    bc::DeviceVBoundary{T1, T2}, ::Val{:HIP}; sort_interval::Integer=10
) where {T1, T2}
    dim = size(mp.ξ, 2)
    lin = grid.NIC in (4, 8)
    cfg = MPMConfig(Int32(dim), T2 == Float32 ? Int32(1) : Int32(0), lin ? Int32(1) : Int32(0),
        Int32(sort_interval), Int64(mp.np), Int64(length(attr.Ks)),
        Int64(grid.nnx), Int64(grid.nny), dim == 3 ? Int64(grid.nnz) : Int64(1),
        Float64(grid.x1), Float64(grid.y1), dim == 3 ? Float64(grid.z1) : 0.0,
        Float64(grid.dx), Float64(grid.dy), dim == 3 ? Float64(grid.dz) : 1.0,
        Float64(mp.dx), Float64(mp.dy), dim == 3 ? Float64(mp.dz) : 1.0)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(C_NULL, ccall((:mpm_create, LIB), Cint, (Ref{MPMConfig}, Ref{Ptr{Cvoid}}), cfg, out))
    ctx = HIPContext(out[])
    p = ctx.ptr
    check(p, ccall((:mpm_upload_particles, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Int64}),
        p, f64(mp.ξ), f64(mp.vs), f64(mp.σij), f64(mp.ϵijs), f64(mp.F), f64(mp.Ω),
        f64(mp.Ω0), f64(mp.ρs), f64(mp.ρs0), f64(mp.ϵq), f64(mp.ϵk), i64(attr.nid)))
    check(p, ccall((:mpm_upload_materials, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        p, f64(attr.Ks), f64(attr.Gs), f64(attr.ϕ), f64(attr.ψ), f64(attr.c), f64(attr.σt),
        f64(attr.Hp), f64(attr.cr), f64(attr.ϕr)))
    vz_idx = dim == 3 ? i64(bc.vz_s_idx) : C_NULL
    vz_val = dim == 3 ? f64(bc.vz_s_val) : C_NULL
    check(p, ccall((:mpm_upload_bc, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Int64}, Ptr{Float64}, Ptr{Int64}, Ptr{Float64}, Ptr{Int64},
         Ptr{Float64}),
        p, i64(bc.vx_s_idx), f64(bc.vx_s_val), i64(bc.vy_s_idx), f64(bc.vy_s_val),
        vz_idx, vz_val))
    return (HIPGrid{T1, T2}(ctx), HIPParticle{T1, T2}(ctx, mp.np, dim),
            HIPProperty{T1, T2}(ctx), HIPVBoundary{T1, T2}(ctx))
end

function hip_step!(args, grid::HIPGrid, ΔT, Ti, scheme::Symbol, va::Symbol)
    a = MPMArgs(SCHEME[scheme], va == :a ? Int32(0) : Int32(1), CONSTITUTIVE[args.constitutive],
        Int32(args.MVL), Float64(args.FLIP), Float64(args.PIC), Float64(args.ζs),
        Float64(args.gravity), Float64(args.Te), extra_gravity(args.ext, :gravity_x),
        extra_gravity(args.ext, :gravity_y))
    p = grid.ctx.ptr
    check(p, ccall((:mpm_step, LIB), Cint, (Ptr{Cvoid}, Ref{MPMArgs}, Cdouble, Cdouble),
        p, a, Float64(ΔT), Float64(Ti)))
    return nothing
end

# OS_MUSL.jl:12-102, OS_USL.jl:102-144,325-366, OS_USF.jl:11-97
for S in (:MUSL, :USL, :USF), VA in (:a, :v)
    @eval function procedure!(
        args::DeviceArgs{T1, T2}, grid::HIPGrid{T1, T2}, mp::HIPParticle{T1, T2},
        attr::HIPProperty{T1, T2}, bc::HIPVBoundary{T1, T2}, ΔT::T2, Ti::T2,
        ::Val{:OS}, ::Val{$(QuoteNode(S))}, ::Val{$(QuoteNode(VA))}
    ) where {T1, T2}
        return hip_step!(args, grid, ΔT, Ti, $(QuoteNode(S)), $(QuoteNode(VA)))
    end
end

function device2host!(
    args::DeviceArgs{T1, T2}, mp::DeviceParticle{T1, T2}, dev_mp::HIPParticle{T1, T2},
    ::Val{:HIP}; verbose::Bool=false
) where {T1, T2}
    np, d = Int(dev_mp.np), dev_mp.dim
    ns = d == 3 ? 6 : 4
    ξ, vs = zeros(np, d), zeros(np, d)
    σij, ϵijs = zeros(np, ns), zeros(np, ns)
    σm, Ω, ϵq, ϵk, ϵv, ms = (zeros(np) for _ in 1:6)
    p = dev_mp.ctx.ptr
    check(p, ccall((:mpm_download_particles, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        p, ξ, vs, σij, σm, Ω, ϵq, ϵk, ϵv, ϵijs, ms))
    copyto!(mp.σm, σm); copyto!(mp.vs, vs); copyto!(mp.ξ, ξ); copyto!(mp.Ω, Ω)
    copyto!(mp.σij, σij); copyto!(mp.ϵk, ϵk); copyto!(mp.ϵq, ϵq); copyto!(mp.ϵv, ϵv)
    copyto!(mp.ϵijs, ϵijs); copyto!(mp.ms, ms)
    verbose && println("[▼ I/O: downloading from :HIP → host]")
    return nothing
end

function clean_device!(
    grid::HIPGrid{T1, T2}, mp::HIPParticle{T1, T2}, attr::HIPProperty{T1, T2},
    bc::HIPVBoundary{T1, T2}, ::Val{:HIP}
) where {T1, T2}
    if grid.ctx.ptr != C_NULL
        ccall((:mpm_destroy, LIB), Cvoid, (Ptr{Cvoid},), grid.ctx.ptr)
        grid.ctx.ptr = C_NULL
    end
    return nothing
end

getBackend(::Val{:HIP}) = CPU()

function Tpeak(::Val{:HIP}; datatype::Symbol=:FP32, ID::Int=0)
    gbps = Ref{Float64}(0.0)
    check(C_NULL, ccall((:mpm_tpeak, LIB), Cint, (Int64, Cint, Ref{Float64}),
        Int64(1) << 30, Cint(10), gbps))
    println("[ Tpeak: $(round(gbps[] / 1.073741824, digits=2)) GiB/s on :HIP ]")
    return gbps[]
end

end # module
