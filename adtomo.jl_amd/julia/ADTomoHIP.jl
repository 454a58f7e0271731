# ADTomoHIP: Julia bindings of libadtomo_hip.so (MI355X / gfx950).
#
# Drop-in for the two ops of src/eikonal_op.jl:
#   eikonal(f, srcx, srcy, h)                    src/eikonal_op.jl:3-21
#   eikonal3d(u0, f, h, m, n, l, tol, verbose)   src/eikonal_op.jl:24-32
# with the same argument meaning and result layout, and ChainRulesCore
# rrules standing in for the ADCME gradient ops EikonalGrad
# (deps/CustomOps/Eikonal/Eikonal.cpp:48-63) and EikonalThreeDGrad
# (deps/CustomOps/Eikonal3D/EikonalThreeD.cpp:47-65).
#
# Arrays are plain Julia `Array{Float64}` (host memory, staged by the library
# with ADT_HOST_POINTERS).  The reference's TF tensors are row-major, so the
# 3D op sees flat index i*n*l + j*l + k: a Julia (m,n,l) array is passed as
# permutedims(A, (3,2,1)) and results are permuted back; the 2D (n+1, m+1)
# matrix likewise goes through permutedims(f) (flat j*(m+1)+i).
module ADTomoHIP

using ChainRulesCore

export eikonal, eikonal3d

const libadt = normpath(joinpath(@__DIR__, "..", "libadtomo_hip.so"))
const ADT_HOST_POINTERS = Int32(0x1)

# Mirrors `adt_info` in include/adtomo_hip.h.
struct AdtInfo
    iterations::Int32
    converged::Int32
    err::Float64
    ambiguous::Int32
    adj_sweeps::Int32
    adj_converged::Int32
    reserved::Int32
    n_singular::Int64
    err_log::NTuple{100,Float64}
end
AdtInfo() = AdtInfo(0, 0, 0.0, 0, 0, 0, 0, 0, ntuple(_ -> 0.0, 100))

function check(rc::Integer, what::AbstractString)
    if rc != 0
        msg = unsafe_string(ccall((:adt_last_error, libadt), Cstring, ()))
        error("$what failed (rc=$rc): $msg")
    end
end

# Julia column-major <-> reference row-major flat layout.
rowmajor(A::AbstractArray{Float64,3}) = permutedims(A, (3, 2, 1))
rowmajor(A::AbstractMatrix{Float64}) = permutedims(A, (2, 1))
colmajor3(B::Array{Float64,3}) = permutedims(B, (3, 2, 1))
colmajor2(B::Matrix{Float64}) = permutedims(B, (2, 1))

# ------------------------------------------------------------------ 2D ---

"""
    eikonal(f, srcx, srcy, h)

Travel times on the (n+1) x (m+1) grid `f` (rows = depth), 1-based source
column `srcx` and row `srcy`, spacing `h` (src/eikonal_op.jl:3-21).
"""
function eikonal(f::Matrix{Float64}, srcx::Int64, srcy::Int64, h::Float64)
    n_, m_ = size(f)
    m, n = m_ - 1, n_ - 1
    fr = rowmajor(f)                      # flat j*(m+1)+i
    ur = similar(fr)
    info = Ref(AdtInfo())
    rc = ccall((:adt_eikonal2d_fwd_f64, libadt), Cint,
               (Ptr{Float64}, Int64, Int64, Int64, Int64, Float64, Ptr{Float64},
                Ref{AdtInfo}, Int32, Ptr{Cvoid}),
               fr, m, n, srcx, srcy, h, ur, info, ADT_HOST_POINTERS, C_NULL)
    check(rc, "eikonal")
    info[].converged == 0 && @warn "Eikonal does not converge!"
    return colmajor2(ur)
end

function eikonal_grad(grad_u::Matrix{Float64}, u::Matrix{Float64}, f::Matrix{Float64},
                      srcx::Int64, srcy::Int64, h::Float64)
    n_, m_ = size(f)
    m, n = m_ - 1, n_ - 1
    gr, ur, fr = rowmajor(grad_u), rowmajor(u), rowmajor(f)
    gfr = similar(fr)
    info = Ref(AdtInfo())
    rc = ccall((:adt_eikonal2d_bwd_f64, libadt), Cint,
               (Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Int64, Int64, Int64, Float64,
                Ptr{Float64}, Ref{AdtInfo}, Int32, Ptr{Cvoid}),
               gr, ur, fr, m, n, srcx, srcy, h, gfr, info, ADT_HOST_POINTERS, C_NULL)
    check(rc, "eikonal_grad")
    return colmajor2(gfr)
end

function ChainRulesCore.rrule(::typeof(eikonal), f::Matrix{Float64}, srcx::Int64, srcy::Int64,
                              h::Float64)
    u = eikonal(f, srcx, srcy, h)
    function eikonal_pullback(ū)
        gf = eikonal_grad(Matrix{Float64}(unthunk(ū)), u, f, srcx, srcy, h)
        # The reference op leaves the gradients of srcx, srcy and h unset.
        return NoTangent(), gf, NoTangent(), NoTangent(), NoTangent()
    end
    return u, eikonal_pullback
end

# ------------------------------------------------------------------ 3D ---

"""
    eikonal3d(u0, f, h, m, n, l, tol, verbose)

Fast-sweeping travel times from initial times `u0` and slowness `f`, both
m x n x l (src/eikonal_op.jl:24-32).  Returns an m x n x l array.
"""
function eikonal3d(u0::AbstractArray{Float64}, f::AbstractArray{Float64}, h::Float64,
                   m::Integer, n::Integer, l::Integer, tol::Float64, verbose::Bool)
    u0r = rowmajor(reshape(u0, m, n, l))
    fr = rowmajor(reshape(f, m, n, l))
    ur = similar(fr)
    info = Ref(AdtInfo())
    rc = ccall((:adt_eikonal3d_fwd_f64, libadt), Cint,
               (Ptr{Float64}, Ptr{Float64}, Float64, Int64, Int64, Int64, Float64, Int32,
                Ptr{Float64}, Ref{AdtInfo}, Int32, Ptr{Cvoid}),
               u0r, fr, h, m, n, l, tol, Int32(verbose), ur, info, ADT_HOST_POINTERS, C_NULL)
    check(rc, "eikonal3d")
    return colmajor3(ur)
end

function eikonal3d_grad(grad_u, u, u0, f, h::Float64, m::Integer, n::Integer, l::Integer)
    gr = rowmajor(reshape(Array{Float64}(grad_u), m, n, l))
    ur, u0r, fr = rowmajor(u), rowmajor(reshape(u0, m, n, l)), rowmajor(reshape(f, m, n, l))
    gu0r, gfr = similar(fr), similar(fr)
    info = Ref(AdtInfo())
    rc = ccall((:adt_eikonal3d_bwd_f64, libadt), Cint,
               (Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Int64, Int64,
                Int64, Ptr{Float64}, Ptr{Float64}, Ref{AdtInfo}, Int32, Ptr{Cvoid}),
               gr, ur, u0r, fr, h, m, n, l, gu0r, gfr, info, ADT_HOST_POINTERS, C_NULL)
    check(rc, "eikonal3d_grad")
    return colmajor3(gu0r), colmajor3(gfr)
end

function ChainRulesCore.rrule(::typeof(eikonal3d), u0, f, h::Float64, m::Integer, n::Integer,
                              l::Integer, tol::Float64, verbose::Bool)
    u = eikonal3d(u0, f, h, m, n, l, tol, verbose)
    function eikonal3d_pullback(ū)
        gu0, gf = eikonal3d_grad(unthunk(ū), u, u0, f, h, m, n, l)
        # Only grad_u0 and grad_f are meaningful in the reference
        # (EikonalThreeD.cpp:154-250); h, dims, tol and verbose get none.
        return (NoTangent(), reshape(gu0, size(u0)), reshape(gf, size(f)), NoTangent(),
                NoTangent(), NoTangent(), NoTangent(), NoTangent(), NoTangent())
    end
    return u, eikonal3d_pullback
end

end # module
