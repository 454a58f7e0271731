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
# The single-event ops take plain Julia `Array{Float64}` (host memory, staged
# by the library with ADT_HOST_POINTERS) through the library's _strided entry
# points: element (i, j, k) of a column-major (m, n, l) Array sits at
# strides (1, m, m*n), and (row j, column i) of an (n+1, m+1) Matrix at
# (i: n+1, j: 1), so arrays pass as they are, without the transposed copies
# of the reference's row-major tensors.  The batched, graph and objective
# calls below take device arrays in the reference's flat layout.
module ADTomoHIP

using ChainRulesCore
using Libdl

export eikonal, eikonal3d, eikonal3d_batched!, EikonalGraph, launch!, close!, source_u0!,
       misfit!, model!, model_grad!, slowness!, slowness_grad!, smooth_l1!, sum_arrays!,
       AdtTune, AdtTuneClass, get_tune, set_tune!, set_tune_class!, tune_classes, save_tune,
       load_tune, autotune!, eikonal3d_misfit_batched!

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
    fault::Int32
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

# Element strides of column-major arrays, in the library's axis order.
strides3(m, n) = Int64[1, m, m * n]                 # (i, j, k) of an (m, n, l) Array
strides2(n) = Int64[n + 1, 1]                       # (i, j) of an (n+1, m+1) Matrix
dense3(A, m, n, l) = A isa Array{Float64,3} && size(A) == (m, n, l) ? A :
                     Array{Float64,3}(reshape(A, m, n, l))
dense2(A) = A isa Matrix{Float64} ? A : Matrix{Float64}(A)

# ------------------------------------------------------------------ 2D ---

"""
    eikonal(f, srcx, srcy, h)

Travel times on the (n+1) x (m+1) grid `f` (rows = depth), 1-based source
column `srcx` and row `srcy`, spacing `h` (src/eikonal_op.jl:3-21).
"""
function eikonal(f::Matrix{Float64}, srcx::Int64, srcy::Int64, h::Float64)
    n_, m_ = size(f)
    m, n = m_ - 1, n_ - 1
    u = similar(f)
    info = Ref(AdtInfo())
    rc = ccall((:adt_eikonal2d_fwd_strided_f64, libadt), Cint,
               (Ptr{Float64}, Int64, Int64, Int64, Int64, Float64, Ptr{Int64}, Ptr{Float64},
                Ref{AdtInfo}, Int32, Ptr{Cvoid}),
               f, m, n, srcx, srcy, h, strides2(n), u, info, ADT_HOST_POINTERS, C_NULL)
    check(rc, "eikonal")
    info[].converged == 0 && @warn "Eikonal does not converge!"
    return u
end

function eikonal_grad(grad_u::Matrix{Float64}, u::Matrix{Float64}, f::Matrix{Float64},
                      srcx::Int64, srcy::Int64, h::Float64)
    n_, m_ = size(f)
    m, n = m_ - 1, n_ - 1
    g = dense2(grad_u)
    gf = similar(f)
    info = Ref(AdtInfo())
    rc = ccall((:adt_eikonal2d_bwd_strided_f64, libadt), Cint,
               (Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Int64, Int64, Int64, Float64,
                Ptr{Int64}, Ptr{Float64}, Ref{AdtInfo}, Int32, Ptr{Cvoid}),
               g, u, f, m, n, srcx, srcy, h, strides2(n), gf, info, ADT_HOST_POINTERS, C_NULL)
    check(rc, "eikonal_grad")
    return gf
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
    u0a, fa = dense3(u0, m, n, l), dense3(f, m, n, l)
    u = similar(fa)
    info = Ref(AdtInfo())
    rc = ccall((:adt_eikonal3d_fwd_strided_f64, libadt), Cint,
               (Ptr{Float64}, Ptr{Float64}, Float64, Int64, Int64, Int64, Float64, Int32,
                Ptr{Int64}, Ptr{Float64}, Ref{AdtInfo}, Int32, Ptr{Cvoid}),
               u0a, fa, h, m, n, l, tol, Int32(verbose), strides3(m, n), u, info,
               ADT_HOST_POINTERS, C_NULL)
    check(rc, "eikonal3d")
    return u
end

function eikonal3d_grad(grad_u, u, u0, f, h::Float64, m::Integer, n::Integer, l::Integer)
    g, ua = dense3(grad_u, m, n, l), dense3(u, m, n, l)
    u0a, fa = dense3(u0, m, n, l), dense3(f, m, n, l)
    gu0, gf = similar(fa), similar(fa)
    info = Ref(AdtInfo())
    rc = ccall((:adt_eikonal3d_bwd_strided_f64, libadt), Cint,
               (Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Int64, Int64,
                Int64, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ref{AdtInfo}, Int32, Ptr{Cvoid}),
               g, ua, u0a, fa, h, m, n, l, strides3(m, n), gu0, gf, info, ADT_HOST_POINTERS,
               C_NULL)
    check(rc, "eikonal3d_grad")
    return gu0, gf
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

# ---------------------------------------------- batched, device arrays ---
#
# The per-station op loop of scripts/inversion.jl:47-62 as one launch over
# device arrays (e.g. AMDGPU.jl `ROCArray`s; anything whose `pointer` is a
# device address).  Device arrays carry the reference's flat layout: a grid
# (m, n, l) is a column-major array of size (l, n, m), so its flat index is
# i*n*l + j*l + k.  `stream` is a HIP stream handle (C_NULL: the null stream).

const ADT_FORWARD = Int32(0x2)
const ADT_ADJOINT = Int32(0x4)
const ADT_F32_COMPUTE_F64 = Int32(0x8)

# Mirrors `adt_event3d` in include/adtomo_hip.h.
struct AdtEvent3d
    m::Int64
    n::Int64
    l::Int64
    h::Float64
    tol::Float64
    u0::Ptr{Cvoid}
    f::Ptr{Cvoid}
    grad_u::Ptr{Cvoid}
    u::Ptr{Cvoid}
    grad_u0::Ptr{Cvoid}
    grad_f::Ptr{Cvoid}
end

devptr(A) = A === nothing ? C_NULL : Ptr{Cvoid}(UInt(pointer(A)))
griddims(A) = (size(A, 3), size(A, 2), size(A, 1))   # (m, n, l) of an (l, n, m) array
sfx(A) = eltype(A) == Float32 ? "_f32" : "_f64"

function event_table(u0s, fs, us, hs, tols, grad_us, grad_u0s, grad_fs)
    return [AdtEvent3d(griddims(fs[e])..., Float64(hs[e]), Float64(tols[e]), devptr(u0s[e]),
                       devptr(fs[e]), grad_us === nothing ? C_NULL : devptr(grad_us[e]),
                       devptr(us[e]), grad_u0s === nothing ? C_NULL : devptr(grad_u0s[e]),
                       grad_fs === nothing ? C_NULL : devptr(grad_fs[e]))
            for e in eachindex(fs)]
end

"""
    eikonal3d_batched!(us, u0s, fs, hs, tols; grad_us, grad_u0s, grad_fs,
                       forward=true, compute_f64=false, stream=C_NULL)

Forward (and, with `grad_us`/`grad_fs`, adjoint) of every event in one
launch; returns the per-event `AdtInfo`s.  Boxes may differ per event.
"""
function eikonal3d_batched!(us, u0s, fs, hs, tols; grad_us=nothing, grad_u0s=nothing,
                            grad_fs=nothing, forward::Bool=true, compute_f64::Bool=false,
                            stream::Ptr{Cvoid}=C_NULL)
    tab = event_table(u0s, fs, us, hs, tols, grad_us, grad_u0s, grad_fs)
    flags = (forward ? ADT_FORWARD : Int32(0)) | (grad_us === nothing ? Int32(0) : ADT_ADJOINT) |
            (compute_f64 ? ADT_F32_COMPUTE_F64 : Int32(0))
    infos = [AdtInfo() for _ in tab]
    fn = Symbol("adt_eikonal3d_batched", sfx(fs[1]))
    rc = ccall(Libdl.dlsym(Libdl.dlopen(libadt), fn), Cint,
               (Ptr{AdtEvent3d}, Int32, Int32, Ptr{AdtInfo}, Ptr{Cvoid}),
               tab, length(tab), flags, infos, stream)
    check(rc, "eikonal3d_batched")
    return infos
end

"""
    EikonalGraph(us, u0s, fs, hs, tols; grad_us, grad_u0s, grad_fs, forward, compute_f64)

hipGraph of the batched launch over fixed device buffers; `launch!(g)`
replays it, `close!(g)` frees it.  The buffers must outlive the graph.
"""
mutable struct EikonalGraph
    ptr::Ptr{Cvoid}
end

function EikonalGraph(us, u0s, fs, hs, tols; grad_us=nothing, grad_u0s=nothing, grad_fs=nothing,
                      forward::Bool=true, compute_f64::Bool=false, stream::Ptr{Cvoid}=C_NULL)
    tab = event_table(u0s, fs, us, hs, tols, grad_us, grad_u0s, grad_fs)
    flags = (forward ? ADT_FORWARD : Int32(0)) | (grad_us === nothing ? Int32(0) : ADT_ADJOINT) |
            (compute_f64 ? ADT_F32_COMPUTE_F64 : Int32(0))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:adt_graph_create, libadt), Cint,
               (Ptr{AdtEvent3d}, Int32, Int32, Int32, Ref{Ptr{Cvoid}}, Ptr{Cvoid}),
               tab, length(tab), flags, Int32(eltype(fs[1]) == Float32), out, stream)
    check(rc, "adt_graph_create")
    g = EikonalGraph(out[])
    finalizer(close!, g)
    return g
end

launch!(g::EikonalGraph; stream::Ptr{Cvoid}=C_NULL) =
    check(ccall((:adt_graph_launch, libadt), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), g.ptr, stream),
          "adt_graph_launch")

function close!(g::EikonalGraph)
    if g.ptr != C_NULL
        ccall((:adt_graph_destroy, libadt), Cint, (Ptr{Cvoid},), g.ptr)
        g.ptr = C_NULL
    end
    return nothing
end

# ------------------------------------------- objective, device arrays ---

# Mirrors `adt_source3d` / `adt_misfit3d` in include/adtomo_hip.h.
struct AdtSource3d
    x::Float64
    y::Float64
    z::Float64
    scale::Float64
    u0::Ptr{Cvoid}
end

struct AdtMisfit3d
    m::Int64
    n::Int64
    l::Int64
    u::Ptr{Cvoid}
    tobs::Ptr{Cvoid}
    weight::Ptr{Cvoid}
    tcal::Ptr{Cvoid}
    grad_u::Ptr{Cvoid}
    loss::Ptr{Cvoid}
    huber_delta::Float64
end

"""
    source_u0!(u0s, stations, vel0, h; scale=1.0, stream=C_NULL)

Initial travel times of each station (x, y, z), 1-based reals, on the
device (scripts/inversion.jl:48-60; `scale` = vp/vs for S).
"""
function source_u0!(u0s, stations, vel0, h; scale=1.0, stream::Ptr{Cvoid}=C_NULL)
    tab = [AdtSource3d(Float64(s[1]), Float64(s[2]), Float64(s[3]), Float64(scale), devptr(u0s[e]))
           for (e, s) in enumerate(stations)]
    m, n, l = griddims(vel0)
    fn = Symbol("adt_source_u0_3d_batched", sfx(vel0))
    T = eltype(vel0)
    rc = ccall(Libdl.dlsym(Libdl.dlopen(libadt), fn), Cint,
               (Ptr{AdtSource3d}, Int32, Ptr{Cvoid}, Int64, Int64, Int64, T, Ptr{Cvoid}),
               tab, length(tab), devptr(vel0), m, n, l, T(h), stream)
    check(rc, "source_u0")
    return u0s
end

"""
    misfit!(loss, us, rx, ry, rz, tobs, weights; grad_us=nothing, tcals=nothing,
            huber_delta=0.0, stream=C_NULL)

Receiver gather + weighted misfit per station (scripts/inversion.jl:64-105):
`loss` is a device vector (one entry per station), `rx, ry, rz` host event
coordinates; `grad_us` receive dL/du for the adjoint.
"""
function misfit!(loss, us, rx, ry, rz, tobs, weights; grad_us=nothing, tcals=nothing,
                 huber_delta=0.0, stream::Ptr{Cvoid}=C_NULL)
    T = eltype(us[1])
    lp = devptr(loss)
    tab = [AdtMisfit3d(griddims(us[e])..., devptr(us[e]), devptr(tobs[e]), devptr(weights[e]),
                       tcals === nothing ? C_NULL : devptr(tcals[e]),
                       grad_us === nothing ? C_NULL : devptr(grad_us[e]),
                       lp + (e - 1) * sizeof(T), Float64(huber_delta)) for e in eachindex(us)]
    fn = Symbol("adt_misfit3d_batched", sfx(us[1]))
    rc = ccall(Libdl.dlsym(Libdl.dlopen(libadt), fn), Cint,
               (Ptr{AdtMisfit3d}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64,
                Ptr{Cvoid}),
               tab, length(tab), Vector{Float64}(rx), Vector{Float64}(ry), Vector{Float64}(rz),
               length(rx), stream)
    check(rc, "misfit")
    return loss
end

"""
    eikonal3d_misfit_batched!(us, loss, u0s, fs, hs, tols, grad_us, rx, ry, rz, tobs, weights;
                              grad_u0s, grad_fs, huber_delta=0.0, compute_f64=false,
                              stream=C_NULL)

One inversion step in one launch (scripts/inversion.jl:47-105): for each
event its forward, the receiver misfit of its travel times (`loss[e]`, its
gradient written to `grad_us[e]` as the adjoint source) and the adjoint into
`grad_fs` / `grad_u0s`.  Device arrays as for `eikonal3d_batched!`; `rx, ry,
rz` host event coordinates (1-based); returns the per-event `AdtInfo`s.
"""
function eikonal3d_misfit_batched!(us, loss, u0s, fs, hs, tols, grad_us, rx, ry, rz, tobs,
                                   weights; grad_u0s=nothing, grad_fs, huber_delta=0.0,
                                   compute_f64::Bool=false, stream::Ptr{Cvoid}=C_NULL)
    tab = event_table(u0s, fs, us, hs, tols, grad_us, grad_u0s, grad_fs)
    T = eltype(fs[1])
    lp = devptr(loss)
    items = [AdtMisfit3d(griddims(us[e])..., devptr(us[e]), devptr(tobs[e]), devptr(weights[e]),
                         C_NULL, devptr(grad_us[e]), lp + (e - 1) * sizeof(T),
                         Float64(huber_delta)) for e in eachindex(us)]
    flags = ADT_FORWARD | ADT_ADJOINT | (compute_f64 ? ADT_F32_COMPUTE_F64 : Int32(0))
    infos = [AdtInfo() for _ in tab]
    fn = Symbol("adt_eikonal3d_misfit_batched", sfx(fs[1]))
    rc = ccall(Libdl.dlsym(Libdl.dlopen(libadt), fn), Cint,
               (Ptr{AdtEvent3d}, Ptr{AdtMisfit3d}, Int32, Ptr{Float64}, Ptr{Float64},
                Ptr{Float64}, Int64, Int32, Ptr{AdtInfo}, Ptr{Cvoid}),
               tab, items, length(tab), Vector{Float64}(rx), Vector{Float64}(ry),
               Vector{Float64}(rz), length(rx), flags, infos, stream)
    check(rc, "eikonal3d_misfit_batched")
    return infos
end

# ------------------------------------- model, regulariser, sum, tuning ---

objfn(base, A) = Libdl.dlsym(Libdl.dlopen(libadt), Symbol(base, sfx(A)))

"""
    model!(fvar, f, x, vel0; stream=C_NULL)

fvar = 2 sigmoid(x) - 1 + vel0 and f = 1 ./ fvar (scripts/inversion.jl:42-43,61)
on device arrays of one length (`f` may be `nothing`).
"""
function model!(fvar, f, x, vel0; stream::Ptr{Cvoid}=C_NULL)
    rc = ccall(objfn("adt_model3d", x), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
               devptr(x), devptr(vel0), length(x), devptr(fvar), devptr(f), stream)
    check(rc, "model")
    return fvar
end

"""
    model_grad!(grad_x, x, fvar; grad_f=nothing, grad_fvar=nothing, stream=C_NULL)

Chain rule of `model!`: grad_x = (grad_fvar - grad_f ./ fvar.^2) .* 2 sigmoid'(x).
"""
function model_grad!(grad_x, x, fvar; grad_f=nothing, grad_fvar=nothing,
                     stream::Ptr{Cvoid}=C_NULL)
    rc = ccall(objfn("adt_model3d_grad", x), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}),
               devptr(x), devptr(fvar), devptr(grad_f), devptr(grad_fvar), length(x),
               devptr(grad_x), stream)
    check(rc, "model_grad")
    return grad_x
end

"""
    slowness!(f_s, fvar, scale; stream=C_NULL)

S slowness f_s = scale ./ fvar (scripts/inversion_joint.jl:79, scale = vp/vs).
"""
function slowness!(f_s, fvar, scale; stream::Ptr{Cvoid}=C_NULL)
    T = eltype(fvar)
    rc = ccall(objfn("adt_slowness3d", fvar), Cint, (Ptr{Cvoid}, Int64, T, Ptr{Cvoid}, Ptr{Cvoid}),
               devptr(fvar), length(fvar), T(scale), devptr(f_s), stream)
    check(rc, "slowness")
    return f_s
end

"""
    slowness_grad!(grad_fvar, fvar, scale, grad_fs; grad_scale=nothing, stream=C_NULL)

Adjoint of `slowness!`: grad_fvar .+= -scale .* grad_fs ./ fvar.^2 (in place);
`grad_scale` (a one-element device array) receives sum(grad_fs ./ fvar).
"""
function slowness_grad!(grad_fvar, fvar, scale, grad_fs; grad_scale=nothing,
                        stream::Ptr{Cvoid}=C_NULL)
    T = eltype(fvar)
    rc = ccall(objfn("adt_slowness3d_grad", fvar), Cint,
               (Ptr{Cvoid}, Int64, T, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
               devptr(fvar), length(fvar), T(scale), devptr(grad_fs), devptr(grad_fvar),
               devptr(grad_scale), stream)
    check(rc, "slowness_grad")
    return grad_fvar
end

"""
    smooth_l1!(grad_fvar, value, fvar, sh, sv, lambda; stream=C_NULL)

Box-smoothing L1 regulariser (scripts/inversion.jl:107-121) of an (l, n, m)
device array: `value` (one-element device array) = lambda * sum |fvar - box(fvar)|,
`grad_fvar` its gradient (may be `nothing`).
"""
function smooth_l1!(grad_fvar, value, fvar, sh::Integer, sv::Integer, lambda;
                    stream::Ptr{Cvoid}=C_NULL)
    T = eltype(fvar)
    m, n, l = griddims(fvar)
    rc = ccall(objfn("adt_smooth3d_l1", fvar), Cint,
               (Ptr{Cvoid}, Int64, Int64, Int64, Int32, Int32, T, Ptr{Cvoid}, Ptr{Cvoid},
                Ptr{Cvoid}),
               devptr(fvar), m, n, l, Int32(sh), Int32(sv), T(lambda), devptr(grad_fvar),
               devptr(value), stream)
    check(rc, "smooth_l1")
    return value
end

"""
    sum_arrays!(out, arrays; stream=C_NULL)

out = sum of the device arrays in order (the per-station dL/df sum).
"""
function sum_arrays!(out, arrays; stream::Ptr{Cvoid}=C_NULL)
    ptrs = Ptr{Cvoid}[devptr(a) for a in arrays]
    rc = ccall(objfn("adt_sum_arrays", out), Cint,
               (Ptr{Ptr{Cvoid}}, Int32, Int64, Ptr{Cvoid}, Ptr{Cvoid}),
               ptrs, length(ptrs), length(out), devptr(out), stream)
    check(rc, "sum_arrays")
    return out
end

# Mirrors `adt_tune` in include/adtomo_hip.h.
struct AdtTune
    diag_max_l::Int32
    adj_wg::Int32
    order::Int32
    fuse_max_n::Int32
    fwd_ms::Float32
    adj_ms::Float32
    adj_algo::Int32
    fwd_segs::Int32
end

# Mirrors `adt_tune_class` in include/adtomo_hip.h.
struct AdtTuneClass
    l_blocks::Int32
    bands::Int32
    log2n::Int32
    tune::AdtTune
end

dtype_code(::Type{Float32}) = Int32(1)
dtype_code(::Type{Float64}) = Int32(0)

"""
    get_tune(T=Float64)

Launch setting of the 3D kernels for element type `T`.
"""
function get_tune(::Type{T}=Float64) where {T}
    t = Ref(AdtTune(0, 0, 0, 0, 0f0, 0f0, 0, 0))
    check(ccall((:adt_get_tune, libadt), Cint, (Int32, Ref{AdtTune}), dtype_code(T), t),
          "get_tune")
    return t[]
end

"""
    set_tune!(t::AdtTune, T=Float64)

Stores a launch setting (validated by the library) for element type `T`.
"""
function set_tune!(t::AdtTune, ::Type{T}=Float64) where {T}
    check(ccall((:adt_set_tune, libadt), Cint, (Int32, Ref{AdtTune}), dtype_code(T), Ref(t)),
          "set_tune")
    return t
end

"""
    set_tune_class!(dims, t::Union{AdtTune,Nothing}, T=Float64)

Gives the box-shape class of `dims` = (m, n, l) the per-event fields of `t`
(diag_max_l, fwd_segs, adj_algo), or removes its setting (`nothing`).
"""
function set_tune_class!(dims, t::Union{AdtTune,Nothing}, ::Type{T}=Float64) where {T}
    m, n, l = Int64.(dims)
    if t === nothing
        rc = ccall((:adt_tune_class_set, libadt), Cint, (Int32, Int64, Int64, Int64, Ptr{AdtTune}),
                   dtype_code(T), m, n, l, C_NULL)
    else
        rc = ccall((:adt_tune_class_set, libadt), Cint,
                   (Int32, Int64, Int64, Int64, Ref{AdtTune}), dtype_code(T), m, n, l, Ref(t))
    end
    check(rc, "tune_class_set")
    return t
end

"""
    tune_classes(T=Float64)

The class settings of element type `T` (`AdtTuneClass` values).
"""
function tune_classes(::Type{T}=Float64) where {T}
    n = ccall((:adt_tune_class_count, libadt), Cint, (Int32,), dtype_code(T))
    check(min(n, 0), "tune_class_count")
    out = AdtTuneClass[]
    for i in 0:n-1
        c = Ref(AdtTuneClass(0, 0, 0, AdtTune(0, 0, 0, 0, 0f0, 0f0, 0, 0)))
        check(ccall((:adt_tune_class_get, libadt), Cint, (Int32, Int32, Ref{AdtTuneClass}),
                    dtype_code(T), Int32(i), c), "tune_class_get")
        push!(out, c[])
    end
    return out
end

"""
    save_tune(path), load_tune(path)

Writes every launch setting (global and per class) to a JSON file, or reads
one back.
"""
save_tune(path::AbstractString) =
    check(ccall((:adt_tune_save, libadt), Cint, (Cstring,), path), "tune_save")
load_tune(path::AbstractString) =
    check(ccall((:adt_tune_load, libadt), Cint, (Cstring,), path), "tune_load")

"""
    autotune!(us, u0s, fs, hs, tols; grad_us, grad_u0s, grad_fs, reps=3,
              compute_f64=false, stream=C_NULL)

Times the launch settings on these events (device arrays, as for
`eikonal3d_batched!`): per box-shape class the forward kernel, its segment
count and the adjoint traversal, over all events the launch-wide fields;
keeps the fastest for their element type and returns the global setting;
the events' outputs hold a valid solve afterwards.
"""
function autotune!(us, u0s, fs, hs, tols; grad_us=nothing, grad_u0s=nothing, grad_fs=nothing,
                   reps::Integer=3, compute_f64::Bool=false, stream::Ptr{Cvoid}=C_NULL)
    tab = event_table(u0s, fs, us, hs, tols, grad_us, grad_u0s, grad_fs)
    flags = ADT_FORWARD | (grad_us === nothing ? Int32(0) : ADT_ADJOINT) |
            (compute_f64 ? ADT_F32_COMPUTE_F64 : Int32(0))
    best = Ref(AdtTune(0, 0, 0, 0, 0f0, 0f0, 0, 0))
    rc = ccall((:adt_autotune, libadt), Cint,
               (Ptr{AdtEvent3d}, Int32, Int32, Int32, Int32, Ref{AdtTune}, Ptr{Cvoid}),
               tab, length(tab), flags, dtype_code(eltype(fs[1])), Int32(reps), best, stream)
    check(rc, "autotune")
    return best[]
end

end # module
