#=
SinogramsMI355X.jl
Methods that put Sinograms.jl's hot loops on libsgx (the MI355X HIP library,
C ABI in include/sgx.h).  Loaded after Sinograms, these Float32 methods take
precedence for plain host arrays; every other call stays on the reference's
own Julia code.

Replaced functions (reference file:line):
  fbp_back_par!    src/fbp2/back-par.jl:206-253
  fbp_back_fan!    src/fbp2/back-fan.jl:215-254
  cbct_back_fan!   src/fbp3/cbct-back.jl:95-144
  fbp_sino_filter  src/fbp/filter.jl:122-147
=#
module SinogramsMI355X

using Sinograms
import Sinograms: fbp_back_par!, fbp_back_fan!, cbct_back_fan!, fbp_sino_filter

const lib = joinpath(@__DIR__, "..", "libsgx.so")
const SGX_HOST = UInt32(1)
const SGX_PREC_FP32 = Cint(0)

const CTX = Ref{Ptr{Cvoid}}(C_NULL)

function ctx()
    if CTX[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:sgx_ctx_create, lib), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, r)
        rc == 0 || error("sgx_ctx_create failed ($rc)")
        CTX[] = r[]
        atexit(() -> ccall((:sgx_ctx_destroy, lib), Cvoid, (Ptr{Cvoid},), CTX[]))
    end
    return CTX[]
end

function check(rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:sgx_last_error, lib), Cstring, (Ptr{Cvoid},), ctx()))
    rc == -2 && throw("size mismatch")
    error(msg)
end

_f64(v) = collect(Float64, v)

function fbp_back_par!(image::Matrix{Float32}, sino::Matrix{Float32},
        sinϕ::AbstractVector{<:Real}, cosϕ::AbstractVector{<:Real},
        ds::Real, offset::Real, xc::AbstractVector{<:Real}, yc::AbstractVector{<:Real},
        mask::AbstractMatrix{Bool}; ia_skip::Int = 1)
    length.((xc, yc)) == size(image) == size(mask) || throw("size mismatch")
    nb, na = size(sino)
    s, c, x, y, m = _f64(sinϕ), _f64(cosϕ), _f64(xc), _f64(yc), UInt8.(mask)
    GC.@preserve image sino s c x y m check(ccall((:sgx_back_par, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64},
         Float64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{UInt8},
         Ptr{Float32}, UInt32),
        ctx(), sino, nb, na, ia_skip, s, c, ds, offset, x, length(x), y, length(y), m,
        image, SGX_HOST))
    return image
end

function fbp_back_fan!(image::Matrix{Float32}, sino::Matrix{Float32},
        sinβ::AbstractVector{<:Real}, cosβ::AbstractVector{<:Real},
        dsd::Real, dso::Real, source_offset::Real, is_arc::Bool,
        ds::Real, offset::Real, xc::AbstractVector{<:Real}, yc::AbstractVector{<:Real},
        mask::AbstractMatrix{Bool}; ia_skip::Int = 1)
    length.((xc, yc)) == size(image) == size(mask) || throw("size mismatch")
    nb, na = size(sino)
    s, c, x, y, m = _f64(sinβ), _f64(cosβ), _f64(xc), _f64(yc), UInt8.(mask)
    GC.@preserve image sino s c x y m check(ccall((:sgx_back_fan, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64},
         Float64, Float64, Float64, Cint, Float64, Float64, Ptr{Float64}, Int64,
         Ptr{Float64}, Int64, Ptr{UInt8}, Ptr{Float32}, UInt32),
        ctx(), sino, nb, na, ia_skip, s, c, dsd, dso, source_offset, is_arc, ds, offset,
        x, length(x), y, length(y), m, image, SGX_HOST))
    return image
end

function cbct_back_fan!(image::Array{Float32,3}, proj::Array{Float32,3},
        betas::AbstractVector{<:Real}, dsd::Real, dso::Real, ds::Real, dt::Real,
        offset_s::Real, offset_t::Real, is_arc::Bool,
        xc::AbstractVector{<:Real}, yc::AbstractVector{<:Real}, zc::AbstractVector{<:Real},
        mask::AbstractArray{Bool,3}; ia_skip::Int = 1)
    length.((xc, yc, zc)) == size(image) == size(mask) || throw("size mismatch")
    (ns, nt, na) = size(proj)
    b, x, y, z, m = _f64(betas), _f64(xc), _f64(yc), _f64(zc), UInt8.(mask)
    GC.@preserve image proj b x y z m check(ccall((:sgx_cbct_back, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Int64, Int64, Ptr{Float64},
         Float64, Float64, Float64, Float64, Float64, Float64, Cint,
         Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
         Ptr{UInt8}, Ptr{Float32}, UInt32),
        ctx(), proj, ns, nt, na, ia_skip, b, dsd, dso, ds, dt, offset_s, offset_t,
        is_arc, x, length(x), y, length(y), z, length(z), m, image, SGX_HOST))
    return image
end

function fbp_sino_filter(sino::Array{Float32}, filter::AbstractVector{Float32};
        extra::Int = 0)
    npad = length(filter)
    nb = size(sino, 1)
    nb + extra > npad && error("nb=$nb + extra=$extra > npad=$npad")
    ncols = div(length(sino), nb)
    out = Array{Float32}(undef, nb + extra, size(sino)[2:end]...)
    GC.@preserve sino filter out check(ccall((:sgx_sino_filter, lib), Cint,
        (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Ptr{Float32}, Int64, Int64, Cint,
         Ptr{Float32}, UInt32),
        ctx(), sino, nb, ncols, filter, npad, extra, SGX_PREC_FP32, out, SGX_HOST))
    return out
end

end # module
