#=
SinogramsMI355X.jl
Methods that put Sinograms.jl's hot path on libsgx (the MI355X HIP library,
C ABI in include/sgx.h).  Loaded after Sinograms, these Float32 methods take
precedence for plain host arrays; every other call stays on the reference's
own Julia code.

Replaced functions (reference file:line):
  fdk(plan, proj)            src/fbp3/fdk.jl:65-81     one sgx_plan_execute
  fbp(plan, sino)            src/fbp2/fbp.jl:24-43     one sgx_plan_execute
  fbp(plan, stack)           src/fbp2/fbp.jl:46-68     one sgx_plan_execute per slice
  fbp_back_par!              src/fbp2/back-par.jl:206-253
  fbp_back_fan!              src/fbp2/back-fan.jl:215-254
  cbct_back_fan!             src/fbp3/cbct-back.jl:95-144
  fbp_sino_filter            src/fbp/filter.jl:122-147

fdk/fbp take the reference's own plan (plan_fbp, src/fbp3/plan3.jl:60-72,
src/fbp2/plan2.jl:106-138) and build the device plan from its geometry, image
grid and filter once (sgx_plan_create); later calls with the same plan and
precision reuse it, so a reconstruction is one ccall with one transfer each
way.  The keyword `precision` (:fp32 default, or :bf16 / :f16s for cone-beam
plans) selects the plan's precision form; it has no counterpart in the
reference.

Threads.  A libsgx context must not be used by two threads at once.  The
free functions take a context from a pool (one per concurrent caller); a
device plan owns its context and a lock held for each call.
=#
module SinogramsMI355X

using Sinograms
import Sinograms: fbp_back_par!, fbp_back_fan!, cbct_back_fan!, fbp_sino_filter, fdk, fbp
using Sinograms: SinoGeom, SinoPar, SinoFan, SinoFanArc, SinoMoj, CtFan, CtFanArc,
    FBPNormalPlan, FDKplan, _ar, _dso, _gamma_max, _angle_weights, parker_weight
using ImageGeoms: ImageGeom

export gpu_plan, release!

const lib = joinpath(@__DIR__, "..", "libsgx.so")
const SGX_HOST = UInt32(1)
const SGX_ERR_SIZE = Cint(-2)
const SGX_PREC = Dict(:fp32 => Cint(0), :bf16 => Cint(1), :f16s => Cint(2))
const SGX_SINO_PAR, SGX_SINO_FAN_ARC, SGX_SINO_FAN_FLAT = Int32(0), Int32(1), Int32(2)
const SGX_CT_FAN_ARC, SGX_CT_FAN_FLAT, SGX_SINO_MOJ = Int32(3), Int32(4), Int32(5)

# ---- contexts ---------------------------------------------------------------

function new_ctx()
    r = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:sgx_ctx_create, lib), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, r)
    rc == 0 || error("sgx_ctx_create failed ($rc)")
    return r[]
end

destroy_ctx(c::Ptr{Cvoid}) = ccall((:sgx_ctx_destroy, lib), Cvoid, (Ptr{Cvoid},), c)

const POOL = Ptr{Cvoid}[]             # idle contexts of the free functions
const POOL_ALL = Ptr{Cvoid}[]
const POOL_LOCK = ReentrantLock()

# run f(ctx) with a context no other thread uses meanwhile
function with_ctx(f)
    c = lock(POOL_LOCK) do
        if isempty(POOL)
            c = new_ctx()
            push!(POOL_ALL, c)
            c
        else
            pop!(POOL)
        end
    end
    try
        return f(c)
    finally
        lock(() -> push!(POOL, c), POOL_LOCK)
    end
end

function check(c::Ptr{Cvoid}, rc::Cint)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:sgx_last_error, lib), Cstring, (Ptr{Cvoid},), c))
    rc == SGX_ERR_SIZE && throw("size mismatch")
    error(msg)
end

_f64(v) = collect(Float64, v)

# ---- device plans -----------------------------------------------------------

# the C layout of sgx_geom (include/sgx.h)
struct SgxGeom
    kind::Int32
    parker::Int32
    ns::Int64
    nt::Int64
    na::Int64
    ds::Float64
    dt::Float64
    offset_s::Float64
    offset_t::Float64
    dsd::Float64
    dso::Float64
    source_offset::Float64
    gamma_max::Float64
    angle_weight::Float64
end

_kind(rg::SinoPar) = SGX_SINO_PAR
_kind(rg::SinoMoj) = SGX_SINO_MOJ
_kind(rg::SinoFan) = rg isa SinoFanArc ? SGX_SINO_FAN_ARC : SGX_SINO_FAN_FLAT
_kind(rg::CtFan) = rg isa CtFanArc ? SGX_CT_FAN_ARC : SGX_CT_FAN_FLAT

# parker_weight's own test for a full orbit (src/fbp/parker.jl:112, 148)
_short(rg::Union{SinoFan,CtFan}) = (rg.orbit ÷ 360) * 360 != rg.orbit
_short(rg) = false

function sgx_geom(rg::SinoGeom)
    fan = rg isa SinoFan
    return SgxGeom(_kind(rg), Int32(_short(rg)), rg.nb, 1, rg.na, rg.d, 0, rg.offset, 0,
        fan ? rg.dsd : 0, fan ? _dso(rg) : 0, fan ? rg.source_offset : 0,
        fan ? _gamma_max(rg) : 0, _angle_weights(_ar(rg)))
end

sgx_geom(rg::CtFan) = SgxGeom(_kind(rg), Int32(_short(rg)), rg.ns, rg.nt, rg.na, rg.ds,
    rg.dt, rg.offset_s, rg.offset_t, rg.dsd, _dso(rg), rg.source_offset, _gamma_max(rg),
    _angle_weights(_ar(rg)))

"""
    MI355XPlan
A device plan of libsgx built from a reference plan (`FDKplan` or
`FBPNormalPlan`): geometry tables, filter, weights and the filtered-data
buffer live on the GPU.  It owns a context and a lock.
"""
mutable struct MI355XPlan{P}
    plan::P
    precision::Symbol
    ctx::Ptr{Cvoid}
    h::Ptr{Cvoid}
    lock::ReentrantLock
end

# (the finalizer: nothing else holds the plan then, so no lock)
function _free!(p::MI355XPlan)
    p.h != C_NULL && ccall((:sgx_plan_destroy, lib), Cvoid, (Ptr{Cvoid},), p.h)
    p.ctx != C_NULL && destroy_ctx(p.ctx)
    p.h = C_NULL
    p.ctx = C_NULL
    return nothing
end

"""
    release!(p::MI355XPlan)
Free the device plan and its context now (the finalizer does it otherwise).
"""
release!(p::MI355XPlan) = lock(() -> _free!(p), p.lock)

"""
    gpu_plan(plan; precision = :fp32)
The device plan of a reference plan (`plan_fbp(rg, ig; ...)`), with its
geometry, image grid (axes and mask) and filter `plan.filter`.  The weights
are those of the reference's `plan.view_weight`, formed on the device from
the geometry (Parker, fan or cone-beam, angle weights).
"""
function gpu_plan(plan::Union{FDKplan,FBPNormalPlan}; precision::Symbol = :fp32)
    haskey(SGX_PREC, precision) || throw(ArgumentError("precision $precision"))
    rg, ig = plan.rg, plan.ig
    G = Ref(sgx_geom(rg))
    betas = _f64(_ar(rg))
    ax = map(_f64, axes(ig))
    length(ax) == (rg isa CtFan ? 3 : 2) || throw("size mismatch")
    zc = length(ax) == 3 ? ax[3] : nothing
    mask = all(ig.mask) ? nothing : UInt8.(ig.mask)
    Hk = collect(Float32, plan.filter)
    c = new_ctx()
    h = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve betas ax zc mask Hk begin
        rc = ccall((:sgx_plan_create, lib), Cint,
            (Ptr{Cvoid}, Ref{SgxGeom}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64},
             Int64, Ptr{Float64}, Int64, Ptr{UInt8}, Ptr{Float32}, Int64, Cint,
             Ref{Ptr{Cvoid}}),
            c, G, betas, ax[1], length(ax[1]), ax[2], length(ax[2]),
            zc === nothing ? C_NULL : pointer(zc), zc === nothing ? 1 : length(zc),
            mask === nothing ? C_NULL : pointer(mask), Hk, length(Hk), SGX_PREC[precision], h)
        if rc != 0
            try
                check(c, rc)
            finally
                destroy_ctx(c)
            end
        end
    end
    p = MI355XPlan(plan, precision, c, h[], ReentrantLock())
    finalizer(_free!, p)
    if rg isa SinoPar     # parker_weight_par (parker.jl:26-58) for orbits not a multiple of 180
        pw = parker_weight(rg)
        if length(pw) > 1
            w = collect(Float32, vec(pw))
            GC.@preserve w check(c, ccall((:sgx_plan_set_view_weight, lib), Cint,
                (Ptr{Cvoid}, Ptr{Float32}), p.h, w))
        end
    end
    return p
end

# device plans of reference plans, reused by later calls (8 at most)
const PLANS = IdDict{Any,MI355XPlan}()
const PLAN_ORDER = Any[]
const PLANS_LOCK = ReentrantLock()

function cached_plan(plan, precision::Symbol)
    key = (plan, precision)
    lock(PLANS_LOCK) do
        p = get(PLANS, key, nothing)
        if p === nothing
            p = gpu_plan(plan; precision)
            PLANS[key] = p
            push!(PLAN_ORDER, key)
            if length(PLAN_ORDER) > 8
                old = popfirst!(PLAN_ORDER)
                release!(pop!(PLANS, old))
            end
        end
        return p
    end
end

# one sgx_plan_execute on host arrays: in, out column-major Float32
function execute!(p::MI355XPlan, out::Array{Float32}, in::Array{Float32})
    lock(p.lock) do
        p.h == C_NULL && error("released plan")
        GC.@preserve in out check(p.ctx, ccall((:sgx_plan_execute, lib), Cint,
            (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, UInt32), p.h, in, out, SGX_HOST))
    end
    return out
end

function fdk(plan::FDKplan, proj::Array{Float32,3}; precision::Symbol = :fp32)
    rg = plan.rg
    size(proj) == (rg.ns, rg.nt, rg.na) ||
        error("size mismatch $(size(proj)) $((rg.ns, rg.nt, rg.na))")
    p = cached_plan(plan, precision)
    return execute!(p, Array{Float32}(undef, size(plan.ig)...), proj)
end

fdk(p::MI355XPlan{<:FDKplan}, proj::Array{Float32,3}) =
    execute!(p, Array{Float32}(undef, size(p.plan.ig)...), proj)

# one method per plan type of the reference's fbp(plan, sino) (fbp.jl:24-43),
# so that neither is ambiguous with these
function _fbp(plan::FBPNormalPlan, sino::Matrix{Float32}, precision::Symbol)
    size(sino) == (plan.rg.nb, plan.rg.na) || error("bad sino size")
    p = cached_plan(plan, precision)
    return execute!(p, Array{Float32}(undef, size(plan.ig)...), sino)
end

fbp(plan::FBPNormalPlan{<:SinoPar}, sino::Matrix{Float32}; precision::Symbol = :fp32) =
    _fbp(plan, sino, precision)
fbp(plan::FBPNormalPlan{<:SinoFan}, sino::Matrix{Float32}; precision::Symbol = :fp32) =
    _fbp(plan, sino, precision)

fbp(p::MI355XPlan{<:FBPNormalPlan}, sino::Matrix{Float32}) =
    execute!(p, Array{Float32}(undef, size(p.plan.ig)...), sino)

# a stack of sinograms (nb, na, ...) (fbp.jl:46-68), and SinoMoj plans: the
# reference's serial slice loop, one execute per slice
function fbp(plan::FBPNormalPlan{<:SinoGeom}, aa::Array{Float32,N};
        precision::Symbol = :fp32) where {N}
    N == 2 && return _fbp(plan, aa, precision)
    nb, na = size(aa, 1), size(aa, 2)
    (nb, na) == (plan.rg.nb, plan.rg.na) || error("bad sino size")
    p = cached_plan(plan, precision)
    dims = size(plan.ig)
    L = prod(size(aa)[3:end])
    out = Array{Float32}(undef, dims..., size(aa)[3:end]...)
    s3 = reshape(aa, nb, na, L)
    o3 = reshape(out, prod(dims), L)
    slice = Array{Float32}(undef, dims...)
    for iz in 1:L
        execute!(p, slice, s3[:, :, iz])
        o3[:, iz] .= vec(slice)
    end
    return out
end

# ---- the reference's inner kernels (one ccall each) --------------------------

function fbp_back_par!(image::Matrix{Float32}, sino::Matrix{Float32},
        sinϕ::AbstractVector{<:Real}, cosϕ::AbstractVector{<:Real},
        ds::Real, offset::Real, xc::AbstractVector{<:Real}, yc::AbstractVector{<:Real},
        mask::AbstractMatrix{Bool}; ia_skip::Int = 1)
    length.((xc, yc)) == size(image) == size(mask) || throw("size mismatch")
    nb, na = size(sino)
    s, c, x, y, m = _f64(sinϕ), _f64(cosϕ), _f64(xc), _f64(yc), UInt8.(mask)
    with_ctx() do ctx
        GC.@preserve image sino s c x y m check(ctx, ccall((:sgx_back_par, lib), Cint,
            (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64},
             Float64, Float64, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{UInt8},
             Ptr{Float32}, UInt32),
            ctx, sino, nb, na, ia_skip, s, c, ds, offset, x, length(x), y, length(y), m,
            image, SGX_HOST))
    end
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
    with_ctx() do ctx
        GC.@preserve image sino s c x y m check(ctx, ccall((:sgx_back_fan, lib), Cint,
            (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Int64, Ptr{Float64}, Ptr{Float64},
             Float64, Float64, Float64, Cint, Float64, Float64, Ptr{Float64}, Int64,
             Ptr{Float64}, Int64, Ptr{UInt8}, Ptr{Float32}, UInt32),
            ctx, sino, nb, na, ia_skip, s, c, dsd, dso, source_offset, is_arc, ds, offset,
            x, length(x), y, length(y), m, image, SGX_HOST))
    end
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
    with_ctx() do ctx
        GC.@preserve image proj b x y z m check(ctx, ccall((:sgx_cbct_back, lib), Cint,
            (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Int64, Int64, Ptr{Float64},
             Float64, Float64, Float64, Float64, Float64, Float64, Cint,
             Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Int64,
             Ptr{UInt8}, Ptr{Float32}, UInt32),
            ctx, proj, ns, nt, na, ia_skip, b, dsd, dso, ds, dt, offset_s, offset_t,
            is_arc, x, length(x), y, length(y), z, length(z), m, image, SGX_HOST))
    end
    return image
end

# the context keeps the device form of recent filters, so repeated calls
# with one Hk do no host transform
function fbp_sino_filter(sino::Array{Float32}, filter::AbstractVector{Float32};
        extra::Int = 0)
    npad = length(filter)
    nb = size(sino, 1)
    nb + extra > npad && error("nb=$nb + extra=$extra > npad=$npad")
    ncols = div(length(sino), nb)
    out = Array{Float32}(undef, nb + extra, size(sino)[2:end]...)
    Hk = collect(filter)
    with_ctx() do ctx
        GC.@preserve sino Hk out check(ctx, ccall((:sgx_sino_filter, lib), Cint,
            (Ptr{Cvoid}, Ptr{Float32}, Int64, Int64, Ptr{Float32}, Int64, Int64, Cint,
             Ptr{Float32}, UInt32),
            ctx, sino, nb, ncols, Hk, npad, extra, SGX_PREC[:fp32], out, SGX_HOST))
    end
    return out
end

function __init__()
    atexit() do
        lock(POOL_LOCK) do
            foreach(destroy_ctx, POOL_ALL)
            empty!(POOL_ALL)
            empty!(POOL)
        end
    end
end

end # module
