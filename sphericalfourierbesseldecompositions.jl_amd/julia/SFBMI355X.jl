# SFBMI355X -- drop-in cat2amln for SphericalFourierBesselDecompositions.jl on an
# MI355X, through the C ABI of libsfb_mi355x.so (include/sfb.h).
#
#     import SphericalFourierBesselDecompositions as SFB
#     import SFBMI355X
#     anlm = SFBMI355X.cat2amln(rθϕ, amodes, nbar, win_rhat_ln, weights)   # was SFB.cat2amln
#
# Same positional signature and return value as SFB.cat2amln
# (src/cat2anlm.jl:257-310): rθϕ is 3 x N, win_rhat_ln is the dense
# (npix, lmax+1, nmax) array or the SeparableArray(mask, w_ln) that SFB.win_rhat_ln
# returns, weight defaults to ones(N).  The basis-function splines are passed
# straight from amodes.basisfunctions.gnl[n, l+1] (x, y, ypp), so the knots are
# the caller's own.  A new function, not a method of SFB.cat2amln (no type piracy).
# Only argument marshalling lives here; every computation runs in the library.
# cat2amln_vjp and the ChainRulesCore.rrule give the adjoint (no counterpart in the
# reference, which has no derivative of cat2amln).
module SFBMI355X

import ChainRulesCore
import SphericalFourierBesselDecompositions as SFB

export cat2amln, cat2nlm, cat2amln_vjp, cat2amln_batch, win_rhat_ln, field2anlm, anlm2field, amln2clnn, pixwin

const libsfb = get(ENV, "SFB_MI355X_LIB",
                   joinpath(@__DIR__, "..", "sfb_amd", "libsfb_mi355x.so"))

const SFB_OK = 0
const SFB_ERR_NONFINITE = 3

function check(rc)
    rc == SFB_OK && return nothing
    msg = unsafe_string(ccall((:sfb_last_error, libsfb), Cstring, ()))
    rc == SFB_ERR_NONFINITE && throw(AssertionError(msg))
    error("libsfb_mi355x: " * msg)
end

mutable struct Plan
    ptr::Ptr{Cvoid}
    nlmsize::Int
    npix::Int
    function Plan(ptr, nlmsize, npix)
        p = new(ptr, nlmsize, npix)
        finalizer(p) do q
            q.ptr == C_NULL || ccall((:sfb_plan_destroy, libsfb), Cvoid, (Ptr{Cvoid},), q.ptr)
            q.ptr = C_NULL
        end
        return p
    end
end

const plans = IdDict{Any,Dict{Tuple{Int,Symbol,Int},Plan}}()

"""
    plan(amodes; niter=3, precision=:fp64, device=0)

Device plan for `amodes` (spline tables, lambda tables, work buffers), cached per
amodes, niter, precision and device.
"""
function plan(amodes; niter::Integer=3, precision::Symbol=:fp64, device::Integer=0)
    cache = get!(() -> Dict{Tuple{Int,Symbol,Int},Plan}(), plans, amodes)
    key = (Int(niter), precision, Int(device))
    haskey(cache, key) && return cache[key]
    gnl = amodes.basisfunctions.gnl
    xs, ys, ypps, offs = Float64[], Float64[], Float64[], Int64[0]
    for n in 1:amodes.nmax, l in 0:amodes.lmax_n[n]
        s = gnl[n, l+1]
        append!(xs, s.x); append!(ys, s.y); append!(ypps, s.ypp)
        push!(offs, length(xs))
    end
    lmax_n = Int64.(amodes.lmax_n)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    prec = precision === :fp64 ? Int32(0) : precision === :mixed ? Int32(1) : error("precision must be :fp64 or :mixed")
    rc = ccall((:sfb_plan_create, libsfb), Cint,
               (Ptr{Ptr{Cvoid}}, Int64, Int64, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}, Ptr{Float64},
                Ptr{Float64}, Int32, Int32, Int32),
               out, amodes.nside, amodes.lmax, amodes.nmax, lmax_n, offs, xs, ys, ypps,
               Int32(niter), prec, Int32(device))
    check(rc)
    nlmsize = sum(l -> (l + 1) * (l + 2) ÷ 2, amodes.lmax_n)
    return cache[key] = Plan(out[], nlmsize, 12 * amodes.nside^2)
end

# (kind, a, b): dense W[npix, lmax+1, nmax], or separable mask[npix] and w_ln[lmax+1, nmax]
function window_arrays(win)
    if hasfield(typeof(win), :arr1) && hasfield(typeof(win), :arr2)   # SeparableArray
        return Int32(1), Vector{Float64}(vec(getfield(win, :arr1))), Matrix{Float64}(getfield(win, :arr2))
    end
    return Int32(0), Array{Float64}(win), Float64[]
end

"""
    cat2amln(rθϕ, amodes, nbar, win_rhat_ln, weight=ones(N); niter=3, precision=:fp64, device=0)

Catalogue -> SFB coefficients on the GPU; the reference's cat2amln
(src/cat2anlm.jl:257-310) with the same arguments and result.
"""
function cat2amln(rθϕ, amodes, nbar, win_rhat_ln, weight=ones(eltype(rθϕ), size(rθϕ, 2));
                  niter::Integer=3, precision::Symbol=:fp64, device::Integer=0)
    p = plan(amodes; niter, precision, device)
    rtp = Matrix{Float64}(rθϕ)
    N = size(rtp, 2)
    w = isempty(weight) ? Float64[] : Vector{Float64}(weight)
    kind, a, b = window_arrays(win_rhat_ln)
    out = Vector{Float64}(undef, 2 * p.nlmsize)
    GC.@preserve rtp w a b out begin
        rc = ccall((:sfb_cat2amln, libsfb), Cint,
                   (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Float64, Int32, Ptr{Float64},
                    Ptr{Float64}, Ptr{Float64}),
                   p.ptr, N, N > 0 ? pointer(rtp) : C_NULL, isempty(w) ? C_NULL : pointer(w),
                   Float64(nbar), kind, pointer(a), kind == 1 ? pointer(b) : C_NULL, pointer(out))
    end
    check(rc)
    return copy(reinterpret(ComplexF64, out))
end

const cat2nlm = cat2amln

"""
    cat2amln_vjp(rθϕ, amodes, nbar, dL_danlm, weight=ones(N); niter=3, precision=:fp64, device=0)

(∂L/∂r, ∂L/∂w, ∂L/∂nbar) of cat2amln for the cotangent `dL_danlm` of its output
(∂L/∂Re a + i ∂L/∂Im a).  θ and ϕ get no gradient (the pixel is piecewise constant);
g'(r) is the spline derivative of src/Splines.jl:367-377.
"""
function cat2amln_vjp(rθϕ, amodes, nbar, dL_danlm, weight=ones(eltype(rθϕ), size(rθϕ, 2));
                      niter::Integer=3, precision::Symbol=:fp64, device::Integer=0)
    p = plan(amodes; niter, precision, device)
    rtp = Matrix{Float64}(rθϕ)
    N = size(rtp, 2)
    w = isempty(weight) ? Float64[] : Vector{Float64}(weight)
    cot = Vector{ComplexF64}(dL_danlm)
    length(cot) == p.nlmsize || throw(DimensionMismatch("dL_danlm must have getnlmsize(amodes) entries"))
    dr, dw, dnbar = zeros(N), zeros(N), Ref(0.0)
    GC.@preserve rtp w cot dr dw begin
        rc = ccall((:sfb_cat2amln_vjp, libsfb), Cint,
                   (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Float64, Ptr{ComplexF64}, Ptr{Float64},
                    Ptr{Float64}, Ptr{Float64}),
                   p.ptr, N, N > 0 ? pointer(rtp) : C_NULL, isempty(w) ? C_NULL : pointer(w), Float64(nbar),
                   pointer(cot), pointer(dr), pointer(dw), dnbar)
    end
    check(rc)
    return dr, dw, dnbar[]
end

"""
    cat2amln_batch(catalogs, amodes, nbars, win_rhat_ln, weights=nothing; niter=3, precision=:fp64, device=0)

cat2amln of every 3 x N_b catalogue in `catalogs` (with its nbar) through one plan;
returns a vector of a_nlm vectors.  Equivalent to calling cat2amln per mock.
"""
function cat2amln_batch(catalogs, amodes, nbars, win_rhat_ln, weights=nothing;
                        niter::Integer=3, precision::Symbol=:fp64, device::Integer=0)
    p = plan(amodes; niter, precision, device)
    B = length(catalogs)
    Nb = Int64[size(c, 2) for c in catalogs]
    rtp = Matrix{Float64}(reduce(hcat, catalogs; init=zeros(3, 0)))
    w = weights === nothing ? Float64[] : Vector{Float64}(reduce(vcat, weights))
    nb = Vector{Float64}(collect(nbars))
    kind, a, b = window_arrays(win_rhat_ln)
    out = Vector{Float64}(undef, 2 * p.nlmsize * max(B, 1))
    GC.@preserve Nb rtp w nb a b out begin
        rc = ccall((:sfb_cat2amln_batch, libsfb), Cint,
                   (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int32, Ptr{Float64},
                    Ptr{Float64}, Ptr{Float64}),
                   p.ptr, Int32(B), Nb, isempty(rtp) ? C_NULL : pointer(rtp), isempty(w) ? C_NULL : pointer(w),
                   nb, kind, pointer(a), kind == 1 ? pointer(b) : C_NULL, out)
    end
    check(rc)
    z = reinterpret(ComplexF64, out)
    return [z[(i - 1) * p.nlmsize + 1:i * p.nlmsize] for i in 1:B]
end

"""
    win_rhat_ln(win, wmodes, amodes; device=0)

The reference's win_rhat_ln (src/windows.jl:244-270) on the GPU: a dense nr x npix
window gives the (npix, lmax+1, nmax) array (NaN where l > lmax_n[n]); a SeparableArray
(phi, mask) gives SeparableArray(mask, w_ln).
"""
function win_rhat_ln(win, wmodes, amodes; device::Integer=0)
    p = plan(amodes; device)
    r = Vector{Float64}(collect(wmodes.r))
    if hasfield(typeof(win), :arr1) && hasfield(typeof(win), :arr2)
        phi = Vector{Float64}(vec(getfield(win, :arr1)))
        out = Matrix{Float64}(undef, amodes.lmax + 1, amodes.nmax)
        rc = ccall((:sfb_win_rhat_ln, libsfb), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Float64, Int32, Ptr{Float64},
                   Ptr{Float64}), p.ptr, length(r), r, Float64(wmodes.Δr), Int32(1), phi, out)
        check(rc)
        return SFB.SeparableArray(getfield(win, :arr2), out, name1=:mask, name2=:w_ln)
    end
    w = Matrix{Float64}(win)
    out = Array{Float64}(undef, p.npix, amodes.lmax + 1, amodes.nmax)
    rc = ccall((:sfb_win_rhat_ln, libsfb), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Float64, Int32, Ptr{Float64},
               Ptr{Float64}), p.ptr, length(r), r, Float64(wmodes.Δr), Int32(0), w, out)
    check(rc)
    return out
end

"""field2anlm(f_xyz, wmodes, amodes) (src/cat2anlm.jl:361-370) on the GPU."""
function field2anlm(f_xyz, wmodes, amodes; device::Integer=0)
    p = plan(amodes; device)
    r = Vector{Float64}(collect(wmodes.r))
    f = Matrix{Float64}(f_xyz)
    out = Vector{Float64}(undef, 2 * p.nlmsize)
    rc = ccall((:sfb_field2anlm, libsfb), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Float64, Ptr{Float64}, Ptr{Float64}),
               p.ptr, length(r), r, Float64(wmodes.Δr), f, out)
    check(rc)
    return copy(reinterpret(ComplexF64, out))
end

"""anlm2field(f_nlm, wmodes, amodes) (src/cat2anlm.jl:373-422) on the GPU."""
function anlm2field(f_nlm, wmodes, amodes; device::Integer=0)
    p = plan(amodes; device)
    r = Vector{Float64}(collect(wmodes.r))
    c = Vector{ComplexF64}(f_nlm)
    out = Matrix{Float64}(undef, length(r), p.npix)
    rc = ccall((:sfb_anlm2field, libsfb), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{ComplexF64}, Ptr{Float64}),
               p.ptr, length(r), r, c, out)
    check(rc)
    return out
end

"""amln2clnn(anlm1, anlm2, cmodes) (src/SphericalFourierBesselDecompositions.jl:130-150) on
the GPU, for a ClnnModes over one AnlmModes or two (anlm1 of amodesA, anlm2 of amodesB)."""
function amln2clnn(anlm1, anlm2, cmodes; device::Integer=0)
    pa = plan(cmodes.amodesA; device)
    pb = plan(cmodes.amodesB; device)  # the same plan for a ClnnModes over one AnlmModes
    lnn = Matrix{Int64}(cmodes.lnn)
    a1, a2 = Vector{ComplexF64}(anlm1), Vector{ComplexF64}(anlm2)
    out = Vector{Float64}(undef, size(lnn, 2))
    rc = ccall((:sfb_amln2clnn_cross, libsfb), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{ComplexF64},
               Ptr{ComplexF64}, Ptr{Float64}), pa.ptr, pb.ptr, size(lnn, 2), lnn, a1, a2, out)
    check(rc)
    return out
end
amln2clnn(anlm, cmodes; kw...) = amln2clnn(anlm, anlm, cmodes; kw...)

"""pixwin(cmodes) (src/SphericalFourierBesselDecompositions.jl:158-167): the pixel window w_l of
every (l, n1, n2) entry, computed on the GPU from the window's definition (HEALPix's tables,
which the reference reads, are not used); `depth` sets the sub-pixel depth (0: by nside)."""
function pixwin(cmodes; depth::Integer=0, device::Integer=0)
    lnn = Matrix{Int64}(cmodes.lnn)
    lmax = maximum(lnn[1, :])
    w = Vector{Float64}(undef, lmax + 1)
    check(ccall((:sfb_pixwin, libsfb), Cint, (Int64, Int64, Int32, Int32, Ptr{Float64}),
                cmodes.amodesA.nside, lmax, Int32(depth), Int32(device), w))
    return [w[lnn[1, i] + 1] for i in 1:size(lnn, 2)]
end

function ChainRulesCore.rrule(::typeof(cat2amln), rθϕ, amodes, nbar, win_rhat_ln,
                              weight=ones(eltype(rθϕ), size(rθϕ, 2)); kw...)
    anlm = cat2amln(rθϕ, amodes, nbar, win_rhat_ln, weight; kw...)
    function cat2amln_pullback(ȳ)
        dr, dw, dnbar = cat2amln_vjp(rθϕ, amodes, nbar, ChainRulesCore.unthunk(ȳ), weight; kw...)
        drθϕ = zeros(size(rθϕ))
        drθϕ[1, :] .= dr
        return (ChainRulesCore.NoTangent(), drθϕ, ChainRulesCore.NoTangent(), dnbar,
                ChainRulesCore.@not_implemented("cotangent of win_rhat_ln"), isempty(weight) ? ChainRulesCore.NoTangent() : dw)
    end
    return anlm, cat2amln_pullback
end

end # module
