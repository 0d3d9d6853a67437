"""
    CrowdPhotAMD

Julia front-end of libcrowdphot_hip.so (MI355X / gfx950).  It keeps the names,
keyword arguments and result types of CrowdPhot.jl's fitting entry points and
forwards the work to the HIP library through `ccall`:

- `fit_all_stars_simultaneous`  (src/photometry/psf_photometry_simultaneous.jl:953)
- `fit_all_stars`               (sequential multi-pass, src/photometry/psf_photometry_single.jl:267)
- `fit_star`                    (per-star Levenberg-Marquardt, src/psf/psf_fitting.jl:229-308)
- `matched_filter`              (src/detection.jl:87)

Depends on the standard library only (`Libdl`), so it loads without `Pkg`
network access.  The PSF model types are CrowdPhot.jl's own
(`CircularGaussianPSF`, `CircularMoffatPSF`, `GriddedPSFModel{T,<:ImagePSF}`);
this module only reads their fields.  When CrowdPhot is loaded, the fits
return its `MultiPassPhotResult` and `LMResult` and `fit_star` its fitted
model (through ConstructionBase, which CrowdPhot loads); otherwise they
return NamedTuples with the same fields.  `lm_irls` with a user-written
`accum!` callback is Julia code and stays on the CPU: the GPU binding is
`fit_star`, whose accumulators the library implements for these models.
"""
module CrowdPhotAMD

using Libdl

const LIB = Ref{Ptr{Cvoid}}(C_NULL)
libpath() = get(ENV, "CROWDPHOT_HIP_LIB", joinpath(@__DIR__, "..", "libcrowdphot_hip.so"))

function __init__()
    LIB[] = Libdl.dlopen(libpath())
end

sym(name) = Libdl.dlsym(LIB[], name)

# ---- C structures (field for field with include/crowdphot_hip.h) ----------
struct CpPsfDesc
    kind::Int32
    fwhm::Float64
    alpha::Float64
    beta::Float64
    n_nodes::Int32
    gny::Int32
    gnx::Int32
    grids::Ptr{Float32}
    os_x::Int32
    os_y::Int32
    origin_y::Float64
    origin_x::Float64
    fill_value::Float64
    node_y::Ptr{Float64}
    node_x::Ptr{Float64}
end

mutable struct CpSimOpts
    solver::Int32
    inner_iterations::Int32
    linear_tol::Float64
    model_rad::Float64
    model_rad_max::Float64
    model_rad_nsigma::Float64
    max_step::Float64
    max_trials::Int32
    max_iter::Int32
    x_tol::Float64
    f_tol::Float64
    g_tol::Float64
    lambda_init::Float64
    lambda_up::Float64
    lambda_down::Float64
    lambda_min::Float64
    lambda_max::Float64
    covariance::Int32
    free_mask::Int32
    fixed_y::Float64
    fixed_x::Float64
    fixed_flux::Float64
    fixed_bkg::Float64
    precision::Int32
    tile::Int32
    layout::Int32
    spread_model_fwhm::Float64
end

mutable struct CpSimResult
    y::Ptr{Float64}; x::Ptr{Float64}; flux::Ptr{Float64}; bkg::Ptr{Float64}
    y_err::Ptr{Float64}; x_err::Ptr{Float64}; flux_err::Ptr{Float64}; bkg_err::Ptr{Float64}
    chisq::Ptr{Float64}
    converged::Ptr{UInt8}; valid::Ptr{UInt8}
    n_iter::Ptr{Int32}
    active::Ptr{UInt8}
    residual::Ptr{Float32}
    n_passes::Int32; n_failed::Int32
    n_active::Int32
    n_union::Int64
    cost_final::Float64
    n_trials_total::Int32; n_inner_total::Int32
    trace::Ptr{Float64}
    trace_cap::Int32; trace_len::Int32
    qfit::Ptr{Float64}; qfit_expected::Ptr{Float64}; qfit_z::Ptr{Float64}; crowding::Ptr{Float64}
    spread_model::Ptr{Float64}; spread_model_err::Ptr{Float64}
end

mutable struct CpLMOpts
    lambda_init::Float64; lambda_up::Float64; lambda_down::Float64; lambda_min::Float64; lambda_max::Float64
    max_iter::Int32
    x_tol::Float64; f_tol::Float64; g_tol::Float64
    damping::Int32
    min_diagonal::Float64
    covariance::Int32
end

mutable struct CpMFResult
    smoothed::Ptr{Float64}; smoothed_inv_var::Ptr{Float64}; significance::Ptr{Float64}; kernel::Ptr{Float64}
    kernel_norm::Float64
    peaks_cap::Int32; n_peaks::Int32
    peak_y::Ptr{Int32}; peak_x::Ptr{Int32}
    peak_significance::Ptr{Float64}; peak_flux::Ptr{Float64}
end

mutable struct Context
    handle::Ptr{Cvoid}
end

function Context(device::Integer = 0; hbm_budget::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall(sym(:cp_create), Cint, (Ref{Ptr{Cvoid}}, Cint, UInt64), h, device, hbm_budget)
    rc == 0 || error("cp_create failed ($rc)")
    ctx = Context(h[])
    finalizer(c -> ccall(sym(:cp_destroy), Cvoid, (Ptr{Cvoid},), c.handle), ctx)
    return ctx
end

function check(ctx::Context, rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall(sym(:cp_last_error), Cstring, (Ptr{Cvoid},), ctx.handle))
    rc == -1 && throw(ArgumentError(msg))
    error("libcrowdphot_hip error $rc: $msg")
end

# Launch shapes of a context (cp_tuning: blocks per CU and the simultaneous
# fit's work-tile sizes, 0 = built-in default).
mutable struct CpTuning
    hbuild_blocks_per_cu::Int32; cg_blocks_per_cu::Int32; render_blocks_per_cu::Int32; lm_blocks_per_cu::Int32
    render_tile::Int32; hbuild_lanes::Int32; fill_lanes::Int32; lm_work_tile::Int32; gk_fold_blocks::Int32
end
CpTuning() = CpTuning(0, 0, 0, 0, 0, 0, 0, 0, 0)

function get_tuning(ctx::Context)
    t = CpTuning()
    check(ctx, ccall(sym(:cp_get_tuning), Cint, (Ptr{Cvoid}, Ref{CpTuning}), ctx.handle, t))
    return t
end

set_tuning!(ctx::Context, t::CpTuning) =
    check(ctx, ccall(sym(:cp_set_tuning), Cint, (Ptr{Cvoid}, Ref{CpTuning}), ctx.handle, t))

# A loaded package module by name (CrowdPhot, ConstructionBase), or nothing.
function loaded(name::AbstractString)
    for (k, m) in Base.loaded_modules
        k.name == name && return m
    end
    return nothing
end

# _extract_source_catalog (psf_photometry_single.jl:152-197): a
# MatchedFilterResult (peaks, peak_fluxes), a Vector{<:NamedTuple} from
# measure_star_shapes (centroid.y, centroid.x, matched_filter_flux) or a
# NamedTuple (y, x, optional flux = 1 and bkg = 0).
function source_catalog(sources)
    if hasproperty(sources, :peaks) && hasproperty(sources, :peak_fluxes)
        y = Float64[Tuple(p)[1] for p in sources.peaks]
        x = Float64[Tuple(p)[2] for p in sources.peaks]
        return (y = y, x = x, flux = Float64.(sources.peak_fluxes), bkg = zeros(length(y)))
    elseif sources isa AbstractVector
        y = Float64[s.centroid.y for s in sources]
        x = Float64[s.centroid.x for s in sources]
        return (y = y, x = x, flux = Float64[s.matched_filter_flux for s in sources], bkg = zeros(length(y)))
    end
    n = length(sources.y)
    return (y = Float64.(sources.y), x = Float64.(sources.x),
            flux = haskey(sources, :flux) ? Float64.(sources.flux) : ones(n),
            bkg = haskey(sources, :bkg) ? Float64.(sources.bkg) : zeros(n))
end

# CrowdPhot's MultiPassPhotResult{FT} (psf_photometry_single.jl:99-122) from
# the result fields when CrowdPhot is loaded, else the NamedTuple itself.
function photometry_result(r, FT)
    cp = loaded("CrowdPhot")
    cp === nothing && return r
    v(a) = Vector{FT}(a)
    return cp.MultiPassPhotResult{FT}(v(r.y), v(r.x), v(r.y_err), v(r.x_err), v(r.flux), v(r.flux_err), v(r.bkg),
                                      v(r.bkg_err), r.converged, r.valid, v(r.chisq), v(r.qfit),
                                      v(r.qfit_expected), v(r.qfit_z), v(r.crowding), v(r.spread_model),
                                      v(r.spread_model_err), r.n_iter, r.n_passes, r.n_failed, r.failure_msgs,
                                      Matrix{FT}(r.residual))
end

# PSF descriptor of a CrowdPhot model (fields read by name; `fixed` shape
# values override the model's own, as model_from_vector does).  `shift`
# moves a GriddedPSFModel's node grid with a cutout's origin.
function psf_desc(psf, fixed, keep; shift = (0.0, 0.0))
    T = typeof(psf)
    name = nameof(T)
    g(k) = haskey(fixed, k) ? Float64(getfield(fixed, k)) : Float64(getproperty(psf, k))
    if name === :CircularGaussianPSF
        return CpPsfDesc(0, g(:fwhm), 0, 0, 0, 0, 0, C_NULL, 1, 1, 0, 0, 0, C_NULL, C_NULL)
    elseif name === :CircularMoffatPSF
        return CpPsfDesc(1, 0, g(:α), g(:β), 0, 0, 0, C_NULL, 1, 1, 0, 0, 0, C_NULL, C_NULL)
    elseif name === :GriddedPSFModel || name === :ImagePSF
        nodes = name === :ImagePSF ? [psf] : psf.psfs
        gy = name === :ImagePSF ? [0.0] : Float64.(psf.grid_y) .- shift[1]
        gx = name === :ImagePSF ? [0.0] : Float64.(psf.grid_x) .- shift[2]
        n0 = nodes[1]
        grids = reduce(vcat, [vec(Float32.(n.data)) for n in nodes])   # column-major per node
        push!(keep, grids, gy, gx)
        return CpPsfDesc(2, 0, 0, 0, length(nodes), size(n0.data, 1), size(n0.data, 2), pointer(grids),
                         n0.oversampling[1], n0.oversampling[2], n0.origin.y, n0.origin.x, n0.fill_value,
                         pointer(gy), pointer(gx))
    end
    throw(ArgumentError("model type $name is not built on the GPU"))
end

"""
    fit_all_stars_simultaneous(image, psf, sources, fit_rad; kws...)

Same contract, keywords and `MultiPassPhotResult` as CrowdPhot.jl
(psf_photometry_simultaneous.jl:953-1431), computed on the GPU; `solver` is
`:lsqr` (default) or `:lsmr`, as in the reference, or the build's `:cg` (CG on
the damped normal equations).  Extra keywords: `precision = :mixed | :fp64`,
`ctx::Context`.  Returns CrowdPhot's `MultiPassPhotResult` when CrowdPhot is
loaded, else a NamedTuple with its fields.
"""
function fit_all_stars_simultaneous(image::AbstractMatrix, psf, sources, fit_rad::Real;
        fixed::NamedTuple = (;), inv_var = nothing, solver::Symbol = :lsqr, inner_iterations::Int = 10,
        linear_tol::Real = 1.0e-6, model_rad::Union{Symbol, Real} = :auto, model_rad_max::Real = 10 * fit_rad,
        model_rad_nsigma::Real = 1.0, max_step::Real = 0.25, max_trials::Int = 8, max_iter::Integer = 40,
        x_tol::Real = 1.0e-5, f_tol::Real = 1.0e-4, g_tol::Real = 1.0e-4, λ_init::Real = 1.0e-4,
        λ_up::Real = 10.0, λ_down::Real = 10.0, λ_min::Real = 1.0e-12, λ_max::Real = 1.0e12,
        precision::Symbol = :mixed, spread_model_fwhm::Union{Nothing, Real} = nothing,
        show_trace::Bool = false, ctx::Context = Context(0), kws...)
    solver in (:lsqr, :lsmr, :cg) ||
        throw(ArgumentError("`solver` must be :lsqr or :lsmr (or the GPU build's :cg), got $(repr(solver))"))
    code = solver === :lsqr ? Int32(0) : solver === :cg ? Int32(1) : Int32(2)
    (model_rad === :auto || (model_rad isa Real && model_rad > 0)) ||
        throw(ArgumentError("model_rad must be :auto or a positive real, got $(repr(model_rad))"))
    props = propertynames_fit(psf)
    free = Tuple(k for k in props if !haskey(fixed, k))
    isempty(free) && throw(ArgumentError("all model parameters are fixed; nothing to fit"))
    off = setdiff(free, (:y, :x, :flux))
    isempty(off) || throw(ArgumentError("fit_all_stars_simultaneous fits only (y, x, flux) per star; got free parameters $(off)"))
    mask = Int32((:y in free ? 1 : 0) | (:x in free ? 2 : 0) | (:flux in free ? 4 : 0))
    cat = source_catalog(sources)
    y0, x0, f0 = cat.y, cat.x, cat.flux
    n = length(y0)
    ny, nx = size(image)
    img = Matrix{Float32}(image)                       # column-major already
    iv = inv_var === nothing ? nothing : Matrix{Float32}(inv_var)
    keep = Any[]
    desc = psf_desc(psf, fixed, keep)
    o = CpSimOpts(code, inner_iterations, linear_tol, model_rad === :auto ? 0.0 : Float64(model_rad),
                  model_rad_max, model_rad_nsigma, max_step, max_trials, max_iter, x_tol, f_tol, g_tol,
                  λ_init, λ_up, λ_down, λ_min, λ_max, 0, mask,
                  Float64(get(fixed, :y, 0)), Float64(get(fixed, :x, 0)), Float64(get(fixed, :flux, 0)),
                  Float64(get(fixed, :bkg, 0)), precision === :fp64 ? 1 : 0, 0, 0,
                  spread_model_fwhm === nothing ? 0.0 : Float64(spread_model_fwhm))
    out = (y = zeros(n), x = zeros(n), flux = zeros(n), bkg = zeros(n), y_err = zeros(n), x_err = zeros(n),
           flux_err = zeros(n), bkg_err = zeros(n), chisq = zeros(n), converged = zeros(UInt8, n),
           valid = zeros(UInt8, n), n_iter = zeros(Int32, n), active = zeros(UInt8, n),
           qfit = zeros(n), qfit_expected = zeros(n), qfit_z = zeros(n), crowding = zeros(n),
           spread_model = zeros(n), spread_model_err = zeros(n),
           residual = zeros(Float32, ny, nx))
    trace = zeros(8 * max(1, show_trace ? max_iter * max_trials : 1))
    r = CpSimResult(pointer(out.y), pointer(out.x), pointer(out.flux), pointer(out.bkg), pointer(out.y_err),
                    pointer(out.x_err), pointer(out.flux_err), pointer(out.bkg_err), pointer(out.chisq),
                    pointer(out.converged), pointer(out.valid), pointer(out.n_iter), pointer(out.active),
                    pointer(out.residual), 0, 0, 0, 0, 0.0, 0, 0,
                    show_trace ? pointer(trace) : C_NULL, show_trace ? Int32(length(trace) ÷ 8) : Int32(0), 0,
                    pointer(out.qfit), pointer(out.qfit_expected), pointer(out.qfit_z), pointer(out.crowding),
                    pointer(out.spread_model), pointer(out.spread_model_err))
    GC.@preserve keep out img iv trace begin
        rc = ccall(sym(:cp_fit_simultaneous), Cint,
                   (Ptr{Cvoid}, Ref{CpPsfDesc}, Ptr{Float32}, Ptr{Float32}, Int32, Int32, Ptr{Float64}, Ptr{Float64},
                    Ptr{Float64}, Int32, Float64, Ref{CpSimOpts}, Ref{CpSimResult}),
                   ctx.handle, desc, img, iv === nothing ? C_NULL : iv, ny, nx, y0, x0, f0, n, fit_rad, o, r)
        check(ctx, rc)
    end
    if show_trace
        # the reference's per-trial line (psf_photometry_simultaneous.jl:1238-1246)
        na = max(Int(r.n_active), 1)
        for k in 1:Int(r.trace_len)
            t = view(trace, 8k - 7:8k)
            conv = floor(100 * (1 - t[6] / na) * 100) / 100
            status = t[4] < t[3] ? "accepted" : "rejected"
            println("Iter ", lpad(Int(t[1]), 4), " trial ", lpad(Int(t[2]), 2), " | converged = ",
                    lpad(string(round(conv; digits = 2)), 5), "% | cost = ", t[3], " -> ", t[4],
                    " | λ = ", t[5], " | ||g|| = ", round(t[8]; digits = 4), " | ", status)
            Int(t[2]) == max_trials && status == "rejected" &&
                println("*** Max damping trials (`max_trials`) reached during outer iteration $(Int(t[1])) ***")
        end
    end
    msgs = ["star $i: excluded before the solve (no usable pixels or non-finite initial parameters)"
            for i in 1:n if out.active[i] == 0]
    res = (y = out.y, x = out.x, y_err = out.y_err, x_err = out.x_err, flux = out.flux, flux_err = out.flux_err,
           bkg = out.bkg, bkg_err = out.bkg_err, converged = BitVector(out.converged .!= 0),
           valid = BitVector(out.valid .!= 0), chisq = out.chisq, qfit = out.qfit, qfit_expected = out.qfit_expected,
           qfit_z = out.qfit_z, crowding = out.crowding, spread_model = out.spread_model,
           spread_model_err = out.spread_model_err,
           n_iter = Int.(out.n_iter), n_passes = Int(r.n_passes), n_failed = Int(r.n_failed),
           failure_msgs = msgs, residual = out.residual)
    return photometry_result(res, float(eltype(image)))
end

"""
    fit_all_stars(image, psf, sources, fit_rad; n_passes = 3, fixed = (;), inv_var = nothing,
                  spread_model_fwhm = nothing, kws...)

Sequential multi-pass photometry (psf_photometry_single.jl:267-469) on the GPU,
with fit_star's lm_irls keywords (`max_iter`, `x_tol`, `f_tol`, `g_tol`, `λ_*`).
`sources` is a NamedTuple (y, x[, flux, bkg]), a Vector{<:NamedTuple} from
measure_star_shapes or a MatchedFilterResult.  Returns CrowdPhot's
`MultiPassPhotResult` when CrowdPhot is loaded, else a NamedTuple with its
fields.
"""
function fit_all_stars(image::AbstractMatrix, psf, sources, fit_rad::Real; n_passes::Integer = 3,
        fixed::NamedTuple = (;), inv_var = nothing, spread_model_fwhm::Union{Nothing, Real} = nothing,
        max_iter::Integer = 200, x_tol::Real = 1.0e-8, f_tol::Real = 1.0e-4, g_tol::Real = 1.0e-4,
        λ_init::Real = 1.0e-4, λ_up::Real = 10.0, λ_down::Real = 10.0, λ_min::Real = 1.0e-12,
        λ_max::Real = 1.0e12, precision::Symbol = :mixed, ctx::Context = Context(0), kws...)
    n_passes > 0 || throw(ArgumentError("n_passes must be positive, got $n_passes"))
    props = propertynames_fit(psf)
    free = Tuple(k for k in props if !haskey(fixed, k))
    isempty(free) && throw(ArgumentError("all model parameters are fixed; nothing to fit"))
    mask = UInt32(sum(k in free ? (1 << (i - 1)) : 0 for (i, k) in enumerate(props)))
    fixv = Float64[haskey(fixed, k) ? Float64(getfield(fixed, k)) : NaN for k in props]
    cat = source_catalog(sources)
    n = length(cat.y)
    theta0 = zeros(length(props), n)                       # column per star = row-major n x npf
    for (i, k) in enumerate(props)
        theta0[i, :] .= k === :y ? cat.y : k === :x ? cat.x : k === :flux ? cat.flux : k === :bkg ? cat.bkg :
                        fill(Float64(getproperty(psf, k)), n)
    end
    ny, nx = size(image)
    img = Matrix{Float32}(image)
    iv = inv_var === nothing ? nothing : Matrix{Float32}(inv_var)
    keep = Any[]
    desc = psf_desc(psf, (;), keep)
    o = CpLMOpts(λ_init, λ_up, λ_down, λ_min, λ_max, max_iter, x_tol, f_tol, g_tol, 0, 1.0e-6, 0)
    out = (y = zeros(n), x = zeros(n), flux = zeros(n), bkg = zeros(n), y_err = zeros(n), x_err = zeros(n),
           flux_err = zeros(n), bkg_err = zeros(n), chisq = zeros(n), converged = zeros(UInt8, n),
           valid = zeros(UInt8, n), n_iter = zeros(Int32, n), qfit = zeros(n), qfit_expected = zeros(n),
           qfit_z = zeros(n), crowding = zeros(n), spread_model = zeros(n), spread_model_err = zeros(n),
           residual = zeros(Float32, ny, nx))
    r = CpSimResult(pointer(out.y), pointer(out.x), pointer(out.flux), pointer(out.bkg), pointer(out.y_err),
                    pointer(out.x_err), pointer(out.flux_err), pointer(out.bkg_err), pointer(out.chisq),
                    pointer(out.converged), pointer(out.valid), pointer(out.n_iter), C_NULL,
                    pointer(out.residual), 0, 0, 0, 0, 0.0, 0, 0, C_NULL, 0, 0,
                    pointer(out.qfit), pointer(out.qfit_expected), pointer(out.qfit_z), pointer(out.crowding),
                    pointer(out.spread_model), pointer(out.spread_model_err))
    GC.@preserve keep out img iv theta0 fixv begin
        rc = ccall(sym(:cp_fit_all_stars), Cint,
                   (Ptr{Cvoid}, Ref{CpPsfDesc}, Ptr{Float32}, Ptr{Float32}, Int32, Int32, Ptr{Float64}, Int32,
                    Float64, Int32, UInt32, Ptr{Float64}, Ref{CpLMOpts}, Float64, Int32, Ref{CpSimResult}),
                   ctx.handle, desc, img, iv === nothing ? C_NULL : iv, ny, nx, theta0, n, fit_rad, n_passes, mask,
                   fixv, o, spread_model_fwhm === nothing ? 0.0 : Float64(spread_model_fwhm),
                   precision === :fp64 ? 1 : 0, r)
        check(ctx, rc)
    end
    res = (y = out.y, x = out.x, y_err = out.y_err, x_err = out.x_err, flux = out.flux, flux_err = out.flux_err,
           bkg = out.bkg, bkg_err = out.bkg_err, converged = BitVector(out.converged .!= 0),
           valid = BitVector(out.valid .!= 0), chisq = out.chisq, qfit = out.qfit, qfit_expected = out.qfit_expected,
           qfit_z = out.qfit_z, crowding = out.crowding, spread_model = out.spread_model,
           spread_model_err = out.spread_model_err, n_iter = Int.(out.n_iter), n_passes = Int(r.n_passes),
           n_failed = Int(r.n_failed), failure_msgs = String[],
           residual = n == 0 ? Matrix{Float32}(undef, 0, 0) : out.residual)
    return photometry_result(res, float(eltype(image)))
end

mutable struct CpLMStat
    minimum::Float64; cost_init::Float64; lambda_final::Float64; chisq::Float64
    iterations::Int32
    converged::UInt8; x_converged::UInt8; f_converged::UInt8; g_converged::UInt8
    status::Int32
end
CpLMStat() = CpLMStat(0, 0, 0, 0, 0, 0, 0, 0, 0, 0)

"""
    fit_star(model, image, inds = axes(image); fixed = (;), inv_var = nothing, kws...) -> (best_model, LMResult)

Per-star Levenberg-Marquardt fit (psf_fitting.jl:229-308, lm_irls
levenberg_marquardt.jl:431-680) on the GPU over the pixels `inds` of `image`,
in the image's own coordinates; `fixed` values override the model's, and
`inv_var` (same size as `image`) weights the pixels, zero or non-finite
values masking them.  Keywords: `λ_init, λ_up, λ_down, λ_min, λ_max,
max_iter, x_tol, f_tol, g_tol`.  `best_model` is the model with the fitted
parameters (ConstructionBase.setproperties); `LMResult` is CrowdPhot's when it
is loaded, else a NamedTuple with its fields.  IRLS reweighting
(`reweight`) is not built on the GPU.
"""
function fit_star(model, image::AbstractMatrix, inds = axes(image); fixed::NamedTuple = (;), inv_var = nothing,
        λ_init::Real = 1.0e-4, λ_up::Real = 10.0, λ_down::Real = 10.0, λ_min::Real = 1.0e-12, λ_max::Real = 1.0e12,
        max_iter::Integer = 200, x_tol::Real = 1.0e-8, f_tol::Real = 1.0e-4, g_tol::Real = 1.0e-4,
        reweight = nothing, ctx::Context = Context(0), kws...)
    inv_var === nothing || size(inv_var) == size(image) ||
        throw(ArgumentError("`inv_var` must be the same size as `image`"))
    reweight === nothing || throw(ArgumentError("IRLS reweighting is not built on the GPU"))
    props = propertynames_fit(model)
    free = Tuple(k for k in props if !haskey(fixed, k))
    isempty(free) && throw(ArgumentError("all model parameters are fixed; nothing to fit"))
    mask = UInt32(sum(k in free ? (1 << (i - 1)) : 0 for (i, k) in enumerate(props)))
    rows, cols = inds isa CartesianIndices ? inds.indices : inds
    r0, c0 = first(rows) - 1, first(cols) - 1
    cut = Matrix{Float32}(image[rows, cols])               # column-major cutout
    ivc = inv_var === nothing ? nothing : Matrix{Float32}(map(v -> isfinite(v) && v > 0 ? v : zero(v),
                                                              inv_var[rows, cols]))
    if ivc !== nothing
        nv = count(>(0), ivc)
        nv > length(free) || throw(ArgumentError("too few valid inv_var pixels ($nv) for $(length(free)) free parameters"))
    end
    val(k) = haskey(fixed, k) ? Float64(getfield(fixed, k)) : Float64(getproperty(model, k))
    theta0 = Float64[val(k) - (k === :y ? r0 : k === :x ? c0 : 0.0) for k in props]
    keep = Any[]
    desc = psf_desc(model, fixed, keep; shift = (Float64(r0), Float64(c0)))
    o = CpLMOpts(λ_init, λ_up, λ_down, λ_min, λ_max, max_iter, x_tol, f_tol, g_tol, 0, 1.0e-6, 0)
    npf, np = length(props), length(free)
    theta = zeros(npf)
    cov = zeros(np, np)
    st = Ref(CpLMStat())
    off = Int64[0]
    cny, cnx = Int32[size(cut, 1)], Int32[size(cut, 2)]
    GC.@preserve keep cut ivc theta0 theta cov off cny cnx begin
        rc = ccall(sym(:cp_fit_star_batch), Cint,
                   (Ptr{Cvoid}, Ref{CpPsfDesc}, Ptr{Float32}, Ptr{Int64}, Ptr{Int32}, Ptr{Int32}, Ptr{Float32},
                    Ptr{Float64}, UInt32, Int32, Ref{CpLMOpts}, Ptr{Float64}, Ptr{Float64}, Ref{CpLMStat}),
                   ctx.handle, desc, cut, off, cny, cnx, ivc === nothing ? C_NULL : ivc, theta0, mask, 1, o, theta,
                   cov, st)
        check(ctx, rc)
    end
    s = st[]
    s.status == 0 || throw(ArgumentError("fit_star rejected the cutout (too few observations for the free parameters)"))
    theta[1] += r0
    theta[2] += c0
    fitted = NamedTuple{props}(Tuple(theta))
    cb = loaded("ConstructionBase")
    best = cb === nothing ? fitted : cb.setproperties(model, fitted)
    FT = float(eltype(image))
    free_idx = [i for (i, k) in enumerate(props) if !haskey(fixed, k)]
    mini = Vector{FT}(theta[free_idx])
    cp = loaded("CrowdPhot")
    lm = cp === nothing ?
        (minimizer = mini, minimum = FT(s.minimum), cost_init = FT(s.cost_init), converged = s.converged != 0,
         x_converged = s.x_converged != 0, f_converged = s.f_converged != 0, g_converged = s.g_converged != 0,
         iterations = Int(s.iterations), λ_final = FT(s.lambda_final), σ_final = FT(NaN), cov = Matrix{FT}(cov),
         chisq = FT(s.chisq)) :
        cp.LMResult(mini, FT(s.minimum), FT(s.cost_init), s.converged != 0, s.x_converged != 0, s.f_converged != 0,
                    s.g_converged != 0, Int(s.iterations), FT(s.lambda_final), FT(NaN), Matrix{FT}(cov), FT(s.chisq))
    return best, lm
end

"""
    matched_filter(image, kernel::AbstractMatrix; inv_var = nothing, normalize_zerosum = true,
                   sigma = 5.0, border = :replicate)

Matched-filter detection (detection.jl:87-205) on the GPU.  Returns a
NamedTuple with the MatchedFilterResult fields; `peaks` are CartesianIndex.
"""
function matched_filter(image::AbstractMatrix, kernel::AbstractMatrix; inv_var = nothing,
        normalize_zerosum::Bool = true, sigma::Real = 5.0, border::Symbol = :replicate, ctx::Context = Context(0))
    border in (:replicate, :zero) || throw(ArgumentError("unknown border mode :$border; use :replicate or :zero"))
    inv_var === nothing || size(inv_var) == size(image) ||
        throw(DimensionMismatch("inv_var size $(size(inv_var)) must match image size $(size(image))"))
    ny, nx = size(image)
    kr, kc = size(kernel)
    img = Matrix{Float32}(image)
    iv = inv_var === nothing ? nothing : Matrix{Float32}(map(v -> isfinite(v) ? v : zero(v), inv_var))
    K = Matrix{Float64}(kernel)
    cap = (div(ny + 1, 2) + 1) * (div(nx + 1, 2) + 1)
    sm, sig = zeros(ny, nx), zeros(ny, nx)
    den = inv_var === nothing ? nothing : zeros(ny, nx)
    kout = zeros(kr, kc)
    py, px = zeros(Int32, cap), zeros(Int32, cap)
    ps, pf = zeros(cap), zeros(cap)
    r = CpMFResult(pointer(sm), den === nothing ? C_NULL : pointer(den), pointer(sig), pointer(kout), 0.0,
                   cap, 0, pointer(py), pointer(px), pointer(ps), pointer(pf))
    GC.@preserve img iv K sm sig den kout py px ps pf begin
        rc = ccall(sym(:cp_matched_filter), Cint,
                   (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Int32, Int32, Ptr{Float64}, Int32, Int32, Int32,
                    Float64, Int32, Ref{CpMFResult}),
                   ctx.handle, img, iv === nothing ? C_NULL : iv, ny, nx, K, kr, kc, normalize_zerosum, sigma,
                   border === :zero ? 1 : 0, r)
        check(ctx, rc)
    end
    m = Int(r.n_peaks)
    return (image = image, inv_var = iv, smoothed_image = sm, smoothed_inv_var = den, significance_map = sig,
            kernel = kout, kernel_norm = r.kernel_norm,
            peaks = [CartesianIndex(Int(py[i]), Int(px[i])) for i in 1:m],
            peak_significances = ps[1:m], peak_fluxes = pf[1:m])
end

propertynames_fit(psf) = begin
    name = nameof(typeof(psf))
    name === :CircularGaussianPSF ? (:y, :x, :fwhm, :flux, :bkg) :
    name === :CircularMoffatPSF ? (:y, :x, :α, :β, :flux, :bkg) : (:y, :x, :flux, :bkg)
end

end # module
