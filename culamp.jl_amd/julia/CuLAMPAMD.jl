# CuLAMPAMD.jl -- the Julia side of the drop-in: CuLAMP.jl's hot-path entry
# points (same names, arguments and returns) calling libculamp's C ABI
# (include/culamp.h) with `ccall`, so that the LAMP / Sinkhorn solvers run on
# an MI355X.  Host arrays replace the reference's CuArrays (same memory
# layout: 3 x N point clouds, column-major feature matrices); the library owns
# every device buffer.  Included from the CuLAMP module, which provides
# EOTArgs (src/Utilities.jl:18-31); any struct with EOTArgs' fields works.
module CuLAMPAMD

import ChainRulesCore
using ChainRulesCore: NoTangent

const libculamp = joinpath(@__DIR__, "..", "libculamp.so")

const CULAMP_FP64 = Int32(0)
const CULAMP_MIXED = Int32(1)
const CULAMP_BF16 = Int32(2)
const CULAMP_BF16X3 = Int32(3)
const PRECISION = Dict("fp64" => CULAMP_FP64, "mixed" => CULAMP_MIXED,
                       "bf16" => CULAMP_BF16, "bf16x3" => CULAMP_BF16X3)
const STOP = Dict("reference" => Int32(0), "infeas" => Int32(1))
const METRIC = Dict(:l1 => Int32(0), :l2 => Int32(1), :cosine => Int32(2), :pearson => Int32(3),
                    :correlation => Int32(3))

# culamp_args: EOTArgs field for field plus precision / stop
Base.@kwdef mutable struct CArgs
    eta_p::Float64 = 0.0
    eta_mu::Float64 = 0.0
    tau_p::Float64 = 1.0
    tau_mu::Float64 = 1.0
    alpha::Float64 = 0.01
    B::Float64 = 1.0
    epsilon::Float64 = 1e-4
    anneal_mult::Float64 = 0.95
    itermax::Int64 = 10_000
    inner_iter::Int64 = 10
    tmax::Float64 = Inf
    verbose::Int32 = 1
    precision::Int32 = 0
    stop::Int32 = 0
    reserved::Int32 = 0
end
CArgs(a; precision="fp64", stop="reference") = CArgs(Float64(a.eta_p), Float64(a.eta_mu),
    Float64(a.tau_p), Float64(a.tau_mu), Float64(a.alpha), Float64(a.B), Float64(a.epsilon),
    Float64(a.anneal_mult), Int64(a.itermax), Int64(a.inner_iter), Float64(a.tmax),
    Int32(a.verbose), PRECISION[string(precision)], STOP[string(stop)], Int32(0))

# culamp_result
struct CResult
    iterations::Int64
    objective::Float64
    residual::Float64
    primal::Float64
    dual::Float64
    winf::Float64
    eta_t::Float64
    elapsed_s::Float64
    pairs::Float64
    passes::Int64
    converged::Int32
    nonfinite::Int32
    fp64_from::Int64
end

# culamp_instance / culamp_batch_tuning / culamp_divergence_opts
struct CInstance
    n::Int64
    m::Int64
    x::Ptr{Float64}
    y::Ptr{Float64}
    r::Ptr{Float64}
    c::Ptr{Float64}
end
mutable struct CTuning
    rows_per_thread::Int32
    graph::Int32
    max_split::Int64
    min_items::Int64
end
mutable struct DivOpts
    eps::Float64
    eta_p::Float64
    tol::Float64
    polish_tol::Float64
    itermax::Int64
    frequency::Int64
    unroll::Int32
    checkpoint::Int32
    precision::Int32
    reserved::Int32
    tmax::Float64
end

last_error() = unsafe_string(ccall((:culamp_last_error, libculamp), Cstring, ()))
check(rc) = rc == 0 || error("libculamp: ", last_error())

mutable struct Ctx
    h::Ptr{Cvoid}
    function Ctx(device::Integer=0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:culamp_ctx_create, libculamp), Cint, (Cint, Ref{Ptr{Cvoid}}), device, r))
        c = new(r[])
        finalizer(c -> ccall((:culamp_ctx_destroy, libculamp), Cint, (Ptr{Cvoid},), c.h), c)
    end
end
const CTX = Ref{Ctx}()
ctx() = isassigned(CTX) ? CTX[] : (CTX[] = Ctx(0))

"FP64 takeover of reduced-precision solves that stall (default on)."
set_fp64_fallback(on::Bool) =
    check(ccall((:culamp_ctx_set_fp64_fallback, libculamp), Cint, (Ptr{Cvoid}, Cint), ctx().h, on))

# log lines arrive already formatted exactly like the reference's @printf
function _print_line(::Ptr{Cvoid}, line::Cstring)::Cvoid
    print(unsafe_string(line))
    nothing
end
const LOGFN = Ref{Ptr{Cvoid}}(C_NULL)
logfn() = LOGFN[] == C_NULL ? (LOGFN[] = @cfunction(_print_line, Cvoid, (Ptr{Cvoid}, Cstring))) : LOGFN[]

# ---- problems ---------------------------------------------------------------

function problem(img1::Matrix{Float64}, img2::Matrix{Float64}, r::Vector{Float64}, c::Vector{Float64})
    @assert size(img1, 1) == 3 && size(img2, 1) == 3
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:culamp_problem_create, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64},
         Ref{Ptr{Cvoid}}),
        ctx().h, img1, size(img1, 2), img2, size(img2, 2), r, c, p))
    p[]
end
destroy(p) = ccall((:culamp_problem_destroy, libculamp), Cint, (Ptr{Cvoid},), p)

# rows of X (n x d, column-major) are the points; Y === nothing: the self cost
function feature_problem(X::Matrix{Float64}, Y::Union{Matrix{Float64},Nothing}, scale, metric,
                         r::Vector{Float64}, c::Vector{Float64})
    p = Ref{Ptr{Cvoid}}(C_NULL)
    m = Y === nothing ? size(X, 1) : size(Y, 1)
    check(ccall((:culamp_feature_problem_create, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Int64, Cint, Float64,
         Ptr{Float64}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
        ctx().h, X, size(X, 1), Y === nothing ? C_NULL : Y, m, size(X, 2), METRIC[Symbol(metric)],
        Float64(scale), r, c, p))
    p[]
end

# ---- solvers (src/solvers/*.jl, same names and returns) ---------------------

# extragradient_color_transfer (src/solvers/MirrorProxKernel.jl:1019-1132)
function extragradient_color_transfer(img1::Matrix{Float64}, img2::Matrix{Float64},
        marginal1::Vector{Float64}, marginal2::Vector{Float64}, args, frequency::Int=100,
        p::Float64=2.0; precision="fp64", stop="reference")
    N, M = size(img1, 2), size(img2, 2)
    ν, φ, ψ = zeros(M), zeros(N), zeros(M)
    out1, out2 = zeros(3, N), zeros(3, M)
    res = Ref{CResult}()
    pb = problem(img1, img2, marginal1, marginal2)
    try
        check(ccall((:culamp_lamp_solve, libculamp), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ref{CArgs}, Int64, Float64, Ptr{Cvoid}, Ptr{Cvoid},
             Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{CResult}),
            ctx().h, pb, CArgs(args; precision, stop), frequency, p, logfn(), C_NULL,
            ν, φ, ψ, out1, out2, res))
    finally
        destroy(pb)
    end
    return ν, φ, ψ, out1, out2
end

# sinkhorn_color_transfer (src/solvers/SinkhornKernel.jl:477-559)
function sinkhorn_color_transfer(img1::Matrix{Float64}, img2::Matrix{Float64},
        marginal1::Vector{Float64}, marginal2::Vector{Float64}, args, frequency::Int=100, p=2;
        return_cuda::Bool=false, return_assignments::Bool=true,
        φ0::Union{Vector{Float64},Nothing}=nothing, ψ0::Union{Vector{Float64},Nothing}=nothing,
        precision="fp64")
    N, M = size(img1, 2), size(img2, 2)
    φ, ψ = zeros(N), zeros(M)
    out1 = return_assignments ? zeros(3, N) : Float64[]
    out2 = return_assignments ? zeros(3, M) : Float64[]
    res = Ref{CResult}()
    pb = problem(img1, img2, marginal1, marginal2)
    try
        check(ccall((:culamp_sinkhorn_solve, libculamp), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ref{CArgs}, Int64, Float64, Ptr{Float64}, Ptr{Float64},
             Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
             Ref{CResult}),
            ctx().h, pb, CArgs(args; precision), frequency, Float64(p),
            φ0 === nothing ? C_NULL : φ0, ψ0 === nothing ? C_NULL : ψ0, logfn(), C_NULL,
            φ, ψ, return_assignments ? out1 : C_NULL, return_assignments ? out2 : C_NULL, res))
    finally
        destroy(pb)
    end
    return_assignments && return φ, ψ, out1, out2
    return_cuda && return φ, ψ, res[].objective, res[].residual, res[].iterations
    return φ, ψ, res[].objective, res[].residual
end

# the solve of annealed_sinkhorn_euclidean (src/solvers/AnnealedSinkhorn.jl:65-145)
function annealed_sinkhorn(img1::Matrix{Float64}, img2::Matrix{Float64},
        marginal1::Vector{Float64}, marginal2::Vector{Float64}, args, frequency::Int=100,
        p::Float64=2.0; precision="fp64")
    N, M = size(img1, 2), size(img2, 2)
    φ, ψ = zeros(N), zeros(M)
    out1, out2 = zeros(3, N), zeros(3, M)
    res = Ref{CResult}()
    pb = problem(img1, img2, marginal1, marginal2)
    try
        check(ccall((:culamp_annealed_sinkhorn_solve, libculamp), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ref{CArgs}, Int64, Float64, Ptr{Cvoid}, Ptr{Cvoid},
             Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{CResult}),
            ctx().h, pb, CArgs(args; precision), frequency, p, logfn(), C_NULL,
            φ, ψ, out1, out2, res))
    finally
        destroy(pb)
    end
    return φ, ψ, out1, out2, res[]
end

# extragradient_cellsim (src/solvers/MirrorProxCellSim.jl:94-188): `data` is the
# reference's features x points matrix; its transpose holds the points as rows
function extragradient_cellsim(data::Matrix{Float64}, scale, metric, marginal1, marginal2, args,
                               frequency::Int=100; precision="fp64", stop="reference")
    X = Matrix(permutedims(data))
    N = size(X, 1)
    ν, φ, ψ = zeros(N), zeros(N), zeros(N)
    res = Ref{CResult}()
    pb = feature_problem(X, nothing, scale, metric, marginal1, marginal2)
    try
        check(ccall((:culamp_lamp_cellsim_solve, libculamp), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ref{CArgs}, Int64, Ptr{Cvoid}, Ptr{Cvoid},
             Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{CResult}),
            ctx().h, pb, CArgs(args; precision, stop), frequency, logfn(), C_NULL, ν, φ, ψ, res))
    finally
        destroy(pb)
    end
    return ν, φ, ψ, res[].objective
end

# sinkhorn_cellsim (src/solvers/SinkhornCellSim.jl:240-310)
function sinkhorn_cellsim(data1::Matrix{Float64}, data2::Matrix{Float64}, scale, metric,
                          marginal1, marginal2, args, frequency::Int=100;
                          φ0=nothing, ψ0=nothing, precision="fp64")
    X, Y = Matrix(permutedims(data1)), Matrix(permutedims(data2))
    φ, ψ = zeros(size(X, 1)), zeros(size(Y, 1))
    res = Ref{CResult}()
    pb = feature_problem(X, Y, scale, metric, marginal1, marginal2)
    try
        check(ccall((:culamp_sinkhorn_cellsim_solve, libculamp), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ref{CArgs}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid},
             Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ref{CResult}),
            ctx().h, pb, CArgs(args; precision), frequency,
            φ0 === nothing ? C_NULL : φ0, ψ0 === nothing ? C_NULL : ψ0, logfn(), C_NULL,
            φ, ψ, res))
    finally
        destroy(pb)
    end
    return φ, ψ, res[].objective
end

# annealed_sinkhorn_cellsim (src/solvers/AnnealedSinkhornCellSim.jl:1-63)
function annealed_sinkhorn_cellsim(data1::Matrix{Float64}, data2::Matrix{Float64}, scale, metric,
                                   marginal1, marginal2, args, frequency::Int=100;
                                   precision="fp64")
    X, Y = Matrix(permutedims(data1)), Matrix(permutedims(data2))
    φ, ψ = zeros(size(X, 1)), zeros(size(Y, 1))
    res = Ref{CResult}()
    pb = feature_problem(X, Y, scale, metric, marginal1, marginal2)
    try
        check(ccall((:culamp_annealed_sinkhorn_cellsim_solve, libculamp), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ref{CArgs}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64},
             Ptr{Float64}, Ref{CResult}),
            ctx().h, pb, CArgs(args; precision), frequency, logfn(), C_NULL, φ, ψ, res))
    finally
        destroy(pb)
    end
    return φ, ψ, res[].objective
end

# ---- kernel level: the reference's @cuda kernels on device pointers ---------
# (dx, dy: 3 x N device arrays from dev(); every call is stream-ordered)

function dev(a::Array{Float64})
    p = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:culamp_malloc, libculamp), Cint, (Ptr{Cvoid}, Csize_t, Ref{Ptr{Cvoid}}),
                ctx().h, sizeof(a), p))
    check(ccall((:culamp_memcpy_htod, libculamp), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Csize_t),
                ctx().h, p[], a, sizeof(a)))
    Ptr{Float64}(p[])
end
function host!(a::Array{Float64}, d::Ptr{Float64})
    check(ccall((:culamp_memcpy_dtoh, libculamp), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Csize_t),
                ctx().h, a, d, sizeof(a)))
    a
end
free(d) = check(ccall((:culamp_free, libculamp), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx().h, d))

# warp_logsumexp_spp_ct_fused! / _opt! (MirrorProxKernel.jl:833-938 / 702-829)
warp_logsumexp_spp_ct_fused!(out, img1, img2, θ, N, M, reg, st, W∞, p; precision=CULAMP_FP64) =
    check(ccall((:culamp_lse_rows, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64},
         Float64, Float64, Float64, Float64, Cint, Ptr{Float64}),
        ctx().h, img1, N, img2, M, θ, reg, st, W∞, p, precision, out))
# residual_spp_c! (MirrorProxKernel.jl:7-94); cost may be C_NULL
residual_spp_c!(res, cost, img1, img2, r, θ, logZ, N, M, reg, st, W∞, p; precision=CULAMP_FP64) =
    check(ccall((:culamp_col_marginal_cost, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Float64, Float64, Float64, Float64, Cint, Ptr{Float64}, Ptr{Float64}),
        ctx().h, img1, N, img2, M, r, θ, logZ, reg, st, W∞, p, precision, res, cost))
# max_logsumexp_spp_ct! (MirrorProxKernel.jl:280-326)
max_logsumexp_spp_ct!(out, img1, img2, N, M, p) =
    check(ccall((:culamp_max_cost, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Float64, Ptr{Float64}),
        ctx().h, img1, N, img2, M, p, out))
# warp_min_reduce! (MirrorProxKernel.jl:941-1016)
warp_min_reduce!(out, img1, img2, θ, N, M, W∞, p) =
    check(ccall((:culamp_minplus_rows, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Float64, Float64,
         Ptr{Float64}),
        ctx().h, img1, N, img2, M, θ, W∞, p, out))
# update_θ_residual (MirrorProx.jl:159-180)
update_θ_residual!(θ, θ0, res, c, eta_mu, M, eta_mu_s, adjust, minv, maxv, W∞) =
    check(ccall((:culamp_lamp_theta_update, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64,
         Float64, Cint, Float64, Float64, Float64),
        ctx().h, θ, θ0, res, c, eta_mu, M, eta_mu_s, adjust, minv, maxv, W∞))
# warp_logsumexp_ct_fused! / _ct_opt! (SinkhornKernel.jl:341-444 / 230-338; p = 2 for _opt!)
warp_logsumexp_ct_fused!(out, img1, img2, marg, ψ, N, M, reg, W∞, p; precision=CULAMP_FP64) =
    check(ccall((:culamp_sinkhorn_half_step, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64},
         Float64, Float64, Float64, Cint, Ptr{Float64}),
        ctx().h, img1, N, img2, M, marg, ψ, reg, W∞, p, precision, out))
# residual_opt! (SinkhornKernel.jl:9-81)
residual_opt!(res, cost, img1, img2, marg, φ, ψ, N, M, reg, W∞, p; precision=CULAMP_FP64) =
    check(ccall((:culamp_row_residual_cost, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Float64, Float64, Float64, Cint, Ptr{Float64}, Ptr{Float64}),
        ctx().h, img1, N, img2, M, marg, φ, ψ, reg, W∞, p, precision, res, cost))
# naive_findmaxindex_ct! (SinkhornKernel.jl:124-162)
naive_findmaxindex_ct!(out, img1, img2, φ, ψ, N, M, reg, W∞, p) =
    check(ccall((:culamp_barycentric_sinkhorn, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64},
         Float64, Float64, Float64, Ptr{Float64}),
        ctx().h, img1, N, img2, M, φ, ψ, reg, W∞, p, out))
# naive_findmaxindex_spp_ct! (MirrorProxKernel.jl:97-142): rows, out is 3 x N
naive_findmaxindex_spp_ct!(out, img1, img2, ν, r, logZ, N, M, reg, W∞, p) =
    check(ccall((:culamp_barycentric_lamp_rows, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Float64, Float64, Float64, Ptr{Float64}),
        ctx().h, img1, N, img2, M, ν, r, logZ, reg, W∞, p, out))
# naive_findmaxindex_spp_ct_t! (MirrorProxKernel.jl:144-189): columns, out is 3 x M
naive_findmaxindex_spp_ct_t!(out, img1, img2, r, ν, logZ, N, M, reg, W∞, p) =
    check(ccall((:culamp_barycentric_lamp_cols, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Float64, Float64, Float64, Ptr{Float64}),
        ctx().h, img1, N, img2, M, r, ν, logZ, reg, W∞, p, out))

# HIP graph of kernel-level calls (issue each call once before capturing it)
capture_begin() = check(ccall((:culamp_capture_begin, libculamp), Cint, (Ptr{Cvoid},), ctx().h))
function capture_end()
    g = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:culamp_capture_end, libculamp), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), ctx().h, g))
    g[]
end
replay(g) = check(ccall((:culamp_graph_launch, libculamp), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), ctx().h, g))
destroy_graph(g) = ccall((:culamp_graph_destroy, libculamp), Cint, (Ptr{Cvoid},), g)
synchronize() = check(ccall((:culamp_ctx_synchronize, libculamp), Cint, (Ptr{Cvoid},), ctx().h))

# ---- dense cost matrix (src/solvers/MirrorProx.jl) ---------------------------
# Wt: the reference's column-major transpose, i.e. W (n x m) in row-major bytes
function dense_matrix(Wt::Matrix{Float64}; precision="fp64")
    m, n = size(Wt)
    dm = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:culamp_dense_create, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Int64, Cint, Ref{Ptr{Cvoid}}),
        ctx().h, Wt, n, m, PRECISION[precision], dm))
    dm[]
end
destroy_dense(dm) = ccall((:culamp_dense_destroy, libculamp), Cint, (Ptr{Cvoid},), dm)
# warp_logsumexp_fused! (MirrorProx.jl:236-296) on device vectors
warp_logsumexp_fused!(sumvals, dm, θ, η, s, W∞) =
    check(ccall((:culamp_dense_lse_rows, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Float64, Float64, Float64, Ptr{Float64}),
        ctx().h, dm, θ, η, s, W∞, sumvals))
# residual_c! (MirrorProx.jl:299-333) on device vectors
residual_c!(out, dm, r, θ, logZ, η, s, W∞) =
    check(ccall((:culamp_dense_col_marginal, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64,
         Float64, Ptr{Float64}),
        ctx().h, dm, r, θ, logZ, η, s, W∞, out))
# LAMP(r, c, W::CuArray, args, frequency; _θ) (MirrorProx.jl:397-503): returns (r .* p, θ)
function LAMP(r::Vector{Float64}, c::Vector{Float64}, W::Matrix{Float64}, args, frequency::Int=100;
              _θ::Union{Vector{Float64},Nothing}=nothing, precision="fp64", stop="reference")
    n, m = size(W)
    dm = dense_matrix(Matrix(permutedims(W)); precision)
    θ, plan = zeros(m), zeros(n, m)   # plan: n x m row-major = W's transpose layout
    res = Ref{CResult}()
    try
        check(ccall((:culamp_dense_lamp_solve, libculamp), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ref{CArgs}, Int64, Ptr{Float64},
             Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ref{CResult}),
            ctx().h, dm, r, c, CArgs(args; precision, stop), frequency,
            _θ === nothing ? C_NULL : _θ, logfn(), C_NULL, θ, plan, res))
    finally
        destroy_dense(dm)
    end
    return permutedims(reshape(plan, m, n)), θ
end

# ---- batches: many instances per launch (no reference symbol) -------------

function batch(Xs, Ys, rs, cs, scale, metric)
    inst = [CInstance(size(X, 1), size(Y, 1), pointer(X), pointer(Y), pointer(r), pointer(c))
            for (X, Y, r, c) in zip(Xs, Ys, rs, cs)]
    b = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve Xs Ys rs cs check(ccall((:culamp_batch_create, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{CInstance}, Int32, Int64, Cint, Float64, Ref{Ptr{Cvoid}}),
        ctx().h, inst, length(inst), size(Xs[1], 2), METRIC[Symbol(metric)], Float64(scale), b))
    b[]
end
destroy_batch(b) = ccall((:culamp_batch_destroy, libculamp), Cint, (Ptr{Cvoid},), b)
function autotune!(b, precision="bf16"; reps=5)
    t, ms = CTuning(0, 0, 0, 0), Ref{Float64}()
    check(ccall((:culamp_batch_autotune, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Float64, Cint, Ref{CTuning}, Ref{Float64}),
        ctx().h, b, PRECISION[precision], 2.0, reps, t, ms))
    t, ms[]
end
# extragradient_cellsim on every instance of a batch; results in instance order
function batch_lamp_cellsim(Xs, Ys, rs, cs, scale, metric, args, frequency;
                            precision="bf16", stop="infeas")
    b = batch(Xs, Ys, rs, cs, scale, metric)
    N, M = sum(size.(Xs, 1)), sum(size.(Ys, 1))
    ν, φ, ψ = zeros(M), zeros(N), zeros(M)
    res = Vector{CResult}(undef, length(Xs))
    try
        autotune!(b, precision)
        check(ccall((:culamp_batch_lamp_solve, libculamp), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ref{CArgs}, Int64, Float64, Ptr{Float64}, Ptr{Float64},
             Ptr{Float64}, Ptr{CResult}),
            ctx().h, b, CArgs(args; precision, stop), frequency, 2.0, ν, φ, ψ, res))
    finally
        destroy_batch(b)
    end
    return ν, φ, ψ, res
end
function batch_sinkhorn_cellsim(Xs, Ys, rs, cs, scale, metric, args, frequency; precision="bf16")
    b = batch(Xs, Ys, rs, cs, scale, metric)
    N, M = sum(size.(Xs, 1)), sum(size.(Ys, 1))
    φ, ψ = zeros(N), zeros(M)
    res = Vector{CResult}(undef, length(Xs))
    try
        autotune!(b, precision)
        check(ccall((:culamp_batch_sinkhorn_solve, libculamp), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ref{CArgs}, Int64, Float64, Ptr{Float64}, Ptr{Float64},
             Ptr{Float64}, Ptr{Float64}, Ptr{CResult}),
            ctx().h, b, CArgs(args; precision), frequency, 2.0, C_NULL, C_NULL, φ, ψ, res))
    finally
        destroy_batch(b)
    end
    return φ, ψ, res
end

# ---- Sinkhorn divergence and its gradient (ChainRulesCore rrule) ------------

function div_opts(; eps=0.0, eta_p=1e-2, tol=1e-6, polish_tol=1e-9, itermax=100_000,
                  frequency=10, unroll=64, checkpoint=16, precision="mixed", tmax=Inf)
    o = DivOpts(0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)
    ccall((:culamp_divergence_opts_default, libculamp), Cvoid, (Ref{DivOpts},), o)
    o.eps, o.eta_p, o.tol, o.polish_tol = eps, eta_p, tol, polish_tol
    o.itermax, o.frequency, o.unroll, o.checkpoint = itermax, frequency, unroll, checkpoint
    o.precision, o.tmax = PRECISION[precision], tmax
    o
end
function sinkhorn_divergence(x::Matrix{Float64}, a::Vector{Float64}, y::Matrix{Float64},
                             b::Vector{Float64}; o::DivOpts=div_opts())
    n, m = size(x, 2), size(y, 2)
    val = Ref{Float64}()
    gx, ga, gy, gb = zeros(3, n), zeros(n), zeros(3, m), zeros(m)
    check(ccall((:culamp_sinkhorn_divergence, libculamp), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64},
         Ref{DivOpts}, Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{CResult}),
        ctx().h, x, n, a, y, m, b, o, val, gx, ga, gy, gb, C_NULL))
    return val[], (gx, ga, gy, gb)
end
function ChainRulesCore.rrule(::typeof(sinkhorn_divergence), x, a, y, b; o::DivOpts=div_opts())
    S, (gx, ga, gy, gb) = sinkhorn_divergence(x, a, y, b; o)
    pullback(dS) = (NoTangent(), dS .* gx, dS .* ga, dS .* gy, dS .* gb)
    return S, pullback
end

end # module
