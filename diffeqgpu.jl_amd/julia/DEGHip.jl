# DEGHip.jl -- the Julia side of the drop-in boundary: routes DiffEqGPU's EnsembleGPUKernel path to
# libdeghip.so (the C ABI in include/deghip.h).
#
# The reference reaches the device through batch_solve_up_kernel (src/solve.jl:382-419):
# adapt(dev, probs) (:400) moves the ImmutableODEProblem vector to the backend, then
# vectorized_solve / vectorized_asolve (src/ensemblegpukernel/lowerlevel_solve.jl:53-131, :253-346)
# size the outputs and launch ode_solve_kernel / ode_asolve_kernel with prob.f compiled in
# (src/algorithms.jl:146-156). This module keeps the user-facing calls unchanged,
#
#     solve(monteprob, GPUTsit5(), EnsembleGPUKernel(HIPDevice()); trajectories, adaptive, dt, saveat)
#
# and specialises three things on a new backend object:
#   Adapt.adapt_storage(::HIPDevice, probs)   packs SoA u0 / p / tspans into device buffers
#   KernelAbstractions.get_backend(v::HIPProbVector) = v.dev
#   vectorized_solve / vectorized_asolve(::HIPProbVector, ...)
#                                             trace f(u, p, t) to HIP C (trace_rhs), compile it once
#                                             with deg_kernel_compile, launch deg_solve
# The host rules stay the reference's: saveat normalisation (normalized_saveat), output sizing,
# the t0 prefill of ts, Array(ts) / Array(us) in the reference's (nsave, N) layout.
#
# The tracer calls f once with a symbolic number type (Sym <: Real) that records every operation
# on a tape in evaluation order and hash-conses it; the tape is printed as straight-line Float32
# HIP C in the form diffeqgpu.jl_amd/tracer.py emits (same node order, names and statement forms),
# so both hosts hand the library the same program for the same f. Jacobians for the implicit
# methods are symbolic forward-mode derivatives of the tape (the role ForwardDiff plays in
# nlsolve/type.jl:129-157), or forward differences (alg_utils.jl:17-27) for autodiff = false.
#
# Not covered here (the Python host covers them): dual-number u0 / p, SDE problems, cpu_offload > 0.
# Float64 literals inside f are rounded to Float32 constants (the reference would promote that
# operation to Float64).
module DEGHip

using StaticArrays, Adapt, KernelAbstractions, SciMLBase, DiffEqGPU, AMDGPU, LinearAlgebra
import DiffEqGPU: vectorized_solve, vectorized_asolve

const libdeghip = joinpath(@__DIR__, "..", "_build", "libdeghip.so")

# ---- C ABI constants (include/deghip.h) -----------------------------------------------------
const DEG_TSIT5, DEG_VERN7, DEG_RB23, DEG_VERN9 = 0, 1, 2, 3
const DEG_RODAS4, DEG_RODAS5P, DEG_KVAERNO3, DEG_KVAERNO5 = 4, 5, 6, 7
const F_ADAPTIVE, F_PARITY, F_DETPOW, F_KAHAN, F_PERSISTENT = 0x0001, 0x0002, 0x0004, 0x0008, 0x0010
const F_FASTPOW, F_SAVEAT_GLOBAL, F_TSTOPS, F_CALLBACKS = 0x0040, 0x0100, 0x4000, 0x8000
const SAVEAT_LDS_CAP = 4096

# mirror of deg_solve_args; its size is checked against deg_solve_args_size() at first use
Base.@kwdef mutable struct SolveArgs
    struct_size::UInt32 = 280
    n::Int32 = 0
    n_traj::Int64 = 0
    stride::Int64 = 0
    id_offset::Int64 = 0
    u0::Ptr{Float32} = C_NULL
    p::Ptr{Float32} = C_NULL
    t0s::Ptr{Float32} = C_NULL
    tfs::Ptr{Float32} = C_NULL
    t0::Float32 = 0
    tf::Float32 = 0
    dt::Float32 = 0
    abstol::Float32 = 0
    reltol::Float32 = 0
    saveat::Ptr{Float32} = C_NULL
    nsaveat::Int32 = 0
    nsave::Int32 = 0
    save_everystep::Int32 = 0
    block_size::Int32 = 0
    ts::Ptr{Float32} = C_NULL
    us::Ptr{Float32} = C_NULL
    retcode::Ptr{Int32} = C_NULL
    naccept::Ptr{Int32} = C_NULL
    nreject::Ptr{Int32} = C_NULL
    max_attempts::Int64 = 0
    stream::Ptr{Cvoid} = C_NULL
    grid_blocks::Int32 = 0
    refill_batch::Int32 = 0
    elapsed_ms::Ptr{Float32} = C_NULL
    mlp::Ptr{Cvoid} = C_NULL
    steplog::Ptr{Float32} = C_NULL
    steplog_count::Ptr{UInt32} = C_NULL
    steplog_cap::Int64 = 0
    chunk::Int32 = 0
    mlp_every::Int32 = 0
    order::Ptr{Int32} = C_NULL
    tstops::Ptr{Float32} = C_NULL
    ntstops::Int32 = 0
    reserved0::Int32 = 0
    seeds::Ptr{UInt64} = C_NULL
    seed::UInt64 = 0
    traj_base::Int64 = 0
    inputs_ordered::Ptr{Float32} = C_NULL
end

last_error() = unsafe_string(ccall((:deg_last_error, libdeghip), Cstring, ()))
check(rc::Integer, what) = rc == 0 || error("deghip: $what failed ($rc): ", last_error())

function check_layout()
    sz = ccall((:deg_solve_args_size, libdeghip), UInt32, ())
    sz == sizeof(SolveArgs) || error("deghip: deg_solve_args is $sz bytes in the library, $(sizeof(SolveArgs)) here")
    return nothing
end

# ---- backend object, context and compiled-kernel cache --------------------------------------
"""    HIPDevice(; device = 0, kahan = false, parity = false)

The backend of `EnsembleGPUKernel(HIPDevice())`. `parity = true` compiles the reference's unfused
Float32 sequence (and the deterministic controller power for adaptive solves)."""
struct HIPDevice
    device::Int
    kahan::Bool
    parity::Bool
end
HIPDevice(; device = 0, kahan = false, parity = false) = HIPDevice(device, kahan, parity)

const CTXS = Dict{Int, Ptr{Cvoid}}()
const KERNELS = Dict{Tuple{String, Int, Int, Int, UInt32}, Ptr{Cvoid}}()
const LOCK = ReentrantLock()

function context(dev::HIPDevice)
    lock(LOCK) do
        get!(CTXS, dev.device) do
            check_layout()
            c = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:deg_ctx_create, libdeghip), Cint, (Cint, Ptr{Ptr{Cvoid}}), dev.device, c), "deg_ctx_create")
            c[]
        end
    end
end

function kernel(dev::HIPDevice, src::String, n::Int, np::Int, method::Int, flags::UInt32)
    ctx = context(dev)
    lock(LOCK) do
        get!(KERNELS, (src, n, np, method, flags)) do
            k = Ref{Ptr{Cvoid}}(C_NULL)
            check(ccall((:deg_kernel_compile, libdeghip), Cint,
                        (Ptr{Cvoid}, Cstring, Cint, Cint, Cint, UInt32, Ptr{Ptr{Cvoid}}),
                        ctx, src, n, np, method, flags, k), "deg_kernel_compile")
            k[]
        end
    end
end

method_id(::GPUTsit5) = DEG_TSIT5
method_id(::GPUVern7) = DEG_VERN7
method_id(::GPURosenbrock23) = DEG_RB23
method_id(::GPUVern9) = DEG_VERN9
method_id(::GPURodas4) = DEG_RODAS4
method_id(::GPURodas5P) = DEG_RODAS5P
method_id(::GPUKvaerno3) = DEG_KVAERNO3
method_id(::GPUKvaerno5) = DEG_KVAERNO5
implicit(alg) = alg isa DiffEqGPU.GPUODEImplicitAlgorithm
autodiff(alg) = implicit(alg) ? DiffEqGPU.alg_autodiff(alg) : true

# ---- SoA problem vector: what adapt(HIPDevice(), probs) returns ------------------------------
struct HIPProbVector{F, P}
    dev::HIPDevice            # the backend of EnsembleGPUKernel(dev), carried to vectorized_*solve
    f::F                      # the problems' right-hand side (shared, as kernels.jl assumes)
    mass_matrix::Any
    n::Int
    np::Int
    N::Int
    u0::ROCMatrix{Float32}    # (N, n): trajectory fastest in memory, i.e. SoA [state][traj]
    p::ROCMatrix{Float32}     # (N, np)
    t0s::ROCVector{Float32}
    tfs::ROCVector{Float32}
    cpu::P                    # the host problems (batch_solve trims and builds solutions from them)
end

function Adapt.adapt_storage(dev::HIPDevice, probs::AbstractVector{<:SciMLBase.AbstractODEProblem})
    N = length(probs)
    n = length(probs[1].u0)
    np = probs[1].p === nothing || probs[1].p isa SciMLBase.NullParameters ? 0 : length(probs[1].p)
    u0 = Matrix{Float32}(undef, N, n)
    p = Matrix{Float32}(undef, N, max(np, 1))
    for i in 1:N
        eltype(probs[i].u0) <: Real && !(eltype(probs[i].u0) <: AbstractFloat) &&
            error("DEGHip: dual-number or symbolic u0 is not supported by this shim")
        for c in 1:n
            u0[i, c] = Float32(probs[i].u0[c])
        end
        for j in 1:np
            p[i, j] = Float32(probs[i].p[j])
        end
    end
    t0s = Float32[pr.tspan[1] for pr in probs]
    tfs = Float32[pr.tspan[2] for pr in probs]
    return HIPProbVector(dev, probs[1].f, probs[1].f.mass_matrix, n, np, N, ROCMatrix(u0), ROCMatrix(p),
                         ROCVector(t0s), ROCVector(tfs), probs)
end
KernelAbstractions.get_backend(v::HIPProbVector) = v.dev
Base.length(v::HIPProbVector) = v.N

# ---- saveat normalisation (lowerlevel_solve.jl:89-102 fixed step, :271-306 adaptive) --------
"""    normalized_saveat(saveat, t0, tf; adaptive) -> Vector{Float32} or nothing

A number s becomes range(t0, tf, length = ceil(|tf - t0| / |s|) + 1) (s == 0: [t0, tf]; adaptive
solves refuse more than 100,000 points); a range is re-made from its first and last entry and its
length in Float32; a vector is converted elementwise."""
function normalized_saveat(saveat, t0::Float32, tf::Float32; adaptive::Bool)
    saveat === nothing && return nothing
    if saveat isa Number
        Float32(saveat) == 0.0f0 && return Float32[t0, tf]
        num_points = Int(ceil(abs(tf - t0) / abs(Float32(saveat)))) + 1
        if adaptive && num_points > 100_000
            error("saveat would create too many save points ($num_points). Consider using a larger saveat value.")
        end
        return Float32.(collect(range(t0, tf, length = num_points)))
    elseif saveat isa AbstractRange && !adaptive
        return Float32.(collect(range(Float32(first(saveat)), Float32(last(saveat)), length = length(saveat))))
    else
        return Float32.(collect(saveat))
    end
end

# ---- outputs: SoA on the device, the reference's layout on Array() --------------------------
struct SoAOutput{A}
    data::A      # ts: (N, nsave) = [save][traj]; us: (N, n, nsave) = [save][state][traj]
    nsave::Int
end
function Base.Array(o::SoAOutput{<:ROCArray{Float32, 3}})
    h = Array(o.data)
    N, n, ns = size(h)
    return [SVector{n, Float32}(ntuple(c -> h[i, c, s], n)) for s in 1:ns, i in 1:N]
end
Base.Array(o::SoAOutput{<:ROCMatrix{Float32}}) = permutedims(Array(o.data))   # (nsave, N)

# ===========================================================================================
# Tracing f(u, p, t) to HIP C
# ===========================================================================================
struct Node
    op::Symbol          # :const :u :p :t :+ :- :* :/ :neg :sin ... :pow :select :bselect :sbsel
    args::Vector{Int}   # 1-based ids of the operands
    val::Float64        # constant value, or the input index of :u / :p
    cmp::Symbol         # the comparison of :cmp and :select nodes, else :none
end

mutable struct Tape
    nodes::Vector{Node}
    memo::Dict{Tuple{Symbol, Vector{Int}, Float64, Symbol}, Int}
end
Tape() = Tape(Node[], Dict{Tuple{Symbol, Vector{Int}, Float64, Symbol}, Int}())

const TAPE = Ref{Union{Nothing, Tape}}(nothing)

"A traced Float32 value."
struct Sym <: Real
    id::Int
end
"A traced comparison (or a combination of comparisons)."
struct SymBool
    id::Int
end

function mk(op::Symbol, args::Vector{Int}, val = 0.0, cmp::Symbol = :none)
    tp = TAPE[]
    tp === nothing && error("DEGHip: a traced value was used outside its trace")
    key = (op, args, Float64(val), cmp)
    return get!(tp.memo, key) do
        push!(tp.nodes, Node(op, args, Float64(val), cmp))
        length(tp.nodes)
    end
end

lift(x::Sym) = x
lift(::Bool) = error("DEGHip: booleans are not arithmetic values")
lift(x::Real) = Sym(mk(:const, Int[], Float64(x)))
liftb(x::SymBool) = x
liftb(x::Bool) = SymBool(mk(:bconst, Int[], x ? 1.0 : 0.0))

Base.promote_rule(::Type{Sym}, ::Type{<:Real}) = Sym
Base.convert(::Type{Sym}, x::Real) = lift(x)
Base.convert(::Type{Sym}, x::Sym) = x
Base.zero(::Type{Sym}) = lift(0.0)
Base.one(::Type{Sym}) = lift(1.0)

Base.:+(a::Sym, b::Sym) = Sym(mk(:+, [a.id, b.id]))
Base.:-(a::Sym, b::Sym) = Sym(mk(:-, [a.id, b.id]))
Base.:*(a::Sym, b::Sym) = Sym(mk(:*, [a.id, b.id]))
Base.:/(a::Sym, b::Sym) = Sym(mk(:/, [a.id, b.id]))
Base.:-(a::Sym) = Sym(mk(:neg, [a.id]))
Base.:+(a::Sym) = a

# x^2 and x^3 with literal exponents are repeated products in Julia (Base.literal_pow)
Base.literal_pow(::typeof(^), x::Sym, ::Val{0}) = lift(1.0)
Base.literal_pow(::typeof(^), x::Sym, ::Val{1}) = x
Base.literal_pow(::typeof(^), x::Sym, ::Val{2}) = x * x
Base.literal_pow(::typeof(^), x::Sym, ::Val{3}) = (x * x) * x
Base.literal_pow(::typeof(^), x::Sym, ::Val{-1}) = lift(1.0) / x
Base.literal_pow(::typeof(^), x::Sym, ::Val{p}) where {p} = x^lift(p)
Base.:^(x::Sym, y::Sym) = Sym(mk(:pow, [x.id, y.id]))
Base.:^(x::Sym, y::Integer) = x^lift(y)
Base.:^(x::Sym, y::AbstractFloat) = x^lift(y)
Base.:^(x::Sym, y::Rational) = x^lift(y)
Base.:^(x::Integer, y::Sym) = lift(x)^y
Base.:^(x::AbstractFloat, y::Sym) = lift(x)^y

for f in (:sin, :cos, :exp, :log, :sqrt, :tanh, :abs)
    @eval Base.$f(x::Sym) = Sym(mk($(QuoteNode(f)), [x.id]))
end

for (f, c) in ((:<, "<"), (:<=, "<="), (:>, ">"), (:>=, ">="), (:(==), "=="), (:!=, "!="))
    @eval begin
        Base.$f(a::Sym, b::Sym) = SymBool(mk(:cmp, [a.id, b.id], 0.0, Symbol($c)))
        Base.$f(a::Sym, b::Real) = $f(a, lift(b))
        Base.$f(a::Real, b::Sym) = $f(lift(a), b)
    end
end
Base.:&(a::SymBool, b::SymBool) = SymBool(mk(:and, [a.id, b.id]))
Base.:&(a::SymBool, b::Bool) = a & liftb(b)
Base.:&(a::Bool, b::SymBool) = liftb(a) & b
Base.:|(a::SymBool, b::SymBool) = SymBool(mk(:or, [a.id, b.id]))
Base.:|(a::SymBool, b::Bool) = a | liftb(b)
Base.:|(a::Bool, b::SymBool) = liftb(a) | b
Base.:!(a::SymBool) = SymBool(mk(:not, [a.id]))

# ifelse on a single comparison is a select on that comparison (both branches are evaluated)
function Base.ifelse(c::SymBool, a, b)
    nd = TAPE[].nodes[c.id]
    x, y = lift(a), lift(b)
    nd.op === :cmp && return Sym(mk(:select, [nd.args[1], nd.args[2], x.id, y.id], 0.0, nd.cmp))
    return Sym(mk(:bselect, [c.id, x.id, y.id]))
end

function with_tape(fn)
    old = TAPE[]
    tp = Tape()
    TAPE[] = tp
    try
        return fn(tp)
    finally
        TAPE[] = old
    end
end

function new_inputs(n::Int, np::Int)
    u = SVector{n, Sym}(ntuple(i -> Sym(mk(:u, Int[], i - 1)), n))
    p = SVector{np, Sym}(ntuple(j -> Sym(mk(:p, Int[], j - 1)), np))
    t = Sym(mk(:t, Int[]))
    return u, p, t
end

# -- symbolic forward-mode differentiation (zeros are `nothing`; the rules of tracer._Diff) --
struct Diff
    var::Int
    memo::Dict{Int, Union{Nothing, Sym}}
end
Diff(var::Int) = Diff(var, Dict{Int, Union{Nothing, Sym}}())

dadd(a, b) = a === nothing ? b : b === nothing ? a : a + b
dsub(a, b) = b === nothing ? a : a === nothing ? -b : a - b
isone_const(x::Sym) = (nd = TAPE[].nodes[x.id]; nd.op === :const && nd.val == 1.0)
function dmul(a, b)
    (a === nothing || b === nothing) && return nothing
    isone_const(a) && return b
    isone_const(b) && return a
    return a * b
end

function d(D::Diff, x::Int)
    haskey(D.memo, x) && return D.memo[x]
    r = d_(D, x)
    D.memo[x] = r
    return r
end

function d_(D::Diff, x::Int)
    x == D.var && return lift(1.0)
    nd = TAPE[].nodes[x]
    op = nd.op
    op in (:const, :u, :p, :t) && return nothing
    a = [Sym(i) for i in nd.args]
    X = Sym(x)
    if op === :+
        return dadd(d(D, a[1].id), d(D, a[2].id))
    elseif op === :-
        return dsub(d(D, a[1].id), d(D, a[2].id))
    elseif op === :*
        return dadd(dmul(d(D, a[1].id), a[2]), dmul(a[1], d(D, a[2].id)))
    elseif op === :/
        da, db = d(D, a[1].id), d(D, a[2].id)
        t1 = da === nothing ? nothing : da / a[2]
        t2 = db === nothing ? nothing : (X * db) / a[2]
        return dsub(t1, t2)
    elseif op === :neg
        da = d(D, a[1].id)
        return da === nothing ? nothing : -da
    elseif op === :sin
        return dmul(Sym(mk(:cos, [a[1].id])), d(D, a[1].id))
    elseif op === :cos
        da = d(D, a[1].id)
        return da === nothing ? nothing : -(Sym(mk(:sin, [a[1].id])) * da)
    elseif op === :exp
        return dmul(X, d(D, a[1].id))
    elseif op === :log
        da = d(D, a[1].id)
        return da === nothing ? nothing : da / a[1]
    elseif op === :sqrt
        da = d(D, a[1].id)
        return da === nothing ? nothing : da / (lift(2.0) * X)
    elseif op === :tanh
        da = d(D, a[1].id)
        return da === nothing ? nothing : (lift(1.0) - X * X) * da
    elseif op === :abs
        # DiffRules: abs'(x) = signbit(x) ? -1 : 1
        da = d(D, a[1].id)
        da === nothing && return nothing
        return Sym(mk(:sbsel, [a[1].id, (-da).id, da.id]))
    elseif op === :pow
        # ForwardDiff's ^ with its guards: no base term for a zero exponent, no log term where the
        # base is 0 and the exponent positive
        base, ex = a
        db, de = d(D, base.id), d(D, ex.id)
        zero_ = lift(0.0)
        r = nothing
        if db !== nothing
            bt = (ex * (base^(ex - lift(1.0)))) * db
            if de === nothing
                bt = Sym(mk(:bselect, [(ex == zero_).id, zero_.id, bt.id]))
            end
            r = bt
        end
        if de !== nothing
            guard = (base == zero_) & (ex > zero_)
            r = dadd(r, Sym(mk(:bselect, [guard.id, zero_.id, (X * log(base)).id])) * de)
        end
        return r
    elseif op === :select
        da, db = d(D, nd.args[3]), d(D, nd.args[4])
        (da === nothing && db === nothing) && return nothing
        return Sym(mk(:select, [nd.args[1], nd.args[2], (da === nothing ? lift(0.0) : da).id,
                                (db === nothing ? lift(0.0) : db).id], 0.0, nd.cmp))
    end
    error("DEGHip: cannot differentiate $op")
end

# -- emission (the statement forms of tracer._emit_values) --
const FN = Dict(:sin => "sinf", :cos => "cosf", :exp => "expf", :log => "logf", :sqrt => "__builtin_sqrtf",
                :tanh => "tanhf", :abs => "__builtin_fabsf")

clit(v::Float64) = (v == trunc(v) && abs(v) < 2.0^24) ? "$(Int(v)).0f" : "(float)$(repr(v))"

function emit_values(tp::Tape, outs::Vector{Int}, indent = "    ")
    need = falses(length(tp.nodes))
    stack = copy(outs)
    while !isempty(stack)
        x = pop!(stack)
        need[x] && continue
        need[x] = true
        append!(stack, tp.nodes[x].args)
    end
    names = Dict{Int, String}()
    lines = String[]
    flops = 0
    for (id, nd) in enumerate(tp.nodes)
        need[id] || continue
        op = nd.op
        v = "v$(id - 1)"
        if op === :const
            names[id] = clit(nd.val)
            continue
        elseif op === :u
            names[id] = "u[$(Int(nd.val))]"
            continue
        elseif op === :p
            names[id] = "p[$(Int(nd.val))]"
            continue
        elseif op === :t
            names[id] = "t"
            continue
        elseif op === :bconst
            names[id] = nd.val != 0 ? "true" : "false"
            continue
        end
        a = [names[y] for y in nd.args]
        if op in (:cmp, :and, :or, :not)
            expr = op === :cmp ? "($(a[1]) $(nd.cmp) $(a[2]))" :
                   op === :and ? "($(a[1]) && $(a[2]))" :
                   op === :or ? "($(a[1]) || $(a[2]))" : "!($(a[1]))"
            push!(lines, "$(indent)const bool $v = $expr;")
            names[id] = v
            continue
        end
        if op in (:+, :-, :*, :/)
            expr = "$(a[1]) $op $(a[2])"
            flops += 1
        elseif op === :neg
            expr = "-($(a[1]))"
        elseif haskey(FN, op)
            expr = "$(FN[op])($(a[1]))"
            op === :sqrt && (flops += 1)
        elseif op === :pow
            expr = "powf($(a[1]), $(a[2]))"
        elseif op === :select
            expr = "($(a[1]) $(nd.cmp) $(a[2])) ? $(a[3]) : $(a[4])"
        elseif op === :bselect
            expr = "$(a[1]) ? $(a[2]) : $(a[3])"
        elseif op === :sbsel
            expr = "__builtin_signbit($(a[1])) ? $(a[2]) : $(a[3])"
        else
            error("DEGHip: unknown op $op")
        end
        push!(lines, "$(indent)const float $v = $expr;")
        names[id] = v
    end
    return lines, names, flops
end

function emit_body(tp::Tape, outs::Vector{Int}, out_name::String)
    lines, names, flops = emit_values(tp, outs)
    for (i, o) in enumerate(outs)
        push!(lines, "    $(out_name)[$(i - 1)] = $(names[o]);")
    end
    return lines, flops
end

const SIG_RHS = ("__device__ __forceinline__ void deg_rhs(const float* __restrict__ u, const float* __restrict__ p, float t,",
                 "                                        float* __restrict__ du) {")
const SIG_JAC = ("__device__ __forceinline__ void deg_jac(const float* __restrict__ u, const float* __restrict__ p, float t,",
                 "                                        float* __restrict__ J) {")
const SIG_TGRAD = ("__device__ __forceinline__ void deg_tgrad(const float* __restrict__ u, const float* __restrict__ p, float t,",
                   "                                          float* __restrict__ dT) {")
# sqrt(eps(Float32)) as Float32: the forward-difference step of alg_utils.jl:18 and nlsolve/type.jl:152
const FD_STEP_BITS = "0x39b504f3"

fd_jac_lines(n) = [SIG_JAC...,
    "    const float dx = __builtin_bit_cast(float, $(FD_STEP_BITS)u);",
    "    float f0[$n], f1[$n], ux[$n];",
    "    deg_rhs(u, p, t, f0);",
    "    _Pragma(\"unroll\") for (int i = 0; i < $n; ++i) {",
    "        _Pragma(\"unroll\") for (int r = 0; r < $n; ++r) ux[r] = u[r];",
    "        ux[i] = u[i] + dx;",
    "        deg_rhs(ux, p, t, f1);",
    "        _Pragma(\"unroll\") for (int r = 0; r < $n; ++r) J[r * $n + i] = (f1[r] - f0[r]) / dx;",
    "    }",
    "}"]

fd_tgrad_lines(n) = [SIG_TGRAD...,
    "    const float dx = __builtin_bit_cast(float, $(FD_STEP_BITS)u);",
    "    float f0[$n], f1[$n];",
    "    deg_rhs(u, p, t, f0);",
    "    deg_rhs(u, p, t + dx, f1);",
    "    _Pragma(\"unroll\") for (int r = 0; r < $n; ++r) dT[r] = (f1[r] - f0[r]) / dx;",
    "}"]

outs_of(tp, out, nout) = begin
    ids = [lift(x).id for x in out]
    length(ids) == nout || error("DEGHip: expected $nout outputs, got $(length(ids))")
    ids
end

"""    trace_rhs(f, n, np; jac = false, autodiff = true) -> (source, flops)

HIP C defining deg_rhs for the out-of-place f(u, p, t); with `jac`, also deg_jac (row-major
J[r*n + c] = d du_r / d u_c) and deg_tgrad, taken from f.jac / f.tgrad when the ODEFunction has
them, else by symbolic forward mode (autodiff = true) or forward differences (autodiff = false),
the order build_J_W / build_tgrad use (nlsolve/type.jl:129-157)."""
function trace_rhs(f, n::Int, np::Int; jac::Bool = false, autodiff::Bool = true)
    fun = f isa SciMLBase.AbstractODEFunction ? f.f : f
    with_tape() do tp
        u, p, t = new_inputs(n, np)
        outs = outs_of(tp, fun(u, p, t), n)
        body, flops = emit_body(tp, outs, "du")
        src = String[SIG_RHS..., "    (void)u; (void)p; (void)t;", body..., "}"]
        if jac
            has_jac = f isa SciMLBase.AbstractODEFunction && SciMLBase.has_jac(f)
            has_tgrad = f isa SciMLBase.AbstractODEFunction && SciMLBase.has_tgrad(f)
            if has_jac
                J = f.jac(u, p, t)
                jbody, _ = emit_body(tp, [lift(J[r, c]).id for r in 1:n for c in 1:n], "J")
                append!(src, [SIG_JAC..., "    (void)u; (void)p; (void)t;", jbody..., "}"])
            elseif autodiff
                Jn = [something(d(Diff(u[c].id), o), lift(0.0)).id for o in outs for c in 1:n]
                jbody, _ = emit_body(tp, Jn, "J")
                append!(src, [SIG_JAC..., "    (void)u; (void)p; (void)t;", jbody..., "}"])
            else
                append!(src, fd_jac_lines(n))
            end
            if has_tgrad
                tbody, _ = emit_body(tp, [lift(x).id for x in f.tgrad(u, p, t)], "dT")
                append!(src, [SIG_TGRAD..., "    (void)u; (void)p; (void)t;", tbody..., "}"])
            elseif autodiff
                Tn = [something(d(Diff(t.id), o), lift(0.0)).id for o in outs]
                tbody, _ = emit_body(tp, Tn, "dT")
                append!(src, [SIG_TGRAD..., "    (void)u; (void)p; (void)t;", tbody..., "}"])
            else
                append!(src, fd_tgrad_lines(n))
            end
        end
        return join(src, "\n") * "\n", flops
    end
end

"""    emit_mass(M, n) -> source

A constant mass matrix (ODEFunction(f; mass_matrix = M)) as the DEG_MASS defines of
deg_kernel_compile; the identity needs none."""
function emit_mass(M, n::Int)
    M isa LinearAlgebra.UniformScaling && M.λ == 1 && return ""
    entries = join([clit(Float64(Float32(M[r, c]))) for r in 1:n for c in 1:n], ", ")
    return "#define DEG_MASS 1\n#define DEG_MASS_ENTRIES $entries\n"
end

# ---- callbacks (callbacks.jl:1-124; applied by the kernel as integrator_utils.jl:274-479) ----
"What a traced condition or affect! receives as `integrator`."
mutable struct SymIntegrator
    u::Any
    p::Any
    t::Sym
    flags::Int
end
SciMLBase.terminate!(i::SymIntegrator, retcode = SciMLBase.ReturnCode.Terminated) = (i.flags |= 1; nothing)
SciMLBase.u_modified!(i::SymIntegrator, b::Bool) = (i.flags = b ? (i.flags & ~2) : (i.flags | 2); nothing)

function trace_condition(fn, n, np, boolean::Bool)
    with_tape() do tp
        u, p, t = new_inputs(n, np)
        out = fn(u, t, SymIntegrator(u, p, t, 0))
        id = boolean ? liftb(out).id : lift(out).id
        lines, names, _ = emit_values(tp, [id], "        ")
        return [lines..., "        return $(names[id]);"]
    end
end

function trace_affect(fn, n, np)
    with_tape() do tp
        u, p, t = new_inputs(n, np)
        integ = SymIntegrator(u, p, t, 0)
        fn(integ)
        outs, targets = Int[], String[]
        for i in 1:n
            x = lift(integ.u[i])
            x.id == u[i].id || (push!(outs, x.id); push!(targets, "u[$(i - 1)]"))
        end
        for j in 1:np
            x = lift(integ.p[j])
            x.id == p[j].id || (push!(outs, x.id); push!(targets, "p[$(j - 1)]"))
        end
        lines, names, _ = emit_values(tp, outs, "        ")
        # every value is computed before any entry is overwritten
        append!(lines, ["        $tg = $(names[o]);" for (tg, o) in zip(targets, outs)])
        return [lines..., "        return $(integ.flags);"]
    end
end

function switch(sig, cases, default)
    out = ["__device__ __forceinline__ $sig {", "    (void)j;", "    switch (j) {"]
    for (j, body) in enumerate(cases)
        append!(out, ["    case $(j - 1): {", body..., "    }"])
    end
    append!(out, ["    default: break;", "    }", "    return $default;", "}"])
    return out
end

to_gpu(cb::DiffEqGPU.GPUDiscreteCallback) = cb
to_gpu(cb::DiffEqGPU.GPUContinuousCallback) = cb
to_gpu(cb::SciMLBase.DiscreteCallback) = convert(DiffEqGPU.GPUDiscreteCallback, cb)
to_gpu(cb::SciMLBase.ContinuousCallback) = convert(DiffEqGPU.GPUContinuousCallback, cb)

const ROOTFIND_CODE = Dict(SciMLBase.NoRootFind => 0, SciMLBase.LeftRootFind => 1, SciMLBase.RightRootFind => 2)

"""    trace_callbacks(cbs::CallbackSet, n, np) -> source

DEG_NDCB / DEG_NCCB and the deg_dcb_* / deg_ccb_* functions of deg_kernel_compile for the
set's discrete and continuous callbacks, each in set order."""
function trace_callbacks(cbs::SciMLBase.CallbackSet, n::Int, np::Int)
    dcs = [to_gpu(c) for c in cbs.discrete_callbacks]
    ccs = [to_gpu(c) for c in cbs.continuous_callbacks]
    head = "        (void)u; (void)p; (void)t;"
    src = ["// ---- traced callbacks ----", "#define DEG_NDCB $(length(dcs))", "#define DEG_NCCB $(length(ccs))"]
    if !isempty(dcs)
        append!(src, switch("bool deg_dcb_cond(int j, const float* __restrict__ u, const float* __restrict__ p, float t)",
                            [[head; trace_condition(c.condition, n, np, true)] for c in dcs], "false"))
        append!(src, switch("int deg_dcb_affect(int j, float* __restrict__ u, float* __restrict__ p, float t)",
                            [[head; trace_affect(c.affect!, n, np)] for c in dcs], "0"))
    end
    if !isempty(ccs)
        append!(src, switch("float deg_ccb_cond(int j, const float* __restrict__ u, const float* __restrict__ p, float t)",
                            [[head; trace_condition(c.condition, n, np, false)] for c in ccs], "0.0f"))
        cases = Vector{Vector{String}}()
        for c in ccs
            pos = c.affect! === nothing ? ["        return 0;"] : trace_affect(c.affect!, n, np)
            neg = c.affect_neg! === nothing ? ["        return 0;"] : trace_affect(c.affect_neg!, n, np)
            push!(cases, [head; "        if (!neg) {"; ["    " * x for x in pos]; "        }"; neg])
        end
        append!(src, switch("int deg_ccb_affect(int j, int neg, float* __restrict__ u, float* __restrict__ p, float t)",
                            cases, "0"))
        mode(c) = (c.affect! === nothing ? 0 : 1) | (c.affect_neg! === nothing ? 0 : 2) | (ROOTFIND_CODE[c.rootfind] << 2)
        f32(x) = clit(Float64(Float32(x)))
        append!(src, switch("int deg_ccb_mode(int j)", [["        return $(mode(c));"] for c in ccs], "0"))
        append!(src, switch("float deg_ccb_abstol(int j)", [["        return $(f32(c.abstol));"] for c in ccs], "0.0f"))
        append!(src, switch("float deg_ccb_nudge(int j)", [["        return $(f32(c.repeat_nudge));"] for c in ccs], "0.0f"))
        append!(src, switch("float deg_ccb_dtrelax(int j)", [["        return $(f32(c.dtrelax));"] for c in ccs], "1.0f"))
    end
    return join(src, "\n") * "\n"
end

# ===========================================================================================
# vectorized_solve / vectorized_asolve on the HIP backend
# ===========================================================================================
function kernel_source(probs::HIPProbVector, alg, callback)
    src, _ = trace_rhs(probs.f, probs.n, probs.np; jac = implicit(alg), autodiff = autodiff(alg))
    implicit(alg) && (src = emit_mass(probs.mass_matrix, probs.n) * src)
    cbs = callback isa SciMLBase.CallbackSet ? callback : SciMLBase.CallbackSet(callback)
    has_cb = !isempty(cbs.discrete_callbacks) || !isempty(cbs.continuous_callbacks)
    has_cb && (src *= trace_callbacks(cbs, probs.n, probs.np))
    return src, has_cb
end

function launch!(probs::HIPProbVector, alg, src, flags, t0, tf, dt, abstol, reltol, sv, nsave, everystep, tstops)
    dsv = sv === nothing ? nothing : ROCVector{Float32}(sv)
    dsv !== nothing && length(dsv) > SAVEAT_LDS_CAP && (flags |= F_SAVEAT_GLOBAL)
    dts = tstops === nothing || isempty(tstops) ? nothing : ROCVector{Float32}(Float32.(collect(tstops)))
    dts === nothing || (flags |= F_TSTOPS)
    k = kernel(probs.dev, src, probs.n, probs.np, method_id(alg), UInt32(flags))
    ts = ROCMatrix{Float32}(undef, probs.N, nsave)
    fill!(ts, t0)                                           # lowerlevel_solve.jl:81, :317
    us = ROCArray{Float32}(undef, probs.N, probs.n, nsave)
    retcode = ROCVector{Int32}(undef, probs.N)
    a = SolveArgs(n = probs.n, n_traj = probs.N, stride = probs.N,
                  u0 = pointer(probs.u0), p = probs.np > 0 ? pointer(probs.p) : C_NULL,
                  t0s = pointer(probs.t0s), tfs = pointer(probs.tfs), t0 = t0, tf = tf, dt = Float32(dt),
                  abstol = Float32(abstol), reltol = Float32(reltol),
                  saveat = dsv === nothing ? C_NULL : pointer(dsv), nsaveat = dsv === nothing ? 0 : length(dsv),
                  nsave = nsave, save_everystep = Int32(everystep), ts = pointer(ts), us = pointer(us),
                  retcode = pointer(retcode), tstops = dts === nothing ? C_NULL : pointer(dts),
                  ntstops = dts === nothing ? 0 : length(dts), stream = AMDGPU.stream().stream)
    GC.@preserve probs ts us retcode dsv dts begin
        check(ccall((:deg_solve, libdeghip), Cint, (Ptr{Cvoid}, Ref{SolveArgs}), k, a), "deg_solve")
    end
    # Array(ts) / Array(us) synchronise on the task-local stream and give the reference's layout
    return SoAOutput(ts, nsave), SoAOutput(us, nsave)
end

"""The fixed-step solve of lowerlevel_solve.jl:53-131 on the HIP backend."""
function vectorized_solve(probs::HIPProbVector, prob::ODEProblem, alg::DiffEqGPU.GPUODEAlgorithm;
                          dt, saveat = nothing, save_everystep = true, debug = false,
                          callback = SciMLBase.CallbackSet(nothing), tstops = nothing, kwargs...)
    t0, tf = Float32.(prob.tspan)
    dt32 = Float32(dt)
    sv = normalized_saveat(saveat, t0, tf; adaptive = false)
    if sv === nothing
        timeseries = t0:dt32:tf
        nsave = save_everystep ? length(timeseries) : 2
        if save_everystep && tstops !== nothing
            nsave += length(tstops) - count(x -> x in tstops, timeseries)   # lowerlevel_solve.jl:73-77
        end
    else
        nsave = length(sv)
    end
    src, has_cb = kernel_source(probs, alg, callback)
    flags = (has_cb ? F_CALLBACKS : 0x0000) | (probs.dev.kahan ? F_KAHAN : 0x0000) |
            (probs.dev.parity ? F_PARITY : 0x0000)
    return launch!(probs, alg, src, flags, t0, tf, dt32, 0.0f0, 0.0f0, sv, nsave,
                   save_everystep && sv === nothing, tstops)
end

"""The adaptive solve of lowerlevel_solve.jl:253-346 on the HIP backend (persistent grid, PI controller)."""
function vectorized_asolve(probs::HIPProbVector, prob::ODEProblem, alg::DiffEqGPU.GPUODEAlgorithm;
                           dt = 0.1f0, saveat = nothing, save_everystep = false, abstol = 1.0f-6, reltol = 1.0f-3,
                           debug = false, callback = SciMLBase.CallbackSet(nothing), tstops = nothing, kwargs...)
    t0, tf = Float32.(prob.tspan)
    sv = normalized_saveat(saveat, t0, tf; adaptive = true)
    nsave = sv === nothing ? (save_everystep ? ceil(Int, (tf - t0) / Float32(dt)) + 1 : 2) : length(sv)
    src, has_cb = kernel_source(probs, alg, callback)
    dev = probs.dev
    flags = F_ADAPTIVE | F_PERSISTENT | (dev.parity ? (F_PARITY | F_DETPOW) : F_FASTPOW) |
            (has_cb ? F_CALLBACKS : 0x0000) | (dev.kahan ? F_KAHAN : 0x0000)
    return launch!(probs, alg, src, flags, t0, tf, Float32(dt), abstol, reltol, sv, nsave, false, tstops)
end

end # module
