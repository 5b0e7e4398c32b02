# NeuralQuantumHIP.jl -- Julia glue that plugs libnqhip (include/nqhip.h) into
# NeuralQuantum.jl as a drop-in backend for the steady-state VMC loop.
#
# Not executed in this repository (no Julia toolchain here); the same C ABI is
# exercised from Python through ctypes by the tests (tests/test_gpu_vmc_api.py
# replays the loop below call for call through the entry points these methods
# use).  The methods extend the reference's own generic functions, so the
# driver loop of examples/dissipative_ising1d.jl:33-55 runs unchanged once
# `BatchedSampler` is taken from this module:
#
#     using NeuralQuantum, NeuralQuantumHIP
#     is = NeuralQuantumHIP.BatchedSampler(net, sampl, liouv, algo; batch_sz=16)
#     add_observable!(is, "Sx", Sx)
#     for i = 1:200
#         ldata, prec = sample!(is)
#         ob = compute_observables(is)
#         grad = precondition!(prec, algo, i)
#         Optimisers.update!(optimizer, net, grad)
#     end
#
# Networks: NDM (af_softplus or af_logcosh, NDM.jl:50), RBMSplit (density
# matrices) with a KLocalLiouvillian; the ket RBM with a Hamiltonian
# (BatchedValSampler, build_Batched.jl:9-10).  Parameters live in the Julia
# network: they are uploaded at every sample! (17 KB at the 4x4 benchmark
# shape) and read back by update!, so the network stays the canonical copy.
module NeuralQuantumHIP

using NeuralQuantum
import NeuralQuantum: sample!, precondition!, add_observable!, compute_observables,
                      AbstractIterativeSampler, Measurement
import NeuralQuantum.Optimisers

const libnqhip = get(ENV, "NQHIP_LIB", joinpath(@__DIR__, "..", "nqhip", "libnqhip.so"))

const NQH_OK = Int32(0)
const NQH_ACT_SOFTPLUS = Int32(0)
const NQH_ACT_LOGCOSH = Int32(1)
const NQH_SHIFT_ADD = Int32(0)
const NQH_SHIFT_DIAG = Int32(1)

struct NQHError <: Exception
    status::Int32
    msg::String
end

check(ctx, st) = st == NQH_OK ? nothing :
    throw(NQHError(st, unsafe_string(ccall((:nqh_last_error, libnqhip), Cstring, (Ptr{Cvoid},), ctx))))

# nqh_optable, field for field
struct NQHOpTable
    n_sites::Int32; n_terms::Int32; n_ops::Int32; n_rows::Int32; n_entries::Int32
    term_kind::Ptr{Int32}; term_op_l::Ptr{Int32}; term_op_r::Ptr{Int32}
    op_site_off::Ptr{Int32}; op_sites::Ptr{Int32}; op_row_off::Ptr{Int32}
    row_ent_off::Ptr{Int32}; ent_mel::Ptr{Float64}; ent_flip::Ptr{UInt64}
end

# ---------------------------------------------------------------- tables
mutable struct FlatTable
    n_sites::Int32
    kinds::Vector{Int32}; opl::Vector{Int32}; opr::Vector{Int32}
    site_off::Vector{Int32}; sites::Vector{Int32}; row_off::Vector{Int32}; ent_off::Vector{Int32}
    mel::Vector{Float64}; flip::Vector{UInt64}
end
FlatTable(n) = FlatTable(Int32(n), Int32[], Int32[], Int32[], Int32[0], Int32[], Int32[0], Int32[0],
                         Float64[], UInt64[])

# one KLocalOperator: its sites and its rows' (mel, flipped sites), diagonal
# first as the reference stores them (KLocalOperator.jl:54-112)
function add_op!(t::FlatTable, op)
    append!(t.sites, Int32.(op.sites) .- Int32(1)); push!(t.site_off, length(t.sites))
    for conns in op.op_conns
        for (m, ch) in conns
            push!(t.mel, real(m)); push!(t.mel, imag(m))
            mask = UInt64(0)
            for s in ch.to_change; mask |= UInt64(1) << (s - 1); end
            push!(t.flip, mask)
        end
        push!(t.ent_off, length(t.flip))
    end
    push!(t.row_off, length(t.ent_off) - 1)
    return Int32(length(t.site_off) - 2)
end

ops_of(x) = x isa NeuralQuantum.KLocalOperatorSum ? x.operators : (x,)

"""The connection table of a KLocalLiouvillian, in the order
accumulate_connections! walks it (KLocalLiouvillian.jl:46-52)."""
function flatten_liouvillian(L)
    t = FlatTable(NeuralQuantum.nsites(NeuralQuantum.physical(L.hilb)))
    for o in ops_of(L.HnH_l); push!(t.kinds, 0); push!(t.opl, add_op!(t, o.op_l)); push!(t.opr, -1); end
    for o in ops_of(L.HnH_r); push!(t.kinds, 1); push!(t.opl, -1); push!(t.opr, add_op!(t, o.op_r)); end
    if !(L.LLdag isa NeuralQuantum.KLocalOperatorZero)
        for o in ops_of(L.LLdag)
            push!(t.kinds, 2); push!(t.opl, add_op!(t, o.op_l)); push!(t.opr, add_op!(t, o.op_r))
        end
    end
    return t
end

"""The table of a ket operator (a Hamiltonian or an observable, a
KLocalOperator or KLocalOperatorSum): kind-3 terms."""
function flatten_ket(op, n_sites)
    t = FlatTable(n_sites)
    for o in ops_of(op); push!(t.kinds, 3); push!(t.opl, add_op!(t, o)); push!(t.opr, -1); end
    return t
end

with_optable(f, t::FlatTable) = GC.@preserve t f(NQHOpTable(t.n_sites, length(t.kinds), length(t.site_off) - 1,
    length(t.ent_off) - 1, length(t.flip), pointer(t.kinds), pointer(t.opl), pointer(t.opr),
    pointer(t.site_off), pointer(t.sites), pointer(t.row_off), pointer(t.ent_off),
    pointer(t.mel), pointer(t.flip)))

# ---------------------------------------------------------------- parameters
# the flat order of the ABI: NDM functor order (NDM.jl:17-19) as real
# doubles; RBMSplit (ar ac b Wr Wc) and RBM (a b W) complex, interleaved
const HIPNet = Union{NDM, RBMSplit, RBM}
fields(net::NDM) = (net.b_μ, net.h_μ, net.w_μ, net.u_μ, net.b_λ, net.h_λ, net.d_λ, net.w_λ, net.u_λ)
fields(net::RBMSplit) = (net.ar, net.ac, net.b, net.Wr, net.Wc)
fields(net::RBM) = (net.a, net.b, net.W)

function flat_params(net::HIPNet)
    v = vcat((vec(f) for f in fields(net))...)
    return eltype(v) <: Complex ? collect(reinterpret(Float64, ComplexF64.(v))) : Float64.(v)
end

function unflatten!(net::HIPNet, p::Vector{Float64})
    cplx = eltype(first(fields(net))) <: Complex
    vals = cplx ? reinterpret(ComplexF64, p) : p
    i = 1
    for f in fields(net)
        n = length(f); f .= reshape(vals[i:i+n-1], size(f)); i += n
    end
    return net
end

function activation_code(net::NDM)
    net.f === NeuralQuantum.logℒ && return NQH_ACT_SOFTPLUS
    net.f === NeuralQuantum.logℒ2 && return NQH_ACT_LOGCOSH
    throw(ArgumentError("NDM activation $(net.f): only af_softplus and af_logcosh are provided"))
end

# ---------------------------------------------------------------- the sampler
mutable struct HIPBatchedSampler <: AbstractIterativeSampler
    ctx::Ptr{Cvoid}
    net::HIPNet
    algo::Any
    sampl::MetropolisSampler
    batch_sz::Int
    np::Int
    tables::Vector{FlatTable}                 # kept alive while the ctx may read them
    observables::Vector{Pair{String,FlatTable}}
    as_run::Bool
end

"""The preconditioning cache sample! returns (BatchedGradSampler.jl:122):
precondition! on it solves on the device."""
struct HIPPrecondCache
    is::HIPBatchedSampler
end

"""The preconditioned gradient precondition! returns: the step (host copy,
reference order), owned by the cache and overwritten by the next call as
SRDirect.jl:89 does, plus the sampler whose device state holds it."""
struct HIPGrad
    is::HIPBatchedSampler
    delta::Vector{Float64}
end

"""
    BatchedSampler(net, sampl, prob, algo; batch_sz=16, backend=:hip, device=0,
                   precision=:fp64, as_run=false)

`NeuralQuantum.BatchedSampler` (build_Batched.jl:1-14) on libnqhip: an NDM or
RBMSplit with a KLocalLiouvillian (the L†L problem), or a ket RBM with a
Hamiltonian (the energy problem).  `precision`: :fp64 (plain form), :fp32
(FP32 network and S), :mixed (FP32 network, BF16 S build, FP64 solve).
`as_run=true` reproduces the reference's sr_multiplicative + sr_cholesky /
sr_div, which factor the unshifted S (SRDirect.jl:66-81).  `backend=:cpu`
returns the reference's own sampler.
"""
function BatchedSampler(net, sampl, prob, algo=NeuralQuantum.Gradient(); backend::Symbol=:hip, kwargs...)
    backend === :cpu && return NeuralQuantum.BatchedSampler(net, sampl, prob, algo; kwargs...)
    return HIPBatchedSampler(net, sampl, prob, algo; kwargs...)
end

function HIPBatchedSampler(net::HIPNet, sampl::MetropolisSampler, prob, algo::SR;
                           batch_sz::Int=16, local_batch_sz::Int=batch_sz, obs_sampler=sampl,
                           device::Int=0, precision::Symbol=:fp64, as_run::Bool=false)
    sampl.rule isa LocalRule || throw(ArgumentError("only LocalRule is provided"))
    obs_sampler === sampl || throw(ArgumentError("observables use the gradient sampler's settings"))
    algo.algorithm == NeuralQuantum.sr_pivcholesky &&
        throw(ArgumentError("sr_pivcholesky: \"Alg not known\" in the reference (SRDirect.jl:85-86)"))
    algo.algorithm == NeuralQuantum.sr_diag && algo.precondition_type == NeuralQuantum.sr_none &&
        throw(ArgumentError("sr_diag with sr_none is the pseudo-inverse of a possibly singular S: use a shift"))
    ctxref = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:nqh_create, libnqhip), Int32, (Ptr{Ptr{Cvoid}}, Int32, UInt64), ctxref, device, sampl.seed)
    st == NQH_OK || throw(NQHError(st, "nqh_create"))
    ctx = ctxref[]
    ket = !(prob isa NeuralQuantum.KLocalLiouvillian)
    if ket
        net isa RBM || throw(ArgumentError("a Hamiltonian takes a ket RBM"))
        t = flatten_ket(prob, length(net.a))
    else
        net isa RBM && throw(ArgumentError("the L†L problem takes an NDM or an RBMSplit"))
        t = flatten_liouvillian(prob)
    end
    with_optable(t) do tab
        check(ctx, ccall((:nqh_set_problem, libnqhip), Int32, (Ptr{Cvoid}, Ref{NQHOpTable}), ctx, tab))
    end
    if net isa NDM
        N = length(net.b_μ); nh = length(net.h_μ); na = length(net.d_λ)
        check(ctx, ccall((:nqh_set_ndm, libnqhip), Int32, (Ptr{Cvoid}, Int32, Int32, Int32, Int32),
                         ctx, N, nh, na, activation_code(net)))
    elseif net isa RBMSplit
        check(ctx, ccall((:nqh_set_rbmsplit, libnqhip), Int32, (Ptr{Cvoid}, Int32, Int32),
                         ctx, length(net.ar), length(net.b)))
    else
        check(ctx, ccall((:nqh_set_rbm, libnqhip), Int32, (Ptr{Cvoid}, Int32, Int32),
                         ctx, length(net.a), length(net.b)))
    end
    np = Int(ccall((:nqh_num_params, libnqhip), Int64, (Ptr{Cvoid},), ctx))
    prec = precision == :fp64 ? (0, 0) : precision == :mixed ? (1, 2) : (1, 1)
    check(ctx, ccall((:nqh_set_precision, libnqhip), Int32, (Ptr{Cvoid}, Int32, Int32), ctx, prec...))
    # passes: MetropolisSampler already bumped an even value to odd (Metropolis.jl:30-38)
    check(ctx, ccall((:nqh_set_sampler, libnqhip), Int32, (Ptr{Cvoid}, Int64, Int32, Int32, Int32, Int32),
                     ctx, batch_sz, sampl.chain_length, sampl.passes, sampl.burn_length, 0))
    # SR(algorithm=sr_cg / sr_minres / sr_qlp / sr_lsq): the matrix-free iterative
    # solves (SRIterative.jl:71-152); sr_cholesky, sr_div and (shifted) sr_diag factor S
    a = algo.algorithm
    solver = a == NeuralQuantum.sr_cg ? 1 : a == NeuralQuantum.sr_minres ? 2 :
             a == NeuralQuantum.sr_qlp ? 3 : a == NeuralQuantum.sr_lsq ? 4 : 0
    check(ctx, ccall((:nqh_set_sr_solver, libnqhip), Int32, (Ptr{Cvoid}, Int32, Float64, Int32),
                     ctx, solver, Float64(algo.sr_precision), 0))
    is = HIPBatchedSampler(ctx, net, algo, sampl, batch_sz, np, [t], Pair{String,FlatTable}[], as_run)
    finalizer(x -> ccall((:nqh_destroy, libnqhip), Cvoid, (Ptr{Cvoid},), x.ctx), is)
    return is
end

# sample!(is) -> (L_stat, precond_cache)   (BatchedGradSampler.jl:77-123,
# BatchedValSampler.jl:117-139): the host network is uploaded first
function sample!(is::HIPBatchedSampler; sample=true)
    p = flat_params(is.net)
    check(is.ctx, ccall((:nqh_set_params, libnqhip), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int64), is.ctx, p, is.np))
    st = zeros(8)
    check(is.ctx, ccall((:nqh_sample_grad, libnqhip), Int32, (Ptr{Cvoid}, Ptr{Float64}), is.ctx, st))
    if is.net isa RBM   # the energy is complex: out[5] holds its imaginary part
        check(is.ctx, ccall((:nqh_get_stats, libnqhip), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int32), is.ctx, st, 6))
        return Measurement(complex(st[1], st[6]), st[2], st[3], st[4], st[5]), HIPPrecondCache(is)
    end
    return Measurement(st[1], st[2], st[3], st[4], st[5]), HIPPrecondCache(is)
end

# precondition!(cache, SR, iter) -> gradient (SRDirect.jl:51-90)
# sr_shift: eps I; sr_multiplicative: max(λ0 b^iter, λmin) diag(S) (SRDirect.jl:56-70),
# or nothing with as_run and a factorization (the reference factors S itself,
# SRDirect.jl:76-81); sr_none: no shift
function precondition!(c::HIPPrecondCache, algo::SR, iter)
    is = c.is
    pt = algo.precondition_type
    direct = algo.algorithm in (NeuralQuantum.sr_cholesky, NeuralQuantum.sr_div)
    eps, kind = pt == NeuralQuantum.sr_shift ? (Float64(algo.sr_diag_shift), NQH_SHIFT_ADD) :
                pt == NeuralQuantum.sr_multiplicative ?
                    (is.as_run && direct ? (0.0, NQH_SHIFT_ADD) :
                     (Float64(max(algo.λ0 * algo.b^iter, algo.λmin)), NQH_SHIFT_DIAG)) :
                (0.0, NQH_SHIFT_ADD)
    delta = zeros(is.np)
    check(is.ctx, ccall((:nqh_precondition, libnqhip), Int32,
                        (Ptr{Cvoid}, Float64, Int32, Int32, Ptr{Float64}),
                        is.ctx, eps, kind, 0, delta))
    return HIPGrad(is, delta)
end

# Optimisers.update!(Descent(η), net, grad) (apply.jl:9, rules.jl:15-17): the
# step is applied on the device by the context that solved it, then the
# network's arrays are refreshed from the device copy
function Optimisers.update!(opt::Optimisers.Descent, net::HIPNet, g::HIPGrad)
    g.is.net === net || throw(ArgumentError("the gradient was computed for another network"))
    is = g.is
    check(is.ctx, ccall((:nqh_update, libnqhip), Int32, (Ptr{Cvoid}, Float64), is.ctx, opt.eta))
    p = zeros(is.np)
    check(is.ctx, ccall((:nqh_get_params, libnqhip), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int64), is.ctx, p, is.np))
    unflatten!(net, p)
    return nothing
end

# add_observable!(is, name, op) / compute_observables(is)
# (BatchedGradSampler.jl:125-131, BatchedObsDMSampler.jl:55-77): a diagonal
# chain with the gradient sampler's settings per observable
function add_observable!(is::HIPBatchedSampler, name::String, op)
    is.net isa RBM && throw(ArgumentError("observables of the density-matrix path"))
    n = is.net isa NDM ? length(is.net.b_μ) : length(is.net.ar)
    push!(is.observables, name => flatten_ket(op, n))
    return is
end

function compute_observables(is::HIPBatchedSampler)
    isempty(is.observables) && return nothing
    res = Dict{String,Measurement}()
    s = is.sampl
    for (name, t) in is.observables
        st = zeros(6)
        with_optable(t) do tab
            check(is.ctx, ccall((:nqh_observable, libnqhip), Int32,
                                (Ptr{Cvoid}, Ref{NQHOpTable}, Int64, Int32, Int32, Int32, Ptr{Float64}),
                                is.ctx, tab, is.batch_sz, s.chain_length, s.passes, s.burn_length, st))
        end
        res[name] = Measurement(complex(st[1], st[6]), st[2], st[3], st[4], st[5])
    end
    return res
end

"""Exact sums over every configuration (the batched ExactSampler with the
accumulator on top, Exact.jl:110-181): (⟨L†L⟩ or E, Σ p L_loc, log Z)."""
function exact_sums(is::HIPBatchedSampler)
    p = flat_params(is.net)
    check(is.ctx, ccall((:nqh_set_params, libnqhip), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int64), is.ctx, p, is.np))
    out = zeros(8)
    check(is.ctx, ccall((:nqh_exact, libnqhip), Int32, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                        is.ctx, out, C_NULL))
    return out
end

"""Accepted moves per chain in the last sampling call (0: a stuck chain)."""
function chain_acceptance(is::HIPBatchedSampler)
    out = zeros(UInt32, is.batch_sz)
    check(is.ctx, ccall((:nqh_get_chain_acceptance, libnqhip), Int32, (Ptr{Cvoid}, Ptr{UInt32}, Int64),
                        is.ctx, out, is.batch_sz))
    return out
end

"""Launch shapes (0 = chosen from the problem): estimator samples and waves per
workgroup, BF16 Gram split-K chunks, sampler waves per workgroup
(`nqh_set_launch_shape`; the values `nqhip.autotune` stores in launch_shapes.json)."""
function set_launch_shape!(is::HIPBatchedSampler; est_samples_per_block=0, est_waves=0, gram_splits=0,
                           sampler_waves=0)
    check(is.ctx, ccall((:nqh_set_launch_shape, libnqhip), Int32, (Ptr{Cvoid}, Int32, Int32, Int32, Int32),
                        is.ctx, est_samples_per_block, est_waves, gram_splits, sampler_waves))
    return is
end

export HIPBatchedSampler, exact_sums, chain_acceptance, set_launch_shape!

end # module
