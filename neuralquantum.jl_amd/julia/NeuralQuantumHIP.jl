# NeuralQuantumHIP.jl -- Julia glue that plugs libnqhip (include/nqhip.h) into
# NeuralQuantum.jl as a drop-in backend for the steady-state VMC loop.
#
# Not executed in this repository (no Julia toolchain here); the same C ABI is
# exercised from Python through ctypes by the tests.  The methods below extend
# the reference's own generic functions so a user's driver loop
# (examples/dissipative_ising1d.jl:45-55) runs unchanged:
#
#     is = HIPBatchedSampler(net, sampl, liouv, SR(ϵ=0.01); batch_sz=4096)
#     for i = 1:300
#         ldata, prec = sample!(is)
#         grad = precondition!(prec, algo, i)
#         Optimisers.update!(optimizer, net, grad)
#     end
module NeuralQuantumHIP

using NeuralQuantum
import NeuralQuantum: sample!, precondition!, AbstractIterativeSampler, Measurement
import NeuralQuantum.Optimisers

const libnqhip = get(ENV, "NQHIP_LIB", joinpath(@__DIR__, "..", "nqhip", "libnqhip.so"))

const NQH_OK = Int32(0)

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

"""Flatten a KLocalLiouvillian into the arrays of `nqh_optable`, walking the
operators exactly as `accumulate_connections!` does (KLocalLiouvillian.jl:46-52)."""
function flatten_liouvillian(L)
    kinds = Int32[]; opl = Int32[]; opr = Int32[]
    site_off = Int32[0]; sites = Int32[]; row_off = Int32[0]; ent_off = Int32[0]
    mel = Float64[]; flip = UInt64[]
    function add_op(op)
        append!(sites, Int32.(op.sites) .- 1); push!(site_off, length(sites))
        for conns in op.op_conns
            for (m, ch) in conns
                push!(mel, real(m)); push!(mel, imag(m))
                mask = UInt64(0)
                for s in ch.to_change; mask |= UInt64(1) << (s - 1); end
                push!(flip, mask)
            end
            push!(ent_off, length(flip))
        end
        push!(row_off, length(ent_off) - 1)
        return Int32(length(site_off) - 2)
    end
    ops(x) = x isa NeuralQuantum.KLocalOperatorSum ? x.operators : (x,)
    for t in ops(L.HnH_l); push!(kinds, 0); push!(opl, add_op(t.op_l)); push!(opr, -1); end
    for t in ops(L.HnH_r); push!(kinds, 1); push!(opl, -1); push!(opr, add_op(t.op_r)); end
    if !(L.LLdag isa NeuralQuantum.KLocalOperatorZero)
        for t in ops(L.LLdag)
            push!(kinds, 2); push!(opl, add_op(t.op_l)); push!(opr, add_op(t.op_r))
        end
    end
    return (; n_sites=Int32(NeuralQuantum.nsites(NeuralQuantum.physical(L.hilb))),
              kinds, opl, opr, site_off, sites, row_off, ent_off, mel, flip)
end

flat_params(net::NDM) = vcat(vec(net.b_μ), vec(net.h_μ), vec(net.w_μ), vec(net.u_μ),
                             vec(net.b_λ), vec(net.h_λ), vec(net.d_λ), vec(net.w_λ), vec(net.u_λ))

function unflatten!(net::NDM, p::Vector{Float64})
    i = 1
    for f in (net.b_μ, net.h_μ, net.w_μ, net.u_μ, net.b_λ, net.h_λ, net.d_λ, net.w_λ, net.u_λ)
        n = length(f); f .= reshape(p[i:i+n-1], size(f)); i += n
    end
    return net
end

mutable struct HIPBatchedSampler <: AbstractIterativeSampler
    ctx::Ptr{Cvoid}
    net::NDM
    algo::SR
    np::Int
    delta::Vector{Float64}
    table_arrays::Any          # keeps the flattened table alive while the ctx may read it
end

function HIPBatchedSampler(net::NDM, sampl::MetropolisSampler, liouv, algo::SR;
                           batch_sz::Int=16, device::Int=0, precision::Symbol=:fp64)
    ctxref = Ref{Ptr{Cvoid}}(C_NULL)
    st = ccall((:nqh_create, libnqhip), Int32, (Ptr{Ptr{Cvoid}}, Int32, UInt64), ctxref, device, sampl.seed)
    st == NQH_OK || throw(NQHError(st, "nqh_create"))
    ctx = ctxref[]
    t = flatten_liouvillian(liouv)
    GC.@preserve t begin
        tab = NQHOpTable(t.n_sites, length(t.kinds), length(t.site_off) - 1, length(t.ent_off) - 1,
                         length(t.flip), pointer(t.kinds), pointer(t.opl), pointer(t.opr),
                         pointer(t.site_off), pointer(t.sites), pointer(t.row_off), pointer(t.ent_off),
                         pointer(t.mel), pointer(t.flip))
        check(ctx, ccall((:nqh_set_problem, libnqhip), Int32, (Ptr{Cvoid}, Ref{NQHOpTable}), ctx, tab))
    end
    N = length(net.b_μ); nh = length(net.h_μ); na = length(net.d_λ)
    check(ctx, ccall((:nqh_set_ndm, libnqhip), Int32, (Ptr{Cvoid}, Int32, Int32, Int32, Int32), ctx, N, nh, na, 0))
    np = Int(ccall((:nqh_num_params, libnqhip), Int64, (Ptr{Cvoid},), ctx))
    prec = precision == :fp64 ? (0, 0) : precision == :mixed ? (1, 2) : (1, 1)
    check(ctx, ccall((:nqh_set_precision, libnqhip), Int32, (Ptr{Cvoid}, Int32, Int32), ctx, prec...))
    p = Float64.(flat_params(net))
    check(ctx, ccall((:nqh_set_params, libnqhip), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int64), ctx, p, np))
    # passes: MetropolisSampler already bumped an even value to odd (Metropolis.jl:30-38)
    check(ctx, ccall((:nqh_set_sampler, libnqhip), Int32, (Ptr{Cvoid}, Int64, Int32, Int32, Int32, Int32),
                     ctx, batch_sz, sampl.chain_length, sampl.passes, sampl.burn_length, 0))
    # SR(algorithm=sr_cg / sr_minres / sr_qlp / sr_lsq): the matrix-free iterative
    # solves (SRIterative.jl:71-152); every other algorithm factors S (SRDirect.jl)
    solver = algo.algorithm == NeuralQuantum.sr_cg ? 1 : algo.algorithm == NeuralQuantum.sr_minres ? 2 :
             algo.algorithm == NeuralQuantum.sr_qlp ? 3 : algo.algorithm == NeuralQuantum.sr_lsq ? 4 : 0
    check(ctx, ccall((:nqh_set_sr_solver, libnqhip), Int32, (Ptr{Cvoid}, Int32, Float64, Int32),
                     ctx, solver, Float64(algo.sr_precision), 0))
    is = HIPBatchedSampler(ctx, net, algo, np, zeros(np), t)
    finalizer(x -> ccall((:nqh_destroy, libnqhip), Cvoid, (Ptr{Cvoid},), x.ctx), is)
    return is
end

# sample!(is) -> (L_stat, precond_cache)   (BatchedGradSampler.jl:77-123)
function sample!(is::HIPBatchedSampler; sample=true)
    st = zeros(5)
    check(is.ctx, ccall((:nqh_sample_grad, libnqhip), Int32, (Ptr{Cvoid}, Ptr{Float64}), is.ctx, st))
    return Measurement(st[1], st[2], st[3], st[4], st[5]), is
end

# precondition!(cache, SR, iter) -> gradient (SRDirect.jl:51-90); owned by the
# cache and overwritten by the next call, as in the reference (SRDirect.jl:89)
# sr_shift: eps I; sr_multiplicative: max(λ0 b^iter, λmin) diag(S) (SRDirect.jl:56-70); sr_none: no shift
function precondition!(is::HIPBatchedSampler, algo::SR, iter)
    pt = algo.precondition_type
    eps, kind = pt == NeuralQuantum.sr_shift ? (Float64(algo.sr_diag_shift), 0) :
                pt == NeuralQuantum.sr_multiplicative ? (Float64(max(algo.λ0 * algo.b^iter, algo.λmin)), 1) :
                (0.0, 0)
    check(is.ctx, ccall((:nqh_precondition, libnqhip), Int32,
                        (Ptr{Cvoid}, Float64, Int32, Int32, Ptr{Float64}),
                        is.ctx, eps, kind, 0, is.delta))
    return is.delta
end

# Optimisers.update!(Descent(η), net, grad) on the device, mirrored to the host arrays
function Optimisers.update!(opt::Optimisers.Descent, is::HIPBatchedSampler, grad)
    check(is.ctx, ccall((:nqh_update, libnqhip), Int32, (Ptr{Cvoid}, Float64), is.ctx, opt.eta))
    p = zeros(is.np)
    check(is.ctx, ccall((:nqh_get_params, libnqhip), Int32, (Ptr{Cvoid}, Ptr{Float64}, Int64), is.ctx, p, is.np))
    unflatten!(is.net, p)
    return is.net
end

export HIPBatchedSampler

end # module
