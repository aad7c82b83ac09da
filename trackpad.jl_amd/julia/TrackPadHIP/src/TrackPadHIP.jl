"""
    TrackPadHIP

MI355X backend for TrackPad.jl's multi-particle tracking path.  A static `TrackPad.Lattice`
is compiled into the flat device element table of `include/trackpad_table.h`
(`compile_table`), uploaded once, and tracked by `libtrackpad_hip` through its C ABI
(`include/trackpad_hip.h`).  Device-resident particles (`HIPParticles`) get methods of the
reference's `linepass!` / `ringpass!` (src/lattice.jl:260-317) and a `track!`; host arrays go
through `linepass_hip!` / `ringpass_hip!` (no type piracy on `Matrix`).  Gradients of
`mean(x^2 + y^2)` come from the hand-written adjoint kernel (`ringpass_grad`), which the
ChainRulesCore extension registers as an `rrule`.
"""
module TrackPadHIP

using Libdl
using TrackPad
using TrackPad: AbstractElement, Beam, Lattice, TimeVaryingElement

export HIPContext, HIPTable, HIPParticles, HIPBatch, compile_table, track!, linepass_hip!, ringpass_hip!,
       ringpass_grad, ExactSBendLinearEdge, download, jacobian, autotune!, set_waves!, set_launch!, batch_grad,
       LaunchCfg, read_madx, MadxError

const LIBPATH = Ref{String}(get(ENV, "TRACKPAD_HIP_LIB",
                                joinpath(@__DIR__, "..", "..", "..", "lib", "libtrackpad_hip.so")))
libtp() = LIBPATH[]

const TP_TRACK_FAST = UInt32(1)
const TP_TRACK_ASYNC = UInt32(2)
const TP_TRACK_TIMED = UInt32(4)

struct TrackPadHIPError <: Exception
    code::Cint
    msg::String
end
Base.showerror(io::IO, e::TrackPadHIPError) = print(io, "libtrackpad_hip error ", e.code, ": ", e.msg)

function check(rc::Cint)
    if rc != 0
        msg = unsafe_string(ccall((:tp_last_error, libtp()), Cstring, ()))
        throw(TrackPadHIPError(rc, msg))
    end
    return nothing
end

# ----------------------------------------------------------------------------------------
# Handles (finalizers may run in any order: the library reference-counts the context)
# ----------------------------------------------------------------------------------------
mutable struct HIPContext
    handle::Ptr{Cvoid}
    device::Int
    tables::Dict{Vector{UInt8}, Any}      # uploaded tables by blob (HIPTable values)
    memo::WeakKeyDict{Any, Any}           # lattice element vector -> (options, elements, HIPTable)
    function HIPContext(device::Integer = 0)
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:tp_ctx_create, libtp()), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, h))
        c = new(h[], Int(device), Dict{Vector{UInt8}, Any}(), WeakKeyDict{Any, Any}())
        finalizer(c) do x
            ccall((:tp_ctx_destroy, libtp()), Cint, (Ptr{Cvoid},), x.handle)
        end
        return c
    end
end

const DEFAULT_CTX = Ref{Union{Nothing, HIPContext}}(nothing)
default_context() = (DEFAULT_CTX[] === nothing && (DEFAULT_CTX[] = HIPContext(0)); DEFAULT_CTX[])

set_launch!(ctx::HIPContext; block = 0, ppl = 0, turns_per_launch = 0) =
    check(ccall((:tp_set_launch, libtp()), Cint, (Ptr{Cvoid}, Cint, Cint, Cint), ctx.handle, block, ppl,
                turns_per_launch))

mutable struct HIPTable
    handle::Ptr{Cvoid}
    ctx::HIPContext
    n_seq::Int
    function HIPTable(ctx::HIPContext, blob::Vector{UInt8})
        h = Ref{Ptr{Cvoid}}(C_NULL)
        GC.@preserve blob check(ccall((:tp_table_create, libtp()), Cint,
                                      (Ptr{Cvoid}, Ptr{UInt8}, Csize_t, Ptr{Ptr{Cvoid}}),
                                      ctx.handle, blob, length(blob), h))
        ns = Ref{Int32}(0)
        check(ccall((:tp_table_info, libtp()), Cint, (Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}, Ptr{Float64}),
                    h[], ns, C_NULL, C_NULL))
        t = new(h[], ctx, Int(ns[]))
        finalizer(t) do x
            ccall((:tp_table_destroy, libtp()), Cint, (Ptr{Cvoid},), x.handle)
        end
        return t
    end
end
HIPTable(lat::Lattice, beam::Beam; ctx::HIPContext = default_context(), kw...) =
    HIPTable(ctx, compile_table(lat, beam; kw...))

# An element whose fields are all immutable values is unchanged exactly when it is the same
# (===) object; elements holding vectors (wake samples, slice weights, wiggler harmonics,
# WithMultipoles) could change in place, so their lattices are compiled on every call.
immutable_fields(e) = !any(f -> ismutable(getfield(e, f)), fieldnames(typeof(e)))

"""
    table_for(ctx, lat, beam; aperture = nothing, turn = 0, nturns = nothing, time = 0.0,
              dt_turn = 0.0) -> HIPTable

The device table of `lat` for a tracking call, uploaded once per context and reused:

* a static lattice whose elements are the same objects as in an earlier call with the same
  beam and aperture gets that call's table without being compiled again;
* otherwise the lattice is compiled (time-varying elements resolved per turn for
  `turn .. turn + nturns - 1`, as `ringpass!` resolves them, src/lattice.jl:305-317) and the
  blob is looked up among the context's uploaded tables (at most 64 are kept).
"""
function table_for(ctx::HIPContext, lat::Lattice, beam::Beam; aperture = nothing, turn::Integer = 0,
                   nturns::Union{Nothing, Integer} = nothing, time::Real = 0.0, dt_turn::Real = 0.0)
    els = lat.elements
    varying = any(w -> base(w) isa TimeVaryingElement, els)
    memo_ok = !varying && all(immutable_fields, els)
    opts = (beam, aperture === nothing ? nothing : Float64(aperture))
    if memo_ok
        m = get(ctx.memo, els, nothing)
        if m !== nothing && m[1] == opts && length(m[2]) == length(els) &&
           all(i -> m[2][i] === els[i], eachindex(els))
            return m[3]::HIPTable
        end
    end
    blob = compile_table(lat, beam; aperture = aperture, turn = turn, nturns = nturns, time = time,
                         dt_turn = dt_turn)
    tab = get(ctx.tables, blob, nothing)
    if tab === nothing
        length(ctx.tables) >= 64 && empty!(ctx.tables)
        tab = HIPTable(ctx, blob)
        ctx.tables[blob] = tab
    end
    memo_ok && (ctx.memo[els] = (opts, collect(Any, els), tab))
    return tab::HIPTable
end

mutable struct HIPParticles
    handle::Ptr{Cvoid}
    ctx::HIPContext
    n::Int
    function HIPParticles(n::Integer; ctx::HIPContext = default_context())
        h = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:tp_particles_create, libtp()), Cint, (Ptr{Cvoid}, Int64, Ptr{Ptr{Cvoid}}), ctx.handle, n, h))
        p = new(h[], ctx, Int(n))
        finalizer(p) do x
            ccall((:tp_particles_destroy, libtp()), Cint, (Ptr{Cvoid},), x.handle)
        end
        return p
    end
end

"""
    HIPParticles(coords::Matrix{Float64}, lost_flags = nothing)

Upload an N x 6 beam (the reference's `coords` layout, src/lattice.jl:274-275: the Julia
column-major matrix is exactly the device's six planes).  `lost_flags[i] == 1` marks a
particle as already lost (src/lattice.jl:269-271).
"""
function HIPParticles(coords::Matrix{Float64}, lost_flags::Union{Nothing, Vector{Int}} = nothing;
                      ctx::HIPContext = default_context())
    size(coords, 2) == 6 || throw(ArgumentError("coords must be N x 6"))
    p = HIPParticles(size(coords, 1); ctx = ctx)
    upload!(p, coords, lost_flags)
    return p
end

function upload!(p::HIPParticles, coords::Matrix{Float64}, lost_flags = nothing)
    size(coords) == (p.n, 6) || throw(DimensionMismatch("coords must be $(p.n) x 6"))
    lf = lost_flags === nothing ? C_NULL : pointer(lost_flags)
    GC.@preserve coords lost_flags check(ccall((:tp_particles_upload, libtp()), Cint,
                                                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Ptr{Int32}, Ptr{Int32}),
                                                p.handle, coords, lf, C_NULL, C_NULL))
    return p
end

"""(coords, lost_flags, lost_turn, lost_elem) on the host; lost_turn / lost_elem are 0-based, -1 alive."""
function download(p::HIPParticles)
    coords = Matrix{Float64}(undef, p.n, 6)
    lf = Vector{Int64}(undef, p.n)
    lt = Vector{Int32}(undef, p.n)
    le = Vector{Int32}(undef, p.n)
    check(ccall((:tp_particles_download, libtp()), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}, Ptr{Int32}, Ptr{Int32}), p.handle, coords, lf, lt, le))
    return coords, Vector{Int}(lf), lt, le
end

# ----------------------------------------------------------------------------------------
# Tracking entry points
# ----------------------------------------------------------------------------------------
"""
    track!(p::HIPParticles, lat, beam; nturns = 1, turn = 0, fast = false)

Track device-resident particles for `nturns` turns (the fused whole-turn kernel).
"""
function track!(p::HIPParticles, tab::HIPTable; nturns::Integer = 1, turn::Integer = 0, fast::Bool = false)
    tab.ctx === p.ctx || throw(ArgumentError("table and particles belong to different contexts"))
    check(ccall((:tp_track, libtp()), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, UInt32),
                p.ctx.handle, tab.handle, p.handle, turn, nturns, fast ? TP_TRACK_FAST : UInt32(0)))
    return p
end

"""
    track!(p::HIPParticles, lat, beam; nturns = 1, turn = 0, time = 0.0, dt_turn = 0.0,
           fast = false, aperture = nothing)

Track device-resident particles through `lat` for `nturns` turns from turn `turn`; turn k
sees the lattice at time `time` advanced by `dt_turn` k times (only time-varying elements
depend on it).  The table comes from `table_for` (compiled and uploaded once per lattice).
"""
function track!(p::HIPParticles, lat::Lattice, beam::Beam; nturns::Integer = 1, turn::Integer = 0,
                time::Real = 0.0, dt_turn::Real = 0.0, fast::Bool = false, aperture = nothing)
    tab = table_for(p.ctx, lat, beam; aperture = aperture, turn = turn, nturns = nturns, time = time,
                    dt_turn = dt_turn)
    return track!(p, tab; nturns = nturns, turn = turn, fast = fast)
end

# Methods of the reference's functions for device particles (src/lattice.jl:260, 305), with
# its keywords: linepass! resolves time-varying elements at (time, turn), ringpass! at
# (time + k dt_turn, turn + k) for turn k.
TrackPad.linepass!(p::HIPParticles, lat::Lattice, beam::Beam; time::Real = 0, turn::Integer = 0, kw...) =
    (track!(p, lat, beam; nturns = 1, turn = turn, time = time, kw...); nothing)
TrackPad.ringpass!(p::HIPParticles, lat::Lattice, beam::Beam, nturns::Int; time::Real = 0,
                   dt_turn::Real = 0, turn::Integer = 0, kw...) =
    (track!(p, lat, beam; nturns = nturns, turn = turn, time = time, dt_turn = dt_turn, kw...); nothing)

"""
    ringpass_hip!(coords, lat, beam, lost_flags, nturns; time = 0.0, dt_turn = 0.0, turn = 0,
                  fast = false, aperture = nothing)

`ringpass!` on host arrays through the GPU: mutates `coords` and `lost_flags` in place with the
reference's semantics (flags equal to 1 are skipped; newly lost particles get 1 and keep their
element-entry coordinates; time-varying elements resolved per turn as `ringpass!` does).
"""
function ringpass_hip!(coords::Matrix{Float64}, lat::Lattice, beam::Beam, lost_flags::Vector{Int},
                       nturns::Integer; time::Real = 0.0, dt_turn::Real = 0.0, turn::Integer = 0,
                       fast::Bool = false, aperture = nothing, ctx::HIPContext = default_context())
    n = size(coords, 1)
    size(coords, 2) == 6 || throw(ArgumentError("coords must be N x 6"))
    length(lost_flags) == n || throw(DimensionMismatch("lost_flags must have N entries"))
    nturns >= 0 || throw(ArgumentError("nturns = $nturns < 0"))
    tab = table_for(ctx, lat, beam; aperture = aperture, turn = turn, nturns = nturns, time = time,
                    dt_turn = dt_turn)
    GC.@preserve coords lost_flags check(ccall((:tp_track_host, libtp()), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Int32}, Ptr{Int32}, Int32, Int32, UInt32),
        ctx.handle, tab.handle, coords, n, lost_flags, C_NULL, C_NULL, turn, nturns,
        fast ? TP_TRACK_FAST : UInt32(0)))
    return nothing
end
linepass_hip!(coords::Matrix{Float64}, lat::Lattice, beam::Beam, lost_flags::Vector{Int};
              time::Real = 0.0, turn::Integer = 0, kw...) =
    ringpass_hip!(coords, lat, beam, lost_flags, 1; time = time, turn = turn, kw...)

"""
    ringpass_grad(coords0, lat, beam, nturns; turn = 0, time = 0.0, dt_turn = 0.0, ckpt_stride = 1,
                  fast = false) -> (loss, dL_dk::Vector{Float64}, dL_dx0::Matrix{Float64}, n_alive)

`loss = mean over surviving particles of (x^2 + y^2)` after `nturns` turns; `dL_dk[i]` is the
derivative with respect to the main strength of lattice position `i` (k1 of a Quadrupole, k2
of a Sextupole, k3 of an Octupole, 0 elsewhere; summed over the turns of a time-varying
position); `dL_dx0` with respect to the initial beam.
"""
function ringpass_grad(coords0::Matrix{Float64}, lat::Lattice, beam::Beam, nturns::Integer;
                       turn::Integer = 0, time::Real = 0.0, dt_turn::Real = 0.0, ckpt_stride::Integer = 1,
                       fast::Bool = false, ctx::HIPContext = default_context())
    tab = table_for(ctx, lat, beam; turn = turn, nturns = nturns, time = time, dt_turn = dt_turn)
    p = HIPParticles(coords0; ctx = ctx)
    loss = Ref{Float64}(0)
    na = Ref{Int64}(0)
    dk = zeros(Float64, max(tab.n_seq, 1))
    dx0 = zeros(Float64, p.n, 6)
    check(ccall((:tp_adjoint, libtp()), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Int32, UInt32, Ptr{Float64}, Ptr{Int64},
                 Ptr{Float64}, Ptr{Float64}),
                ctx.handle, tab.handle, p.handle, turn, nturns, ckpt_stride, fast ? TP_TRACK_FAST : UInt32(0),
                loss, na, dk, dx0))
    return loss[], dk[1:tab.n_seq], dx0, Int(na[])
end

"""A copy of `lat` whose main strengths (k1 / k2 / k3) at 1-based `positions` are `k`."""
function set_strengths(lat::Lattice, positions, k)
    els = Any[e for e in lat.elements]
    for (p, v) in zip(positions, k)
        e = els[p]
        f = e isa Quadrupole ? :k1 : e isa Sextupole ? :k2 : e isa Octupole ? :k3 :
            throw(ArgumentError("position $p has no strength"))
        els[p] = typeof(e)((n === f ? oftype(getfield(e, n), v) : getfield(e, n) for n in fieldnames(typeof(e)))...)
    end
    return Lattice(els; name = lat.name)
end

"""
    track_loss(k, coords0, lat, beam, nturns, positions)

`mean(x^2 + y^2)` over surviving particles after `nturns` turns of `lat` with strengths `k` at
`positions`; differentiable through the ChainRulesCore extension (hand-written adjoint).
"""
function track_loss(k::Vector{Float64}, coords0::Matrix{Float64}, lat, beam, nturns::Integer, positions)
    ctx = default_context()
    p = HIPParticles(coords0; ctx = ctx)
    track!(p, set_strengths(lat, positions, k), beam; nturns = nturns)
    c, lf, _, _ = download(p)
    alive = lf .== 0
    return sum(c[alive, 1] .^ 2 .+ c[alive, 3] .^ 2) / count(alive)
end

# ----------------------------------------------------------------------------------------
# Register budgets, autotuner, Jacobians, batches (include/trackpad_hip.h)
# ----------------------------------------------------------------------------------------
"""
    set_waves!(ctx; fwd_waves = 0, back_waves = 0)

Register budget (resident waves per SIMD) of the tracking kernels (0, 5, 6) and of the
reverse kernels (0, 3, 4); 0 leaves it to the compiler.
"""
set_waves!(ctx::HIPContext; fwd_waves::Integer = 0, back_waves::Integer = 0) =
    check(ccall((:tp_set_waves, libtp()), Cint, (Ptr{Cvoid}, Cint, Cint), ctx.handle, fwd_waves, back_waves))

# tp_launch_cfg
struct LaunchCfg
    block::Int32
    ppl::Int32
    turns_per_launch::Int32
    fwd_waves::Int32
    back_waves::Int32
    candidates::Int32
    rate::Float64
end
LaunchCfg() = LaunchCfg(0, 0, 0, 0, 0, 0, 0.0)

const TP_TUNE_ADJOINT = UInt32(0x100)

"""
    autotune!(p, tab; probe_turns = 2, fast = false, target_launch_ms = 0, cache = nothing) -> LaunchCfg

Time every compiled launch shape on a device copy of `p` (left unchanged) and apply the
fastest to the context, with turns per launch for launches of about `target_launch_ms`.
"""
function autotune!(p::HIPParticles, tab::HIPTable; probe_turns::Integer = 2, fast::Bool = false,
                   target_launch_ms::Real = 0, cache::Union{Nothing, AbstractString} = nothing)
    out = Ref(LaunchCfg())
    check(ccall((:tp_autotune, libtp()), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, UInt32, Float64, Cstring, Ref{LaunchCfg}),
                p.ctx.handle, tab.handle, p.handle, probe_turns, fast ? TP_TRACK_FAST : UInt32(0),
                target_launch_ms, cache === nothing ? C_NULL : cache, out))
    return out[]
end

"""
    jacobian(p, tab; nturns = 1, turn = 0, ckpt_stride = 1, fast = false) -> Array{Float64,3}

Exact `nturns`-turn Jacobian d(final)/d(initial) of every particle of `p` (N x 6 x 6,
`J[q, i, j] = d out_i / d in_j`), by six reverse passes with unit seeds.  `p` is tracked.
"""
function jacobian(p::HIPParticles, tab::HIPTable; nturns::Integer = 1, turn::Integer = 0,
                  ckpt_stride::Integer = 1, fast::Bool = false)
    jac = zeros(Float64, 36 * p.n)
    check(ccall((:tp_jacobian, libtp()), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32, Int32, Int32, UInt32, Ptr{Float64}),
                p.ctx.handle, tab.handle, p.handle, turn, nturns, ckpt_stride, fast ? TP_TRACK_FAST : UInt32(0),
                jac))
    # row-major 6 x 6 per particle in C order -> J[q, i, j]
    return permutedims(reshape(jac, 6, 6, p.n), (3, 2, 1))
end

"""
    HIPBatch(ctx, tables, particles)

Lattice variants with their particle sets, tracked (`track!`) or differentiated
(`adjoint`) in one persistent graph launch.  The batch keeps its tables and particle sets
alive.
"""
mutable struct HIPBatch
    handle::Ptr{Cvoid}
    ctx::HIPContext
    tables::Vector{HIPTable}
    parts::Vector{HIPParticles}
    function HIPBatch(ctx::HIPContext, tables::Vector{HIPTable}, parts::Vector{HIPParticles})
        length(tables) == length(parts) || throw(DimensionMismatch("one particle set per table"))
        h = Ref{Ptr{Cvoid}}(C_NULL)
        th = Ptr{Cvoid}[t.handle for t in tables]
        ph = Ptr{Cvoid}[q.handle for q in parts]
        check(ccall((:tp_batch_create, libtp()), Cint, (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                    Ptr{Ptr{Cvoid}}), ctx.handle, length(tables), th, ph, h))
        b = new(h[], ctx, tables, parts)
        finalizer(b) do x
            ccall((:tp_batch_destroy, libtp()), Cint, (Ptr{Cvoid},), x.handle)
        end
        return b
    end
end

set_launch!(b::HIPBatch; block::Integer = 0, fwd_waves::Integer = 0, back_waves::Integer = 0) =
    check(ccall((:tp_batch_set_launch, libtp()), Cint, (Ptr{Cvoid}, Cint, Cint, Cint), b.handle, block,
                fwd_waves, back_waves))

"""
    track!(b::HIPBatch, nturns::Vector{Int32}; turn0 = nothing, fast = false)

Track variant v for `nturns[v]` turns from `turn0[v]`, all variants in one launch.
"""
function track!(b::HIPBatch, nturns::Vector{Int32}; turn0::Union{Nothing, Vector{Int32}} = nothing,
                fast::Bool = false)
    check(ccall((:tp_batch_track, libtp()), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}, UInt32),
                b.ctx.handle, b.handle, turn0 === nothing ? C_NULL : turn0, nturns,
                fast ? TP_TRACK_FAST : UInt32(0)))
    return b
end

"""
    batch_grad(b::HIPBatch, nturns; turn0 = nothing, ckpt_stride = 10, fast = false)
        -> (loss, n_alive, dL_dk per variant, dL_dx0 per variant)

`ringpass_grad` of every variant in one graph launch (forward with checkpoints, reverse
pass, per-variant reduction in tile order).
"""
function batch_grad(b::HIPBatch, nturns::Vector{Int32}; turn0::Union{Nothing, Vector{Int32}} = nothing,
                 ckpt_stride::Integer = 10, fast::Bool = false)
    nv = length(b.tables)
    loss = zeros(Float64, nv)
    na = zeros(Int64, nv)
    nseq = [t.n_seq for t in b.tables]
    ns = [q.n for q in b.parts]
    dk = zeros(Float64, sum(nseq))
    dx = zeros(Float64, 6 * sum(ns))
    check(ccall((:tp_batch_adjoint, libtp()), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int32}, Ptr{Int32}, Int32, UInt32, Ptr{Float64}, Ptr{Int64},
                 Ptr{Float64}, Ptr{Float64}),
                b.ctx.handle, b.handle, turn0 === nothing ? C_NULL : turn0, nturns, ckpt_stride,
                fast ? TP_TRACK_FAST : UInt32(0), loss, na, dk, dx))
    dks = Vector{Vector{Float64}}(undef, nv)
    dxs = Vector{Matrix{Float64}}(undef, nv)
    ok = 0
    ox = 0
    for v in 1:nv
        dks[v] = dk[ok + 1:ok + nseq[v]]
        dxs[v] = reshape(dx[ox + 1:ox + 6 * ns[v]], ns[v], 6)   # column-major N_v x 6 per variant
        ok += nseq[v]
        ox += 6 * ns[v]
    end
    return loss, na, dks, dxs
end

"""
    autotune!(b::HIPBatch, probe_turns; ckpt_stride = 10, adjoint = false, fast = false,
              cache = nothing) -> LaunchCfg

Block size and register budgets of the persistent kernels, timed on the batch itself with
turn counts `probe_turns[v]` (particle sets restored after every probe).
"""
function autotune!(b::HIPBatch, probe_turns::Vector{Int32}; ckpt_stride::Integer = 10, adjoint::Bool = false,
                   fast::Bool = false, cache::Union{Nothing, AbstractString} = nothing)
    out = Ref(LaunchCfg())
    fl = (fast ? TP_TRACK_FAST : UInt32(0)) | (adjoint ? TP_TUNE_ADJOINT : UInt32(0))
    check(ccall((:tp_batch_autotune, libtp()), Cint,
                (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int32}, Int32, UInt32, Cstring, Ref{LaunchCfg}),
                b.ctx.handle, b.handle, probe_turns, ckpt_stride, fl, cache === nothing ? C_NULL : cache, out))
    return out[]
end

include("table.jl")
include("madx.jl")

end # module TrackPadHIP
