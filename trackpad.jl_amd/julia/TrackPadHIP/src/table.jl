# table.jl -- Lattice -> device element table (the layout of include/trackpad_table.h).
#
# Every particle-independent constant the reference recomputes inside pass! (step lengths
# and kicks of symplectic4_pass, src/tracking.jl:324-328; kick-angle corrections :398-414,
# 670-683, 935-942; edge tangents and fringe corrections :160-221; exact-body cos/sin
# :875-876; RF constants :504-511) is formed here once, with the same Float64 operations,
# so the device reproduces the reference's values.  The Python writer
# (trackpad_hip/table.py) produces the same bytes; tests/golden/tables/ holds blobs of both
# checked lattices (test/runtests.jl compares them).

const TP_TABLE_MAGIC = UInt32(0x54504854)
const TP_TABLE_VERSION = UInt32(1)
const TAG = (MARKER = 0, DRIFT = 1, MPOLE = 2, THINMP = 3, SBEND = 4, ESBEND = 5, ESBEND_LE = 6, RFCAV = 7)
const F_T1, F_T2, F_R1, F_R2 = 0x1, 0x2, 0x4, 0x8
const F_FIN_SHIFT, F_FOUT_SHIFT = 4, 8
const F_DRIFTLIMIT, F_RF_DRIFT, F_RF_KICK = 0x1000, 0x2000, 0x4000
const MAXORD = 10
# The reference's decimal literals (src/tracking.jl:29-32).
const DRIFT1 = 0.6756035959798286638
const DRIFT2 = -0.1756035959798286639
const KICK1 = 1.351207191959657328
const KICK2 = -1.702414383919314656

"""
    ExactSBendLinearEdge(L, angle, e1 = 0, e2 = 0; fint1, fint2, gap, fringe_bend_entrance = 1,
                         fringe_bend_exit = 1, num_int_steps = 10, max_order = 0, polynom_a, polynom_b,
                         kick_angle, t1, t2)

The exact sector-bend body of `ExactSBend` (src/tracking.jl:923-953) between the linear edge
maps of `SBend` (src/tracking.jl:160-221).  No reference counterpart.
"""
Base.@kwdef struct ExactSBendLinearEdge <: TrackPad.AbstractMagnet
    L::Float64
    angle::Float64
    e1::Float64 = 0.0
    e2::Float64 = 0.0
    fint1::Float64 = 0.0
    fint2::Float64 = 0.0
    gap::Float64 = 0.0
    fringe_bend_entrance::Int = 1
    fringe_bend_exit::Int = 1
    num_int_steps::Int = 10
    max_order::Int = 0
    polynom_a::Vector{Float64} = zeros(4)
    polynom_b::Vector{Float64} = zeros(4)
    kick_angle::NTuple{2, Float64} = (0.0, 0.0)
    t1::NTuple{6, Float64} = ntuple(_ -> 0.0, 6)
    t2::NTuple{6, Float64} = ntuple(_ -> 0.0, 6)
end
ExactSBendLinearEdge(L, angle, e1 = 0.0, e2 = 0.0; kw...) =
    ExactSBendLinearEdge(; L = Float64(L), angle = Float64(angle), e1 = Float64(e1), e2 = Float64(e2), kw...)

"""
    WithMultipoles(elem, polynom_a, polynom_b, max_order)

Carry polynomials longer than the reference's `SVector{4}` (orders up to 10) for `elem`.
"""
struct WithMultipoles{E} <: AbstractElement
    elem::E
    polynom_a::Vector{Float64}
    polynom_b::Vector{Float64}
    max_order::Int
end

polys(e) = (collect(Float64, e.polynom_a), collect(Float64, e.polynom_b), e.max_order)
polys(w::WithMultipoles) = (w.polynom_a, w.polynom_b, w.max_order)
base(e) = e
base(w::WithMultipoles) = w.elem

function padded(a, b, m)
    (0 <= m <= MAXORD) || throw(ArgumentError("max_order $m outside [0, $MAXORD]"))
    A = zeros(MAXORD + 1); B = zeros(MAXORD + 1)
    A[1:length(a)] .= a
    B[1:length(b)] .= b
    return A, B
end

interleave(A, B, m) = vcat([[B[i], A[i]] for i in 1:m+1]...)

function kick_corrected!(A, B, e)
    if e.L > 0
        B[1] = B[1] - sin(e.kick_angle[1]) / e.L
        A[1] = A[1] + sin(e.kick_angle[2]) / e.L
    end
end

function yoshida(L, n)
    SL = L / n
    return SL * DRIFT1, SL * DRIFT2, SL * KICK1, SL * KICK2
end

function edge_block(irho, e, fint, gap, method)
    fc = (fint == 0 || gap == 0 || method == 0) ? 0.0 : irho * gap * fint * (1 + sin(e)^2) / cos(e)
    t = tan(e)
    return [irho * t, irho * t, e, fc]
end

function misalign(e)
    hasproperty(e, :t1) || return 0, nothing
    t1 = collect(Float64, e.t1); t2 = collect(Float64, e.t2)
    r1 = hasproperty(e, :r1) ? vec(collect(Float64, e.r1)) : zeros(36)
    r2 = hasproperty(e, :r2) ? vec(collect(Float64, e.r2)) : zeros(36)
    f = (any(!iszero, t1) ? F_T1 : 0) | (any(!iszero, t2) ? F_T2 : 0) |
        (any(!iszero, r1) ? F_R1 : 0) | (any(!iszero, r2) ? F_R2 : 0)
    f == 0 && return 0, nothing
    return f, vcat(t1, t2, r1, r2)          # SMatrix storage is column-major, as the device reads it
end

# (tag, nsteps, max_order, flags, gorder, block, misalignment block) of one element
function record(w, bt)
    e = base(w)
    e isa TimeVaryingElement && throw(ArgumentError(
        "time-varying element: call materialize_lattice first (src/time_dependence.jl:137-139)"))
    if e isa Marker
        return TAG.MARKER, 0, 0, 0, -1, Float64[], nothing
    elseif e isa Drift
        f, mb = misalign(e)
        return TAG.DRIFT, 0, 0, f, -1, [e.L, e.L * bt], mb
    elseif e isa Union{Quadrupole, Sextupole, Octupole}
        a, b, m = polys(w)
        A, B = padded(a, b, m)
        if e isa Quadrupole
            B[2] = e.k1; kick_corrected!(A, B, e); g = 1       # src/tracking.jl:395-414
        elseif e isa Sextupole
            B[3] = e.k2 / 2; g = 2                              # :446
        else
            B[4] = e.k3 / 6; g = 3                              # :478
        end
        n = e.num_int_steps
        L1, L2, K1, K2 = yoshida(e.L, n)
        f, mb = misalign(e)
        return TAG.MPOLE, n, m, f, (g > m ? -1 : g), vcat([L1, L2, K1, K2, L1 * bt, L2 * bt], interleave(A, B, m)), mb
    elseif e isa ThinMultipole
        a, b, m = polys(w)
        A, B = padded(a, b, m)
        f, mb = misalign(e)
        return TAG.THINMP, 0, m, f, -1, interleave(A, B, m), mb
    elseif e isa SBend
        a, b, m = polys(w)
        A, B = padded(a, b, m)
        kick_corrected!(A, B, e)
        irho = e.angle / e.L
        n = e.num_int_steps
        L1, L2, K1, K2 = yoshida(e.L, n)
        fin, fout = e.fringe_bend_entrance, e.fringe_bend_exit
        blk = vcat([L1, L2, K1, K2, L1 * bt, L2 * bt, irho, K1 * irho, K2 * irho],
                   edge_block(irho, e.e1, e.fint1, e.gap, fin), edge_block(irho, e.e2, e.fint2, e.gap, fout),
                   interleave(A, B, m))
        f, mb = misalign(e)
        return TAG.SBEND, n, m, f | (fin << F_FIN_SHIFT) | (fout << F_FOUT_SHIFT), -1, blk, mb
    elseif e isa Union{ExactSBend, ExactSBendLinearEdge}
        a, b, m = polys(w)
        A, B = padded(a, b, m)
        n = e.num_int_steps
        irho = e.angle / e.L
        if n == 0
            body = [irho, e.L, 0.0, 0.0, 0.0, e.L * bt, 0.0, cos(irho * e.L), sin(irho * e.L), 0.0, 0.0]
        else
            kick_corrected!(A, B, e)
            L1, L2, K1, K2 = yoshida(e.L, n)
            body = [irho, L1, L2, K1, K2, L1 * bt, L2 * bt, cos(irho * L1), sin(irho * L1), cos(irho * L2),
                    sin(irho * L2)]
        end
        fin, fout = e.fringe_bend_entrance, e.fringe_bend_exit
        f = (fin << F_FIN_SHIFT) | (fout << F_FOUT_SHIFT) | (abs(irho) < 1e-6 ? F_DRIFTLIMIT : 0)
        if e isa ExactSBend
            t1, t2 = -e.e1, -e.e2
            edges = [cos(e.e1), sin(e.e1), cos(t1), sin(t1), sin(2 * t1), t1, cos(t2), sin(t2), sin(2 * t2), t2,
                     cos(e.e2), sin(e.e2), e.gk]
            tag = TAG.ESBEND
        else
            edges = vcat(edge_block(irho, e.e1, e.fint1, e.gap, fin), edge_block(irho, e.e2, e.fint2, e.gap, fout),
                         zeros(5))
            tag = TAG.ESBEND_LE
        end
        fm, mb = misalign(e)
        return tag, n, m, f | fm, -1, vcat(body, edges, interleave(A, B, m)), mb
    elseif e isa RFCavity
        beta = inv(bt)
        f = (e.L > 0 ? F_RF_DRIFT : 0) | (e.energy > 0 ? F_RF_KICK : 0)
        nv = e.energy > 0 ? e.volt / e.energy : 0.0
        return TAG.RFCAV, 0, 0, f, -1, [e.L / 2, (e.L / 2) * bt, nv, 2pi * e.freq, e.lag, e.philag, beta * beta], nothing
    end
    throw(ArgumentError("element type $(typeof(e)) has no device map"))
end

"""
    compile_table(lat, beam; aperture = nothing, exact_hamiltonian = true,
                  coord_limit = 1.0, angle_limit = 1.0) -> Vector{UInt8}

Compile a static lattice into the device table blob.
"""
function compile_table(lat::Lattice, beam::Beam; aperture = nothing, exact_hamiltonian::Bool = true,
                       coord_limit::Real = 1.0, angle_limit::Real = 1.0)
    bt = TrackPad.beti(beam)
    pool = Float64[]
    recs = NTuple{8, Int32}[]
    index = Dict{Any, Int}()
    seq = Int32[]
    for w in lat.elements
        tag, n, m, f, g, blk, mb = record(w, bt)
        key = (tag, n, m, f, g, reinterpret(UInt64, Float64.(blk)), mb === nothing ? nothing : reinterpret(UInt64, mb))
        idx = get(index, key, -1)
        if idx < 0
            off = length(pool)
            append!(pool, blk)
            moff = -1
            if mb !== nothing
                moff = length(pool)
                append!(pool, mb)
            end
            push!(recs, (Int32(tag), Int32(n), Int32(m), Int32(off), reinterpret(Int32, UInt32(f)), Int32(moff),
                         Int32(g), Int32(0)))
            idx = length(recs) - 1
            index[key] = idx
        end
        push!(seq, Int32(idx))
    end
    io = IOBuffer()
    write(io, TP_TABLE_MAGIC, TP_TABLE_VERSION, UInt32(length(seq)), UInt32(length(recs)), UInt32(length(pool)),
          UInt32(exact_hamiltonian ? 1 : 0), UInt32(0), UInt32(0))
    write(io, Float64(bt), Float64(coord_limit), Float64(angle_limit),
          aperture === nothing ? 0.0 : Float64(aperture)^2)
    for s in seq
        write(io, s)
    end
    isodd(length(seq)) && write(io, Int32(0))
    for r in recs
        for v in r
            write(io, v)
        end
    end
    for v in pool
        write(io, v)
    end
    return take!(io)
end
