# table.jl -- Lattice -> device element table (the layout of include/trackpad_table.h).
#
# Every particle-independent constant the reference recomputes inside pass! (step lengths
# and kicks of symplectic4_pass, src/tracking.jl:324-328; kick-angle corrections :398-414,
# 670-683, 935-942; edge tangents and fringe corrections :160-221; exact-body cos/sin
# :875-876; RF constants :504-511; the other families :529-605, 981-1218, 1224-1525,
# 1531-1663) is formed here once, with the same Float64 operations, so the device reproduces
# the reference's values.  The Python writer (trackpad_hip/table.py) produces the same bytes,
# record for record and in the same record order; tests/golden/tables/ holds its blobs and
# test/runtests.jl compares this writer with them.

const TP_TABLE_MAGIC = UInt32(0x54504854)
const TP_TABLE_VERSION = UInt32(1)
const TAG = (MARKER = 0, DRIFT = 1, MPOLE = 2, THINMP = 3, SBEND = 4, ESBEND = 5, ESBEND_LE = 6, RFCAV = 7,
             SOLENOID = 8, CORRECTOR = 9, TRANSLATION = 10, YROTATION = 11, CRAB = 12, ACCEL = 13, LONGRF = 14,
             BOOST = 15, INVBOOST = 16, GAUSSBB = 17, RLCWAKE = 18, TABWAKE = 19, LBEND = 20, NAN = 21,
             WIGGLER = 22)
const F_T1, F_T2, F_R1, F_R2 = 0x1, 0x2, 0x4, 0x8
const F_FIN_SHIFT, F_FOUT_SHIFT = 4, 8
const F_DRIFTLIMIT, F_RF_DRIFT, F_RF_KICK = 0x1000, 0x2000, 0x4000
const H_EXACT_HAMILTONIAN = UInt32(0x1)
const MAXORD = 10
const EPS = eps(Float64)
# The reference's decimal literals (src/tracking.jl:29-32).
const DRIFT1 = 0.6756035959798286638
const DRIFT2 = -0.1756035959798286639
const KICK1 = 1.351207191959657328
const KICK2 = -1.702414383919314656
const C_LIGHT = 2.99792458e8            # src/elements/rf.jl:28
const TWO_PI = 2pi                      # Float64(2pi), the reference's T(2pi)

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
    polynom_a::NTuple{4, Float64} = (0.0, 0.0, 0.0, 0.0)
    polynom_b::NTuple{4, Float64} = (0.0, 0.0, 0.0, 0.0)
    kick_angle::NTuple{2, Float64} = (0.0, 0.0)
    t1::NTuple{6, Float64} = ntuple(_ -> 0.0, 6)
    t2::NTuple{6, Float64} = ntuple(_ -> 0.0, 6)
    name::Symbol = :EXACTSBENDLINEAREDGE
end
TrackPad.get_length(e::ExactSBendLinearEdge) = e.L
# Immutable fields only (as the reference's SVector{4} polynomials), so an unchanged element
# is recognised by identity; orders beyond 3 go through WithMultipoles.
ExactSBendLinearEdge(L, angle, e1 = 0.0, e2 = 0.0; polynom_a = zeros(4), polynom_b = zeros(4),
                     kick_angle = (0.0, 0.0), t1 = zeros(6), t2 = zeros(6), name = :EXACTSBENDLINEAREDGE, kw...) =
    ExactSBendLinearEdge(; L = Float64(L), angle = Float64(angle), e1 = Float64(e1), e2 = Float64(e2),
                         polynom_a = NTuple{4, Float64}(polynom_a), polynom_b = NTuple{4, Float64}(polynom_b),
                         kick_angle = NTuple{2, Float64}(kick_angle), t1 = NTuple{6, Float64}(t1),
                         t2 = NTuple{6, Float64}(t2), name = Symbol(name), kw...)

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

TrackPad.get_length(w::WithMultipoles) = TrackPad.get_length(w.elem)

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

interleave(A, B, m) = Float64[isodd(j) ? B[(j + 1) ÷ 2] : A[j ÷ 2] for j in 1:2 * (m + 1)]

function kick_corrected!(A, B, e)
    if e.L > 0
        B[1] = B[1] - sin(e.kick_angle[1]) / e.L
        A[1] = A[1] + sin(e.kick_angle[2]) / e.L
    end
end

function yoshida(L, n)
    SL = L / n                          # n == 0: +-Inf or NaN, as the reference's L / 0
    return SL * DRIFT1, SL * DRIFT2, SL * KICK1, SL * KICK2
end

function edge_block(irho, e, fint, gap, method)
    fc = (fint == 0 || gap == 0 || method == 0) ? 0.0 : irho * gap * fint * (1 + sin(e)^2) / cos(e)
    t = tan(e)
    return [irho * t, irho * t, e, fc]
end

function fringe_method(v, what)
    (0 <= v <= 15) || throw(ArgumentError("$what = $v is outside the supported range 0..15"))
    return Int(v)
end

function check_steps(n, e)
    n >= 0 || throw(ArgumentError("$(typeof(e)): num_int_steps $n < 0"))
    return Int(n)
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

const NO_MISALIGN = nothing

# (tag, nsteps, max_order, flags, gorder, block, misalignment block) of one element
function record(w, bt)
    e = base(w)
    e isa TimeVaryingElement && throw(ArgumentError(
        "time-varying element: compile_table resolves it per turn when given the turn range " *
        "(turn, nturns, time, dt_turn), or call materialize_lattice first (src/time_dependence.jl:137-139)"))
    if e isa Union{Marker, SpaceCharge}            # SpaceCharge: identity (src/tracking.jl:1176-1180)
        return TAG.MARKER, 0, 0, 0, -1, Float64[], NO_MISALIGN
    elseif e isa Union{Drift, DriftSC}
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
        n = check_steps(e.num_int_steps, e)
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
        fin = fringe_method(e.fringe_bend_entrance, "fringe_bend_entrance")
        fout = fringe_method(e.fringe_bend_exit, "fringe_bend_exit")
        blk = vcat([L1, L2, K1, K2, L1 * bt, L2 * bt, irho, K1 * irho, K2 * irho],
                   edge_block(irho, e.e1, e.fint1, e.gap, fin), edge_block(irho, e.e2, e.fint2, e.gap, fout),
                   interleave(A, B, m))
        f, mb = misalign(e)
        return TAG.SBEND, n, m, f | (fin << F_FIN_SHIFT) | (fout << F_FOUT_SHIFT), -1, blk, mb
    elseif e isa Union{ExactSBend, ExactSBendLinearEdge}
        a, b, m = polys(w)
        A, B = padded(a, b, m)
        n = check_steps(e.num_int_steps, e)
        irho = e.angle / e.L
        if n == 0
            body = [irho, e.L, 0.0, 0.0, 0.0, e.L * bt, 0.0, cos(irho * e.L), sin(irho * e.L), 0.0, 0.0]
        else
            kick_corrected!(A, B, e)
            L1, L2, K1, K2 = yoshida(e.L, n)
            body = [irho, L1, L2, K1, K2, L1 * bt, L2 * bt, cos(irho * L1), sin(irho * L1), cos(irho * L2),
                    sin(irho * L2)]
        end
        fin = fringe_method(e.fringe_bend_entrance, "fringe_bend_entrance")
        fout = fringe_method(e.fringe_bend_exit, "fringe_bend_exit")
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
    end
    extra = record_extra(e, bt)
    extra === nothing || return extra
    if e isa RFCavity
        beta = inv(bt)                                          # src/tracking.jl:504
        f = (e.L > 0 ? F_RF_DRIFT : 0) | (e.energy > 0 ? F_RF_KICK : 0)
        nv = e.energy > 0 ? e.volt / e.energy : 0.0
        return TAG.RFCAV, 0, 0, f, -1, [e.L / 2, (e.L / 2) * bt, nv, TWO_PI * e.freq, e.lag, e.philag, beta * beta],
               NO_MISALIGN
    end
    throw(ArgumentError("element type $(typeof(e)) has no device map"))
end

# polynom (k0, k1, k2/2, k3/6) of the SC multipoles (src/tracking.jl:993-995), max_order
# clamped to [0, 3] (:1008)
function sc_poly(e)
    A = zeros(MAXORD + 1); B = zeros(MAXORD + 1)
    A[1:4] .= collect(Float64, e.polynom_a)
    B[1:4] .= (e.k0, e.k1, e.k2 / 2, e.k3 / 6)
    return A, B, clamp(e.max_order, 0, 3)
end

# Records of the remaining element families (src/tracking.jl:529-605, 981-1218, 1531-1663),
# or nothing.
function record_extra(e, bt)
    if e isa Union{QuadrupoleSC, SextupoleSC, OctupoleSC}
        A, B, m = sc_poly(e)
        e isa QuadrupoleSC && kick_corrected!(A, B, e)         # src/tracking.jl:1004-1007
        n = check_steps(e.num_int_steps, e)
        L1, L2, K1, K2 = yoshida(e.L, n)
        f, mb = misalign(e)
        return TAG.MPOLE, n, m, f, -1, vcat([L1, L2, K1, K2, L1 * bt, L2 * bt], interleave(A, B, m)), mb
    elseif e isa SBendSC
        f, mb = misalign(e)
        e.L == 0 && return TAG.NAN, 0, 0, 0, -1, Float64[], NO_MISALIGN     # src/tracking.jl:1047-1049
        A = zeros(MAXORD + 1); B = zeros(MAXORD + 1)
        A[1:4] .= collect(Float64, e.polynom_a)
        B[1:4] .= collect(Float64, e.polynom_b)
        m = clamp(e.max_order, 0, 3)
        kick_corrected!(A, B, e)
        irho = e.angle / e.L
        n = e.num_int_steps
        L1, L2, K1, K2 = yoshida(e.L, n)
        fin = fringe_method(e.fringe_bend_entrance, "fringe_bend_entrance")
        fout = fringe_method(e.fringe_bend_exit, "fringe_bend_exit")
        blk = vcat([L1, L2, K1, K2, L1 * bt, L2 * bt, irho, K1 * irho, K2 * irho],
                   edge_block(irho, e.e1, e.fint1, e.gap, fin), edge_block(irho, e.e2, e.fint2, e.gap, fout),
                   interleave(A, B, m))
        return TAG.SBEND, n, m, f | (fin << F_FIN_SHIFT) | (fout << F_FOUT_SHIFT), -1, blk, mb
    elseif e isa Solenoid
        f, mb = misalign(e)
        e.ks == 0 && return TAG.DRIFT, 0, 0, f, -1, [e.L, e.L * bt], mb      # src/tracking.jl:539-540
        return TAG.SOLENOID, 0, 0, f, -1, [e.L, e.ks], mb
    elseif e isa Corrector
        f, mb = misalign(e)
        xk, yk = e.xkick, e.ykick                                           # src/tracking.jl:588-596
        return TAG.CORRECTOR, 0, 0, f, -1, [e.L, xk, yk, xk^2 / 3 + yk^2 / 3, xk / 2, yk / 2], mb
    elseif e isa Translation
        return TAG.TRANSLATION, 0, 0, 0, -1, [e.dx, e.dy, e.ds], NO_MISALIGN
    elseif e isa YRotation
        ang = -e.angle                                                      # src/tracking.jl:1195
        ang == 0 && return TAG.MARKER, 0, 0, 0, -1, Float64[], NO_MISALIGN
        return TAG.YROTATION, 0, 0, 0, -1, [cos(ang), sin(ang), tan(ang)], NO_MISALIGN
    elseif e isa CrabCavity
        volt = e.volt * (1 + e.errors[1])                                   # src/tracking.jl:1532-1534
        phi = e.phi + e.errors[2]
        E = max(abs(e.energy), EPS)
        blk = [e.L / 2, (e.L / 2) * bt, volt / E, e.k, phi, e.k * volt / E * bt]
        return TAG.CRAB, 0, 0, (e.L > 0 ? F_RF_DRIFT : 0), -1, blk, NO_MISALIGN
    elseif e isa AccelCavity
        beta = inv(bt)                                                      # src/tracking.jl:1550-1558
        beta2 = beta * beta
        beta2 <= EPS && return TAG.MARKER, 0, 0, 0, -1, Float64[], NO_MISALIGN
        E = max(abs(e.energy), EPS)
        return TAG.ACCEL, 0, 0, 0, -1, [e.volt / (beta2 * E), e.k, e.phis, sin(e.phis)], NO_MISALIGN
    elseif e isa LongitudinalRFMap
        k = Float64(TrackPad._rf_k(e.rf))                                   # src/tracking.jl:1575-1585
        abs(k) <= EPS && return TAG.MARKER, 0, 0, 0, -1, Float64[], NO_MISALIGN
        h = Float64(TrackPad._rf_h(e.rf))
        beta = inv(bt)
        eta = e.alphac - (1 - beta * beta)
        return TAG.LONGRF, 0, 0, 0, -1, [TWO_PI * h * eta / k], NO_MISALIGN
    elseif e isa LorentzBoost
        e.mode != 0 && return TAG.MARKER, 0, 0, 0, -1, Float64[], NO_MISALIGN
        return TAG.BOOST, 0, 0, 0, -1, [e.tanang, inv(e.cosang)], NO_MISALIGN
    elseif e isa InvLorentzBoost
        e.mode != 0 && return TAG.MARKER, 0, 0, 0, -1, Float64[], NO_MISALIGN
        return TAG.INVBOOST, 0, 0, 0, -1, [e.sinang, e.cosang], NO_MISALIGN
    elseif e isa StrongThinGaussianBeam
        sx2 = max(e.rmssizex^2, EPS)                                        # src/tracking.jl:1611-1612
        sy2 = max(e.rmssizey^2, EPS)
        return TAG.GAUSSBB, 1, 0, 0, -1, [sx2, sy2, 1.0, e.amplitude, e.xoffset, e.yoffset], NO_MISALIGN
    elseif e isa StrongGaussianBeam
        n = e.nzslice
        n <= 0 && return TAG.MARKER, 0, 0, 0, -1, Float64[], NO_MISALIGN
        sx2 = max(e.beamsize[1]^2, EPS)                                     # src/tracking.jl:1625-1627
        sy2 = max(e.beamsize[2]^2, EPS)
        blk = [sx2, sy2, e.charge / max(abs(e.total_energy), EPS)]
        for i in 1:n
            push!(blk, i <= length(e.zslice_npar) ? e.zslice_npar[i] : 1.0 / n,
                  i <= length(e.xoffsets) ? e.xoffsets[i] : 0.0,
                  i <= length(e.yoffsets) ? e.yoffsets[i] : 0.0)
        end
        return TAG.GAUSSBB, n, 0, 0, -1, blk, NO_MISALIGN
    elseif e isa LongitudinalRLCWake
        e.scale == 0 && return TAG.MARKER, 0, 0, 0, -1, Float64[], NO_MISALIGN
        # src/elements/wake.jl:27-36 (a Q0 below 1/2 gives NaN here where sqrt would throw)
        Q0p = e.Q0^2 >= 0.25 ? sqrt(e.Q0^2 - 0.25) : NaN
        w0 = TWO_PI * e.freq
        w0p = w0 / e.Q0 * Q0p
        return TAG.RLCWAKE, 0, 0, 0, -1, [e.scale, e.Rshunt * w0 / e.Q0, w0p, 2 * Q0p, w0, 2 * e.Q0], NO_MISALIGN
    elseif e isa LongitudinalWake
        e.scale == 0 && return TAG.MARKER, 0, 0, 0, -1, Float64[], NO_MISALIGN
        n = length(e.times)
        blk = vcat([e.scale, e.fliphalf, e.times[1] * e.fliphalf], collect(Float64, e.times),
                   collect(Float64, e.wakefields))
        return TAG.TABWAKE, n, 0, 0, -1, blk, NO_MISALIGN
    elseif e isa Wiggler
        return wiggler_record(e)
    elseif e isa LBend
        f, mb = misalign(e)
        e.L == 0 && return TAG.NAN, 0, 0, 0, -1, Float64[], NO_MISALIGN     # src/tracking.jl:1158-1160
        irho = e.angle / e.L
        blk = vcat([e.L, e.K, e.angle, e.by_error, e.angle / e.L],
                   edge_block(irho, e.e1, e.fint1, e.full_gap, 1), edge_block(irho, e.e2, e.fint2, e.full_gap, 1))
        return TAG.LBEND, 0, 0, f | (1 << F_FIN_SHIFT) | (1 << F_FOUT_SHIFT), -1, blk, mb
    end
    return nothing
end

# TP_WIGGLER: the step plan of _wig_pass_4th / _wig_pass_4th_rad and the particle-independent
# leading factors of every harmonic term (src/tracking.jl:1224-1525).
function wiggler_record(e)
    PN = e.Nsteps
    Nw = round(Int, e.L / e.lw)
    Nstep = PN * max(1, Nw)
    dl = e.lw / Float64(PN)
    dl1 = dl * 1.3512071919596573
    dl0 = dl * -1.7024143839193146
    gamma = e.energy / TrackPad.M_ELECTRON
    Po = sqrt(gamma^2 - 1)
    Aw = 1e-9 * C_LIGHT / (TrackPad.M_ELECTRON * 1e-9) / TWO_PI * e.lw * e.Bmax
    srCoef = (2.0 / 3.0) * 2.8179403205e-15 * gamma^3
    H = Po * TrackPad.M_ELECTRON / C_LIGHT
    kw = TWO_PI / e.lw
    PB0 = e.Bmax
    nh = min(e.NHharm, length(e.By) ÷ 6)
    nv = min(e.NVharm, length(e.Bx) ÷ 6)
    blk = Float64[Nstep, dl1, dl0, nh, nv, e.rad, H, srCoef, dl]
    for i in 0:nh-1
        b = Float64.(e.By[6i+1:6i+6])
        HCw = b[2] * Aw / Po
        kx, ky, kz, tz = b[3] * kw, b[4] * kw, b[5] * kw, b[6]
        a = HCw * (kw / kz)
        append!(blk, [kx, ky, kz, tz,
                      a, a * ky,                                    # ax, axpy
                      a * (kx / ky), a * ((kx / ky) * (kx / ky)),   # ay, aypx
                      PB0 * b[2] * (kx / ky), PB0 * b[2],           # Bx, By of _wig_B
                      abs(kx / kw) > 1e-6 ? 1.0 : 0.0])
    end
    for i in 0:nv-1
        b = Float64.(e.Bx[6i+1:6i+6])
        VCw = b[2] * Aw / Po
        kx, ky, kz, tz = b[3] * kw, b[4] * kw, b[5] * kw, b[6]
        a = VCw * (kw / kz)
        append!(blk, [kx, ky, kz, tz,
                      a * (ky / kx), a * ((ky / kx) * (ky / kx)),   # ax, axpy
                      a, a * kx,                                    # ay, aypx
                      PB0 * b[2], PB0 * b[2] * (ky / kx),
                      abs(ky / kw) > 1e-6 ? 1.0 : 0.0])
    end
    f, mb = misalign(e)
    return TAG.WIGGLER, Nstep, 0, f, -1, blk, mb
end

"""
    TableInfo

What the writer produced for the first turn row: the record of each lattice position and
its main-strength polynomial index (-1 where the position has no k1 / k2 / k3).
"""
mutable struct TableInfo
    seq::Vector{Int32}
    gorder::Vector{Int32}
    n_rec::Int
    pool_len::Int
end
TableInfo() = TableInfo(Int32[], Int32[], 0, 0)

"""
    compile_table(lat, beam; aperture = nothing, exact_hamiltonian = true, coord_limit = 1.0,
                  angle_limit = 1.0, turn = 0, nturns = nothing, time = 0.0, dt_turn = 0.0,
                  info = nothing) -> Vector{UInt8}

Compile a lattice into the device table blob.

A lattice with `TimeVaryingElement`s is resolved per turn, as `ringpass!` resolves it in every
`linepass!` call (src/lattice.jl:267, 305-317): row k of the table is the lattice materialised
at `TimeContext(time + k dt_turn, turn + k)` (the time accumulated by repeated addition, as
`ringpass!` does), for `nturns` rows (one row at (time, turn) when `nturns` is `nothing`).
Static elements share one record across rows; identical materialised elements share records
too.  Records are created in the Python writer's order (static positions first, then each
row's time-varying positions), so both writers emit the same bytes.
"""
function compile_table(lat::Lattice, beam::Beam; aperture = nothing, exact_hamiltonian::Bool = true,
                       coord_limit::Real = 1.0, angle_limit::Real = 1.0, turn::Integer = 0,
                       nturns::Union{Nothing, Integer} = nothing, time::Real = 0.0, dt_turn::Real = 0.0,
                       info::Union{Nothing, TableInfo} = nothing)
    bt = Float64(TrackPad.beti(beam))
    pool = Float64[]
    recs = NTuple{8, Int32}[]
    index = Dict{Any, Int}()
    function rec_of(w)
        tag, n, m, f, g, blk, mb = record(w, bt)
        blk = Vector{Float64}(blk)
        key = (tag, n, m, f, g, collect(reinterpret(UInt64, blk)),
               mb === nothing ? nothing : collect(reinterpret(UInt64, Vector{Float64}(mb))))
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
        return idx, g
    end
    els = lat.elements
    varying = [base(w) isa TimeVaryingElement for w in els]
    rows = (!any(varying) || nturns === nothing) ? 1 : max(Int(nturns), 1)
    static = [v ? nothing : rec_of(w) for (w, v) in zip(els, varying)]
    seq = Int32[]
    t = Float64(time)
    if info !== nothing
        empty!(info.seq); empty!(info.gorder)
    end
    for k in 0:rows-1
        ctx = TrackPad.TimeContext(t; turn = Int(turn) + k)
        for (w, st) in zip(els, static)
            r = st === nothing ? rec_of(TrackPad.materialize(base(w), ctx)) : st
            push!(seq, Int32(r[1]))
            if info !== nothing && k == 0
                push!(info.seq, Int32(r[1])); push!(info.gorder, Int32(r[2]))
            end
        end
        t += Float64(dt_turn)
    end
    n_rows, row_turn0 = rows > 1 ? (rows, Int(turn)) : (0, 0)
    if info !== nothing
        info.n_rec = length(recs); info.pool_len = length(pool)
    end
    io = IOBuffer()
    write(io, TP_TABLE_MAGIC, TP_TABLE_VERSION, UInt32(length(els)), UInt32(length(recs)), UInt32(length(pool)),
          exact_hamiltonian ? H_EXACT_HAMILTONIAN : UInt32(0), UInt32(n_rows), Int32(row_turn0))
    write(io, bt, Float64(coord_limit), Float64(angle_limit), aperture === nothing ? 0.0 : Float64(aperture)^2)
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
