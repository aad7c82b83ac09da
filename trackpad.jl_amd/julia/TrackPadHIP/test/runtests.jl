# Checks the Julia table writer against the blobs the Python writer committed
# (tests/golden/make_tables.py), and (with a GPU) the device path against TrackPad's own CPU
# linepass! / ringpass! on the reference's test input, the table cache, the time-varying rows
# and the AD rules.
using Test
using TrackPad
using TrackPadHIP

const GOLDEN = joinpath(@__DIR__, "..", "..", "..", "..", "tests", "golden", "tables")

particles_initial = [
    1.0e-4   2.0e-4   3.0e-4  -1.0e-4   5.0e-5   2.0e-4
   -2.2e-4  1.7e-4  -1.1e-4   2.3e-4  -7.0e-5   1.0e-4
    3.5e-4  -2.1e-4  8.0e-5  -1.8e-4   1.4e-4  -2.6e-4
   -4.0e-4  2.9e-4   1.6e-4   9.0e-5  -1.2e-4   3.1e-4
    5.2e-4  -3.3e-4 -2.4e-4   1.1e-4   2.6e-4  -3.7e-4
   -6.1e-4  4.4e-4   2.7e-4  -2.5e-4  -3.0e-4   4.5e-4
    7.0e-4  -5.2e-4 -3.1e-4   3.4e-4   3.3e-4  -5.0e-4
   -8.0e-4  6.0e-4   3.9e-4  -4.2e-4  -3.8e-4   5.8e-4
    9.1e-4  -6.7e-4 -4.6e-4   5.1e-4   4.4e-4  -6.3e-4
   -9.8e-4  7.3e-4   5.3e-4  -5.9e-4  -4.9e-4   7.1e-4
]

verify_dba_line() = Lattice([Drift(2.0), SBend(1.0, pi/9, pi/18, pi/18), Drift(0.5),
                             Quadrupole(0.4, 1.5; num_int_steps = 10), Drift(0.5),
                             Quadrupole(0.4, -1.2; num_int_steps = 10), Drift(0.5),
                             Quadrupole(0.4, 1.5; num_int_steps = 10), Drift(0.5),
                             SBend(1.0, pi/9, pi/18, pi/18), Drift(2.0)])

# 6 x 6 roll about the s axis (tests/lattices.py tilt)
function tilt(theta)
    R = zeros(6, 6)
    for i in 1:6
        R[i, i] = 1.0
    end
    c, s = cos(theta), sin(theta)
    for (i, j) in ((1, 3), (2, 4))
        R[i, i] = c; R[i, j] = s; R[j, i] = -s; R[j, j] = c
    end
    return R
end

# tests/lattices.py families_ring: every element family beyond the hot lattices
function families_ring()
    q(k) = Quadrupole(0.4, k; num_int_steps = 8)
    rolled = Quadrupole(0.4, 1.2; num_int_steps = 8, r1 = tilt(2e-3), r2 = tilt(-2e-3) * 1.0001,
                        t1 = [1e-4, 0, -5e-5, 0, 0, 0], t2 = [-1e-4, 0, 5e-5, 0, 0, 0])
    els = AbstractElement[
        q(1.2), Drift(1.0), Solenoid(0.4, 0.3; t1 = [1e-4, 0, -2e-4, 0, 0, 0], t2 = [-1e-4, 0, 2e-4, 0, 0, 0]),
        Drift(0.5), q(-1.2), Drift(1.0), Corrector(0.2, 1e-5, -2e-5),
        LBend(0.8, 0.05, 0.01, 0.02; K = 0.1, fint1 = 0.3, fint2 = 0.3, full_gap = 0.02), Drift(0.5), rolled,
        Drift(1.0), QuadrupoleSC(0.4; k1 = -1.2), DriftSC(1.0), SextupoleSC(0.2; k2 = 3.0),
        SBendSC(0.8, 0.05, 0.025, 0.025), Drift(0.5), LorentzBoost(0.01),
        StrongThinGaussianBeam(1e-7, 1e-3, 1e-3; xoffset = 1e-4), InvLorentzBoost(0.01),
        StrongGaussianBeam(3e-3, M_ELECTRON, 1.0, 10^6, 3e9, [1e-3, 1.2e-3]; nzslice = 3,
                           xoffsets = [0, 1e-4, -1e-4], zslice_npar = [0.2, 0.5, 0.3]),
        Translation(0.0; dx = 1e-4, dy = -2e-4, ds = 1e-3), YRotation(0.0; angle = 1e-3), Drift(0.3),
        YRotation(0.0; angle = -1.5e-3), CrabCavity(0.3; volt = 1e5, freq = 5e8, phi = 0.1, energy = 3e9),
        AccelCavity(0.0; volt = 1e6, freq = 5e8, h = 400.0, phis = 0.1, energy = 3e9),
        LongitudinalRFMap(1e-3, AccelCavity(0.0; volt = 1e6, freq = 5e8, h = 400.0)),
        LongitudinalRLCWake(; freq = 1e9, Rshunt = 1e4, Q0 = 2.0, scale = 1e-20),
        LongitudinalWake([-4e-12, -1e-12, 0.0, 2e-12, 5e-12], [0.5, -1.0, 3.0, 5.0, 1.0]; scale = 1e-7),
        SpaceCharge(0.0)]
    return Lattice(els)
end

# trackpad_hip/workloads.py dba_cell / dba_ring: 20 DBA cells + one RF cavity (501 elements)
const ENERGY = 3.0e9
function dba_cell(ncell = 20; steps = 10)
    th = 2.0 * pi / (2 * ncell)
    B(nm) = ExactSBendLinearEdge(1.5, th, 0.0, 0.0; fint1 = 0.5, fint2 = 0.5, gap = 0.03, num_int_steps = steps,
                                 name = nm)
    Q(L, k, nm) = Quadrupole(L, k; num_int_steps = steps, name = nm)
    S(k, nm) = Sextupole(0.2, k; num_int_steps = steps, name = nm)
    kf1, kd1, kf2, ksd, ksf = 2.1, -2.4, 0.5, -32.0, 27.0
    return AbstractElement[
        Drift(2.0; name = "DL"), Q(0.3, kf1, "QF1"), Drift(0.25; name = "D1"), Q(0.3, kd1, "QD1"),
        Drift(0.4; name = "D2"), B("B1"), Drift(0.3; name = "D3"), S(ksd, "SD"), Drift(0.2; name = "D4"),
        Q(0.4, kf2, "QF2"), Drift(0.2; name = "D5"), S(ksf, "SF"), Drift(0.1; name = "DM"), S(ksf, "SF"),
        Drift(0.2; name = "D5"), Q(0.4, kf2, "QF2"), Drift(0.2; name = "D4"), S(ksd, "SD"),
        Drift(0.3; name = "D3"), B("B2"), Drift(0.4; name = "D2"), Q(0.3, kd1, "QD1"), Drift(0.25; name = "D1"),
        Q(0.3, kf1, "QF1"), Drift(2.0; name = "DL")]
end
function dba_ring(beam = Beam(ENERGY); ncell = 20)
    els = AbstractElement[]
    for _ in 1:ncell
        append!(els, dba_cell(ncell))
    end
    L = total_length(Lattice(els))
    h = 400.0
    freq = h * beam.beta * 2.99792458e8 / L
    push!(els, RFCavity(0.0, 1.0e6, freq, 0.0; h = h, philag = 0.0, energy = ENERGY, name = "RF"))
    return Lattice(els; name = :DBA)
end

# tests/golden/make_tables.py dba_ring_time_varying and its turn range
const TV_TURN, TV_NTURNS, TV_TIME, TV_DT = 5, 4, 2e-4, 1e-6
function dba_ring_time_varying()
    els = AbstractElement[e for e in dba_ring().elements]
    for (i, e) in enumerate(els)
        if e isa Sextupole
            k0 = e.k2
            els[i] = timed(e; k2 = ctx -> k0 * (1.0 + 0.002 * ctx.turn))
        end
    end
    q = els[2]
    k1 = q.k1
    els[2] = timed(q; k1 = ctx -> k1 * (1.0 + 1e-3 * sin(2 * pi * 50.0 * ctx.real_time)))
    return Lattice(els; name = :DBA_TV)
end

# header, sequence and records byte for byte; pool doubles within 2 ulp (Julia's and glibc's
# sin / cos / tan are each within 1 ulp of the exact value and may differ in the last place)
# (rtol: pool values formed from lengths that a sequence gives as differences of positions,
# where a last-place difference of sin is amplified by the subtraction)
function same_table(blob, gold; rtol = 0.0)
    length(blob) == length(gold) || return false
    npool = Int(reinterpret(UInt32, gold[17:20])[1])
    head = length(gold) - 8 * npool
    blob[1:head] == gold[1:head] || return false
    a = reinterpret(Float64, blob[head+1:end]); b = reinterpret(Float64, gold[head+1:end])
    return all(abs.(a .- b) .<= max.(2 .* eps.(abs.(b)), rtol .* abs.(b)))
end
golden(name) = read(joinpath(GOLDEN, name * ".bin"))
const GOLDEN_MADX = joinpath(@__DIR__, "..", "..", "..", "..", "tests", "golden", "madx")

@testset "table writer matches the committed blobs" begin
    @test same_table(compile_table(verify_dba_line(), Beam(3.5e9)), golden("verify_dba_line"))
    @test same_table(compile_table(families_ring(), Beam(3e9)), golden("families_ring"))
    @test same_table(compile_table(dba_ring(), Beam(ENERGY)), golden("dba_workload"))
    @test same_table(compile_table(dba_ring_time_varying(), Beam(ENERGY); turn = TV_TURN, nturns = TV_NTURNS,
                                   time = TV_TIME, dt_turn = TV_DT), golden("dba_time_varying"))
    # without a turn range a time-varying lattice is one row (header rows = 0) at (time, turn)
    blob = compile_table(dba_ring_time_varying(), Beam(ENERGY); time = TV_TIME, turn = TV_TURN)
    @test reinterpret(UInt32, blob[25:28])[1] == 0
    @test length(blob) < length(golden("dba_time_varying"))
end

@testset "MAD-X reader gives the Python reader's tables" begin
    # the DBA workload ring as the Python writer prints it (every value reads back exactly)
    lat, beam = read_madx(joinpath(GOLDEN_MADX, "dba_workload.madx"); energy = ENERGY)
    @test length(lat.elements) == 501
    @test same_table(compile_table(lat, beam), golden("dba_workload"))
    # a hand-written cell: variables, deferred expressions, inheritance, a sequence with
    # refer = centre and from =, an RBEND chord, a multipole, a monitor, a kicker, RF by harmon
    lat, beam = read_madx(joinpath(GOLDEN_MADX, "hand_cell.madx"); bend = "ExactSBendLinearEdge")
    @test [String(e.name) for e in lat.elements] ==
          ["QF", "DRIFT_1", "QD", "DRIFT_2", "B", "SD", "MP", "DRIFT_3", "QF2", "SF", "BPM", "DRIFT_4", "SF2",
           "DRIFT_5", "RB", "DRIFT_6", "KICK", "DRIFT_7", "RF"]
    @test same_table(compile_table(lat, beam), golden("madx_hand_cell"); rtol = 1e-13)
    @test_throws MadxError read_madx("q: quadrupole, l = 1, k1 = 1, k1s = 0.1; r: line = (q); beam, energy = 3; use, period = r;")
    @test_throws MadxError read_madx("q: quadrupole, l = 1, k1 = a; r: line = (q); beam, energy = 3; use, period = r;")
    # lines: repeat, invert, nest
    lat, beam = read_madx("""
        d: drift, l = 1; q: quadrupole, l = 0.5, k1 = 1; s: sextupole, l = 0.2, k2 = 3;
        half: line = (d, q, s);
        ring: line = (2*half, -half, (d, 2*(q, s)));
        beam, particle = proton, energy = 7000;
        use, period = ring;
        """)
    kinds = join(e isa Drift ? "d" : e isa Quadrupole ? "q" : "s" for e in lat.elements)
    @test kinds == "dqs" ^ 2 * "sqd" * "d" * "qs" ^ 2
    @test beam.mass == TrackPad.M_PROTON
end

# verify_dba_line with both focusing quadrupoles ramped per turn and in time: reference
# element types only, so TrackPad's own CPU ringpass! can track it
function tv_line()
    els = AbstractElement[e for e in verify_dba_line().elements]
    for i in (4, 8)
        els[i] = timed(els[i]; k1 = ctx -> 1.5 * (1.0 + 0.01 * ctx.turn) + 100.0 * ctx.real_time)
    end
    return Lattice(els)
end

if get(ENV, "TRACKPAD_HIP_GPU", "0") == "1"
    @testset "device linepass! == TrackPad linepass! (verify_dba)" begin
        lat = verify_dba_line(); beam = Beam(3.5e9)
        ref = copy(particles_initial); lf = zeros(Int, 10)
        linepass!(ref, lat, beam, lf)
        got = copy(particles_initial); lg = zeros(Int, 10)
        linepass_hip!(got, lat, beam, lg)
        @test lg == lf
        @test isapprox(got, ref; atol = 1e-15)
    end

    @testset "families ring: device ringpass! == TrackPad ringpass!" begin
        lat = families_ring(); beam = Beam(3e9)
        ref = copy(particles_initial); lf = zeros(Int, 10)
        ringpass!(ref, lat, beam, lf, 5)
        got = copy(particles_initial); lg = zeros(Int, 10)
        ringpass_hip!(got, lat, beam, lg, 5)
        @test lg == lf
        @test isapprox(got, ref; rtol = 1e-12, atol = 1e-15)
    end

    @testset "time-varying ring: time and dt_turn reach the device" begin
        lat = tv_line(); beam = Beam(3.5e9)
        ref = copy(particles_initial); lf = zeros(Int, 10)
        ringpass!(ref, lat, beam, lf, 4; time = 1e-4, dt_turn = 2e-6, turn = 3)
        p = HIPParticles(copy(particles_initial))
        ringpass!(p, lat, beam, 4; time = 1e-4, dt_turn = 2e-6, turn = 3)
        got, lg, _, _ = TrackPadHIP.download(p)
        @test lg == lf
        @test isapprox(got, ref; rtol = 1e-12, atol = 1e-15)
        # the same call ignoring time would track another lattice
        p0 = HIPParticles(copy(particles_initial))
        ringpass!(p0, lat, beam, 4; turn = 3)
        @test !isapprox(TrackPadHIP.download(p0)[1], ref; rtol = 1e-12, atol = 1e-15)
    end

    @testset "tables are compiled and uploaded once per lattice" begin
        ctx = TrackPadHIP.default_context()
        lat = dba_ring(); beam = Beam(ENERGY)
        t1 = TrackPadHIP.table_for(ctx, lat, beam)
        @test TrackPadHIP.table_for(ctx, lat, beam) === t1                       # memo by identity
        @test TrackPadHIP.table_for(ctx, Lattice(copy(lat.elements)), beam) === t1  # same blob
        @test TrackPadHIP.table_for(ctx, lat, Beam(2.0e9)) !== t1
    end

    if Base.find_package("Enzyme") !== nothing
        @eval using Enzyme
        @testset "Enzyme rule: the adjoint kernel's gradient" begin
            lat = verify_dba_line(); beam = Beam(3.5e9)
            positions = [4, 6, 8]
            k = [1.5, -1.2, 1.5]
            dk = zeros(3)
            Enzyme.autodiff(Enzyme.Reverse, TrackPadHIP.track_loss, Enzyme.Active, Enzyme.Duplicated(k, dk),
                            Enzyme.Const(copy(particles_initial)), Enzyme.Const(lat), Enzyme.Const(beam),
                            Enzyme.Const(3), Enzyme.Const(positions))
            h = 1e-6
            for i in 1:3
                kp = copy(k); kp[i] += h; km = copy(k); km[i] -= h
                fd = (TrackPadHIP.track_loss(kp, particles_initial, lat, beam, 3, positions) -
                      TrackPadHIP.track_loss(km, particles_initial, lat, beam, 3, positions)) / (2h)
                @test isapprox(dk[i], fd; rtol = 1e-5)
            end
        end
    end
end
