# Checks the Julia table writer against the blobs the Python writer committed, and (with a
# GPU) the device path against TrackPad's own CPU linepass! on the reference's test input.
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

@testset "table writer matches the committed blobs" begin
    # header, sequence and records byte for byte; pool doubles within 1 ulp (tan/sin/cos
    # of Julia's Base.Math and of glibc may differ in the last place)
    blob = compile_table(verify_dba_line(), Beam(3.5e9))
    gold = read(joinpath(GOLDEN, "verify_dba_line.bin"))
    @test length(blob) == length(gold)
    npool = Int(reinterpret(UInt32, gold[17:20])[1])
    head = length(gold) - 8 * npool
    @test blob[1:head] == gold[1:head]
    a = reinterpret(Float64, blob[head+1:end]); b = reinterpret(Float64, gold[head+1:end])
    @test all(abs.(a .- b) .<= eps.(abs.(b)))
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
end
