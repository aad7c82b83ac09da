"""
    TrackPadHIPEnzymeExt

Enzyme reverse rule for `TrackPadHIP.track_loss`: the slot the reference leaves open for
tracking AD rules (ext/TrackPadEnzymeExt.jl:14-17, "EnzymeRules.reverse for gradient
computation").  Enzyme does not differentiate through `ccall`s into device code, so the
augmented primal runs the hand-written adjoint kernel (`tp_adjoint` through
`TrackPadHIP.ringpass_grad`) once -- forward with checkpoints and the reverse pass -- and keeps
the gradient on the tape; the reverse step scales it by the incoming adjoint of the loss.

    using Enzyme, TrackPad, TrackPadHIP
    dk = zero(k)
    Enzyme.autodiff(Reverse, TrackPadHIP.track_loss, Active, Duplicated(k, dk), Const(coords0),
                    Const(lat), Const(beam), Const(nturns), Const(positions))
    # dk == dL/dk at the main strengths (k1 / k2 / k3) of lattice positions `positions`

`coords0` may be `Duplicated` as well, to accumulate dL/d(initial coordinates).
"""
module TrackPadHIPEnzymeExt

using Enzyme
using Enzyme: EnzymeRules
using TrackPadHIP

const TL = typeof(TrackPadHIP.track_loss)

function EnzymeRules.augmented_primal(config::EnzymeRules.RevConfigWidth{1}, ::Const{TL},
                                      ::Type{<:Active}, k::Duplicated{Vector{Float64}},
                                      coords0::Union{Const{Matrix{Float64}}, Duplicated{Matrix{Float64}}},
                                      lat::Const, beam::Const, nturns::Const, positions::Const)
    latk = TrackPadHIP.set_strengths(lat.val, positions.val, k.val)
    loss, dk, dx0, _ = TrackPadHIP.ringpass_grad(coords0.val, latk, beam.val, nturns.val)
    tape = (dk[positions.val], dx0)
    primal = EnzymeRules.needs_primal(config) ? loss : nothing
    return EnzymeRules.AugmentedReturn(primal, nothing, tape)
end

function EnzymeRules.reverse(config::EnzymeRules.RevConfigWidth{1}, ::Const{TL}, dret::Active, tape,
                             k::Duplicated{Vector{Float64}},
                             coords0::Union{Const{Matrix{Float64}}, Duplicated{Matrix{Float64}}},
                             lat::Const, beam::Const, nturns::Const, positions::Const)
    gk, gx0 = tape
    k.dval .+= dret.val .* gk
    if coords0 isa Duplicated
        coords0.dval .+= dret.val .* gx0
    end
    return (nothing, nothing, nothing, nothing, nothing, nothing)
end

end # module
