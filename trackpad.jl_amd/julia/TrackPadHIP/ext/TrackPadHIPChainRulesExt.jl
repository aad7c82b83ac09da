"""
    TrackPadHIPChainRulesExt

Registers the hand-written adjoint kernel as the reverse rule of `TrackPadHIP.track_loss`
(the slot the reference reserves for AD rules, ext/TrackPadEnzymeExt.jl:14-17).  Enzyme users
import it with `Enzyme.@import_rrule(typeof(TrackPadHIP.track_loss), Vector{Float64},
Matrix{Float64}, Any, Any, Any, Any)`.
"""
module TrackPadHIPChainRulesExt

using ChainRulesCore
using TrackPadHIP

function ChainRulesCore.rrule(::typeof(TrackPadHIP.track_loss), k::Vector{Float64}, coords0::Matrix{Float64},
                              lat, beam, nturns::Integer, positions::AbstractVector{<:Integer})
    latk = TrackPadHIP.set_strengths(lat, positions, k)
    loss, dk, dx0, _ = TrackPadHIP.ringpass_grad(coords0, latk, beam, nturns)
    function track_loss_pullback(dy)
        return (NoTangent(), dy .* dk[positions], dy .* dx0, NoTangent(), NoTangent(), NoTangent(), NoTangent())
    end
    return loss, track_loss_pullback
end

end # module
