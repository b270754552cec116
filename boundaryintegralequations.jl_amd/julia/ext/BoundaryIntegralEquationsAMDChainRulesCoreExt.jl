"""
ChainRulesCore rule for `BoundaryIntegralEquationsAMD.scattering_objective`: the pullback
is the adjoint gradient (two vector-Jacobian products of the GPU assembly, no finite
differences).  The tangent of the mesh is a `Tangent` on its `coordinates`; the collocation
nodes move with the coordinates they sit on, which the gradient already includes.  The
rule selection is held fixed (it is piecewise constant in the coordinates).
"""
module BoundaryIntegralEquationsAMDChainRulesCoreExt

import ChainRulesCore
import BoundaryIntegralEquationsAMD
const AMD = BoundaryIntegralEquationsAMD

function ChainRulesCore.rrule(::typeof(AMD.scattering_objective), mesh::AMD.BIE.Mesh3d, k::Real,
                              field_points::AbstractMatrix; n=4, m=4,
                              direction=[0.0, 0.0, 1.0])
    J, grad, dk = AMD.scattering_gradient(mesh, k, field_points; n=n, m=m,
                                          direction=direction)
    function scattering_objective_pullback(dJ)
        s = ChainRulesCore.unthunk(dJ)
        dmesh = ChainRulesCore.Tangent{typeof(mesh)}(; coordinates=s .* grad)
        return (ChainRulesCore.NoTangent(), dmesh, s * dk, ChainRulesCore.NoTangent())
    end
    return J, scattering_objective_pullback
end

end # module
