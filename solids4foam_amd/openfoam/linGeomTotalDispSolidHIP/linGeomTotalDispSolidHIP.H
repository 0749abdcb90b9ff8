/*---------------------------------------------------------------------------*\
    linGeomTotalDispSolidHIP

    Run-time selectable solidModel "linearGeometryTotalDisplacementHIP": the
    same model as solids4foam's linearGeometryTotalDisplacement
    (src/solids4FoamModels/solidModels/linGeomTotalDispSolid/linGeomTotalDispSolid.H:109)
    whose momentum loop (linGeomTotalDispSolid.C:135-209) runs on an MI355X
    through libs4f (include/s4f_abi.h).

    A case switches with one dictionary entry in constant/solidProperties:

        solidModel     linearGeometryTotalDisplacementHIP;

    and reads the same linearGeometryTotalDisplacementCoeffs (nCorrectors,
    solutionTolerance, alternativeTolerance, stabilisation scaleFactor), plus
    optionally `precision fp64|mixed;`.  The inner PCG reads the D (or "D|DD")
    entry of system/fvSolution (tolerance, relTol, maxIter).

    Supported on the device path: one linearElastic law (with its rho and
    constant/g), fixedDisplacement, linearSpatialDisplacement, solidTraction,
    solidSymmetry, fixedDisplacementZeroShear and empty patches, steady state
    or the Euler d2dt2 scheme (with the predictor), fixed or Aitken field
    relaxation.  Anything else aborts with FatalError, as solidModel::New
    does for unknown types.

    After evolve() the OpenFOAM fields D, gradD and sigma are populated, so
    writeFields, end() and the function objects work unchanged.
\*---------------------------------------------------------------------------*/
#ifndef linGeomTotalDispSolidHIP_H
#define linGeomTotalDispSolidHIP_H

#include "solidModel.H"
#include "volFields.H"
#include "s4f_abi.h"

namespace Foam
{
namespace solidModels
{

class linGeomTotalDispSolidHIP
:
    public solidModel
{
    // Private data

        //- Implicit stiffness (2 mu + lambda) and its reciprocal, used by
        //  tractionBoundarySnGrad exactly as linGeomTotalDispSolid does
        const volScalarField impK_;
        const volScalarField rImpK_;

        //- Device handles
        s4f_ctx* ctx_;
        s4f_mesh* gpuMesh_;

        //- Patch types handed to the device (S4F_PATCH_*)
        labelList patchType_;

        //- linGeomTotalDispSolid's linear predictor (transient only)
        Switch predictor_;

    // Private Member Functions

        void createDeviceMesh();
        void uploadBoundaryValues();
        s4f_solve_opts solveOptions() const;
        void check(s4f_status st, const char* what) const;

        linGeomTotalDispSolidHIP(const linGeomTotalDispSolidHIP&);
        void operator=(const linGeomTotalDispSolidHIP&);

protected:

        virtual nonLinearGeometry::nonLinearType nonLinGeom() const
        {
            return nonLinearGeometry::LINEAR_GEOMETRY;
        }

public:

    TypeName("linearGeometryTotalDisplacementHIP");

    linGeomTotalDispSolidHIP
    (
        Time& runTime,
        const word& region = dynamicFvMesh::defaultRegion
    );

    virtual ~linGeomTotalDispSolidHIP();

    virtual volVectorField& solutionD()
    {
        return D();
    }

    virtual bool evolve();

    virtual tmp<vectorField> tractionBoundarySnGrad
    (
        const vectorField& traction,
        const scalarField& pressure,
        const fvPatch& patch
    ) const;
};

} // End namespace solidModels
} // End namespace Foam

#endif
