/*---------------------------------------------------------------------------*\
    HIPPCG

    Run-time selectable lduMatrix::solver "HIPPCG": OpenFOAM's PCG (the
    solver fvMatrix::solveSegregated calls per component from
    linGeomTotalDispSolid.C:164) with the diagonal preconditioner, run on an
    MI355X through libs4f (s4f_ldu_* in include/s4f_abi.h).  Selected in
    system/fvSolution:

        "D|DD" { solver HIPPCG; tolerance 1e-9; relTol 0.1; maxIter 1000; precision fp64; }

    The normalised residuals and the convergence test are OpenFOAM's
    (lduMatrix::solver::normFactor, SolverPerformance::checkConvergence).
    Coupled (processor) interfaces are not handled on the device path.
\*---------------------------------------------------------------------------*/
#ifndef HIPPCG_H
#define HIPPCG_H

#include "lduMatrix.H"
#include "s4f_abi.h"

namespace Foam
{

class HIPPCG
:
    public lduMatrix::solver
{
    mutable s4f_ctx* ctx_;
    mutable s4f_ldu* ldu_;
    mutable label nCellsLdu_;

    HIPPCG(const HIPPCG&);
    void operator=(const HIPPCG&);

public:

    TypeName("HIPPCG");

    HIPPCG
    (
        const word& fieldName,
        const lduMatrix& matrix,
        const FieldField<Field, scalar>& interfaceBouCoeffs,
        const FieldField<Field, scalar>& interfaceIntCoeffs,
        const lduInterfaceFieldPtrsList& interfaces,
        const dictionary& solverControls
    );

    virtual ~HIPPCG();

    virtual solverPerformance solve
    (
        scalarField& psi,
        const scalarField& source,
        const direction cmpt = 0
    ) const;
};

} // End namespace Foam

#endif
