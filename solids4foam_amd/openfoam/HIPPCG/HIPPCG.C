/*---------------------------------------------------------------------------*\
    HIPPCG: see the header.
\*---------------------------------------------------------------------------*/
#include "HIPPCG.H"

#include <vector>

namespace Foam
{

defineTypeNameAndDebug(HIPPCG, 0);

lduMatrix::solver::addsymMatrixConstructorToTable<HIPPCG>
    addHIPPCGSymMatrixConstructorToTable_;


HIPPCG::HIPPCG
(
    const word& fieldName,
    const lduMatrix& matrix,
    const FieldField<Field, scalar>& interfaceBouCoeffs,
    const FieldField<Field, scalar>& interfaceIntCoeffs,
    const lduInterfaceFieldPtrsList& interfaces,
    const dictionary& solverControls
)
:
    lduMatrix::solver
    (
        fieldName,
        matrix,
        interfaceBouCoeffs,
        interfaceIntCoeffs,
        interfaces,
        solverControls
    ),
    ctx_(nullptr),
    ldu_(nullptr),
    nCellsLdu_(-1)
{
    forAll(interfaces, patchI)
    {
        if (interfaces.set(patchI))
        {
            FatalErrorIn("HIPPCG::HIPPCG")
                << "coupled interfaces are not supported by HIPPCG"
                << abort(FatalError);
        }
    }
}


HIPPCG::~HIPPCG()
{
    s4f_ldu_destroy(ldu_);
    s4f_ctx_destroy(ctx_);
}


solverPerformance HIPPCG::solve
(
    scalarField& psi,
    const scalarField& source,
    const direction cmpt
) const
{
    solverPerformance perf(typeName, fieldName_);

    const lduAddressing& addr = matrix_.lduAddr();
    const label nCells = psi.size();
    if (!ctx_)
    {
        if (s4f_ctx_create(0, &ctx_) != S4F_OK)
        {
            FatalErrorIn("HIPPCG::solve") << s4f_last_error() << abort(FatalError);
        }
    }
    if (!ldu_ || nCellsLdu_ != nCells)
    {
        s4f_ldu_destroy(ldu_);
        ldu_ = nullptr;
        std::vector<int32_t> l(addr.lowerAddr().begin(), addr.lowerAddr().end());
        std::vector<int32_t> u(addr.upperAddr().begin(), addr.upperAddr().end());
        if (s4f_ldu_create(ctx_, nCells, l.size(), l.data(), u.data(), &ldu_) != S4F_OK)
        {
            FatalErrorIn("HIPPCG::solve") << s4f_last_error() << abort(FatalError);
        }
        nCellsLdu_ = nCells;
    }

    const word prec(controlDict_.lookupOrDefault<word>("precision", "fp64"));
    double initRes = 0, finalRes = 0;
    int32_t nIter = 0;
    const s4f_status st = s4f_ldu_solve
    (
        ldu_,
        matrix_.diag().begin(),
        matrix_.upper().begin(),
        source.begin(),
        psi.begin(),
        tolerance_,
        relTol_,
        maxIter_,
        prec == "mixed" ? S4F_PRECISION_MIXED : S4F_PRECISION_FP64,
        &initRes,
        &finalRes,
        &nIter
    );
    if (st != S4F_OK)
    {
        FatalErrorIn("HIPPCG::solve") << s4f_last_error() << abort(FatalError);
    }
    perf.initialResidual() = initRes;
    perf.finalResidual() = finalRes;
    perf.nIterations() = nIter;
    perf.checkConvergence(tolerance_, relTol_);
    return perf;
}

} // End namespace Foam
