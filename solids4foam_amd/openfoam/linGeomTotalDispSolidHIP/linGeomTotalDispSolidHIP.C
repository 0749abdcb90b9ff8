/*---------------------------------------------------------------------------*\
    linGeomTotalDispSolidHIP: see the header.  Compiled with wmake against an
    OpenFOAM + solids4foam installation (Make/options), linking libs4f.
\*---------------------------------------------------------------------------*/
#include "linGeomTotalDispSolidHIP.H"
#include "addToRunTimeSelectionTable.H"
#include "fvc.H"
#include "fvm.H"
#include "linearElastic.H"
#include "solidTractionFvPatchVectorField.H"
#include "fixedDisplacementFvPatchVectorField.H"
#include "directionMixedFvPatchFields.H"

#include <vector>

namespace Foam
{
namespace solidModels
{

defineTypeNameAndDebug(linGeomTotalDispSolidHIP, 0);
addToRunTimeSelectionTable(solidModel, linGeomTotalDispSolidHIP, dictionary);


void linGeomTotalDispSolidHIP::check(s4f_status st, const char* what) const
{
    if (st != S4F_OK && st != S4F_WARN_MAXITER)
    {
        FatalErrorIn(type() + "::" + what)
            << "libs4f: " << s4f_last_error() << abort(FatalError);
    }
}


// flat owner/neighbour/geometry arrays in face order (internal faces, then
// every patch in order) -- the layout OpenFOAM's polyMesh uses
void linGeomTotalDispSolidHIP::createDeviceMesh()
{
    const fvMesh& m = mesh();
    const label nCells = m.nCells();
    const label nInt = m.nInternalFaces();
    const label nFaces = m.nFaces();

    patchType_.setSize(m.boundary().size());
    forAll(D().boundaryField(), patchI)
    {
        const word t = D().boundaryField()[patchI].type();
        if
        (
            t == fixedDisplacementFvPatchVectorField::typeName
         || t == "linearSpatialDisplacement"
        )
        {
            // linearSpatialDisplacement: its evaluated values are uploaded
            patchType_[patchI] = S4F_PATCH_FIXED_DISPLACEMENT;
        }
        else if (t == solidTractionFvPatchVectorField::typeName)
        {
            patchType_[patchI] = S4F_PATCH_TRACTION;
        }
        else if (t == "solidSymmetry")
        {
            patchType_[patchI] = S4F_PATCH_SYMMETRY;
        }
        else if (t == "fixedDisplacementZeroShear")
        {
            patchType_[patchI] = S4F_PATCH_ZERO_SHEAR;
        }
        else if (t == "empty")
        {
            patchType_[patchI] = S4F_PATCH_EMPTY;
        }
        else
        {
            FatalErrorIn(type() + "::createDeviceMesh()")
                << "patch " << m.boundary()[patchI].name() << " has type " << t
                << "; the device path supports fixedDisplacement, "
                << "linearSpatialDisplacement, solidTraction, solidSymmetry, "
                << "fixedDisplacementZeroShear and empty"
                << abort(FatalError);
        }
    }

    std::vector<int32_t> owner(nFaces), neighbour(nInt), pStart, pSize, pType;
    std::vector<double> C(3*nCells), V(nCells), Sf(3*nFaces), Cf(3*nFaces);
    forAll(m.owner(), faceI) { owner[faceI] = m.owner()[faceI]; }
    forAll(m.neighbour(), faceI) { neighbour[faceI] = m.neighbour()[faceI]; }
    forAll(m.C(), cellI)
    {
        for (direction d = 0; d < 3; ++d) { C[3*cellI + d] = m.C()[cellI][d]; }
        V[cellI] = m.V()[cellI];
    }
    for (label faceI = 0; faceI < nInt; ++faceI)
    {
        for (direction d = 0; d < 3; ++d)
        {
            Sf[3*faceI + d] = m.Sf()[faceI][d];
            Cf[3*faceI + d] = m.Cf()[faceI][d];
        }
    }
    // patches from the polyMesh (an empty fvPatch has size 0, its faces exist)
    forAll(m.boundaryMesh(), patchI)
    {
        const polyPatch& pp = m.boundaryMesh()[patchI];
        pStart.push_back(pp.start());
        pSize.push_back(pp.size());
        pType.push_back(patchType_[patchI]);
        for (label i = 0; i < pp.size(); ++i)
        {
            const label faceI = pp.start() + i;
            for (direction d = 0; d < 3; ++d)
            {
                Sf[3*faceI + d] = m.faceAreas()[faceI][d];
                Cf[3*faceI + d] = m.faceCentres()[faceI][d];
            }
        }
    }

    s4f_mesh_desc desc;
    desc.nCells = nCells;
    desc.nFaces = nFaces;
    desc.nInternalFaces = nInt;
    desc.owner = owner.data();
    desc.neighbour = neighbour.data();
    desc.C = C.data();
    desc.V = V.data();
    desc.Sf = Sf.data();
    desc.Cf = Cf.data();
    desc.nPatches = pStart.size();
    desc.patchStart = pStart.data();
    desc.patchSize = pSize.data();
    desc.patchType = pType.data();
    // leastSquares -> normal patch delta; extendedLeastSquares 0 -> full delta
    const word gradScheme(m.gradScheme("grad(D)"));
    desc.lsFullDelta = gradScheme.find("extendedLeastSquares") != string::npos;
    check(s4f_mesh_create(ctx_, &desc, &gpuMesh_), "createDeviceMesh()");

    // linearElastic laws (linearElastic.C:62-133): one, or one per cellZone
    // named like the law, each solved on its own sub-mesh
    // (mechanicalModel.C:195-245, solidSubMeshes.C:29-44); the device takes
    // the material per cell and the zones for the bi-material interfaces
    const PtrList<mechanicalLaw>& laws = mechanical();
    std::vector<double> E(nCells, -1.0), nu(nCells, 0.0);
    std::vector<int32_t> zone(nCells, -1);
    forAll(laws, lawI)
    {
        if (!isA<linearElastic>(laws[lawI]))
        {
            FatalErrorIn(type() + "::createDeviceMesh()")
                << "the device path takes linearElastic laws"
                << abort(FatalError);
        }
        const linearElastic& law = refCast<const linearElastic>(laws[lawI]);
        // the pressure-equation formulation (mechanicalLaw.C:1374) and an
        // initial stress (linearElastic.C:158-166) change the reference's
        // stress; the device path builds neither
        if (law.solvePressureEqn())
        {
            FatalErrorIn(type() + "::createDeviceMesh()")
                << "solvePressureEqn is not supported by the device path"
                << abort(FatalError);
        }
        if (gMax(mag(law.sigma0())()) > SMALL)
        {
            FatalErrorIn(type() + "::createDeviceMesh()")
                << "a nonzero sigma0 is not supported by the device path"
                << abort(FatalError);
        }
        labelList cells;
        if (laws.size() == 1)
        {
            cells = identity(nCells);
        }
        else
        {
            const label zoneID = m.cellZones().findZoneID(law.name());
            if (zoneID < 0)
            {
                FatalErrorIn(type() + "::createDeviceMesh()")
                    << "no cellZone named " << law.name() << " for its law"
                    << abort(FatalError);
            }
            cells = m.cellZones()[zoneID];
        }
        forAll(cells, i)
        {
            E[cells[i]] = law.E().value();
            nu[cells[i]] = law.nu().value();
            zone[cells[i]] = lawI;
        }
    }
    for (label cellI = 0; cellI < nCells; ++cellI)
    {
        if (zone[cellI] < 0)
        {
            FatalErrorIn(type() + "::createDeviceMesh()")
                << "cell " << cellI << " belongs to no law's cellZone"
                << abort(FatalError);
        }
    }
    // planeStress (mechanicalLaw.C:485-497): linearElastic's 2-D lambda and K
    const Switch planeStress
    (
        m.lookupObject<IOdictionary>("mechanicalProperties")
            .lookupOrDefault<Switch>("planeStress", false)
    );
    check(s4f_plane_stress_set(gpuMesh_, planeStress ? 1 : 0), "createDeviceMesh()");
    check(s4f_material_set(gpuMesh_, E.data(), nu.data(), laws.size() > 1 ? 1 : 0), "createDeviceMesh()");
    if (laws.size() > 1)
    {
        check(s4f_mesh_zones_set(gpuMesh_, zone.data(), Cf.data()), "createDeviceMesh()");
    }

    // rho (inertia and body force) and g (solidModel::g(), constant/g)
    std::vector<double> rho(nCells);
    forAll(rho, cellI) { rho[cellI] = this->rho()[cellI]; }
    check(s4f_density_set(gpuMesh_, rho.data(), 1), "createDeviceMesh()");
    const vector gv = g().value();
    const double gg[3] = {gv.x(), gv.y(), gv.z()};
    check(s4f_gravity_set(gpuMesh_, gg), "createDeviceMesh()");
}


void linGeomTotalDispSolidHIP::uploadBoundaryValues()
{
    const fvMesh& m = mesh();
    forAll(m.boundary(), patchI)
    {
        if (patchType_[patchI] == S4F_PATCH_EMPTY)
        {
            continue;  // empty faces carry no values
        }
        const label n = m.boundary()[patchI].size();
        std::vector<double> v(3*n), p(n, 0.0);
        if (patchType_[patchI] == S4F_PATCH_TRACTION)
        {
            const solidTractionFvPatchVectorField& bc =
                refCast<const solidTractionFvPatchVectorField>(D().boundaryField()[patchI]);
            forAll(bc.traction(), i)
            {
                for (direction d = 0; d < 3; ++d) { v[3*i + d] = bc.traction()[i][d]; }
                p[i] = bc.pressure()[i];
            }
        }
        else if (patchType_[patchI] == S4F_PATCH_ZERO_SHEAR)
        {
            // the prescribed displacement is the direction-mixed refValue
            const directionMixedFvPatchVectorField& bc =
                refCast<const directionMixedFvPatchVectorField>(D().boundaryField()[patchI]);
            forAll(bc, i)
            {
                for (direction d = 0; d < 3; ++d) { v[3*i + d] = bc.refValue()[i][d]; }
            }
        }
        else
        {
            const fvPatchVectorField& bc = D().boundaryField()[patchI];
            forAll(bc, i)
            {
                for (direction d = 0; d < 3; ++d) { v[3*i + d] = bc[i][d]; }
            }
        }
        check(s4f_bc_set(gpuMesh_, patchI, v.data(), p.data()), "uploadBoundaryValues()");
    }
}


s4f_solve_opts linGeomTotalDispSolidHIP::solveOptions() const
{
    s4f_solve_opts o;
    s4f_solve_opts_default(&o);
    o.nCorrectors = nCorr();
    o.solutionTolerance = solutionTol();
    o.alternativeTolerance =
        solidModelDict().lookupOrDefault<scalar>("alternativeTolerance", 1e-7);
    const dictionary& stab = solidModelDict().subDict("stabilisation");
    if (word(stab.lookup("type")) != "RhieChow")
    {
        FatalErrorIn(type() + "::solveOptions()")
            << "the device path implements the RhieChow stabilisation"
            << abort(FatalError);
    }
    o.rhieChowScale = readScalar(stab.lookup("scaleFactor"));
    o.interfaceScale = stab.lookupOrDefault<scalar>("interfaceScaleFactor", 0.01);
    const dictionary& sol = mesh().solverDict(D().name());
    o.pcgTolerance = sol.lookupOrDefault<scalar>("tolerance", 1e-6);
    o.pcgRelTol = sol.lookupOrDefault<scalar>("relTol", 0);
    o.pcgMaxIter = sol.lookupOrDefault<label>("maxIter", 1000);
    o.precision =
        word(solidModelDict().lookupOrDefault<word>("precision", "fp64")) == "mixed"
      ? S4F_PRECISION_MIXED : S4F_PRECISION_FP64;
    // solidModel::relaxField (solidModel.C:823-906): fixed or Aitken, with
    // the fvSolution field relaxation factor of D
    const word method
    (
        solidModelDict().lookupOrDefault<word>("relaxationMethod", "fixed")
    );
    if (method != "fixed" && method != "Aitken")
    {
        FatalErrorIn(type() + "::solveOptions()")
            << "relaxationMethod " << method << " is not on the device path"
            << abort(FatalError);
    }
    o.relaxationMethod = method == "Aitken" ? S4F_RELAX_AITKEN : S4F_RELAX_FIXED;
    if (mesh().relaxField(D().name()))
    {
        o.relaxD = mesh().fieldRelaxationFactor(D().name());
    }
    // rho*fvm::d2dt2(D): steadyState, Euler or backward (linGeomTotalDispSolid.C:143)
    const word d2dt2Scheme(mesh().d2dt2Scheme("d2dt2(" + D().name() + ')'));
    if (d2dt2Scheme == "Euler" || d2dt2Scheme == "backward")
    {
        o.d2dt2 = d2dt2Scheme == "Euler" ? S4F_D2DT2_EULER : S4F_D2DT2_BACKWARD;
        o.deltaT = runTime().deltaTValue();
        o.deltaT0 = runTime().deltaT0Value();
    }
    else if (d2dt2Scheme != "steadyState")
    {
        FatalErrorIn(type() + "::solveOptions()")
            << "d2dt2 scheme " << d2dt2Scheme << " is not on the device path "
            << "(steadyState, Euler, backward)" << abort(FatalError);
    }
    o.warmStart = 1;
    return o;
}


linGeomTotalDispSolidHIP::linGeomTotalDispSolidHIP
(
    Time& runTime,
    const word& region
)
:
    solidModel(typeName, runTime, region),
    impK_(mechanical().impK()),
    rImpK_(1.0/impK_),
    ctx_(nullptr),
    gpuMesh_(nullptr),
    patchType_(),
    predictor_(false)
{
    DisRequired();
    // old-time fields of D exist from the start (linGeomTotalDispSolid.C:75-76)
    fvm::d2dt2(D());
    predictor_ = solidModelDict().lookupOrDefault<Switch>("predictor", false);
    check(s4f_ctx_create(0, &ctx_), "linGeomTotalDispSolidHIP()");
    createDeviceMesh();
    D().correctBoundaryConditions();
    D().storePrevIter();
    mechanical().grad(D(), gradD());
}


linGeomTotalDispSolidHIP::~linGeomTotalDispSolidHIP()
{
    s4f_mesh_destroy(gpuMesh_);
    s4f_ctx_destroy(ctx_);
}


bool linGeomTotalDispSolidHIP::evolve()
{
    Info<< "Evolving solid solver (MI355X)" << endl;

    uploadBoundaryValues();
    const s4f_solve_opts o = solveOptions();
    if (predictor_)
    {
        if (o.d2dt2 == S4F_D2DT2_STEADY)
        {
            FatalErrorIn(type() + "::evolve()")
                << "the predictor needs a transient d2dt2 scheme" << abort(FatalError);
        }
        // linGeomTotalDispSolid::predict (linGeomTotalDispSolid.C:47-59):
        // the step starts from D.oldTime() + U deltaT (the device recomputes
        // gradD and sigma from it)
        D() = D().oldTime() + U()*runTime().deltaT();
    }

    const label nCells = mesh().nCells();
    std::vector<double> Dv(3*nCells), sig(6*nCells), G(9*nCells);
    forAll(D(), cellI)
    {
        for (direction d = 0; d < 3; ++d) { Dv[3*cellI + d] = D()[cellI][d]; }
    }
    if (o.d2dt2 == S4F_D2DT2_EULER)
    {
        // D.oldTime() and D.oldTime().oldTime() of this time step
        std::vector<double> Do(3*nCells), Doo(3*nCells);
        const volVectorField& D0 = D().oldTime();
        const volVectorField& D00 = D().oldTime().oldTime();
        forAll(D0, cellI)
        {
            for (direction d = 0; d < 3; ++d)
            {
                Do[3*cellI + d] = D0[cellI][d];
                Doo[3*cellI + d] = D00[cellI][d];
            }
        }
        check(s4f_old_time_set(gpuMesh_, Do.data(), Doo.data()), "evolve()");
    }
    else if (o.d2dt2 == S4F_D2DT2_BACKWARD)
    {
        // backwardD2dt2Scheme::fvmD2dt2 reads D.oldTime() down to its fourth
        // level (backward ddt of D.oldTime() and D.oldTime().oldTime())
        std::vector<double> lv(4*3*nCells);
        const volVectorField* Dl = &D().oldTime();
        for (label l = 0; l < 4; ++l)
        {
            forAll(*Dl, cellI)
            {
                for (direction d = 0; d < 3; ++d) { lv[(l*nCells + cellI)*3 + d] = (*Dl)[cellI][d]; }
            }
            if (l < 3) { Dl = &Dl->oldTime(); }
        }
        check(s4f_old_times_set(gpuMesh_, 4, lv.data()), "evolve()");
    }
    s4f_solve_report rep;
    const s4f_status st =
        s4f_solve(gpuMesh_, &o, Dv.data(), sig.data(), G.data(), nullptr, &rep);
    check(st, "evolve()");
    if (st == S4F_WARN_MAXITER)
    {
        Warning<< "Max iterations reached within momentum loop" << endl;
    }
    Info<< "    Corr, res, relRes, iters: " << rep.nOuter << ", "
        << rep.initialResidual << ", " << rep.residualvf << ", "
        << rep.innerTotal << " (" << rep.seconds << " s on the device)" << endl;

    D().storePrevIter();
    forAll(D(), cellI)
    {
        for (direction d = 0; d < 3; ++d) { D()[cellI][d] = Dv[3*cellI + d]; }
    }
    D().correctBoundaryConditions();

    // boundary values of gradD and sigma are recomputed by the reference's own
    // operators so every patch and function object sees consistent fields
    mechanical().grad(D(), gradD());
    mechanical().correct(sigma());

    DD() = D() - D().oldTime();
    mechanical().interpolate(D(), gradD(), pointD());
    pointDD() = pointD() - pointD().oldTime();
    U() = fvc::ddt(D());

    return true;
}


tmp<vectorField> linGeomTotalDispSolidHIP::tractionBoundarySnGrad
(
    const vectorField& traction,
    const scalarField& pressure,
    const fvPatch& patch
) const
{
    // same expression as linGeomTotalDispSolid.C:235-271
    const label patchID = patch.index();
    const scalarField& impK = impK_.boundaryField()[patchID];
    const scalarField& rImpK = rImpK_.boundaryField()[patchID];
    const tensorField& pGradD = gradD().boundaryField()[patchID];
    const symmTensorField& pSigma = sigma().boundaryField()[patchID];
    const vectorField n(patch.nf());
    return tmp<vectorField>
    (
        new vectorField
        (
            ((traction - n*pressure) - (n & (pSigma - impK*pGradD)))*rImpK
        )
    );
}

} // End namespace solidModels
} // End namespace Foam
