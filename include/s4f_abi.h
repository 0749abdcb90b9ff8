/*
 * s4f_abi.h -- C ABI of the MI355X implementation of solids4foam's segregated
 * small-strain solid solver (solidModel "linearGeometryTotalDisplacement").
 *
 * Everything crossing this boundary is plain C: opaque handles, int32/int64
 * sizes, FP64 host arrays in OpenFOAM's layouts (vector = xyz per cell,
 * tensor = XX XY XZ YX YY YZ ZX ZY ZZ per cell, symmTensor = XX XY XZ YY YZ
 * ZZ per cell).  Precision modes are internal; every value at the ABI is FP64.
 *
 * Conventions
 *   - every function returns s4f_status (0 = OK); s4f_last_error() returns a
 *     thread-local message for the last failure on the calling thread;
 *   - no C++ exception or abort crosses the ABI;
 *   - input host buffers are borrowed for the duration of the call only;
 *     output buffers are caller-owned; device memory is owned by the handles;
 *   - calls on one context are serialised by the caller (a context is not
 *     thread-safe); separate contexts are independent.
 *
 * Reference interfaces replaced (paths relative to the solids4foam tree):
 *   solidModel run-time selection      src/solids4FoamModels/solidModels/solidModel/solidModel.H:376-386
 *   linGeomTotalDispSolid::evolve      src/solids4FoamModels/solidModels/linGeomTotalDispSolid/linGeomTotalDispSolid.C:111-232
 *   tractionBoundarySnGrad             linGeomTotalDispSolid.C:235-271
 *   lduMatrix::solver::solve (PCG)     OpenFOAM lduMatrix solver table, selected by fvSolution "solver"
 */
#ifndef S4F_ABI_H
#define S4F_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    S4F_OK = 0,
    S4F_ERR_INVALID = 1,      /* bad argument, mesh or dictionary value */
    S4F_ERR_HIP = 2,          /* HIP runtime failure (device missing, launch failure) */
    S4F_ERR_OOM = 3,          /* device allocation failed */
    S4F_ERR_DIVERGED = 4,     /* NaN/Inf in the outer residual norms */
    S4F_WARN_MAXITER = 5,     /* nCorrectors exhausted before convergence (result is valid) */
    S4F_ERR_ADMISSION = 6,    /* serving: request larger than the HBM budget */
    S4F_ERR_NOT_READY = 7     /* serving: no finished result to return yet */
} s4f_status;

typedef struct s4f_ctx s4f_ctx;
typedef struct s4f_polymesh s4f_polymesh;
typedef struct s4f_mesh s4f_mesh;
typedef struct s4f_batch s4f_batch;
typedef struct s4f_server s4f_server;

/* boundary patch types (solidModels/fvPatchFields): */
enum {
    S4F_PATCH_FIXED_DISPLACEMENT = 0, /* fixedDisplacementFvPatchVectorField.C:258-356 */
    S4F_PATCH_TRACTION = 1,           /* solidTractionFvPatchVectorField.C:321-463 */
    S4F_PATCH_EMPTY = 2,              /* [OF] emptyFvPatch: 2-D; its faces carry no values and the
                                         component along their normal is not solved */
    S4F_PATCH_SYMMETRY = 3,           /* solidSymmetryFvPatchVectorField.C:147-250 (first order) */
    S4F_PATCH_ZERO_SHEAR = 4          /* fixedDisplacementZeroShearFvPatchVectorField.C:189-216 over
                                         solidDirectionMixedFvPatchVectorField.C:154-378 (first order,
                                         non-orthogonal corrections, no limiter): normal displacement
                                         prescribed, zero shear traction */
};

/* field relaxation method */
enum { S4F_RELAX_FIXED = 0, S4F_RELAX_AITKEN = 1 };

/* d2dt2Schemes of rho*fvm::d2dt2(D) (linGeomTotalDispSolid.C:143) */
enum { S4F_D2DT2_STEADY = 0, S4F_D2DT2_EULER = 1,
       S4F_D2DT2_BACKWARD = 2 /* backwardD2dt2Scheme.C:227-300, constant time step only */ };

/* precision of the inner linear solve */
enum {
    S4F_PRECISION_FP64 = 0,  /* plain path: everything FP64 */
    S4F_PRECISION_MIXED = 1  /* FP32 PCG sweeps inside the FP64 defect-correction loop */
};

/* inner preconditioner (fvSolution "preconditioner") */
enum {
    S4F_PRECOND_DIAGONAL = 0,  /* OpenFOAM "diagonal" (Jacobi) */
    S4F_PRECOND_GAMG = 2,      /* aggregation multigrid V(1,1) cycle (PCG with preconditioner GAMG) */
    S4F_PRECOND_GAMG_SMOOTHED = 3  /* the same cycle with a smoothed prolongation P = (I - 2/3 D^-1 A) P_agg
                                      at the first coarsening (smoothed aggregation): fewer PCG iterations
                                      on large meshes; s4f_solve and s4f_batch_solve (the server refuses it) */
};

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
/* Creates a context on HIP device `device` with its own stream. */
s4f_status s4f_ctx_create(int device, s4f_ctx **out);
s4f_status s4f_ctx_destroy(s4f_ctx *ctx);
/* Thread-local message describing the last non-OK status on this thread. */
const char *s4f_last_error(void);
/* Library version string. */
const char *s4f_version(void);

/* ------------------------------------------------------------------ */
/* polyMesh generation and geometry (host)                             */
/* ------------------------------------------------------------------ */
/* Single-block box mesh in blockMesh numbering (points and cells i-fastest,
   internal faces in upper-triangular owner/neighbour order, patches x-min,
   x-max, y-min, y-max, z-min, z-max).  jitter > 0 moves every interior vertex
   by a seeded uniform +/- jitter*spacing per axis.  tets != 0 splits every
   hex into 6 tetrahedra around its (0,0,0)-(1,1,1) diagonal. */
s4f_status s4f_polymesh_box(int nx, int ny, int nz, const double L[3], double jitter,
                            uint64_t seed, int tets, s4f_polymesh **out);
/* Borrows external polyMesh arrays (copied): faceOffsets has nFaces+1 entries. */
s4f_status s4f_polymesh_from_arrays(int32_t nPoints, const double *points, int32_t nCells,
                                    int32_t nFaces, int32_t nInternalFaces,
                                    const int32_t *faceOffsets, const int32_t *faceVerts,
                                    const int32_t *owner, const int32_t *neighbour,
                                    int32_t nPatches, const int32_t *patchStart,
                                    const int32_t *patchSize, s4f_polymesh **out);
s4f_status s4f_polymesh_free(s4f_polymesh *pm);
/* sizes: nPoints, nCells, nFaces, nInternalFaces, nPatches, nFaceVerts */
s4f_status s4f_polymesh_sizes(const s4f_polymesh *pm, int64_t out[6]);
/* copies addressing out (any pointer may be NULL) */
s4f_status s4f_polymesh_addressing(const s4f_polymesh *pm, double *points, int32_t *faceOffsets,
                                   int32_t *faceVerts, int32_t *owner, int32_t *neighbour,
                                   int32_t *patchStart, int32_t *patchSize);
/* OpenFOAM primitiveMesh geometry: face centres Cf, area vectors Sf (nFaces*3),
   cell centres C (nCells*3), volumes V (nCells). */
s4f_status s4f_polymesh_geometry(const s4f_polymesh *pm, double *Cf, double *Sf, double *C, double *V);
/* pointCells (primitiveMesh::pointCells): offsets (nPoints+1) and the cells of
   each point in ascending order; cells = NULL fills only the offsets. */
s4f_status s4f_polymesh_point_cells(const s4f_polymesh *pm, int32_t *offsets, int32_t *cells);

/* ------------------------------------------------------------------ */
/* solver mesh: addressing + geometry on the device                    */
/* ------------------------------------------------------------------ */
typedef struct {
    int32_t nCells;
    int32_t nFaces;          /* internal + boundary */
    int32_t nInternalFaces;
    const int32_t *owner;    /* nFaces */
    const int32_t *neighbour;/* nInternalFaces */
    const double *C;         /* nCells*3 cell centres */
    const double *V;         /* nCells volumes */
    const double *Sf;        /* nFaces*3 area vectors (owner -> neighbour / outward) */
    const double *Cf;        /* nFaces*3 face centres */
    int32_t nPatches;
    const int32_t *patchStart; /* absolute face index of each patch's first face */
    const int32_t *patchSize;
    const int32_t *patchType;  /* S4F_PATCH_* */
    int32_t lsFullDelta;       /* 0: OpenFOAM leastSquares (normal patch delta), 1: extendedLeastSquares 0 */
} s4f_mesh_desc;

/* Uploads addressing and geometry, builds the sliced-ELL cell->face lists,
   interpolation weights, delta coefficients, correction vectors and the
   FP64 least-squares inverse matrices.  Host buffers may be freed on return.
   Replaces the fvMesh / surfaceInterpolation / leastSquaresVectors set-up the
   linGeomTotalDispSolid constructor triggers (linGeomTotalDispSolid.C:64-105). */
s4f_status s4f_mesh_create(s4f_ctx *ctx, const s4f_mesh_desc *desc, s4f_mesh **out);
/* Convenience: computes the geometry of `pm` on the host and creates the mesh. */
s4f_status s4f_mesh_create_from_polymesh(s4f_ctx *ctx, const s4f_polymesh *pm,
                                         const int32_t *patchType, int32_t lsFullDelta,
                                         s4f_mesh **out);
s4f_status s4f_mesh_destroy(s4f_mesh *mesh);
/* nCells, nInternalFaces, nBoundaryFaces, ELL entries, ELL slices, device bytes */
s4f_status s4f_mesh_info(const s4f_mesh *mesh, int64_t out[6]);
/* Multigrid hierarchy of the mesh's current matrix, built by its first GAMG
   solve (S4F_ERR_INVALID before): the level count, and for the first
   maxLevels levels their rows and stored off-diagonal entries; *qEntries
   (may be NULL): stored entries of the level-0 operator offdiag(A) P grouped
   per coarse row.  No counterpart in the reference: OpenFOAM's
   GAMGAgglomeration reports its level sizes only in debug output. */
s4f_status s4f_mesh_mg_info(const s4f_mesh *mesh, int32_t maxLevels, int32_t *nLevels, int64_t *rows,
                            int64_t *entries, int64_t *qEntries);

/* Workgroup tiles of a mesh: at most maxTiles tiles of at least minSlices
   64-cell slices each (defaults 8192 and 4; one slice per wave of a
   256-thread workgroup).  Every kernel over the mesh's cells runs one
   workgroup per tile, and its reductions sum per tile, so the tiles decide
   occupancy, cache locality and the finalize kernels' work; an instance keeps
   its own rule in batches and the server.  Set by the autotuner
   (solids4foam_amd/autotune.py).  No counterpart in the reference. */
s4f_status s4f_mesh_tiles_set(s4f_mesh *mesh, int32_t maxTiles, int32_t minSlices);

/* Material zones of a multi-law case: mechanicalProperties with one
   mechanical law per cellZone (mechanicalModel.C:571-615 grad(), :477-498
   correct(), :790-877 interpolate(); solidSubMeshes.C:47-110,1550-2031).
   Each zone is solved on its own sub-mesh: an internal face between two
   zones becomes a boundary face of both sub-meshes with an interface
   displacement, updated every outer iteration toward traction balance, that
   each side's least-squares gradient and boundary stress see; pointD is the
   mean of the sub-meshes' interpolations at points shared by materials.
   cellZone: the zone of every cell (nCells values; the material itself is
   still set per cell by s4f_material_set), or NULL for one material.
   faceCentres: the mesh's face centres, 3 values per face in face order (at
   least the nInternalFaces internal ones).  The interface state starts at
   zero (and the post-loop fields wait for the next solve); a warm or
   transient solve continues it, and s4f_grad uses it as
   the last solve left it.  A mesh with zones is solved alone (batches and
   the server refuse it; zones set after a request was submitted do not
   reach that request). */
s4f_status s4f_mesh_zones_set(s4f_mesh *mesh, const int32_t *cellZone, const double *faceCentres);
/* linearElastic (E, nu) per cell (perCell != 0) or uniform (E[0], nu[0]);
   linearElastic.C:62-133 (mu, lambda, K), :204-245 (impK). */
s4f_status s4f_material_set(s4f_mesh *mesh, const double *E, const double *nu, int32_t perCell);
/* mechanicalProperties planeStress (mechanicalLaw.C:485-497): nonzero selects
   linearElastic's plane-stress moduli (lambda = nu E/((1+nu)(1-nu)),
   K = E/(3(1-nu)), linearElastic.C:109-113) for 2-D cases.  Takes effect at
   the next s4f_material_set. */
s4f_status s4f_plane_stress_set(s4f_mesh *mesh, int32_t planeStress);

/* Density (mechanicalProperties rho, linearElastic.C:62-104): perCell ? nCells
   values : one.  Needed for the inertia term and the body force. */
s4f_status s4f_density_set(s4f_mesh *mesh, const double *rho, int32_t perCell);

/* Gravitational acceleration g (constant/g; solidModel's rho*g body force in
   linGeomTotalDispSolid.C:146): 3 values, zero by default. */
s4f_status s4f_gravity_set(s4f_mesh *mesh, const double *g);

/* Old-time fields of a transient step (S4F_D2DT2_EULER): D.oldTime() and
   D.oldTime().oldTime() (nCells*3 each, AoS; DoldOld NULL: equal to Dold).
   The step starts from Dold unless the solve is given a warm-start field, and
   the convergence test's old-time field is Dold (solidModelTemplates.C:26-188). */
s4f_status s4f_old_time_set(s4f_mesh *mesh, const double *Dold, const double *DoldOld);
/* Old-time levels D.oldTime(), D.oldTime().oldTime(), ... (nLevels = 1..4
   fields of nCells*3, AoS, one after the other; levels not given repeat the
   deepest one, as OpenFOAM creates a missing old-time level as a copy).
   S4F_D2DT2_BACKWARD reads four levels: backwardD2dt2Scheme::fvmD2dt2
   (backwardD2dt2Scheme.C:227-300) applies backward ddt to D.oldTime() and
   D.oldTime().oldTime(), which read their own two old levels. */
s4f_status s4f_old_times_set(s4f_mesh *mesh, int32_t nLevels, const double *levels);
/* Per-face boundary data of patch `patch`: fixedDisplacement and
   fixedDisplacementZeroShear -> displacement (size*3; only its normal
   component binds the zero-shear patch); solidTraction -> traction (size*3)
   and pressure (size, or NULL = 0). */
s4f_status s4f_bc_set(s4f_mesh *mesh, int32_t patch, const double *value, const double *pressure);

/* ------------------------------------------------------------------ */
/* solve                                                               */
/* ------------------------------------------------------------------ */
typedef struct {
    /* solidProperties linearGeometryTotalDisplacementCoeffs (solidModel.C:1147-1177) */
    int32_t nCorrectors;           /* default 10000 */
    double solutionTolerance;      /* default 1e-6 */
    double alternativeTolerance;   /* default 1e-7 */
    /* stabilisation { type RhieChow; scaleFactor; } (solidModel.C:1307-1315) */
    double rhieChowScale;          /* default 0.1 */
    double interfaceScale;         /* default 0.01 (momentumStabilisation.C:122-125) */
    /* fvSolution solvers { D { solver PCG; tolerance; relTol; maxIter; } } */
    double pcgTolerance;           /* default 1e-9 */
    double pcgRelTol;              /* default 0.1 */
    int32_t pcgMaxIter;            /* default 1000 */
    int32_t preconditioner;        /* S4F_PRECOND_* */
    int32_t eagerLaunch;           /* 0: the per-iteration launch sequences are captured in HIP graphs
                                      and replayed; 1: every kernel is launched individually */
    int32_t precision;             /* S4F_PRECISION_* */
    int32_t warmStart;             /* 1: D_inout holds the initial field (old-time field = it) */
    double mgCoarseScale;          /* GAMG: scale of the coarse-grid correction, in (0, 2]; 0 = 1.8 */
    /* fvSolution relaxationFactors { fields { D a; } }: with relaxationMethod
       fixed (solidModel::relaxField, solidModel.C:837-841) D = D_prev + a (D - D_prev)
       after each solve, 0 or 1 = no relaxation; with Aitken the first factor */
    double relaxD;
    /* solidProperties relaxationMethod (solidModel.C:1175-1178): S4F_RELAX_FIXED or
       S4F_RELAX_AITKEN (solidModel.C:842-897: per cell and boundary face,
       D -= alpha (D_prev - D), alpha updated by Aitken's formula, clipped to [0, 2]) */
    int32_t relaxationMethod;
    /* mixed precision's fallback to FP64 sweeps: when an iterating instance's
       |initRes| finds no new minimum 1 % below the previous one within this
       many outer iterations, the solve continues with FP64 sweeps from where it
       stands (report: mixedFallback); 0 = 200, negative = never */
    int32_t mixedStallWindow;
    /* transient: S4F_D2DT2_EULER ([OF] EulerD2dt2Scheme: matrix diagonal
       coefft rDeltaT2 rho V, source rDeltaT2 rho V (coefft0 D_old - coefft00 D_oldOld))
       with the time step deltaT and the previous one deltaT0 (0: = deltaT);
       S4F_D2DT2_BACKWARD (backwardD2dt2Scheme.C:227-300: diagonal coefft^2 rho V/dt^2,
       four old-time levels, deltaT0 must be 0 or deltaT) */
    int32_t d2dt2;
    int32_t progressSeconds;       /* > 0: a progress line on stderr (outer iteration, |initRes|,
                                      residualvf) at most this often, like the reference's
                                      per-iteration log; 0 = silent (default) */
    double deltaT;
    double deltaT0;
} s4f_solve_opts;

typedef struct {
    int32_t nOuter;          /* outer iterations run (iCorr + 1) */
    int32_t converged;       /* solidModel::converged became true */
    int64_t innerTotal;      /* PCG iterations summed over components and outer iterations */
    double initialResidual;  /* |initRes| of the last outer iteration */
    double residualvf;       /* relative displacement change of the last outer iteration */
    double seconds;          /* wall time of the solve on the device stream */
    int32_t mixedFallback;   /* mixed precision fell back to FP64 inner sweeps (mixedStallWindow;
                                solves and batches: the server keeps FP32 sweeps) */
    int32_t warmStarted;     /* the solve started from a resident / supplied field */
    int64_t pcgIterations;   /* PCG iterations (3-RHS sweeps) summed over outer iterations
                                (solves and batches; 0 from the server, which reports innerTotal) */
} s4f_solve_report;

/* Fills opts with the reference defaults. */
void s4f_solve_opts_default(s4f_solve_opts *opts);

/* Runs the momentum loop of linGeomTotalDispSolid::evolve (:135-209) to
   convergence.  D_inout (nCells*3): initial field when opts->warmStart, the
   converged displacement on return.  sigma_out (nCells*6), gradD_out
   (nCells*9) and Db_out (nBoundaryFaces*3) may be NULL.  Returns S4F_OK,
   S4F_WARN_MAXITER or S4F_ERR_DIVERGED. */
s4f_status s4f_solve(s4f_mesh *mesh, const s4f_solve_opts *opts, double *D_inout,
                     double *sigma_out, double *gradD_out, double *Db_out,
                     s4f_solve_report *report);
/* Per-outer-iteration history of the last solve: initRes, rvf (doubles) and
   inner iterations (int32) for the first `cap` iterations. */
/* Post-loop fields of the last s4f_solve on this mesh (device-resident D, gradD):
   solidModel::writeFields (solidModel.C:1936-1962): epsilon = symm(gradD)
   (C*6, XX XY XZ YY YZ ZZ), epsilonEq = sqrt(2/3 |dev epsilon|^2), sigmaEq =
   sqrt(3/2 |dev sigma|^2); any output may be NULL. */
s4f_status s4f_strain_fields(s4f_mesh *mesh, double *epsilon, double *epsilonEq, double *sigmaEq);
/* pointD of the last s4f_solve (linGeomTotalDispSolid.C:212-214 ->
   mechanicalModel::interpolate, mechanicalModel.C:829-845 ->
   enhancedVolPointInterpolation with gradient, enhancedVolPointInterpolate.C:350-420):
   inverse-distance weighted, gradient-extrapolated cell values over each
   point's cells.  points (nPoints*3), pointCells CSR (s4f_polymesh_point_cells). */
s4f_status s4f_point_displacement(s4f_mesh *mesh, int32_t nPoints, const double *points,
                                  const int32_t *pointCellOffsets, const int32_t *pointCells, double *pointD);

s4f_status s4f_solve_history(s4f_mesh *mesh, int32_t cap, double *initRes, double *rvf,
                             int32_t *inner, int32_t *count);

/* ------------------------------------------------------------------ */
/* fine-grained operators (host buffers in, host buffers out)          */
/* ------------------------------------------------------------------ */
/* least-squares grad(D) with the boundary correction: D (C*3), Db (B*3),
   snGrad_b (B*3) -> gradD (C*9), gradD_b (B*9, may be NULL).
   [OF leastSquaresGrad + gaussGrad::correctBoundaryConditions] */
s4f_status s4f_grad(s4f_mesh *mesh, const double *D, const double *Db, const double *snb,
                    double *gradD, double *gradDb);
/* linearElastic::correct: gradD (C*9) -> sigma (C*6) */
s4f_status s4f_sigma(s4f_mesh *mesh, const double *gradD, double *sigma);
/* explicit force imbalance fvc::div(sigma) + Rhie-Chow, integrated per cell:
   D (C*3), gradD (C*9), sigma_b (B*6) -> R (C*3) */
s4f_status s4f_residual(s4f_mesh *mesh, const double *D, const double *gradD,
                        const double *sigmab, double *R);
/* y = M x for three right-hand sides (x, y: C*3, vector layout); precision
   S4F_PRECISION_MIXED runs the FP32 kernel */
s4f_status s4f_spmv3(s4f_mesh *mesh, const double *x, double *y, int32_t precision);
/* PCG on M x = b for three right-hand sides from x = 0 with the given
   tolerance/relTol/maxIter; normFactor per component supplied by the caller
   (OpenFOAM lduMatrix::solver::normFactor); iters (3) may be NULL. */
s4f_status s4f_pcg3(s4f_mesh *mesh, const s4f_solve_opts *opts, const double *b,
                    const double normFactor[3], double *x, int32_t iters[3]);
/* s4f_pcg3 for callers that cannot allocate device memory (code under HIP
   graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
   counterpart in the reference.  b and x are the caller's DEVICE arrays in
   component-major layout (component j of cell c at j*nCells + c, FP64); the
   control block, reduction partials and PCG vectors live in a fixed device
   array of the library (about 26 MiB per GPU, held until the process ends),
   so the call allocates nothing.  Limit: single meshes of at most 262144
   cells (larger meshes return S4F_ERR_INVALID); diagonal preconditioner.
   Every call on one GPU shares that array: calls that overlap on one GPU
   (several contexts or threads) race.  Same kernels as s4f_pcg3, same result. */
s4f_status s4f_pcg3_noalloc(s4f_mesh *mesh, const s4f_solve_opts *opts, const double *b_dev,
                            const double normFactor[3], double *x_dev, int32_t iters[3]);
/* M's diagonal (C) and per-face coefficient a_f (F, upper = -a_f) */
s4f_status s4f_matrix(s4f_mesh *mesh, double *diag, double *faceCoeff);

/* ------------------------------------------------------------------ */
/* measurement (no counterpart in the reference: its timing is the      */
/* per-time-step ExecutionTime/ClockTime print, solids4Foam.C:68-72)    */
/* ------------------------------------------------------------------ */
/* kernel ids for s4f_kernel_time */
enum { S4F_KERNEL_RESIDUAL = 0,   /* div(sigma) + Rhie-Chow face loop */
       S4F_KERNEL_SPMV = 1,       /* PCG direction update fused with the 3-RHS SpMV */
       S4F_KERNEL_PCG_UPDATE = 2, /* PCG solution / residual update */
       S4F_KERNEL_FINALIZE = 3,   /* one-workgroup reduction finalize (alpha) */
       S4F_KERNEL_GRAD = 4,       /* boundary update + least-squares gradient */
       S4F_KERNEL_MG0_DOWN = 5,   /* multigrid: fine-level smoothing + residual */
       S4F_KERNEL_MG0_UP = 6,     /* multigrid: fine-level prolongation + smoothing */
       S4F_KERNEL_VCYCLE = 7,     /* multigrid: every launch of one V-cycle */
       S4F_KERNEL_SPMV_MG = 8 };  /* multigrid PCG: direction update from the V-cycle's z + SpMV */
/* Replays `kernel` back to back `launches` times on the mesh's current state
   (e.g. right after a solve) with all three components active, on the
   context's stream between one pair of hipEvents, and returns the average
   duration in ms.  Launch-by-launch events would perturb the caches the
   kernels rely on; this is the steady per-launch time. */
s4f_status s4f_kernel_time(s4f_mesh *mesh, int32_t kernel, int32_t precision, int32_t launches, double *avgMs);
/* Runs a plain streaming copy kernel (k_copy) of `bytes` read + `bytes`
   written twice and returns the second launch's time: the known-size
   reference for HBM traffic counters (FETCH_SIZE / WRITE_SIZE). */
s4f_status s4f_calibrate_copy(s4f_ctx *ctx, int64_t bytes, double *ms);


/* ------------------------------------------------------------------ */
/* batching: many meshes packed into one sliced-ELL structure          */
/* ------------------------------------------------------------------ */
s4f_status s4f_batch_create(s4f_ctx *ctx, s4f_batch **out);
/* adds a mesh (with its material and BCs set) to the batch; returns its index */
s4f_status s4f_batch_add(s4f_batch *batch, s4f_mesh *mesh, int32_t *index);
/* Solves every instance to its own convergence with shared launches;
   D_inout[i] (nCells_i*3) per instance; reports (count entries). */
s4f_status s4f_batch_solve(s4f_batch *batch, const s4f_solve_opts *opts, double **D_inout,
                           double **sigma_out, s4f_solve_report *reports);
s4f_status s4f_batch_destroy(s4f_batch *batch);

/* ------------------------------------------------------------------ */
/* lduMatrix::solver seam: PCG on a caller's symmetric LDU matrix      */
/* (replaces OpenFOAM's PCG selected by fvSolution "solver", called    */
/* per component from fvMatrix::solveSegregated,                       */
/* linGeomTotalDispSolid.C:164)                                        */
/* ------------------------------------------------------------------ */
typedef struct s4f_ldu s4f_ldu;
/* Addressing of an lduMatrix: lowerAddr (owner) and upperAddr (neighbour) of
   its nFaces off-diagonal pairs; builds the device sliced-ELL structure once. */
s4f_status s4f_ldu_create(s4f_ctx *ctx, int32_t nCells, int32_t nFaces, const int32_t *lowerAddr,
                          const int32_t *upperAddr, s4f_ldu **out);
/* PCG with the diagonal preconditioner on (diag, upper) for one component,
   from psi (updated in place), with OpenFOAM's normalised residual
   (lduMatrix::solver::normFactor) and convergence test (tolerance, relTol,
   maxIter).  Returns the initial and final normalised residuals and the
   iteration count, as SolverPerformance reports them. */
s4f_status s4f_ldu_solve(s4f_ldu *ldu, const double *diag, const double *upper, const double *source,
                         double *psi, double tolerance, double relTol, int32_t maxIter,
                         int32_t precision, double *initRes, double *finalRes, int32_t *nIter);
s4f_status s4f_ldu_destroy(s4f_ldu *ldu);

/* ------------------------------------------------------------------ */
/* serving: continuous batching of solve requests (no counterpart in   */
/* the reference, which runs one case per process)                     */
/* ------------------------------------------------------------------ */
typedef struct {
    s4f_mesh *mesh;            /* mesh with material set; must outlive the request */
    const double *bcValue;     /* nBoundaryFaces*3: fixed displacement / traction per boundary face (copied) */
    const double *bcPressure;  /* nBoundaryFaces or NULL (copied) */
    int32_t warmStart;         /* 1: start from this mesh's last converged field if it is still resident */
} s4f_request;

/* Starts a server thread that owns an HBM arena of `budgetBytes` and up to
   `maxSlots` concurrent solves, all with the same options.  Requests are
   admitted in arrival order as soon as their cells, faces and slots fit;
   in-flight solves of different sizes advance together, one outer iteration
   per shared launch sequence, and retire on convergence.  Meshes stay
   resident after their solve (evicted least-recently-used when space is
   needed), so a repeated mesh skips the upload and can warm start. */
s4f_status s4f_server_create(s4f_ctx *ctx, int64_t budgetBytes, int32_t maxSlots,
                             const s4f_solve_opts *opts, s4f_server **out);
/* Queues a request; S4F_ERR_ADMISSION if it can never fit the budget. */
s4f_status s4f_server_submit(s4f_server *srv, const s4f_request *req, int64_t *id);
/* Id of the next finished request (waits up to waitMs); S4F_ERR_NOT_READY if none. */
s4f_status s4f_server_poll(s4f_server *srv, int32_t waitMs, int64_t *id);
/* Copies a finished result out and releases it: D (nCells*3), sigma
   (nCells*6), report, timing = {submit, start, finish} seconds on the host
   monotonic clock.  Returns the solve's status (OK, WARN_MAXITER, ERR_DIVERGED). */
s4f_status s4f_server_result(s4f_server *srv, int64_t id, double *D_out, double *sigma_out,
                             s4f_solve_report *report, double timing[3]);
/* queued, finished, arena bytes resident, budget, cell capacity, slots */
s4f_status s4f_server_stats(s4f_server *srv, int64_t out[6]);
/* Stops the server thread (in-flight requests are dropped) and frees the arena. */
s4f_status s4f_server_destroy(s4f_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* S4F_ABI_H */
