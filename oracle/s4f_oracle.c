/*
 * s4f_oracle.c -- FP64 CPU restatement of solids4foam's segregated
 * small-strain solid solver, the `linearGeometryTotalDisplacement` outer loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by the package's smoke() self-check and by the CPU-baseline leg of
 * bench.py, and nowhere else.  The product library (libs4f) never links,
 * loads or calls it, and nothing here runs on the GPU.
 *
 * Reference files restated here (paths relative to the solids4foam tree):
 *   src/solids4FoamModels/solidModels/linGeomTotalDispSolid/
 *       linGeomTotalDispSolid.C:111-232 (evolve), :235-271 (traction snGrad)
 *   src/solids4FoamModels/solidModels/solidModel/solidModelTemplates.C:26-188
 *   src/solids4FoamModels/solidModels/solidModel/solidModel.C:1147-1177,1307-1315
 *   src/solids4FoamModels/solidModels/solidModel/momentumStabilisation/
 *       momentumStabilisation.C:43-216
 *   src/solids4FoamModels/materialModels/mechanicalModel/mechanicalModel.C:409-415,571-582
 *   src/solids4FoamModels/materialModels/mechanicalModel/mechanicalLaws/
 *       linearGeometryLaws/linearElastic/linearElastic.C:62-133,204-245,318-339
 *       mechanicalLaw/mechanicalLaw.C:983-1001,1466-1475
 *   src/solids4FoamModels/solidModels/fvPatchFields/solidTraction/
 *       solidTractionFvPatchVectorField.C:321-463
 *   src/solids4FoamModels/solidModels/fvPatchFields/fixedDisplacement/
 *       fixedDisplacementFvPatchVectorField.C:258-356
 *   src/solids4FoamModels/numerics/patchCorrectionVectors/patchCorrectionVectors.C:24-36
 *   src/solids4FoamModels/materialModels/mechanicalModel/mechanicalModel.C:571-615,790-823
 *   src/solids4FoamModels/materialModels/mechanicalModel/solidSubMeshes/
 *       solidSubMeshes.C:1550-1950,1952-2031, solidSubMeshesTemplates.C:30-92,271-336
 *   src/solids4FoamModels/numerics/extendedLeastSquaresGrad/
 *       extendedLeastSquaresVectors.C:123-285, extendedLeastSquaresGrad.C:103-167
 *
 * The finite-volume arithmetic itself lives in OpenFOAM (not vendored in
 * solids4foam).  Functions marked [OF] restate the published OpenFOAM v2312
 * algorithms: primitiveMesh face/cell geometry, surfaceInterpolation weights
 * and delta coefficients, leastSquaresVectors/leastSquaresGrad,
 * gaussGrad::correctBoundaryConditions, gaussLaplacianScheme (corrected),
 * linear interpolation, surfaceIntegrate, lduMatrix::Amul/sumA,
 * lduMatrix::solver::normFactor, PCG, DIC and diagonal preconditioners, and
 * blockMesh's single-block point/cell/face numbering.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define OR_MAX_PATCH 16
#define OR_SMALL 1e-15   /* OpenFOAM SMALL for double */
#define OR_VSMALL 1e-300 /* OpenFOAM VSMALL for double */
#define OR_ROOTVSMALL 1e-150

enum { OR_FIXED = 0, OR_TRACTION = 1, OR_EMPTY = 2, OR_SYMMETRY = 3, OR_ZERO_SHEAR = 4 };
enum { OR_PRECOND_DIC = 0, OR_PRECOND_DIAGONAL = 1, OR_PRECOND_NONE = 2 };

typedef struct {
    int nPoints, nCells, nFaces, nInternalFaces, nPatches;
    double *points; /* nPoints*3 */
    int *fOff;      /* nFaces+1 */
    int *fVerts;    /* fOff[nFaces] */
    int *owner;     /* nFaces */
    int *neighbour; /* nInternalFaces */
    int patchStart[OR_MAX_PATCH], patchSize[OR_MAX_PATCH];
    /* geometry [OF primitiveMesh] */
    double *Cf, *Sf, *magSf; /* nFaces*3, nFaces*3, nFaces */
    double *C, *V;           /* nCells*3, nCells */
} or_mesh;

typedef struct {
    int nCorr;              /* nCorrectors */
    double solTol, altTol;  /* solutionTolerance, alternativeTolerance */
    int precond;            /* OR_PRECOND_* */
    double pcgTol, pcgRelTol;
    int pcgMaxIter;
    double rcScale;         /* stabilisation scaleFactor (RhieChow) */
    double interfaceScale;  /* stabilisation interfaceScaleFactor */
    int historyCap;         /* capacity of the history arrays below */
    double *histInitRes, *histRvf; int *histInner;
    double relaxD;          /* fvSolution relaxationFactors { fields { D } }: 0 or 1 = none */
    int relaxAitken;        /* solidProperties relaxationMethod: 0 fixed, 1 Aitken */
    int d2dt2;              /* d2dt2Schemes: 0 steadyState, 1 Euler, 2 backward (constant step) */
    double deltaT, deltaT0; /* time step and the previous one (Euler d2dt2) */
    double g[3];            /* gravitational acceleration (constant/g), body force rho g */
} or_opts;

typedef struct {
    int nOuter, converged, innerTotal, maxIterReached;
    double initRes, rvf;
} or_report;

typedef struct {
    const or_mesh *m;
    int nB;                 /* boundary faces */
    int emptyDir;           /* 2-D: axis normal to the empty patches (not solved), -1 in 3-D */
    int lsFullDelta;        /* 0: OpenFOAM leastSquares (normal patch delta); 1: extendedLeastSquares full delta */
    int *patchType;         /* per patch */
    int *bType;             /* per boundary face */
    double *bValue;         /* nB*3: fixed displacement or traction */
    double *bPressure;      /* nB */
    /* material */
    double *mu, *lambda, *K, *impK; /* per cell */
    /* interpolation / geometry helpers */
    double *w;              /* weights (internal faces) */
    double *nonOrthDelta;   /* internal faces */
    double *corrVec;        /* internal faces *3 */
    double *bDelta;         /* boundary deltaCoeffs (1/|delta_n|) */
    double *bK;             /* boundary correction vectors k_b, nB*3 */
    double *bN;             /* boundary unit normals, nB*3 */
    double *lsP, *lsN;      /* internal faces *3 */
    double *lsB;            /* boundary faces *3 */
    double *impKf, *gammaf; /* internal faces */
    /* matrix */
    double *diag, *upper;   /* M (SPD): diag>0, upper<0 */
    /* state */
    double *D, *Dprev, *Dold, *Db;   /* C*3, C*3, C*3, nB*3 */
    double *G, *Gb;                  /* C*9, nB*9 */
    double *sig, *sigb;              /* C*6, nB*6 */
    double *g;                       /* traction snGrad (solidTraction gradient()), nB*3 */
    double *snb;                     /* boundary snGrad used by the gradient boundary correction, nB*3 */
    double *R;                       /* explicit residual, C*3 (integrated) */
    double *aAlpha, *aRes;           /* Aitken: aitkenAlpha (C), aitkenResidual (C*3) */
    double *rho;                     /* density per cell (mechanicalProperties rho) */
    double *Doo;                     /* D.oldTime().oldTime(), C*3 */
    double *Dooo, *Doooo;            /* the third and fourth old-time levels (backward d2dt2), C*3 */
    double *diag0;                   /* the steady matrix diagonal (the inertia term is added per solve) */
    double *aAlphaB, *aResB;         /* their boundary values (nB, nB*3) */
    int iCorr;
    /* bi-material interfaces of a multi-law case (solidSubMeshes; set by
       or_case_set_zones): each material is its own sub-mesh, the internal
       faces between cells of different zones are boundary faces of both */
    int *zone;                       /* material zone of each cell, NULL: one material */
    int nIf;                         /* interface faces */
    int *ifFace;                     /* nIf: their internal-face indices */
    int *faceIf;                     /* internal faces: interface index or -1 */
    double *ifLs;                    /* nIf*6: least-squares boundary vectors, side a (owner) then side b */
    double *ifDc;                    /* nIf*2: sub-mesh patch deltaCoeffs 1/|n.(Cf - C)|, side a then b */
    double *ifDI;                    /* nIf*3: interface displacement (the two sub-meshes hold equal values) */
    double *ifSig;                   /* nIf*12: sub-mesh boundary stress at the face, side a then b */
} or_case;

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */
static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); return p; }

static inline double dot3(const double *a, const double *b) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static inline double mag3(const double *a) { return sqrt(dot3(a, a)); }

/* vector & tensor (OpenFOAM): (v & T)_j = v_i T_ij */
static inline void vdotT(const double *v, const double *T, double *out)
{
    for (int j = 0; j < 3; ++j) out[j] = v[0]*T[0*3+j] + v[1]*T[1*3+j] + v[2]*T[2*3+j];
}

/* counter-based uniform draw in [0,1) for the vertex jitter:
   splitmix64 of (seed, 3*point + axis).  Not Foam::Random: the jitter is an
   input generator, shared bit-for-bit with the product's mesh generator. */
static inline double jitter_u01(uint64_t seed, uint64_t ctr)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (ctr + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------ */
/* [OF] primitiveMeshFaceCentresAndAreas / primitiveMeshCellCentresAndVols */
/* ------------------------------------------------------------------ */
static void or_geometry(or_mesh *m)
{
    const double *p = m->points;
    m->Cf = xcalloc((size_t)m->nFaces*3, sizeof(double));
    m->Sf = xcalloc((size_t)m->nFaces*3, sizeof(double));
    m->magSf = xcalloc((size_t)m->nFaces, sizeof(double));
    m->C = xcalloc((size_t)m->nCells*3, sizeof(double));
    m->V = xcalloc((size_t)m->nCells, sizeof(double));

    for (int f = 0; f < m->nFaces; ++f) {
        const int *fv = m->fVerts + m->fOff[f];
        int np = m->fOff[f+1] - m->fOff[f];
        double *cf = m->Cf + 3*f, *sf = m->Sf + 3*f;
        if (np == 3) {
            const double *a = p + 3*fv[0], *b = p + 3*fv[1], *c = p + 3*fv[2];
            double e1[3], e2[3];
            for (int d = 0; d < 3; ++d) { cf[d] = (1.0/3.0)*(a[d] + b[d] + c[d]); e1[d] = b[d]-a[d]; e2[d] = c[d]-a[d]; }
            sf[0] = 0.5*(e1[1]*e2[2] - e1[2]*e2[1]);
            sf[1] = 0.5*(e1[2]*e2[0] - e1[0]*e2[2]);
            sf[2] = 0.5*(e1[0]*e2[1] - e1[1]*e2[0]);
        } else {
            double fc[3] = { p[3*fv[0]], p[3*fv[0]+1], p[3*fv[0]+2] };
            for (int i = 1; i < np; ++i) for (int d = 0; d < 3; ++d) fc[d] += p[3*fv[i]+d];
            for (int d = 0; d < 3; ++d) fc[d] /= np;
            double sumN[3] = {0,0,0}, sumA = 0.0, sumAc[3] = {0,0,0};
            for (int i = 0; i < np; ++i) {
                const double *pi = p + 3*fv[i], *pn = p + 3*fv[(i+1) % np];
                double c[3], e1[3], e2[3], n[3];
                for (int d = 0; d < 3; ++d) { c[d] = pi[d] + pn[d] + fc[d]; e1[d] = pn[d]-pi[d]; e2[d] = fc[d]-pi[d]; }
                n[0] = e1[1]*e2[2] - e1[2]*e2[1];
                n[1] = e1[2]*e2[0] - e1[0]*e2[2];
                n[2] = e1[0]*e2[1] - e1[1]*e2[0];
                double a = mag3(n);
                for (int d = 0; d < 3; ++d) { sumN[d] += n[d]; sumAc[d] += a*c[d]; }
                sumA += a;
            }
            if (sumA < OR_ROOTVSMALL) {
                for (int d = 0; d < 3; ++d) { cf[d] = fc[d]; sf[d] = 0.0; }
            } else {
                for (int d = 0; d < 3; ++d) { cf[d] = (1.0/3.0)*sumAc[d]/sumA; sf[d] = 0.5*sumN[d]; }
            }
        }
        m->magSf[f] = mag3(sf);
    }

    /* cell centre estimate: average of face centres, owner loop then neighbour loop */
    double *cEst = xcalloc((size_t)m->nCells*3, sizeof(double));
    int *nCF = xcalloc((size_t)m->nCells, sizeof(int));
    for (int f = 0; f < m->nFaces; ++f) { int o = m->owner[f]; for (int d = 0; d < 3; ++d) cEst[3*o+d] += m->Cf[3*f+d]; nCF[o]++; }
    for (int f = 0; f < m->nInternalFaces; ++f) { int n = m->neighbour[f]; for (int d = 0; d < 3; ++d) cEst[3*n+d] += m->Cf[3*f+d]; nCF[n]++; }
    for (int c = 0; c < m->nCells; ++c) for (int d = 0; d < 3; ++d) cEst[3*c+d] /= nCF[c];

    for (int f = 0; f < m->nFaces; ++f) {
        int o = m->owner[f];
        double dd[3]; for (int d = 0; d < 3; ++d) dd[d] = m->Cf[3*f+d] - cEst[3*o+d];
        double pyr3Vol = dot3(m->Sf + 3*f, dd);
        for (int d = 0; d < 3; ++d) {
            double pc = (3.0/4.0)*m->Cf[3*f+d] + (1.0/4.0)*cEst[3*o+d];
            m->C[3*o+d] += pyr3Vol*pc;
        }
        m->V[o] += pyr3Vol;
    }
    for (int f = 0; f < m->nInternalFaces; ++f) {
        int n = m->neighbour[f];
        double dd[3]; for (int d = 0; d < 3; ++d) dd[d] = cEst[3*n+d] - m->Cf[3*f+d];
        double pyr3Vol = dot3(m->Sf + 3*f, dd);
        for (int d = 0; d < 3; ++d) {
            double pc = (3.0/4.0)*m->Cf[3*f+d] + (1.0/4.0)*cEst[3*n+d];
            m->C[3*n+d] += pyr3Vol*pc;
        }
        m->V[n] += pyr3Vol;
    }
    for (int c = 0; c < m->nCells; ++c) {
        if (fabs(m->V[c]) > OR_VSMALL) { for (int d = 0; d < 3; ++d) m->C[3*c+d] /= m->V[c]; }
        else { for (int d = 0; d < 3; ++d) m->C[3*c+d] = cEst[3*c+d]; }
        m->V[c] *= (1.0/3.0);
    }
    free(cEst); free(nCF);
}

/* ------------------------------------------------------------------ */
/* [OF] blockMesh single hex block: point, cell, face and patch numbering */
/* ------------------------------------------------------------------ */
static inline int ptIdx(int nx, int ny, int i, int j, int k) { return i + (nx+1)*(j + (ny+1)*k); }

/* Kuhn split of a hex into 6 tets sharing the (0,0,0)-(1,1,1) diagonal.
   Local hex vertex ids follow the OpenFOAM hex model: 0:(0,0,0) 1:(1,0,0)
   2:(1,1,0) 3:(0,1,0) 4:(0,0,1) 5:(1,0,1) 6:(1,1,1) 7:(0,1,1). */
static const int kuhnTets[6][4] = {
    /* path 0 -> +x -> +y -> 6 */ {0, 1, 2, 6},
    /* 0 -> +y -> +x */          {0, 2, 3, 6},
    /* 0 -> +y -> +z */          {0, 3, 7, 6},
    /* 0 -> +z -> +y */          {0, 7, 4, 6},
    /* 0 -> +z -> +x */          {0, 4, 5, 6},
    /* 0 -> +x -> +z */          {0, 5, 1, 6},
};
/* OpenFOAM tet model faces (outward for a positively oriented tet) */
static const int tetFaces[4][3] = { {1,2,3}, {0,3,2}, {0,1,3}, {0,2,1} };
static const int hexFaces[6][4] = { {0,4,7,3}, {1,2,6,5}, {0,1,5,4}, {3,7,6,2}, {0,3,2,1}, {4,5,6,7} };

typedef struct { uint64_t key0, key1; int cell, lf; } facekey;
static int cmp_facekey(const void *a, const void *b)
{
    const facekey *x = a, *y = b;
    if (x->key0 != y->key0) return x->key0 < y->key0 ? -1 : 1;
    if (x->key1 != y->key1) return x->key1 < y->key1 ? -1 : 1;
    if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
    return x->lf - y->lf;
}
static void sort3(int *v) { int t; if (v[0]>v[1]){t=v[0];v[0]=v[1];v[1]=t;} if (v[1]>v[2]){t=v[1];v[1]=v[2];v[2]=t;} if (v[0]>v[1]){t=v[0];v[0]=v[1];v[1]=t;} }

typedef struct { int nb, lf; } nbrent;
static int cmp_nbr(const void *a, const void *b) { return ((const nbrent*)a)->nb - ((const nbrent*)b)->nb; }

or_mesh *or_box_mesh(int nx, int ny, int nz, double Lx, double Ly, double Lz,
                     double jitter, uint64_t seed, int tets)
{
    if (nx < 1 || ny < 1 || nz < 1) return NULL;
    or_mesh *m = xcalloc(1, sizeof(or_mesh));
    m->nPoints = (nx+1)*(ny+1)*(nz+1);
    m->points = xcalloc((size_t)m->nPoints*3, sizeof(double));
    const double hx = Lx/nx, hy = Ly/ny, hz = Lz/nz;
    for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) {
        int pI = ptIdx(nx, ny, i, j, k);
        double *x = m->points + 3*pI;
        x[0] = Lx*i/nx; x[1] = Ly*j/ny; x[2] = Lz*k/nz;
        if (jitter != 0.0 && i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz) {
            x[0] += jitter*hx*(2.0*jitter_u01(seed, 3ull*pI + 0) - 1.0);
            x[1] += jitter*hy*(2.0*jitter_u01(seed, 3ull*pI + 1) - 1.0);
            x[2] += jitter*hz*(2.0*jitter_u01(seed, 3ull*pI + 2) - 1.0);
        }
    }
    const int nHex = nx*ny*nz;
    m->nPatches = 6;

    if (!tets) {
        m->nCells = nHex;
        int nIF = (nx-1)*ny*nz + nx*(ny-1)*nz + nx*ny*(nz-1);
        int nBF = 2*(ny*nz + nx*nz + nx*ny);
        m->nInternalFaces = nIF;
        m->nFaces = nIF + nBF;
        m->fOff = xcalloc((size_t)m->nFaces+1, sizeof(int));
        m->fVerts = xcalloc((size_t)m->nFaces*4, sizeof(int));
        m->owner = xcalloc((size_t)m->nFaces, sizeof(int));
        m->neighbour = xcalloc((size_t)nIF, sizeof(int));
        int f = 0;
        for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int c = i + nx*(j + ny*k);
            int hv[8] = { ptIdx(nx,ny,i,j,k), ptIdx(nx,ny,i+1,j,k), ptIdx(nx,ny,i+1,j+1,k), ptIdx(nx,ny,i,j+1,k),
                          ptIdx(nx,ny,i,j,k+1), ptIdx(nx,ny,i+1,j,k+1), ptIdx(nx,ny,i+1,j+1,k+1), ptIdx(nx,ny,i,j+1,k+1) };
            int nbrs[3] = { i < nx-1 ? c+1 : -1, j < ny-1 ? c+nx : -1, k < nz-1 ? c+nx*ny : -1 };
            int mf[3] = { 1, 3, 5 };
            for (int q = 0; q < 3; ++q) if (nbrs[q] >= 0) {
                m->owner[f] = c; m->neighbour[f] = nbrs[q];
                m->fOff[f] = 4*f;
                for (int v = 0; v < 4; ++v) m->fVerts[4*f+v] = hv[hexFaces[mf[q]][v]];
                ++f;
            }
        }
        /* boundary faces, blockMesh createBoundary order */
        int pi = 0;
        #define BFACE(ci, a, b, c_, d) do { m->owner[f] = (ci); m->fOff[f] = 4*f; \
            m->fVerts[4*f] = (a); m->fVerts[4*f+1] = (b); m->fVerts[4*f+2] = (c_); m->fVerts[4*f+3] = (d); ++f; } while (0)
        m->patchStart[pi] = f; /* x-min */
        for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j)
            BFACE(0 + nx*(j + ny*k), ptIdx(nx,ny,0,j,k), ptIdx(nx,ny,0,j,k+1), ptIdx(nx,ny,0,j+1,k+1), ptIdx(nx,ny,0,j+1,k));
        m->patchSize[pi] = f - m->patchStart[pi]; ++pi;
        m->patchStart[pi] = f; /* x-max */
        for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j)
            BFACE(nx-1 + nx*(j + ny*k), ptIdx(nx,ny,nx,j,k), ptIdx(nx,ny,nx,j+1,k), ptIdx(nx,ny,nx,j+1,k+1), ptIdx(nx,ny,nx,j,k+1));
        m->patchSize[pi] = f - m->patchStart[pi]; ++pi;
        m->patchStart[pi] = f; /* y-min */
        for (int i = 0; i < nx; ++i) for (int k = 0; k < nz; ++k)
            BFACE(i + nx*(0 + ny*k), ptIdx(nx,ny,i,0,k), ptIdx(nx,ny,i+1,0,k), ptIdx(nx,ny,i+1,0,k+1), ptIdx(nx,ny,i,0,k+1));
        m->patchSize[pi] = f - m->patchStart[pi]; ++pi;
        m->patchStart[pi] = f; /* y-max */
        for (int i = 0; i < nx; ++i) for (int k = 0; k < nz; ++k)
            BFACE(i + nx*(ny-1 + ny*k), ptIdx(nx,ny,i,ny,k), ptIdx(nx,ny,i,ny,k+1), ptIdx(nx,ny,i+1,ny,k+1), ptIdx(nx,ny,i+1,ny,k));
        m->patchSize[pi] = f - m->patchStart[pi]; ++pi;
        m->patchStart[pi] = f; /* z-min */
        for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j)
            BFACE(i + nx*(j + ny*0), ptIdx(nx,ny,i,j,0), ptIdx(nx,ny,i,j+1,0), ptIdx(nx,ny,i+1,j+1,0), ptIdx(nx,ny,i+1,j,0));
        m->patchSize[pi] = f - m->patchStart[pi]; ++pi;
        m->patchStart[pi] = f; /* z-max */
        for (int i = 0; i < nx; ++i) for (int j = 0; j < ny; ++j)
            BFACE(i + nx*(j + ny*(nz-1)), ptIdx(nx,ny,i,j,nz), ptIdx(nx,ny,i+1,j,nz), ptIdx(nx,ny,i+1,j+1,nz), ptIdx(nx,ny,i,j+1,nz));
        m->patchSize[pi] = f - m->patchStart[pi]; ++pi;
        #undef BFACE
        m->fOff[m->nFaces] = 4*m->nFaces;
    } else {
        /* 6-tet split.  Cells: 6*hex + local tet.  Internal faces in
           upper-triangular order (owner ascending, neighbour ascending);
           boundary triangles grouped by the hex patch order above, two per
           quad, in the order of the tets that own them. */
        const int nC = 6*nHex;
        m->nCells = nC;
        facekey *keys = xcalloc((size_t)nC*4, sizeof(facekey));
        int *cellPts = xcalloc((size_t)nC*4, sizeof(int));
        for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            int h = i + nx*(j + ny*k);
            int hv[8] = { ptIdx(nx,ny,i,j,k), ptIdx(nx,ny,i+1,j,k), ptIdx(nx,ny,i+1,j+1,k), ptIdx(nx,ny,i,j+1,k),
                          ptIdx(nx,ny,i,j,k+1), ptIdx(nx,ny,i+1,j,k+1), ptIdx(nx,ny,i+1,j+1,k+1), ptIdx(nx,ny,i,j+1,k+1) };
            for (int t = 0; t < 6; ++t) {
                int c = 6*h + t;
                int v[4];
                for (int q = 0; q < 4; ++q) v[q] = hv[kuhnTets[t][q]];
                /* orient positively: det[v1-v0, v2-v0, v3-v0] > 0 on the undistorted hex */
                const double *p0 = m->points + 3*v[0], *p1 = m->points + 3*v[1], *p2 = m->points + 3*v[2], *p3 = m->points + 3*v[3];
                double a[3], b[3], cc[3];
                for (int d = 0; d < 3; ++d) { a[d] = p1[d]-p0[d]; b[d] = p2[d]-p0[d]; cc[d] = p3[d]-p0[d]; }
                double det = a[0]*(b[1]*cc[2]-b[2]*cc[1]) - a[1]*(b[0]*cc[2]-b[2]*cc[0]) + a[2]*(b[0]*cc[1]-b[1]*cc[0]);
                if (det < 0) { int tmp = v[1]; v[1] = v[2]; v[2] = tmp; }
                for (int q = 0; q < 4; ++q) cellPts[4*c+q] = v[q];
                for (int lf = 0; lf < 4; ++lf) {
                    int fv[3] = { v[tetFaces[lf][0]], v[tetFaces[lf][1]], v[tetFaces[lf][2]] };
                    sort3(fv);
                    facekey *fk = keys + 4*c + lf;
                    fk->key0 = (uint64_t)(uint32_t)fv[0] << 32 | (uint32_t)fv[1];
                    fk->key1 = (uint64_t)(uint32_t)fv[2];
                    fk->cell = c; fk->lf = lf;
                }
            }
        }
        int *nbrOf = xcalloc((size_t)nC*4, sizeof(int));
        for (size_t q = 0; q < (size_t)nC*4; ++q) nbrOf[q] = -1;
        qsort(keys, (size_t)nC*4, sizeof(facekey), cmp_facekey);
        int nIF = 0;
        for (size_t q = 0; q + 1 < (size_t)nC*4; ++q) {
            if (keys[q].key0 == keys[q+1].key0 && keys[q].key1 == keys[q+1].key1) {
                nbrOf[4*keys[q].cell + keys[q].lf] = keys[q+1].cell;
                nbrOf[4*keys[q+1].cell + keys[q+1].lf] = keys[q].cell;
                ++nIF; ++q;
            }
        }
        free(keys);
        int nBF = 4*nC - 2*nIF;
        m->nInternalFaces = nIF;
        m->nFaces = nIF + nBF;
        m->fOff = xcalloc((size_t)m->nFaces+1, sizeof(int));
        m->fVerts = xcalloc((size_t)m->nFaces*3, sizeof(int));
        m->owner = xcalloc((size_t)m->nFaces, sizeof(int));
        m->neighbour = xcalloc((size_t)nIF, sizeof(int));
        int f = 0;
        for (int c = 0; c < nC; ++c) {
            nbrent ent[4]; int ne = 0;
            for (int lf = 0; lf < 4; ++lf) { int nb = nbrOf[4*c+lf]; if (nb > c) { ent[ne].nb = nb; ent[ne].lf = lf; ++ne; } }
            qsort(ent, ne, sizeof(nbrent), cmp_nbr);
            for (int q = 0; q < ne; ++q) {
                m->owner[f] = c; m->neighbour[f] = ent[q].nb; m->fOff[f] = 3*f;
                for (int v = 0; v < 3; ++v) m->fVerts[3*f+v] = cellPts[4*c + tetFaces[ent[q].lf][v]];
                ++f;
            }
        }
        /* boundary: classify each boundary tet face by the box side it lies on */
        for (int pi = 0; pi < 6; ++pi) {
            m->patchStart[pi] = f;
            for (int c = 0; c < nC; ++c) for (int lf = 0; lf < 4; ++lf) {
                if (nbrOf[4*c+lf] >= 0) continue;
                int fv[3] = { cellPts[4*c + tetFaces[lf][0]], cellPts[4*c + tetFaces[lf][1]], cellPts[4*c + tetFaces[lf][2]] };
                /* side from integer point coordinates */
                int side = -1;
                for (int ax = 0; ax < 3 && side < 0; ++ax) {
                    int all0 = 1, allN = 1;
                    int N = ax == 0 ? nx : (ax == 1 ? ny : nz);
                    for (int v = 0; v < 3; ++v) {
                        int pid = fv[v];
                        int ii = pid % (nx+1), jj = (pid / (nx+1)) % (ny+1), kk = pid / ((nx+1)*(ny+1));
                        int cI = ax == 0 ? ii : (ax == 1 ? jj : kk);
                        if (cI != 0) all0 = 0;
                        if (cI != N) allN = 0;
                    }
                    if (all0) side = 2*ax; else if (allN) side = 2*ax + 1;
                }
                if (side != pi) continue;
                m->owner[f] = c; m->fOff[f] = 3*f;
                for (int v = 0; v < 3; ++v) m->fVerts[3*f+v] = fv[v];
                ++f;
            }
            m->patchSize[pi] = f - m->patchStart[pi];
        }
        m->fOff[m->nFaces] = 3*m->nFaces;
        free(nbrOf); free(cellPts);
    }
    or_geometry(m);
    return m;
}

/* Build an oracle mesh from external polyMesh arrays (copied). */
or_mesh *or_mesh_from_arrays(int nPoints, const double *points, int nCells, int nFaces, int nInternalFaces,
                             const int *fOff, const int *fVerts, const int *owner, const int *neighbour,
                             int nPatches, const int *patchStart, const int *patchSize)
{
    if (nPatches > OR_MAX_PATCH) return NULL;
    or_mesh *m = xcalloc(1, sizeof(or_mesh));
    m->nPoints = nPoints; m->nCells = nCells; m->nFaces = nFaces; m->nInternalFaces = nInternalFaces; m->nPatches = nPatches;
    m->points = xcalloc((size_t)nPoints*3, sizeof(double)); memcpy(m->points, points, sizeof(double)*3*nPoints);
    m->fOff = xcalloc((size_t)nFaces+1, sizeof(int)); memcpy(m->fOff, fOff, sizeof(int)*(nFaces+1));
    m->fVerts = xcalloc((size_t)fOff[nFaces], sizeof(int)); memcpy(m->fVerts, fVerts, sizeof(int)*fOff[nFaces]);
    m->owner = xcalloc((size_t)nFaces, sizeof(int)); memcpy(m->owner, owner, sizeof(int)*nFaces);
    m->neighbour = xcalloc((size_t)nInternalFaces, sizeof(int)); memcpy(m->neighbour, neighbour, sizeof(int)*nInternalFaces);
    for (int p = 0; p < nPatches; ++p) { m->patchStart[p] = patchStart[p]; m->patchSize[p] = patchSize[p]; }
    or_geometry(m);
    return m;
}

void or_mesh_free(or_mesh *m)
{
    if (!m) return;
    free(m->points); free(m->fOff); free(m->fVerts); free(m->owner); free(m->neighbour);
    free(m->Cf); free(m->Sf); free(m->magSf); free(m->C); free(m->V);
    free(m);
}

void or_mesh_info(const or_mesh *m, int *out)
{
    out[0] = m->nPoints; out[1] = m->nCells; out[2] = m->nFaces; out[3] = m->nInternalFaces;
    out[4] = m->nPatches; out[5] = m->fOff[m->nFaces];
}

void or_mesh_get(const or_mesh *m, double *points, int *fOff, int *fVerts, int *owner, int *neighbour,
                 int *patchStart, int *patchSize)
{
    if (points) memcpy(points, m->points, sizeof(double)*3*m->nPoints);
    if (fOff) memcpy(fOff, m->fOff, sizeof(int)*(m->nFaces+1));
    if (fVerts) memcpy(fVerts, m->fVerts, sizeof(int)*m->fOff[m->nFaces]);
    if (owner) memcpy(owner, m->owner, sizeof(int)*m->nFaces);
    if (neighbour) memcpy(neighbour, m->neighbour, sizeof(int)*m->nInternalFaces);
    for (int p = 0; p < m->nPatches; ++p) { if (patchStart) patchStart[p] = m->patchStart[p]; if (patchSize) patchSize[p] = m->patchSize[p]; }
}

void or_mesh_geometry(const or_mesh *m, double *Cf, double *Sf, double *C, double *V)
{
    if (Cf) memcpy(Cf, m->Cf, sizeof(double)*3*m->nFaces);
    if (Sf) memcpy(Sf, m->Sf, sizeof(double)*3*m->nFaces);
    if (C) memcpy(C, m->C, sizeof(double)*3*m->nCells);
    if (V) memcpy(V, m->V, sizeof(double)*m->nCells);
}

/* ------------------------------------------------------------------ */
/* case set-up                                                         */
/* ------------------------------------------------------------------ */
static inline int bCell(const or_case *c, int b) { return c->m->owner[c->m->nInternalFaces + b]; }

/* inverse of a symmetric tensor (xx xy xz yy yz zz) by cofactors [OF inv(symmTensor)] */
static void inv_symm(const double *s, double *o)
{
    double xx = s[0], xy = s[1], xz = s[2], yy = s[3], yz = s[4], zz = s[5];
    double det = xx*yy*zz + xy*yz*xz + xz*xy*yz - xx*yz*yz - xy*xy*zz - xz*yy*xz;
    o[0] = (yy*zz - yz*yz)/det;
    o[1] = (xz*yz - xy*zz)/det;
    o[2] = (xy*yz - xz*yy)/det;
    o[3] = (xx*zz - xz*xz)/det;
    o[4] = (xy*xz - xx*yz)/det;
    o[5] = (xx*yy - xy*xy)/det;
}
static inline void symm_dot_v(const double *s, const double *v, double *o)
{
    o[0] = s[0]*v[0] + s[1]*v[1] + s[2]*v[2];
    o[1] = s[1]*v[0] + s[3]*v[1] + s[4]*v[2];
    o[2] = s[2]*v[0] + s[4]*v[1] + s[5]*v[2];
}

/* patch delta [OF fvPatch::delta for non-coupled patches: normal component];
   full = 1 gives Cf - Cn (extendedLeastSquaresVectors.C:147-151). */
static void patch_delta(const or_case *c, int b, int full, double *d)
{
    const or_mesh *m = c->m;
    int f = m->nInternalFaces + b, P = m->owner[f];
    double dl[3]; for (int q = 0; q < 3; ++q) dl[q] = m->Cf[3*f+q] - m->C[3*P+q];
    if (full) { for (int q = 0; q < 3; ++q) d[q] = dl[q]; return; }
    const double *n = c->bN + 3*b;
    double nd = dot3(n, dl);
    for (int q = 0; q < 3; ++q) d[q] = n[q]*nd;
}

/* cell of side a (owner, side 0) or side b (neighbour, side 1) of interface face i */
static inline int if_cell(const or_case *c, int i, int side)
{
    int f = c->ifFace[i];
    return side ? c->m->neighbour[f] : c->m->owner[f];
}

/* the sub-mesh patch delta of interface face i seen from one side: Cf - C
   (lsFullDelta) or its normal part n (n.(Cf - C)), as patch_delta does for
   boundary faces; the sign of the sub-mesh normal cancels in n n */
static void if_delta(const or_case *c, int i, int side, double *d)
{
    const or_mesh *m = c->m;
    int f = c->ifFace[i], P = if_cell(c, i, side);
    double dl[3]; for (int q = 0; q < 3; ++q) dl[q] = m->Cf[3*f+q] - m->C[3*P+q];
    if (c->lsFullDelta) { for (int q = 0; q < 3; ++q) d[q] = dl[q]; return; }
    double n[3]; for (int q = 0; q < 3; ++q) n[q] = m->Sf[3*f+q]/m->magSf[f];
    double nd = dot3(n, dl);
    for (int q = 0; q < 3; ++q) d[q] = n[q]*nd;
}

/* [OF leastSquaresVectors] (the patch-delta switch: lsFullDelta) of the
   mesh, or of every sub-mesh when zones are set: a bi-material interface
   face is then a boundary face of the sub-mesh on each side, with that
   side's patch delta (solidSubMeshes.C:47-110 builds the sub-meshes,
   mechanicalModel.C:584-593 takes fvc::grad on each) */
static void ls_vectors(or_case *c)
{
    const or_mesh *m = c->m;
    const int C = m->nCells, F = m->nInternalFaces, B = c->nB;
    double *dd = xcalloc((size_t)C*6, sizeof(double));
    for (int f = 0; f < F; ++f) {
        if (c->faceIf && c->faceIf[f] >= 0) continue;  /* a boundary face of each sub-mesh, below */
        int o = m->owner[f], n = m->neighbour[f];
        double d[3]; for (int q = 0; q < 3; ++q) d[q] = m->C[3*n+q] - m->C[3*o+q];
        double md2 = dot3(d, d);
        double wdd[6] = { d[0]*d[0]/md2, d[0]*d[1]/md2, d[0]*d[2]/md2, d[1]*d[1]/md2, d[1]*d[2]/md2, d[2]*d[2]/md2 };
        for (int q = 0; q < 6; ++q) { dd[6*o+q] += wdd[q]; dd[6*n+q] += wdd[q]; }
    }
    for (int b = 0; b < B; ++b) {
        if (c->bType[b] == OR_EMPTY) continue;  /* empty faces carry no values */
        int P = bCell(c, b);
        double d[3]; patch_delta(c, b, c->lsFullDelta, d);
        double md2 = dot3(d, d);
        double wdd[6] = { d[0]*d[0]/md2, d[0]*d[1]/md2, d[0]*d[2]/md2, d[1]*d[1]/md2, d[1]*d[2]/md2, d[2]*d[2]/md2 };
        for (int q = 0; q < 6; ++q) dd[6*P+q] += wdd[q];
    }
    for (int i = 0; i < c->nIf; ++i)
        for (int side = 0; side < 2; ++side) {
            int P = if_cell(c, i, side);
            double d[3]; if_delta(c, i, side, d);
            double md2 = dot3(d, d);
            double wdd[6] = { d[0]*d[0]/md2, d[0]*d[1]/md2, d[0]*d[2]/md2, d[1]*d[1]/md2, d[1]*d[2]/md2, d[2]*d[2]/md2 };
            for (int q = 0; q < 6; ++q) dd[6*P+q] += wdd[q];
        }
    double *invDd = xcalloc((size_t)C*6, sizeof(double));
    /* 2-D: dd has a zero row and column along the empty direction; invert
       dd + e e (e that axis) and take e e off the inverse again, so the
       gradient has no component along it */
    static const int diagIdx[3] = {0, 3, 5};
    for (int i = 0; i < C; ++i) {
        if (c->emptyDir >= 0) dd[6*i + diagIdx[c->emptyDir]] += 1.0;
        inv_symm(dd + 6*i, invDd + 6*i);
        if (c->emptyDir >= 0) invDd[6*i + diagIdx[c->emptyDir]] -= 1.0;
    }
    for (int f = 0; f < F; ++f) {
        if (c->faceIf && c->faceIf[f] >= 0) {
            for (int q = 0; q < 3; ++q) { c->lsP[3*f+q] = 0.0; c->lsN[3*f+q] = 0.0; }
            continue;
        }
        int o = m->owner[f], n = m->neighbour[f];
        double d[3]; for (int q = 0; q < 3; ++q) d[q] = m->C[3*n+q] - m->C[3*o+q];
        double r = 1.0/dot3(d, d);
        double t[3];
        symm_dot_v(invDd + 6*o, d, t); for (int q = 0; q < 3; ++q) c->lsP[3*f+q] = r*t[q];
        symm_dot_v(invDd + 6*n, d, t); for (int q = 0; q < 3; ++q) c->lsN[3*f+q] = -r*t[q];
    }
    for (int b = 0; b < B; ++b) {
        if (c->bType[b] == OR_EMPTY) continue;
        int P = bCell(c, b);
        double d[3]; patch_delta(c, b, c->lsFullDelta, d);
        double r = 1.0/dot3(d, d), t[3];
        symm_dot_v(invDd + 6*P, d, t);
        for (int q = 0; q < 3; ++q) c->lsB[3*b+q] = r*t[q];
    }
    for (int i = 0; i < c->nIf; ++i)
        for (int side = 0; side < 2; ++side) {
            int P = if_cell(c, i, side);
            double d[3]; if_delta(c, i, side, d);
            double r = 1.0/dot3(d, d), t[3];
            symm_dot_v(invDd + 6*P, d, t);
            for (int q = 0; q < 3; ++q) c->ifLs[6*i + 3*side + q] = r*t[q];
        }
    free(dd); free(invDd);
}

or_case *or_case_create(const or_mesh *m, const double *E, const double *nu, const int *patchType,
                        const double *bValue, const double *bPressure, int lsFullDelta,
                        double rcScale, double interfaceScale, int planeStress)
{
    /* linearElastic.C:101-147: -1 <= nu <= 0.5; nu = 0.5 sets lambda = K =
       GREAT for the pressure-equation formulation, which is not restated:
       refused (NULL) like an unphysical nu or E <= 0 */
    for (int i = 0; i < m->nCells; ++i)
        if (!(E[i] > 0.0) || !(nu[i] > -1.0 && nu[i] < 0.5)) return NULL;
    or_case *c = xcalloc(1, sizeof(or_case));
    c->m = m;
    const int C = m->nCells, F = m->nInternalFaces, B = m->nFaces - m->nInternalFaces;
    c->nB = B; c->lsFullDelta = lsFullDelta;
    c->patchType = xcalloc(OR_MAX_PATCH, sizeof(int));
    c->bType = xcalloc((size_t)B, sizeof(int));
    for (int p = 0; p < m->nPatches; ++p) {
        c->patchType[p] = patchType[p];
        for (int q = 0; q < m->patchSize[p]; ++q) c->bType[m->patchStart[p] - F + q] = patchType[p];
    }
    /* 2-D (empty patches, [OF] emptyFvPatch): the empty direction is the
       axis along the empty faces' normal; that component is not solved */
    c->emptyDir = -1;
    for (int b = 0; b < B; ++b) if (c->bType[b] == OR_EMPTY) {
        const double *S = m->Sf + 3*(F+b);
        int ax = 0;
        for (int q = 1; q < 3; ++q) if (fabs(S[q]) > fabs(S[ax])) ax = q;
        c->emptyDir = ax;
        break;
    }
    c->bValue = xcalloc((size_t)B*3, sizeof(double)); memcpy(c->bValue, bValue, sizeof(double)*3*B);
    c->bPressure = xcalloc((size_t)B, sizeof(double)); if (bPressure) memcpy(c->bPressure, bPressure, sizeof(double)*B);

    /* linearElastic.C:62-133 (E,nu branch; planeStress :109-113 changes
       lambda and K), impK linearElastic.C:204-245 */
    c->mu = xcalloc(C, sizeof(double)); c->lambda = xcalloc(C, sizeof(double));
    c->K = xcalloc(C, sizeof(double)); c->impK = xcalloc(C, sizeof(double));
    for (int i = 0; i < C; ++i) {
        double e = E[i], v = nu[i];
        c->mu[i] = e/(2.0*(1.0 + v));
        c->lambda[i] = planeStress ? v*e/((1.0 + v)*(1.0 - v)) : v*e/((1.0 + v)*(1.0 - 2.0*v));
        c->K[i] = planeStress ? e/(3.0*(1.0 - v)) : e/(3.0*(1.0 - 2.0*v));
        c->impK[i] = 2.0*c->mu[i] + c->lambda[i];
    }

    /* [OF surfaceInterpolation] weights, nonOrthDeltaCoeffs, nonOrthCorrectionVectors */
    c->w = xcalloc(F, sizeof(double)); c->nonOrthDelta = xcalloc(F, sizeof(double));
    c->corrVec = xcalloc((size_t)F*3, sizeof(double));
    for (int f = 0; f < F; ++f) {
        int o = m->owner[f], n = m->neighbour[f];
        const double *Sf = m->Sf + 3*f, *Cf = m->Cf + 3*f;
        double a[3], b[3], d[3], uA[3];
        for (int q = 0; q < 3; ++q) { a[q] = Cf[q] - m->C[3*o+q]; b[q] = m->C[3*n+q] - Cf[q]; d[q] = m->C[3*n+q] - m->C[3*o+q]; uA[q] = Sf[q]/m->magSf[f]; }
        double SfdOwn = fabs(dot3(Sf, a)), SfdNei = fabs(dot3(Sf, b));
        c->w[f] = SfdNei/(SfdOwn + SfdNei);
        double und = dot3(uA, d), md = mag3(d);
        c->nonOrthDelta[f] = 1.0/fmax(und, 0.05*md);
        for (int q = 0; q < 3; ++q) c->corrVec[3*f+q] = uA[q] - d[q]*c->nonOrthDelta[f];
    }
    c->bN = xcalloc((size_t)B*3, sizeof(double)); c->bK = xcalloc((size_t)B*3, sizeof(double));
    c->bDelta = xcalloc(B, sizeof(double));
    for (int b = 0; b < B; ++b) {
        int f = F + b, P = m->owner[f];
        double n[3], dl[3];
        for (int q = 0; q < 3; ++q) { n[q] = m->Sf[3*f+q]/m->magSf[f]; dl[q] = m->Cf[3*f+q] - m->C[3*P+q]; }
        for (int q = 0; q < 3; ++q) c->bN[3*b+q] = n[q];
        /* patchCorrectionVectors.C:24-36: k = (I - n n) & (Cf - Cn) */
        double nd = dot3(n, dl);
        for (int q = 0; q < 3; ++q) c->bK[3*b+q] = dl[q] - n[q]*nd;
        /* deltaCoeffs = 1/|delta|, delta = n (n & (Cf - Cn)) */
        double dn[3]; for (int q = 0; q < 3; ++q) dn[q] = n[q]*nd;
        c->bDelta[b] = 1.0/mag3(dn);
    }

    c->lsP = xcalloc((size_t)F*3, sizeof(double)); c->lsN = xcalloc((size_t)F*3, sizeof(double));
    c->lsB = xcalloc((size_t)B*3, sizeof(double));
    ls_vectors(c);

    /* impKf = interpolate(impK) (mechanicalModel.C:409-415);
       gammaf: momentumStabilisation.C:111-150 */
    c->impKf = xcalloc(F, sizeof(double)); c->gammaf = xcalloc(F, sizeof(double));
    for (int f = 0; f < F; ++f) {
        int o = m->owner[f], n = m->neighbour[f];
        double wf = c->w[f];
        c->impKf[f] = wf*(c->impK[o] - c->impK[n]) + c->impK[n];
        double go = rcScale*c->impK[o], gn = rcScale*c->impK[n];
        c->gammaf[f] = wf*(go - gn) + gn;
        if (fabs(c->impK[o] - c->impK[n]) > OR_SMALL) c->gammaf[f] = interfaceScale*0.5*(c->impK[o] + c->impK[n]);
    }

    /* M: [OF gaussLaplacianScheme] of -(laplacian(impKf, D)); traction patches
       (fixedGradient) add nothing to the diagonal, fixed patches add
       impK_b |S_b| deltaCoeffs_b (fixedValue gradientInternalCoeffs). */
    c->diag = xcalloc(C, sizeof(double)); c->upper = xcalloc(F, sizeof(double));
    for (int f = 0; f < F; ++f) {
        double a = c->nonOrthDelta[f]*(c->impKf[f]*m->magSf[f]);
        c->upper[f] = -a;
        c->diag[m->owner[f]] += a;
        c->diag[m->neighbour[f]] += a;
    }
    for (int b = 0; b < B; ++b)
        if (c->bType[b] == OR_FIXED || c->bType[b] == OR_SYMMETRY || c->bType[b] == OR_ZERO_SHEAR) {
        /* symmetry and fixedDisplacementZeroShear: [OF] basicSymmetry's and
           directionMixed's implicit parts are per component (|n_j|); one
           matrix serves the three components here, so the full coefficient
           is taken (M only steers the defect correction; the residual carries
           the exact condition) */
        int P = bCell(c, b);
        c->diag[P] += c->impK[P]*m->magSf[F+b]*c->bDelta[b];
    }

    c->D = xcalloc((size_t)C*3, sizeof(double)); c->Dprev = xcalloc((size_t)C*3, sizeof(double));
    c->Dold = xcalloc((size_t)C*3, sizeof(double)); c->Db = xcalloc((size_t)B*3, sizeof(double));
    c->G = xcalloc((size_t)C*9, sizeof(double)); c->Gb = xcalloc((size_t)B*9, sizeof(double));
    c->sig = xcalloc((size_t)C*6, sizeof(double)); c->sigb = xcalloc((size_t)B*6, sizeof(double));
    c->g = xcalloc((size_t)B*3, sizeof(double)); c->snb = xcalloc((size_t)B*3, sizeof(double));
    c->aAlpha = xcalloc((size_t)C, sizeof(double)); c->aRes = xcalloc((size_t)C*3, sizeof(double));
    c->aAlphaB = xcalloc((size_t)B, sizeof(double)); c->aResB = xcalloc((size_t)B*3, sizeof(double));
    c->R = xcalloc((size_t)C*3, sizeof(double));
    c->rho = xcalloc((size_t)C, sizeof(double)); c->Doo = xcalloc((size_t)C*3, sizeof(double));
    c->Dooo = xcalloc((size_t)C*3, sizeof(double)); c->Doooo = xcalloc((size_t)C*3, sizeof(double));
    c->diag0 = xcalloc((size_t)C, sizeof(double)); memcpy(c->diag0, c->diag, sizeof(double)*C);
    return c;
}

void or_case_free(or_case *c)
{
    if (!c) return;
    free(c->patchType); free(c->bType); free(c->bValue); free(c->bPressure);
    free(c->rho); free(c->Doo); free(c->Dooo); free(c->Doooo); free(c->diag0);
    free(c->mu); free(c->lambda); free(c->K); free(c->impK);
    free(c->w); free(c->nonOrthDelta); free(c->corrVec); free(c->bDelta); free(c->bK); free(c->bN);
    free(c->lsP); free(c->lsN); free(c->lsB); free(c->impKf); free(c->gammaf);
    free(c->diag); free(c->upper);
    free(c->D); free(c->Dprev); free(c->Dold); free(c->Db); free(c->G); free(c->Gb);
    free(c->sig); free(c->sigb); free(c->g); free(c->snb); free(c->R);
    free(c->aAlpha); free(c->aRes); free(c->aAlphaB); free(c->aResB);
    free(c->zone); free(c->ifFace); free(c->faceIf); free(c->ifLs); free(c->ifDc); free(c->ifDI); free(c->ifSig);
    free(c);
}

/* Material zones of a multi-law case (mechanicalModel with one law per
   cellZone): every zone becomes a sub-mesh (solidSubMeshes.C:47-110) and
   the internal faces between zones become bi-material interface faces,
   boundary faces of the sub-mesh on either side.  The interface
   displacement and the sub-mesh boundary stress start at zero, as the
   sub-mesh fields are constructed (solidSubMeshes.C:274-321).  zone NULL
   returns to one material.  Returns the number of interface faces. */
int or_case_set_zones(or_case *c, const int *zone)
{
    const or_mesh *m = c->m;
    const int C = m->nCells, F = m->nInternalFaces;
    free(c->zone); free(c->ifFace); free(c->faceIf); free(c->ifLs); free(c->ifDc); free(c->ifDI); free(c->ifSig);
    c->zone = NULL; c->ifFace = NULL; c->faceIf = NULL; c->ifLs = c->ifDc = c->ifDI = c->ifSig = NULL;
    c->nIf = 0;
    if (zone) {
        c->zone = xcalloc(C, sizeof(int)); memcpy(c->zone, zone, sizeof(int)*C);
        c->faceIf = xcalloc(F, sizeof(int));
        for (int f = 0; f < F; ++f) c->faceIf[f] = zone[m->owner[f]] != zone[m->neighbour[f]] ? c->nIf++ : -1;
        c->ifFace = xcalloc(c->nIf, sizeof(int));
        for (int f = 0; f < F; ++f) if (c->faceIf[f] >= 0) c->ifFace[c->faceIf[f]] = f;
        c->ifLs = xcalloc((size_t)c->nIf*6, sizeof(double));
        c->ifDc = xcalloc((size_t)c->nIf*2, sizeof(double));
        c->ifDI = xcalloc((size_t)c->nIf*3, sizeof(double));
        c->ifSig = xcalloc((size_t)c->nIf*12, sizeof(double));
        for (int i = 0; i < c->nIf; ++i) {
            int f = c->ifFace[i];
            double n[3]; for (int q = 0; q < 3; ++q) n[q] = m->Sf[3*f+q]/m->magSf[f];
            for (int side = 0; side < 2; ++side) {
                int P = if_cell(c, i, side);
                double dl[3]; for (int q = 0; q < 3; ++q) dl[q] = m->Cf[3*f+q] - m->C[3*P+q];
                /* [OF] fvPatch deltaCoeffs of a non-coupled patch: 1/|n (n.(Cf - Cn))| */
                c->ifDc[2*i+side] = 1.0/fabs(dot3(n, dl));
            }
        }
        if (c->nIf == 0) { free(c->faceIf); c->faceIf = NULL; }
    }
    ls_vectors(c);
    return c->nIf;
}

/* ------------------------------------------------------------------ */
/* operators                                                           */
/* ------------------------------------------------------------------ */

/* [OF leastSquaresGrad::calcGrad] interior gradient G_ij = dD_j/dx_i,
   then [OF gaussGrad::correctBoundaryConditions]: G_b = G_P + n (snGrad_b - n & G_P).
   snb must hold the boundary snGrad (solidTraction: gradient(); fixedDisplacement:
   fixedDisplacementFvPatchVectorField.C:297-326). */
void or_grad(or_case *c)
{
    const or_mesh *m = c->m;
    const int C = m->nCells, F = m->nInternalFaces, B = c->nB;
    memset(c->G, 0, sizeof(double)*9*C);
    for (int f = 0; f < F; ++f) {
        if (c->faceIf && c->faceIf[f] >= 0) continue;
        int o = m->owner[f], n = m->neighbour[f];
        double dv[3]; for (int q = 0; q < 3; ++q) dv[q] = c->D[3*n+q] - c->D[3*o+q];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) {
            c->G[9*o + 3*i + j] += c->lsP[3*f+i]*dv[j];
            c->G[9*n + 3*i + j] -= c->lsN[3*f+i]*dv[j];
        }
    }
    for (int b = 0; b < B; ++b) {
        if (c->bType[b] == OR_EMPTY) continue;
        int P = bCell(c, b);
        double dv[3]; for (int q = 0; q < 3; ++q) dv[q] = c->Db[3*b+q] - c->D[3*P+q];
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) c->G[9*P + 3*i + j] += c->lsB[3*b+i]*dv[j];
    }
    /* sub-meshes: each side of an interface face sees the interface
       displacement as a boundary value (mechanicalModel.C:584-593) */
    for (int k = 0; k < c->nIf; ++k)
        for (int side = 0; side < 2; ++side) {
            int P = if_cell(c, k, side);
            double dv[3]; for (int q = 0; q < 3; ++q) dv[q] = c->ifDI[3*k+q] - c->D[3*P+q];
            const double *l = c->ifLs + 6*k + 3*side;
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) c->G[9*P + 3*i + j] += l[i]*dv[j];
        }
    for (int b = 0; b < B; ++b) {
        if (c->bType[b] == OR_EMPTY) continue;
        int P = bCell(c, b);
        const double *n = c->bN + 3*b, *GP = c->G + 9*P;
        double nG[3]; vdotT(n, GP, nG);
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j)
            c->Gb[9*b + 3*i + j] = GP[3*i+j] + n[i]*(c->snb[3*b+j] - nG[j]);
    }
}

/* linearElastic.C:318-339, mechanicalLaw.C:983-1001,1474:
   eps = symm(G); sigma = 2 mu dev(eps) + K tr(eps) I */
static void hooke(const double *G, double mu, double K, double *s)
{
    double e[6] = { G[0], 0.5*(G[1]+G[3]), 0.5*(G[2]+G[6]), G[4], 0.5*(G[5]+G[7]), G[8] };
    double tr = e[0] + e[3] + e[5];
    double sh = K*tr;
    s[0] = 2.0*mu*(e[0] - tr/3.0) + sh;
    s[1] = 2.0*mu*e[1];
    s[2] = 2.0*mu*e[2];
    s[3] = 2.0*mu*(e[3] - tr/3.0) + sh;
    s[4] = 2.0*mu*e[4];
    s[5] = 2.0*mu*(e[5] - tr/3.0) + sh;
}

void or_sigma(or_case *c)
{
    const or_mesh *m = c->m;
    for (int i = 0; i < m->nCells; ++i) hooke(c->G + 9*i, c->mu[i], c->K[i], c->sig + 6*i);
    for (int b = 0; b < c->nB; ++b) { int P = bCell(c, b); hooke(c->Gb + 9*b, c->mu[P], c->K[P], c->sigb + 6*b); }
}

/* solidSubMeshes::interpolateDtoSubMeshD (solidSubMeshes.C:1550-1950), a
   bi-material interface face with side a the base owner, side b the
   neighbour, n = Sf/|Sf| (a to b):
     D_I += wab (n.sigma_b - n.sigma_a) + wa (D_a - Dprev_a) + wb (D_b - Dprev_b)
   da = |n.(Cf - C_a)|, db = |n.(C_b - Cf)|, Ka, Kb = impK of the two cells,
   wab = da db/(db Ka + da Kb), wa = db Ka/(db Ka + da Kb), wb = 1 - wa;
   sigma_a, sigma_b are the sub-meshes' boundary stresses at the face from
   their last correct() (the shadow stress is the other side's), D the field
   after the solve and relaxation, Dprev = D.prevIter().  Seen from sub-mesh
   b the normal, the weights and the sides swap and the update is the same,
   so one value serves both sub-meshes (mechanicalModel.C:584-587 calls it
   from grad(), mechanicalModel.C:790-823 once more after the loop). */
static void iface_update(or_case *c, const double *D, const double *Dprev)
{
    const or_mesh *m = c->m;
    for (int i = 0; i < c->nIf; ++i) {
        const int f = c->ifFace[i], a = m->owner[f], b = m->neighbour[f];
        double n[3]; for (int q = 0; q < 3; ++q) n[q] = m->Sf[3*f+q]/m->magSf[f];
        const double da = 1.0/c->ifDc[2*i], db = 1.0/c->ifDc[2*i+1];
        const double Ka = c->impK[a], Kb = c->impK[b];
        const double wab = da*db/(db*Ka + da*Kb);
        const double wa = db*Ka/(db*Ka + da*Kb), wb = 1.0 - wa;
        double ta[3], tb[3];
        symm_dot_v(c->ifSig + 12*i, n, ta);
        symm_dot_v(c->ifSig + 12*i + 6, n, tb);
        for (int q = 0; q < 3; ++q)
            c->ifDI[3*i+q] += wab*(tb[q] - ta[q]) + wa*(D[3*a+q] - Dprev[3*a+q]) + wb*(D[3*b+q] - Dprev[3*b+q]);
    }
}

/* The sub-meshes' boundary stress at the interface faces: gradient
   G_I = G_P + n_s ((D_I - D_P) deltaCoeffs - n_s.G_P) (fvc::grad's
   boundary value, then solidSubMeshes::correctBoundarySnGrad,
   solidSubMeshes.C:1952-2031, whose correction vector
   k = delta - n (n.delta) vanishes for the normal patch delta), n_s the
   sub-mesh's outward normal (+n on side a, -n on side b), then the law's
   Hooke's law on that sub-mesh (mechanicalModel.C:477-498). */
static void iface_sigma(or_case *c)
{
    const or_mesh *m = c->m;
    for (int i = 0; i < c->nIf; ++i) {
        const int f = c->ifFace[i];
        for (int side = 0; side < 2; ++side) {
            const int P = if_cell(c, i, side);
            const double sg = side ? -1.0 : 1.0;
            double n[3]; for (int q = 0; q < 3; ++q) n[q] = sg*m->Sf[3*f+q]/m->magSf[f];
            const double *GP = c->G + 9*P;
            double nG[3]; vdotT(n, GP, nG);
            double GI[9];
            for (int j = 0; j < 3; ++j) {
                const double sn = (c->ifDI[3*i+j] - c->D[3*P+j])*c->ifDc[2*i+side];
                for (int k = 0; k < 3; ++k) GI[3*k+j] = GP[3*k+j] + n[k]*(sn - nG[j]);
            }
            hooke(GI, c->mu[P], c->K[P], c->ifSig + 12*i + 6*side);
        }
    }
}

/* the two interface steps on their own, for tests: the sub-meshes'
   interface stress from the current gradient and D_I, and one interface
   update from D.prevIter() (Dprev) to D */
void or_iface_sigma(or_case *c) { iface_sigma(c); }
void or_iface_update(or_case *c) { iface_update(c, c->D, c->Dprev); }

/* solidTractionFvPatchVectorField::updateCoeffs (:321-395) with
   linGeomTotalDispSolid::tractionBoundarySnGrad (:235-271), relaxFac 1.
   fixedDisplacementZeroShear::updateCoeffs
   (fixedDisplacementZeroShearFvPatchVectorField.C:189-216) sets its refGrad
   to the same snGrad for zero traction and zero pressure. */
void or_update_traction(or_case *c)
{
    for (int b = 0; b < c->nB; ++b) if (c->bType[b] == OR_TRACTION || c->bType[b] == OR_ZERO_SHEAR) {
        int P = bCell(c, b);
        const int zs = c->bType[b] == OR_ZERO_SHEAR;
        const double *n = c->bN + 3*b, *s = c->sigb + 6*b, *Gb = c->Gb + 9*b;
        double ik = c->impK[P], rik = 1.0/ik;
        double T[9] = { s[0] - ik*Gb[0], s[1] - ik*Gb[1], s[2] - ik*Gb[2],
                        s[1] - ik*Gb[3], s[3] - ik*Gb[4], s[4] - ik*Gb[5],
                        s[2] - ik*Gb[6], s[4] - ik*Gb[7], s[5] - ik*Gb[8] };
        double nT[3]; vdotT(n, T, nT);
        for (int q = 0; q < 3; ++q)
            c->g[3*b+q] = (zs ? -nT[q] : (c->bValue[3*b+q] - n[q]*c->bPressure[b]) - nT[q])*rik;
    }
}

/* fixedDisplacementZeroShear value (solidDirectionMixedFvPatchVectorField.C:
   154-259, nonOrthogonalCorrections on, first order, limitCoeff 1) from the
   cell value DP: valueFraction nn, refValue the patch displacement, refGrad g;
   n (n.refValue) + (I - nn).(DP + k & G_P + g/deltaCoeffs) */
static void zero_shear_value(const or_case *c, int b, const double *DP, const double *kG, double *out)
{
    const double *n = c->bN + 3*b, *g = c->g + 3*b, *Dv = c->bValue + 3*b;
    double gv[3];
    for (int q = 0; q < 3; ++q) gv[q] = DP[q] + kG[q] + g[q]/c->bDelta[b];
    double nD = dot3(n, Dv), ng = dot3(n, gv);
    for (int q = 0; q < 3; ++q) out[q] = n[q]*nD + (gv[q] - n[q]*ng);
}

/* D.correctBoundaryConditions() after the solve:
   solidTraction::evaluate (:398-463, nonOrthogonalCorrections on, first order):
       D_b = D_P + (k & G_P) + gradient()/deltaCoeffs
   fixedDisplacement: D_b = prescribed value.
   Also stores the boundary snGrad used by the following gradient's boundary
   correction: traction -> gradient(); fixed -> (D_b - (D_P + k & G_P)) deltaCoeffs
   with G_P the registered grad(D) (the previous iterate) as looked up by
   fixedDisplacementFvPatchVectorField::snGrad (:297-326). */
void or_correct_boundary(or_case *c)
{
    for (int b = 0; b < c->nB; ++b) {
        if (c->bType[b] == OR_EMPTY) continue;
        int P = bCell(c, b);
        double kG[3]; vdotT(c->bK + 3*b, c->G + 9*P, kG);
        if (c->bType[b] == OR_TRACTION) {
            for (int q = 0; q < 3; ++q) {
                c->Db[3*b+q] = c->D[3*P+q] + kG[q] + c->g[3*b+q]/c->bDelta[b];
                c->snb[3*b+q] = c->g[3*b+q];
            }
        } else if (c->bType[b] == OR_ZERO_SHEAR) {
            /* value from the solved cell value; snGrad (:262-378) recomputes
               the same value from the cell value at its call:
               (value - (D_P + k & G_P)) deltaCoeffs */
            double v[3]; zero_shear_value(c, b, c->D + 3*P, kG, v);
            for (int q = 0; q < 3; ++q) {
                c->Db[3*b+q] = v[q];
                c->snb[3*b+q] = (v[q] - (c->D[3*P+q] + kG[q]))*c->bDelta[b];
            }
        } else if (c->bType[b] == OR_SYMMETRY) {
            /* solidSymmetry, first order (solidSymmetryFvPatchVectorField.C:147-250):
               DP = D_P + k & G_P with k = (I - nn) & patch delta, and the patch
               delta is the normal one ([OF.com] fvPatch::delta), so k = 0;
               value (DP + (I - 2nn) DP)/2 = DP - n (n.DP), snGrad -n (n.DP) deltaCoeffs */
            const double *n = c->bN + 3*b, *DP = c->D + 3*P;
            double nD = dot3(n, DP);
            for (int q = 0; q < 3; ++q) {
                c->Db[3*b+q] = DP[q] - n[q]*nD;
                c->snb[3*b+q] = -n[q]*nD*c->bDelta[b];
            }
        } else {
            for (int q = 0; q < 3; ++q) {
                c->Db[3*b+q] = c->bValue[3*b+q];
                c->snb[3*b+q] = (c->Db[3*b+q] - (c->D[3*P+q] + kG[q]))*c->bDelta[b];
            }
        }
    }
}

/* interior face-interpolated gradient, [OF linear]: w (G_o - G_n) + G_n */
static inline void interpG(const or_case *c, int f, double *Gf)
{
    int o = c->m->owner[f], n = c->m->neighbour[f];
    double w = c->w[f];
    for (int q = 0; q < 9; ++q) Gf[q] = w*(c->G[9*o+q] - c->G[9*n+q]) + c->G[9*n+q];
}

/* Explicit force imbalance, integrated over each cell (the volume factor of
   surfaceIntegrate and of the fvMatrix source cancel):
     fvc::div(sigma) [OF Gauss linear]            linGeomTotalDispSolid.C:146
   + RhieChow stabilisation                        momentumStabilisation.C:210-217
       fvc::laplacian(gammaf, D) [Gauss linear corrected] - fvc::div(gammaf*(Sf & interpolate(gradD)))
   (gammaf_b = 0 on non-coupled patches: momentumStabilisation.C:198-206). */
void or_residual(or_case *c)
{
    const or_mesh *m = c->m;
    const int C = m->nCells, F = m->nInternalFaces, B = c->nB;
    memset(c->R, 0, sizeof(double)*3*C);
    /* div(sigma) */
    for (int f = 0; f < F; ++f) {
        int o = m->owner[f], n = m->neighbour[f];
        double w = c->w[f], sf[6];
        for (int q = 0; q < 6; ++q) sf[q] = w*(c->sig[6*o+q] - c->sig[6*n+q]) + c->sig[6*n+q];
        const double *S = m->Sf + 3*f;
        double fl[3] = { S[0]*sf[0] + S[1]*sf[1] + S[2]*sf[2],
                         S[0]*sf[1] + S[1]*sf[3] + S[2]*sf[4],
                         S[0]*sf[2] + S[1]*sf[4] + S[2]*sf[5] };
        for (int q = 0; q < 3; ++q) { c->R[3*o+q] += fl[q]; c->R[3*n+q] -= fl[q]; }
    }
    for (int b = 0; b < B; ++b) {
        if (c->bType[b] == OR_EMPTY) continue;
        int P = bCell(c, b);
        const double *S = m->Sf + 3*(F+b), *s = c->sigb + 6*b;
        double fl[3] = { S[0]*s[0] + S[1]*s[1] + S[2]*s[2],
                         S[0]*s[1] + S[1]*s[3] + S[2]*s[4],
                         S[0]*s[2] + S[1]*s[4] + S[2]*s[5] };
        for (int q = 0; q < 3; ++q) c->R[3*P+q] += fl[q];
    }
    /* RhieChow: laplacian(gammaf, D) - div(gammaf (Sf & interpolate(gradD))) */
    for (int f = 0; f < F; ++f) {
        int o = m->owner[f], n = m->neighbour[f];
        double Gf[9]; interpG(c, f, Gf);
        double kG[3]; vdotT(c->corrVec + 3*f, Gf, kG);
        double SG[3]; vdotT(m->Sf + 3*f, Gf, SG);
        double gm = c->gammaf[f]*m->magSf[f];
        for (int q = 0; q < 3; ++q) {
            double sn = c->nonOrthDelta[f]*(c->D[3*n+q] - c->D[3*o+q]) + kG[q];
            double fl = gm*sn - c->gammaf[f]*SG[q];
            c->R[3*o+q] += fl; c->R[3*n+q] -= fl;
        }
    }
}

/* [OF lduMatrix::Amul] symmetric: y = diag x + offdiag */
static void amul(const or_case *c, const double *x, double *y)
{
    const or_mesh *m = c->m;
    for (int i = 0; i < m->nCells; ++i) y[i] = c->diag[i]*x[i];
    for (int f = 0; f < m->nInternalFaces; ++f) {
        int l = m->owner[f], u = m->neighbour[f];
        y[u] += c->upper[f]*x[l];
        y[l] += c->upper[f]*x[u];
    }
}

/* [OF DIC] reciprocal preconditioned diagonal and preconditioning */
static void dic_rD(const or_case *c, double *rD)
{
    const or_mesh *m = c->m;
    memcpy(rD, c->diag, sizeof(double)*m->nCells);
    for (int f = 0; f < m->nInternalFaces; ++f) {
        int l = m->owner[f], u = m->neighbour[f];
        rD[u] -= c->upper[f]*c->upper[f]/rD[l];
    }
    for (int i = 0; i < m->nCells; ++i) rD[i] = 1.0/rD[i];
}
static void precondition(const or_case *c, int type, const double *rD, const double *r, double *w)
{
    const or_mesh *m = c->m;
    const int C = m->nCells, F = m->nInternalFaces;
    if (type == OR_PRECOND_NONE) { memcpy(w, r, sizeof(double)*C); return; }
    for (int i = 0; i < C; ++i) w[i] = rD[i]*r[i];
    if (type == OR_PRECOND_DIAGONAL) return;
    for (int f = 0; f < F; ++f) { int l = m->owner[f], u = m->neighbour[f]; w[u] -= rD[u]*c->upper[f]*w[l]; }
    for (int f = F - 1; f >= 0; --f) { int l = m->owner[f], u = m->neighbour[f]; w[l] -= rD[l]*c->upper[f]*w[u]; }
}

/* [OF PCG::solve] for one component, with [OF normFactor]; returns nIter. */
static int pcg(const or_case *c, const or_opts *o, const double *rD, double *psi, const double *source,
               double *initRes, double *finalRes, double *work)
{
    const or_mesh *m = c->m;
    const int N = m->nCells;
    double *pA = work, *wA = work + N, *rA = work + 2*N, *tmp = work + 3*N;
    amul(c, psi, wA);
    for (int i = 0; i < N; ++i) rA[i] = source[i] - wA[i];
    /* normFactor: sumA * gAverage(psi) */
    for (int i = 0; i < N; ++i) tmp[i] = c->diag[i];
    for (int f = 0; f < m->nInternalFaces; ++f) { tmp[m->neighbour[f]] += c->upper[f]; tmp[m->owner[f]] += c->upper[f]; }
    double avg = 0.0; for (int i = 0; i < N; ++i) avg += psi[i];
    avg /= N;
    double nf = 0.0;
    for (int i = 0; i < N; ++i) { double t = tmp[i]*avg; nf += fabs(wA[i] - t) + fabs(source[i] - t); }
    nf += 1e-20;
    double s = 0.0; for (int i = 0; i < N; ++i) s += fabs(rA[i]);
    *initRes = s/nf; *finalRes = *initRes;
    int nIter = 0;
    #define CONVERGED(fr, ir) ((fr) < o->pcgTol || (o->pcgRelTol > 1e-20 && (fr) < o->pcgRelTol*(ir)))
    if (!CONVERGED(*finalRes, *initRes)) {
        double wArA = 1e20, wArAold;
        do {
            wArAold = wArA;
            precondition(c, o->precond, rD, rA, wA);
            wArA = 0.0; for (int i = 0; i < N; ++i) wArA += wA[i]*rA[i];
            if (nIter == 0) { for (int i = 0; i < N; ++i) pA[i] = wA[i]; }
            else { double beta = wArA/wArAold; for (int i = 0; i < N; ++i) pA[i] = wA[i] + beta*pA[i]; }
            amul(c, pA, wA);
            double wApA = 0.0; for (int i = 0; i < N; ++i) wApA += wA[i]*pA[i];
            if (fabs(wApA)/nf < OR_VSMALL) break;
            double alpha = wArA/wApA;
            for (int i = 0; i < N; ++i) { psi[i] += alpha*pA[i]; rA[i] -= alpha*wA[i]; }
            s = 0.0; for (int i = 0; i < N; ++i) s += fabs(rA[i]);
            *finalRes = s/nf;
        } while (++nIter < o->pcgMaxIter && !CONVERGED(*finalRes, *initRes));
    }
    #undef CONVERGED
    return nIter;
}

/* reference-form source of the assembled DEqn for component q (literal
   fvm::laplacian - fvc::laplacian + explicit terms, integrated), for the test
   that pins b == M D + R: returns b in out (C values). */
void or_reference_source(or_case *c, int q, double *out)
{
    const or_mesh *m = c->m;
    const int C = m->nCells, F = m->nInternalFaces, B = c->nB;
    /* -(fvm source): fvm.source -= V*div(gammaMagSf*correction); boundary source = boundaryCoeffs */
    for (int i = 0; i < C; ++i) out[i] = c->R[3*i+q];
    for (int f = 0; f < F; ++f) {
        int o = m->owner[f], n = m->neighbour[f];
        double Gf[9]; interpG(c, f, Gf);
        double kG[3]; vdotT(c->corrVec + 3*f, Gf, kG);
        double gm = c->impKf[f]*m->magSf[f];
        /* implicit-part correction of fvm::laplacian, moved to the DEqn source with a minus sign */
        double corr = gm*kG[q];
        out[o] += corr; out[n] -= corr;
        /* - fvc::laplacian(impKf, D) (explicit, same corrected snGrad) */
        double fl = gm*(c->nonOrthDelta[f]*(c->D[3*n+q] - c->D[3*o+q]) + kG[q]);
        out[o] -= fl; out[n] += fl;
    }
    for (int b = 0; b < B; ++b) {
        if (c->bType[b] == OR_EMPTY || c->bType[b] == OR_SYMMETRY || c->bType[b] == OR_ZERO_SHEAR)
            continue;  /* literal form: fixed and traction */
        int P = bCell(c, b);
        double gm = c->impK[P]*m->magSf[F+b];
        if (c->bType[b] == OR_TRACTION) {
            /* fixedGradient: boundaryCoeffs = gm*gradient, fvc snGrad = gradient */
            out[P] += gm*c->g[3*b+q];
            out[P] -= gm*c->g[3*b+q];
        } else {
            double kG[3]; vdotT(c->bK + 3*b, c->G + 9*P, kG);
            /* fixedDisplacement gradientBoundaryCoeffs = deltaCoeffs (D_b - k & G_P) */
            out[P] += gm*c->bDelta[b]*(c->Db[3*b+q] - kG[q]);
            double sn = (c->Db[3*b+q] - (c->D[3*P+q] + kG[q]))*c->bDelta[b];
            out[P] -= gm*sn;
        }
    }
}

static double maxmag(const double *a, const double *b, int n)
{
    double mx = 0.0;
    for (int i = 0; i < n; ++i) {
        double d[3] = { a[3*i] - (b ? b[3*i] : 0.0), a[3*i+1] - (b ? b[3*i+1] : 0.0), a[3*i+2] - (b ? b[3*i+2] : 0.0) };
        double v = mag3(d);
        if (isnan(v)) return NAN;
        if (v > mx) mx = v;
    }
    return mx;
}

/* Initialise: D, D_b, G, sigma as in the linGeomTotalDispSolid constructor
   (:76-84): D from the caller (zero by default), D.correctBoundaryConditions()
   with the traction gradient at its initial value (zero), grad(D). sigma is
   zero until the first correct(). */
void or_init(or_case *c, const double *D0)
{
    const or_mesh *m = c->m;
    const int C = m->nCells, B = c->nB;
    if (D0) memcpy(c->D, D0, sizeof(double)*3*C); else memset(c->D, 0, sizeof(double)*3*C);
    memcpy(c->Dold, c->D, sizeof(double)*3*C);
    memset(c->G, 0, sizeof(double)*9*C);
    memset(c->g, 0, sizeof(double)*3*B);
    or_correct_boundary(c);
    /* sub-meshes: the interface displacement and boundary stress start at
       zero; the constructor's grad() leaves D_I unchanged (D == prevIter,
       zero stress) */
    if (c->nIf) {
        memset(c->ifDI, 0, sizeof(double)*3*c->nIf);
        memset(c->ifSig, 0, sizeof(double)*12*c->nIf);
    }
    or_grad(c);
    /* the constructor computes no stress: a cold start keeps the zero sigma
       field until the first mechanical().correct(), so the first residual
       and traction update see sigma = 0; a restart (D0 given) reads the
       stress written with D, i.e. Hooke's law of its gradient */
    if (D0) { or_sigma(c); iface_sigma(c); }
    else {
        memset(c->sig, 0, sizeof(double)*6*C);
        memset(c->sigb, 0, sizeof(double)*6*B);
    }
    c->iCorr = 0;
}

/* pointD (linGeomTotalDispSolid.C:212-214 -> mechanicalModel::interpolate,
   mechanicalModel.C:829-845 -> enhancedVolPointInterpolation::interpolate with
   the gradient, enhancedVolPointInterpolate.C:350-420): for every point p,
   sum over its cells c (ascending) of w (D_c + (x_p - C_c).G_c), w = 1/|x_p - C_c|,
   divided by the sum of w.  D (C*3) and G (C*9) are given; out (P*3).
   With zones (mechanicalModel.C:829-845 with more than one law) each
   sub-mesh interpolates from its own cells, and a point takes the mean of
   the values of the materials around it (mapSubMeshPointFields,
   solidSubMeshesTemplates.C:271-336: value/noMat summed over the
   sub-meshes, here in the order of their first cell in the point's list). */
void or_point_interpolate(const or_case *c, const double *D, const double *G, double *out)
{
    const or_mesh *m = c->m;
    const int P = m->nPoints;
    int *cnt = calloc((size_t)P + 1, sizeof(int));
    for (int f = 0; f < m->nFaces; ++f)
        for (int k = m->fOff[f]; k < m->fOff[f+1]; ++k) cnt[m->fVerts[k] + 1] += (f < m->nInternalFaces) ? 2 : 1;
    for (int p = 0; p < P; ++p) cnt[p+1] += cnt[p];
    int *lst = malloc(sizeof(int)*(size_t)(cnt[P] > 0 ? cnt[P] : 1));
    int *fill = malloc(sizeof(int)*(size_t)(P > 0 ? P : 1));
    for (int p = 0; p < P; ++p) fill[p] = cnt[p];
    for (int f = 0; f < m->nFaces; ++f)
        for (int k = m->fOff[f]; k < m->fOff[f+1]; ++k) {
            int p = m->fVerts[k];
            lst[fill[p]++] = m->owner[f];
            if (f < m->nInternalFaces) lst[fill[p]++] = m->neighbour[f];
        }
    for (int p = 0; p < P; ++p) {
        int a = cnt[p], b = cnt[p+1];
        /* insertion sort (a point has few cells), then skip repeats */
        for (int i = a + 1; i < b; ++i) {
            int v = lst[i], j = i - 1;
            while (j >= a && lst[j] > v) { lst[j+1] = lst[j]; --j; }
            lst[j+1] = v;
        }
        const double *x = m->points + 3*(size_t)p;
        /* the materials around the point, in the order of their first cell */
        int zl[64], nz = 0;
        for (int i = a; i < b; ++i) {
            const int z = c->zone ? c->zone[lst[i]] : 0;
            int seen = 0;
            for (int k = 0; k < nz; ++k) seen |= zl[k] == z;
            if (!seen && nz < 64) zl[nz++] = z;
        }
        double res[3] = {0, 0, 0};
        for (int k = 0; k < nz; ++k) {
            double acc[3] = {0, 0, 0}, sw = 0.0;
            for (int i = a; i < b; ++i) {
                if (i > a && lst[i] == lst[i-1]) continue;
                int cc = lst[i];
                if (c->zone && c->zone[cc] != zl[k]) continue;
                double d[3] = {x[0] - m->C[3*cc], x[1] - m->C[3*cc+1], x[2] - m->C[3*cc+2]};
                double w = 1.0/sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2]);
                for (int j = 0; j < 3; ++j)
                    acc[j] += w*(D[3*cc+j] + d[0]*G[9*cc+j] + d[1]*G[9*cc+3+j] + d[2]*G[9*cc+6+j]);
                sw += w;
            }
            for (int j = 0; j < 3; ++j) {
                const double v = sw > 1e-300 ? acc[j]/sw : acc[j];
                res[j] = nz == 1 ? v : res[j] + v/nz;
            }
        }
        for (int j = 0; j < 3; ++j) out[3*(size_t)p+j] = res[j];
    }
    free(cnt); free(lst); free(fill);
}

/* One outer iteration of linGeomTotalDispSolid::evolve (:138-209).
   Returns 1 if converged per solidModel::converged. */
int or_iterate(or_case *c, const or_opts *o, double *initResOut, double *rvfOut, int *innerOut)
{
    const or_mesh *m = c->m;
    const int C = m->nCells;
    memcpy(c->Dprev, c->D, sizeof(double)*3*C);           /* storePrevIter */
    or_update_traction(c);                                  /* BC updateCoeffs at assembly */
    or_residual(c);                                         /* explicit terms */
    /* rho*fvm::d2dt2(D) on the left and rho*g on the right
       (linGeomTotalDispSolid.C:141-149).  [OF] EulerD2dt2Scheme::fvmD2dt2:
       diag = coefft rDeltaT2 V rho, source = rDeltaT2 V rho (coefft0 D.oldTime()
       - coefft00 D.oldTime().oldTime()), coefft = (dt + dt0)/(2 dt),
       coefft00 = (dt + dt0)/(2 dt0), coefft0 = coefft + coefft00,
       rDeltaT2 = 4/(dt + dt0)^2.  In defect form the matrix gains the diagonal
       and the explicit residual gains source - diag D + rho g V. */
    /* backwardD2dt2Scheme::fvmD2dt2 (backwardD2dt2Scheme.C:227-300; constant
       step only, :238-243): coefft rDeltaT fvmDdt(D) [OF backwardDdtScheme]
       + rDeltaT V (coefft0 fvcDdt(D.oldTime()) - coefft00 fvcDdt(D.oldTime().oldTime())),
       coefft = 1 + dt/(dt + dt0), coefft00 = dt^2/(dt0 (dt + dt0)),
       coefft0 = coefft + coefft00 (:112-117); [OF] backward fvmDdt(X): diag
       coefft V/dt, source V/dt (coefft0 X0 - coefft00 X00); fvcDdt(X) =
       (coefft X - coefft0 X0 + coefft00 X00)/dt.  Diagonal coefft^2 V rho/dt^2. */
    {
        const double *V = m->V;
        double ct = 0.0, ct0 = 0.0, ct00 = 0.0, r2 = 0.0;
        double bt = 0.0, bt0 = 0.0, bt00 = 0.0, rdt = 0.0;
        if (o->d2dt2 == 1) {
            const double dt = o->deltaT, dt0 = o->deltaT0 > 0.0 ? o->deltaT0 : o->deltaT;
            ct = (dt + dt0)/(2.0*dt); ct00 = (dt + dt0)/(2.0*dt0); ct0 = ct + ct00;
            r2 = 4.0/((dt + dt0)*(dt + dt0));
        } else if (o->d2dt2 == 2) {
            const double dt = o->deltaT, dt0 = dt;
            bt = 1.0 + dt/(dt + dt0); bt00 = dt*dt/(dt0*(dt + dt0)); bt0 = bt + bt00;
            rdt = 1.0/dt;
        }
        for (int i = 0; i < C; ++i) {
            const double rv = c->rho[i]*V[i];
            const double iner = o->d2dt2 == 2 ? bt*rdt*bt*rdt*rv : ct*r2*rv;
            c->diag[i] = c->diag0[i] + iner;
            for (int q = 0; q < 3; ++q) {
                const int k = 3*i+q;
                double add = rv*o->g[q];
                if (o->d2dt2 == 1)
                    add += r2*rv*(ct0*c->Dold[k] - ct00*c->Doo[k]) - iner*c->D[k];
                else if (o->d2dt2 == 2) {
                    const double fvmSrc = bt*rdt*rdt*rv*(bt0*c->Dold[k] - bt00*c->Doo[k]);
                    const double ddt0 = rdt*(bt*c->Dold[k] - bt0*c->Doo[k] + bt00*c->Dooo[k]);
                    const double ddt00 = rdt*(bt*c->Doo[k] - bt0*c->Dooo[k] + bt00*c->Doooo[k]);
                    add += fvmSrc + rdt*rv*(bt0*ddt0 - bt00*ddt00) - iner*c->D[k];
                }
                c->R[k] += add;
            }
        }
    }
    /* DEqn.solve(): segregated PCG on M x = M D + R, warm start x = D */
    double *rD = malloc(sizeof(double)*C);
    if (o->precond == OR_PRECOND_DIC) dic_rD(c, rD);
    else for (int i = 0; i < C; ++i) rD[i] = 1.0/c->diag[i];
    double *psi = malloc(sizeof(double)*C), *src = malloc(sizeof(double)*C), *Ap = malloc(sizeof(double)*C);
    double *work = malloc(sizeof(double)*4*C);
    double ir[3], fr[3]; int nit[3];
    for (int q = 0; q < 3; ++q) {
        if (q == c->emptyDir) { ir[q] = 0.0; fr[q] = 0.0; nit[q] = 0; continue; }  /* not a solved component */
        for (int i = 0; i < C; ++i) psi[i] = c->D[3*i+q];
        amul(c, psi, Ap);
        for (int i = 0; i < C; ++i) src[i] = Ap[i] + c->R[3*i+q];
        nit[q] = pcg(c, o, rD, psi, src, &ir[q], &fr[q], work);
        for (int i = 0; i < C; ++i) c->D[3*i+q] = psi[i];
    }
    free(rD); free(psi); free(src); free(Ap); free(work);
    const int aitken = o->relaxAitken != 0;
    const int relax = aitken || (o->relaxD > 0.0 && o->relaxD != 1.0);
    double *DbPrev = NULL;
    if (relax) {                                            /* boundary values of D.prevIter() */
        DbPrev = malloc(sizeof(double)*3*c->nB);
        memcpy(DbPrev, c->Db, sizeof(double)*3*c->nB);
    }
    or_correct_boundary(c);                                 /* psi.correctBoundaryConditions() */
    if (relax) {
        if (aitken) {
            /* relaxField, Aitken (solidModel.C:842-897); aitkenAlpha is a
               volScalarField and aitkenResidual a volVectorField, so both act
               per cell and per boundary face:
                 r = D.prevIter() - D
                 alpha = a0 (iCorr 0: the D relaxation factor, or 1), else
                 alpha = clip(alpha (r_prev . (r_prev - r)) / (|r_prev - r|^2 + SMALL), 0, 2)
                 D -= alpha r
               Fixed-value patches ignore the subtraction (fixedValue disables
               its assignment operators); the other patch values take it. */
            const double a0 = o->relaxD > 0.0 ? o->relaxD : 1.0;
            for (int i = 0; i < C + c->nB; ++i) {
                const int bnd = i >= C, b = i - C;
                if (bnd && (c->bType[b] == OR_EMPTY || c->bType[b] == OR_FIXED)) continue;
                double *Dv = bnd ? c->Db + 3*b : c->D + 3*i;
                const double *Dp = bnd ? DbPrev + 3*b : c->Dprev + 3*i;
                double *rv = bnd ? c->aResB + 3*b : c->aRes + 3*i;
                double *al = bnd ? c->aAlphaB + b : c->aAlpha + i;
                double r[3] = {Dp[0] - Dv[0], Dp[1] - Dv[1], Dp[2] - Dv[2]};
                double alpha = a0;
                if (c->iCorr != 0) {
                    double d[3] = {rv[0] - r[0], rv[1] - r[1], rv[2] - r[2]};
                    alpha = *al*dot3(rv, d)/(dot3(d, d) + OR_SMALL);
                    alpha = alpha > 2.0 ? 2.0 : (alpha < 0.0 ? 0.0 : alpha);
                }
                *al = alpha;
                for (int q = 0; q < 3; ++q) { rv[q] = r[q]; Dv[q] -= alpha*r[q]; }
            }
        } else {
            /* relaxField, fixed method (solidModel.C:837-841): D.relax(), i.e.
               D == prevIter + a (D - prevIter) on cells and boundary values */
            const double a = o->relaxD;
            for (int i = 0; i < 3*C; ++i) c->D[i] = c->Dprev[i] + a*(c->D[i] - c->Dprev[i]);
            for (int b = 0; b < c->nB; ++b) {
                if (c->bType[b] == OR_EMPTY || c->bType[b] == OR_FIXED) continue;
                for (int q = 0; q < 3; ++q) c->Db[3*b+q] = DbPrev[3*b+q] + a*(c->Db[3*b+q] - DbPrev[3*b+q]);
            }
        }
        /* the snGrads of fixedDisplacement (fixedDisplacementFvPatchVectorField.C:
           315-319), solidSymmetry and the zero-shear patch are evaluated
           afterwards from the relaxed cell value */
        for (int b = 0; b < c->nB; ++b) {
            int P = bCell(c, b);
            if (c->bType[b] == OR_EMPTY || c->bType[b] == OR_TRACTION) continue;
            double kG[3]; vdotT(c->bK + 3*b, c->G + 9*P, kG);
            if (c->bType[b] == OR_ZERO_SHEAR) {
                double v[3]; zero_shear_value(c, b, c->D + 3*P, kG, v);
                for (int q = 0; q < 3; ++q) c->snb[3*b+q] = (v[q] - (c->D[3*P+q] + kG[q]))*c->bDelta[b];
            } else if (c->bType[b] == OR_SYMMETRY) {
                const double *n = c->bN + 3*b, *DP = c->D + 3*P;
                double nD = dot3(n, DP);
                for (int q = 0; q < 3; ++q) c->snb[3*b+q] = -n[q]*nD*c->bDelta[b];
            } else {
                for (int q = 0; q < 3; ++q)
                    c->snb[3*b+q] = (c->Db[3*b+q] - (c->D[3*P+q] + kG[q]))*c->bDelta[b];
            }
        }
        free(DbPrev);
    }
    iface_update(c, c->D, c->Dprev);                        /* sub-meshes: interface displacement */
    or_grad(c);                                             /* mechanical().grad(D, gradD) */
    or_sigma(c);                                            /* mechanical().correct(sigma) */
    iface_sigma(c);                                         /* sub-meshes: their interface stress */
    /* solidModel::converged (solidModelTemplates.C:26-188), total approach */
    double initRes = sqrt(ir[0]*ir[0] + ir[1]*ir[1] + ir[2]*ir[2]);
    int nIter = nit[0] > nit[1] ? nit[0] : nit[1]; if (nit[2] > nIter) nIter = nit[2];
    double denom = maxmag(c->D, c->Dold, C);
    if (denom < OR_SMALL) { denom = maxmag(c->D, NULL, C); if (denom < OR_SMALL) denom = OR_SMALL; }
    double rvf = maxmag(c->D, c->Dprev, C)/denom;
    int conv = 0;
    if (c->iCorr > 1) {   /* materialResidual = 0 for linearElastic */
        if (initRes < o->solTol && rvf < o->solTol) conv = 1;
        else if (rvf < o->altTol) conv = 1;
        else if (initRes < o->altTol) conv = 1;
    }
    if (initResOut) *initResOut = initRes;
    if (rvfOut) *rvfOut = rvf;
    if (innerOut) *innerOut = nit[0] + nit[1] + nit[2];
    (void)nIter;
    return conv;
}

/* Full momentum loop (:135-209) from the current state. */
int or_solve(or_case *c, const or_opts *o, or_report *rep)
{
    memset(rep, 0, sizeof(*rep));
    int conv = 0;
    do {
        double ir, rvf; int inner;
        conv = or_iterate(c, o, &ir, &rvf, &inner);
        if (o->historyCap > c->iCorr) {
            if (o->histInitRes) o->histInitRes[c->iCorr] = ir;
            if (o->histRvf) o->histRvf[c->iCorr] = rvf;
            if (o->histInner) o->histInner[c->iCorr] = inner;
        }
        rep->initRes = ir; rep->rvf = rvf; rep->innerTotal += inner;
        rep->nOuter = c->iCorr + 1;
        if (!isfinite(ir) || !isfinite(rvf)) { rep->converged = 0; return -1; }
        if (!conv && c->iCorr == o->nCorr - 1) rep->maxIterReached = 1;
    } while (!conv && ++c->iCorr < o->nCorr);
    /* mechanical().interpolate(D, gradD, pointD) after the loop
       (linGeomTotalDispSolid.C:211-212) updates the interface displacement
       once more (mechanicalModel.C:851-852) */
    iface_update(c, c->D, c->Dprev);
    rep->converged = conv;
    return conv;
}

/* start a new time step: the converged field becomes D.oldTime() and the
   previous old-time field D.oldTime().oldTime(); the loop counter restarts */
void or_advance_time(or_case *c)
{
    const int C = c->m->nCells;
    memcpy(c->Doooo, c->Dooo, sizeof(double)*3*C);
    memcpy(c->Dooo, c->Doo, sizeof(double)*3*C);
    memcpy(c->Doo, c->Dold, sizeof(double)*3*C);
    memcpy(c->Dold, c->D, sizeof(double)*3*C);
    c->iCorr = 0;
}

/* timing helper for the CPU baseline: n outer iterations, wall seconds */
double or_time_iterations(or_case *c, const or_opts *o, int n)
{
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < n; ++i) { or_iterate(c, o, NULL, NULL, NULL); c->iCorr++; }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9*(double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------------------ */
/* field access                                                         */
/* ------------------------------------------------------------------ */
static double *field_ptr(or_case *c, const char *name, size_t *n)
{
    const int C = c->m->nCells, F = c->m->nInternalFaces, B = c->nB;
    #define F_(nm, p, cnt) if (!strcmp(name, nm)) { *n = (size_t)(cnt); return p; }
    F_("D", c->D, 3*C) F_("Dprev", c->Dprev, 3*C) F_("Dold", c->Dold, 3*C) F_("Db", c->Db, 3*B)
    F_("G", c->G, 9*C) F_("Gb", c->Gb, 9*B) F_("sigma", c->sig, 6*C) F_("sigmab", c->sigb, 6*B)
    F_("g", c->g, 3*B) F_("snb", c->snb, 3*B) F_("R", c->R, 3*C)
    F_("diag", c->diag, C) F_("upper", c->upper, F) F_("w", c->w, F) F_("nonOrthDelta", c->nonOrthDelta, F)
    F_("corrVec", c->corrVec, 3*F) F_("bDelta", c->bDelta, B) F_("bK", c->bK, 3*B) F_("bN", c->bN, 3*B)
    F_("lsP", c->lsP, 3*F) F_("lsN", c->lsN, 3*F) F_("lsB", c->lsB, 3*B)
    F_("impK", c->impK, C) F_("impKf", c->impKf, F) F_("gammaf", c->gammaf, F)
    F_("mu", c->mu, C) F_("lambda", c->lambda, C) F_("K", c->K, C)
    F_("aitkenAlpha", c->aAlpha, C) F_("aitkenResidual", c->aRes, 3*C)
    F_("rho", c->rho, C) F_("Doo", c->Doo, 3*C) F_("Dooo", c->Dooo, 3*C) F_("Doooo", c->Doooo, 3*C)
    F_("aitkenAlphab", c->aAlphaB, B) F_("aitkenResidualb", c->aResB, 3*B)
    F_("ifDI", c->ifDI, 3*c->nIf) F_("ifSigma", c->ifSig, 12*c->nIf) F_("ifLs", c->ifLs, 6*c->nIf)
    F_("ifDc", c->ifDc, 2*c->nIf)
    #undef F_
    *n = 0; return NULL;
}

long or_field_size(or_case *c, const char *name) { size_t n; return field_ptr(c, name, &n) ? (long)n : -1; }
int or_get_field(or_case *c, const char *name, double *out)
{ size_t n; double *p = field_ptr(c, name, &n); if (!p) return -1; memcpy(out, p, n*sizeof(double)); return 0; }
int or_set_field(or_case *c, const char *name, const double *in)
{ size_t n; double *p = field_ptr(c, name, &n); if (!p) return -1; memcpy(p, in, n*sizeof(double)); return 0; }

/* y = M x for one scalar field (C values) */
void or_amul(or_case *c, const double *x, double *y) { amul(c, x, y); }

/* von Mises equivalent stress sqrt(3/2 |dev sigma|^2) (solidModel.C:1957-1961) */
void or_sigma_eq(or_case *c, double *out)
{
    for (int i = 0; i < c->m->nCells; ++i) {
        const double *s = c->sig + 6*i;
        double tr3 = (s[0] + s[3] + s[5])/3.0;
        double d0 = s[0]-tr3, d3 = s[3]-tr3, d5 = s[5]-tr3;
        double ms = d0*d0 + d3*d3 + d5*d5 + 2.0*(s[1]*s[1] + s[2]*s[2] + s[4]*s[4]);
        out[i] = sqrt(1.5*ms);
    }
}


/* ------------------------------------------------------------------ */
/* OpenMP restatement of the outer iteration, for the CPU baseline      */
/* ------------------------------------------------------------------ */
/* The outer iteration of or_iterate (steady; fixedDisplacement, solidTraction
   and empty patches; no field relaxation) doing the work the reference does
   per iteration, on nThreads cores as a decomposed run would:
     * grad(D) four times: the explicit mechanical().grad (linGeomTotalDispSolid.C:173)
       and the non-orthogonal corrections of fvm::laplacian (:144),
       fvc::laplacian (:145) and the stabilisation (momentumStabilisation.C:214);
     * the implicit Laplacian's coefficients re-assembled (:144);
     * the DEqn source from the literal face fluxes: fvm correction,
       fvc::laplacian, Gauss div(sigma), Rhie-Chow, boundary coefficients (the
       source or_reference_source forms, equal to M D + R);
     * PCG with DIC restricted to each thread's contiguous block of cells (the
       block-diagonal DIC of a decomposed PCG/DIC run), rebuilt per component
       as PCG::solve constructs its preconditioner;
     * boundary update, grad(D), Hooke, convergence norms.
   Face fluxes are formed per face and gathered per cell (no atomics), so for
   a fixed thread count the result is deterministic.  PCG iteration counts
   differ from the serial DIC oracle (block DIC); the fixed point does not. */
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int nT;
    int *cOff, *cFace;      /* cell -> internal faces in face order: f+1 (owner side) or -(f+1) */
    int *bOff, *bList;      /* cell -> boundary faces */
    int *blk0;              /* nT+1 block starts (cells) */
    int *bfOff, *bfList;    /* per block: its internal faces (both cells inside), face order */
    double *G2, *flux, *upper, *diag, *rowSum, *rD, *b3, *work;
} or_omp;

void or_omp_free(or_omp *p)
{
    if (!p) return;
    free(p->cOff); free(p->cFace); free(p->bOff); free(p->bList); free(p->blk0); free(p->bfOff); free(p->bfList);
    free(p->G2); free(p->flux); free(p->upper); free(p->diag); free(p->rowSum); free(p->rD); free(p->b3);
    free(p->work);
    free(p);
}

/* addressing and scratch of the parallel path; NULL if the case has patch
   types it does not cover */
or_omp *or_omp_create(const or_case *c, int nThreads)
{
    const or_mesh *m = c->m;
    const int C = m->nCells, F = m->nInternalFaces, B = c->nB;
    for (int b = 0; b < B; ++b)
        if (c->bType[b] != OR_FIXED && c->bType[b] != OR_TRACTION && c->bType[b] != OR_EMPTY) return NULL;
    or_omp *p = xcalloc(1, sizeof(or_omp));
    p->nT = nThreads < 1 ? 1 : nThreads;
    p->cOff = xcalloc((size_t)C + 1, sizeof(int));
    p->cFace = xcalloc((size_t)2*F, sizeof(int));
    for (int f = 0; f < F; ++f) { p->cOff[m->owner[f]+1]++; p->cOff[m->neighbour[f]+1]++; }
    for (int i = 0; i < C; ++i) p->cOff[i+1] += p->cOff[i];
    int *fill = xcalloc(C, sizeof(int));
    for (int f = 0; f < F; ++f) {
        int o = m->owner[f], n = m->neighbour[f];
        p->cFace[p->cOff[o] + fill[o]++] = f + 1;
        p->cFace[p->cOff[n] + fill[n]++] = -(f + 1);
    }
    p->bOff = xcalloc((size_t)C + 1, sizeof(int));
    p->bList = xcalloc(B, sizeof(int));
    for (int b = 0; b < B; ++b) p->bOff[bCell(c, b)+1]++;
    for (int i = 0; i < C; ++i) p->bOff[i+1] += p->bOff[i];
    memset(fill, 0, sizeof(int)*C);
    for (int b = 0; b < B; ++b) { int P = bCell(c, b); p->bList[p->bOff[P] + fill[P]++] = b; }
    free(fill);
    p->blk0 = xcalloc((size_t)p->nT + 1, sizeof(int));
    for (int t = 0; t <= p->nT; ++t) p->blk0[t] = (int)((long)C*t/p->nT);
    /* owner < neighbour: a face lies in block t iff its owner is in t and its
       neighbour below the block's end */
    p->bfOff = xcalloc((size_t)p->nT + 1, sizeof(int));
    p->bfList = xcalloc(F, sizeof(int));
    int k = 0, t = 0;
    for (int f = 0; f < F; ++f) {
        int o = m->owner[f], n = m->neighbour[f];
        while (o >= p->blk0[t+1]) p->bfOff[++t] = k;
        if (n < p->blk0[t+1]) p->bfList[k++] = f;
    }
    while (t < p->nT) p->bfOff[++t] = k;
    p->G2 = xcalloc((size_t)9*C, sizeof(double));
    p->flux = xcalloc((size_t)3*F, sizeof(double));
    p->upper = xcalloc(F, sizeof(double));
    p->diag = xcalloc(C, sizeof(double)); p->rowSum = xcalloc(C, sizeof(double));
    p->rD = xcalloc(C, sizeof(double));
    p->b3 = xcalloc((size_t)3*C, sizeof(double));
    p->work = xcalloc((size_t)6*C, sizeof(double));
    return p;
}

/* leastSquares grad(D) with the boundary values Db into G (cell gather) */
static void omp_grad(const or_case *c, const or_omp *p, double *G)
{
    const or_mesh *m = c->m;
    const int C = m->nCells;
    #pragma omp parallel for schedule(static) num_threads(p->nT)
    for (int P = 0; P < C; ++P) {
        double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int k = p->cOff[P]; k < p->cOff[P+1]; ++k) {
            const int s = p->cFace[k], f = (s > 0 ? s : -s) - 1;
            const int o = m->owner[f], n = m->neighbour[f];
            double dv[3];
            for (int q = 0; q < 3; ++q) dv[q] = c->D[3*n+q] - c->D[3*o+q];
            const double *l = s > 0 ? c->lsP + 3*f : c->lsN + 3*f;
            const double sg = s > 0 ? 1.0 : -1.0;
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) g[3*i+j] += sg*l[i]*dv[j];
        }
        for (int k = p->bOff[P]; k < p->bOff[P+1]; ++k) {
            const int b = p->bList[k];
            if (c->bType[b] == OR_EMPTY) continue;
            double dv[3];
            for (int q = 0; q < 3; ++q) dv[q] = c->Db[3*b+q] - c->D[3*P+q];
            for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) g[3*i+j] += c->lsB[3*b+i]*dv[j];
        }
        memcpy(G + 9*P, g, sizeof g);
    }
}

static void omp_amul(const or_case *c, const or_omp *p, const double *x, double *y)
{
    const or_mesh *m = c->m;
    #pragma omp parallel for schedule(static) num_threads(p->nT)
    for (int P = 0; P < m->nCells; ++P) {
        double s = p->diag[P]*x[P];
        for (int k = p->cOff[P]; k < p->cOff[P+1]; ++k) {
            const int sf = p->cFace[k], f = (sf > 0 ? sf : -sf) - 1;
            s += p->upper[f]*x[sf > 0 ? m->neighbour[f] : m->owner[f]];
        }
        y[P] = s;
    }
}

/* [OF PCG::solve] + [OF normFactor] for one component, block DIC */
static int omp_pcg(const or_case *c, or_omp *p, const or_opts *o, double *psi, const double *src, double *initRes)
{
    const or_mesh *m = c->m;
    const int N = m->nCells;
    double *pA = p->work, *wA = p->work + N, *rA = p->work + 2*N;
    /* DIC factor per block ([OF DICPreconditioner::calcReciprocalD]) */
    #pragma omp parallel for schedule(static, 1) num_threads(p->nT)
    for (int t = 0; t < p->nT; ++t) {
        for (int i = p->blk0[t]; i < p->blk0[t+1]; ++i) p->rD[i] = p->diag[i];
        for (int k = p->bfOff[t]; k < p->bfOff[t+1]; ++k) {
            const int f = p->bfList[k], l = m->owner[f], u = m->neighbour[f];
            p->rD[u] -= p->upper[f]*p->upper[f]/p->rD[l];
        }
        for (int i = p->blk0[t]; i < p->blk0[t+1]; ++i) p->rD[i] = 1.0/p->rD[i];
    }
    omp_amul(c, p, psi, wA);
    double avg = 0.0;
    #pragma omp parallel for schedule(static) num_threads(p->nT) reduction(+:avg)
    for (int i = 0; i < N; ++i) { rA[i] = src[i] - wA[i]; avg += psi[i]; }
    avg /= N;
    double nf = 0.0, s = 0.0;
    #pragma omp parallel for schedule(static) num_threads(p->nT) reduction(+:nf, s)
    for (int i = 0; i < N; ++i) {
        const double t = p->rowSum[i]*avg;
        nf += fabs(wA[i] - t) + fabs(src[i] - t);
        s += fabs(rA[i]);
    }
    nf += 1e-20;
    *initRes = s/nf;
    double fr = *initRes;
    int nIter = 0;
    #define OMP_CONV(fr, ir) ((fr) < o->pcgTol || (o->pcgRelTol > 1e-20 && (fr) < o->pcgRelTol*(ir)))
    if (!OMP_CONV(fr, *initRes)) {
        double wArA = 1e20, wArAold;
        do {
            wArAold = wArA;
            #pragma omp parallel for schedule(static, 1) num_threads(p->nT)
            for (int t = 0; t < p->nT; ++t) {
                for (int i = p->blk0[t]; i < p->blk0[t+1]; ++i) wA[i] = p->rD[i]*rA[i];
                for (int k = p->bfOff[t]; k < p->bfOff[t+1]; ++k) {
                    const int f = p->bfList[k], l = m->owner[f], u = m->neighbour[f];
                    wA[u] -= p->rD[u]*p->upper[f]*wA[l];
                }
                for (int k = p->bfOff[t+1] - 1; k >= p->bfOff[t]; --k) {
                    const int f = p->bfList[k], l = m->owner[f], u = m->neighbour[f];
                    wA[l] -= p->rD[l]*p->upper[f]*wA[u];
                }
            }
            wArA = 0.0;
            #pragma omp parallel for schedule(static) num_threads(p->nT) reduction(+:wArA)
            for (int i = 0; i < N; ++i) wArA += wA[i]*rA[i];
            const double beta = nIter == 0 ? 0.0 : wArA/wArAold;
            #pragma omp parallel for schedule(static) num_threads(p->nT)
            for (int i = 0; i < N; ++i) pA[i] = wA[i] + beta*pA[i];
            omp_amul(c, p, pA, wA);
            double wApA = 0.0;
            #pragma omp parallel for schedule(static) num_threads(p->nT) reduction(+:wApA)
            for (int i = 0; i < N; ++i) wApA += wA[i]*pA[i];
            if (fabs(wApA)/nf < OR_VSMALL) break;
            const double alpha = wArA/wApA;
            s = 0.0;
            #pragma omp parallel for schedule(static) num_threads(p->nT) reduction(+:s)
            for (int i = 0; i < N; ++i) { psi[i] += alpha*pA[i]; rA[i] -= alpha*wA[i]; s += fabs(rA[i]); }
            fr = s/nf;
        } while (++nIter < o->pcgMaxIter && !OMP_CONV(fr, *initRes));
    }
    #undef OMP_CONV
    return nIter;
}

/* One outer iteration on nThreads (or_omp_create); returns the convergence
   flag of solidModel::converged, -1 for options it does not cover. */
int or_iterate_omp(or_case *c, or_omp *p, const or_opts *o, double *initResOut, double *rvfOut, int *innerOut)
{
    const or_mesh *m = c->m;
    const int C = m->nCells, F = m->nInternalFaces, B = c->nB;
    if (o->d2dt2 || o->relaxAitken || (o->relaxD > 0.0 && o->relaxD != 1.0) || c->emptyDir >= 0 || c->nIf) return -1;
    #pragma omp parallel for schedule(static) num_threads(p->nT)
    for (int i = 0; i < 3*C; ++i) c->Dprev[i] = c->D[i];
    or_update_traction(c);
    /* grad(D) of the corrected Laplacians' and the stabilisation's own
       non-orthogonal corrections (the same D: the same field three times) */
    omp_grad(c, p, p->G2);
    omp_grad(c, p, p->G2);
    omp_grad(c, p, p->G2);
    #pragma omp parallel for schedule(static) num_threads(p->nT)
    for (int f = 0; f < F; ++f) p->upper[f] = -c->impKf[f]*m->magSf[f]*c->nonOrthDelta[f];
    /* per face: fvm correction - fvc::laplacian + Gauss div(sigma) + Rhie-Chow */
    #pragma omp parallel for schedule(static) num_threads(p->nT)
    for (int f = 0; f < F; ++f) {
        const int ow = m->owner[f], n = m->neighbour[f];
        const double w = c->w[f];
        double Gf[9], sf[6], kG[3], SG[3];
        for (int q = 0; q < 9; ++q) Gf[q] = w*(p->G2[9*ow+q] - p->G2[9*n+q]) + p->G2[9*n+q];
        for (int q = 0; q < 6; ++q) sf[q] = w*(c->sig[6*ow+q] - c->sig[6*n+q]) + c->sig[6*n+q];
        const double *S = m->Sf + 3*f;
        vdotT(c->corrVec + 3*f, Gf, kG);
        vdotT(S, Gf, SG);
        const double dl = c->nonOrthDelta[f], gm = c->impKf[f]*m->magSf[f], gr = c->gammaf[f]*m->magSf[f];
        const double dv[3] = { S[0]*sf[0] + S[1]*sf[1] + S[2]*sf[2], S[0]*sf[1] + S[1]*sf[3] + S[2]*sf[4],
                               S[0]*sf[2] + S[1]*sf[4] + S[2]*sf[5] };
        for (int q = 0; q < 3; ++q) {
            const double dD = c->D[3*n+q] - c->D[3*ow+q];
            const double corr = gm*kG[q];
            const double lap = gm*(dl*dD + kG[q]);
            const double rc = gr*(dl*dD + kG[q]) - c->gammaf[f]*SG[q];
            p->flux[3*f+q] = dv[q] + rc + corr - lap;
        }
    }
    /* gather: source, matrix diagonal (fvm::laplacian + fixed-patch
       internalCoeffs) and row sums (normFactor) */
    #pragma omp parallel for schedule(static) num_threads(p->nT)
    for (int P = 0; P < C; ++P) {
        double s[3] = {0, 0, 0}, dsum = 0.0, bsum = 0.0;
        for (int k = p->cOff[P]; k < p->cOff[P+1]; ++k) {
            const int sf = p->cFace[k], f = (sf > 0 ? sf : -sf) - 1;
            const double sg = sf > 0 ? 1.0 : -1.0;
            for (int q = 0; q < 3; ++q) s[q] += sg*p->flux[3*f+q];
            dsum -= p->upper[f];
        }
        for (int k = p->bOff[P]; k < p->bOff[P+1]; ++k) {
            const int b = p->bList[k];
            if (c->bType[b] == OR_EMPTY) continue;
            const double *S = m->Sf + 3*(F+b), *sb = c->sigb + 6*b;
            s[0] += S[0]*sb[0] + S[1]*sb[1] + S[2]*sb[2];
            s[1] += S[0]*sb[1] + S[1]*sb[3] + S[2]*sb[4];
            s[2] += S[0]*sb[2] + S[1]*sb[4] + S[2]*sb[5];
            const double gm = c->impK[P]*m->magSf[F+b];
            if (c->bType[b] == OR_FIXED) {
                double kG[3]; vdotT(c->bK + 3*b, c->G + 9*P, kG);
                bsum += gm*c->bDelta[b];
                for (int q = 0; q < 3; ++q) {
                    s[q] += gm*c->bDelta[b]*(c->Db[3*b+q] - kG[q]);
                    s[q] -= gm*(c->Db[3*b+q] - (c->D[3*P+q] + kG[q]))*c->bDelta[b];
                }
            } else {
                for (int q = 0; q < 3; ++q) { s[q] += gm*c->g[3*b+q]; s[q] -= gm*c->g[3*b+q]; }
            }
        }
        p->diag[P] = dsum + bsum;
        p->rowSum[P] = bsum;
        for (int q = 0; q < 3; ++q) p->b3[3*P+q] = s[q];
    }
    /* DEqn.solve(): segregated PCG/DIC, warm start x = D */
    double *psi = p->work + 3*C, *src = p->work + 4*C;
    double ir[3];
    int inner = 0;
    for (int q = 0; q < 3; ++q) {
        #pragma omp parallel for schedule(static) num_threads(p->nT)
        for (int i = 0; i < C; ++i) { psi[i] = c->D[3*i+q]; src[i] = p->b3[3*i+q]; }
        inner += omp_pcg(c, p, o, psi, src, &ir[q]);
        #pragma omp parallel for schedule(static) num_threads(p->nT)
        for (int i = 0; i < C; ++i) c->D[3*i+q] = psi[i];
    }
    or_correct_boundary(c);
    omp_grad(c, p, c->G);
    #pragma omp parallel for schedule(static) num_threads(p->nT)
    for (int b = 0; b < B; ++b) {
        if (c->bType[b] == OR_EMPTY) continue;
        const int P = bCell(c, b);
        const double *n = c->bN + 3*b, *GP = c->G + 9*P;
        double nG[3]; vdotT(n, GP, nG);
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j)
            c->Gb[9*b + 3*i + j] = GP[3*i+j] + n[i]*(c->snb[3*b+j] - nG[j]);
    }
    #pragma omp parallel for schedule(static) num_threads(p->nT)
    for (int i = 0; i < C; ++i) hooke(c->G + 9*i, c->mu[i], c->K[i], c->sig + 6*i);
    #pragma omp parallel for schedule(static) num_threads(p->nT)
    for (int b = 0; b < B; ++b) { const int P = bCell(c, b); hooke(c->Gb + 9*b, c->mu[P], c->K[P], c->sigb + 6*b); }
    /* solidModel::converged (total approach) */
    double dOld = 0.0, dNow = 0.0, dPrev = 0.0;
    #pragma omp parallel for schedule(static) num_threads(p->nT) reduction(max:dOld, dNow, dPrev)
    for (int i = 0; i < C; ++i) {
        double a[3], b[3], e[3];
        for (int q = 0; q < 3; ++q) {
            a[q] = c->D[3*i+q] - c->Dold[3*i+q]; b[q] = c->D[3*i+q]; e[q] = c->D[3*i+q] - c->Dprev[3*i+q];
        }
        const double va = mag3(a), vb = mag3(b), ve = mag3(e);
        if (va > dOld || va != va) dOld = va;
        if (vb > dNow || vb != vb) dNow = vb;
        if (ve > dPrev || ve != ve) dPrev = ve;
    }
    double denom = dOld;
    if (denom < OR_SMALL) { denom = dNow; if (denom < OR_SMALL) denom = OR_SMALL; }
    const double rvf = dPrev/denom;
    const double initRes = sqrt(ir[0]*ir[0] + ir[1]*ir[1] + ir[2]*ir[2]);
    int conv = 0;
    if (c->iCorr > 1) {
        if (initRes < o->solTol && rvf < o->solTol) conv = 1;
        else if (rvf < o->altTol) conv = 1;
        else if (initRes < o->altTol) conv = 1;
    }
    if (initResOut) *initResOut = initRes;
    if (rvfOut) *rvfOut = rvf;
    if (innerOut) *innerOut = inner;
    return conv;
}

/* CPU-baseline timing: n outer iterations of or_iterate_omp on nThreads,
   wall seconds (negative: options not covered) */
double or_time_iterations_omp(or_case *c, const or_opts *o, int n, int nThreads)
{
    or_omp *p = or_omp_create(c, nThreads);
    if (!p) return -1.0;
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int i = 0; i < n; ++i) {
        if (or_iterate_omp(c, p, o, NULL, NULL, NULL) < 0) { or_omp_free(p); return -1.0; }
        c->iCorr++;
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    or_omp_free(p);
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9*(double)(t1.tv_nsec - t0.tv_nsec);
}

/* the outer-iteration counter (iCorr) of a loop driven from outside */
void or_bump_icorr(or_case *c) { c->iCorr++; }
