// s4f_device.h -- device-side data structures shared by the HIP sources.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace s4f {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kSlice = 64;           // rows per sliced-ELL slice (one wavefront)
constexpr int kPad = INT32_MIN;      // padding entry in the ELL column array
constexpr int kMaxPartials = 16;     // reduction values per block
constexpr double kSmall = 1e-15;     // OpenFOAM SMALL (double)
constexpr double kVSmall = 1e-300;   // OpenFOAM VSMALL (double)

// Streamed data (matrix and face entries, read once per launch; outputs that
// the same launch does not read again) goes through non-temporal accesses,
// so the L2 keeps the neighbour rows that the gathers reuse.
template <typename T> __device__ __forceinline__ T ldS(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T> __device__ __forceinline__ void stS(T* p, T v) { __builtin_nontemporal_store(v, p); }

// Device view of one solver mesh (or of a packed batch of meshes).
// Cell-indexed arrays are structure-of-arrays: component k of cell c at
// [k*nCells + c].  ELL entries of slice s, slot k, lane l sit at
// sliceStart[s] + k*64 + l, so one wavefront reads one slot of 64 rows with a
// single coalesced access.  A column >= 0 is the neighbour cell, -1-b a
// boundary face b, kPad an empty slot.  Entries of a row are ordered as
// OpenFOAM accumulates face contributions into that cell: internal faces in
// face order (= ascending other-cell label), then boundary faces.
struct DevMesh {
    int nCells, nBoundary, nSlices, slicesPerBlock, nBlocks;
    size_t nEntries;
    const int* sliceStart;  // first ELL entry of each slice
    const int* sliceWidth;  // slots of each slice
    const int* col;
    const double* eS;      // 3*nEntries: area vector pointing out of the row cell
    const double* eW;      // interpolation weight of the row cell
    const double* eA;      // matrix coefficient a_f = impKf |S| deltaCoeffs (off-diagonal = -a_f)
    const double* eC;      // Rhie-Chow coefficient gamma_f |S| deltaCoeffs
    const float* eA32;
    const double* cC;      // 3*nCells cell centres
    // explicit sources of the momentum equation per cell (nullptr: none):
    // fext = rho V g + inertia source (3*nCells), iner = inertia diagonal (nCells)
    const double* fext;
    const double* iner;
    const double* minv;    // 6*nCells inverse least-squares matrices (xx xy xz yy yz zz)
    const double* diag;
    const double* invDiag;
    const float* diag32;
    const float* invDiag32;
    // material: uniform (mu == nullptr) or per cell
    double mu0, K0, impK0;
    const double* mu;
    const double* K;
    const double* impK;
    // boundary faces
    const int* bCell;
    const int* bType;      // 0 fixed displacement, 1 traction
    const double* bN;      // 3*B unit normals
    const double* bKv;     // 3*B correction vectors k_b = (I - nn).(Cf - C)
    const double* bDelta;  // B deltaCoeffs 1/|n (n.(Cf - C))|
    const double* bLs;     // 3*B least-squares boundary vectors d/|d|^2
    const double* bVal;    // 3*B fixed displacement or traction
    const double* bPres;   // B pressure
    // batching: per-instance segment of cells (nullptr for a single mesh)
    const int* cellInst;   // instance of each cell
    int nInst;
    // material zones of a multi-law case, each solved on its own sub-mesh:
    // the bi-material interface faces between zones (nIf = 0: one material)
    int nIf;
    const int* ifCell;            // 2 nIf: side a (the face owner), side b (the neighbour)
    const double* ifN;            // 3 nIf (SoA) unit normals from a to b
    const double* ifDc;           // 2 nIf sub-mesh patch deltaCoeffs 1/|n.(Cf - C)|, side a then b
    const double* ifLs;           // 3 x 2 nIf (SoA) least-squares boundary vectors of each side
    const int* eIf;               // per ELL entry: interface side 2 i + (0 a, 1 b), or -1
    const unsigned char* sliceIf; // per slice: 1 if it holds interface entries
    const int* zone;              // material zone of each cell (nullptr: one material)
};

// Solve control block, written by the last workgroup of each reducing kernel
// and read by every later kernel in the stream.
struct Ctl {
    int outerDone, innerDone, converged, diverged;
    int iCorr, nOuter, firstInner, warm;
    int dpar;        // which displacement buffer holds the current field
    int sigma0;      // cold start: the stress field is zero until the first outer iteration ends
    int active[4];
    int nIter[4];
    double initRes[4], normFactor[4], finalRes[4];
    double rho[4], alpha[4], beta[4];
    double xbar[4];
    double initResMag, rvf;
    long long innerTotal;
    unsigned counter[8];
    // settings
    double pcgTol, pcgRelTol, solTol, altTol;
    int pcgMaxIter, nCorr, histCap, pcMode;  // pcMode 1: z comes from the V-cycle (k_fin_rho sets beta)
    double relaxD;                           // fixed field under-relaxation factor (1: none)
    int solveMask;                           // solved components (bit j); 2-D leaves the empty axis out
    int relaxAitken;                         // relaxationMethod Aitken (relaxD: its first factor)
};

}  // namespace s4f
