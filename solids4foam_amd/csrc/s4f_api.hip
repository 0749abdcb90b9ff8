// s4f_api.hip -- host side of libs4f: device mesh set-up, material and
// boundary data, the outer defect-correction loop of
// linGeomTotalDispSolid::evolve (linGeomTotalDispSolid.C:111-232) driving the
// HIP kernels, the fine-grained operator entry points and batching.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <exception>
#include <map>
#include <mutex>
#include <thread>
#include <type_traits>
#include <unordered_map>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/s4f_abi.h"
#include "s4f_device.h"
#include "s4f_internal.h"
#include "s4f_kernels.h"
#include "s4f_amg.h"

namespace s4f {

thread_local std::string gLastError;
void setError(const std::string& msg) { gLastError = msg; }

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct OomError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct InvalidError : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

#define HIP_CHECK(x)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (x);                                                                            \
        if (e_ != hipSuccess) {                                                                         \
            if (e_ == hipErrorOutOfMemory) throw OomError(std::string("out of device memory: ") + #x);   \
            throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));                             \
        }                                                                                               \
    } while (0)

// Device allocation owned by a handle.
template <typename T>
struct DArr {
    T* p = nullptr;
    size_t n = 0;
    DArr() = default;
    DArr(const DArr&) = delete;
    DArr& operator=(const DArr&) = delete;
    ~DArr() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
        n = count;
    }
    void upload(const std::vector<T>& h, hipStream_t s) {
        alloc(h.size());
        if (!h.empty()) HIP_CHECK(hipMemcpyAsync(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
    }
    size_t bytes() const { return n * sizeof(T); }
};

}  // namespace s4f

using namespace s4f;

// A context lives until it has been destroyed AND its last mesh/batch is
// gone: destroying it earlier only marks it, so handles never outlive it.
struct s4f_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t capture = nullptr;  // private stream on which launch sequences are captured into graphs
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int nCU = 0;
    int children = 0;
    bool destroyPending = false;
};

// ---------------------------------------------------------------------------
// Scratch of the allocation-free PCG entry point (s4f_pcg3_noalloc): one
// file-scope device array per GPU, holding the control block, the per-
// workgroup reduction partials and the PCG vectors of a mesh of at most
// kNoAllocMaxCells cells.  In host code the array's name is only a stand-in
// for the symbol: its device address is taken once per device with
// hipGetSymbolAddress (at context creation on that device) and only that
// address is ever passed to kernels.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxCells = 262144;  // 64^3, the largest batched-box instance
constexpr size_t kNoAllocCtlBytes = 1024;
constexpr size_t kNoAllocPartBytes = (size_t)8192 * kMaxPartials * sizeof(double);
// r, p, p2, q as FP64 (FP32 sweeps need r, p, p2, q, x: 60 of these 96 bytes per cell)
constexpr size_t kNoAllocVecBytes = (size_t)4 * 3 * kNoAllocMaxCells * sizeof(double);
constexpr size_t kNoAllocScratchBytes = kNoAllocCtlBytes + kNoAllocPartBytes + kNoAllocVecBytes;
static_assert(sizeof(Ctl) <= kNoAllocCtlBytes, "control block exceeds its scratch slot");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex gScratchMu;
void* gScratchPtr[kMaxDevices] = {};

namespace {
// with `device` current: the device address of this device's scratch array
void scratchInit(int device) {
    if (device >= kMaxDevices) throw InvalidError("device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(gScratchMu);
    if (gScratchPtr[device]) return;
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned no address for the scratch array");
    gScratchPtr[device] = p;
}
unsigned char* scratchPtr(int device) {
    std::lock_guard<std::mutex> lk(gScratchMu);
    void* p = (device >= 0 && device < kMaxDevices) ? gScratchPtr[device] : nullptr;
    if (!p) throw HipError("no scratch address for this device (context not created on it)");
    return (unsigned char*)p;
}
}  // namespace

namespace {
void ctxRelease(s4f_ctx* ctx) {
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipStreamDestroy(ctx->stream);
    if (ctx->capture) (void)hipStreamDestroy(ctx->capture);
    delete ctx;
}
void ctxChildGone(s4f_ctx* ctx) {
    if (--ctx->children == 0 && ctx->destroyPending) ctxRelease(ctx);
}
}  // namespace

struct s4f_polymesh {
    std::unique_ptr<PolyMesh> pm;
};

// Per-mesh host staging (kept for batching and for re-uploading material data)
struct HostMesh {
    int C = 0, F = 0, B = 0, nPatches = 0;
    std::vector<int> owner, neighbour, patchStart, patchSize, patchType;
    std::vector<double> w, nonOrthDelta, magSf, bMagSf, bDelta;  // F, F, F, B, B
    // ELL
    int nSlices = 0;
    std::vector<int> sliceStart, sliceWidth, col;
    std::vector<int> entOwn, entNei, entB;  // ELL entry of each face side
    std::vector<double> eS, eW;             // 3E, E
    std::vector<double> cC, minv;           // SoA 3C, 6C
    std::vector<int> bCell, bType;
    std::vector<double> bN, bKv, bLs;       // SoA 3B
    // material / boundary values
    bool uniformMaterial = true;
    bool planeStress = false;  // mechanicalProperties planeStress: the 2-D moduli of linearElastic.C:109-113
    bool lsFullDelta = false;  // least-squares boundary delta Cf - C (extendedLeastSquares) instead of its normal part
    std::vector<double> matE, matNu;        // the E and nu the moduli were computed from (C or 1)
    std::vector<double> mu, K, impK;        // C (or 1)
    std::vector<double> eA, eC;             // E
    std::vector<double> diag;               // C
    std::vector<double> bVal, bPres;        // SoA 3B, B
    // instances packed in this structure (one for a plain mesh): cells start
    // on a 64-row slice boundary, padding rows carry no faces
    std::vector<int> instCell0, instNCell, instB0, instNB;
    std::vector<int> instMask;  // per instance: solved components (bit j: component j; 2-D leaves one out)
    // per instance: workgroup-tile rule (at most tileCap tiles of at least
    // tileMin slices; s4f_mesh_tiles_set, the autotuner's knob)
    std::vector<int> instTileCap, instTileMin;
    // dynamics and body force (single meshes)
    std::vector<double> V;        // cell volumes
    std::vector<double> rho;      // density, C values or 1 (empty: not set)
    double g[3] = {0.0, 0.0, 0.0};
    std::vector<double> diagBase; // the steady diagonal (diag may carry the inertia term)
    std::vector<double> Do, Doo;  // old-time fields (SoA 3C) of a transient step
    std::vector<double> Dooo, Doooo;  // third and fourth old-time levels (backward d2dt2)
    bool hasOld = false;
    // material zones of a multi-law case and their bi-material interface
    // faces (s4f_mesh_zones_set); zone empty: one material
    std::vector<int> zone;
    int nIf = 0;
    std::vector<int> ifCell;              // 2 nIf: side a (owner), side b (neighbour)
    std::vector<double> ifN, ifDc, ifLs;  // SoA 3 nIf; 2 nIf; SoA 3 x 2 nIf
    std::vector<int> eIf;                 // per ELL entry: interface side, -1 (empty: no interfaces)
    std::vector<unsigned char> sliceIf;   // per slice: holds interface entries
    std::vector<double> minv0;            // whole-mesh LS inverses, kept once zones replace some
};

struct SolveState {
    DArr<double> D0, D1, Dold, G, Db, g, snb, Gb, sigb, sig;
    DArr<double> DI, sigI;  // sub-meshes: interface displacement (3 nIf), boundary stress (6 x 2 nIf)
    DArr<double> aAlpha, aRes, aAlphaB, aResB;  // Aitken relaxation state (cells, boundary faces)
    DArr<double> r64, p64, q64, x64, pb64, z64;
    DArr<float> r32, p32, q32, x32, pb32, z32;
    DArr<Ctl> ctl;
    DArr<double> part, hist;
    int histCap = 0;
    std::vector<double> hHist;
    int lastNOuter = 0;
    int lastDpar = -1;  // buffer (0: D0, 1: D1) holding the last solve's field; -1: no solve yet
};

// Aggregation multigrid hierarchy of a mesh's matrix M (host build + device arrays)
struct AmgHostLevel {
    int n = 0, nSlices = 0;
    std::vector<int> sliceStart, sliceWidth, col;
    std::vector<double> a, diag;
    std::vector<int> agg;          // rows -> rows of the next coarser level (-1: padding row; empty on the coarsest)
    std::vector<int> memOff, mem;  // rows -> member rows of the finer level (empty on level 0)
    std::vector<int> inst0, instN; // per instance: first row and row count (instances are contiguous)
    std::vector<int> rowInst;      // row -> instance (-1: padding row)
    std::vector<int> instPass;     // per instance: 1 if it passes unchanged to the next level (already coarsest)
    std::vector<double> memW;      // member weights (smoothed prolongation into this level; empty: all 1)
    // the weighted member lists again as sliced ELL over this level's rows
    // (slot k of a slice: member k of each row; an empty slot repeats the
    // row's first member with a zero weight): the restriction's coalesced form
    std::vector<int> rStart, rWidth, rMem;
    std::vector<double> rW;
    // level 0 of a smoothed hierarchy: the up-sweep operator G = (D/w) P - A P
    // in its own sliced-ELL layout (per slice: first entry and width)
    std::vector<int> gStart, gWidth, gCol;
    std::vector<double> g;
};
struct AmgDevLevel {
    int n = 0, nSlices = 0;
    DArr<int> sliceStart, sliceWidth, col, agg, memOff, mem, rowInst, instPass;
    DArr<int> blkRow0, blkN, invOff;          // coarsest level: per-instance dense blocks
    DArr<double> a64, d64, iD64;
    DArr<float> a32, d32, iD32;
    DArr<double> x64, b64, r64, xs64, inv64;  // inv: explicit inverses (coarsest level only)
    DArr<float> x32, b32, r32, xs32, inv32;
    // level 0 only: q = offdiag(A) P, the off-diagonal magnitudes of each row
    // summed per coarse row, in the fine matrix's own sliced-ELL slots
    // (qCol < 0: empty slot)
    DArr<int> qCol, qWidth;  // qWidth: per slice, the filled q slots (coarse levels)
    DArr<float> q32;
    DArr<double> q64;
    DArr<float> memW32;      // restriction weights into this level (smoothed prolongation)
    DArr<double> memW64;
    DArr<int> rStart, rWidth, rMem;  // the same lists as sliced ELL
    DArr<float> rW32;
    DArr<double> rW64;
    DArr<int> gStart, gWidth, gCol;  // level 0 of a smoothed hierarchy: G = (D/w) P - A P
    DArr<float> g32;
    DArr<double> g64;
    MgLevel view() const {
        MgLevel L{};
        L.n = n; L.nSlices = nSlices;
        L.sliceStart = sliceStart.p; L.sliceWidth = sliceWidth.p; L.col = col.p;
        L.a32 = a32.p; L.a64 = a64.p; L.diag32 = d32.p; L.diag64 = d64.p; L.invD32 = iD32.p; L.invD64 = iD64.p;
        L.agg = agg.p; L.memOff = memOff.p; L.mem = mem.p; L.memW32 = memW32.p; L.memW64 = memW64.p;
        L.rStart = rStart.p; L.rWidth = rWidth.p; L.rMem = rMem.p; L.rW32 = rW32.p; L.rW64 = rW64.p;
        L.rowInst = rowInst.p; L.blkRow0 = blkRow0.p; L.blkN = blkN.p; L.invOff = invOff.p; L.pass = instPass.p;
        L.qCol = qCol.p; L.q32 = q32.p; L.q64 = q64.p; L.qWidth = qWidth.p;
        return L;
    }
};
struct Amg {
    unsigned gen = ~0u;  // material generation the hierarchy was built for
    bool smoothed = false;  // smoothed prolongation at the first coarsening
    std::vector<AmgHostLevel> h;
    std::vector<std::unique_ptr<AmgDevLevel>> d;
    double omega = kMgOmega;
    double alpha = 1.0;  // coarse-correction scale (set per solve)
    double buildSeconds = 0.0;
    long long qEntries = 0;  // stored entries of the level-0 grouped operator q
};

struct s4f_mesh {
    s4f_ctx* ctx = nullptr;
    HostMesh h;
    Amg amg;     // plain aggregation (S4F_PRECOND_GAMG; the server packs these)
    Amg amgSm;   // smoothed first prolongation (S4F_PRECOND_GAMG_SMOOTHED)
    Amg* amgLast = nullptr;  // hierarchy of the last multigrid solve
    // device geometry / addressing
    DArr<int> sliceStart, sliceWidth, col, bCell, bType;
    DArr<double> eS, eW, eA, eC, cC, minv, diag, invDiag, mu, K, impK, bN, bKv, bDelta, bLs, bVal, bPres;
    DArr<float> eA32, diag32, invDiag32;
    DArr<int> ifCell, eIf, zone;           // bi-material interfaces (HostMesh)
    DArr<double> ifN, ifDc, ifLs;
    DArr<unsigned char> sliceIf;
    // block map (single instance)
    DArr<int> bmInst, bmS0, bmS1, iBlk0, iNBlk, iCell0, iNCell, iB0, iNB;
    int nBlocks = 0, slicesPerBlock = 0, maxInstNB = 0;
    std::vector<int> tInst, tS0, tS1, tBlk0, tNBlk;  // host copy of the workgroup tiles
    DArr<int> aInst, aS0, aS1, aBlk0, aNBlk, aFin;   // tiles of the instances still iterating (batches)
    bool materialSet = false;
    unsigned generation = 0;  // bumped when material data changes (serving cache key)
    unsigned tileGen = 0;     // bumped when the workgroup-tile rule changes
    double rcScale = 0.1, ifScale = 0.01;  // stabilisation scales baked into eC
    std::vector<char> bcSet;
    SolveState st;
    // momentum-equation sources of the current solve (DevMesh fext / iner)
    DArr<double> fext, iner;
    bool srcOn = false, inerOn = false;
    double inerKey = 0.0;     // coefft rDeltaT2 the device diagonal carries (0: steady)
    unsigned rhoGen = 0, inerRhoGen = 0;
    DevMesh dm() const;
    BlockMap bmap() const;
};

DevMesh s4f_mesh::dm() const {
    DevMesh m{};
    m.nCells = h.C;
    m.nBoundary = h.B;
    m.nSlices = h.nSlices;
    m.slicesPerBlock = slicesPerBlock;
    m.nBlocks = nBlocks;
    m.nEntries = h.col.size();
    m.sliceStart = sliceStart.p;
    m.sliceWidth = sliceWidth.p;
    m.col = col.p;
    m.eS = eS.p;
    m.eW = eW.p;
    m.eA = eA.p;
    m.eC = eC.p;
    m.eA32 = eA32.p;
    m.cC = cC.p;
    m.fext = srcOn ? fext.p : nullptr;
    m.iner = inerOn ? iner.p : nullptr;
    m.minv = minv.p;
    m.diag = diag.p;
    m.invDiag = invDiag.p;
    m.diag32 = diag32.p;
    m.invDiag32 = invDiag32.p;
    if (h.uniformMaterial && !h.mu.empty()) {
        m.mu0 = h.mu[0];
        m.K0 = h.K[0];
        m.impK0 = h.impK[0];
        m.mu = m.K = m.impK = nullptr;
    } else {
        m.mu = mu.p;
        m.K = K.p;
        m.impK = impK.p;
    }
    m.bCell = bCell.p;
    m.bType = bType.p;
    m.bN = bN.p;
    m.bKv = bKv.p;
    m.bDelta = bDelta.p;
    m.bLs = bLs.p;
    m.bVal = bVal.p;
    m.bPres = bPres.p;
    m.cellInst = nullptr;
    m.nInst = 1;
    if (h.nIf > 0) {
        m.nIf = h.nIf;
        m.ifCell = ifCell.p;
        m.ifN = ifN.p;
        m.ifDc = ifDc.p;
        m.ifLs = ifLs.p;
        m.eIf = eIf.p;
        m.sliceIf = sliceIf.p;
    }
    m.zone = h.zone.empty() ? nullptr : zone.p;
    return m;
}

BlockMap s4f_mesh::bmap() const {
    BlockMap b{};
    b.inst = bmInst.p;
    b.s0 = bmS0.p;
    b.s1 = bmS1.p;
    b.iBlk0 = iBlk0.p;
    b.iNBlk = iNBlk.p;
    b.iCell0 = iCell0.p;
    b.iNCell = iNCell.p;
    b.iB0 = iB0.p;
    b.iNB = iNB.p;
    b.finInst = nullptr;
    b.nBlocks = nBlocks;
    return b;
}

namespace {

s4f_status fail(s4f_status s, const std::string& msg) {
    setError(msg);
    return s;
}

template <typename F>
s4f_status guarded(F&& f) {
    try {
        return f();
    } catch (const OomError& e) {
        return fail(S4F_ERR_OOM, e.what());
    } catch (const HipError& e) {
        return fail(S4F_ERR_HIP, e.what());
    } catch (const std::invalid_argument& e) {
        return fail(S4F_ERR_INVALID, e.what());
    } catch (const std::bad_alloc&) {
        return fail(S4F_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return fail(S4F_ERR_HIP, e.what());
    }
}

inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Upper bound on the workgroup tiles of one instance (the reduction partial
// rows its finalize kernel reads): 32 per CU on a 256-CU MI355X.  Smaller
// tiles than 8 per CU keep workgroups that stream neighbouring planes closer
// in time (their shared rows hit L2 more often): at 200^3 the residual,
// gradient, SpMV and multigrid sweeps run 5-10 % faster than with 2048
// tiles, for about 10 us more per finalize (its one workgroup reads 4x the
// partial rows).
int gMaxBlocks = 8192;
// default smallest tile: one slice per wave of a workgroup
constexpr int kMinTileSlices = kBlock / 64;

// tile rule of instance i of h
inline int tileSlices(const HostMesh& h, int i, int nS) {
    const int cap = h.instTileCap.empty() ? gMaxBlocks : h.instTileCap[i];
    const int mn = h.instTileMin.empty() ? kMinTileSlices : h.instTileMin[i];
    return std::max(mn, (nS + cap - 1) / cap);
}


// Host-side precomputation shared by every entry point: OpenFOAM
// surfaceInterpolation weights and deltas, boundary normals/correction
// vectors, least-squares matrices (inverted in FP64) and the sliced-ELL
// cell->face lists.
void setupHostMesh(HostMesh& h, const s4f_mesh_desc& d) {
    const int C = d.nCells, F = d.nInternalFaces, NF = d.nFaces, B = NF - F;
    if (C <= 0 || F < 0 || NF < F || !d.owner || (F && !d.neighbour) || !d.C || !d.V || !d.Sf || !d.Cf)
        throw InvalidError("invalid mesh descriptor");
    if (d.nPatches <= 0 || !d.patchStart || !d.patchSize || !d.patchType) throw InvalidError("mesh needs patches");
    h.C = C; h.F = F; h.B = B; h.nPatches = d.nPatches;
    h.V.assign(d.V, d.V + C);
    h.owner.assign(d.owner, d.owner + NF);
    h.neighbour.assign(d.neighbour, d.neighbour + F);
    h.patchStart.assign(d.patchStart, d.patchStart + d.nPatches);
    h.patchSize.assign(d.patchSize, d.patchSize + d.nPatches);
    h.patchType.assign(d.patchType, d.patchType + d.nPatches);
    {
        int f = F;
        for (int p = 0; p < d.nPatches; ++p) {
            if (h.patchStart[p] != f || h.patchSize[p] < 0) throw InvalidError("patches must tile the boundary faces in order");
            if (h.patchType[p] < S4F_PATCH_FIXED_DISPLACEMENT || h.patchType[p] > S4F_PATCH_ZERO_SHEAR)
                throw InvalidError("unsupported patch type");
            f += h.patchSize[p];
        }
        if (f != NF) throw InvalidError("patches do not cover the boundary faces");
    }
    for (int f = 0; f < NF; ++f)
        if (h.owner[f] < 0 || h.owner[f] >= C) throw InvalidError("owner out of range");
    for (int f = 0; f < F; ++f)
        if (h.neighbour[f] <= h.owner[f] || h.neighbour[f] >= C) throw InvalidError("faces must be upper-triangular (owner < neighbour < nCells)");

    const double* Cc = d.C;
    const double* Sf = d.Sf;
    const double* Cf = d.Cf;
    h.w.resize(F); h.nonOrthDelta.resize(F); h.magSf.resize(F);
    for (int f = 0; f < F; ++f) {
        const int o = h.owner[f], n = h.neighbour[f];
        double a[3], b[3], dd[3], uA[3];
        const double ms = std::sqrt(dot3(Sf + 3 * (size_t)f, Sf + 3 * (size_t)f));
        h.magSf[f] = ms;
        for (int q = 0; q < 3; ++q) {
            a[q] = Cf[3 * (size_t)f + q] - Cc[3 * (size_t)o + q];
            b[q] = Cc[3 * (size_t)n + q] - Cf[3 * (size_t)f + q];
            dd[q] = Cc[3 * (size_t)n + q] - Cc[3 * (size_t)o + q];
            uA[q] = Sf[3 * (size_t)f + q] / ms;
        }
        const double SfdOwn = std::fabs(dot3(Sf + 3 * (size_t)f, a)), SfdNei = std::fabs(dot3(Sf + 3 * (size_t)f, b));
        h.w[f] = SfdNei / (SfdOwn + SfdNei);
        h.nonOrthDelta[f] = 1.0 / std::max(dot3(uA, dd), 0.05 * std::sqrt(dot3(dd, dd)));
    }
    h.bCell.resize(B); h.bType.resize(B); h.bN.resize(3 * (size_t)B); h.bKv.resize(3 * (size_t)B);
    h.bLs.resize(3 * (size_t)B); h.bDelta.resize(B); h.bMagSf.resize(B);
    for (int p = 0; p < d.nPatches; ++p)
        for (int q = 0; q < h.patchSize[p]; ++q) h.bType[h.patchStart[p] - F + q] = h.patchType[p];
    // 2-D: the axis along the (first) empty face's normal is not solved
    int emptyDir = -1;
    for (int b = 0; b < B && emptyDir < 0; ++b)
        if (h.bType[b] == S4F_PATCH_EMPTY) {
            const double* S = Sf + 3 * (size_t)(F + b);
            emptyDir = 0;
            for (int q = 1; q < 3; ++q)
                if (std::fabs(S[q]) > std::fabs(S[emptyDir])) emptyDir = q;
        }
    // least-squares matrices
    std::vector<double> dd6((size_t)C * 6, 0.0);
    auto addDD = [&](int c, const double* v) {
        const double md2 = dot3(v, v);
        double* t = &dd6[6 * (size_t)c];
        t[0] += v[0] * v[0] / md2; t[1] += v[0] * v[1] / md2; t[2] += v[0] * v[2] / md2;
        t[3] += v[1] * v[1] / md2; t[4] += v[1] * v[2] / md2; t[5] += v[2] * v[2] / md2;
    };
    for (int f = 0; f < F; ++f) {
        double v[3];
        for (int q = 0; q < 3; ++q) v[q] = Cc[3 * (size_t)h.neighbour[f] + q] - Cc[3 * (size_t)h.owner[f] + q];
        addDD(h.owner[f], v);
        addDD(h.neighbour[f], v);
    }
    std::vector<double> bdl((size_t)B * 3);
    for (int b = 0; b < B; ++b) {
        const int f = F + b, P = h.owner[f];
        h.bCell[b] = P;
        const double ms = std::sqrt(dot3(Sf + 3 * (size_t)f, Sf + 3 * (size_t)f));
        h.bMagSf[b] = ms;
        double n[3], dl[3];
        for (int q = 0; q < 3; ++q) {
            n[q] = Sf[3 * (size_t)f + q] / ms;
            dl[q] = Cf[3 * (size_t)f + q] - Cc[3 * (size_t)P + q];
        }
        const double nd = dot3(n, dl);
        double dn[3];
        for (int q = 0; q < 3; ++q) {
            h.bN[(size_t)q * B + b] = n[q];
            h.bKv[(size_t)q * B + b] = dl[q] - n[q] * nd;  // (I - nn).(Cf - C)
            dn[q] = n[q] * nd;
        }
        h.bDelta[b] = 1.0 / std::sqrt(dot3(dn, dn));
        h.lsFullDelta = d.lsFullDelta != 0;
        const double* ld = d.lsFullDelta ? dl : dn;
        for (int q = 0; q < 3; ++q) bdl[3 * (size_t)b + q] = ld[q];
        if (h.bType[b] == S4F_PATCH_EMPTY) {  // empty faces carry no values
            for (int q = 0; q < 3; ++q) h.bLs[(size_t)q * B + b] = 0.0;
            continue;
        }
        addDD(P, ld);
        const double r = 1.0 / dot3(ld, ld);
        for (int q = 0; q < 3; ++q) h.bLs[(size_t)q * B + b] = ld[q] * r;
    }
    h.minv.resize(6 * (size_t)C);
    h.cC.resize(3 * (size_t)C);
    // 2-D: the LS matrix has a zero row and column along the empty axis:
    // invert it with a unit on that diagonal entry, then take the unit off
    // the inverse, so gradients have no component along the axis
    const int emptyDiag = emptyDir < 0 ? -1 : (emptyDir == 0 ? 0 : (emptyDir == 1 ? 3 : 5));
    for (int c = 0; c < C; ++c) {
        double* s = &dd6[6 * (size_t)c];
        if (emptyDiag >= 0) s[emptyDiag] += 1.0;
        const double xx = s[0], xy = s[1], xz = s[2], yy = s[3], yz = s[4], zz = s[5];
        const double det = xx * yy * zz + xy * yz * xz + xz * xy * yz - xx * yz * yz - xy * xy * zz - xz * yy * xz;
        if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) throw InvalidError("singular least-squares matrix in cell " + std::to_string(c));
        double inv[6] = {(yy * zz - yz * yz) / det, (xz * yz - xy * zz) / det, (xy * yz - xz * yy) / det,
                         (xx * zz - xz * xz) / det, (xy * xz - xx * yz) / det, (xx * yy - xy * xy) / det};
        if (emptyDiag >= 0) inv[emptyDiag] -= 1.0;
        for (int q = 0; q < 6; ++q) h.minv[(size_t)q * C + c] = inv[q];
        for (int q = 0; q < 3; ++q) h.cC[(size_t)q * C + c] = Cc[3 * (size_t)c + q];
    }
    // sliced ELL: row lists in face order, internal first then boundary
    std::vector<int> cnt(C, 0);
    for (int f = 0; f < NF; ++f)
        if (f < F || h.bType[f - F] != S4F_PATCH_EMPTY) cnt[h.owner[f]]++;
    for (int f = 0; f < F; ++f) cnt[h.neighbour[f]]++;
    h.nSlices = (C + kSlice - 1) / kSlice;
    h.sliceStart.resize((size_t)h.nSlices + 1);
    size_t E = 0;
    for (int s = 0; s < h.nSlices; ++s) {
        h.sliceStart[s] = (int)E;
        int wmax = 0;
        for (int l = 0; l < kSlice; ++l) {
            const int c = s * kSlice + l;
            if (c < C) wmax = std::max(wmax, cnt[c]);
        }
        E += (size_t)wmax * kSlice;
        if (E > (size_t)INT32_MAX) throw InvalidError("mesh too large for int32 ELL offsets");
    }
    h.sliceStart[h.nSlices] = (int)E;
    h.sliceWidth.resize(h.nSlices);
    for (int s = 0; s < h.nSlices; ++s) h.sliceWidth[s] = (h.sliceStart[s + 1] - h.sliceStart[s]) / kSlice;
    h.col.assign(E, kPad);
    h.eS.assign(3 * E, 0.0);
    h.eW.assign(E, 0.0);
    h.entOwn.resize(F); h.entNei.resize(F); h.entB.resize(B);
    std::vector<int> fill(C, 0);
    auto slot = [&](int c) {
        const int s = c / kSlice, l = c % kSlice;
        return h.sliceStart[s] + fill[c]++ * kSlice + l;
    };
    // internal faces in face order: each cell receives its entries sorted by face index
    for (int f = 0; f < F; ++f) {
        const int o = h.owner[f], n = h.neighbour[f];
        // neighbour side first only if its face index order requires: both lists are
        // filled in face order, which is the accumulation order of each cell
        const int eo = slot(o);
        h.col[eo] = n;
        h.entOwn[f] = eo;
        h.eW[eo] = h.w[f];
        for (int q = 0; q < 3; ++q) h.eS[(size_t)q * E + eo] = Sf[3 * (size_t)f + q];
        const int en = slot(n);
        h.col[en] = o;
        h.entNei[f] = en;
        h.eW[en] = 1.0 - h.w[f];
        for (int q = 0; q < 3; ++q) h.eS[(size_t)q * E + en] = -Sf[3 * (size_t)f + q];
    }
    for (int b = 0; b < B; ++b) {
        const int f = F + b, P = h.owner[f];
        if (h.bType[b] == S4F_PATCH_EMPTY) {  // no row slot: no kernel touches an empty face
            h.entB[b] = -1;
            continue;
        }
        const int e = slot(P);
        h.col[e] = -1 - b;
        h.entB[b] = e;
        for (int q = 0; q < 3; ++q) h.eS[(size_t)q * E + e] = Sf[3 * (size_t)f + q];
    }
    h.bVal.assign(3 * (size_t)B, 0.0);
    h.bPres.assign(B, 0.0);
    h.instCell0 = {0};
    h.instNCell = {C};
    h.instB0 = {0};
    h.instNB = {B};
    h.instMask = {emptyDir < 0 ? 7 : (7 & ~(1 << emptyDir))};
    h.instTileCap = {gMaxBlocks};
    h.instTileMin = {kMinTileSlices};
}

// Rhie-Chow coefficient gamma_f |S_f| deltaCoeffs of every ELL entry
// (momentumStabilisation.C:112-150): gamma_f = rcScale * interpolate(impK),
// on material-interface faces interfaceScale * mean(impK)
std::vector<double> stabilisationCoeffs(const HostMesh& h, double rcScale, double ifScale) {
    std::vector<double> eC(h.col.size(), 0.0);
    auto ik = [&](int c) { return h.uniformMaterial ? h.impK[0] : h.impK[c]; };
    for (int f = 0; f < h.F; ++f) {
        const int o = h.owner[f], nn = h.neighbour[f];
        const double wf = h.w[f];
        const double go = rcScale * ik(o), gn = rcScale * ik(nn);
        double gam = wf * (go - gn) + gn;
        if (std::fabs(ik(o) - ik(nn)) > kSmall) gam = ifScale * 0.5 * (ik(o) + ik(nn));
        const double c = gam * h.magSf[f] * h.nonOrthDelta[f];
        eC[h.entOwn[f]] = c;
        eC[h.entNei[f]] = c;
    }
    return eC;
}

constexpr double kRhieChowScaleDefault = 0.1, kInterfaceScaleDefault = 0.01;

void computeMaterial(HostMesh& h, const double* E, const double* nu, bool perCell) {
    const int C = h.C, F = h.F, B = h.B;
    const int n = perCell ? C : 1;
    // linearElastic.C:101-133: nu = 0.5 gives lambda = K = GREAT, a case the
    // reference solves with its pressure equation (solvePressureEqn); the
    // device path does not build that formulation
    for (int i = 0; i < n; ++i) {
        const double e = E[i], v = nu[i];
        if (!(e > 0.0) || !std::isfinite(e) || !std::isfinite(v) || !(v > -1.0 && v <= 0.5))
            throw InvalidError("linearElastic needs E > 0 and -1 < nu <= 0.5");
        if (v == 0.5)
            throw InvalidError("nu = 0.5 (incompressible) needs linearElastic's pressure equation (solvePressureEqn), "
                               "which the device path does not build");
    }
    h.uniformMaterial = !perCell;
    h.matE.assign(E, E + n);
    h.matNu.assign(nu, nu + n);
    h.mu.resize(n); h.K.resize(n); h.impK.resize(n);
    for (int i = 0; i < n; ++i) {
        const double e = E[i], v = nu[i];
        const double mu = e / (2.0 * (1.0 + v));
        // linearElastic.C:109-128: plane stress changes lambda and K, not mu
        const double lam = h.planeStress ? v * e / ((1.0 + v) * (1.0 - v)) : v * e / ((1.0 + v) * (1.0 - 2.0 * v));
        h.mu[i] = mu;
        h.K[i] = h.planeStress ? e / (3.0 * (1.0 - v)) : e / (3.0 * (1.0 - 2.0 * v));
        h.impK[i] = 2.0 * mu + lam;
        if (!std::isfinite(h.K[i]) || !std::isfinite(h.impK[i])) throw InvalidError("non-finite elastic moduli");
    }
    auto ik = [&](int c) { return perCell ? h.impK[c] : h.impK[0]; };
    const size_t E_ = h.col.size();
    h.eA.assign(E_, 0.0);
    h.diag.assign(C, 0.0);
    for (int f = 0; f < F; ++f) {
        const int o = h.owner[f], nn = h.neighbour[f];
        const double wf = h.w[f];
        const double impKf = wf * (ik(o) - ik(nn)) + ik(nn);
        const double a = h.nonOrthDelta[f] * (impKf * h.magSf[f]);
        h.eA[h.entOwn[f]] = a;
        h.eA[h.entNei[f]] = a;
        h.diag[o] += a;
        h.diag[nn] += a;
    }
    // the defaults; a solve with other scales rebuilds the device copy
    h.eC = stabilisationCoeffs(h, kRhieChowScaleDefault, kInterfaceScaleDefault);
    for (int b = 0; b < B; ++b)
        // symmetry and zero shear: the full fixed-value coefficient (their
        // implicit parts are per component; one matrix serves the three
        // components, and the residual carries the exact condition)
        if (h.bType[b] == S4F_PATCH_FIXED_DISPLACEMENT || h.bType[b] == S4F_PATCH_SYMMETRY ||
            h.bType[b] == S4F_PATCH_ZERO_SHEAR) {
            const int P = h.bCell[b];
            h.diag[P] += ik(P) * h.bMagSf[b] * h.bDelta[b];
        }
}

void uploadMaterial(s4f_mesh* M) {
    HostMesh& h = M->h;
    hipStream_t s = M->ctx->stream;
    h.diagBase = h.diag;  // the steady matrix
    M->inerKey = 0.0;
    M->eA.upload(h.eA, s);
    M->eC.upload(h.eC, s);
    M->diag.upload(h.diag, s);
    std::vector<double> inv(h.C);
    std::vector<float> a32(h.eA.size()), d32(h.C), i32(h.C);
    for (int c = 0; c < h.C; ++c) {
        inv[c] = 1.0 / h.diag[c];
        d32[c] = (float)h.diag[c];
        i32[c] = 1.0f / d32[c];
    }
    for (size_t e = 0; e < h.eA.size(); ++e) a32[e] = (float)h.eA[e];
    M->invDiag.upload(inv, s);
    M->eA32.upload(a32, s);
    M->diag32.upload(d32, s);
    M->invDiag32.upload(i32, s);
    if (!h.uniformMaterial) {
        M->mu.upload(h.mu, s);
        M->K.upload(h.K, s);
        M->impK.upload(h.impK, s);
    }
    HIP_CHECK(hipStreamSynchronize(s));
}

// rhieChowScale / interfaceScale other than the defaults rescale eC
void applyStabilisationScale(s4f_mesh* M, double rcScale, double ifScale) {
    if (rcScale == M->rcScale && ifScale == M->ifScale) return;
    HostMesh& h = M->h;
    // a packed batch is packed with its solve's scales (s4f_batch_solve repacks when they change)
    if (h.instCell0.size() != 1) throw InvalidError("packed batch with other stabilisation scales than its solve");
    M->eC.upload(stabilisationCoeffs(h, rcScale, ifScale), M->ctx->stream);
    M->rcScale = rcScale;
    M->ifScale = ifScale;
}

void uploadTiles(s4f_mesh* M);

void uploadGeometry(s4f_mesh* M) {
    HostMesh& h = M->h;
    hipStream_t s = M->ctx->stream;
    M->sliceStart.upload(h.sliceStart, s);
    M->sliceWidth.upload(h.sliceWidth, s);
    M->col.upload(h.col, s);
    M->eS.upload(h.eS, s);
    M->eW.upload(h.eW, s);
    M->cC.upload(h.cC, s);
    M->minv.upload(h.minv, s);
    M->bCell.upload(h.bCell, s);
    M->bType.upload(h.bType, s);
    M->bN.upload(h.bN, s);
    M->bKv.upload(h.bKv, s);
    M->bDelta.upload(h.bDelta, s);
    M->bLs.upload(h.bLs, s);
    M->bVal.upload(h.bVal, s);
    M->bPres.upload(h.bPres, s);
    uploadTiles(M);
}

// Workgroups: each instance is cut into workgroup tiles by a rule that
// depends on that instance alone (tileSlices: at least one slice per wave,
// at most its tile cap), so its reductions -- and therefore its result -- are
// bitwise the same whether it is solved alone or packed with others.
void uploadTiles(s4f_mesh* M) {
    HostMesh& h = M->h;
    hipStream_t s = M->ctx->stream;
    const int nInst = (int)h.instCell0.size();
    std::vector<int> inst, s0, s1, blk0(nInst), nblk(nInst);
    int spb = 0;
    for (int i = 0; i < nInst; ++i) {
        const int a = h.instCell0[i] / kSlice;
        const int nS = (h.instNCell[i] + kSlice - 1) / kSlice;
        const int e = a + nS;
        const int spbI = tileSlices(h, i, nS);
        spb = std::max(spb, spbI);
        blk0[i] = (int)inst.size();
        for (int x = a; x < e; x += spbI) {
            inst.push_back(i);
            s0.push_back(x);
            s1.push_back(std::min(x + spbI, e));
        }
        if ((int)inst.size() == blk0[i]) {  // empty instance: one idle workgroup keeps the reductions defined
            inst.push_back(i);
            s0.push_back(a);
            s1.push_back(a);
        }
        nblk[i] = (int)inst.size() - blk0[i];
    }
    M->slicesPerBlock = spb;
    M->nBlocks = (int)inst.size();
    M->tInst = inst; M->tS0 = s0; M->tS1 = s1; M->tBlk0 = blk0; M->tNBlk = nblk;
    M->bmInst.upload(inst, s);
    M->bmS0.upload(s0, s);
    M->bmS1.upload(s1, s);
    M->iBlk0.upload(blk0, s);
    M->iNBlk.upload(nblk, s);
    M->iCell0.upload(h.instCell0, s);
    M->iNCell.upload(h.instNCell, s);
    M->iB0.upload(h.instB0, s);
    M->iNB.upload(h.instNB, s);
    M->maxInstNB = h.instNB.empty() ? 0 : *std::max_element(h.instNB.begin(), h.instNB.end());
    HIP_CHECK(hipStreamSynchronize(s));
}

void aosToSoa(const double* a, int n, int k, std::vector<double>& out) {
    out.resize((size_t)n * k);
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < k; ++q) out[(size_t)q * n + i] = a[(size_t)i * k + q];
}
void soaToAos(const std::vector<double>& s, int n, int k, double* out) {
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < k; ++q) out[(size_t)i * k + q] = s[(size_t)q * n + i];
}
std::vector<double> download(const DArr<double>& d, size_t n, hipStream_t s) {
    std::vector<double> h(n);
    if (n) HIP_CHECK(hipMemcpyAsync(h.data(), d.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return h;
}

Ctl makeCtl(const s4f_solve_opts& o, int histCap) {
    Ctl c;
    std::memset(&c, 0, sizeof(c));
    c.pcgTol = o.pcgTolerance;
    c.pcgRelTol = o.pcgRelTol;
    c.solTol = o.solutionTolerance;
    c.altTol = o.alternativeTolerance;
    c.pcgMaxIter = o.pcgMaxIter;
    c.nCorr = o.nCorrectors;
    c.histCap = histCap;
    c.relaxD = (o.relaxD > 0.0) ? o.relaxD : 1.0;
    c.relaxAitken = o.relaxationMethod == S4F_RELAX_AITKEN ? 1 : 0;
    c.solveMask = 7;
    return c;
}

void singleMeshOnly(const s4f_solve_opts& o, const char* entry) {
    if (o.preconditioner == S4F_PRECOND_GAMG || o.preconditioner == S4F_PRECOND_GAMG_SMOOTHED)
        throw InvalidError(std::string(entry) + ": the multigrid preconditioner is not available here "
                           "(s4f_solve, s4f_batch_solve and the server have it)");
}

void checkOpts(const s4f_solve_opts& o) {
    if (o.nCorrectors < 1) throw InvalidError("nCorrectors must be >= 1");
    if (o.pcgMaxIter < 0) throw InvalidError("maxIter must be >= 0");
    if (o.precision != S4F_PRECISION_FP64 && o.precision != S4F_PRECISION_MIXED) throw InvalidError("unknown precision");
    if (o.preconditioner != S4F_PRECOND_DIAGONAL && o.preconditioner != S4F_PRECOND_GAMG &&
        o.preconditioner != S4F_PRECOND_GAMG_SMOOTHED)
        throw InvalidError("unknown preconditioner");
    if (!(o.mgCoarseScale >= 0.0 && o.mgCoarseScale <= 2.0)) throw InvalidError("mgCoarseScale must be in [0, 2]");
    if (!(o.relaxD >= 0.0 && o.relaxD <= 1.0)) throw InvalidError("relaxD must be in [0, 1]");
    if (o.eagerLaunch != 0 && o.eagerLaunch != 1) throw InvalidError("eagerLaunch must be 0 or 1");
    if (o.relaxationMethod != S4F_RELAX_FIXED && o.relaxationMethod != S4F_RELAX_AITKEN)
        throw InvalidError("relaxationMethod must be fixed (0) or Aitken (1)");
    if (o.d2dt2 != S4F_D2DT2_STEADY && o.d2dt2 != S4F_D2DT2_EULER && o.d2dt2 != S4F_D2DT2_BACKWARD)
        throw InvalidError("d2dt2 must be steady (0), Euler (1) or backward (2)");
    if (o.d2dt2 == S4F_D2DT2_BACKWARD && o.deltaT0 > 0.0 && o.deltaT0 != o.deltaT)
        throw InvalidError("backward d2dt2 needs a constant time step (backwardD2dt2Scheme.C:238-243)");
    if (!(o.pcgTolerance >= 0) || !(o.pcgRelTol >= 0) || !(o.solutionTolerance >= 0) || !(o.alternativeTolerance >= 0))
        throw InvalidError("tolerances must be >= 0");
}

constexpr int kHistCap = 100000;

template <typename T>
struct PcgVecs {
    T *r, *p, *q, *x, *p2;  // p and p2 alternate as old/new search direction
    T* z = nullptr;         // preconditioned residual (V-cycle output)
};

template <typename T>
PcgVecs<T> pcgVecs(SolveState& st, int C);
template <>
PcgVecs<double> pcgVecs<double>(SolveState& st, int C) {
    st.r64.alloc(3 * (size_t)C); st.p64.alloc(3 * (size_t)C); st.q64.alloc(3 * (size_t)C); st.x64.alloc(3 * (size_t)C);
    st.pb64.alloc(3 * (size_t)C);
    st.z64.alloc(3 * (size_t)C);
    return {st.r64.p, st.p64.p, st.q64.p, st.x64.p, st.pb64.p, st.z64.p};
}
template <>
PcgVecs<float> pcgVecs<float>(SolveState& st, int C) {
    st.r32.alloc(3 * (size_t)C); st.p32.alloc(3 * (size_t)C); st.q32.alloc(3 * (size_t)C); st.x32.alloc(3 * (size_t)C);
    st.pb32.alloc(3 * (size_t)C);
    st.z32.alloc(3 * (size_t)C);
    return {st.r32.p, st.p32.p, st.q32.p, st.x32.p, st.pb32.p, st.z32.p};
}

void allocState(s4f_mesh* M) {
    SolveState& st = M->st;
    const int C = M->h.C, B = std::max(M->h.B, 1);
    const int nInst = (int)M->h.instCell0.size();
    st.D0.alloc(3 * (size_t)C); st.D1.alloc(3 * (size_t)C); st.Dold.alloc(3 * (size_t)C);
    st.G.alloc(9 * (size_t)C);
    st.Db.alloc(3 * (size_t)B); st.g.alloc(3 * (size_t)B); st.snb.alloc(3 * (size_t)B);
    st.Gb.alloc(9 * (size_t)B); st.sigb.alloc(6 * (size_t)B);
    st.ctl.alloc(nInst);
    st.part.alloc((size_t)M->nBlocks * kMaxPartials);
    st.histCap = nInst == 1 ? kHistCap : 4096;
    st.hist.alloc(3 * (size_t)st.histCap * nInst);
}

// `n` inner PCG iterations: fused direction update + SpMV, alpha, solution /
// residual update, convergence + beta.  v.p / v.p2 swap roles every iteration.
template <typename T>
void launchPcgChunk(const DevMesh& dm, const BlockMap& bm, int nInst, PcgVecs<T>& v, Ctl* ctl, double* part, int n,
                    hipStream_t s) {
    for (int i = 0; i < n; ++i) {
        hipLaunchKernelGGL((k_spmvp<T, false>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const T*)v.r, (const T*)v.p,
                           v.p2, v.q, ctl, part, 1, (const T*)nullptr);
        hipLaunchKernelGGL(k_fin_spmv, dim3(nInst), dim3(kFinBlock), 0, s, bm, ctl, (const double*)part);
        hipLaunchKernelGGL((k_update<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, v.x, (const T*)v.p2, v.r,
                           (const T*)v.q, (const Ctl*)ctl, part);
        hipLaunchKernelGGL(k_fin_update, dim3(nInst), dim3(kFinBlock), 0, s, bm, ctl, (const double*)part);
        std::swap(v.p, v.p2);
    }
    HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// aggregation multigrid: host build
// ---------------------------------------------------------------------------
// f(begin, end) over [0, n) split into contiguous ranges on up to 16 host
// threads (one range below `grain`); every output element is written by one
// range, so the result does not depend on the thread count
template <typename F>
void parallelFor(int n, int grain, F&& f) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min({hw, 16, n / std::max(1, grain)}));
    if (nt <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    std::exception_ptr err;
    std::mutex em;
    auto run = [&](int b, int e) {
        try {
            f(b, e);
        } catch (...) {
            std::lock_guard<std::mutex> g(em);
            if (!err) err = std::current_exception();
        }
    };
    for (int t = 1; t < nt; ++t) th.emplace_back(run, (int)((long long)n * t / nt), (int)((long long)n * (t + 1) / nt));
    run(0, (int)((long long)n / nt));
    for (std::thread& x : th) x.join();
    if (err) std::rethrow_exception(err);
}

template <typename F>
inline void forRow(const std::vector<int>& ss, const std::vector<int>& sw, const std::vector<int>& col, int i, F&& f) {
    const int sI = i / kSlice, l = i % kSlice;
    for (int k = 0; k < sw[sI]; ++k) {
        const int e = ss[sI] + k * kSlice + l;
        if (col[e] >= 0) f(e, col[e]);
    }
}

// Greedy aggregation on strong connections (a_ij >= theta max_k a_ik) of the
// rows [r0, r1) (one instance; no connection leaves it): seed aggregates from
// rows whose strong neighbours are all free, attach the rest to their
// strongest aggregated neighbour, leftovers become singletons.  Aggregates
// are numbered from nc0 in seed order; returns the count.
int aggregate(const AmgHostLevel& L, std::vector<int>& agg, int r0, int r1, int nc0) {
    const double theta = 0.25;
    std::vector<double> amax(r1 - r0, 0.0);
    parallelFor(r1 - r0, 65536, [&](int a0, int a1) {
        for (int i = r0 + a0; i < r0 + a1; ++i)
            forRow(L.sliceStart, L.sliceWidth, L.col, i,
                   [&](int e, int) { amax[i - r0] = std::max(amax[i - r0], L.a[e]); });
    });
    int nc = nc0;
    for (int i = r0; i < r1; ++i) {
        if (agg[i] >= 0) continue;
        bool free = true;
        forRow(L.sliceStart, L.sliceWidth, L.col, i, [&](int e, int j) {
            if (L.a[e] >= theta * amax[i - r0] && agg[j] >= 0) free = false;
        });
        if (!free) continue;
        agg[i] = nc;
        forRow(L.sliceStart, L.sliceWidth, L.col, i, [&](int e, int j) {
            if (L.a[e] >= theta * amax[i - r0]) agg[j] = nc;
        });
        ++nc;
    }
    std::vector<int> joined(r1 - r0, -1);
    for (int i = r0; i < r1; ++i) {
        if (agg[i] >= 0) continue;
        double best = -1.0;
        int to = -1;
        forRow(L.sliceStart, L.sliceWidth, L.col, i, [&](int e, int j) {
            if (agg[j] >= 0 && L.a[e] >= theta * amax[i - r0] && L.a[e] > best) { best = L.a[e]; to = agg[j]; }
        });
        joined[i - r0] = to;
    }
    for (int i = r0; i < r1; ++i)
        if (agg[i] < 0 && joined[i - r0] >= 0) agg[i] = joined[i - r0];
    for (int i = r0; i < r1; ++i)
        if (agg[i] < 0) agg[i] = nc++;
    return nc - nc0;
}

// Galerkin product with piecewise-constant prolongation: coarse diagonal
// = sum of member diagonals minus intra-aggregate couplings, off-diagonal
// magnitudes summed per coarse pair; stored as sliced ELL sorted by column.
void galerkin(const AmgHostLevel& F, const std::vector<int>& agg, int nc, AmgHostLevel& Cl) {
    Cl = AmgHostLevel();
    Cl.n = nc;
    Cl.memOff.assign(nc + 1, 0);
    for (int i = 0; i < F.n; ++i)
        if (agg[i] >= 0) Cl.memOff[agg[i] + 1]++;
    for (int I = 0; I < nc; ++I) Cl.memOff[I + 1] += Cl.memOff[I];
    Cl.mem.resize(Cl.memOff[nc]);
    {
        std::vector<int> fill(Cl.memOff.begin(), Cl.memOff.end() - 1);
        for (int i = 0; i < F.n; ++i)
            if (agg[i] >= 0) Cl.mem[fill[agg[i]]++] = i;
    }
    Cl.diag.assign(nc, 0.0);
    // coarse rows in contiguous ranges per thread: each range's rows (sorted
    // by column, duplicates summed) go to that range's own entry list
    struct Part {
        int I0 = 0;
        std::vector<int> off, col;
        std::vector<double> a;
    };
    const int nParts = std::max(1, std::min(64, nc / 4096));
    std::vector<Part> parts(nParts);
    std::vector<int> width(nc, 0);
    parallelFor(nParts, 1, [&](int p0, int p1) {
        std::vector<std::pair<int, double>> r;
        for (int p = p0; p < p1; ++p) {
            Part& P = parts[p];
            P.I0 = (int)((long long)nc * p / nParts);
            const int I1 = (int)((long long)nc * (p + 1) / nParts);
            P.off.assign(I1 - P.I0 + 1, 0);
            for (int I = P.I0; I < I1; ++I) {
                r.clear();
                for (int k = Cl.memOff[I]; k < Cl.memOff[I + 1]; ++k) {
                    const int i = Cl.mem[k];
                    Cl.diag[I] += F.diag[i];
                    forRow(F.sliceStart, F.sliceWidth, F.col, i, [&](int e, int j) {
                        const int J = agg[j];
                        if (J == I) Cl.diag[I] -= F.a[e];
                        else r.push_back({J, F.a[e]});
                    });
                }
                std::sort(r.begin(), r.end(), [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
                    return x.first < y.first;
                });
                int w = 0;
                for (size_t k = 0; k < r.size(); ++k) {
                    if (w > 0 && P.col.back() == r[k].first) {
                        P.a.back() += r[k].second;
                    } else {
                        P.col.push_back(r[k].first);
                        P.a.push_back(r[k].second);
                        ++w;
                    }
                }
                width[I] = w;
                P.off[I - P.I0 + 1] = P.off[I - P.I0] + w;
            }
        }
    });
    Cl.nSlices = (nc + kSlice - 1) / kSlice;
    Cl.sliceStart.resize((size_t)Cl.nSlices + 1);
    Cl.sliceWidth.resize(Cl.nSlices);
    size_t E = 0;
    for (int sI = 0; sI < Cl.nSlices; ++sI) {
        Cl.sliceStart[sI] = (int)E;
        int wmax = 0;
        for (int l = 0; l < kSlice; ++l) {
            const int I = sI * kSlice + l;
            if (I < nc) wmax = std::max(wmax, width[I]);
        }
        Cl.sliceWidth[sI] = wmax;
        E += (size_t)wmax * kSlice;
    }
    Cl.sliceStart[Cl.nSlices] = (int)E;
    Cl.col.assign(E, kPad);
    Cl.a.assign(E, 0.0);
    parallelFor(nParts, 1, [&](int p0, int p1) {
        for (int p = p0; p < p1; ++p) {
            const Part& P = parts[p];
            const int n = (int)P.off.size() - 1;
            for (int t = 0; t < n; ++t) {
                const int I = P.I0 + t, sI = I / kSlice, l = I % kSlice;
                for (int k = P.off[t]; k < P.off[t + 1]; ++k) {
                    const size_t e = (size_t)Cl.sliceStart[sI] + (size_t)(k - P.off[t]) * kSlice + l;
                    Cl.col[e] = P.col[k];
                    Cl.a[e] = P.a[k];
                }
            }
        }
    });
}

// Explicit inverse of an SPD matrix (row-major n x n) by Cholesky:
// A = L L^T, A^-1 = L^-T L^-1.  Returns false on a non-positive pivot.
bool spdInverse(const AmgHostLevel& H, int r0, int n, std::vector<double>& inv) {
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        A[(size_t)i * n + i] = H.diag[r0 + i];
        forRow(H.sliceStart, H.sliceWidth, H.col, r0 + i,
               [&](int e, int j) { A[(size_t)i * n + (j - r0)] -= H.a[e]; });
    }
    for (int j = 0; j < n; ++j) {  // in-place Cholesky, lower triangle
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double v = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) v -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = v / d;
        }
    }
    // W = L^-1 row by row: W_i = (e_i - sum_{k<i} L_ik W_k) / L_ii (contiguous row updates)
    std::vector<double> W((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double* wi = &W[(size_t)i * n];
        wi[i] = 1.0;
        for (int k = 0; k < i; ++k) {
            const double l = A[(size_t)i * n + k];
            const double* wk = &W[(size_t)k * n];
            for (int j = 0; j <= k; ++j) wi[j] -= l * wk[j];
        }
        const double rd = 1.0 / A[(size_t)i * n + i];
        for (int j = 0; j <= i; ++j) wi[j] *= rd;
    }
    // A^-1 = W^T W, accumulated by rows of W (lower triangle, then mirrored)
    inv.assign((size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) {
        const double* wk = &W[(size_t)k * n];
        for (int i = 0; i <= k; ++i) {
            const double a = wk[i];
            if (a == 0.0) continue;
            double* row = &inv[(size_t)i * n];
            for (int j = 0; j <= i; ++j) row[j] += a * wk[j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) inv[(size_t)j * n + i] = inv[(size_t)i * n + j];
    return true;
}

// q = offdiag(A) P of a level: per row, the magnitudes of its neighbour
// entries summed per coarse row of the neighbour, in ascending coarse-row
// order, in the row's own sliced-ELL slots (the rest empty: column -1);
// qw = per slice, the largest filled slot count
void groupByAggregate(const AmgHostLevel& F, std::vector<int>& qc, std::vector<double>& q, std::vector<int>& qw) {
    qc.assign(F.col.size(), -1);
    q.assign(F.col.size(), 0.0);
    qw.assign(F.nSlices, 0);
    parallelFor(F.nSlices, 256, [&](int sA, int sB) {
    std::vector<std::pair<int, double>> row;
    for (int i = sA * kSlice; i < std::min(F.n, sB * kSlice); ++i) {
        row.clear();
        forRow(F.sliceStart, F.sliceWidth, F.col, i, [&](int e, int j) { row.push_back({F.agg[j], F.a[e]}); });
        // stable insertion sort by aggregate (a row has few entries; no allocation)
        for (size_t u = 1; u < row.size(); ++u) {
            const std::pair<int, double> t = row[u];
            size_t v = u;
            while (v > 0 && row[v - 1].first > t.first) {
                row[v] = row[v - 1];
                --v;
            }
            row[v] = t;
        }
        const int sI = i / kSlice, l = i % kSlice;
        int k = -1, last = -1;
        for (const std::pair<int, double>& t : row) {
            if (t.first != last) {
                ++k;
                last = t.first;
                qc[(size_t)F.sliceStart[sI] + (size_t)k * kSlice + l] = t.first;
            }
            q[(size_t)F.sliceStart[sI] + (size_t)k * kSlice + l] += t.second;
        }
        qw[sI] = std::max(qw[sI], k + 1);
    }
    });
}

// Row-wise sparse vectors of the smoothed first coarsening (CSR by row).
struct SparseRows {
    std::vector<long long> off;
    std::vector<int> col;
    std::vector<double> val;
};

// entries of one row sorted by column (stable: equal columns keep their
// order) and equal columns summed in that order, so the result depends only
// on the row
inline void sortMergeRow(std::vector<std::pair<int, double>>& r) {
    for (size_t u = 1; u < r.size(); ++u) {
        const std::pair<int, double> t = r[u];
        size_t v = u;
        while (v > 0 && r[v - 1].first > t.first) {
            r[v] = r[v - 1];
            --v;
        }
        r[v] = t;
    }
    size_t w = 0;
    for (size_t u = 0; u < r.size(); ++u) {
        if (w > 0 && r[w - 1].first == r[u].first) r[w - 1].second += r[u].second;
        else r[w++] = r[u];
    }
    r.resize(w);
}

// rows [0, n) of a sparse operator from a row builder, in two threaded passes
// (sizes, then entries)
template <typename F>
SparseRows buildRows(int n, F&& row) {
    SparseRows S;
    S.off.assign((size_t)n + 1, 0);
    parallelFor(n, 4096, [&](int a, int b) {
        std::vector<std::pair<int, double>> r;
        for (int i = a; i < b; ++i) {
            row(i, r);
            S.off[(size_t)i + 1] = (long long)r.size();
        }
    });
    for (int i = 0; i < n; ++i) S.off[(size_t)i + 1] += S.off[i];
    S.col.resize((size_t)S.off[n]);
    S.val.resize((size_t)S.off[n]);
    parallelFor(n, 4096, [&](int a, int b) {
        std::vector<std::pair<int, double>> r;
        for (int i = a; i < b; ++i) {
            row(i, r);
            for (size_t k = 0; k < r.size(); ++k) {
                S.col[(size_t)S.off[i] + k] = r[k].first;
                S.val[(size_t)S.off[i] + k] = r[k].second;
            }
        }
    });
    return S;
}

// Smoothed prolongation of the first coarsening (level 0 -> 1), the
// smoothed-aggregation construction: P = (I - wp D^-1 A) P_t with P_t the
// aggregation and wp = 2/3 (4/3 over 2, the bound of the Jacobi-scaled
// spectrum of a diagonally dominant M-matrix), so row i of P is
// (1 - wp) e_agg(i) + (wp / D_i) q_i with q = offdiag(A) P_t.  Builds the
// coarse operator A_c = P^T (A P) in Cl (sliced ELL, off-diagonal entries as
// magnitudes, the restriction P^T as weighted member lists in fine-row order)
// and level 0's up-sweep operator G = (D / w) P - A P in F: with it the
// prolongation fused into the post-smoothing sweep is
//   z = w D^-1 (b + r0) + alpha w D^-1 G e.
// Off-diagonal entries of A_c below filter * sqrt(|a_II a_JJ|) are dropped
// and added to the row's diagonal (row sums kept; the criterion is symmetric,
// so is the filtered operator): the smoothed stencil's many small far entries
// cost coarse-level bandwidth and barely help the cycle.
// F.agg must hold the aggregation.  Every sum runs in an order fixed by the
// rows alone, so an instance's hierarchy does not depend on its batch.
void galerkinSmoothed(AmgHostLevel& F, int nc, double omega, AmgHostLevel& Cl, double filter) {
    const double wp = 2.0 / 3.0;
    const std::vector<int>& agg = F.agg;
    const int N = F.n;
    std::vector<int> qc, qw;
    std::vector<double> q;
    groupByAggregate(F, qc, q, qw);
    const SparseRows P = buildRows(N, [&](int i, std::vector<std::pair<int, double>>& r) {
        r.clear();
        if (agg[i] < 0) return;
        const double d = F.diag[i];
        r.push_back({agg[i], 1.0 - wp});
        const int sI = i / kSlice, l = i % kSlice;
        for (int k = 0; k < F.sliceWidth[sI]; ++k) {
            const size_t e = (size_t)F.sliceStart[sI] + (size_t)k * kSlice + l;
            if (qc[e] >= 0) r.push_back({qc[e], wp * q[e] / d});
        }
        sortMergeRow(r);
    });
    // A P: row i = D_i P_i - sum over neighbours j of a_ij P_j
    const SparseRows AP = buildRows(N, [&](int i, std::vector<std::pair<int, double>>& r) {
        r.clear();
        if (agg[i] < 0) return;
        for (long long k = P.off[i]; k < P.off[i + 1]; ++k) r.push_back({P.col[k], F.diag[i] * P.val[k]});
        forRow(F.sliceStart, F.sliceWidth, F.col, i, [&](int e, int j) {
            for (long long k = P.off[j]; k < P.off[j + 1]; ++k) r.push_back({P.col[k], -F.a[e] * P.val[k]});
        });
        sortMergeRow(r);
    });
    // G = (D / w) P - A P on level 0, sliced ELL of its own; an empty slot
    // repeats the row's first column with a zero weight (every load in bounds
    // and within the row's instance)
    const SparseRows Gr = buildRows(N, [&](int i, std::vector<std::pair<int, double>>& r) {
        r.clear();
        if (agg[i] < 0) return;
        const double s = F.diag[i] / omega;
        for (long long k = P.off[i]; k < P.off[i + 1]; ++k) r.push_back({P.col[k], s * P.val[k]});
        for (long long k = AP.off[i]; k < AP.off[i + 1]; ++k) r.push_back({AP.col[k], -AP.val[k]});
        sortMergeRow(r);
    });
    {
        const int nS = F.nSlices;
        F.gStart.assign((size_t)nS + 1, 0);
        F.gWidth.assign(nS, 0);
        long long E = 0;
        for (int sI = 0; sI < nS; ++sI) {
            int wmax = 0;
            for (int l = 0; l < kSlice; ++l) {
                const int i = sI * kSlice + l;
                if (i < N) wmax = std::max(wmax, (int)(Gr.off[i + 1] - Gr.off[i]));
            }
            if (E > INT32_MAX) throw InvalidError("smoothed multigrid operator exceeds 2^31 entries");
            F.gStart[sI] = (int)E;
            F.gWidth[sI] = wmax;
            E += (long long)wmax * kSlice;
        }
        if (E > INT32_MAX) throw InvalidError("smoothed multigrid operator exceeds 2^31 entries");
        F.gStart[nS] = (int)E;
        F.gCol.assign((size_t)E, 0);
        F.g.assign((size_t)E, 0.0);
        parallelFor(nS, 256, [&](int s0, int s1) {
            for (int sI = s0; sI < s1; ++sI)
                for (int l = 0; l < kSlice; ++l) {
                    const int i = sI * kSlice + l;
                    if (i >= N || Gr.off[i + 1] == Gr.off[i]) continue;
                    const int w = (int)(Gr.off[i + 1] - Gr.off[i]);
                    for (int k = 0; k < F.gWidth[sI]; ++k) {
                        const size_t e = (size_t)F.gStart[sI] + (size_t)k * kSlice + l;
                        F.gCol[e] = Gr.col[(size_t)Gr.off[i] + (k < w ? k : 0)];
                        F.g[e] = k < w ? Gr.val[(size_t)Gr.off[i] + k] : 0.0;
                    }
                }
        });
    }
    // restriction P^T: per coarse row its fine rows (ascending) and weights
    Cl = AmgHostLevel();
    Cl.n = nc;
    Cl.memOff.assign((size_t)nc + 1, 0);
    for (long long k = 0; k < P.off[N]; ++k) Cl.memOff[(size_t)P.col[k] + 1]++;
    for (int I = 0; I < nc; ++I) Cl.memOff[I + 1] += Cl.memOff[I];
    if ((long long)Cl.memOff[nc] != P.off[N]) throw InvalidError("smoothed prolongation: inconsistent transpose");
    Cl.mem.resize((size_t)P.off[N]);
    Cl.memW.resize((size_t)P.off[N]);
    {
        std::vector<int> fill(Cl.memOff.begin(), Cl.memOff.end() - 1);
        for (int i = 0; i < N; ++i)
            for (long long k = P.off[i]; k < P.off[i + 1]; ++k) {
                const int I = P.col[k];
                Cl.mem[fill[I]] = i;
                Cl.memW[fill[I]] = P.val[k];
                ++fill[I];
            }
    }
    {
        const int nS = (nc + kSlice - 1) / kSlice;
        Cl.rStart.assign((size_t)nS + 1, 0);
        Cl.rWidth.assign(nS, 0);
        long long E = 0;
        for (int sI = 0; sI < nS; ++sI) {
            int wmax = 0;
            for (int l = 0; l < kSlice; ++l) {
                const int I = sI * kSlice + l;
                if (I < nc) wmax = std::max(wmax, Cl.memOff[I + 1] - Cl.memOff[I]);
            }
            if (E > INT32_MAX) throw InvalidError("smoothed multigrid restriction exceeds 2^31 entries");
            Cl.rStart[sI] = (int)E;
            Cl.rWidth[sI] = wmax;
            E += (long long)wmax * kSlice;
        }
        if (E > INT32_MAX) throw InvalidError("smoothed multigrid restriction exceeds 2^31 entries");
        Cl.rStart[nS] = (int)E;
        Cl.rMem.assign((size_t)E, 0);
        Cl.rW.assign((size_t)E, 0.0);
        parallelFor(nS, 256, [&](int s0, int s1) {
            for (int sI = s0; sI < s1; ++sI)
                for (int l = 0; l < kSlice; ++l) {
                    const int I = sI * kSlice + l;
                    if (I >= nc) continue;
                    const int k0 = Cl.memOff[I], w = Cl.memOff[I + 1] - k0;
                    for (int k = 0; k < Cl.rWidth[sI]; ++k) {
                        const size_t e = (size_t)Cl.rStart[sI] + (size_t)k * kSlice + l;
                        Cl.rMem[e] = w > 0 ? Cl.mem[k0 + (k < w ? k : 0)] : 0;
                        Cl.rW[e] = k < w ? Cl.memW[k0 + k] : 0.0;
                    }
                }
        });
    }
    // A_c = P^T (A P), coarse rows in contiguous ranges per thread (a dense
    // accumulator per thread; contributions summed in member, then column order)
    Cl.diag.assign(nc, 0.0);
    std::vector<int> width(nc, 0);
    struct Part {
        int I0 = 0;
        std::vector<int> off, col;
        std::vector<double> a;
    };
    const int nParts = std::max(1, std::min(64, nc / 4096));
    std::vector<Part> parts(nParts);
    parallelFor(nParts, 1, [&](int p0, int p1) {
        std::vector<double> acc(nc, 0.0);
        std::vector<char> mark(nc, 0);
        std::vector<int> touched;
        for (int p = p0; p < p1; ++p) {
            Part& Pt = parts[p];
            Pt.I0 = (int)((long long)nc * p / nParts);
            const int I1 = (int)((long long)nc * (p + 1) / nParts);
            Pt.off.assign(I1 - Pt.I0 + 1, 0);
            for (int I = Pt.I0; I < I1; ++I) {
                touched.clear();
                for (int k = Cl.memOff[I]; k < Cl.memOff[I + 1]; ++k) {
                    const int i = Cl.mem[k];
                    const double wgt = Cl.memW[k];
                    for (long long m = AP.off[i]; m < AP.off[i + 1]; ++m) {
                        const int J = AP.col[m];
                        if (!mark[J]) {
                            mark[J] = 1;
                            acc[J] = 0.0;
                            touched.push_back(J);
                        }
                        acc[J] += wgt * AP.val[m];
                    }
                }
                std::sort(touched.begin(), touched.end());
                int w = 0;
                for (int J : touched) {
                    mark[J] = 0;
                    if (J == I) {
                        Cl.diag[I] = acc[J];
                        continue;
                    }
                    Pt.col.push_back(J);
                    Pt.a.push_back(-acc[J]);  // stored as the magnitude of an M-matrix entry (entry = -a)
                    ++w;
                }
                width[I] = w;
                Pt.off[I - Pt.I0 + 1] = Pt.off[I - Pt.I0] + w;
            }
        }
    });
    // filter against the unfiltered diagonals, then lump
    std::vector<double> lump(nc, 0.0);
    std::vector<std::vector<char>> keep(nParts);
    parallelFor(nParts, 1, [&](int p0, int p1) {
        for (int p = p0; p < p1; ++p) {
            const Part& Pt = parts[p];
            keep[p].assign(Pt.col.size(), 1);
            const int n = (int)Pt.off.size() - 1;
            for (int t = 0; t < n; ++t) {
                const int I = Pt.I0 + t;
                int w = 0;
                for (int k = Pt.off[t]; k < Pt.off[t + 1]; ++k) {
                    const int J = Pt.col[k];
                    if (std::fabs(Pt.a[k]) < filter * std::sqrt(std::fabs(Cl.diag[I] * Cl.diag[J]))) {
                        keep[p][k] = 0;
                        lump[I] -= Pt.a[k];  // the entry (-a) joins the diagonal
                    } else {
                        ++w;
                    }
                }
                width[I] = w;
            }
        }
    });
    for (int I = 0; I < nc; ++I) Cl.diag[I] += lump[I];
    Cl.nSlices = (nc + kSlice - 1) / kSlice;
    Cl.sliceStart.resize((size_t)Cl.nSlices + 1);
    Cl.sliceWidth.resize(Cl.nSlices);
    size_t E = 0;
    for (int sI = 0; sI < Cl.nSlices; ++sI) {
        Cl.sliceStart[sI] = (int)E;
        int wmax = 0;
        for (int l = 0; l < kSlice; ++l) {
            const int I = sI * kSlice + l;
            if (I < nc) wmax = std::max(wmax, width[I]);
        }
        Cl.sliceWidth[sI] = wmax;
        E += (size_t)wmax * kSlice;
    }
    if (E > (size_t)INT32_MAX) throw InvalidError("smoothed multigrid coarse operator exceeds 2^31 entries");
    Cl.sliceStart[Cl.nSlices] = (int)E;
    Cl.col.assign(E, kPad);
    Cl.a.assign(E, 0.0);
    parallelFor(nParts, 1, [&](int p0, int p1) {
        for (int p = p0; p < p1; ++p) {
            const Part& Pt = parts[p];
            const int n = (int)Pt.off.size() - 1;
            for (int t = 0; t < n; ++t) {
                const int I = Pt.I0 + t, sI = I / kSlice, l = I % kSlice;
                int slot = 0;
                for (int k = Pt.off[t]; k < Pt.off[t + 1]; ++k) {
                    if (!keep[p][k]) continue;
                    const size_t e = (size_t)Cl.sliceStart[sI] + (size_t)slot * kSlice + l;
                    Cl.col[e] = Pt.col[k];
                    Cl.a[e] = Pt.a[k];
                    ++slot;
                }
            }
        }
    });
}

void buildAmg(s4f_mesh* M, bool smoothed) {
    Amg& A = smoothed ? M->amgSm : M->amg;
    if (A.gen == M->generation && !A.d.empty()) return;
    A.smoothed = smoothed;
    const auto t0 = std::chrono::steady_clock::now();
    const HostMesh& h = M->h;
    const int nInst = (int)h.instCell0.size();
    A.h.clear();
    A.d.clear();
    AmgHostLevel L0;
    L0.n = h.C; L0.nSlices = h.nSlices;
    L0.sliceStart = h.sliceStart; L0.sliceWidth = h.sliceWidth; L0.col = h.col;
    L0.a = h.eA; L0.diag = h.diag;
    L0.inst0 = h.instCell0; L0.instN = h.instNCell;
    L0.rowInst.assign(h.C, -1);
    for (int i = 0; i < nInst; ++i)
        for (int r = h.instCell0[i]; r < h.instCell0[i] + h.instNCell[i]; ++r) L0.rowInst[r] = i;
    A.h.push_back(std::move(L0));
    // coarsen every instance until its rows fit one dense block; an instance
    // already there passes to the next level unchanged (identity aggregation)
    // and its rows skip the smoothing on the way, so its cycle is the one it
    // has alone.  At least two levels: level 0 is smoothed, the coarsest solved.
    for (;;) {
        AmgHostLevel& F = A.h.back();
        const int big = *std::max_element(F.instN.begin(), F.instN.end());
        if (A.h.size() > 1 && big <= kCoarseBlock) break;
        if (A.h.size() > 24) throw InvalidError("multigrid hierarchy deeper than 24 levels");
        std::vector<int> agg(F.n, -1);
        AmgHostLevel Cl;
        F.instPass.assign(nInst, 0);
        int nc = 0;
        for (int i = 0; i < nInst; ++i) {
            const int r0 = F.inst0[i], n = F.instN[i];
            Cl.inst0.push_back(nc);
            int k;
            if (n <= kCoarseBlock) {
                for (int r = 0; r < n; ++r) agg[r0 + r] = nc + r;
                k = n;
                F.instPass[i] = 1;
            } else {
                k = aggregate(F, agg, r0, r0 + n, nc);
                if (k >= n) throw InvalidError("multigrid coarsening stalled");
            }
            Cl.instN.push_back(k);
            nc += k;
        }
        AmgHostLevel G;
        if (smoothed && A.h.size() == 1) {
            F.agg = agg;
            galerkinSmoothed(F, nc, A.omega, G, kMgSmoothedFilter);
        } else {
            galerkin(F, agg, nc, G);
        }
        G.inst0 = std::move(Cl.inst0);
        G.instN = std::move(Cl.instN);
        G.rowInst.assign(nc, -1);
        for (int i = 0; i < nInst; ++i)
            for (int r = G.inst0[i]; r < G.inst0[i] + G.instN[i]; ++r) G.rowInst[r] = i;
        A.h.back().agg = std::move(agg);
        A.h.push_back(std::move(G));
    }
    // coarsest: one dense inverse per instance block
    std::vector<double> inv;
    std::vector<int> invOff(nInst, 0);
    if (A.h.size() > 1) {
        const AmgHostLevel& H = A.h.back();
        std::vector<std::vector<double>> blks(nInst);
        std::vector<char> spd(nInst, 1);
        parallelFor(nInst, 1, [&](int i0, int i1) {
            for (int i = i0; i < i1; ++i)
                if (H.instN[i] > 0) spd[i] = spdInverse(H, H.inst0[i], H.instN[i], blks[i]) ? 1 : 0;
        });
        for (int i = 0; i < nInst; ++i) {
            if (!spd[i]) throw InvalidError("multigrid coarsest matrix is not positive definite (no fixed boundary?)");
            invOff[i] = (int)inv.size();
            inv.insert(inv.end(), blks[i].begin(), blks[i].end());
        }
    }
    hipStream_t s = M->ctx->stream;
    for (size_t l = 0; l < A.h.size(); ++l) {
        const AmgHostLevel& H = A.h[l];
        auto D = std::make_unique<AmgDevLevel>();
        D->n = H.n;
        D->nSlices = H.nSlices;
        if (!H.agg.empty()) D->agg.upload(H.agg, s);
        if (!H.gCol.empty()) {  // level 0 of a smoothed hierarchy: G replaces q
            std::vector<float> g32(H.g.begin(), H.g.end());
            D->gStart.upload(H.gStart, s);
            D->gWidth.upload(H.gWidth, s);
            D->gCol.upload(H.gCol, s);
            D->g64.upload(H.g, s);
            D->g32.upload(g32, s);
            long long cnt = 0;  // stored (non-padding) entries
            for (int sI = 0; sI < H.nSlices; ++sI)
                for (int lane = 0; lane < kSlice; ++lane) {
                    const int i = sI * kSlice + lane;
                    if (i >= H.n || H.agg[i] < 0) continue;
                    for (int k = 0; k < H.gWidth[sI]; ++k)
                        if (H.g[(size_t)H.gStart[sI] + (size_t)k * kSlice + lane] != 0.0) ++cnt;
                }
            A.qEntries = cnt;
        } else if (!H.agg.empty()) {  // every level with a coarser one: q = offdiag(A) P
            std::vector<int> qc, qw;
            std::vector<double> q;
            groupByAggregate(H, qc, q, qw);
            if (l == 0) A.qEntries = (long long)std::count_if(qc.begin(), qc.end(), [](int c) { return c >= 0; });
            std::vector<float> q32(q.begin(), q.end());
            D->qCol.upload(qc, s);
            D->q64.upload(q, s);
            D->q32.upload(q32, s);
            D->qWidth.upload(qw, s);
        }
        if (!H.memW.empty()) {
            std::vector<float> w32(H.memW.begin(), H.memW.end());
            D->memW64.upload(H.memW, s);
            D->memW32.upload(w32, s);
        }
        if (!H.rMem.empty()) {
            std::vector<float> w32(H.rW.begin(), H.rW.end());
            D->rStart.upload(H.rStart, s);
            D->rWidth.upload(H.rWidth, s);
            D->rMem.upload(H.rMem, s);
            D->rW64.upload(H.rW, s);
            D->rW32.upload(w32, s);
        }
        if (l > 0) {
            D->sliceStart.upload(H.sliceStart, s);
            D->sliceWidth.upload(H.sliceWidth, s);
            D->col.upload(H.col, s);
            D->a64.upload(H.a, s);
            D->d64.upload(H.diag, s);
            std::vector<double> iD(H.n);
            std::vector<float> a32(H.a.size()), d32(H.n), iD32(H.n);
            for (int i = 0; i < H.n; ++i) { iD[i] = 1.0 / H.diag[i]; d32[i] = (float)H.diag[i]; iD32[i] = 1.0f / d32[i]; }
            for (size_t e = 0; e < H.a.size(); ++e) a32[e] = (float)H.a[e];
            D->iD64.upload(iD, s);
            D->a32.upload(a32, s);
            D->d32.upload(d32, s);
            D->iD32.upload(iD32, s);
            D->memOff.upload(H.memOff, s);
            D->mem.upload(H.mem, s);
            D->rowInst.upload(H.rowInst, s);
            if (!H.instPass.empty()) D->instPass.upload(H.instPass, s);
        }
        if (l + 1 == A.h.size() && l > 0) {
            std::vector<float> inv32(inv.size());
            for (size_t k = 0; k < inv.size(); ++k) inv32[k] = (float)inv[k];
            D->inv64.upload(inv, s);
            D->inv32.upload(inv32, s);
            D->blkRow0.upload(H.inst0, s);
            D->blkN.upload(H.instN, s);
            D->invOff.upload(invOff, s);
        }
        A.d.push_back(std::move(D));
    }
    HIP_CHECK(hipStreamSynchronize(s));
    A.gen = M->generation;
    A.buildSeconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// V-cycle work vectors: level 0 holds only the residual (component-major,
// like the PCG vectors); coarser levels hold rows of four
template <typename T>
struct MgVec;
template <>
struct MgVec<double> {
    static double* r0(AmgDevLevel& L) { L.r64.alloc(3 * (size_t)L.n); return L.r64.p; }
    static Row4<double>* x(AmgDevLevel& L) { return v4(L.x64, L.n); }
    static Row4<double>* b(AmgDevLevel& L) { return v4(L.b64, L.n); }
    static Row4<double>* r(AmgDevLevel& L) { return v4(L.r64, L.n); }
    static Row4<double>* xs(AmgDevLevel& L) { return v4(L.xs64, L.n); }
    static const double* inv(AmgDevLevel& L) { return L.inv64.p; }
    static const double* q(AmgDevLevel& L) { return L.q64.p; }
    static const double* g(AmgDevLevel& L) { return L.g64.p; }
    static Row4<double>* v4(DArr<double>& a, int n) {
        a.alloc(4 * (size_t)n);
        return reinterpret_cast<Row4<double>*>(a.p);
    }
};
template <>
struct MgVec<float> {
    static float* r0(AmgDevLevel& L) { L.r32.alloc(3 * (size_t)L.n); return L.r32.p; }
    static Row4<float>* x(AmgDevLevel& L) { return v4(L.x32, L.n); }
    static Row4<float>* b(AmgDevLevel& L) { return v4(L.b32, L.n); }
    static Row4<float>* r(AmgDevLevel& L) { return v4(L.r32, L.n); }
    static Row4<float>* xs(AmgDevLevel& L) { return v4(L.xs32, L.n); }
    static const float* inv(AmgDevLevel& L) { return L.inv32.p; }
    static const float* q(AmgDevLevel& L) { return L.q32.p; }
    static const float* g(AmgDevLevel& L) { return L.g32.p; }
    static Row4<float>* v4(DArr<float>& a, int n) {
        a.alloc(4 * (size_t)n);
        return reinterpret_cast<Row4<float>*>(a.p);
    }
};

inline int gridRows(int n) { return std::max(1, std::min(8192, (n + kBlock - 1) / kBlock)); }
inline int gridSlices(int nSlices) { return std::max(1, std::min(8192, (nSlices + kBlock / 64 - 1) / (kBlock / 64))); }

// the levels a V-cycle runs on: a mesh's own hierarchy, or the server's
// packed hierarchy of its in-flight requests
struct MgRun {
    std::vector<std::unique_ptr<AmgDevLevel>>* levels;
    double omega, alpha;
    int nFin;  // finalize workgroups (instances)
};

MgRun meshRun(s4f_mesh* M, bool smoothed) {
    Amg& A = smoothed ? M->amgSm : M->amg;
    return MgRun{&A.d, A.omega, A.alpha, (int)M->h.instCell0.size()};
}

// level-0 prolongation + post-smoothing: the aggregation form (q) or, on a
// smoothed hierarchy, the G form
template <typename T>
void launchMg0Up(std::vector<std::unique_ptr<AmgDevLevel>>& d, const DevMesh& dm, const BlockMap& bm, PcgVecs<T>& v,
                 double omega, double alpha, Ctl* ctl, double* part, hipStream_t s) {
    AmgDevLevel& D0 = *d[0];
    const Row4<T>* e1 = d.size() > 1 ? (const Row4<T>*)MgVec<T>::xs(*d[1]) : nullptr;
    if (D0.gCol.p && e1)
        hipLaunchKernelGGL((k_mg0_up_g<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const T*)v.r,
                           (const T*)MgVec<T>::r0(D0), (const int*)D0.gStart.p, (const int*)D0.gWidth.p,
                           (const int*)D0.gCol.p, MgVec<T>::g(D0), e1, v.z, omega, alpha, (const Ctl*)ctl, part);
    else
        hipLaunchKernelGGL((k_mg0_up<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const T*)v.r,
                           (const T*)MgVec<T>::r0(D0), (const int*)D0.agg.p, (const int*)D0.qCol.p, MgVec<T>::q(D0),
                           e1, v.z, omega, alpha, (const Ctl*)ctl, part);
}

// z = V(r): one V(1,1) cycle on the 3-RHS residual, then k_fin_rho
template <typename T>
void launchVcycle(const MgRun& R, const DevMesh& dm, const BlockMap& bm, PcgVecs<T>& v, Ctl* ctl, double* part,
                  hipStream_t s) {
    std::vector<std::unique_ptr<AmgDevLevel>>& Ad = *R.levels;
    struct {
        std::vector<std::unique_ptr<AmgDevLevel>>& d;
        double alpha;
    } A{Ad, R.alpha};
    const int nl = (int)A.d.size();
    const double w = R.omega;
    AmgDevLevel& D0 = *A.d[0];
    T* r0 = MgVec<T>::r0(D0);
    hipLaunchKernelGGL((k_mg0_down<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const T*)v.r, r0, w,
                       (const Ctl*)ctl);
    if (nl > 1) {
        // down: restrict the finer level's residual (with the coarse
        // pre-smoothing), then the coarse residual
        for (int l = 1; l < nl; ++l) {
            AmgDevLevel& L = *A.d[l];
            AmgDevLevel& Lf = *A.d[l - 1];
            Row4<T>* b = MgVec<T>::b(L);
            Row4<T>* x = MgVec<T>::x(L);
            if (l == 1 && L.rMem.p)  // smoothed prolongation: weighted lists in sliced ELL
                hipLaunchKernelGGL((k_mg_restrict_w<T>), dim3(gridSlices(L.nSlices)), dim3(kBlock), 0, s, L.view(),
                                   (const T*)r0, Lf.n, b, x, w, (const Ctl*)ctl);
            else if (l == 1)
                hipLaunchKernelGGL((k_mg_restrict<T, true>), dim3(gridRows(L.n)), dim3(kBlock), 0, s, L.view(),
                                   (const T*)r0, (const Row4<T>*)nullptr, Lf.n, b, x, w, (const Ctl*)ctl);
            else
                hipLaunchKernelGGL((k_mg_restrict<T, false>), dim3(gridRows(L.n)), dim3(kBlock), 0, s, L.view(),
                                   (const T*)nullptr, (const Row4<T>*)MgVec<T>::r(Lf), Lf.n, b, x, w,
                                   (const Ctl*)ctl);
            if (l < nl - 1)
                hipLaunchKernelGGL((k_mg_down<T>), dim3(gridSlices(L.nSlices)), dim3(kBlock), 0, s, L.view(),
                                   (const Row4<T>*)b, (const Row4<T>*)x, MgVec<T>::r(L), (const Ctl*)ctl);
        }
        // coarsest: exact solve into xs
        {
            AmgDevLevel& L = *A.d[nl - 1];
            hipLaunchKernelGGL((k_mg_dense<T>), dim3((L.n + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), 0, s,
                               L.view(), MgVec<T>::inv(L), (const Row4<T>*)MgVec<T>::b(L), MgVec<T>::xs(L),
                               (const Ctl*)ctl);
        }
        // up: prolongate the coarser correction (its xs) into the
        // post-smoothing sweep; a level's result goes to its xs (x keeps the
        // pre-smoothed values).  No buffer changes role, so a captured launch
        // sequence replays with the same pointers.
        for (int l = nl - 2; l >= 1; --l) {
            AmgDevLevel& L = *A.d[l];
            AmgDevLevel& Lc = *A.d[l + 1];
            hipLaunchKernelGGL((k_mg_up<T>), dim3(gridSlices(L.nSlices)), dim3(kBlock), 0, s, L.view(),
                               (const Row4<T>*)MgVec<T>::x(L), (const Row4<T>*)MgVec<T>::r(L),
                               (const Row4<T>*)MgVec<T>::xs(Lc), MgVec<T>::xs(L), w, A.alpha, (const Ctl*)ctl);
        }
    }
    launchMg0Up<T>(A.d, dm, bm, v, w, A.alpha, ctl, part, s);
    hipLaunchKernelGGL(k_fin_rho, dim3(R.nFin), dim3(kFinBlock), 0, s, bm, ctl, (const double*)part);
    HIP_CHECK(hipGetLastError());
}

// PCG iterations with the V-cycle: spmv(p = z + beta p), alpha, update,
// convergence, z = V(r), beta
template <typename T>
void launchPcgChunkMg(const MgRun& R, const DevMesh& dm, const BlockMap& bm, PcgVecs<T>& v, Ctl* ctl, double* part,
                      int n, hipStream_t s) {
    for (int i = 0; i < n; ++i) {
        hipLaunchKernelGGL((k_spmvp<T, true>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const T*)v.r, (const T*)v.p,
                           v.p2, v.q, ctl, part, 1, (const T*)v.z);
        const dim3 nI(R.nFin);
        hipLaunchKernelGGL(k_fin_spmv, nI, dim3(kFinBlock), 0, s, bm, ctl, (const double*)part);
        hipLaunchKernelGGL((k_update<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, v.x, (const T*)v.p2, v.r,
                           (const T*)v.q, (const Ctl*)ctl, part);
        hipLaunchKernelGGL(k_fin_update, nI, dim3(kFinBlock), 0, s, bm, ctl, (const double*)part);
        launchVcycle<T>(R, dm, bm, v, ctl, part, s);
        std::swap(v.p, v.p2);
    }
    HIP_CHECK(hipGetLastError());
}

// device copies of the (effective) diagonal and its inverse, FP64 and FP32
void uploadDiagonal(s4f_mesh* M) {
    const HostMesh& h = M->h;
    hipStream_t s = M->ctx->stream;
    std::vector<double> inv(h.C);
    std::vector<float> d32(h.C), i32(h.C);
    for (int c = 0; c < h.C; ++c) {
        inv[c] = 1.0 / h.diag[c];
        d32[c] = (float)h.diag[c];
        i32[c] = 1.0f / d32[c];
    }
    M->diag.upload(h.diag, s);
    M->invDiag.upload(inv, s);
    M->diag32.upload(d32, s);
    M->invDiag32.upload(i32, s);
    HIP_CHECK(hipStreamSynchronize(s));
}

// rho*fvm::d2dt2(D) and rho*g of linGeomTotalDispSolid.C:141-149 for this
// solve.  [OF] EulerD2dt2Scheme::fvmD2dt2: diagonal coefft rDeltaT2 rho V,
// source rDeltaT2 rho V (coefft0 D_old - coefft00 D_oldOld), coefft =
// (dt + dt0)/(2 dt), coefft00 = (dt + dt0)/(2 dt0), coefft0 = coefft +
// coefft00, rDeltaT2 = 4/(dt + dt0)^2.  In defect form the matrix carries the
// diagonal (uploaded, and the multigrid hierarchy rebuilt, when it changes)
// and the residual adds fext - iner D with fext = rho V g + source.
// D.oldTime() .. its fourth level from AoS host fields (levels[l], nCells*3)
void setOldTimes(HostMesh& h, const double* const* levels) {
    const int C = h.C;
    std::vector<double>* dst[4] = {&h.Do, &h.Doo, &h.Dooo, &h.Doooo};
    for (int l = 0; l < 4; ++l) {
        std::vector<double>& d = *dst[l];
        d.resize(3 * (size_t)C);
        for (int c = 0; c < C; ++c)
            for (int q = 0; q < 3; ++q) {
                const double v = levels[l][3 * (size_t)c + q];
                if (!std::isfinite(v)) throw InvalidError("non-finite old-time field");
                d[(size_t)q * C + c] = v;
            }
    }
    h.hasOld = true;
}

void configureSources(s4f_mesh* M, const s4f_solve_opts& o) {
    HostMesh& h = M->h;
    const bool transient = o.d2dt2 != S4F_D2DT2_STEADY;
    const bool backward = o.d2dt2 == S4F_D2DT2_BACKWARD;
    const bool grav = h.g[0] != 0.0 || h.g[1] != 0.0 || h.g[2] != 0.0;
    M->srcOn = M->inerOn = false;
    double key = 0.0, ct = 0.0, ct0 = 0.0, ct00 = 0.0, r2 = 0.0, rdt = 0.0;
    if (transient || grav) {
        if (h.instCell0.size() != 1) throw InvalidError("transient and body-force solves take a single mesh");
        if (h.rho.empty()) throw InvalidError("density not set (s4f_density_set)");
    }
    if (transient) {
        if (!(o.deltaT > 0.0) || !(o.deltaT0 >= 0.0)) throw InvalidError("deltaT must be > 0 (and deltaT0 >= 0)");
        if (!h.hasOld) throw InvalidError("a transient solve needs the old-time fields (s4f_old_time_set)");
        const double dt = o.deltaT, dt0 = o.deltaT0 > 0.0 ? o.deltaT0 : o.deltaT;
        if (backward) {
            // backwardD2dt2Scheme.C:112-117 (constant step: dt0 = dt)
            ct = 1.0 + dt / (dt + dt0);
            ct00 = dt * dt / (dt0 * (dt + dt0));
            ct0 = ct + ct00;
            rdt = 1.0 / dt;
            key = ct * rdt * ct * rdt;
        } else {
            // [OF] EulerD2dt2Scheme
            ct = (dt + dt0) / (2.0 * dt);
            ct00 = (dt + dt0) / (2.0 * dt0);
            ct0 = ct + ct00;
            r2 = 4.0 / ((dt + dt0) * (dt + dt0));
            key = ct * r2;
        }
    }
    const int C = h.C;
    auto rhoV = [&](int c) { return (h.rho.size() == 1 ? h.rho[0] : h.rho[c]) * h.V[c]; };
    if (key != M->inerKey || (key != 0.0 && M->inerRhoGen != M->rhoGen)) {
        h.diag = h.diagBase;
        std::vector<double> iner(C, 0.0);
        if (key != 0.0)
            for (int c = 0; c < C; ++c) {
                iner[c] = key * rhoV(c);
                h.diag[c] += iner[c];
            }
        uploadDiagonal(M);
        M->iner.upload(iner, M->ctx->stream);
        M->inerKey = key;
        M->inerRhoGen = M->rhoGen;
        M->generation++;  // the multigrid hierarchy follows the matrix
    }
    if (!(transient || grav)) return;
    std::vector<double> f(3 * (size_t)C);
    for (int c = 0; c < C; ++c) {
        const double rv = rhoV(c);
        for (int q = 0; q < 3; ++q) {
            const size_t k = (size_t)q * C + c;
            double v = rv * h.g[q];
            if (transient && !backward) v += r2 * rv * (ct0 * h.Do[k] - ct00 * h.Doo[k]);
            if (backward) {
                // coefft rDeltaT fvmDdt(D) source + rDeltaT V (coefft0 fvcDdt(D.oldTime())
                // - coefft00 fvcDdt(D.oldTime().oldTime())), backward ddt of each level
                const double fvmSrc = ct * rdt * rdt * rv * (ct0 * h.Do[k] - ct00 * h.Doo[k]);
                const double ddt0 = rdt * (ct * h.Do[k] - ct0 * h.Doo[k] + ct00 * h.Dooo[k]);
                const double ddt00 = rdt * (ct * h.Doo[k] - ct0 * h.Dooo[k] + ct00 * h.Doooo[k]);
                v += fvmSrc + rdt * rv * (ct0 * ddt0 - ct00 * ddt00);
            }
            f[k] = v;
        }
    }
    M->fext.upload(f, M->ctx->stream);
    HIP_CHECK(hipStreamSynchronize(M->ctx->stream));
    M->srcOn = true;
    M->inerOn = transient;
}

// The fields the gradient step reads and writes.
struct GradBufs {
    double *D0, *D1, *G, *Db, *g, *snb, *Gb, *sigb;
    double *aAlpha = nullptr, *aRes = nullptr, *aAlphaB = nullptr, *aResB = nullptr;  // Aitken state
    double *DI = nullptr, *sigI = nullptr;  // sub-meshes: interface displacement and boundary stress
};

// the gradient-step buffers of a single mesh's solve state
GradBufs meshGradBufs(s4f_mesh* M) {
    SolveState& st = M->st;
    GradBufs f{st.D0.p, st.D1.p, st.G.p, st.Db.p, st.g.p, st.snb.p, st.Gb.p, st.sigb.p};
    if (M->h.nIf > 0) {
        f.DI = st.DI.p;
        f.sigI = st.sigI.p;
    }
    return f;
}

// boundary values, cell gradients, boundary gradients and stress (modes as
// k_grad); nInst instances (bm.finInst, or 0..nInst-1), at most maxNB
// boundary faces each
template <typename T>
void launchGradStep(const DevMesh& dm, const BlockMap& bm, int nInst, int maxNB, const GradBufs& f, const T* x,
                    const double* Dold, Ctl* ctl, double* part, int mode, hipStream_t s) {
    const dim3 fg(std::max(1, std::min(1024, (maxNB + kBlock - 1) / kBlock)), std::max(1, nInst));
    if (mode == 0 && f.aAlpha)
        hipLaunchKernelGGL((k_aitken<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, f.D0, x, f.D1, f.aAlpha,
                           f.aRes, (const Ctl*)ctl);
    if (mode != 2)
        hipLaunchKernelGGL((k_grad_bnd_pre<T>), fg, dim3(kBlock), 0, s, dm, bm, (const double*)f.D0, x,
                           (const double*)f.D1, (const double*)f.G, f.Db, (const double*)f.g, f.snb,
                           (const Ctl*)ctl, f.aAlphaB, f.aResB, mode);
    // sub-meshes: the interface displacement before the gradient, the
    // sub-meshes' interface stress after it (a mode-2 gradient uses the
    // interface state as it is)
    const bool sub = dm.nIf > 0 && mode != 2;
    const dim3 ig(std::max(1, std::min(1024, (dm.nIf + kBlock - 1) / kBlock)));
    if (sub)
        hipLaunchKernelGGL((k_iface_pre<T>), ig, dim3(kBlock), 0, s, dm, (const double*)f.D0, x, (const double*)f.D1,
                           f.DI, (const double*)f.sigI, (const Ctl*)ctl, mode);
    hipLaunchKernelGGL((k_grad<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, f.D0, x, f.D1, f.G,
                       (const double*)f.Db, (const double*)f.DI, Dold, ctl, part, mode);
    if (sub)
        hipLaunchKernelGGL(k_iface_post, ig, dim3(kBlock), 0, s, dm, (const double*)f.D0, (const double*)f.D1,
                           (const double*)f.G, (const double*)f.DI, f.sigI, (const Ctl*)ctl, mode);
    hipLaunchKernelGGL(k_grad_bnd_post, fg, dim3(kBlock), 0, s, dm, bm, (const double*)f.G, (const double*)f.snb,
                       f.Gb, f.sigb, f.g, (const Ctl*)ctl, mode);
}

// An instantiated HIP graph of a captured launch sequence.
struct GraphExec {
    hipGraphExec_t e = nullptr;
    GraphExec() = default;
    GraphExec(const GraphExec&) = delete;
    GraphExec& operator=(const GraphExec&) = delete;
    ~GraphExec() { reset(); }
    void reset() {
        if (e) (void)hipGraphExecDestroy(e);
        e = nullptr;
    }
    // records what f launches on stream `cap` (nothing runs) and instantiates it
    template <typename F>
    void capture(hipStream_t cap, F&& f) {
        reset();
        HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
            f(cap);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(cap, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        hipGraph_t g = nullptr;
        HIP_CHECK(hipStreamEndCapture(cap, &g));
        const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) {
            e = nullptr;
            HIP_CHECK(err);
        }
    }
    void launch(hipStream_t s) const { HIP_CHECK(hipGraphLaunch(e, s)); }
};

struct InstanceIO {
    double* D_inout = nullptr;  // AoS nCells*3
    double* sigma_out = nullptr;
    double* gradD_out = nullptr;
    double* Db_out = nullptr;
    s4f_solve_report* rep = nullptr;
    s4f_status status = S4F_OK;
};

// Mixed precision's fallback to FP64 sweeps: the FP32 run stops at an
// outer-iteration boundary (the last PCG finished, the gradient step not yet
// launched) and an FP64 run continues from the device state as it stands.
struct Fallback {
    bool stalled = false;            // the FP32 run stopped for the fallback
    std::vector<char> live;          // instances still iterating at the switch
    std::vector<long long> instPcg;  // 3-RHS PCG sweeps so far, per instance
};

// Default window of mixed precision's stall test: outer iterations within
// which every iterating instance's |initRes| must reach a new minimum at
// least 1 % below its previous one.
constexpr int kMixedStallWindow = 200;

// The outer momentum loop of linGeomTotalDispSolid::evolve on every instance
// packed in M, with shared launches; instances finish independently.  FP32
// sweeps (T = float) given `fb` stop for the FP64 fallback when an iterating
// instance's |initRes| stalls for the window (fb->stalled); `resume` runs the
// FP64 continuation of such a run.
template <typename T>
void solveT(s4f_mesh* M, const s4f_solve_opts& o, std::vector<InstanceIO>& io, Fallback* fb = nullptr,
            bool resume = false) {
    s4f_ctx* ctx = M->ctx;
    hipStream_t s = ctx->stream;
    HostMesh& h = M->h;
    const int C = h.C;
    const int nInst = (int)h.instCell0.size();
    allocState(M);
    SolveState& st = M->st;
    PcgVecs<T> v = pcgVecs<T>(st, C);
    applyStabilisationScale(M, o.rhieChowScale, o.interfaceScale);
    configureSources(M, o);
    struct SourceOff {  // the sources belong to this solve only
        s4f_mesh* m;
        ~SourceOff() { m->srcOn = m->inerOn = false; }
    } sourceOff{M};
    const bool transient = o.d2dt2 != S4F_D2DT2_STEADY;
    const bool mg = o.preconditioner == S4F_PRECOND_GAMG || o.preconditioner == S4F_PRECOND_GAMG_SMOOTHED;
    const bool mgSmoothed = o.preconditioner == S4F_PRECOND_GAMG_SMOOTHED;
    if (mg) {
        buildAmg(M, mgSmoothed);
        Amg& A = mgSmoothed ? M->amgSm : M->amg;
        // over-correction of the coarse correction (plain aggregation
        // under-represents smooth modes); any scale in (0, 2) keeps the
        // cycle symmetric positive definite
        A.alpha = o.mgCoarseScale > 0.0 ? o.mgCoarseScale : kMgCoarseScaleDefault;
        M->amgLast = &A;
    }
    const DevMesh dm = M->dm();
    const BlockMap bmFull = M->bmap();
    BlockMap bm = bmFull;  // narrowed to the iterating instances as a batch's instances finish
    int nFin = nInst;
    MgRun mgRun = meshRun(M, mgSmoothed);
    // initial and old-time fields (a resumed run keeps the device state)
    std::vector<double> D0(resume ? 0 : 3 * (size_t)C, 0.0);
    bool warm = false;
    for (int i = 0; i < nInst; ++i) {
        if (!(o.warmStart && io[i].D_inout)) continue;
        warm = true;
        if (resume) continue;
        const int c0 = h.instCell0[i], n = h.instNCell[i];
        for (int c = 0; c < n; ++c)
            for (int q = 0; q < 3; ++q) D0[(size_t)q * C + c0 + c] = io[i].D_inout[3 * (size_t)c + q];
    }
    const bool warmIn = warm;
    if (transient) warm = true;
    if (!resume) {
        if (transient && !warmIn) D0 = h.Do;  // a time step starts from the old-time field
        HIP_CHECK(hipMemcpyAsync(st.D0.p, D0.data(), D0.size() * sizeof(double), hipMemcpyHostToDevice, s));
        if (transient) {  // the convergence test's old-time field
            HIP_CHECK(hipMemcpyAsync(st.Dold.p, h.Do.data(), h.Do.size() * sizeof(double), hipMemcpyHostToDevice, s));
        } else if (warm) {
            HIP_CHECK(hipMemcpyAsync(st.Dold.p, D0.data(), D0.size() * sizeof(double), hipMemcpyHostToDevice, s));
        }
        HIP_CHECK(hipMemsetAsync(st.G.p, 0, st.G.bytes(), s));
        HIP_CHECK(hipMemsetAsync(st.g.p, 0, st.g.bytes(), s));
        if (h.nIf > 0 && !warm) {  // sub-meshes: a new run's interface state starts at zero
            HIP_CHECK(hipMemsetAsync(st.DI.p, 0, st.DI.bytes(), s));
            HIP_CHECK(hipMemsetAsync(st.sigI.p, 0, st.sigI.bytes(), s));
        }
        std::vector<Ctl> c0(nInst, makeCtl(o, st.histCap));
        for (int i = 0; i < nInst; ++i) {
            c0[i].warm = ((o.warmStart && io[i].D_inout) || transient) ? 1 : 0;  // a time step continues
            c0[i].sigma0 = c0[i].warm ? 0 : 1;  // a cold start's stress field is zero until the first update
            c0[i].pcMode = mg ? 1 : 0;
            c0[i].solveMask = M->h.instMask.empty() ? 7 : M->h.instMask[i];
        }
        HIP_CHECK(hipMemcpyAsync(st.ctl.p, c0.data(), nInst * sizeof(Ctl), hipMemcpyHostToDevice, s));
        HIP_CHECK(hipMemsetAsync(st.hist.p, 0, st.hist.bytes(), s));
    }
    Ctl* hctl = nullptr;
    HIP_CHECK(hipHostMalloc((void**)&hctl, nInst * sizeof(Ctl), hipHostMallocDefault));
    std::unique_ptr<Ctl, void (*)(Ctl*)> hctlGuard(hctl, [](Ctl* p) { (void)hipHostFree(p); });
    const double* Dold = warm ? st.Dold.p : nullptr;
    if (!resume) HIP_CHECK(hipEventRecord(ctx->ev0, s));  // a resumed run's time counts from the FP32 start
    // constructor: D.correctBoundaryConditions(); grad(D) (linGeomTotalDispSolid.C:82-84)
    GradBufs gbuf = meshGradBufs(M);
    if (o.relaxationMethod == S4F_RELAX_AITKEN) {
        st.aAlpha.alloc(C); st.aRes.alloc(3 * (size_t)C);
        st.aAlphaB.alloc(std::max(1, h.B)); st.aResB.alloc(3 * (size_t)std::max(1, h.B));
        gbuf.aAlpha = st.aAlpha.p; gbuf.aRes = st.aRes.p; gbuf.aAlphaB = st.aAlphaB.p; gbuf.aResB = st.aResB.p;
    }
    if (!resume) {
        launchGradStep<T>(dm, bm, nInst, M->maxInstNB, gbuf, nullptr, Dold, st.ctl.p, st.part.p, 1, s);
        hipLaunchKernelGGL(k_fin_grad, dim3(nInst), dim3(kFinBlock), 0, s, bm, st.ctl.p, (const double*)st.part.p,
                           st.hist.p, 1);
        HIP_CHECK(hipGetLastError());
    }
    // Batches: when instances finish, launch only the tiles of those still
    // iterating.  Each instance keeps its own tiles (and the order of its
    // partials), so its result does not change.
    auto narrow = [&](const std::vector<int>& live) {
        std::vector<int> inst, s0, s1, blk0(nInst, 0), nblk(nInst, 0);
        for (int i : live) {
            blk0[i] = (int)inst.size();
            for (int b = M->tBlk0[i]; b < M->tBlk0[i] + M->tNBlk[i]; ++b) {
                inst.push_back(M->tInst[b]);
                s0.push_back(M->tS0[b]);
                s1.push_back(M->tS1[b]);
            }
            nblk[i] = (int)inst.size() - blk0[i];
        }
        M->aInst.upload(inst, s); M->aS0.upload(s0, s); M->aS1.upload(s1, s);
        M->aBlk0.upload(blk0, s); M->aNBlk.upload(nblk, s); M->aFin.upload(live, s);
        bm = bmFull;
        bm.inst = M->aInst.p; bm.s0 = M->aS0.p; bm.s1 = M->aS1.p; bm.iBlk0 = M->aBlk0.p; bm.iNBlk = M->aNBlk.p;
        bm.finInst = M->aFin.p;
        bm.nBlocks = (int)inst.size();
        nFin = (int)live.size();
        mgRun.nFin = nFin;
    };
    // launch sequences: the explicit terms of an outer iteration (after the
    // previous one's gradient), and PCG iterations.  After the first outer
    // iteration (which allocates every buffer) they are captured into HIP
    // graphs and replayed; PCG iterations go in pairs so the two search-
    // direction buffers are where the captured launches expect them.
    auto launchGradResidual = [&](hipStream_t q) {
        launchGradStep<T>(dm, bm, nFin, M->maxInstNB, gbuf, (const T*)v.x, Dold, st.ctl.p, st.part.p, 0, q);
        hipLaunchKernelGGL(k_fin_grad, dim3(nFin), dim3(kFinBlock), 0, q, bm, st.ctl.p, (const double*)st.part.p,
                           st.hist.p, 0);
    };
    auto launchResidual = [&](hipStream_t q) {
        hipLaunchKernelGGL((k_residual<T>), dim3(bm.nBlocks), dim3(kBlock), 0, q, dm, bm, (const double*)st.D0.p,
                           (const double*)st.D1.p, st.G.p, st.sigb.p, v.r, st.ctl.p, st.part.p);
        hipLaunchKernelGGL(k_fin_residual, dim3(nFin), dim3(kFinBlock), 0, q, bm, st.ctl.p, (const double*)st.part.p);
        if (mg) launchVcycle<T>(mgRun, dm, bm, v, st.ctl.p, st.part.p, q);  // z0 = V(r0), rho0
    };
    auto launchPcg = [&](int n, hipStream_t q) {
        if (mg) launchPcgChunkMg<T>(mgRun, dm, bm, v, st.ctl.p, st.part.p, n, q);
        else launchPcgChunk<T>(dm, bm, nFin, v, st.ctl.p, st.part.p, n, q);
    };
    bool graphs = o.eagerLaunch == 0;
    GraphExec gOuter, gPair;
    bool captured = false;
    int chunk = 4;
    std::vector<long long> instPcg(nInst, 0);
    if (resume) {
        // the FP32 run's last PCG solution in FP64, then the outer iteration it stopped before
        instPcg = fb->instPcg;
        hipLaunchKernelGGL(k_widen, dim3(gridRows(3 * C)), dim3(kBlock), 0, s, (const float*)st.x32.p, (double*)v.x,
                           3 * (size_t)C);
        launchGradResidual(s);
    }
    launchResidual(s);
    HIP_CHECK(hipGetLastError());
    const int stallWindow = o.mixedStallWindow == 0 ? kMixedStallWindow : o.mixedStallWindow;
    const auto tStart = std::chrono::steady_clock::now();
    double tShown = 0.0;
    std::vector<double> bestRes(nInst, HUGE_VAL);
    std::vector<int> bestIt(nInst, 0);
    for (int it = 0;; ++it) {
        bool allDone = false;
        int total = 0;
        for (;;) {
            if (captured) {
                const int pairs = (chunk + 1) / 2;
                for (int k = 0; k < pairs; ++k) gPair.launch(s);
                total += 2 * pairs;
            } else {
                launchPcg(chunk, s);
                total += chunk;
            }
            chunk = std::max(2, chunk / 4);  // top-up chunks after an under-estimate
            HIP_CHECK(hipMemcpyAsync(hctl, st.ctl.p, nInst * sizeof(Ctl), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            allDone = true;
            bool anyInner = false;
            for (int i = 0; i < nInst; ++i) {
                if (!hctl[i].outerDone) allDone = false;
                if (!hctl[i].outerDone && !hctl[i].innerDone) anyInner = true;
            }
            if (allDone || !anyInner) break;
            if (total > o.pcgMaxIter + 2) break;  // maxIter deactivates every component before this
        }
        if (allDone) break;
        if (o.progressSeconds > 0) {  // the per-iteration log, rate-limited
            const double now = std::chrono::duration<double>(std::chrono::steady_clock::now() - tStart).count();
            if (now - tShown >= o.progressSeconds) {
                tShown = now;
                int live = 0, iMax = 0;
                double rMax = 0.0;
                for (int i = 0; i < nInst; ++i) {
                    if (hctl[i].outerDone) continue;
                    ++live;
                    iMax = std::max(iMax, hctl[i].iCorr);
                    rMax = std::max(rMax, hctl[i].initResMag);
                }
                fprintf(stderr, "[s4f] %.0f s: outer iteration %d, %d of %d instances iterating, |initRes| %.3e, "
                        "residualvf %.3e\n", now, iMax, live, nInst, rMax, hctl[0].rvf);
                fflush(stderr);
            }
        }
        int used = 1;
        for (int i = 0; i < nInst; ++i)
            if (!hctl[i].outerDone) used = std::max({used, hctl[i].nIter[0], hctl[i].nIter[1], hctl[i].nIter[2]});
        for (int i = 0; i < nInst; ++i)
            if (!hctl[i].outerDone) instPcg[i] += std::max({hctl[i].nIter[0], hctl[i].nIter[1], hctl[i].nIter[2]});
        // mixed precision's stall test: an iterating instance whose |initRes|
        // found no new minimum 1 % below the last one within the window
        if (std::is_same<T, float>::value && fb && stallWindow > 0) {
            bool stall = false;
            for (int i = 0; i < nInst; ++i) {
                if (hctl[i].outerDone || hctl[i].iCorr < 2) continue;
                const double r = hctl[i].initResMag;
                if (r < 0.99 * bestRes[i]) {
                    bestRes[i] = r;
                    bestIt[i] = it;
                } else if (it - bestIt[i] >= stallWindow) {
                    stall = true;
                }
            }
            if (stall) {
                HIP_CHECK(hipStreamSynchronize(s));
                fb->stalled = true;
                fb->live.assign(nInst, 0);
                for (int i = 0; i < nInst; ++i) fb->live[i] = hctl[i].outerDone ? 0 : 1;
                fb->instPcg = instPcg;
                return;
            }
        }
        // next outer iteration: launch just under the last count, then top up
        chunk = std::max(1, used - 1);
        if (nInst > 1) {  // instances that finished at the previous outer iteration leave the launch
            std::vector<int> live;
            for (int i = 0; i < nInst; ++i)
                if (!hctl[i].outerDone) live.push_back(i);
            if ((int)live.size() < nFin && !live.empty()) {
                HIP_CHECK(hipStreamSynchronize(s));  // queued kernels still read the current tile arrays
                narrow(live);
                captured = false;
            }
        }
        if (graphs && !captured) {
            try {
                gOuter.capture(ctx->capture, [&](hipStream_t q) { launchGradResidual(q); launchResidual(q); });
                gPair.capture(ctx->capture, [&](hipStream_t q) { launchPcg(2, q); });
                captured = true;
            } catch (const HipError&) {  // a runtime without capture: the same launches, one by one
                gOuter.reset();
                gPair.reset();
                (void)hipGetLastError();
                graphs = false;
            }
        }
        if (captured) {
            gOuter.launch(s);
        } else {
            launchGradResidual(s);
            launchResidual(s);
        }
        HIP_CHECK(hipGetLastError());
        if (it > o.nCorrectors + 2) break;  // the device loop ends at nCorrectors
    }
    if (h.nIf > 0)  // sub-meshes: the interface update of the point interpolation after the loop
        hipLaunchKernelGGL((k_iface_pre<T>), dim3(std::max(1, std::min(1024, (h.nIf + kBlock - 1) / kBlock))),
                           dim3(kBlock), 0, s, dm, (const double*)st.D0.p, (const T*)nullptr, (const double*)st.D1.p,
                           st.DI.p, (const double*)st.sigI.p, (const Ctl*)st.ctl.p, 3);
    HIP_CHECK(hipEventRecord(ctx->ev1, s));
    HIP_CHECK(hipMemcpyAsync(hctl, st.ctl.p, nInst * sizeof(Ctl), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    const bool anySigma = std::any_of(io.begin(), io.end(), [](const InstanceIO& x) { return x.sigma_out != nullptr; });
    std::vector<double> hD0 = download(st.D0, 3 * (size_t)C, s), hD1 = download(st.D1, 3 * (size_t)C, s);
    std::vector<double> hG, hSig, hDb;
    if (std::any_of(io.begin(), io.end(), [](const InstanceIO& x) { return x.gradD_out != nullptr; }))
        hG = download(st.G, 9 * (size_t)C, s);
    if (anySigma) {
        st.sig.alloc(6 * (size_t)C);
        hipLaunchKernelGGL(k_sigma, dim3(std::min(4096, (C + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, dm, st.G.p,
                           st.sig.p, 0, C);
        HIP_CHECK(hipGetLastError());
        hSig = download(st.sig, 6 * (size_t)C, s);
    }
    if (std::any_of(io.begin(), io.end(), [](const InstanceIO& x) { return x.Db_out != nullptr; }) && h.B)
        hDb = download(st.Db, 3 * (size_t)h.B, s);
    st.hHist = download(st.hist, st.hist.n, s);
    st.lastNOuter = hctl[0].nOuter;
    st.lastDpar = nInst == 1 ? hctl[0].dpar : -1;
    for (int i = 0; i < nInst; ++i) {
        const Ctl& cf = hctl[i];
        const int c0i = h.instCell0[i], n = h.instNCell[i], b0 = h.instB0[i], nb = h.instNB[i];
        // the final field of instance i is in the buffer its parity flag names
        const std::vector<double>& Df = cf.dpar ? hD1 : hD0;
        auto copyOut = [&](const std::vector<double>& src, int k, int off, int cnt, size_t stride, double* out) {
            for (int c = 0; c < cnt; ++c)
                for (int q = 0; q < k; ++q) out[(size_t)c * k + q] = src[(size_t)q * stride + off + c];
        };
        if (io[i].D_inout) copyOut(Df, 3, c0i, n, C, io[i].D_inout);
        if (io[i].gradD_out) copyOut(hG, 9, c0i, n, C, io[i].gradD_out);
        if (io[i].sigma_out) copyOut(hSig, 6, c0i, n, C, io[i].sigma_out);
        if (io[i].Db_out && nb) copyOut(hDb, 3, b0, nb, h.B, io[i].Db_out);
        if (io[i].rep) {
            s4f_solve_report* rep = io[i].rep;
            rep->nOuter = cf.nOuter;
            rep->converged = cf.converged;
            rep->innerTotal = cf.innerTotal;
            rep->initialResidual = cf.initResMag;
            rep->residualvf = cf.rvf;
            rep->seconds = ms * 1e-3;
            rep->mixedFallback = (resume && fb->live[i]) ? 1 : 0;
            rep->warmStarted = cf.warm;
            rep->pcgIterations = instPcg[i];
        }
        io[i].status = cf.diverged ? S4F_ERR_DIVERGED : (cf.converged ? S4F_OK : S4F_WARN_MAXITER);
    }
}

void solveDispatch(s4f_mesh* M, const s4f_solve_opts& o, std::vector<InstanceIO>& io) {
    if (o.precision == S4F_PRECISION_MIXED) {
        Fallback fb;
        solveT<float>(M, o, io, &fb);
        if (fb.stalled) solveT<double>(M, o, io, &fb, true);
    } else {
        solveT<double>(M, o, io);
    }
}

// Concatenates meshes (their material and boundary values included) into one
// packed structure: each instance's cells start on a slice boundary; column,
// boundary-face and cell indices are offset; material data is per cell.  The
// steady matrix (diagBase: no inertia term of an earlier transient solve) and
// the Rhie-Chow coefficients of the given stabilisation scales are packed.
void packMeshes(const std::vector<const HostMesh*>& ms, HostMesh& P, double rcScale, double ifScale) {
    P = HostMesh();
    const int n = (int)ms.size();
    for (const HostMesh* m : ms)
        if (!m->zone.empty()) throw InvalidError("a mesh with material zones (sub-meshes) is solved alone");
    for (const HostMesh* m : ms)
        if (m->g[0] != 0.0 || m->g[1] != 0.0 || m->g[2] != 0.0)
            throw InvalidError("body-force solves take a single mesh");
    const bool defaultScales = rcScale == kRhieChowScaleDefault && ifScale == kInterfaceScaleDefault;
    long long C = 0, B = 0, E = 0, S = 0;
    for (const HostMesh* m : ms) {
        P.instCell0.push_back((int)C);
        P.instNCell.push_back(m->C);
        P.instB0.push_back((int)B);
        P.instNB.push_back(m->B);
        P.instMask.push_back(m->instMask.empty() ? 7 : m->instMask[0]);
        P.instTileCap.push_back(m->instTileCap.empty() ? gMaxBlocks : m->instTileCap[0]);
        P.instTileMin.push_back(m->instTileMin.empty() ? kMinTileSlices : m->instTileMin[0]);
        C += (long long)m->nSlices * kSlice;
        B += m->B;
        E += (long long)m->col.size();
        S += m->nSlices;
    }
    if (C > INT32_MAX / 9 || E > INT32_MAX || B > INT32_MAX / 9) throw InvalidError("batch too large for int32 indexing");
    P.C = (int)C; P.B = (int)B; P.nSlices = (int)S; P.F = 0;
    P.uniformMaterial = false;
    P.sliceStart.resize(S + 1);
    P.sliceWidth.resize(S);
    P.col.resize(E);
    P.eS.resize(3 * E); P.eW.resize(E); P.eA.resize(E); P.eC.resize(E);
    P.cC.assign(3 * C, 0.0);
    P.minv.assign(6 * C, 0.0);
    P.diag.assign(C, 1.0);
    P.mu.assign(C, 1.0); P.K.assign(C, 1.0); P.impK.assign(C, 1.0);
    P.bCell.resize(B); P.bType.resize(B);
    P.bN.resize(3 * B); P.bKv.resize(3 * B); P.bLs.resize(3 * B); P.bDelta.resize(B);
    P.bVal.resize(3 * B); P.bPres.resize(B);
    long long e0 = 0, s0 = 0;
    for (int i = 0; i < n; ++i) {
        const HostMesh& m = *ms[i];
        const int c0 = P.instCell0[i], b0 = P.instB0[i];
        const size_t Em = m.col.size();
        const std::vector<double> eCs = defaultScales ? std::vector<double>() : stabilisationCoeffs(m, rcScale, ifScale);
        const std::vector<double>& eC = defaultScales ? m.eC : eCs;
        for (int s = 0; s < m.nSlices; ++s) {
            P.sliceStart[s0 + s] = (int)(e0 + m.sliceStart[s]);
            P.sliceWidth[s0 + s] = m.sliceWidth[s];
        }
        for (size_t e = 0; e < Em; ++e) {
            const int c = m.col[e];
            P.col[e0 + e] = c >= 0 ? c + c0 : (c == kPad ? kPad : -1 - ((-1 - c) + b0));
            for (int q = 0; q < 3; ++q) P.eS[(size_t)q * E + e0 + e] = m.eS[(size_t)q * Em + e];
            P.eW[e0 + e] = m.eW[e];
            P.eA[e0 + e] = m.eA[e];
            P.eC[e0 + e] = eC[e];
        }
        for (int c = 0; c < m.C; ++c) {
            for (int q = 0; q < 3; ++q) P.cC[(size_t)q * C + c0 + c] = m.cC[(size_t)q * m.C + c];
            for (int q = 0; q < 6; ++q) P.minv[(size_t)q * C + c0 + c] = m.minv[(size_t)q * m.C + c];
            P.diag[c0 + c] = m.diagBase.empty() ? m.diag[c] : m.diagBase[c];
            const int mc = m.uniformMaterial ? 0 : c;
            P.mu[c0 + c] = m.mu[mc];
            P.K[c0 + c] = m.K[mc];
            P.impK[c0 + c] = m.impK[mc];
        }
        for (int b = 0; b < m.B; ++b) {
            P.bCell[b0 + b] = m.bCell[b] + c0;
            P.bType[b0 + b] = m.bType[b];
            P.bDelta[b0 + b] = m.bDelta[b];
            P.bPres[b0 + b] = m.bPres[b];
            for (int q = 0; q < 3; ++q) {
                P.bN[(size_t)q * B + b0 + b] = m.bN[(size_t)q * m.B + b];
                P.bKv[(size_t)q * B + b0 + b] = m.bKv[(size_t)q * m.B + b];
                P.bLs[(size_t)q * B + b0 + b] = m.bLs[(size_t)q * m.B + b];
                P.bVal[(size_t)q * B + b0 + b] = m.bVal[(size_t)q * m.B + b];
            }
        }
        e0 += (long long)Em;
        s0 += m.nSlices;
    }
    P.sliceStart[S] = (int)E;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* s4f_last_error(void) { return gLastError.c_str(); }
const char* s4f_version(void) { return "s4f-mi355x 0.1 (gfx950)"; }

s4f_status s4f_ctx_create(int device, s4f_ctx** out) {
    return guarded([&]() -> s4f_status {
        if (!out) throw InvalidError("null output");
        int n = 0;
        HIP_CHECK(hipGetDeviceCount(&n));
        if (device < 0 || device >= n) throw InvalidError("no HIP device " + std::to_string(device));
        HIP_CHECK(hipSetDevice(device));
        scratchInit(device);
        auto c = std::make_unique<s4f_ctx>();
        c->device = device;
        HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HIP_CHECK(hipStreamCreateWithFlags(&c->capture, hipStreamNonBlocking));
        HIP_CHECK(hipEventCreate(&c->ev0));
        HIP_CHECK(hipEventCreate(&c->ev1));
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, device));
        c->nCU = prop.multiProcessorCount;
        *out = c.release();
        return S4F_OK;
    });
}

s4f_status s4f_ctx_destroy(s4f_ctx* ctx) {
    if (!ctx) return S4F_OK;
    if (ctx->destroyPending) return fail(S4F_ERR_INVALID, "context already destroyed");
    ctx->destroyPending = true;
    if (ctx->children == 0) ctxRelease(ctx);
    return S4F_OK;
}

s4f_status s4f_polymesh_box(int nx, int ny, int nz, const double L[3], double jitter, uint64_t seed, int tets,
                            s4f_polymesh** out) {
    return guarded([&]() -> s4f_status {
        if (!out || !L) throw InvalidError("null argument");
        auto p = std::make_unique<s4f_polymesh>();
        p->pm.reset(makeBoxMesh(nx, ny, nz, L, jitter, seed, tets != 0));
        *out = p.release();
        return S4F_OK;
    });
}

s4f_status s4f_polymesh_from_arrays(int32_t nPoints, const double* points, int32_t nCells, int32_t nFaces,
                                    int32_t nInternalFaces, const int32_t* faceOffsets, const int32_t* faceVerts,
                                    const int32_t* owner, const int32_t* neighbour, int32_t nPatches,
                                    const int32_t* patchStart, const int32_t* patchSize, s4f_polymesh** out) {
    return guarded([&]() -> s4f_status {
        if (!out || !points || !faceOffsets || !faceVerts || !owner || (nInternalFaces && !neighbour))
            throw InvalidError("null argument");
        if (nCells <= 0 || nFaces <= 0 || nInternalFaces < 0 || nInternalFaces > nFaces || nPatches < 0)
            throw InvalidError("invalid sizes");
        auto p = std::make_unique<s4f_polymesh>();
        p->pm = std::make_unique<PolyMesh>();
        PolyMesh& m = *p->pm;
        m.nPoints = nPoints; m.nCells = nCells; m.nFaces = nFaces; m.nInternalFaces = nInternalFaces; m.nPatches = nPatches;
        m.points.assign(points, points + 3 * (size_t)nPoints);
        m.fOff.assign(faceOffsets, faceOffsets + nFaces + 1);
        for (int f = 0; f < nFaces; ++f)
            if (m.fOff[f + 1] < m.fOff[f] + 3) throw InvalidError("faces need >= 3 points");
        m.fVerts.assign(faceVerts, faceVerts + m.fOff[nFaces]);
        for (int v : m.fVerts)
            if (v < 0 || v >= nPoints) throw InvalidError("face point out of range");
        m.owner.assign(owner, owner + nFaces);
        m.neighbour.assign(neighbour, neighbour + nInternalFaces);
        for (int f = 0; f < nFaces; ++f)
            if (m.owner[f] < 0 || m.owner[f] >= nCells) throw InvalidError("owner out of range");
        for (int f = 0; f < nInternalFaces; ++f)
            if (m.neighbour[f] < 0 || m.neighbour[f] >= nCells) throw InvalidError("neighbour out of range");
        m.patchStart.assign(patchStart, patchStart + nPatches);
        m.patchSize.assign(patchSize, patchSize + nPatches);
        m.computeGeometry();
        *out = p.release();
        return S4F_OK;
    });
}

s4f_status s4f_polymesh_free(s4f_polymesh* pm) {
    delete pm;
    return S4F_OK;
}

s4f_status s4f_polymesh_sizes(const s4f_polymesh* p, int64_t out[6]) {
    if (!p || !out) return fail(S4F_ERR_INVALID, "null argument");
    const PolyMesh& m = *p->pm;
    out[0] = m.nPoints; out[1] = m.nCells; out[2] = m.nFaces; out[3] = m.nInternalFaces; out[4] = m.nPatches;
    out[5] = m.fOff.empty() ? 0 : m.fOff[m.nFaces];
    return S4F_OK;
}

s4f_status s4f_polymesh_addressing(const s4f_polymesh* p, double* points, int32_t* faceOffsets, int32_t* faceVerts,
                                   int32_t* owner, int32_t* neighbour, int32_t* patchStart, int32_t* patchSize) {
    if (!p) return fail(S4F_ERR_INVALID, "null argument");
    const PolyMesh& m = *p->pm;
    if (points) std::memcpy(points, m.points.data(), m.points.size() * sizeof(double));
    if (faceOffsets) std::memcpy(faceOffsets, m.fOff.data(), m.fOff.size() * sizeof(int));
    if (faceVerts) std::memcpy(faceVerts, m.fVerts.data(), m.fVerts.size() * sizeof(int));
    if (owner) std::memcpy(owner, m.owner.data(), m.owner.size() * sizeof(int));
    if (neighbour) std::memcpy(neighbour, m.neighbour.data(), m.neighbour.size() * sizeof(int));
    if (patchStart) std::memcpy(patchStart, m.patchStart.data(), m.patchStart.size() * sizeof(int));
    if (patchSize) std::memcpy(patchSize, m.patchSize.data(), m.patchSize.size() * sizeof(int));
    return S4F_OK;
}

s4f_status s4f_polymesh_geometry(const s4f_polymesh* p, double* Cf, double* Sf, double* C, double* V) {
    if (!p) return fail(S4F_ERR_INVALID, "null argument");
    const PolyMesh& m = *p->pm;
    if (Cf) std::memcpy(Cf, m.Cf.data(), m.Cf.size() * sizeof(double));
    if (Sf) std::memcpy(Sf, m.Sf.data(), m.Sf.size() * sizeof(double));
    if (C) std::memcpy(C, m.C.data(), m.C.size() * sizeof(double));
    if (V) std::memcpy(V, m.V.data(), m.V.size() * sizeof(double));
    return S4F_OK;
}

s4f_status s4f_mesh_create(s4f_ctx* ctx, const s4f_mesh_desc* desc, s4f_mesh** out) {
    return guarded([&]() -> s4f_status {
        if (!ctx || !desc || !out) throw InvalidError("null argument");
        if (ctx->destroyPending) throw InvalidError("context was destroyed");
        HIP_CHECK(hipSetDevice(ctx->device));
        auto M = std::make_unique<s4f_mesh>();
        M->ctx = ctx;
        setupHostMesh(M->h, *desc);
        M->bcSet.assign(M->h.nPatches, 0);
        uploadGeometry(M.get());
        ctx->children++;
        *out = M.release();
        return S4F_OK;
    });
}

s4f_status s4f_mesh_create_from_polymesh(s4f_ctx* ctx, const s4f_polymesh* p, const int32_t* patchType,
                                         int32_t lsFullDelta, s4f_mesh** out) {
    if (!p || !patchType) return fail(S4F_ERR_INVALID, "null argument");
    const PolyMesh& m = *p->pm;
    s4f_mesh_desc d{};
    d.nCells = m.nCells;
    d.nFaces = m.nFaces;
    d.nInternalFaces = m.nInternalFaces;
    d.owner = m.owner.data();
    d.neighbour = m.neighbour.data();
    d.C = m.C.data();
    d.V = m.V.data();
    d.Sf = m.Sf.data();
    d.Cf = m.Cf.data();
    d.nPatches = m.nPatches;
    d.patchStart = m.patchStart.data();
    d.patchSize = m.patchSize.data();
    d.patchType = patchType;
    d.lsFullDelta = lsFullDelta;
    return s4f_mesh_create(ctx, &d, out);
}

s4f_status s4f_mesh_destroy(s4f_mesh* mesh) {
    if (!mesh) return S4F_OK;
    s4f_ctx* ctx = mesh->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete mesh;
    ctxChildGone(ctx);
    return S4F_OK;
}

s4f_status s4f_mesh_tiles_set(s4f_mesh* M, int32_t maxTiles, int32_t minSlices) {
    return guarded([&]() -> s4f_status {
        if (!M) throw InvalidError("null argument");
        if (M->h.instCell0.size() != 1) throw InvalidError("tile rules belong to single meshes");
        if (maxTiles < 1 || minSlices < 1 || minSlices > 1024) throw InvalidError("maxTiles >= 1, 1 <= minSlices <= 1024");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        HIP_CHECK(hipStreamSynchronize(M->ctx->stream));
        M->h.instTileCap = {maxTiles};
        M->h.instTileMin = {minSlices};
        uploadTiles(M);
        M->tileGen++;  // a batch packing this mesh re-packs
        return S4F_OK;
    });
}

// Least-squares inverse of cell P over its ELL row, the bi-material
// interface entries taken as boundary faces of the cell's sub-mesh
// (interface delta dIf(entry)); boundary entries from their vectors
// bLs = d/|d|^2 (d d^T/|d|^2 = bLs bLs^T/|bLs|^2).  Returns the inverse and,
// through dd, nothing else: the caller forms the interface vectors.
static void subMeshLsInverse(const HostMesh& h, int P, int emptyDir,
                             const std::function<void(int, double*)>& dIf, double inv[6]) {
    const int s = P / kSlice, lane = P % kSlice, E = (int)h.col.size();
    (void)E;
    double t[6] = {0, 0, 0, 0, 0, 0};
    auto add = [&](const double* v, double md2) {
        t[0] += v[0] * v[0] / md2; t[1] += v[0] * v[1] / md2; t[2] += v[0] * v[2] / md2;
        t[3] += v[1] * v[1] / md2; t[4] += v[1] * v[2] / md2; t[5] += v[2] * v[2] / md2;
    };
    for (int k = 0; k < h.sliceWidth[s]; ++k) {
        const int e = h.sliceStart[s] + k * kSlice + lane;
        const int n = h.col[e];
        if (n == kPad) continue;
        double v[3];
        if (n >= 0 && h.eIf[e] >= 0) {
            dIf(e, v);
        } else if (n >= 0) {
            for (int q = 0; q < 3; ++q) v[q] = h.cC[(size_t)q * h.C + n] - h.cC[(size_t)q * h.C + P];
        } else {
            const int b = -1 - n;
            for (int q = 0; q < 3; ++q) v[q] = h.bLs[(size_t)q * h.B + b];
            if (v[0] == 0.0 && v[1] == 0.0 && v[2] == 0.0) continue;
        }
        add(v, v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    const int emptyDiag = emptyDir < 0 ? -1 : (emptyDir == 0 ? 0 : (emptyDir == 1 ? 3 : 5));
    if (emptyDiag >= 0) t[emptyDiag] += 1.0;
    const double xx = t[0], xy = t[1], xz = t[2], yy = t[3], yz = t[4], zz = t[5];
    const double det = xx * yy * zz + xy * yz * xz + xz * xy * yz - xx * yz * yz - xy * xy * zz - xz * yy * xz;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
        throw InvalidError("singular sub-mesh least-squares matrix in cell " + std::to_string(P));
    inv[0] = (yy * zz - yz * yz) / det; inv[1] = (xz * yz - xy * zz) / det; inv[2] = (xy * yz - xz * yy) / det;
    inv[3] = (xx * zz - xz * xz) / det; inv[4] = (xy * xz - xx * yz) / det; inv[5] = (xx * yy - xy * xy) / det;
    if (emptyDiag >= 0) inv[emptyDiag] -= 1.0;
}

// Material zones of a multi-law case (mechanicalModel with one law per
// cellZone, solidSubMeshes.C:47-110): the internal faces between zones
// become bi-material interface faces, each a boundary face of the sub-mesh
// on either side (interface data, sub-mesh least-squares inverses of the
// cells next to them, the ELL entries that carry them).
s4f_status s4f_mesh_zones_set(s4f_mesh* M, const int32_t* cellZone, const double* faceCentres) {
    return guarded([&]() -> s4f_status {
        if (!M) throw InvalidError("null argument");
        HostMesh& h = M->h;
        if (h.instCell0.size() != 1) throw InvalidError("material zones belong to single meshes");
        if (cellZone && !faceCentres) throw InvalidError("material zones need the mesh's face centres");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        hipStream_t s = M->ctx->stream;
        HIP_CHECK(hipStreamSynchronize(s));
        const int C = h.C, F = h.F;
        const size_t E = h.col.size();
        if (!h.minv0.empty()) h.minv = h.minv0;  // the whole-mesh inverses back
        M->st.lastDpar = -1;  // the last solve's fields belong to the old zones
        h.zone.clear();
        h.nIf = 0;
        h.ifCell.clear(); h.ifN.clear(); h.ifDc.clear(); h.ifLs.clear(); h.eIf.clear(); h.sliceIf.clear();
        if (cellZone) {
            h.zone.assign(cellZone, cellZone + C);
            std::vector<int> ifFace;
            for (int f = 0; f < F; ++f)
                if (h.zone[h.owner[f]] != h.zone[h.neighbour[f]]) ifFace.push_back(f);
            const int nI = (int)ifFace.size(), nI2 = 2 * nI;
            h.nIf = nI;
            if (nI > 0) {
                if (h.minv0.empty()) h.minv0 = h.minv;
                int emptyDir = -1;
                const int mask = h.instMask.empty() ? 7 : h.instMask[0];
                for (int q = 0; q < 3; ++q)
                    if (!(mask & (1 << q))) emptyDir = q;
                h.ifCell.resize(nI2); h.ifN.resize(3 * (size_t)nI); h.ifDc.resize(nI2); h.ifLs.assign(3 * (size_t)nI2, 0.0);
                h.eIf.assign(E, -1);
                h.sliceIf.assign(h.nSlices, 0);
                std::vector<double> dIf(3 * (size_t)nI2);  // each side's patch delta
                for (int i = 0; i < nI; ++i) {
                    const int f = ifFace[i];
                    const int eo = h.entOwn[f], en = h.entNei[f];
                    double n[3];
                    for (int q = 0; q < 3; ++q) n[q] = h.eS[(size_t)q * E + eo];
                    const double ms = std::sqrt(dot3(n, n));
                    for (int q = 0; q < 3; ++q) {
                        n[q] /= ms;
                        h.ifN[(size_t)q * nI + i] = n[q];
                    }
                    for (int side = 0; side < 2; ++side) {
                        const int P = side ? h.neighbour[f] : h.owner[f];
                        double dl[3];
                        for (int q = 0; q < 3; ++q) dl[q] = faceCentres[3 * (size_t)f + q] - h.cC[(size_t)q * C + P];
                        const double nd = dot3(n, dl);
                        h.ifDc[2 * i + side] = 1.0 / std::fabs(nd);
                        for (int q = 0; q < 3; ++q) dIf[3 * (size_t)(2 * i + side) + q] = h.lsFullDelta ? dl[q] : n[q] * nd;
                        h.ifCell[2 * i + side] = P;
                        const int e = side ? en : eo;
                        h.eIf[e] = 2 * i + side;
                        h.sliceIf[P / kSlice] = 1;
                    }
                }
                auto dOf = [&](int e, double* v) {
                    for (int q = 0; q < 3; ++q) v[q] = dIf[3 * (size_t)h.eIf[e] + q];
                };
                std::vector<char> done(C, 0);
                for (int k = 0; k < nI2; ++k) {
                    const int P = h.ifCell[k];
                    if (done[P]) continue;
                    done[P] = 1;
                    double inv[6];
                    subMeshLsInverse(h, P, emptyDir, dOf, inv);
                    for (int q = 0; q < 6; ++q) h.minv[(size_t)q * C + P] = inv[q];
                }
                for (int k = 0; k < nI2; ++k) {  // interface vectors d/|d|^2 (k_grad applies minv, as for bLs)
                    const double* d = &dIf[3 * (size_t)k];
                    const double r = 1.0 / dot3(d, d);
                    for (int q = 0; q < 3; ++q) h.ifLs[(size_t)q * nI2 + k] = d[q] * r;
                }
            }
        }
        M->minv.upload(h.minv, s);
        M->zone.upload(h.zone, s);
        M->ifCell.upload(h.ifCell, s);
        M->ifN.upload(h.ifN, s);
        M->ifDc.upload(h.ifDc, s);
        M->ifLs.upload(h.ifLs, s);
        M->eIf.upload(h.eIf, s);
        M->sliceIf.upload(h.sliceIf, s);
        M->st.DI.alloc(3 * (size_t)h.nIf);
        M->st.sigI.alloc(12 * (size_t)h.nIf);
        if (h.nIf > 0) {
            HIP_CHECK(hipMemsetAsync(M->st.DI.p, 0, M->st.DI.bytes(), s));
            HIP_CHECK(hipMemsetAsync(M->st.sigI.p, 0, M->st.sigI.bytes(), s));
        }
        HIP_CHECK(hipStreamSynchronize(s));
        return S4F_OK;
    });
}

s4f_status s4f_mesh_info(const s4f_mesh* m, int64_t out[6]) {
    if (!m || !out) return fail(S4F_ERR_INVALID, "null argument");
    out[0] = m->h.C; out[1] = m->h.F; out[2] = m->h.B; out[3] = (int64_t)m->h.col.size(); out[4] = m->h.nSlices;
    size_t bytes = m->sliceStart.bytes() + m->col.bytes() + m->eS.bytes() + m->eW.bytes() + m->eA.bytes() + m->eC.bytes() +
                   m->eA32.bytes() + m->cC.bytes() + m->minv.bytes() + m->diag.bytes() + m->invDiag.bytes() +
                   m->diag32.bytes() + m->invDiag32.bytes() + m->bN.bytes() + m->bKv.bytes() + m->bLs.bytes() +
                   m->bDelta.bytes() + m->bVal.bytes() + m->bPres.bytes() + m->bCell.bytes() + m->bType.bytes() +
                   m->ifCell.bytes() + m->eIf.bytes() + m->zone.bytes() + m->ifN.bytes() + m->ifDc.bytes() +
                   m->ifLs.bytes() + m->sliceIf.bytes();
    out[5] = (int64_t)bytes;
    return S4F_OK;
}

s4f_status s4f_mesh_mg_info(const s4f_mesh* m, int32_t maxLevels, int32_t* nLevels, int64_t* rows, int64_t* entries,
                            int64_t* qEntries) {
    if (!m || !nLevels || maxLevels < 0 || (maxLevels > 0 && (!rows || !entries)))
        return fail(S4F_ERR_INVALID, "null argument");
    const Amg& A = m->amgLast ? *m->amgLast : m->amg;
    if (A.d.empty() || A.gen != m->generation)
        return fail(S4F_ERR_INVALID, "no multigrid hierarchy for the current matrix (run a GAMG solve first)");
    *nLevels = (int32_t)A.h.size();
    for (int l = 0; l < std::min<int>(maxLevels, (int)A.h.size()); ++l) {
        const AmgHostLevel& H = A.h[l];
        rows[l] = H.n;
        entries[l] = (int64_t)std::count_if(H.col.begin(), H.col.end(), [](int c) { return c >= 0; });
    }
    if (qEntries) *qEntries = A.qEntries;
    return S4F_OK;
}

s4f_status s4f_material_set(s4f_mesh* M, const double* E, const double* nu, int32_t perCell) {
    return guarded([&]() -> s4f_status {
        if (!M || !E || !nu) throw InvalidError("null argument");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        computeMaterial(M->h, E, nu, perCell != 0);
        uploadMaterial(M);
        M->generation++;
        M->rcScale = kRhieChowScaleDefault;
        M->ifScale = kInterfaceScaleDefault;
        M->materialSet = true;
        return S4F_OK;
    });
}

s4f_status s4f_plane_stress_set(s4f_mesh* M, int32_t planeStress) {
    return guarded([&]() -> s4f_status {
        if (!M) throw InvalidError("null argument");
        const bool on = planeStress != 0;
        if (M->h.planeStress == on) return S4F_OK;
        M->h.planeStress = on;
        if (M->materialSet) {  // the moduli change now, from the E and nu last set
            HIP_CHECK(hipSetDevice(M->ctx->device));
            const std::vector<double> E = M->h.matE, nu = M->h.matNu;
            computeMaterial(M->h, E.data(), nu.data(), !M->h.uniformMaterial);
            uploadMaterial(M);
            M->generation++;
            M->rcScale = kRhieChowScaleDefault;
            M->ifScale = kInterfaceScaleDefault;
        }
        return S4F_OK;
    });
}

s4f_status s4f_density_set(s4f_mesh* M, const double* rho, int32_t perCell) {
    return guarded([&]() -> s4f_status {
        if (!M || !rho) throw InvalidError("null argument");
        const int n = perCell ? M->h.C : 1;
        for (int i = 0; i < n; ++i)
            if (!(rho[i] > 0.0) || !std::isfinite(rho[i])) throw InvalidError("density must be > 0");
        M->h.rho.assign(rho, rho + n);
        M->rhoGen++;
        return S4F_OK;
    });
}

s4f_status s4f_gravity_set(s4f_mesh* M, const double* g) {
    return guarded([&]() -> s4f_status {
        if (!M || !g) throw InvalidError("null argument");
        for (int q = 0; q < 3; ++q) {
            if (!std::isfinite(g[q])) throw InvalidError("non-finite gravity");
            M->h.g[q] = g[q];
        }
        return S4F_OK;
    });
}

s4f_status s4f_old_time_set(s4f_mesh* M, const double* Dold, const double* DoldOld) {
    return guarded([&]() -> s4f_status {
        if (!M || !Dold) throw InvalidError("null argument");
        HostMesh& h = M->h;
        if (h.instCell0.size() != 1) throw InvalidError("old-time fields belong to single meshes");
        const double* lv[4] = {Dold, DoldOld ? DoldOld : Dold, nullptr, nullptr};
        lv[2] = lv[3] = lv[1];
        setOldTimes(h, lv);
        return S4F_OK;
    });
}

s4f_status s4f_old_times_set(s4f_mesh* M, int32_t nLevels, const double* levels) {
    return guarded([&]() -> s4f_status {
        if (!M || !levels) throw InvalidError("null argument");
        if (nLevels < 1 || nLevels > 4) throw InvalidError("nLevels must be 1..4");
        HostMesh& h = M->h;
        if (h.instCell0.size() != 1) throw InvalidError("old-time fields belong to single meshes");
        const double* lv[4];
        for (int l = 0; l < 4; ++l) lv[l] = levels + (size_t)std::min(l, nLevels - 1) * 3 * (size_t)h.C;
        setOldTimes(h, lv);
        return S4F_OK;
    });
}

s4f_status s4f_bc_set(s4f_mesh* M, int32_t patch, const double* value, const double* pressure) {
    return guarded([&]() -> s4f_status {
        if (!M || !value) throw InvalidError("null argument");
        HostMesh& h = M->h;
        if (patch < 0 || patch >= h.nPatches) throw InvalidError("patch index out of range");
        const int b0 = h.patchStart[patch] - h.F, n = h.patchSize[patch];
        for (int i = 0; i < n; ++i) {
            for (int q = 0; q < 3; ++q) {
                const double v = value[3 * (size_t)i + q];
                if (!std::isfinite(v)) throw InvalidError("non-finite boundary value");
                h.bVal[(size_t)q * h.B + b0 + i] = v;
            }
            h.bPres[b0 + i] = pressure ? pressure[i] : 0.0;
        }
        HIP_CHECK(hipSetDevice(M->ctx->device));
        M->bVal.upload(h.bVal, M->ctx->stream);
        M->bPres.upload(h.bPres, M->ctx->stream);
        HIP_CHECK(hipStreamSynchronize(M->ctx->stream));
        M->bcSet[patch] = 1;
        return S4F_OK;
    });
}

void s4f_solve_opts_default(s4f_solve_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->nCorrectors = 10000;
    o->solutionTolerance = 1e-6;
    o->alternativeTolerance = 1e-7;
    o->rhieChowScale = 0.1;
    o->interfaceScale = 0.01;
    o->pcgTolerance = 1e-9;
    o->pcgRelTol = 0.1;
    o->pcgMaxIter = 1000;
    o->preconditioner = S4F_PRECOND_DIAGONAL;
    o->precision = S4F_PRECISION_FP64;
    o->warmStart = 0;
}

s4f_status s4f_solve(s4f_mesh* M, const s4f_solve_opts* opts, double* D_inout, double* sigma_out, double* gradD_out,
                     double* Db_out, s4f_solve_report* report) {
    return guarded([&]() -> s4f_status {
        if (!M || !opts) throw InvalidError("null argument");
        if (!M->materialSet) throw InvalidError("material not set");
        checkOpts(*opts);
        HIP_CHECK(hipSetDevice(M->ctx->device));
        std::vector<InstanceIO> io(1);
        io[0].D_inout = D_inout;
        io[0].sigma_out = sigma_out;
        io[0].gradD_out = gradD_out;
        io[0].Db_out = Db_out;
        io[0].rep = report;
        solveDispatch(M, *opts, io);
        if (io[0].status == S4F_ERR_DIVERGED) return fail(S4F_ERR_DIVERGED, "outer residual is not finite (diverged)");
        if (io[0].status == S4F_WARN_MAXITER) return fail(S4F_WARN_MAXITER, "Max iterations reached within momentum loop");
        return S4F_OK;
    });
}

s4f_status s4f_solve_history(s4f_mesh* M, int32_t cap, double* initRes, double* rvf, int32_t* inner, int32_t* count) {
    if (!M) return fail(S4F_ERR_INVALID, "null argument");
    const SolveState& st = M->st;
    const int n = std::min({cap, st.lastNOuter, st.histCap});
    for (int i = 0; i < n; ++i) {
        if (initRes) initRes[i] = st.hHist[i];
        if (rvf) rvf[i] = st.hHist[(size_t)st.histCap + i];
        if (inner) inner[i] = (int32_t)st.hHist[2 * (size_t)st.histCap + i];
    }
    if (count) *count = n;
    return S4F_OK;
}

s4f_status s4f_grad(s4f_mesh* M, const double* D, const double* Db, const double* snb, double* gradD, double* gradDb) {
    return guarded([&]() -> s4f_status {
        if (!M || !D || !Db || !snb || !gradD) throw InvalidError("null argument");
        if (!M->materialSet) throw InvalidError("material not set");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        hipStream_t s = M->ctx->stream;
        const int C = M->h.C, B = M->h.B;
        allocState(M);
        SolveState& st = M->st;
        std::vector<double> t;
        aosToSoa(D, C, 3, t);
        HIP_CHECK(hipMemcpyAsync(st.D0.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, s));
        if (B) {
            aosToSoa(Db, B, 3, t);
            HIP_CHECK(hipMemcpyAsync(st.Db.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, s));
            aosToSoa(snb, B, 3, t);
            HIP_CHECK(hipMemcpyAsync(st.snb.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, s));
        }
        Ctl c0;
        std::memset(&c0, 0, sizeof(c0));
        HIP_CHECK(hipMemcpyAsync(st.ctl.p, &c0, sizeof(Ctl), hipMemcpyHostToDevice, s));
        const DevMesh dm = M->dm();
        const BlockMap bm = M->bmap();
        const GradBufs gbuf = meshGradBufs(M);
        launchGradStep<double>(dm, bm, (int)M->h.instCell0.size(), M->maxInstNB, gbuf, nullptr, nullptr, st.ctl.p,
                               st.part.p, 2, s);
        HIP_CHECK(hipGetLastError());
        soaToAos(download(st.G, 9 * (size_t)C, s), C, 9, gradD);
        if (gradDb && B) soaToAos(download(st.Gb, 9 * (size_t)B, s), B, 9, gradDb);
        return S4F_OK;
    });
}

s4f_status s4f_sigma(s4f_mesh* M, const double* gradD, double* sigma) {
    return guarded([&]() -> s4f_status {
        if (!M || !gradD || !sigma) throw InvalidError("null argument");
        if (!M->materialSet) throw InvalidError("material not set");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        hipStream_t s = M->ctx->stream;
        const int C = M->h.C;
        allocState(M);
        SolveState& st = M->st;
        st.sig.alloc(6 * (size_t)C);
        std::vector<double> t;
        aosToSoa(gradD, C, 9, t);
        HIP_CHECK(hipMemcpyAsync(st.G.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_sigma, dim3(std::min(4096, (C + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, M->dm(), st.G.p,
                           st.sig.p, 0, C);
        HIP_CHECK(hipGetLastError());
        soaToAos(download(st.sig, 6 * (size_t)C, s), C, 6, sigma);
        return S4F_OK;
    });
}

s4f_status s4f_residual(s4f_mesh* M, const double* D, const double* gradD, const double* sigmab, double* R) {
    return guarded([&]() -> s4f_status {
        if (!M || !D || !gradD || !sigmab || !R) throw InvalidError("null argument");
        if (!M->materialSet) throw InvalidError("material not set");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        hipStream_t s = M->ctx->stream;
        const int C = M->h.C, B = M->h.B;
        allocState(M);
        SolveState& st = M->st;
        PcgVecs<double> v = pcgVecs<double>(st, C);
        std::vector<double> t;
        aosToSoa(D, C, 3, t);
        HIP_CHECK(hipMemcpyAsync(st.D0.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, s));
        aosToSoa(gradD, C, 9, t);
        HIP_CHECK(hipMemcpyAsync(st.G.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, s));
        if (B) {
            aosToSoa(sigmab, B, 6, t);
            HIP_CHECK(hipMemcpyAsync(st.sigb.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, s));
        }
        Ctl c0;
        std::memset(&c0, 0, sizeof(c0));
        HIP_CHECK(hipMemcpyAsync(st.ctl.p, &c0, sizeof(Ctl), hipMemcpyHostToDevice, s));
        const DevMesh dm = M->dm();
        const BlockMap bm = M->bmap();
        hipLaunchKernelGGL((k_residual<double>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const double*)st.D0.p,
                           (const double*)st.D1.p, st.G.p, st.sigb.p, v.r, st.ctl.p, st.part.p);
        HIP_CHECK(hipGetLastError());
        soaToAos(download(st.r64, 3 * (size_t)C, s), C, 3, R);
        return S4F_OK;
    });
}

s4f_status s4f_spmv3(s4f_mesh* M, const double* x, double* y, int32_t precision) {
    return guarded([&]() -> s4f_status {
        if (!M || !x || !y) throw InvalidError("null argument");
        if (!M->materialSet) throw InvalidError("material not set");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        hipStream_t s = M->ctx->stream;
        const int C = M->h.C;
        allocState(M);
        SolveState& st = M->st;
        const DevMesh dm = M->dm();
        const BlockMap bm = M->bmap();
        std::vector<double> t;
        aosToSoa(x, C, 3, t);
        if (precision == S4F_PRECISION_MIXED) {
            PcgVecs<float> v = pcgVecs<float>(st, C);
            std::vector<float> tf(t.begin(), t.end());
            HIP_CHECK(hipMemcpyAsync(v.p, tf.data(), tf.size() * 4, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL((k_spmvp<float, false>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const float*)nullptr,
                               (const float*)v.p, (float*)nullptr, v.q, st.ctl.p, st.part.p, 0, (const float*)nullptr);
            HIP_CHECK(hipGetLastError());
            std::vector<float> yf(3 * (size_t)C);
            HIP_CHECK(hipMemcpyAsync(yf.data(), v.q, yf.size() * 4, hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            std::vector<double> yd(yf.begin(), yf.end());
            soaToAos(yd, C, 3, y);
        } else {
            PcgVecs<double> v = pcgVecs<double>(st, C);
            HIP_CHECK(hipMemcpyAsync(v.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL((k_spmvp<double, false>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const double*)nullptr,
                               (const double*)v.p, (double*)nullptr, v.q, st.ctl.p, st.part.p, 0, (const double*)nullptr);
            HIP_CHECK(hipGetLastError());
            soaToAos(download(st.q64, 3 * (size_t)C, s), C, 3, y);
        }
        return S4F_OK;
    });
}

s4f_status s4f_pcg3(s4f_mesh* M, const s4f_solve_opts* o, const double* b, const double normFactor[3], double* x,
                    int32_t iters[3]) {
    return guarded([&]() -> s4f_status {
        if (!M || !o || !b || !normFactor || !x) throw InvalidError("null argument");
        if (!M->materialSet) throw InvalidError("material not set");
        checkOpts(*o);
        singleMeshOnly(*o, "s4f_pcg3");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        hipStream_t s = M->ctx->stream;
        const int C = M->h.C;
        allocState(M);
        SolveState& st = M->st;
        const DevMesh dm = M->dm();
        const BlockMap bm = M->bmap();
        std::vector<double> t;
        aosToSoa(b, C, 3, t);
        DArr<double> bd;
        bd.upload(t, s);
        Ctl c0 = makeCtl(*o, 0);
        for (int j = 0; j < 3; ++j) c0.normFactor[j] = normFactor[j];
        HIP_CHECK(hipMemcpyAsync(st.ctl.p, &c0, sizeof(Ctl), hipMemcpyHostToDevice, s));
        Ctl hc;
        auto run = [&](auto v, auto tag) {
            using T = decltype(tag);
            hipLaunchKernelGGL((k_pcg_start<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, bd.p, v.r, v.p, st.ctl.p,
                               st.part.p);
            hipLaunchKernelGGL(k_fin_pcg_start, dim3(1), dim3(kFinBlock), 0, s, bm, st.ctl.p, (const double*)st.part.p);
            HIP_CHECK(hipGetLastError());
            // x = 0 when no iteration runs
            HIP_CHECK(hipMemsetAsync(v.x, 0, 3 * (size_t)C * sizeof(T), s));
            int total = 0;
            for (;;) {
                launchPcgChunk<T>(dm, bm, 1, v, st.ctl.p, st.part.p, 8, s);
                total += 8;
                HIP_CHECK(hipMemcpyAsync(&hc, st.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipStreamSynchronize(s));
                if (hc.innerDone || total > o->pcgMaxIter + 8) break;
            }
            std::vector<T> xs(3 * (size_t)C);
            HIP_CHECK(hipMemcpyAsync(xs.data(), v.x, xs.size() * sizeof(T), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            std::vector<double> xd(xs.begin(), xs.end());
            soaToAos(xd, C, 3, x);
        };
        if (o->precision == S4F_PRECISION_MIXED) run(pcgVecs<float>(st, C), 0.0f);
        else run(pcgVecs<double>(st, C), 0.0);
        if (iters)
            for (int j = 0; j < 3; ++j) iters[j] = hc.nIter[j];
        return S4F_OK;
    });
}

// s4f_pcg3 without any allocation: b and x are the caller's device arrays
// (component-major, j*nCells + c), the control block, reduction partials and
// PCG vectors live in this device's scratch array.  Same kernels and
// workgroup tiles as s4f_pcg3, so the same bits.
s4f_status s4f_pcg3_noalloc(s4f_mesh* M, const s4f_solve_opts* o, const double* b, const double normFactor[3],
                            double* x, int32_t iters[3]) {
    return guarded([&]() -> s4f_status {
        if (!M || !o || !b || !normFactor || !x) throw InvalidError("null argument");
        if (!M->materialSet) throw InvalidError("material not set");
        checkOpts(*o);
        singleMeshOnly(*o, "s4f_pcg3_noalloc");
        const int C = M->h.C;
        if (C > kNoAllocMaxCells)
            throw InvalidError("s4f_pcg3_noalloc: " + std::to_string(C) + " cells exceed its scratch limit of " +
                               std::to_string(kNoAllocMaxCells) + " cells (use s4f_pcg3)");
        if ((size_t)M->nBlocks * kMaxPartials * sizeof(double) > kNoAllocPartBytes)
            throw InvalidError("s4f_pcg3_noalloc: too many workgroup tiles for its scratch");
        if (M->h.instCell0.size() != 1) throw InvalidError("s4f_pcg3_noalloc: single meshes only");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        hipStream_t s = M->ctx->stream;
        unsigned char* base = scratchPtr(M->ctx->device);
        Ctl* ctl = (Ctl*)base;
        double* part = (double*)(base + kNoAllocCtlBytes);
        unsigned char* vec = base + kNoAllocCtlBytes + kNoAllocPartBytes;
        const DevMesh dm = M->dm();
        const BlockMap bm = M->bmap();
        Ctl c0 = makeCtl(*o, 0);
        for (int j = 0; j < 3; ++j) c0.normFactor[j] = normFactor[j];
        hipLaunchKernelGGL(k_ctl_set, dim3(1), dim3(64), 0, s, ctl, c0);
        const size_t n3 = 3 * (size_t)C;
        Ctl hc;
        auto run = [&](auto tag, auto* xv) {
            using T = decltype(tag);
            PcgVecs<T> v;
            T* r = (T*)vec;
            v.r = r;
            v.p = r + n3;
            v.p2 = r + 2 * n3;
            v.q = r + 3 * n3;
            v.x = xv;
            hipLaunchKernelGGL((k_pcg_start<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, b, v.r, v.p, ctl, part);
            hipLaunchKernelGGL(k_fin_pcg_start, dim3(1), dim3(kFinBlock), 0, s, bm, ctl, (const double*)part);
            // x = 0 when no iteration runs
            hipLaunchKernelGGL((k_fill<T>), dim3(gridRows((int)n3)), dim3(kBlock), 0, s, v.x, n3, (T)0);
            HIP_CHECK(hipGetLastError());
            int total = 0;
            for (;;) {
                launchPcgChunk<T>(dm, bm, 1, v, ctl, part, 8, s);
                total += 8;
                HIP_CHECK(hipMemcpyAsync(&hc, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipStreamSynchronize(s));
                if (hc.innerDone || total > o->pcgMaxIter + 8) break;
            }
        };
        if (o->precision == S4F_PRECISION_MIXED) {
            float* x32 = (float*)vec + 4 * n3;
            run(0.0f, x32);
            hipLaunchKernelGGL(k_widen, dim3(gridRows((int)n3)), dim3(kBlock), 0, s, (const float*)x32, x, n3);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipStreamSynchronize(s));
        } else {
            run(0.0, x);
        }
        if (iters)
            for (int j = 0; j < 3; ++j) iters[j] = hc.nIter[j];
        return S4F_OK;
    });
}

s4f_status s4f_matrix(s4f_mesh* M, double* diag, double* faceCoeff) {
    if (!M) return fail(S4F_ERR_INVALID, "null argument");
    if (!M->materialSet) return fail(S4F_ERR_INVALID, "material not set");
    if (diag) std::memcpy(diag, M->h.diag.data(), M->h.diag.size() * sizeof(double));
    if (faceCoeff)
        for (int f = 0; f < M->h.F; ++f) faceCoeff[f] = M->h.eA[M->h.entOwn[f]];
    return S4F_OK;
}

s4f_status s4f_calibrate_copy(s4f_ctx* ctx, int64_t bytes, double* ms) {
    return guarded([&]() -> s4f_status {
        if (!ctx || bytes < 32) throw InvalidError("bad argument");
        HIP_CHECK(hipSetDevice(ctx->device));
        const size_t n = (size_t)bytes / 32;
        if ((n + kBlock - 1) / kBlock > (size_t)INT32_MAX) throw InvalidError("copy too large");
        DArr<double> a, b;
        a.alloc(4 * n);
        b.alloc(4 * n);
        HIP_CHECK(hipMemsetAsync(a.p, 0, a.bytes(), ctx->stream));
        const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
        hipLaunchKernelGGL(k_copy, grid, dim3(kBlock), 0, ctx->stream, (const dbl4*)a.p, (dbl4*)b.p, n);
        HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
        hipLaunchKernelGGL(k_copy, grid, dim3(kBlock), 0, ctx->stream, (const dbl4*)a.p, (dbl4*)b.p, n);
        HIP_CHECK(hipEventRecord(ctx->ev1, ctx->stream));
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipEventSynchronize(ctx->ev1));
        float t = 0.f;
        HIP_CHECK(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        if (ms) *ms = t;
        return S4F_OK;
    });
}

}  // extern "C"

// Back-to-back replay of one solver kernel on the mesh's current state (all
// three components active), timed by one event pair around `launches`
// launches on the library stream: the kernel's steady duration without the
// perturbation per-launch events cause.
template <typename T>
double replayT(s4f_mesh* M, int kernel, int launches) {
    hipStream_t s = M->ctx->stream;
    allocState(M);
    SolveState& st = M->st;
    PcgVecs<T> v = pcgVecs<T>(st, M->h.C);
    const DevMesh dm = M->dm();
    const BlockMap bm = M->bmap();
    Ctl c;
    std::memset(&c, 0, sizeof(c));
    for (int j = 0; j < 3; ++j) { c.active[j] = 1; c.alpha[j] = 0.5; c.beta[j] = 0.5; c.rho[j] = 1.0; c.normFactor[j] = 1.0; }
    DArr<Ctl> rc;
    rc.alloc(1);
    HIP_CHECK(hipMemcpyAsync(rc.p, &c, sizeof(Ctl), hipMemcpyHostToDevice, s));
    // the multigrid kernels replay the hierarchy of the mesh's last multigrid solve
    const bool smoothedK = M->amgLast == &M->amgSm;
    Amg& AK = smoothedK ? M->amgSm : M->amg;
    auto launch = [&]() {
        switch (kernel) {
            case S4F_KERNEL_RESIDUAL:
                hipLaunchKernelGGL((k_residual<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const double*)st.D0.p,
                                   (const double*)st.D1.p, st.G.p, st.sigb.p, v.r, rc.p, st.part.p);
                break;
            case S4F_KERNEL_SPMV:
                hipLaunchKernelGGL((k_spmvp<T, false>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const T*)v.r,
                                   (const T*)v.p, v.p2, v.q, rc.p, st.part.p, 1, (const T*)nullptr);
                std::swap(v.p, v.p2);
                break;
            case S4F_KERNEL_SPMV_MG:
                hipLaunchKernelGGL((k_spmvp<T, true>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const T*)v.r,
                                   (const T*)v.p, v.p2, v.q, rc.p, st.part.p, 1, (const T*)v.z);
                std::swap(v.p, v.p2);
                break;
            case S4F_KERNEL_PCG_UPDATE:
                hipLaunchKernelGGL((k_update<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, v.x, (const T*)v.p, v.r,
                                   (const T*)v.q, (const Ctl*)rc.p, st.part.p);
                break;
            case S4F_KERNEL_FINALIZE:
                hipLaunchKernelGGL(k_fin_spmv, dim3(1), dim3(kFinBlock), 0, s, bm, rc.p, (const double*)st.part.p);
                break;
            case S4F_KERNEL_GRAD:
                launchGradStep<T>(dm, bm, 1, M->maxInstNB,
                                  meshGradBufs(M),
                                  (const T*)v.x, nullptr, rc.p, st.part.p, 0, s);
                break;
            case S4F_KERNEL_MG0_DOWN: {
                AmgDevLevel& D0 = *AK.d[0];
                hipLaunchKernelGGL((k_mg0_down<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const T*)v.r,
                                   MgVec<T>::r0(D0), AK.omega, (const Ctl*)rc.p);
                break;
            }
            case S4F_KERNEL_MG0_UP:
                launchMg0Up<T>(AK.d, dm, bm, v, AK.omega, AK.alpha, rc.p, st.part.p, s);
                break;
            case S4F_KERNEL_VCYCLE:
                launchVcycle<T>(meshRun(M, smoothedK), dm, bm, v, rc.p, st.part.p, s);
                break;
            default:
                throw InvalidError("unknown kernel id");
        }
    };
    if (kernel >= S4F_KERNEL_MG0_DOWN && kernel != S4F_KERNEL_SPMV_MG) {
        buildAmg(M, smoothedK);
        for (size_t l = 0; l < AK.d.size(); ++l) {  // defined inputs on every level
            AmgDevLevel& L = *AK.d[l];
            if (l == 0) {  // the fine level keeps only the residual
                HIP_CHECK(hipMemsetAsync(MgVec<T>::r0(L), 0, 3 * (size_t)L.n * sizeof(T), s));
                continue;
            }
            const size_t n4 = 4 * (size_t)L.n * sizeof(T);
            HIP_CHECK(hipMemsetAsync(MgVec<T>::x(L), 0, n4, s));
            HIP_CHECK(hipMemsetAsync(MgVec<T>::xs(L), 0, n4, s));
            HIP_CHECK(hipMemsetAsync(MgVec<T>::b(L), 0, n4, s));
            HIP_CHECK(hipMemsetAsync(MgVec<T>::r(L), 0, n4, s));
        }
    }
    launch();  // warm
    HIP_CHECK(hipEventRecord(M->ctx->ev0, s));
    for (int i = 0; i < launches; ++i) launch();
    HIP_CHECK(hipEventRecord(M->ctx->ev1, s));
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventSynchronize(M->ctx->ev1));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, M->ctx->ev0, M->ctx->ev1));
    return ms / launches;
}

extern "C" {

s4f_status s4f_kernel_time(s4f_mesh* M, int32_t kernel, int32_t precision, int32_t launches, double* avgMs) {
    return guarded([&]() -> s4f_status {
        if (!M || !avgMs || launches < 1) throw InvalidError("bad argument");
        if (!M->materialSet) throw InvalidError("material not set");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        *avgMs = precision == S4F_PRECISION_MIXED ? replayT<float>(M, kernel, launches)
                                                  : replayT<double>(M, kernel, launches);
        return S4F_OK;
    });
}

s4f_status s4f_strain_fields(s4f_mesh* M, double* epsilon, double* epsilonEq, double* sigmaEq) {
    return guarded([&]() -> s4f_status {
        if (!M) throw InvalidError("null argument");
        if (M->st.lastDpar < 0) throw InvalidError("no converged field on the device: call s4f_solve first");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        hipStream_t s = M->ctx->stream;
        const int C = M->h.C;
        DArr<double> e, ee, se;
        if (epsilon) e.alloc(6 * (size_t)C);
        if (epsilonEq) ee.alloc(C);
        if (sigmaEq) se.alloc(C);
        hipLaunchKernelGGL(k_strain_fields, dim3(std::max(1, std::min(4096, (C + kBlock - 1) / kBlock))), dim3(kBlock),
                           0, s, M->dm(), (const double*)M->st.G.p, e.p, ee.p, se.p);
        HIP_CHECK(hipGetLastError());
        if (epsilon) {
            std::vector<double> h = download(e, 6 * (size_t)C, s);
            for (int c = 0; c < C; ++c)
                for (int q = 0; q < 6; ++q) epsilon[6 * (size_t)c + q] = h[(size_t)q * C + c];
        }
        if (epsilonEq) {
            std::vector<double> h = download(ee, C, s);
            std::memcpy(epsilonEq, h.data(), C * sizeof(double));
        }
        if (sigmaEq) {
            std::vector<double> h = download(se, C, s);
            std::memcpy(sigmaEq, h.data(), C * sizeof(double));
        }
        return S4F_OK;
    });
}

s4f_status s4f_point_displacement(s4f_mesh* M, int32_t nPoints, const double* points, const int32_t* pointCellOffsets,
                                  const int32_t* pointCells, double* pointD) {
    return guarded([&]() -> s4f_status {
        if (!M || !points || !pointCellOffsets || !pointCells || !pointD || nPoints < 0)
            throw InvalidError("null argument");
        if (M->st.lastDpar < 0) throw InvalidError("no converged field on the device: call s4f_solve first");
        const int C = M->h.C;
        const int nnz = pointCellOffsets[nPoints];
        if (pointCellOffsets[0] != 0 || nnz < 0) throw InvalidError("bad point-cell offsets");
        for (int p = 0; p < nPoints; ++p)
            if (pointCellOffsets[p + 1] < pointCellOffsets[p]) throw InvalidError("bad point-cell offsets");
        for (int k = 0; k < nnz; ++k)
            if (pointCells[k] < 0 || pointCells[k] >= C) throw InvalidError("point-cell index out of range");
        HIP_CHECK(hipSetDevice(M->ctx->device));
        hipStream_t s = M->ctx->stream;
        DArr<double> pts, out;
        DArr<int> off, cells;
        pts.upload(std::vector<double>(points, points + 3 * (size_t)nPoints), s);
        off.upload(std::vector<int>(pointCellOffsets, pointCellOffsets + nPoints + 1), s);
        cells.upload(std::vector<int>(pointCells, pointCells + nnz), s);
        out.alloc(3 * (size_t)std::max(nPoints, 1));
        const double* D = M->st.lastDpar ? M->st.D1.p : M->st.D0.p;
        hipLaunchKernelGGL(k_point_interpolate, dim3(std::max(1, std::min(4096, (nPoints + kBlock - 1) / kBlock))),
                           dim3(kBlock), 0, s, M->dm(), D, (const double*)M->st.G.p, nPoints, (const double*)pts.p,
                           (const int*)off.p, (const int*)cells.p, out.p);
        HIP_CHECK(hipGetLastError());
        std::vector<double> h = download(out, 3 * (size_t)nPoints, s);
        std::memcpy(pointD, h.data(), 3 * (size_t)nPoints * sizeof(double));
        return S4F_OK;
    });
}

s4f_status s4f_polymesh_point_cells(const s4f_polymesh* p, int32_t* offsets, int32_t* cells) {
    if (!p || !offsets) return fail(S4F_ERR_INVALID, "null argument");
    const PolyMesh& m = *p->pm;
    // cells of each point through the faces of each cell, ascending cell order
    std::vector<std::vector<int>> pc(m.nPoints);
    for (int f = 0; f < m.nFaces; ++f)
        for (int k = m.fOff[f]; k < m.fOff[f + 1]; ++k) {
            const int pt = m.fVerts[k];
            pc[pt].push_back(m.owner[f]);
            if (f < m.nInternalFaces) pc[pt].push_back(m.neighbour[f]);
        }
    offsets[0] = 0;
    for (int pt = 0; pt < m.nPoints; ++pt) {
        std::vector<int>& v = pc[pt];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        offsets[pt + 1] = offsets[pt] + (int)v.size();
        if (cells) std::copy(v.begin(), v.end(), cells + offsets[pt]);
    }
    return S4F_OK;
}

struct s4f_batch {
    s4f_ctx* ctx = nullptr;
    std::vector<s4f_mesh*> members;   // borrowed; must outlive the solve
    std::unique_ptr<s4f_mesh> packed; // rebuilt when membership, a member's material or the scales change
    std::vector<unsigned> packedGen, packedTileGen;  // member generations the packing was built from
    double packedRc = 0.0, packedIf = 0.0;
    bool dirty = true;
};

s4f_status s4f_batch_create(s4f_ctx* ctx, s4f_batch** out) {
    return guarded([&]() -> s4f_status {
        if (!ctx || !out) throw InvalidError("null argument");
        if (ctx->destroyPending) throw InvalidError("context was destroyed");
        auto b = std::make_unique<s4f_batch>();
        b->ctx = ctx;
        ctx->children++;
        *out = b.release();
        return S4F_OK;
    });
}

s4f_status s4f_batch_add(s4f_batch* b, s4f_mesh* m, int32_t* index) {
    return guarded([&]() -> s4f_status {
        if (!b || !m) throw InvalidError("null argument");
        if (!m->materialSet) throw InvalidError("material not set");
        if (m->ctx != b->ctx) throw InvalidError("mesh belongs to another context");
        if (m->h.instCell0.size() != 1) throw InvalidError("cannot nest batches");
        if (!m->h.zone.empty()) throw InvalidError("a mesh with material zones (sub-meshes) is solved alone");
        if (m->h.g[0] != 0.0 || m->h.g[1] != 0.0 || m->h.g[2] != 0.0)
            throw InvalidError("batched solves are steady solves without body force (gravity is set on this mesh)");
        b->members.push_back(m);
        b->dirty = true;
        if (index) *index = (int32_t)b->members.size() - 1;
        return S4F_OK;
    });
}

s4f_status s4f_batch_solve(s4f_batch* b, const s4f_solve_opts* opts, double** D_inout, double** sigma_out,
                           s4f_solve_report* reports) {
    return guarded([&]() -> s4f_status {
        if (!b || !opts) throw InvalidError("null argument");
        if (b->members.empty()) throw InvalidError("empty batch");
        checkOpts(*opts);
        if (opts->d2dt2 != S4F_D2DT2_STEADY) throw InvalidError("batched solves are steady solves");
        HIP_CHECK(hipSetDevice(b->ctx->device));
        bool stale = b->dirty || !b->packed || opts->rhieChowScale != b->packedRc || opts->interfaceScale != b->packedIf;
        for (size_t i = 0; !stale && i < b->members.size(); ++i)
            stale = b->members[i]->generation != b->packedGen[i] || b->members[i]->tileGen != b->packedTileGen[i];
        if (stale) {
            std::vector<const HostMesh*> hs;
            for (s4f_mesh* m : b->members) {
                if (m->h.g[0] != 0.0 || m->h.g[1] != 0.0 || m->h.g[2] != 0.0)
                    throw InvalidError("batched solves are steady solves without body force (gravity is set on a member)");
                hs.push_back(&m->h);
            }
            auto P = std::make_unique<s4f_mesh>();
            P->ctx = b->ctx;
            packMeshes(hs, P->h, opts->rhieChowScale, opts->interfaceScale);
            uploadGeometry(P.get());
            uploadMaterial(P.get());
            P->rcScale = opts->rhieChowScale;
            P->ifScale = opts->interfaceScale;
            P->materialSet = true;
            b->packed = std::move(P);
            b->packedGen.clear();
            b->packedTileGen.clear();
            for (s4f_mesh* m : b->members) b->packedGen.push_back(m->generation);
            for (s4f_mesh* m : b->members) b->packedTileGen.push_back(m->tileGen);
            b->packedRc = opts->rhieChowScale;
            b->packedIf = opts->interfaceScale;
            b->dirty = false;
        } else {
            // refresh boundary values (loads may change between solves)
            HostMesh& P = b->packed->h;
            for (size_t i = 0; i < b->members.size(); ++i) {
                const HostMesh& m = b->members[i]->h;
                for (int x = 0; x < m.B; ++x) {
                    P.bPres[P.instB0[i] + x] = m.bPres[x];
                    for (int q = 0; q < 3; ++q) P.bVal[(size_t)q * P.B + P.instB0[i] + x] = m.bVal[(size_t)q * m.B + x];
                }
            }
            b->packed->bVal.upload(P.bVal, b->ctx->stream);
            b->packed->bPres.upload(P.bPres, b->ctx->stream);
        }
        const int n = (int)b->members.size();
        std::vector<InstanceIO> io(n);
        for (int i = 0; i < n; ++i) {
            io[i].D_inout = D_inout ? D_inout[i] : nullptr;
            io[i].sigma_out = sigma_out ? sigma_out[i] : nullptr;
            io[i].rep = reports ? reports + i : nullptr;
        }
        solveDispatch(b->packed.get(), *opts, io);
        s4f_status worst = S4F_OK;
        for (const InstanceIO& x : io) {
            if (x.status == S4F_ERR_DIVERGED) worst = S4F_ERR_DIVERGED;
            else if (x.status == S4F_WARN_MAXITER && worst == S4F_OK) worst = S4F_WARN_MAXITER;
        }
        if (worst != S4F_OK) return fail(worst, worst == S4F_ERR_DIVERGED ? "an instance diverged"
                                                                          : "an instance reached nCorrectors");
        return S4F_OK;
    });
}

s4f_status s4f_batch_destroy(s4f_batch* b) {
    if (!b) return S4F_OK;
    s4f_ctx* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete b;
    ctxChildGone(ctx);
    return S4F_OK;
}

}  // extern "C"

// ===========================================================================
// Serving: continuous batching over an HBM arena
// ===========================================================================
namespace {

double nowSeconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, cap) (units: cells, entries or faces).
struct RangeAlloc {
    std::map<long long, long long> freeMap;  // start -> length
    long long cap = 0, used = 0;
    void init(long long c) { cap = c; used = 0; freeMap.clear(); if (c > 0) freeMap[0] = c; }
    long long alloc(long long n, long long align) {
        if (n <= 0) return 0;
        for (auto it = freeMap.begin(); it != freeMap.end(); ++it) {
            const long long st = (it->first + align - 1) / align * align;
            const long long end = it->first + it->second;
            if (st + n > end) continue;
            const long long fs = it->first, fl = it->second;
            freeMap.erase(it);
            if (st > fs) freeMap[fs] = st - fs;
            if (st + n < fs + fl) freeMap[st + n] = fs + fl - (st + n);
            used += n;
            return st;
        }
        return -1;
    }
    void release(long long st, long long n) {
        if (n <= 0) return;
        used -= n;
        auto it = freeMap.emplace(st, n).first;
        auto nx = std::next(it);
        if (nx != freeMap.end() && it->first + it->second == nx->first) { it->second += nx->second; freeMap.erase(nx); }
        if (it != freeMap.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) { pv->second += it->second; freeMap.erase(it); }
        }
    }
};

struct Placement {
    const s4f_mesh* mesh = nullptr;
    unsigned gen = 0;
    long long c0 = -1, nCellsPad = 0, e0 = -1, nE = 0, b0 = -1, nB = 0;
    int nCells = 0, slice0 = 0, nSlices = 0;
    int dpar = 0;          // buffer holding the resident field
    bool hasField = false; // a converged field is resident (warm start possible)
    double lastUse = 0;
};

struct Request {
    long long id = 0;
    s4f_mesh* mesh = nullptr;
    std::vector<double> bVal, bPres;  // SoA 3B, B (the mesh's boundary-face order)
    int warm = 0;
    double tSubmit = 0, tStart = 0, tFinish = 0;
    int slot = -1;
    int place = -1;
};

struct Result {
    std::vector<double> D, sigma;  // AoS
    s4f_solve_report rep{};
    double tSubmit = 0, tStart = 0, tFinish = 0;
    s4f_status status = S4F_OK;
};

}  // namespace

// host -> device copy on the server's own stream, complete on return (no
// legacy-stream copy: another server's thread may be capturing a graph)
template <typename T>
void hcopy(hipStream_t s, T* dst, const T* src, size_t n) {
    if (!n) return;
    HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
}

struct s4f_server {
    s4f_ctx* ctx = nullptr;
    s4f_solve_opts opts{};
    int maxSlots = 0;
    long long budget = 0;
    long long Ccap = 0, Ecap = 0, Bcap = 0;
    RangeAlloc cells, ents, bnds;
    // arena: device geometry and state, laid out like one packed s4f_mesh
    DArr<int> sliceStart, sliceWidth, col, bCell, bType;
    DArr<double> eS, eW, eA, eC, cC, minv, diag, invDiag, mu, K, impK, bN, bKv, bDelta, bLs, bVal, bPres;
    DArr<float> eA32, diag32, invDiag32;
    DArr<double> D0, D1, Dold, G, Db, g, snb, Gb, sigb, sig;
    DArr<double> aAlpha, aRes, aAlphaB, aResB;  // Aitken relaxation state
    DArr<double> r64, p64, pb64, q64, x64, z64;
    DArr<float> r32, p32, pb32, q32, x32, z32;
    DArr<Ctl> ctl;
    // GAMG: packed block-diagonal hierarchy of the in-flight requests (level 0
    // is the arena: agg per arena cell; levels >= 1 packed per request, rebuilt
    // on the device from each mesh's own hierarchy whenever the set changes)
    bool mgOn = false;
    std::vector<std::unique_ptr<AmgDevLevel>> mg;
    DArr<SegInt> segI;
    DArr<SegWords> segW;
    DArr<double> part;
    DArr<int> bmInst, bmS0, bmS1, iBlk0, iNBlk, iCell0, iNCell, iB0, iNB, finList;
    DArr<int> nbmInst, nbmS0, nbmS1, nfinList;  // init map of newly admitted instances
    std::vector<int> hiBlk0, hiNBlk, hiCell0, hiNCell, hiB0, hiNB;
    int nTiles = 0, nActive = 0, chunk = 4;
    // launch sequences of a step replayed from HIP graphs while the packed
    // map they were captured on stands (mapGen counts rebuilds)
    GraphExec gStart, gPair, gEnd;
    unsigned mapGen = 0, capGen = ~0u, ranGen = ~0u;
    Ctl* hctl = nullptr;
    // bookkeeping (server thread only)
    std::vector<Placement> places;
    std::vector<int> slotReq;  // request index per slot (-1 free)
    std::vector<Request> live;
    // shared with clients
    std::mutex mu_;
    std::condition_variable cvIn, cvOut;
    std::deque<Request> incoming;
    std::map<long long, Result> results;
    std::deque<long long> doneOrder;
    long long nextId = 1;
    bool stop = false;
    std::string threadError;
    std::thread th;
    std::atomic<long long> nQueued{0}, nDone{0};
    std::atomic<long long> bytesInUse{0};
};

namespace {

constexpr double kCellBytes = 24 + 48 + 8 + 8 + 4 + 4 + 24 + 72 + 72 + 48 + 5 * 3 * 8;  // geometry + state + PCG (FP64)
constexpr double kEntryBytes = 4 + 24 + 8 + 8 + 8 + 4;
constexpr double kBndBytes = 4 + 4 + 24 + 24 + 8 + 24 + 24 + 8 + 24 + 24 + 24 + 72 + 48;

long long footprintBytes(const HostMesh& h) {
    const long long C = (long long)h.nSlices * kSlice;
    return (long long)(C * kCellBytes + (double)h.col.size() * kEntryBytes + h.B * kBndBytes);
}

void serverAllocate(s4f_server* S) {
    // split the budget over cells, ELL entries (7 per cell) and boundary faces (1 per cell)
    const bool aitken = S->opts.relaxationMethod == S4F_RELAX_AITKEN;
    // multigrid: the level-0 grouped operator q (column + value per entry)
    const double qBytes = S->opts.preconditioner != S4F_PRECOND_GAMG ? 0.0
                          : S->opts.precision == S4F_PRECISION_MIXED ? 8.0 : 12.0;
    const double perCell = kCellBytes + 7.0 * (kEntryBytes + qBytes) + 1.0 * kBndBytes + (aitken ? 64.0 : 0.0);
    long long C = (long long)(S->budget / perCell);
    C = std::max<long long>(C / kSlice * kSlice, kSlice);
    if (C > INT32_MAX / 16) C = INT32_MAX / 16 / kSlice * kSlice;
    S->Ccap = C;
    S->Ecap = 7 * C;
    S->Bcap = C;
    S->cells.init(S->Ccap);
    S->ents.init(S->Ecap);
    S->bnds.init(S->Bcap);
    const size_t Cc = (size_t)S->Ccap, Ec = (size_t)S->Ecap, Bc = (size_t)S->Bcap;
    const size_t nS = Cc / kSlice;
    S->sliceStart.alloc(nS + 1); S->sliceWidth.alloc(nS); S->col.alloc(Ec);
    S->eS.alloc(3 * Ec); S->eW.alloc(Ec); S->eA.alloc(Ec); S->eC.alloc(Ec); S->eA32.alloc(Ec);
    S->cC.alloc(3 * Cc); S->minv.alloc(6 * Cc); S->diag.alloc(Cc); S->invDiag.alloc(Cc);
    S->diag32.alloc(Cc); S->invDiag32.alloc(Cc); S->mu.alloc(Cc); S->K.alloc(Cc); S->impK.alloc(Cc);
    S->bCell.alloc(Bc); S->bType.alloc(Bc); S->bN.alloc(3 * Bc); S->bKv.alloc(3 * Bc); S->bDelta.alloc(Bc);
    S->bLs.alloc(3 * Bc); S->bVal.alloc(3 * Bc); S->bPres.alloc(Bc);
    S->D0.alloc(3 * Cc); S->D1.alloc(3 * Cc); S->Dold.alloc(3 * Cc); S->G.alloc(9 * Cc); S->sig.alloc(6 * Cc);
    S->Db.alloc(3 * Bc); S->g.alloc(3 * Bc); S->snb.alloc(3 * Bc); S->Gb.alloc(9 * Bc); S->sigb.alloc(6 * Bc);
    if (aitken) {
        S->aAlpha.alloc(Cc); S->aRes.alloc(3 * Cc); S->aAlphaB.alloc(Bc); S->aResB.alloc(3 * Bc);
    }
    S->mgOn = S->opts.preconditioner == S4F_PRECOND_GAMG;
    if (S->mgOn) {
        S->mg.clear();
        S->mg.push_back(std::make_unique<AmgDevLevel>());
        S->mg[0]->n = (int)Cc;
        S->mg[0]->agg.alloc(Cc);
        S->mg[0]->qCol.alloc(Ec);
        if (S->opts.precision == S4F_PRECISION_MIXED) S->mg[0]->q32.alloc(Ec);
        else S->mg[0]->q64.alloc(Ec);
    }
    if (S->opts.precision == S4F_PRECISION_MIXED) {
        S->r32.alloc(3 * Cc); S->p32.alloc(3 * Cc); S->pb32.alloc(3 * Cc); S->q32.alloc(3 * Cc); S->x32.alloc(3 * Cc);
        if (S->mgOn) S->z32.alloc(3 * Cc);
    } else {
        S->r64.alloc(3 * Cc); S->p64.alloc(3 * Cc); S->pb64.alloc(3 * Cc); S->q64.alloc(3 * Cc); S->x64.alloc(3 * Cc);
        if (S->mgOn) S->z64.alloc(3 * Cc);
    }
    S->ctl.alloc(S->maxSlots);
    std::vector<Ctl> zero(S->maxSlots);
    std::memset(zero.data(), 0, zero.size() * sizeof(Ctl));
    for (Ctl& c : zero) c.outerDone = 1;
    hcopy(S->ctx->stream, S->ctl.p, zero.data(), zero.size());
    HIP_CHECK(hipHostMalloc((void**)&S->hctl, S->maxSlots * sizeof(Ctl), hipHostMallocDefault));
    S->hiBlk0.assign(S->maxSlots, 0); S->hiNBlk.assign(S->maxSlots, 0);
    S->hiCell0.assign(S->maxSlots, 0); S->hiNCell.assign(S->maxSlots, 0);
    S->hiB0.assign(S->maxSlots, 0); S->hiNB.assign(S->maxSlots, 0);
    S->iBlk0.alloc(S->maxSlots); S->iNBlk.alloc(S->maxSlots); S->iCell0.alloc(S->maxSlots); S->iNCell.alloc(S->maxSlots);
    S->iB0.alloc(S->maxSlots); S->iNB.alloc(S->maxSlots);
    S->slotReq.assign(S->maxSlots, -1);
}

DevMesh serverDm(const s4f_server* S) {
    DevMesh m{};
    m.nCells = (int)S->Ccap;
    m.nBoundary = (int)S->Bcap;
    m.nSlices = (int)(S->Ccap / kSlice);
    m.nEntries = (size_t)S->Ecap;
    m.sliceStart = S->sliceStart.p; m.sliceWidth = S->sliceWidth.p; m.col = S->col.p;
    m.eS = S->eS.p; m.eW = S->eW.p; m.eA = S->eA.p; m.eC = S->eC.p; m.eA32 = S->eA32.p;
    m.cC = S->cC.p; m.minv = S->minv.p; m.diag = S->diag.p; m.invDiag = S->invDiag.p;
    m.diag32 = S->diag32.p; m.invDiag32 = S->invDiag32.p;
    m.mu = S->mu.p; m.K = S->K.p; m.impK = S->impK.p;
    m.bCell = S->bCell.p; m.bType = S->bType.p; m.bN = S->bN.p; m.bKv = S->bKv.p; m.bDelta = S->bDelta.p;
    m.bLs = S->bLs.p; m.bVal = S->bVal.p; m.bPres = S->bPres.p;
    m.nInst = S->maxSlots;
    return m;
}

// writes one mesh's geometry/material into its arena ranges (offsets applied)
void placeMesh(s4f_server* S, const HostMesh& h, const Placement& pl) {
    const size_t Cc = (size_t)S->Ccap, Ec = (size_t)S->Ecap, Bc = (size_t)S->Bcap;
    const size_t E = h.col.size();
    const int c0 = (int)pl.c0, e0 = (int)pl.e0, b0 = (int)pl.b0;
    std::vector<int> ss(h.nSlices), col(E);
    for (int s = 0; s < h.nSlices; ++s) ss[s] = e0 + h.sliceStart[s];
    hcopy(S->ctx->stream, S->sliceStart.p + pl.slice0, ss.data(), ss.size());
    hcopy(S->ctx->stream, S->sliceWidth.p + pl.slice0, h.sliceWidth.data(), h.sliceWidth.size());
    for (size_t e = 0; e < E; ++e) {
        const int c = h.col[e];
        col[e] = c >= 0 ? c + c0 : (c == kPad ? kPad : -1 - ((-1 - c) + b0));
    }
    hcopy(S->ctx->stream, S->col.p + e0, col.data(), E);
    for (int q = 0; q < 3; ++q) hcopy(S->ctx->stream, S->eS.p + (size_t)q * Ec + e0, h.eS.data() + (size_t)q * E, E);
    hcopy(S->ctx->stream, S->eW.p + e0, h.eW.data(), E);
    hcopy(S->ctx->stream, S->eA.p + e0, h.eA.data(), E);
    // Rhie-Chow coefficients of the server's stabilisation scales
    if (S->opts.rhieChowScale == kRhieChowScaleDefault && S->opts.interfaceScale == kInterfaceScaleDefault) {
        hcopy(S->ctx->stream, S->eC.p + e0, h.eC.data(), E);
    } else {
        const std::vector<double> eC = stabilisationCoeffs(h, S->opts.rhieChowScale, S->opts.interfaceScale);
        hcopy(S->ctx->stream, S->eC.p + e0, eC.data(), E);
    }
    std::vector<float> a32(E);
    for (size_t e = 0; e < E; ++e) a32[e] = (float)h.eA[e];
    hcopy(S->ctx->stream, S->eA32.p + e0, a32.data(), E);
    const int C = h.C, Cp = (int)pl.nCellsPad;
    // cell arrays (padding rows: unit diagonal, zero geometry)
    std::vector<double> t(Cp);
    auto putCell = [&](double* base, const std::vector<double>& src, int k, int strideSrc, double pad) {
        for (int q = 0; q < k; ++q) {
            for (int c = 0; c < Cp; ++c) t[c] = c < C ? src[(size_t)q * strideSrc + c] : pad;
            hcopy(S->ctx->stream, base + (size_t)q * Cc + c0, t.data(), (size_t)Cp);
        }
    };
    putCell(S->cC.p, h.cC, 3, C, 0.0);
    // the whole-mesh LS inverses (zones set after submission do not reach a request)
    putCell(S->minv.p, h.minv0.empty() ? h.minv : h.minv0, 6, C, 0.0);
    putCell(S->diag.p, h.diagBase, 1, C, 1.0);
    std::vector<double> inv(C), muv(C), Kv(C), ikv(C);
    std::vector<float> d32(Cp, 1.0f), i32(Cp, 1.0f);
    for (int c = 0; c < C; ++c) {
        inv[c] = 1.0 / h.diagBase[c];
        d32[c] = (float)h.diagBase[c];
        i32[c] = 1.0f / d32[c];
        const int mc = h.uniformMaterial ? 0 : c;
        muv[c] = h.mu[mc]; Kv[c] = h.K[mc]; ikv[c] = h.impK[mc];
    }
    putCell(S->invDiag.p, inv, 1, C, 1.0);
    putCell(S->mu.p, muv, 1, C, 1.0);
    putCell(S->K.p, Kv, 1, C, 1.0);
    putCell(S->impK.p, ikv, 1, C, 1.0);
    hcopy(S->ctx->stream, S->diag32.p + c0, d32.data(), (size_t)Cp);
    hcopy(S->ctx->stream, S->invDiag32.p + c0, i32.data(), (size_t)Cp);
    const int B = h.B;
    std::vector<int> bc(B);
    for (int b = 0; b < B; ++b) bc[b] = h.bCell[b] + c0;
    hcopy(S->ctx->stream, S->bCell.p + b0, bc.data(), (size_t)B);
    hcopy(S->ctx->stream, S->bType.p + b0, h.bType.data(), (size_t)B);
    hcopy(S->ctx->stream, S->bDelta.p + b0, h.bDelta.data(), (size_t)B);
    for (int q = 0; q < 3; ++q) {
        hcopy(S->ctx->stream, S->bN.p + (size_t)q * Bc + b0, h.bN.data() + (size_t)q * B, (size_t)B);
        hcopy(S->ctx->stream, S->bKv.p + (size_t)q * Bc + b0, h.bKv.data() + (size_t)q * B, (size_t)B);
        hcopy(S->ctx->stream, S->bLs.p + (size_t)q * Bc + b0, h.bLs.data() + (size_t)q * B, (size_t)B);
    }
}

void placeLoads(s4f_server* S, const Request& rq, const Placement& pl) {
    const size_t Bc = (size_t)S->Bcap;
    const int B = (int)pl.nB;
    for (int q = 0; q < 3; ++q) hcopy(S->ctx->stream, S->bVal.p + (size_t)q * Bc + pl.b0, rq.bVal.data() + (size_t)q * B, (size_t)B);
    hcopy(S->ctx->stream, S->bPres.p + pl.b0, rq.bPres.data(), (size_t)B);
}

void releasePlacement(s4f_server* S, int pi) {
    Placement& pl = S->places[pi];
    S->cells.release(pl.c0, pl.nCellsPad);
    S->ents.release(pl.e0, pl.nE);
    S->bnds.release(pl.b0, pl.nB);
    S->bytesInUse -= footprintBytes(pl.mesh->h);
    pl.mesh = nullptr;
}

bool placementBusy(const s4f_server* S, int pi) {
    for (const Request& r : S->live)
        if (r.place == pi) return true;
    return false;
}

// finds or creates the placement of a request's mesh; -1 when it does not fit
int acquirePlacement(s4f_server* S, const Request& rq) {
    const HostMesh& h = rq.mesh->h;
    for (int i = 0; i < (int)S->places.size(); ++i) {
        Placement& p = S->places[i];
        if (p.mesh == rq.mesh && p.gen == rq.mesh->generation && !placementBusy(S, i)) return i;
    }
    Placement pl;
    pl.mesh = rq.mesh;
    pl.gen = rq.mesh->generation;
    pl.nCells = h.C;
    pl.nSlices = h.nSlices;
    pl.nCellsPad = (long long)h.nSlices * kSlice;
    pl.nE = (long long)h.col.size();
    pl.nB = h.B;
    for (;;) {
        const long long c0 = S->cells.alloc(pl.nCellsPad, kSlice);
        const long long e0 = c0 >= 0 ? S->ents.alloc(pl.nE, kSlice) : -1;
        const long long b0 = e0 >= 0 ? S->bnds.alloc(pl.nB, 1) : -1;
        if (b0 >= 0) {
            pl.c0 = c0; pl.e0 = e0; pl.b0 = b0;
            pl.slice0 = (int)(c0 / kSlice);
            break;
        }
        if (e0 >= 0) S->ents.release(e0, pl.nE);
        if (c0 >= 0) S->cells.release(c0, pl.nCellsPad);
        // evict the least recently used resident mesh that no solve is using
        int victim = -1;
        for (int i = 0; i < (int)S->places.size(); ++i) {
            const Placement& p = S->places[i];
            if (!p.mesh || placementBusy(S, i)) continue;
            if (victim < 0 || p.lastUse < S->places[victim].lastUse) victim = i;
        }
        if (victim < 0) return -1;
        releasePlacement(S, victim);
    }
    placeMesh(S, h, pl);
    S->bytesInUse += footprintBytes(h);
    for (int i = 0; i < (int)S->places.size(); ++i)
        if (!S->places[i].mesh) { S->places[i] = pl; return i; }
    S->places.push_back(pl);
    return (int)S->places.size() - 1;
}

// tile list of the given slots (the same per-instance rule as a plain solve)
void buildTiles(const s4f_server* S, const std::vector<int>& slots, std::vector<int>& inst, std::vector<int>& s0,
                std::vector<int>& s1, std::vector<int>& blk0, std::vector<int>& nblk) {
    inst.clear(); s0.clear(); s1.clear();
    for (int slot : slots) {
        const Request& rq = S->live[S->slotReq[slot]];
        const Placement& pl = S->places[rq.place];
        const int a = pl.slice0, nS = pl.nSlices, e = a + nS;
        const int spbI = tileSlices(rq.mesh->h, 0, nS);
        blk0[slot] = (int)inst.size();
        for (int x = a; x < e; x += spbI) {
            inst.push_back(slot);
            s0.push_back(x);
            s1.push_back(std::min(x + spbI, e));
        }
        nblk[slot] = (int)inst.size() - blk0[slot];
    }
}

template <typename T>
PcgVecs<T> serverVecs(s4f_server* S);
template <>
PcgVecs<double> serverVecs<double>(s4f_server* S) {
    return {S->r64.p, S->p64.p, S->q64.p, S->x64.p, S->pb64.p, S->z64.p};
}
template <>
PcgVecs<float> serverVecs<float>(s4f_server* S) {
    return {S->r32.p, S->p32.p, S->q32.p, S->x32.p, S->pb32.p, S->z32.p};
}

GradBufs serverGradBufs(s4f_server* S) {
    GradBufs f{S->D0.p, S->D1.p, S->G.p, S->Db.p, S->g.p, S->snb.p, S->Gb.p, S->sigb.p};
    f.aAlpha = S->aAlpha.p; f.aRes = S->aRes.p; f.aAlphaB = S->aAlphaB.p; f.aResB = S->aResB.p;  // null unless Aitken
    return f;
}

int serverMaxNB(const s4f_server* S) { return *std::max_element(S->hiNB.begin(), S->hiNB.end()); }

BlockMap makeMap(const s4f_server* S, const int* inst, const int* s0, const int* s1, const int* fin, int nBlocks) {
    BlockMap b{};
    b.inst = inst; b.s0 = s0; b.s1 = s1;
    b.iBlk0 = S->iBlk0.p; b.iNBlk = S->iNBlk.p; b.iCell0 = S->iCell0.p; b.iNCell = S->iNCell.p;
    b.iB0 = S->iB0.p; b.iNB = S->iNB.p;
    b.finInst = fin;
    b.nBlocks = nBlocks;
    return b;
}

// admits queued requests in arrival order while they fit
template <typename T>
void serverAdmit(s4f_server* S) {
    std::vector<int> newSlots;
    for (;;) {
        Request rq;
        {
            std::lock_guard<std::mutex> lk(S->mu_);
            if (S->incoming.empty()) break;
            rq = S->incoming.front();
        }
        int slot = -1;
        for (int i = 0; i < S->maxSlots; ++i)
            if (S->slotReq[i] < 0) { slot = i; break; }
        if (slot < 0) break;
        const int pi = acquirePlacement(S, rq);
        if (pi < 0) break;  // head of the queue waits for memory (arrival order is kept)
        {
            std::lock_guard<std::mutex> lk(S->mu_);
            S->incoming.pop_front();
        }
        S->nQueued--;
        Placement& pl = S->places[pi];
        rq.place = pi;
        rq.slot = slot;
        rq.tStart = nowSeconds();
        const int warm = (rq.warm && pl.hasField) ? 1 : 0;
        placeLoads(S, rq, pl);
        Ctl c = makeCtl(S->opts, 0);
        c.dpar = warm ? pl.dpar : 0;
        c.solveMask = rq.mesh->h.instMask.empty() ? 7 : rq.mesh->h.instMask[0];
        c.pcMode = S->mgOn ? 1 : 0;
        c.warm = warm;
        c.sigma0 = warm ? 0 : 1;
        HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
        hcopy(S->ctx->stream, S->ctl.p + slot, &c, 1);
        const int n = (int)pl.nCells;
        hipLaunchKernelGGL(k_slot_init, dim3(std::max(1, std::min(2048, (n + kBlock - 1) / kBlock))), dim3(kBlock), 0,
                           S->ctx->stream, S->D0.p, S->D1.p, S->Dold.p, S->G.p, S->g.p, (int)S->Ccap, (int)S->Bcap,
                           (int)pl.c0, n, (int)pl.b0, (int)pl.nB, warm, c.dpar);
        HIP_CHECK(hipGetLastError());
        S->hiCell0[slot] = (int)pl.c0;
        S->hiNCell[slot] = n;
        S->hiB0[slot] = (int)pl.b0;
        S->hiNB[slot] = (int)pl.nB;
        pl.hasField = false;
        S->live.push_back(rq);
        S->slotReq[slot] = (int)S->live.size() - 1;
        newSlots.push_back(slot);
    }
    if (newSlots.empty()) return;
    // initialisation (constructor grad) of the new instances only
    std::vector<int> inst, s0, s1;
    buildTiles(S, newSlots, inst, s0, s1, S->hiBlk0, S->hiNBlk);
    hipStream_t s = S->ctx->stream;
    S->nbmInst.upload(inst, s); S->nbmS0.upload(s0, s); S->nbmS1.upload(s1, s); S->nfinList.upload(newSlots, s);
    S->iBlk0.upload(S->hiBlk0, s); S->iNBlk.upload(S->hiNBlk, s);
    S->iCell0.upload(S->hiCell0, s); S->iNCell.upload(S->hiNCell, s);
    S->iB0.upload(S->hiB0, s); S->iNB.upload(S->hiNB, s);
    S->part.alloc(std::max<size_t>(S->part.n, inst.size() * kMaxPartials));
    const BlockMap bm = makeMap(S, S->nbmInst.p, S->nbmS0.p, S->nbmS1.p, S->nfinList.p, (int)inst.size());
    const DevMesh dm = serverDm(S);
    launchGradStep<T>(dm, bm, (int)newSlots.size(), serverMaxNB(S),
                      serverGradBufs(S), nullptr,
                      (const double*)S->Dold.p, S->ctl.p, S->part.p, 1, s);
    hipLaunchKernelGGL(k_fin_grad, dim3((int)newSlots.size()), dim3(kFinBlock), 0, s, bm, S->ctl.p,
                       (const double*)S->part.p, (double*)nullptr, 1);
    HIP_CHECK(hipGetLastError());
    S->nTiles = -1;  // active map must be rebuilt
}

template <typename D>
void ensureCap(DArr<D>& a, size_t n) {
    if (a.n < n || !a.p) a.alloc(std::max<size_t>(n, 1));
}

// Packs the in-flight requests' own hierarchies into one block-diagonal
// hierarchy: per level, each request's rows start on a slice boundary (its
// slices, entries and member lists copied with index offsets); a request
// whose hierarchy is shallower passes through the deeper levels (identity
// aggregation, marked in instPass) to its dense block on the coarsest level.
// All copies are segmented device-to-device kernels.
void serverPackMg(s4f_server* S, const std::vector<int>& slots) {
    hipStream_t s = S->ctx->stream;
    const bool f32 = S->opts.precision == S4F_PRECISION_MIXED;
    const int nI = (int)slots.size();
    std::vector<s4f_mesh*> ms(nI);
    std::vector<int> c0(nI), dk(nI);
    int D = 2;
    for (int k = 0; k < nI; ++k) {
        const Request& rq = S->live[S->slotReq[slots[k]]];
        ms[k] = rq.mesh;
        c0[k] = (int)S->places[rq.place].c0;
        dk[k] = (int)ms[k]->amg.h.size();
        if (dk[k] < 2 || (int)ms[k]->amg.d.size() != dk[k]) throw InvalidError("request without a multigrid hierarchy");
        D = std::max(D, dk[k]);
    }
    while ((int)S->mg.size() > D) S->mg.pop_back();
    while ((int)S->mg.size() < D) S->mg.push_back(std::make_unique<AmgDevLevel>());
    auto rowsOf = [&](int k, int l) { return ms[k]->amg.h[std::min(l, dk[k] - 1)].n; };
    std::vector<std::vector<int>> R(D, std::vector<int>(nI, 0)), Sl = R, E = R, Mo = R;
    std::vector<int> Rt(D, 0), St(D, 0), Et(D, 0), Mt(D, 0);
    for (int l = 1; l < D; ++l)
        for (int k = 0; k < nI; ++k) {
            const bool real = l < dk[k];
            const int rows = rowsOf(k, l), slices = (rows + kSlice - 1) / kSlice;
            const AmgHostLevel& H = ms[k]->amg.h[std::min(l, dk[k] - 1)];
            const int ents = real ? H.sliceStart[H.nSlices] : 0;
            const int mem = real ? (int)H.mem.size() : rows;
            R[l][k] = Rt[l]; Sl[l][k] = St[l]; E[l][k] = Et[l]; Mo[l][k] = Mt[l];
            Rt[l] += slices * kSlice; St[l] += slices; Et[l] += ents; Mt[l] += mem;
        }
    std::vector<SegInt> si;
    std::vector<SegWords> sw;
    auto addI = [&](const int* src, int* dst, int n, int add, int mode) {
        if (n > 0) si.push_back(SegInt{src, dst, n, add, mode, 0});
    };
    auto addW = [&](const void* src, void* dst, long long nWords) {
        if (nWords > 0) sw.push_back(SegWords{(const unsigned*)src, (unsigned*)dst, nWords});
    };
    std::vector<int> blkRow0(S->maxSlots, 0), blkN(S->maxSlots, 0), invOff(S->maxSlots, 0);
    long long invTotal = 0;
    for (int k = 0; k < nI; ++k) {
        const int nc = rowsOf(k, D - 1);
        blkRow0[slots[k]] = R[D - 1][k];
        blkN[slots[k]] = nc;
        invOff[slots[k]] = (int)invTotal;
        invTotal += (long long)nc * nc;
    }
    if (invTotal > INT32_MAX) throw InvalidError("coarsest blocks too large");
    for (int l = 1; l < D; ++l) {
        AmgDevLevel& L = *S->mg[l];
        L.n = Rt[l];
        L.nSlices = St[l];
        ensureCap(L.sliceStart, (size_t)St[l] + 1);
        ensureCap(L.sliceWidth, St[l]);
        ensureCap(L.col, Et[l]);
        if (f32) { ensureCap(L.a32, Et[l]); ensureCap(L.d32, Rt[l]); ensureCap(L.iD32, Rt[l]); }
        else { ensureCap(L.a64, Et[l]); ensureCap(L.d64, Rt[l]); ensureCap(L.iD64, Rt[l]); }
        ensureCap(L.memOff, (size_t)Rt[l] + 1);
        ensureCap(L.mem, Mt[l]);
        ensureCap(L.rowInst, Rt[l]);
        if (l < D - 1) {
            ensureCap(L.agg, Rt[l]);
            ensureCap(L.qCol, Et[l]);
            ensureCap(L.qWidth, St[l]);
            if (f32) ensureCap(L.q32, Et[l]);
            else ensureCap(L.q64, Et[l]);
        }
        std::vector<int> pass(S->maxSlots, 0);
        for (int k = 0; k < nI; ++k) {
            const bool real = l < dk[k];
            const int rows = rowsOf(k, l), slices = (rows + kSlice - 1) / kSlice;
            const int finerOff = l == 1 ? c0[k] : R[l - 1][k];
            pass[slots[k]] = l >= dk[k] - 1 ? 1 : 0;
            if (real) {
                const AmgDevLevel& Dl = *ms[k]->amg.d[l];
                const AmgHostLevel& H = ms[k]->amg.h[l];
                const int ents = H.sliceStart[H.nSlices];
                addI(Dl.sliceStart.p, L.sliceStart.p + Sl[l][k], slices, E[l][k], 0);
                addI(Dl.sliceWidth.p, L.sliceWidth.p + Sl[l][k], slices, 0, 0);
                addI(Dl.col.p, L.col.p + E[l][k], ents, R[l][k], 1);
                if (f32) {
                    addW(Dl.a32.p, L.a32.p + E[l][k], ents);
                    addW(Dl.d32.p, L.d32.p + R[l][k], rows);
                    addW(Dl.iD32.p, L.iD32.p + R[l][k], rows);
                } else {
                    addW(Dl.a64.p, L.a64.p + E[l][k], 2LL * ents);
                    addW(Dl.d64.p, L.d64.p + R[l][k], 2LL * rows);
                    addW(Dl.iD64.p, L.iD64.p + R[l][k], 2LL * rows);
                }
                addI(Dl.memOff.p, L.memOff.p + R[l][k], rows, Mo[l][k], 0);
                addI(Dl.mem.p, L.mem.p + Mo[l][k], (int)H.mem.size(), finerOff, 0);
                if (l < D - 1) {
                    if (l < dk[k] - 1) {
                        addI(Dl.agg.p, L.agg.p + R[l][k], rows, R[l + 1][k], 0);
                        addI(Dl.qCol.p, L.qCol.p + E[l][k], ents, R[l + 1][k], 1);
                        if (f32) addW(Dl.q32.p, L.q32.p + E[l][k], ents);
                        else addW(Dl.q64.p, L.q64.p + E[l][k], 2LL * ents);
                        addI(Dl.qWidth.p, L.qWidth.p + Sl[l][k], slices, 0, 0);
                    } else {
                        addI(nullptr, L.agg.p + R[l][k], rows, R[l + 1][k], 2);
                        addI(nullptr, L.qWidth.p + Sl[l][k], slices, 0, 3);
                    }
                }
            } else {  // pass-through rows: no matrix, identity members and aggregation
                addI(nullptr, L.sliceStart.p + Sl[l][k], slices, E[l][k], 3);
                addI(nullptr, L.sliceWidth.p + Sl[l][k], slices, 0, 3);
                addI(nullptr, L.memOff.p + R[l][k], rows, Mo[l][k], 2);
                addI(nullptr, L.mem.p + Mo[l][k], rows, finerOff, 2);
                if (l < D - 1) {
                    addI(nullptr, L.agg.p + R[l][k], rows, R[l + 1][k], 2);
                    addI(nullptr, L.qWidth.p + Sl[l][k], slices, 0, 3);
                }
            }
            const int memN = real ? (int)ms[k]->amg.h[l].mem.size() : rows;
            addI(nullptr, L.rowInst.p + R[l][k], rows, slots[k], 3);
            addI(nullptr, L.rowInst.p + R[l][k] + rows, slices * kSlice - rows, -1, 3);
            addI(nullptr, L.memOff.p + R[l][k] + rows, slices * kSlice - rows, Mo[l][k] + memN, 3);
        }
        addI(nullptr, L.sliceStart.p + St[l], 1, Et[l], 3);
        addI(nullptr, L.memOff.p + Rt[l], 1, Mt[l], 3);
        L.instPass.upload(pass, s);
        if (l == D - 1) {
            L.blkRow0.upload(blkRow0, s);
            L.blkN.upload(blkN, s);
            L.invOff.upload(invOff, s);
            if (f32) ensureCap(L.inv32, (size_t)invTotal);
            else ensureCap(L.inv64, (size_t)invTotal);
            for (int k = 0; k < nI; ++k) {
                const AmgDevLevel& Dc = *ms[k]->amg.d[dk[k] - 1];
                const long long nn = (long long)blkN[slots[k]] * blkN[slots[k]];
                if (f32) addW(Dc.inv32.p, L.inv32.p + invOff[slots[k]], nn);
                else addW(Dc.inv64.p, L.inv64.p + invOff[slots[k]], 2 * nn);
            }
        }
    }
    // level 0: arena cell -> packed level-1 row; q in the arena's entry
    // slots with its columns moved to the packed level-1 rows
    AmgDevLevel& L0 = *S->mg[0];
    for (int k = 0; k < nI; ++k) {
        const AmgDevLevel& M0 = *ms[k]->amg.d[0];
        addI(M0.agg.p, L0.agg.p + c0[k], ms[k]->amg.h[0].n, R[1][k], 0);
        const size_t e0 = (size_t)S->places[S->live[S->slotReq[slots[k]]].place].e0;
        const int nE = (int)M0.qCol.n;
        addI(M0.qCol.p, L0.qCol.p + e0, nE, R[1][k], 1);
        if (f32) addW(M0.q32.p, L0.q32.p + e0, nE);
        else addW(M0.q64.p, L0.q64.p + e0, 2LL * nE);
    }
    if (!si.empty()) {
        S->segI.upload(si, s);
        int maxN = 0;
        for (const SegInt& g : si) maxN = std::max(maxN, g.n);
        for (size_t a = 0; a < si.size(); a += 65535) {
            const int nseg = (int)std::min<size_t>(65535, si.size() - a);
            hipLaunchKernelGGL(k_seg_int, dim3(std::max(1, std::min(1024, (maxN + kBlock - 1) / kBlock)), nseg),
                               dim3(kBlock), 0, s, (const SegInt*)(S->segI.p + a));
        }
    }
    if (!sw.empty()) {
        S->segW.upload(sw, s);
        long long maxN = 0;
        for (const SegWords& g : sw) maxN = std::max(maxN, g.n);
        for (size_t a = 0; a < sw.size(); a += 65535) {
            const int nseg = (int)std::min<size_t>(65535, sw.size() - a);
            hipLaunchKernelGGL(k_seg_words,
                               dim3((unsigned)std::max<long long>(1, std::min<long long>(1024, (maxN + kBlock - 1) / kBlock)), nseg),
                               dim3(kBlock), 0, s, (const SegWords*)(S->segW.p + a));
        }
    }
    HIP_CHECK(hipGetLastError());
    // the host tables above are read by the copies: finish them before they go
    HIP_CHECK(hipStreamSynchronize(s));
}

MgRun serverRun(s4f_server* S) {
    return MgRun{&S->mg, kMgOmega, S->opts.mgCoarseScale > 0.0 ? S->opts.mgCoarseScale : kMgCoarseScaleDefault,
                 S->nActive};
}

void serverRebuildMap(s4f_server* S) {
    S->mapGen++;  // tiles, multigrid packing and slot set change: captured graphs are stale
    std::vector<int> slots;
    for (int i = 0; i < S->maxSlots; ++i)
        if (S->slotReq[i] >= 0) slots.push_back(i);
    if (S->mgOn && !slots.empty()) serverPackMg(S, slots);
    std::vector<int> inst, s0, s1;
    buildTiles(S, slots, inst, s0, s1, S->hiBlk0, S->hiNBlk);
    hipStream_t s = S->ctx->stream;
    S->bmInst.upload(inst, s); S->bmS0.upload(s0, s); S->bmS1.upload(s1, s); S->finList.upload(slots, s);
    S->iBlk0.upload(S->hiBlk0, s); S->iNBlk.upload(S->hiNBlk, s);
    S->part.alloc(std::max<size_t>(S->part.n, inst.size() * kMaxPartials));
    S->nTiles = (int)inst.size();
    S->nActive = (int)slots.size();
}

// one outer iteration of every in-flight solve, then retire the converged
template <typename T>
void serverStep(s4f_server* S) {
    hipStream_t s = S->ctx->stream;
    if (S->nTiles < 0) serverRebuildMap(S);
    if (S->nActive == 0) return;
    const DevMesh dm = serverDm(S);
    const BlockMap bm = makeMap(S, S->bmInst.p, S->bmS0.p, S->bmS1.p, S->finList.p, S->nTiles);
    PcgVecs<T> v = serverVecs<T>(S);
    const int nA = S->nActive;
    std::vector<int> slots;
    for (int i = 0; i < S->maxSlots; ++i)
        if (S->slotReq[i] >= 0) slots.push_back(i);
    // one outer iteration of every active request: explicit terms and the
    // first V-cycle, PCG iterations (in pairs, so the search-direction
    // buffers return to their places), the gradient step
    auto launchStart = [&](hipStream_t q) {
        hipLaunchKernelGGL((k_residual<T>), dim3(bm.nBlocks), dim3(kBlock), 0, q, dm, bm, (const double*)S->D0.p,
                           (const double*)S->D1.p, S->G.p, S->sigb.p, v.r, S->ctl.p, S->part.p);
        hipLaunchKernelGGL(k_fin_residual, dim3(nA), dim3(kFinBlock), 0, q, bm, S->ctl.p, (const double*)S->part.p);
        if (S->mgOn) launchVcycle<T>(serverRun(S), dm, bm, v, S->ctl.p, S->part.p, q);  // z0 = V(r0), rho0
    };
    auto launchPcg = [&](int n, hipStream_t q) {
        if (S->mgOn) launchPcgChunkMg<T>(serverRun(S), dm, bm, v, S->ctl.p, S->part.p, n, q);
        else launchPcgChunk<T>(dm, bm, nA, v, S->ctl.p, S->part.p, n, q);
    };
    auto launchEnd = [&](hipStream_t q) {
        launchGradStep<T>(dm, bm, nA, serverMaxNB(S), serverGradBufs(S), (const T*)v.x, (const double*)S->Dold.p,
                          S->ctl.p, S->part.p, 0, q);
        hipLaunchKernelGGL(k_fin_grad, dim3(nA), dim3(kFinBlock), 0, q, bm, S->ctl.p, (const double*)S->part.p,
                           (double*)nullptr, 0);
    };
    // the first step on a new map runs eagerly (it allocates the multigrid
    // vectors of the packing); the next one captures
    bool graphs = false;
    if (S->opts.eagerLaunch == 0 && S->ranGen == S->mapGen) {
        graphs = true;
        if (S->capGen != S->mapGen) {
            try {
                S->gStart.capture(S->ctx->capture, launchStart);
                S->gPair.capture(S->ctx->capture, [&](hipStream_t q) { launchPcg(2, q); });
                S->gEnd.capture(S->ctx->capture, launchEnd);
                S->capGen = S->mapGen;
            } catch (const HipError&) {  // a runtime without capture: the same launches, one by one
                S->gStart.reset();
                S->gPair.reset();
                S->gEnd.reset();
                (void)hipGetLastError();
                S->opts.eagerLaunch = 1;
                graphs = false;
            }
        }
    }
    if (graphs) S->gStart.launch(s);
    else launchStart(s);
    int total = 0, chunk = S->chunk;
    for (;;) {
        if (graphs) {
            const int pairs = (chunk + 1) / 2;
            for (int k = 0; k < pairs; ++k) S->gPair.launch(s);
            chunk = 2 * pairs;
        } else {
            launchPcg(chunk, s);
        }
        total += chunk;
        chunk = std::max(2, chunk / 4);
        HIP_CHECK(hipMemcpyAsync(S->hctl, S->ctl.p, S->maxSlots * sizeof(Ctl), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipStreamSynchronize(s));
        bool any = false;
        for (int sl : slots)
            if (!S->hctl[sl].outerDone && !S->hctl[sl].innerDone) any = true;
        if (!any || total > S->opts.pcgMaxIter + 2) break;
    }
    int used = 1;
    for (int sl : slots)
        if (!S->hctl[sl].outerDone) used = std::max({used, S->hctl[sl].nIter[0], S->hctl[sl].nIter[1], S->hctl[sl].nIter[2]});
    S->chunk = std::max(1, used - 1);
    if (graphs) S->gEnd.launch(s);
    else launchEnd(s);
    HIP_CHECK(hipGetLastError());
    S->ranGen = S->mapGen;
    HIP_CHECK(hipMemcpyAsync(S->hctl, S->ctl.p, S->maxSlots * sizeof(Ctl), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    // retire
    bool changed = false;
    for (int sl : slots) {
        const Ctl& cf = S->hctl[sl];
        if (!cf.outerDone) continue;
        Request& rq = S->live[S->slotReq[sl]];
        Placement& pl = S->places[rq.place];
        const int n = pl.nCells;
        hipLaunchKernelGGL(k_sigma, dim3(std::max(1, std::min(2048, (n + kBlock - 1) / kBlock))), dim3(kBlock), 0, s, dm,
                           S->G.p, S->sig.p, (int)pl.c0, n);
        HIP_CHECK(hipGetLastError());
        Result res;
        res.D.resize(3 * (size_t)n);
        res.sigma.resize(6 * (size_t)n);
        std::vector<double> tmp(n);
        const double* Df = cf.dpar ? S->D1.p : S->D0.p;
        for (int q = 0; q < 3; ++q) {
            HIP_CHECK(hipMemcpyAsync(tmp.data(), Df + (size_t)q * S->Ccap + pl.c0, n * sizeof(double),
                                     hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            for (int c = 0; c < n; ++c) res.D[3 * (size_t)c + q] = tmp[c];
        }
        for (int q = 0; q < 6; ++q) {
            HIP_CHECK(hipMemcpyAsync(tmp.data(), S->sig.p + (size_t)q * S->Ccap + pl.c0, n * sizeof(double),
                                     hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            for (int c = 0; c < n; ++c) res.sigma[6 * (size_t)c + q] = tmp[c];
        }
        res.rep.nOuter = cf.nOuter;
        res.rep.converged = cf.converged;
        res.rep.innerTotal = cf.innerTotal;
        res.rep.initialResidual = cf.initResMag;
        res.rep.residualvf = cf.rvf;
        res.rep.warmStarted = cf.warm;
        res.status = cf.diverged ? S4F_ERR_DIVERGED : (cf.converged ? S4F_OK : S4F_WARN_MAXITER);
        res.tSubmit = rq.tSubmit;
        res.tStart = rq.tStart;
        res.tFinish = nowSeconds();
        res.rep.seconds = res.tFinish - res.tStart;
        pl.dpar = cf.dpar;
        pl.hasField = cf.converged != 0;
        pl.lastUse = res.tFinish;
        const long long id = rq.id;
        rq.place = -1;
        S->slotReq[sl] = -1;
        changed = true;
        {
            std::lock_guard<std::mutex> lk(S->mu_);
            S->results[id] = std::move(res);
            S->doneOrder.push_back(id);
        }
        S->nDone++;
        S->cvOut.notify_all();
    }
    if (changed) {
        // compact the live list (slot -> request indices)
        std::vector<Request> keep;
        for (int i = 0; i < S->maxSlots; ++i)
            if (S->slotReq[i] >= 0) {
                keep.push_back(S->live[S->slotReq[i]]);
                S->slotReq[i] = (int)keep.size() - 1;
            }
        S->live.swap(keep);
        S->nTiles = -1;
    }
}

template <typename T>
void serverLoop(s4f_server* S) {
    try {
        HIP_CHECK(hipSetDevice(S->ctx->device));
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(S->mu_);
                if (S->stop) break;
                if (S->live.empty() && S->incoming.empty())
                    S->cvIn.wait_for(lk, std::chrono::milliseconds(20));
                if (S->stop) break;
            }
            serverAdmit<T>(S);
            if (S->live.empty()) continue;
            serverStep<T>(S);
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(S->mu_);
        S->threadError = e.what();
        S->stop = true;
        S->cvOut.notify_all();
    }
}

}  // namespace

extern "C" {

s4f_status s4f_server_create(s4f_ctx* ctx, int64_t budgetBytes, int32_t maxSlots, const s4f_solve_opts* opts,
                             s4f_server** out) {
    return guarded([&]() -> s4f_status {
        if (!ctx || !opts || !out) throw InvalidError("null argument");
        if (ctx->destroyPending) throw InvalidError("context was destroyed");
        if (budgetBytes < (1 << 20)) throw InvalidError("budget must be at least 1 MiB");
        if (maxSlots < 1 || maxSlots > 4096) throw InvalidError("maxSlots must be in 1..4096");
        checkOpts(*opts);
        if (opts->d2dt2 != S4F_D2DT2_STEADY) throw InvalidError("the server runs steady solves");
        if (opts->preconditioner == S4F_PRECOND_GAMG_SMOOTHED)
            throw InvalidError("the server packs plain-aggregation hierarchies: use S4F_PRECOND_GAMG");
        HIP_CHECK(hipSetDevice(ctx->device));
        auto S = std::make_unique<s4f_server>();
        S->ctx = ctx;
        S->opts = *opts;
        S->opts.warmStart = 0;
        S->maxSlots = maxSlots;
        S->budget = budgetBytes;
        serverAllocate(S.get());
        s4f_server* raw = S.get();
        if (S->opts.precision == S4F_PRECISION_MIXED) S->th = std::thread(serverLoop<float>, raw);
        else S->th = std::thread(serverLoop<double>, raw);
        ctx->children++;
        *out = S.release();
        return S4F_OK;
    });
}

s4f_status s4f_server_submit(s4f_server* S, const s4f_request* rq, int64_t* id) {
    return guarded([&]() -> s4f_status {
        if (!S || !rq || !rq->mesh || !rq->bcValue) throw InvalidError("null argument");
        s4f_mesh* m = rq->mesh;
        if (!m->materialSet) throw InvalidError("material not set");
        if (m->ctx != S->ctx) throw InvalidError("mesh belongs to another context");
        if (m->h.instCell0.size() != 1) throw InvalidError("requests take plain meshes");
        if (!m->h.zone.empty()) throw InvalidError("a mesh with material zones (sub-meshes) is solved alone");
        if (m->inerKey != 0.0 || m->h.g[0] != 0.0 || m->h.g[1] != 0.0 || m->h.g[2] != 0.0)
            throw InvalidError("requests take static meshes without body force (the last solve was transient or "
                               "gravity is set)");
        const HostMesh& h = m->h;
        if ((long long)h.nSlices * kSlice > S->Ccap || (long long)h.col.size() > S->Ecap || h.B > S->Bcap)
            return fail(S4F_ERR_ADMISSION, "request footprint exceeds the server's HBM budget");
        if (S->mgOn) {  // the request's own hierarchy (cached on the mesh), packed by the server thread
            HIP_CHECK(hipSetDevice(S->ctx->device));
            buildAmg(m, false);
        }
        Request r;
        r.mesh = m;
        r.warm = rq->warmStart;
        r.bVal.resize(3 * (size_t)h.B);
        r.bPres.assign(h.B, 0.0);
        for (int b = 0; b < h.B; ++b) {
            for (int q = 0; q < 3; ++q) {
                const double v = rq->bcValue[3 * (size_t)b + q];
                if (!std::isfinite(v)) throw InvalidError("non-finite boundary value");
                r.bVal[(size_t)q * h.B + b] = v;
            }
            if (rq->bcPressure) r.bPres[b] = rq->bcPressure[b];
        }
        r.tSubmit = nowSeconds();
        {
            std::lock_guard<std::mutex> lk(S->mu_);
            if (S->stop) return fail(S4F_ERR_HIP, S->threadError.empty() ? "server stopped" : S->threadError);
            r.id = S->nextId++;
            if (id) *id = r.id;
            S->incoming.push_back(std::move(r));
        }
        S->nQueued++;
        S->cvIn.notify_all();
        return S4F_OK;
    });
}

s4f_status s4f_server_poll(s4f_server* S, int32_t waitMs, int64_t* id) {
    if (!S || !id) return fail(S4F_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(S->mu_);
    if (S->doneOrder.empty() && waitMs > 0)
        S->cvOut.wait_for(lk, std::chrono::milliseconds(waitMs), [&] { return !S->doneOrder.empty() || S->stop; });
    if (!S->threadError.empty()) return fail(S4F_ERR_HIP, S->threadError);
    if (S->doneOrder.empty()) return fail(S4F_ERR_NOT_READY, "no finished request");
    *id = S->doneOrder.front();
    S->doneOrder.pop_front();
    return S4F_OK;
}

s4f_status s4f_server_result(s4f_server* S, int64_t id, double* D_out, double* sigma_out, s4f_solve_report* rep,
                             double timing[3]) {
    if (!S) return fail(S4F_ERR_INVALID, "null argument");
    Result r;
    {
        std::lock_guard<std::mutex> lk(S->mu_);
        auto it = S->results.find(id);
        if (it == S->results.end()) return fail(S4F_ERR_NOT_READY, "no such finished request");
        r = std::move(it->second);
        S->results.erase(it);
    }
    if (D_out) std::memcpy(D_out, r.D.data(), r.D.size() * sizeof(double));
    if (sigma_out) std::memcpy(sigma_out, r.sigma.data(), r.sigma.size() * sizeof(double));
    if (rep) *rep = r.rep;
    if (timing) { timing[0] = r.tSubmit; timing[1] = r.tStart; timing[2] = r.tFinish; }
    return r.status;
}

s4f_status s4f_server_stats(s4f_server* S, int64_t out[6]) {
    if (!S || !out) return fail(S4F_ERR_INVALID, "null argument");
    out[0] = S->nQueued.load();
    out[1] = S->nDone.load();
    out[2] = S->bytesInUse.load();
    out[3] = S->budget;
    out[4] = S->Ccap;
    out[5] = S->maxSlots;
    return S4F_OK;
}

s4f_status s4f_server_destroy(s4f_server* S) {
    if (!S) return S4F_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu_);
        S->stop = true;
    }
    S->cvIn.notify_all();
    if (S->th.joinable()) S->th.join();
    s4f_ctx* ctx = S->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    if (S->hctl) (void)hipHostFree(S->hctl);
    delete S;
    ctxChildGone(ctx);
    return S4F_OK;
}

}  // extern "C"

// ===========================================================================
// lduMatrix::solver entry: PCG on a caller's LDU matrix (OpenFOAM addressing)
// ===========================================================================
struct s4f_ldu {
    s4f_ctx* ctx = nullptr;
    std::unique_ptr<s4f_mesh> M;   // ELL structure + PCG state (no geometry)
    std::vector<int> faceEntOwn, faceEntNei;
};

extern "C" {

s4f_status s4f_ldu_create(s4f_ctx* ctx, int32_t nCells, int32_t nFaces, const int32_t* lowerAddr,
                          const int32_t* upperAddr, s4f_ldu** out) {
    return guarded([&]() -> s4f_status {
        if (!ctx || !out || nCells <= 0 || nFaces < 0 || (nFaces && (!lowerAddr || !upperAddr)))
            throw InvalidError("invalid LDU addressing");
        if (ctx->destroyPending) throw InvalidError("context was destroyed");
        HIP_CHECK(hipSetDevice(ctx->device));
        auto L = std::make_unique<s4f_ldu>();
        L->ctx = ctx;
        L->M = std::make_unique<s4f_mesh>();
        s4f_mesh* M = L->M.get();
        M->ctx = ctx;
        HostMesh& h = M->h;
        h.C = nCells; h.F = nFaces; h.B = 0; h.nPatches = 0;
        std::vector<int> cnt(nCells, 0);
        for (int f = 0; f < nFaces; ++f) {
            const int l = lowerAddr[f], u = upperAddr[f];
            if (l < 0 || u < 0 || l >= nCells || u >= nCells || l == u) throw InvalidError("LDU address out of range");
            cnt[l]++; cnt[u]++;
        }
        h.nSlices = (nCells + kSlice - 1) / kSlice;
        h.sliceStart.resize((size_t)h.nSlices + 1);
        h.sliceWidth.resize(h.nSlices);
        size_t E = 0;
        for (int sI = 0; sI < h.nSlices; ++sI) {
            h.sliceStart[sI] = (int)E;
            int wmax = 0;
            for (int l = 0; l < kSlice; ++l) {
                const int c = sI * kSlice + l;
                if (c < nCells) wmax = std::max(wmax, cnt[c]);
            }
            h.sliceWidth[sI] = wmax;
            E += (size_t)wmax * kSlice;
        }
        h.sliceStart[h.nSlices] = (int)E;
        h.col.assign(E, kPad);
        h.eS.assign(3 * E, 0.0); h.eW.assign(E, 0.0);
        std::vector<int> fill(nCells, 0);
        auto slot = [&](int c) { return h.sliceStart[c / kSlice] + fill[c]++ * kSlice + c % kSlice; };
        L->faceEntOwn.resize(nFaces); L->faceEntNei.resize(nFaces);
        for (int f = 0; f < nFaces; ++f) {
            const int eo = slot(lowerAddr[f]);
            h.col[eo] = upperAddr[f];
            L->faceEntOwn[f] = eo;
            const int en = slot(upperAddr[f]);
            h.col[en] = lowerAddr[f];
            L->faceEntNei[f] = en;
        }
        h.cC.assign(3 * (size_t)nCells, 0.0);
        h.minv.assign(6 * (size_t)nCells, 0.0);
        h.bVal.clear(); h.bPres.clear();
        h.instCell0 = {0}; h.instNCell = {nCells}; h.instB0 = {0}; h.instNB = {0};
        h.mu = {1.0}; h.K = {1.0}; h.impK = {1.0};
        uploadGeometry(M);
        ctx->children++;
        *out = L.release();
        return S4F_OK;
    });
}

s4f_status s4f_ldu_solve(s4f_ldu* L, const double* diag, const double* upper, const double* source, double* psi,
                         double tolerance, double relTol, int32_t maxIter, int32_t precision, double* initRes,
                         double* finalRes, int32_t* nIter) {
    return guarded([&]() -> s4f_status {
        if (!L || !diag || !upper || !source || !psi) throw InvalidError("null argument");
        s4f_mesh* M = L->M.get();
        HostMesh& h = M->h;
        hipStream_t s = L->ctx->stream;
        HIP_CHECK(hipSetDevice(L->ctx->device));
        const int C = h.C;
        // off-diagonal entries of a symmetric lduMatrix: -eA = upper
        h.eA.assign(h.col.size(), 0.0);
        for (int f = 0; f < h.F; ++f) {
            h.eA[L->faceEntOwn[f]] = -upper[f];
            h.eA[L->faceEntNei[f]] = -upper[f];
        }
        h.diag.assign(diag, diag + C);
        h.eC.assign(h.col.size(), 0.0);
        for (int c = 0; c < C; ++c)
            if (!(h.diag[c] != 0.0)) throw InvalidError("zero diagonal");
        uploadMaterial(M);
        M->materialSet = true;
        allocState(M);
        SolveState& st = M->st;
        double avg = 0.0;
        for (int c = 0; c < C; ++c) avg += psi[c];
        avg /= C;
        DArr<double> dpsi, dsrc;
        dpsi.upload(std::vector<double>(psi, psi + C), s);
        dsrc.upload(std::vector<double>(source, source + C), s);
        s4f_solve_opts o;
        s4f_solve_opts_default(&o);
        o.pcgTolerance = tolerance;
        o.pcgRelTol = relTol;
        o.pcgMaxIter = maxIter;
        o.precision = precision;
        checkOpts(o);
        Ctl c0 = makeCtl(o, 0);
        HIP_CHECK(hipMemcpyAsync(st.ctl.p, &c0, sizeof(Ctl), hipMemcpyHostToDevice, s));
        const DevMesh dm = M->dm();
        const BlockMap bm = M->bmap();
        Ctl hc;
        std::vector<double> x(C, 0.0);
        auto run = [&](auto v, auto tag) {
            using T = decltype(tag);
            hipLaunchKernelGGL((k_ldu_start<T>), dim3(bm.nBlocks), dim3(kBlock), 0, s, dm, bm, (const double*)dpsi.p,
                               (const double*)dsrc.p, avg, v.r, st.part.p);
            hipLaunchKernelGGL(k_fin_residual, dim3(1), dim3(kFinBlock), 0, s, bm, st.ctl.p, (const double*)st.part.p);
            HIP_CHECK(hipMemsetAsync(v.x, 0, 3 * (size_t)C * sizeof(T), s));
            HIP_CHECK(hipGetLastError());
            int total = 0, chunk = 8;
            for (;;) {
                HIP_CHECK(hipMemcpyAsync(&hc, st.ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, s));
                HIP_CHECK(hipStreamSynchronize(s));
                if (hc.innerDone || total > maxIter + 8) break;
                launchPcgChunk<T>(dm, bm, 1, v, st.ctl.p, st.part.p, chunk, s);
                total += chunk;
            }
            std::vector<T> xs(C);
            HIP_CHECK(hipMemcpyAsync(xs.data(), v.x, C * sizeof(T), hipMemcpyDeviceToHost, s));
            HIP_CHECK(hipStreamSynchronize(s));
            if (!hc.firstInner)
                for (int c = 0; c < C; ++c) x[c] = (double)xs[c];
        };
        if (precision == S4F_PRECISION_MIXED) run(pcgVecs<float>(st, C), 0.0f);
        else run(pcgVecs<double>(st, C), 0.0);
        for (int c = 0; c < C; ++c) psi[c] += x[c];
        if (initRes) *initRes = hc.initRes[0];
        if (finalRes) *finalRes = hc.finalRes[0];
        if (nIter) *nIter = hc.nIter[0];
        return S4F_OK;
    });
}

s4f_status s4f_ldu_destroy(s4f_ldu* L) {
    if (!L) return S4F_OK;
    s4f_ctx* ctx = L->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete L;
    ctxChildGone(ctx);
    return S4F_OK;
}

}  // extern "C"
