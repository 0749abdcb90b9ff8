// s4f_amg.h -- aggregation multigrid V-cycle used as the PCG preconditioner
// (the device counterpart of running OpenFOAM's PCG with `preconditioner GAMG`).
//
// Hierarchy: plain (piecewise-constant) aggregation of strongly connected
// cells, Galerkin coarse operators A_c = P^T A P, built once per matrix on the
// host.  Cycle: V(1,1) with damped Jacobi smoothing (omega = kMgOmega) and an
// exact solve on the coarsest level (explicit inverse of each instance's
// block, <= kCoarseBlock rows), so the cycle is a fixed symmetric positive
// definite operator and CG stays CG.  A batch of meshes is one block-diagonal
// matrix: aggregates never cross instances, every level keeps each
// instance's rows contiguous, and each row looks up its instance's control
// block (converged instances and components are skipped).
//
// Launch shape per level (level 0: 3 components, SoA, stride = the row
// count, converged components skipped; coarser levels: rows of four):
//   restriction: b_c[I] = sum_{i in I} r[i] (members gathered, no atomics)
//         and the pre-smoothed value x_c = w D_c^-1 b_c in the same pass
//   down: r = b - A x (on level 0 the neighbours' x = w D^-1 b is recomputed
//         from their b and 1/diag)
//   up:   prolongation x += e[agg] fused into the post-smoothing sweep,
//         formed from the down sweep's residual and q = offdiag(A) P (no
//         matrix pass; only the coarser correction is gathered)
// Level 0 is the solver mesh's own sliced-ELL matrix; coarser levels have
// their own sliced-ELL arrays.  Every row's arithmetic runs in a fixed order
// that depends only on its own instance, so results do not depend on which
// other instances share a launch.
#pragma once

#include "s4f_kernels.h"

namespace s4f {

constexpr int kCoarseBlock = 1024;  // rows per instance on the coarsest level (explicit inverse)
constexpr int kMgChunk = 8;         // slots per unconditional load group of the wide (Galerkin / G) rows
// smoothed hierarchy: relative size below which first-coarse-level entries
// are lumped into the diagonal (|a_IJ| < f sqrt(|a_II a_JJ|)).  Off: lumping
// keeps row sums but can make the coarse operator indefinite (it did on a
// 384-tetrahedron mesh at f = 0.04), and the cycle must stay SPD.
constexpr double kMgSmoothedFilter = 0.0;
constexpr double kMgCoarseScaleDefault = 1.8;  // coarse-correction scale when the caller gives 0
// Jacobi damping of the smoothing sweeps: near the best smoothing factor of
// damped Jacobi for 7-point-like operators in 3-D (6/7) and below 1, where
// the sweep converges on every level's diagonally dominant M-matrix, so the
// cycle stays symmetric positive definite
constexpr double kMgOmega = 0.9;

// device view of one coarse level
struct MgLevel {
    int n;             // rows
    int nSlices;
    const int* sliceStart;
    const int* sliceWidth;
    const int* col;    // >= 0 neighbour row, kPad empty
    const float* a32;  // off-diagonal magnitude (entry = -a)
    const double* a64;
    const float* diag32;
    const double* diag64;
    const float* invD32;
    const double* invD64;
    const int* agg;    // row -> row of the next coarser level (nullptr on the coarsest)
    const int* memOff; // (coarse side) row -> first member of the finer level
    const int* mem;    // members (rows of the finer level)
    const float* memW32;   // member weights (smoothed prolongation; nullptr: all 1)
    const double* memW64;
    const int* rStart;     // the weighted member lists as sliced ELL (smoothed prolongation only)
    const int* rWidth;
    const int* rMem;
    const float* rW32;
    const double* rW64;
    const int* rowInst;  // row -> instance
    const int* blkRow0;  // coarsest level, per instance: first row of its block
    const int* blkN;     //   rows of its block
    const int* invOff;   //   offset of its explicit inverse (blkN^2 entries, row-major)
    const int* pass;     // per instance: passes through this level unchanged (no smoothing, no scaling)
    const int* qCol;     // q = offdiag(A) P in the level's slots (< 0: empty; nullptr on the coarsest)
    const float* q32;
    const double* q64;
    const int* qWidth;   // per slice: filled q slots
};

template <typename T> struct MgSel;
template <> struct MgSel<double> {
    __device__ static const double* q(const MgLevel& L) { return L.q64; }
    __device__ static const double* memW(const MgLevel& L) { return L.memW64; }
    __device__ static const double* rW(const MgLevel& L) { return L.rW64; }
    __device__ static const double* a(const MgLevel& L) { return L.a64; }
    __device__ static const double* d(const MgLevel& L) { return L.diag64; }
    __device__ static const double* iD(const MgLevel& L) { return L.invD64; }
};
template <> struct MgSel<float> {
    __device__ static const float* q(const MgLevel& L) { return L.q32; }
    __device__ static const float* memW(const MgLevel& L) { return L.memW32; }
    __device__ static const float* rW(const MgLevel& L) { return L.rW32; }
    __device__ static const float* a(const MgLevel& L) { return L.a32; }
    __device__ static const float* d(const MgLevel& L) { return L.diag32; }
    __device__ static const float* iD(const MgLevel& L) { return L.invD32; }
};

__device__ __forceinline__ bool mgSkip(const Ctl* c) { return c->outerDone || c->innerDone; }

// Coarse-level vectors are rows of four (three components and a zero pad):
// one 16-byte (FP32) or 32-byte (FP64) load per neighbour instead of three
// component gathers.  The fine level keeps the PCG's component-major layout.
template <typename T> struct Row4Sel;
template <> struct Row4Sel<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct Row4Sel<double> { typedef double type __attribute__((ext_vector_type(4))); };
template <typename T> using Row4 = typename Row4Sel<T>::type;

template <typename T>
__device__ __forceinline__ Row4<T> row4(T a, T b, T c) {
    Row4<T> v;
    v[0] = a; v[1] = b; v[2] = c; v[3] = (T)0;
    return v;
}

// ---------------------------------------------------------------------------
// level 0 (the mesh's own matrix, BlockMap iteration)
// ---------------------------------------------------------------------------

// Level-0 rows with all three components active and a compile-time slot
// count W: every load is issued unconditionally (invalid slots gather the
// row's own cell with a zero coefficient) so a wave's gathers overlap.
template <typename T, int W>
__device__ __forceinline__ void mg0DownRowAll(const DevMesh& m, int P, int e0, int lane, int C, const T* __restrict__ b,
                                              T* __restrict__ r, T w) {
    const T* __restrict__ A = MatSel<T>::a(m);
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    int nb[W];
    T a[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int e = e0 + k * kSlice + lane;
        nb[k] = ldS(m.col + e);
        a[k] = ldS(A + e);
    }
    const T iD = invD[P], dP = MatSel<T>::diag(m)[P];
    T bP[3], xP[3], y[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        bP[j] = b[(size_t)j * C + P];
        xP[j] = w * iD * bP[j];
        y[j] = dP * xP[j];
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const bool v = nb[k] >= 0;
        const int c = v ? nb[k] : P;
        const T av = v ? a[k] : (T)0;
        const T wi = w * invD[c];
#pragma unroll
        for (int j = 0; j < 3; ++j) y[j] -= av * (wi * b[(size_t)j * C + c]);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) stS(r + (size_t)j * C + P, bP[j] - y[j]);
}

// level-0 prolongation + post-smoothing of one row (all three components,
// compile-time slot count W, every load unconditional: an empty slot reads
// the row's own coarse row with a zero weight)
template <typename T, int W>
__device__ __forceinline__ void mg0UpRowAll(const DevMesh& m, const int* __restrict__ qCol, const T* __restrict__ q,
                                            int P, int e0, int lane, int C, const T* __restrict__ b,
                                            const T* __restrict__ r0, const int* __restrict__ agg,
                                            const Row4<T>* __restrict__ e, T* __restrict__ z, T w, T w1, T al,
                                            double* acc) {
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    int J[W];
    T qv[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int en = e0 + k * kSlice + lane;
        J[k] = ldS(qCol + en);
        qv[k] = ldS(q + en);
    }
    const int IP = agg[P];
    const T wi = w * invD[P];
    T s[3] = {(T)0, (T)0, (T)0};
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const bool v = J[k] >= 0;
        const Row4<T> ek = e[v ? J[k] : IP];
        const T c = v ? qv[k] : (T)0;
#pragma unroll
        for (int j = 0; j < 3; ++j) s[j] += c * ek[j];
    }
    const Row4<T> eP = e[IP];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const size_t i = (size_t)j * C + P;
        const T bv = b[i];
        const T zv = wi * (bv + r0[i]) + al * (w1 * eP[j] + wi * s[j]);
        stS(z + i, zv);
        acc[j] += (double)zv * (double)bv;
    }
}

// pre-smoothing from zero and the residual for the restriction:
// x = w D^-1 b (not stored: k_mg0_up works from b and r), r = b - A x
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg0_down(const DevMesh m, const BlockMap bm, const T* __restrict__ b,
                                                     T* __restrict__ r, double omega,
                                                     const Ctl* __restrict__ ctl) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    const Ctl* c = ctl + inst;
    if (mgSkip(c)) return;
    const int C = m.nCells;
    const T* __restrict__ A = MatSel<T>::a(m);
    const T* __restrict__ dg = MatSel<T>::diag(m);
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    const bool act[3] = {c->active[0] != 0, c->active[1] != 0, c->active[2] != 0};
    const bool all = act[0] && act[1] && act[2];
    const T w = (T)omega;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        if (P >= cEnd) continue;
        const int e0 = m.sliceStart[s];
        const int width = m.sliceWidth[s];
        if (all && width == 6) {
            mg0DownRowAll<T, 6>(m, P, e0, lane, C, b, r, w);
            continue;
        }
        if (all && width == 4) {
            mg0DownRowAll<T, 4>(m, P, e0, lane, C, b, r, w);
            continue;
        }
        const T iD = invD[P];
        T xP[3], y[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            xP[j] = act[j] ? w * iD * b[(size_t)j * C + P] : (T)0;
            y[j] = dg[P] * xP[j];
        }
        for (int k = 0; k < width; ++k) {
            const int e = e0 + k * kSlice + lane;
            const int n = m.col[e];
            if (n >= 0) {
                const T a = A[e];
                const T iDn = invD[n];
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (act[j]) y[j] -= a * (w * iDn * b[(size_t)j * C + n]);
            }
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (!act[j]) continue;
            const size_t i = (size_t)j * C + P;
            r[i] = b[i] - y[j];
        }
    }
}

// Prolongation + post-smoothing on level 0.  With x = w D^-1 b the
// pre-smoothed value and r0 = b - A x from k_mg0_down, the corrected
// xf = x + alpha P e and its Jacobi sweep z = xf + w D^-1 (b - A xf) are
//   z = w D^-1 (b + r0) + alpha ((1 - w) e[agg] + w D^-1 sum_J q_J e_J)
// with q = offdiag(A) P grouped per coarse row J (host-built in the fine
// matrix's slots): no fine-matrix pass and no neighbour b, 1/diag or agg
// gathers.  e == nullptr: no coarse correction.  Partial z.b per block (b is
// the PCG residual: z.b = rho of the preconditioned residual).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg0_up(const DevMesh m, const BlockMap bm, const T* __restrict__ b,
                                                   const T* __restrict__ r0, const int* __restrict__ agg,
                                                   const int* __restrict__ qCol, const T* __restrict__ q,
                                                   const Row4<T>* __restrict__ e, T* __restrict__ z, double omega,
                                                   double alpha, const Ctl* __restrict__ ctl,
                                                   double* __restrict__ part) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    const Ctl* c = ctl + inst;
    if (mgSkip(c)) return;
    const int C = m.nCells;
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    const bool act[3] = {c->active[0] != 0, c->active[1] != 0, c->active[2] != 0};
    const bool all = act[0] && act[1] && act[2] && e != nullptr;
    const T w = (T)omega, w1 = (T)(1.0 - omega), al = (T)alpha;
    double acc[3] = {0, 0, 0};
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        if (P >= cEnd) continue;
        const int e0 = m.sliceStart[s];
        const int width = m.sliceWidth[s];
        if (all && width == 6) {
            mg0UpRowAll<T, 6>(m, qCol, q, P, e0, lane, C, b, r0, agg, e, z, w, w1, al, acc);
            continue;
        }
        if (all && width == 4) {
            mg0UpRowAll<T, 4>(m, qCol, q, P, e0, lane, C, b, r0, agg, e, z, w, w1, al, acc);
            continue;
        }
        const T wi = w * invD[P];
        T s3[3] = {(T)0, (T)0, (T)0};
        Row4<T> eP = {};
        if (e) {
            eP = e[agg[P]];
            for (int k = 0; k < width; ++k) {
                const int en = e0 + k * kSlice + lane;
                const int J = qCol[en];
                if (J >= 0) {
                    const Row4<T> ek = e[J];
                    const T cq = q[en];
#pragma unroll
                    for (int j = 0; j < 3; ++j) s3[j] += cq * ek[j];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (!act[j]) continue;
            const size_t i = (size_t)j * C + P;
            const T bv = b[i];
            const T zv = e ? wi * (bv + r0[i]) + al * (w1 * eP[j] + wi * s3[j]) : wi * (bv + r0[i]);
            z[i] = zv;
            acc[j] += (double)zv * (double)bv;
        }
    }
    storeBlockPartials<3>(acc, 0u, part, lb, bm.nBlocks);
}

// Prolongation + post-smoothing on level 0 of a smoothed hierarchy (the first
// prolongation P = (I - wp D^-1 A) P_aggregation).  The corrected
// xf = x + alpha P e and its Jacobi sweep reduce as in k_mg0_up to
//   z = w D^-1 (b + r0) + alpha w D^-1 sum_J G_J e_J,   G = (D / w) P - A P,
// with G host-built per row in its own sliced-ELL slots (an empty slot
// repeats the row's first column with a zero weight, so every load is
// unconditional and in bounds).  Partial z.b per block.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg0_up_g(const DevMesh m, const BlockMap bm, const T* __restrict__ b,
                                                     const T* __restrict__ r0, const int* __restrict__ gStart,
                                                     const int* __restrict__ gWidth, const int* __restrict__ gCol,
                                                     const T* __restrict__ g, const Row4<T>* __restrict__ e,
                                                     T* __restrict__ z, double omega, double alpha,
                                                     const Ctl* __restrict__ ctl, double* __restrict__ part) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    const Ctl* c = ctl + inst;
    if (mgSkip(c)) return;
    const int C = m.nCells;
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    const bool act[3] = {c->active[0] != 0, c->active[1] != 0, c->active[2] != 0};
    const T w = (T)omega, al = (T)alpha;
    double acc[3] = {0, 0, 0};
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        if (P >= cEnd) continue;
        const int e0 = gStart[s];
        const int width = gWidth[s];
        T sum[3] = {(T)0, (T)0, (T)0};
        for (int k0 = 0; k0 < width; k0 += kMgChunk) {
            int J[kMgChunk];
            T cw[kMgChunk];
#pragma unroll
            for (int u = 0; u < kMgChunk; ++u) {
                const int k = k0 + u < width ? k0 + u : width - 1;
                const int en = e0 + k * kSlice + lane;
                J[u] = ldS(gCol + en);
                cw[u] = k0 + u < width ? ldS(g + en) : (T)0;
            }
#pragma unroll
            for (int u = 0; u < kMgChunk; ++u) {
                const Row4<T> ev = e[J[u]];
#pragma unroll
                for (int j = 0; j < 3; ++j) sum[j] += cw[u] * ev[j];
            }
        }
        const T wi = w * invD[P];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (!act[j]) continue;
            const size_t i = (size_t)j * C + P;
            const T bv = b[i];
            const T zv = wi * (bv + r0[i]) + al * (wi * sum[j]);
            stS(z + i, zv);
            acc[j] += (double)zv * (double)bv;
        }
    }
    storeBlockPartials<3>(acc, 0u, part, lb, bm.nBlocks);
}

// ---------------------------------------------------------------------------
// coarse levels (grid-stride over the slices of one level)
// ---------------------------------------------------------------------------

// restriction: bc[I] = sum over members i of I of r[i] (r of the finer
// level: component-major on level 0 when FINE, rows of four otherwise; each
// member weighted when the prolongation into I is smoothed), and
// the coarse pre-smoothed value xc[I] = w D^-1 bc[I] (the coarse level's
// smoothing from zero, formed once here instead of at every use)
template <typename T, bool FINE>
__global__ __launch_bounds__(kBlock) void k_mg_restrict(const MgLevel Lc, const T* __restrict__ r0,
                                                        const Row4<T>* __restrict__ r4, int nFine,
                                                        Row4<T>* __restrict__ bc, Row4<T>* __restrict__ xc,
                                                        double omega, const Ctl* __restrict__ ctl) {
    const T* __restrict__ iDg = MgSel<T>::iD(Lc);
    const T* __restrict__ mw = MgSel<T>::memW(Lc);
    const T w = (T)omega;
    for (int I = logicalBlock(gridDim.x) * kBlock + threadIdx.x; I < Lc.n; I += gridDim.x * kBlock) {
        if (Lc.rowInst[I] < 0) continue;  // padding row
        if (mgSkip(ctl + Lc.rowInst[I])) continue;
        T s[3] = {(T)0, (T)0, (T)0};
        const int k1 = Lc.memOff[I + 1];
        // members 8 at a time, loads unconditional (past the end: the last
        // member again, weight zero), in member order
        for (int k0 = Lc.memOff[I]; k0 < k1; k0 += 8) {
            int mi[8];
            T wv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int k = k0 + u < k1 ? k0 + u : k1 - 1;
                mi[u] = Lc.mem[k];
                wv[u] = mw ? mw[k] : (T)1;
            }
            T rv[8][3];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (FINE) {
#pragma unroll
                    for (int j = 0; j < 3; ++j) rv[u][j] = r0[(size_t)j * nFine + mi[u]];
                } else {
                    const Row4<T> v = r4[mi[u]];
#pragma unroll
                    for (int j = 0; j < 3; ++j) rv[u][j] = v[j];
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (k0 + u < k1) s[j] += wv[u] * rv[u][j];
        }
        // all three components stored: a converged component's coarse
        // values are never read for an active one
        const T wi = w * iDg[I];
        bc[I] = row4(s[0], s[1], s[2]);
        xc[I] = row4(wi * s[0], wi * s[1], wi * s[2]);
    }
}

// Coarse-level kernels take their rows in the XCD-aware order of the level-0
// kernels (logicalBlock: each XCD a contiguous range of rows, so gathers of
// neighbouring rows and of a restriction's fine members stay in its L2).
// Coarse-level rows are wider (Galerkin stencils) and of varying width:
// slots are processed kMgChunk at a time with every load unconditional (slots
// past the row's width or empty read an in-bounds slot and the row's own
// value with a zero coefficient).

// restriction into the first coarse level of a smoothed hierarchy: bc = P^T r0
// with the weighted member lists in the coarse level's sliced-ELL slots (one
// wave per 64 coarse rows, each slot one coalesced load of members and
// weights; members in ascending fine-row order), and xc = w D^-1 bc
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_restrict_w(const MgLevel Lc, const T* __restrict__ r0, int nFine,
                                                          Row4<T>* __restrict__ bc, Row4<T>* __restrict__ xc,
                                                          double omega, const Ctl* __restrict__ ctl) {
    const T* __restrict__ iDg = MgSel<T>::iD(Lc);
    const T* __restrict__ rw = MgSel<T>::rW(Lc);
    const T w = (T)omega;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wavesTotal = gridDim.x * (kBlock / 64);
    for (int s = logicalBlock(gridDim.x) * (kBlock / 64) + wave; s < Lc.nSlices; s += wavesTotal) {
        const int I = s * kSlice + lane;
        if (I >= Lc.n) continue;
        const int inst = Lc.rowInst[I];
        if (inst < 0) continue;  // padding row
        if (mgSkip(ctl + inst)) continue;
        const int e0 = Lc.rStart[s];
        const int width = Lc.rWidth[s];
        T sum[3] = {(T)0, (T)0, (T)0};
        for (int k0 = 0; k0 < width; k0 += kMgChunk) {
            int mi[kMgChunk];
            T wv[kMgChunk];
#pragma unroll
            for (int u = 0; u < kMgChunk; ++u) {
                const int k = k0 + u < width ? k0 + u : width - 1;
                const int en = e0 + k * kSlice + lane;
                mi[u] = ldS(Lc.rMem + en);
                wv[u] = k0 + u < width ? ldS(rw + en) : (T)0;
            }
#pragma unroll
            for (int u = 0; u < kMgChunk; ++u)
#pragma unroll
                for (int j = 0; j < 3; ++j) sum[j] += wv[u] * r0[(size_t)j * nFine + mi[u]];
        }
        const T wi = w * iDg[I];
        bc[I] = row4(sum[0], sum[1], sum[2]);
        xc[I] = row4(wi * sum[0], wi * sum[1], wi * sum[2]);
    }
}

// down on a coarse level: r = b - A x with x = w D^-1 b from the restriction
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_down(const MgLevel L, const Row4<T>* __restrict__ b,
                                                    const Row4<T>* __restrict__ x, Row4<T>* __restrict__ r,
                                                    const Ctl* __restrict__ ctl) {
    const T* __restrict__ A = MgSel<T>::a(L);
    const T* __restrict__ dg = MgSel<T>::d(L);
    const int n = L.n;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wavesTotal = gridDim.x * (kBlock / 64);
    for (int s = logicalBlock(gridDim.x) * (kBlock / 64) + wave; s < L.nSlices; s += wavesTotal) {
        const int P = s * kSlice + lane;
        if (P >= n) continue;
        const int inst = L.rowInst[P];
        if (inst < 0) continue;  // padding row
        if (mgSkip(ctl + inst)) continue;
        const Row4<T> bP = b[P];
        if (L.pass[inst]) {  // already coarsest for this instance: hand b down unchanged
            r[P] = bP;
            continue;
        }
        const int e0 = L.sliceStart[s];
        const int width = L.sliceWidth[s];
        const Row4<T> xP = x[P];
        const T d = dg[P];
        T y[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) y[j] = d * xP[j];
        for (int k0 = 0; k0 < width; k0 += kMgChunk) {
            int nb[kMgChunk];
            T a[kMgChunk];
#pragma unroll
            for (int u = 0; u < kMgChunk; ++u) {
                const int k = k0 + u < width ? k0 + u : width - 1;
                const int e = e0 + k * kSlice + lane;
                nb[u] = k0 + u < width ? ldS(L.col + e) : -1;
                a[u] = ldS(A + e);
            }
#pragma unroll
            for (int u = 0; u < kMgChunk; ++u) {
                const bool v = nb[u] >= 0;
                const Row4<T> xc = x[v ? nb[u] : P];
                const T av = v ? a[u] : (T)0;
#pragma unroll
                for (int j = 0; j < 3; ++j) y[j] -= av * xc[j];
            }
        }
        r[P] = row4(bP[0] - y[0], bP[1] - y[1], bP[2] - y[2]);
    }
}

// up on a coarse level: the corrected xf = x + alpha e[agg] and its Jacobi
// sweep out = xf + w D^-1 (b - A xf), formed as on level 0 from the down
// sweep's residual r = b - A x and q = offdiag(A) P:
//   out = x + w D^-1 r + alpha ((1 - w) e[agg] + w D^-1 sum_J q_J e_J)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_up(const MgLevel L, const Row4<T>* __restrict__ x,
                                                  const Row4<T>* __restrict__ r, const Row4<T>* __restrict__ e,
                                                  Row4<T>* __restrict__ out, double omega, double alpha,
                                                  const Ctl* __restrict__ ctl) {
    const T* __restrict__ q = MgSel<T>::q(L);
    const T* __restrict__ iDg = MgSel<T>::iD(L);
    const T w = (T)omega, w1 = (T)(1.0 - omega), al = (T)alpha;
    const int n = L.n;
    const int* __restrict__ agg = L.agg;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wavesTotal = gridDim.x * (kBlock / 64);
    for (int s = logicalBlock(gridDim.x) * (kBlock / 64) + wave; s < L.nSlices; s += wavesTotal) {
        const int P = s * kSlice + lane;
        if (P >= n) continue;
        const int inst = L.rowInst[P];
        if (inst < 0) continue;  // padding row
        if (mgSkip(ctl + inst)) continue;
        const int IP = agg[P];
        const Row4<T> eP = e[IP];
        if (L.pass[inst]) {  // already coarsest for this instance: the coarser solution unchanged
            out[P] = eP;
            continue;
        }
        const int e0 = L.sliceStart[s];
        const int width = L.qWidth[s];
        T sum[3] = {(T)0, (T)0, (T)0};
        for (int k0 = 0; k0 < width; k0 += kMgChunk) {
            int J[kMgChunk];
            T c[kMgChunk];
#pragma unroll
            for (int u = 0; u < kMgChunk; ++u) {
                const int k = k0 + u < width ? k0 + u : width - 1;
                const int en = e0 + k * kSlice + lane;
                J[u] = k0 + u < width ? ldS(L.qCol + en) : -1;
                c[u] = ldS(q + en);
            }
#pragma unroll
            for (int u = 0; u < kMgChunk; ++u) {
                const bool v = J[u] >= 0;
                const Row4<T> ev = e[v ? J[u] : IP];
                const T cv = v ? c[u] : (T)0;
#pragma unroll
                for (int j = 0; j < 3; ++j) sum[j] += cv * ev[j];
            }
        }
        const Row4<T> xv = x[P], rv = r[P];
        const T wi = w * iDg[P];
        out[P] = row4(xv[0] + wi * rv[0] + al * (w1 * eP[0] + wi * sum[0]),
                      xv[1] + wi * rv[1] + al * (w1 * eP[1] + wi * sum[1]),
                      xv[2] + wi * rv[2] + al * (w1 * eP[2] + wi * sum[2]));
    }
}

// coarsest level: x = A^-1 b with the explicit (symmetric) inverse of the
// row's instance block, one wavefront per row (each lane's terms in column
// order, then the wavefront reduction)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mg_dense(const MgLevel L, const T* __restrict__ Ainv,
                                                     const Row4<T>* __restrict__ b, Row4<T>* __restrict__ x,
                                                     const Ctl* __restrict__ ctl) {
    const int row = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= L.n) return;
    const int inst = L.rowInst[row];
    if (inst < 0) return;  // padding row
    if (mgSkip(ctl + inst)) return;
    const int r0 = L.blkRow0[inst], nb = L.blkN[inst];
    double acc[3] = {0, 0, 0};
    const T* __restrict__ a = Ainv + (size_t)L.invOff[inst] + (size_t)(row - r0) * nb;
#pragma unroll 4
    for (int k = lane; k < nb; k += 64) {
        const double v = (double)a[k];
        const Row4<T> bk = b[r0 + k];
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[j] += v * (double)bk[j];
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[j] = waveReduce<false>(acc[j]);
    if (lane == 0) x[row] = row4((T)acc[0], (T)acc[1], (T)acc[2]);
}

// rho' = z.r finalize for a V-cycle-preconditioned PCG: beta = rho'/rho
// (first iteration: no beta), rho = rho'
__global__ __launch_bounds__(kFinBlock) void k_fin_rho(const BlockMap bm, Ctl* __restrict__ ctlAll,
                                                       const double* __restrict__ part) {
    const int inst = bm.finInst ? bm.finInst[blockIdx.x] : (int)blockIdx.x;
    Ctl* ctl = ctlAll + inst;
    if (ctl->outerDone || ctl->innerDone) return;
    __shared__ double tot[3];
    finReduce<3>(part, bm.nBlocks, bm.iBlk0[inst], bm.iNBlk[inst], 0u, tot);
    if (threadIdx.x == 0) {
        for (int j = 0; j < 3; ++j) {
            if (!ctl->active[j]) continue;
            ctl->beta[j] = ctl->firstInner ? 0.0 : tot[j] / ctl->rho[j];
            ctl->rho[j] = tot[j];
        }
    }
}

// ---------------------------------------------------------------------------
// packing of per-mesh hierarchies into one block-diagonal hierarchy (server)
// ---------------------------------------------------------------------------

// int segment: mode 0 dst = src + add; 1 dst = src >= 0 ? src + add : src;
// 2 dst = add + k; 3 dst = add
struct SegInt {
    const int* src;
    int* dst;
    int n, add, mode, pad;
};
// 32-bit word copy segment
struct SegWords {
    const unsigned* src;
    unsigned* dst;
    long long n;
};

__global__ __launch_bounds__(kBlock) void k_seg_int(const SegInt* __restrict__ segs) {
    const SegInt g = segs[blockIdx.y];
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < g.n; k += gridDim.x * kBlock) {
        int v;
        if (g.mode == 0) v = g.src[k] + g.add;
        else if (g.mode == 1) { v = g.src[k]; if (v >= 0) v += g.add; }
        else if (g.mode == 2) v = g.add + k;
        else v = g.add;
        g.dst[k] = v;
    }
}

__global__ __launch_bounds__(kBlock) void k_seg_words(const SegWords* __restrict__ segs) {
    const SegWords g = segs[blockIdx.y];
    for (long long k = (long long)blockIdx.x * kBlock + threadIdx.x; k < g.n; k += (long long)gridDim.x * kBlock)
        g.dst[k] = g.src[k];
}

}  // namespace s4f
