// s4f_kernels.h -- HIP kernels of the segregated small-strain solid solver.
//
// All kernels walk the sliced-ELL cell->face lists: wave w of a workgroup
// owns one 64-row slice at a time, lane l owns one cell, and each slot of the
// slice is one coalesced load per array.  Face contributions are gathered
// per cell in OpenFOAM's face order, so no atomics touch field data and the
// summation order is fixed (run-to-run bitwise deterministic).  Reductions
// are per workgroup (wave shuffles + LDS) into a partial array; a one-block-
// per-instance finalize kernel then reduces the instance's partials in fixed
// order and updates the control block (kernel boundaries order the writes of
// every XCD before the finalize reads them).
//
// Kernel <-> reference map (paths relative to the solids4foam tree):
//   k_residual   linGeomTotalDispSolid.C:141-149: fvc::div(sigma) [Gauss linear]
//                + momentumStabilisation.C:210-217 (RhieChow) + the normalisation
//                of OpenFOAM lduMatrix::solver::normFactor for the PCG initial residual
//   k_spmvp      OpenFOAM PCG search-direction update p = z + beta p fused with
//                lduMatrix::Amul (3 right-hand sides sharing M)
//   k_update     OpenFOAM PCG solution/residual update + diagonal preconditioner
//   k_fin_*      the scalar parts of PCG::solve / normFactor / solidModel::converged
//   k_grad       solidTractionFvPatchVectorField.C:398-463 / fixedDisplacement
//                boundary values, leastSquares grad(D) (mechanicalModel.C:571-582),
//                gaussGrad boundary correction, linearElastic::correct on
//                boundary faces (linearElastic.C:318-339), solidTraction
//                updateCoeffs (:321-395 + linGeomTotalDispSolid.C:235-271) and
//                solidModel::converged (solidModelTemplates.C:26-188)
#pragma once

#include "s4f_device.h"

namespace s4f {

struct BlockMap {
    const int* inst;   // instance of each logical block
    const int* s0;     // first slice
    const int* s1;     // one past last slice
    const int* iBlk0;  // per instance: first logical block
    const int* iNBlk;  // per instance: number of blocks
    const int* iCell0; // per instance: first cell
    const int* iNCell; // per instance: number of cells
    const int* finInst;// instance finalized by finalize-kernel workgroup i (nullptr: i itself)
    const int* iB0;    // per instance: first boundary face
    const int* iNB;    // per instance: number of boundary faces
    int nBlocks;
};

// Workgroups are dispatched round-robin over the 8 XCDs (hardware block b runs
// on XCD b % 8).  Give XCD x a contiguous run of logical blocks so that
// neighbouring cells (and their gathers) stay in one XCD's L2.
__device__ __forceinline__ int logicalBlock(int nb) {
    const int b = blockIdx.x;
    const int x = b & 7, q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// NaN-propagating maximum (fmax would drop a NaN and hide a diverged solve)
__device__ __forceinline__ double nanMax(double a, double b) { return (b > a || b != b) ? b : a; }

template <bool MAX>
__device__ __forceinline__ double waveReduce(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off, 64);
        v = MAX ? nanMax(v, o) : v + o;
    }
    return v;
}

constexpr int kFinBlock = 256;

// Block partials are stored value-major: value i of logical block lb sits at
// part[i * nb + lb] (nb = the launch's logical block count), so a finalize
// reads each value's partials as one contiguous run instead of one 128-byte
// row per block.
//
// Fixed-order reduction of the instance's partials (one workgroup per
// instance); bit i of maxMask selects a maximum instead of a sum.  Result in
// out[0..NV) (shared), valid for every thread after the call.
template <int NV>
__device__ void finReduce(const double* __restrict__ part, int nb, int blk0, int nblk, unsigned maxMask, double* out) {
    double v[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = 0.0;
    for (int b = threadIdx.x; b < nblk; b += kFinBlock) {
        const double* col = part + blk0 + b;
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const double x = col[(size_t)i * nb];
            v[i] = ((maxMask >> i) & 1u) ? nanMax(v[i], x) : v[i] + x;
        }
    }
    __shared__ double sh[NV][kFinBlock / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        v[i] = ((maxMask >> i) & 1u) ? waveReduce<true>(v[i]) : waveReduce<false>(v[i]);
        if (l == 0) sh[i][w] = v[i];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int i = threadIdx.x;
        double r = sh[i][0];
        for (int k = 1; k < kFinBlock / 64; ++k) r = ((maxMask >> i) & 1u) ? nanMax(r, sh[i][k]) : r + sh[i][k];
        out[i] = r;
    }
    __syncthreads();
}

// block-level reduction of NV per-thread values; threads 0..NV-1 store them
// at part[i * nb + lb]
template <int NV>
__device__ __forceinline__ void storeBlockPartials(const double* acc, unsigned maxMask, double* part, int lb, int nb) {
    __shared__ double sh[NV][kBlock / 64];
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double v = ((maxMask >> i) & 1u) ? waveReduce<true>(acc[i]) : waveReduce<false>(acc[i]);
        if (l == 0) sh[i][w] = v;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int i = threadIdx.x;
        double r = sh[i][0];
        for (int k = 1; k < kBlock / 64; ++k) r = ((maxMask >> i) & 1u) ? nanMax(r, sh[i][k]) : r + sh[i][k];
        part[(size_t)i * nb + lb] = r;
    }
}

// linearElastic: sigma = 2 mu dev(symm G) + K tr(symm G) I   (xx xy xz yy yz zz)
__device__ __forceinline__ void hooke(const double* G, double mu, double K, double* s) {
    const double e0 = G[0], e1 = 0.5 * (G[1] + G[3]), e2 = 0.5 * (G[2] + G[6]);
    const double e3 = G[4], e4 = 0.5 * (G[5] + G[7]), e5 = G[8];
    const double tr = e0 + e3 + e5;
    const double sh = K * tr;
    const double t3 = tr / 3.0;
    s[0] = 2.0 * mu * (e0 - t3) + sh;
    s[1] = 2.0 * mu * e1;
    s[2] = 2.0 * mu * e2;
    s[3] = 2.0 * mu * (e3 - t3) + sh;
    s[4] = 2.0 * mu * e4;
    s[5] = 2.0 * mu * (e5 - t3) + sh;
}

__device__ __forceinline__ double cellMu(const DevMesh& m, int c) { return m.mu ? m.mu[c] : m.mu0; }
__device__ __forceinline__ double cellK(const DevMesh& m, int c) { return m.K ? m.K[c] : m.K0; }
__device__ __forceinline__ double cellImpK(const DevMesh& m, int c) { return m.impK ? m.impK[c] : m.impK0; }

template <typename T> struct MatSel;
template <> struct MatSel<double> {
    __device__ static const double* a(const DevMesh& m) { return m.eA; }
    __device__ static const double* diag(const DevMesh& m) { return m.diag; }
    __device__ static const double* invDiag(const DevMesh& m) { return m.invDiag; }
};
template <> struct MatSel<float> {
    __device__ static const float* a(const DevMesh& m) { return m.eA32; }
    __device__ static const float* diag(const DevMesh& m) { return m.diag32; }
    __device__ static const float* invDiag(const DevMesh& m) { return m.invDiag32; }
};

// Internal-face slots of a k_residual row with compile-time slot count W,
// branch-free: a boundary or empty slot gathers the row's own cell and has
// zero area, Rhie-Chow and matrix coefficients, so it adds exactly zero.
// Rows list internal faces before boundary faces, so running the boundary
// slots afterwards keeps the summation order of the slot loop.
__device__ __forceinline__ void residualInternal(const DevMesh& m, int P, int e0, int W, int lane, int C,
                                                  const double* __restrict__ D, const double* __restrict__ G,
                                                  const double* DP, const double* GP, const double* CP,
                                                  const double* sP, double sf, double* div, double* rc,
                                                  double* off, double& asum, bool& hasB) {
    const size_t E = m.nEntries;
#pragma unroll 1
    for (int k = 0; k < W; ++k) {
        const int e = e0 + k * kSlice + lane;
        const int n = ldS(m.col + e);
        const bool v = n >= 0;
        hasB |= (!v && n != kPad);
        const int c = v ? n : P;
        const double S0r = ldS(m.eS + e), S1r = ldS(m.eS + E + e), S2r = ldS(m.eS + 2 * E + e);
        const double w = ldS(m.eW + e), cfr = ldS(m.eC + e), ar = ldS(m.eA + e);
        const double S0 = v ? S0r : 0.0, S1 = v ? S1r : 0.0, S2 = v ? S2r : 0.0;
        const double cf = v ? cfr : 0.0, a = v ? ar : 0.0;
        double GN[9], sN[6], DN[3], d[3];
#pragma unroll
        for (int q = 0; q < 9; ++q) GN[q] = G[(size_t)q * C + c];
#pragma unroll
        for (int q = 0; q < 3; ++q) { DN[q] = D[(size_t)q * C + c]; d[q] = m.cC[(size_t)q * C + c] - CP[q]; }
        hooke(GN, cellMu(m, c) * sf, cellK(m, c) * sf, sN);
        double sfc[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) sfc[q] = w * (sP[q] - sN[q]) + sN[q];
        div[0] += S0 * sfc[0] + S1 * sfc[1] + S2 * sfc[2];
        div[1] += S0 * sfc[1] + S1 * sfc[3] + S2 * sfc[4];
        div[2] += S0 * sfc[2] + S1 * sfc[4] + S2 * sfc[5];
        double Gf[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) Gf[q] = w * (GP[q] - GN[q]) + GN[q];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double dG = d[0] * Gf[j] + d[1] * Gf[3 + j] + d[2] * Gf[6 + j];
            rc[j] += cf * ((DN[j] - DP[j]) - dG);
        }
        asum += a;
#pragma unroll
        for (int j = 0; j < 3; ++j) off[j] += a * DN[j];
    }
}

// ---------------------------------------------------------------------------
// k_residual: R = div(sigma) + RhieChow (integrated), r = R, and per-workgroup
// partials of sum|R|, the two normFactor sums and z.r (k_fin_residual).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_residual(const DevMesh m, const BlockMap bm, const double* __restrict__ D0,
                                                     const double* __restrict__ D1, const double* __restrict__ G, const double* __restrict__ sigb,
                                                     T* __restrict__ r, Ctl* __restrict__ ctlAll,
                                                     double* __restrict__ part) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    Ctl* ctl = ctlAll + inst;
    if (ctl->outerDone) return;
    const double* __restrict__ D = ctl->dpar ? D1 : D0;
    const int C = m.nCells;
    const double xb0 = ctl->xbar[0], xb1 = ctl->xbar[1], xb2 = ctl->xbar[2];
    // a cold start's first residual sees the zero initial stress field (the
    // constructor computes grad(D) but no sigma, linGeomTotalDispSolid.C:64-84)
    const double sf = (ctl->sigma0 && ctl->iCorr == 0) ? 0.0 : 1.0;
    double acc[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) acc[i] = 0.0;
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        const int e0 = m.sliceStart[s];
        const int width = m.sliceWidth[s];
        if (P >= cEnd) continue;
        double DP[3], GP[9], CP[3], sP[6];
#pragma unroll
        for (int k = 0; k < 3; ++k) { DP[k] = D[(size_t)k * C + P]; CP[k] = m.cC[(size_t)k * C + P]; }
#pragma unroll
        for (int k = 0; k < 9; ++k) GP[k] = G[(size_t)k * C + P];
        const double muP = cellMu(m, P) * sf, KP = cellK(m, P) * sf;
        hooke(GP, muP, KP, sP);
        double div[3] = {0, 0, 0}, rc[3] = {0, 0, 0}, off[3] = {0, 0, 0}, asum = 0.0;
        bool hasB = false;
        residualInternal(m, P, e0, width, lane, C, D, G, DP, GP, CP, sP, sf, div, rc, off, asum, hasB);
        for (int k = 0; hasB && k < width; ++k) {
            const int e = e0 + k * kSlice + lane;
            const int n = m.col[e];
            if (n < 0 && n != kPad) {
                const int b = -1 - n;
                const size_t E = m.nEntries;
                const double S0 = m.eS[e], S1 = m.eS[E + e], S2 = m.eS[2 * E + e];
                double sb[6];
#pragma unroll
                for (int q = 0; q < 6; ++q) sb[q] = sigb[(size_t)q * m.nBoundary + b];
                div[0] += S0 * sb[0] + S1 * sb[1] + S2 * sb[2];
                div[1] += S0 * sb[1] + S1 * sb[3] + S2 * sb[4];
                div[2] += S0 * sb[2] + S1 * sb[4] + S2 * sb[5];
            }
        }
        const double dg = m.diag[P];
        const double rowsum = dg - asum;
        const double xb[3] = {xb0, xb1, xb2};
        // body force and the inertia terms (rho g V + source - diag D)
        double src[3] = {0.0, 0.0, 0.0};
        if (m.fext) {
#pragma unroll
            for (int j = 0; j < 3; ++j) src[j] = m.fext[(size_t)j * C + P] - (m.iner ? m.iner[P] * DP[j] : 0.0);
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double R = (div[j] + rc[j]) + src[j];
            const double MD = dg * DP[j] - off[j];
            const double t = rowsum * xb[j];
            acc[j] += fabs(R);
            acc[3 + j] += fabs(MD - t);
            acc[6 + j] += fabs((MD + R) - t);
            const T rT = (T)R;
            stS(r + (size_t)j * C + P, rT);
            const T z = invD[P] * rT;
            acc[9 + j] += (double)z * (double)rT;
        }
    }
    storeBlockPartials<12>(acc, 0u, part, lb, bm.nBlocks);
}

// Row of the fused direction update + SpMV with all three components active
// and a compile-time slot count W: every load of the row is issued without a
// branch (invalid slots gather the row's own cell and contribute zero), so
// the wave keeps all of its gathers in flight at once.
template <typename T, int W, bool MG>
__device__ __forceinline__ void spmvpRowAll(const DevMesh& m, int P, int e0, int lane, int C, const T* __restrict__ r,
                                            const T* __restrict__ pOld, const T* __restrict__ zArr,
                                            T* __restrict__ pNew, T* __restrict__ q, int first, const T* be,
                                            double* acc) {
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
    auto dir = [&](int j, int c, T iDc) -> T {
        const size_t i = (size_t)j * C + c;
        const T z = MG ? zArr[i] : iDc * r[i];
        const T d = z + be[j] * pOld[i];
        return first ? z : d;
    };
    const T iDP = MG ? (T)0 : invD[P];
    T pP[3], y[3];
    const T dP = MatSel<T>::diag(m)[P];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        pP[j] = dir(j, P, iDP);
        y[j] = dP * pP[j];
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const bool v = nb[k] >= 0;
        const int c = v ? nb[k] : P;
        const T av = v ? a[k] : (T)0;
        const T iDc = MG ? (T)0 : invD[c];
#pragma unroll
        for (int j = 0; j < 3; ++j) y[j] -= av * dir(j, c, iDc);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const size_t i = (size_t)j * C + P;
        stS(pNew + i, pP[j]);
        stS(q + i, y[j]);
        acc[j] += (double)pP[j] * (double)y[j];
    }
}

// ---------------------------------------------------------------------------
// k_spmvp: p = z + beta p (z = D^-1 r, or the multigrid output when MG),
// q = M p for the three components; partial p.q per component.  With
// reduce = 0 it is a plain q = M p (s4f_spmv3).
// ---------------------------------------------------------------------------
template <typename T, bool MG>
__global__ __launch_bounds__(kBlock) void k_spmvp(const DevMesh m, const BlockMap bm, const T* __restrict__ r,
                                                  const T* __restrict__ pOld, T* __restrict__ pNew, T* __restrict__ q,
                                                  Ctl* __restrict__ ctlAll, double* __restrict__ part, int reduce,
                                                  const T* __restrict__ zArr) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    const Ctl* ctl = ctlAll + inst;
    if (reduce && (ctl->outerDone || ctl->innerDone)) return;
    const int C = m.nCells;
    const T* __restrict__ A = MatSel<T>::a(m);
    const T* __restrict__ dg = MatSel<T>::diag(m);
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    const int first = reduce ? ctl->firstInner : 0;
    const bool act[3] = {reduce ? ctl->active[0] != 0 : true, reduce ? ctl->active[1] != 0 : true,
                         reduce ? ctl->active[2] != 0 : true};
    const bool all = reduce && act[0] && act[1] && act[2];
    const T be[3] = {reduce ? (T)ctl->beta[0] : (T)0, reduce ? (T)ctl->beta[1] : (T)0, reduce ? (T)ctl->beta[2] : (T)0};
    // direction component of cell c
    auto dir = [&](int j, int c) -> T {
        const size_t i = (size_t)j * C + c;
        if (!reduce) return pOld[i];
        const T z = MG ? zArr[i] : invD[c] * r[i];
        return first ? z : z + be[j] * pOld[i];
    };
    double acc[3] = {0, 0, 0};
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        const int e0 = m.sliceStart[s];
        const int width = m.sliceWidth[s];
        if (P >= cEnd) continue;
        if (all && width == 6) {
            spmvpRowAll<T, 6, MG>(m, P, e0, lane, C, r, pOld, zArr, pNew, q, first, be, acc);
            continue;
        }
        if (all && width == 4) {
            spmvpRowAll<T, 4, MG>(m, P, e0, lane, C, r, pOld, zArr, pNew, q, first, be, acc);
            continue;
        }
        T pP[3], y[3];
        const T d = dg[P];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            pP[j] = act[j] ? dir(j, P) : (T)0;
            y[j] = d * pP[j];
        }
        for (int k = 0; k < width; ++k) {
            const int e = e0 + k * kSlice + lane;
            const int n = m.col[e];
            if (n >= 0) {
                const T a = A[e];
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    if (act[j]) y[j] -= a * dir(j, n);
            }
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            if (!act[j]) continue;
            const size_t i = (size_t)j * C + P;
            if (reduce) pNew[i] = pP[j];
            q[i] = y[j];
            acc[j] += (double)pP[j] * (double)y[j];
        }
    }
    if (!reduce) return;
    storeBlockPartials<3>(acc, 0u, part, lb, bm.nBlocks);
}

// ---------------------------------------------------------------------------
// k_update: x += alpha p, r -= alpha q, z = r/diag; partials z.r and sum|r|
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_update(const DevMesh m, const BlockMap bm, T* __restrict__ x,
                                                   const T* __restrict__ p, T* __restrict__ r, const T* __restrict__ q,
                                                   const Ctl* __restrict__ ctlAll, double* __restrict__ part) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    const Ctl* ctl = ctlAll + inst;
    if (ctl->outerDone || ctl->innerDone) return;
    const int C = m.nCells;
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    const T al[3] = {(T)ctl->alpha[0], (T)ctl->alpha[1], (T)ctl->alpha[2]};
    const bool act[3] = {ctl->active[0] != 0, ctl->active[1] != 0, ctl->active[2] != 0};
    const int first = ctl->firstInner;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        if (P >= cEnd) continue;
        const T iD = invD[P];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const size_t i = (size_t)j * C + P;
            if (!act[j]) {
                if (first) x[i] = (T)0;
                continue;
            }
            const T xv = first ? (T)0 : x[i];
            x[i] = xv + al[j] * p[i];
            const T rv = r[i] - al[j] * q[i];
            r[i] = rv;
            const T z = iD * rv;
            acc[j] += (double)z * (double)rv;
            acc[3 + j] += fabs((double)rv);
        }
    }
    storeBlockPartials<6>(acc, 0u, part, lb, bm.nBlocks);
}

// Internal-face slots of the least-squares gradient sum with compile-time
// slot count W, branch-free (boundary and empty slots gather the row's own
// cell with zero weight); boundary slots follow in the general loop, in slot
// order, as rows list internal faces first.
template <typename T, int W, bool HASX>
__device__ __forceinline__ void gradInternalW(const DevMesh& m, int P, int e0, int lane, int C,
                                              const double* __restrict__ D, const T* __restrict__ x,
                                              double ra, const double* DnP, const double* CP, double* A,
                                              bool& hasB) {
    // two faces at a time: enough gathers in flight at four waves per SIMD
#pragma unroll 2
    for (int k = 0; k < W; ++k) {
        const int e = e0 + k * kSlice + lane;
        const int n = ldS(m.col + e);
        const bool v = n >= 0;
        hasB |= (!v && n != kPad);
        const int c = v ? n : P;
        double d[3], dv[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            d[q] = m.cC[(size_t)q * C + c] - CP[q];
            const double DN = HASX ? D[(size_t)q * C + c] + ra * (double)x[(size_t)q * C + c] : D[(size_t)q * C + c];
            dv[q] = DN - DnP[q];
        }
        const double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double rr = v ? 1.0 / dd : 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double di = d[i] * rr;
#pragma unroll
            for (int j = 0; j < 3; ++j) A[3 * i + j] += di * dv[j];
        }
    }
}

// ---------------------------------------------------------------------------
// The gradient step of an outer iteration, in three launches:
//   k_grad_bnd_pre   boundary faces: D' = D + x at the face cell, boundary
//                    values D_b (solidTraction evaluate / fixedDisplacement /
//                    solidSymmetry) and the snGrad used by the residual
//   k_grad           cells: least-squares grad(D') with the boundary values,
//                    D' stored, partials of the outer convergence test
//   k_grad_bnd_post  boundary faces: boundary gradient and stress, the next
//                    traction snGrad
// mode 0: outer iteration; mode 1: initialisation (x = 0, G_old = 0, g = 0,
// no convergence test); mode 2: gradient from given D_b/snb (no boundary
// update: k_grad_bnd_pre is not launched, k_grad_bnd_post leaves g).
// The face kernels run one workgroup row (blockIdx.y) per instance over that
// instance's boundary faces; empty faces carry no value and are skipped.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int faceInst(const BlockMap& bm) {
    return bm.finInst ? bm.finInst[blockIdx.y] : (int)blockIdx.y;
}

// Aitken factor update of relaxField (solidModel.C:876-892) at one cell or
// boundary face: r = D_prev - D*, alpha = a0 on the first outer iteration,
// else clip(alpha (r_prev . (r_prev - r)) / (|r_prev - r|^2 + SMALL), 0, 2);
// returns alpha and stores it with r (the next r_prev)
__device__ __forceinline__ double aitkenFactor(const double* r, double* alpha, double* res, size_t i, size_t n,
                                               bool first, double a0) {
    double al = a0;
    if (!first) {
        const double rp[3] = {res[i], res[n + i], res[2 * n + i]};
        const double d[3] = {rp[0] - r[0], rp[1] - r[1], rp[2] - r[2]};
        al = alpha[i] * (rp[0] * d[0] + rp[1] * d[1] + rp[2] * d[2]) / ((d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) + kSmall);
        al = al > 2.0 ? 2.0 : (al < 0.0 ? 0.0 : al);
    }
    alpha[i] = al;
    res[i] = r[0];
    res[n + i] = r[1];
    res[2 * n + i] = r[2];
    return al;
}

// relaxField, Aitken (solidModel.C:842-897), on the cells: D' = D* - alpha
// (D_prev - D*) with D* = D + x the solved field, into the buffer receiving D'
template <typename T>
__global__ __launch_bounds__(kBlock) void k_aitken(const DevMesh m, const BlockMap bm, double* __restrict__ D0,
                                                   const T* __restrict__ x, double* __restrict__ D1,
                                                   double* __restrict__ alpha, double* __restrict__ res,
                                                   const Ctl* __restrict__ ctlAll) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    const Ctl* ctl = ctlAll + inst;
    if (ctl->outerDone) return;
    if (ctl->firstInner) x = nullptr;
    const double* D = ctl->dpar ? D1 : D0;
    double* Dn = ctl->dpar ? D0 : D1;
    const bool first = ctl->iCorr == 0;
    const double a0 = ctl->relaxD;
    const int C = m.nCells;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        if (P >= cEnd) continue;
        double Ds[3], r[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double DP = D[(size_t)k * C + P];
            Ds[k] = DP + (x ? (double)x[(size_t)k * C + P] : 0.0);
            r[k] = DP - Ds[k];
        }
        const double al = aitkenFactor(r, alpha, res, P, C, first, a0);
#pragma unroll
        for (int k = 0; k < 3; ++k) Dn[(size_t)k * C + P] = Ds[k] - al * r[k];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_grad_bnd_pre(const DevMesh m, const BlockMap bm,
                                                         const double* __restrict__ D0, const T* __restrict__ x,
                                                         const double* __restrict__ D1, const double* __restrict__ G,
                                                         double* __restrict__ Db, const double* __restrict__ g,
                                                         double* __restrict__ snb, const Ctl* __restrict__ ctlAll,
                                                         double* __restrict__ alphaB, double* __restrict__ resB,
                                                         int mode) {
    const int inst = faceInst(bm);
    const Ctl* ctl = ctlAll + inst;
    if (mode == 0 && ctl->outerDone) return;
    if (mode == 0 && ctl->firstInner) x = nullptr;  // no PCG iteration ran: the update is zero
    const double* __restrict__ D = ctl->dpar ? D1 : D0;
    // Aitken: the relaxed cell values are in the buffer receiving D' (k_aitken)
    const bool aitken = mode == 0 && ctl->relaxAitken;
    const double* __restrict__ Dr = ctl->dpar ? D0 : D1;
    const bool relax = mode == 0 && !aitken && ctl->relaxD != 1.0;
    const double ra = relax ? ctl->relaxD : 1.0;
    const bool first = ctl->iCorr == 0;
    const int C = m.nCells, B = m.nBoundary;
    const int b0 = bm.iB0[inst], b1 = b0 + bm.iNB[inst];
    for (int b = b0 + blockIdx.x * kBlock + threadIdx.x; b < b1; b += gridDim.x * kBlock) {
        const int type = m.bType[b];
        if (type == 2) continue;
        const int P = m.bCell[b];
        // DuP: the solve's update D + x; DnP: after the fixed field
        // under-relaxation D_prev + a (DuP - D_prev) (a = 1: the same)
        double DuP[3], DnP[3], Gold[9];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double DP = D[(size_t)k * C + P];
            const double xv = x ? (double)x[(size_t)k * C + P] : 0.0;
            DuP[k] = DP + xv;
            DnP[k] = aitken ? Dr[(size_t)k * C + P] : DP + ra * xv;
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) Gold[q] = (mode == 1) ? 0.0 : G[(size_t)q * C + P];
        const double k0 = m.bKv[b], k1 = m.bKv[(size_t)B + b], k2 = m.bKv[2 * (size_t)B + b];
        const double dl = m.bDelta[b];
        const bool trac = type == 1;
        const bool sym = type == 3;
        const bool zs = type == 4;
        // solidSymmetry (first order, normal patch delta so no k correction):
        // value D_P - n (n.D_P) after the solve (relaxed like every boundary
        // value), snGrad -n (n.D_P) deltaCoeffs from the relaxed cell value
        const double nb0 = m.bN[b], nb1 = m.bN[(size_t)B + b], nb2 = m.bN[2 * (size_t)B + b];
        const double nDu = nb0 * DuP[0] + nb1 * DuP[1] + nb2 * DuP[2];
        const double nDn = nb0 * DnP[0] + nb1 * DnP[1] + nb2 * DnP[2];
        const double nbv[3] = {nb0, nb1, nb2};
        double kGv[3], gvv[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            kGv[j] = k0 * Gold[j] + k1 * Gold[3 + j] + k2 * Gold[6 + j];
            gvv[j] = (mode == 1 || !(trac || zs)) ? 0.0 : g[(size_t)j * B + b];
        }
        // fixedDisplacementZeroShear (direction-mixed, valueFraction nn, first
        // order): n (n.Dbar) + (I - nn).(D_P + k.G_P + g/deltaCoeffs), the value
        // from the solved cell value, the snGrad from the relaxed one
        double zu[3] = {0, 0, 0}, zn[3] = {0, 0, 0};
        if (zs) {
            double gu[3], gn[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                gu[j] = DuP[j] + kGv[j] + gvv[j] / dl;
                gn[j] = DnP[j] + kGv[j] + gvv[j] / dl;
            }
            const double nD = nb0 * m.bVal[b] + nb1 * m.bVal[(size_t)B + b] + nb2 * m.bVal[2 * (size_t)B + b];
            const double ngu = nb0 * gu[0] + nb1 * gu[1] + nb2 * gu[2];
            const double ngn = nb0 * gn[0] + nb1 * gn[1] + nb2 * gn[2];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                zu[j] = nbv[j] * nD + (gu[j] - nbv[j] * ngu);
                zn[j] = nbv[j] * nD + (gn[j] - nbv[j] * ngn);
            }
        }
        // correctBoundaryConditions after the solve, then relaxField on the
        // patch values (fixed values stay); the snGrads follow the relaxed cell
        double Dbv[3], sn[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double kG = kGv[j];
            if (zs) {
                Dbv[j] = zu[j];
                sn[j] = (zn[j] - (DnP[j] + kG)) * dl;
            } else if (sym) {
                Dbv[j] = DuP[j] - nbv[j] * nDu;
                sn[j] = -nbv[j] * nDn * dl;
            } else if (trac) {
                Dbv[j] = DuP[j] + kG + gvv[j] / dl;
                sn[j] = gvv[j];
            } else {
                Dbv[j] = m.bVal[(size_t)j * B + b];
                sn[j] = (Dbv[j] - (DnP[j] + kG)) * dl;
            }
        }
        if (type != 0 && (relax || aitken)) {
            double Dbp[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) Dbp[j] = Db[(size_t)j * B + b];
            if (aitken) {
                const double r[3] = {Dbp[0] - Dbv[0], Dbp[1] - Dbv[1], Dbp[2] - Dbv[2]};
                const double al = aitkenFactor(r, alphaB, resB, b, B, first, ctl->relaxD);
#pragma unroll
                for (int j = 0; j < 3; ++j) Dbv[j] -= al * r[j];
            } else {
#pragma unroll
                for (int j = 0; j < 3; ++j) Dbv[j] = Dbp[j] + ra * (Dbv[j] - Dbp[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Db[(size_t)j * B + b] = Dbv[j];
            snb[(size_t)j * B + b] = sn[j];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_grad(const DevMesh m, const BlockMap bm, double* __restrict__ D0,
                                                 const T* __restrict__ x, double* __restrict__ D1,
                                                 double* __restrict__ G, const double* __restrict__ Db,
                                                 const double* __restrict__ DI, const double* __restrict__ Dold,
                                                 Ctl* __restrict__ ctlAll, double* __restrict__ part, int mode) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    Ctl* ctl = ctlAll + inst;
    if (mode == 0 && ctl->outerDone) return;
    if (mode == 0 && ctl->firstInner) x = nullptr;  // no PCG iteration ran: the update is zero
    // current field and the buffer receiving D' (per-instance ping-pong)
    const double* __restrict__ D = ctl->dpar ? D1 : D0;
    double* Dn = (mode == 0) ? (ctl->dpar ? D0 : D1) : nullptr;
    // D' = D + a x (fixed relaxation, a = 1 without), or the Aitken-relaxed
    // field k_aitken already stored in Dn
    const bool aitken = mode == 0 && ctl->relaxAitken;
    const bool relax = mode == 0 && !aitken && ctl->relaxD != 1.0;
    const double ra = relax ? ctl->relaxD : 1.0;  // x scaled by exactly 1.0 without relaxation
    const double* Dg = aitken ? Dn : D;
    const T* xg = aitken ? nullptr : x;
    const int C = m.nCells, B = m.nBoundary;
    double acc[6] = {0, 0, 0, 0, 0, 0};  // max|D'-D|, max|D'-Dold|, max|D'|, sum D'
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        const int e0 = m.sliceStart[s];
        const int width = m.sliceWidth[s];
        if (P >= cEnd) continue;
        double DP[3], DnP[3], CP[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            DP[k] = D[(size_t)k * C + P];
            const double xv = xg ? (double)xg[(size_t)k * C + P] : 0.0;
            DnP[k] = aitken ? Dg[(size_t)k * C + P] : DP[k] + ra * xv;
            CP[k] = m.cC[(size_t)k * C + P];
        }
        double A[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) A[q] = 0.0;
        bool hasB = false;
        int kInt = width;  // internal slots left to the general loop below
        // a slice holding bi-material interface faces: every slot goes
        // through the general loop, which gives each side's sub-mesh the
        // interface displacement as a boundary value
        const bool ifs = m.sliceIf && m.sliceIf[s];
        if (ifs) {
            hasB = true;
        } else if (width == 6) {
            if (xg) gradInternalW<T, 6, true>(m, P, e0, lane, C, Dg, xg, ra, DnP, CP, A, hasB);
            else gradInternalW<T, 6, false>(m, P, e0, lane, C, Dg, xg, ra, DnP, CP, A, hasB);
            kInt = 0;
        } else if (width == 4) {
            if (xg) gradInternalW<T, 4, true>(m, P, e0, lane, C, Dg, xg, ra, DnP, CP, A, hasB);
            else gradInternalW<T, 4, false>(m, P, e0, lane, C, Dg, xg, ra, DnP, CP, A, hasB);
            kInt = 0;
        } else {
            hasB = true;
        }
        for (int k = 0; hasB && k < width; ++k) {
            const int e = e0 + k * kSlice + lane;
            const int n = m.col[e];
            if (n >= 0) {
                if (k >= kInt) continue;  // already summed by the fixed-width pass
                if (ifs && m.eIf[e] >= 0) {
                    const int si = m.eIf[e], nI2 = 2 * m.nIf;
                    const double lv[3] = {m.ifLs[si], m.ifLs[(size_t)nI2 + si], m.ifLs[2 * (size_t)nI2 + si]};
                    double dv[3];
#pragma unroll
                    for (int q = 0; q < 3; ++q) dv[q] = DI[(size_t)q * m.nIf + (si >> 1)] - DnP[q];
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) A[3 * i + j] += lv[i] * dv[j];
                    continue;
                }
                double d[3], dv[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    d[q] = m.cC[(size_t)q * C + n] - CP[q];
                    const double DN =
                        xg ? Dg[(size_t)q * C + n] + ra * (double)xg[(size_t)q * C + n] : Dg[(size_t)q * C + n];
                    dv[q] = DN - DnP[q];
                }
                const double rr = 1.0 / (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const double di = d[i] * rr;
#pragma unroll
                    for (int j = 0; j < 3; ++j) A[3 * i + j] += di * dv[j];
                }
            } else if (n != kPad) {
                const int b = -1 - n;
                const double lv[3] = {m.bLs[b], m.bLs[(size_t)B + b], m.bLs[2 * (size_t)B + b]};
                double Dbv[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) Dbv[q] = Db[(size_t)q * B + b];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) A[3 * i + j] += lv[i] * (Dbv[j] - DnP[j]);
            }
        }
        const double M0 = m.minv[P], M1 = m.minv[(size_t)C + P], M2 = m.minv[2 * (size_t)C + P];
        const double M3 = m.minv[3 * (size_t)C + P], M4 = m.minv[4 * (size_t)C + P], M5 = m.minv[5 * (size_t)C + P];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            stS(G + (size_t)j * C + P, M0 * A[j] + M1 * A[3 + j] + M2 * A[6 + j]);
            stS(G + (size_t)(3 + j) * C + P, M1 * A[j] + M3 * A[3 + j] + M4 * A[6 + j]);
            stS(G + (size_t)(6 + j) * C + P, M2 * A[j] + M4 * A[3 + j] + M5 * A[6 + j]);
        }
        if (Dn && !aitken) {
#pragma unroll
            for (int q = 0; q < 3; ++q) stS(Dn + (size_t)q * C + P, DnP[q]);
        }
        if (mode != 2) {
            double dx2 = 0.0, dold2 = 0.0, dn2 = 0.0;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                const double a = DnP[q] - DP[q];
                const double o = Dold ? DnP[q] - Dold[(size_t)q * C + P] : DnP[q];
                dx2 += a * a;
                dold2 += o * o;
                dn2 += DnP[q] * DnP[q];
                acc[3 + q] += DnP[q];
            }
            acc[0] = nanMax(acc[0], sqrt(dx2));
            acc[1] = nanMax(acc[1], sqrt(dold2));
            acc[2] = nanMax(acc[2], sqrt(dn2));
        }
    }
    if (mode == 2) return;
    storeBlockPartials<6>(acc, 7u, part, lb, bm.nBlocks);
}

// ---------------------------------------------------------------------------
// Sub-meshes of a multi-law case (one law per cellZone): the internal faces
// between zones are boundary faces of the sub-mesh on either side, holding
// an interface displacement D_I that is iterated to traction balance.
// k_iface_pre  solidSubMeshes::interpolateDtoSubMeshD (solidSubMeshes.C:
//              1550-1950), side a the owner, b the neighbour, n from a to b:
//                D_I += wab n.(sigma_b - sigma_a) + wa (D_a - Dprev_a) + wb (D_b - Dprev_b)
//              da = 1/deltaCoeffs_a, db = 1/deltaCoeffs_b, K = impK,
//              wab = da db/(db Ka + da Kb), wa = db Ka/(db Ka + da Kb), wb = 1 - wa;
//              sigma_a, sigma_b the sub-meshes' boundary stresses from the
//              previous correct().  Seen from sub-mesh b the normal, weights
//              and sides swap and the update is the same, so one D_I serves
//              both.  Modes 0/1 as k_grad (D = the current field, the new one
//              D + a x or the Aitken-relaxed one); mode 3: after the loop
//              (mechanicalModel::interpolate, mechanicalModel.C:851-852),
//              D the final field and Dprev the one its last iteration began
//              from.
// k_iface_post the sub-meshes' boundary gradient at the face,
//              G_P + n_s ((D_I - D_P) deltaCoeffs - n_s.G_P) (fvc::grad's
//              boundary value corrected by solidSubMeshes::
//              correctBoundarySnGrad, solidSubMeshes.C:1952-2031, whose k
//              vanishes for the normal patch delta), n_s = +n on side a, -n
//              on side b, and the side's Hooke's law (mechanicalModel.C:
//              477-498).  A cold start's first stress is zero, like the
//              cell stress.
// One instance (sub-meshes are solved alone).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(kBlock) void k_iface_pre(const DevMesh m, const double* __restrict__ D0,
                                                      const T* __restrict__ x, const double* __restrict__ D1,
                                                      double* __restrict__ DI, const double* __restrict__ sigI,
                                                      const Ctl* __restrict__ ctl, int mode) {
    if (mode == 0 && ctl->outerDone) return;
    if (mode != 0 || ctl->firstInner) x = nullptr;
    const bool post = mode == 3;
    const double* __restrict__ Dcur = ctl->dpar ? D1 : D0;
    const double* __restrict__ Doth = ctl->dpar ? D0 : D1;
    // the old field and the new one: (D, D + a x or the Aitken-relaxed one),
    // after the loop (Dprev, D)
    const bool aitken = mode == 0 && ctl->relaxAitken;
    const double ra = (mode == 0 && !aitken && ctl->relaxD != 1.0) ? ctl->relaxD : 1.0;
    const double* __restrict__ Dp = post ? Doth : Dcur;
    const int C = m.nCells, nI = m.nIf, nI2 = 2 * m.nIf;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < nI; i += gridDim.x * kBlock) {
        const int a = m.ifCell[2 * i], b = m.ifCell[2 * i + 1];
        const double n[3] = {m.ifN[i], m.ifN[(size_t)nI + i], m.ifN[2 * (size_t)nI + i]};
        const double da = 1.0 / m.ifDc[2 * i], db = 1.0 / m.ifDc[2 * i + 1];
        const double Ka = cellImpK(m, a), Kb = cellImpK(m, b);
        const double wab = da * db / (db * Ka + da * Kb);
        const double wa = db * Ka / (db * Ka + da * Kb), wb = 1.0 - wa;
        double sa[6], sb[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            sa[q] = sigI[(size_t)q * nI2 + 2 * i];
            sb[q] = sigI[(size_t)q * nI2 + 2 * i + 1];
        }
        const double ta[3] = {sa[0] * n[0] + sa[1] * n[1] + sa[2] * n[2], sa[1] * n[0] + sa[3] * n[1] + sa[4] * n[2],
                              sa[2] * n[0] + sa[4] * n[1] + sa[5] * n[2]};
        const double tb[3] = {sb[0] * n[0] + sb[1] * n[1] + sb[2] * n[2], sb[1] * n[0] + sb[3] * n[1] + sb[4] * n[2],
                              sb[2] * n[0] + sb[4] * n[1] + sb[5] * n[2]};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double pa = Dp[(size_t)q * C + a], pb = Dp[(size_t)q * C + b];
            double na, nb;
            if (post) {
                na = Dcur[(size_t)q * C + a];
                nb = Dcur[(size_t)q * C + b];
            } else if (aitken) {
                na = Doth[(size_t)q * C + a];
                nb = Doth[(size_t)q * C + b];
            } else {
                na = pa + (x ? ra * (double)x[(size_t)q * C + a] : 0.0);
                nb = pb + (x ? ra * (double)x[(size_t)q * C + b] : 0.0);
            }
            DI[(size_t)q * nI + i] += wab * (tb[q] - ta[q]) + wa * (na - pa) + wb * (nb - pb);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_iface_post(const DevMesh m, const double* __restrict__ D0,
                                                       const double* __restrict__ D1, const double* __restrict__ G,
                                                       const double* __restrict__ DI, double* __restrict__ sigI,
                                                       const Ctl* __restrict__ ctl, int mode) {
    if (mode == 0 && ctl->outerDone) return;
    // the gradient's cell values: k_grad left D' in the buffer the flip at
    // the end of the step makes current (mode 0), else the current field
    const double* __restrict__ D = (mode == 0) == (ctl->dpar != 0) ? D0 : D1;
    const bool zero = mode == 1 && ctl->sigma0;
    const int C = m.nCells, nI = m.nIf, nI2 = 2 * m.nIf;
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < nI2; k += gridDim.x * kBlock) {
        const int i = k >> 1, side = k & 1;
        const int P = m.ifCell[k];
        const double sg = side ? -1.0 : 1.0;
        const double n[3] = {sg * m.ifN[i], sg * m.ifN[(size_t)nI + i], sg * m.ifN[2 * (size_t)nI + i]};
        double GI[9], sI[6];
        if (zero) {
#pragma unroll
            for (int q = 0; q < 6; ++q) sI[q] = 0.0;
        } else {
            double GP[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) GP[q] = G[(size_t)q * C + P];
            const double dc = m.ifDc[k];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double nG = n[0] * GP[j] + n[1] * GP[3 + j] + n[2] * GP[6 + j];
                const double sn = (DI[(size_t)j * nI + i] - D[(size_t)j * C + P]) * dc;
#pragma unroll
                for (int r = 0; r < 3; ++r) GI[3 * r + j] = GP[3 * r + j] + n[r] * (sn - nG);
            }
            hooke(GI, cellMu(m, P), cellK(m, P), sI);
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) sigI[(size_t)q * nI2 + k] = sI[q];
    }
}

__global__ __launch_bounds__(kBlock) void k_grad_bnd_post(const DevMesh m, const BlockMap bm,
                                                          const double* __restrict__ G,
                                                          const double* __restrict__ snb, double* __restrict__ Gb,
                                                          double* __restrict__ sigb, double* __restrict__ g,
                                                          const Ctl* __restrict__ ctlAll, int mode) {
    const int inst = faceInst(bm);
    const Ctl* ctl = ctlAll + inst;
    if (mode == 0 && ctl->outerDone) return;
    const int C = m.nCells, B = m.nBoundary;
    const int b0 = bm.iB0[inst], b1 = b0 + bm.iNB[inst];
    for (int b = b0 + blockIdx.x * kBlock + threadIdx.x; b < b1; b += gridDim.x * kBlock) {
        const int type = m.bType[b];
        if (type == 2) continue;
        const int P = m.bCell[b];
        double Gn[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) Gn[q] = G[(size_t)q * C + P];
        const double nv[3] = {m.bN[b], m.bN[(size_t)B + b], m.bN[2 * (size_t)B + b]};
        double Gbv[9], sb[6];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double nG = nv[0] * Gn[j] + nv[1] * Gn[3 + j] + nv[2] * Gn[6 + j];
            const double sn = snb[(size_t)j * B + b];
#pragma unroll
            for (int i = 0; i < 3; ++i) Gbv[3 * i + j] = Gn[3 * i + j] + nv[i] * (sn - nG);
        }
        hooke(Gbv, cellMu(m, P), cellK(m, P), sb);
        if (mode == 1 && ctl->sigma0) {  // cold start: the stress field is still zero
#pragma unroll
            for (int q = 0; q < 6; ++q) sb[q] = 0.0;
        }
        if (Gb) {
#pragma unroll
            for (int q = 0; q < 9; ++q) Gb[(size_t)q * B + b] = Gbv[q];
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) sigb[(size_t)q * B + b] = sb[q];
        if (mode != 2 && (type == 1 || type == 4)) {
            // fixedDisplacementZeroShear: the traction snGrad for zero
            // traction and pressure (its shear-free refGrad)
            const bool zs = type == 4;
            const double ik = cellImpK(m, P);
            const double Tt[9] = {sb[0] - ik * Gbv[0], sb[1] - ik * Gbv[1], sb[2] - ik * Gbv[2],
                                  sb[1] - ik * Gbv[3], sb[3] - ik * Gbv[4], sb[4] - ik * Gbv[5],
                                  sb[2] - ik * Gbv[6], sb[4] - ik * Gbv[7], sb[5] - ik * Gbv[8]};
            const double rik = 1.0 / ik;
            const double pr = zs ? 0.0 : m.bPres[b];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double nT = nv[0] * Tt[j] + nv[1] * Tt[3 + j] + nv[2] * Tt[6 + j];
                g[(size_t)j * B + b] = (zs ? -nT : (m.bVal[(size_t)j * B + b] - nv[j] * pr) - nT) * rik;
            }
        }
    }
}

// sigma = Hooke(G) for output (xx xy xz yy yz zz, SoA)
__global__ __launch_bounds__(kBlock) void k_sigma(const DevMesh m, const double* __restrict__ G,
                                                  double* __restrict__ sig, int c0, int n) {
    const int C = m.nCells;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int P = c0 + i;
        double GP[9], s[6];
#pragma unroll
        for (int q = 0; q < 9; ++q) GP[q] = G[(size_t)q * C + P];
        hooke(GP, cellMu(m, P), cellK(m, P), s);
#pragma unroll
        for (int q = 0; q < 6; ++q) sig[(size_t)q * C + P] = s[q];
    }
}

// solidModel::writeFields (solidModel.C:1936-1962): epsilon = symm(gradD)
// (xx xy xz yy yz zz, SoA), epsilonEq = sqrt(2/3 |dev epsilon|^2),
// sigmaEq = sqrt(3/2 |dev sigma|^2) with sigma from Hooke's law
__global__ __launch_bounds__(kBlock) void k_strain_fields(const DevMesh m, const double* __restrict__ G,
                                                          double* __restrict__ eps, double* __restrict__ epsEq,
                                                          double* __restrict__ sigEq) {
    const int C = m.nCells;
    for (int P = blockIdx.x * kBlock + threadIdx.x; P < C; P += gridDim.x * kBlock) {
        double GP[9], s[6];
#pragma unroll
        for (int q = 0; q < 9; ++q) GP[q] = G[(size_t)q * C + P];
        const double e[6] = {GP[0], 0.5 * (GP[1] + GP[3]), 0.5 * (GP[2] + GP[6]), GP[4], 0.5 * (GP[5] + GP[7]), GP[8]};
        if (eps) {
#pragma unroll
            for (int q = 0; q < 6; ++q) eps[(size_t)q * C + P] = e[q];
        }
        // |dev t|^2 of a symmetric tensor (xx xy xz yy yz zz)
        auto magDev2 = [](const double* t) {
            const double tr3 = (t[0] + t[3] + t[5]) / 3.0;
            const double a = t[0] - tr3, b = t[3] - tr3, c = t[5] - tr3;
            return a * a + b * b + c * c + 2.0 * (t[1] * t[1] + t[2] * t[2] + t[4] * t[4]);
        };
        if (epsEq) epsEq[P] = sqrt((2.0 / 3.0) * magDev2(e));
        if (sigEq) {
            hooke(GP, cellMu(m, P), cellK(m, P), s);
            sigEq[P] = sqrt(1.5 * magDev2(s));
        }
    }
}

// mechanicalModel::interpolate -> enhancedVolPointInterpolation with the
// gradient (mechanicalModel.C:829-845, enhancedVolPointInterpolate.C:350-420):
// point value = sum_c w (D_c + (x_p - C_c).gradD_c) / sum_c w over the
// point's cells (ascending cell order), w = 1/|x_p - C_c|
__global__ __launch_bounds__(kBlock) void k_point_interpolate(const DevMesh m, const double* __restrict__ D,
                                                              const double* __restrict__ G, int nPoints,
                                                              const double* __restrict__ pts,
                                                              const int* __restrict__ pcOff,
                                                              const int* __restrict__ pcCell,
                                                              double* __restrict__ pD) {
    const int C = m.nCells;
    for (int p = blockIdx.x * kBlock + threadIdx.x; p < nPoints; p += gridDim.x * kBlock) {
        const double x[3] = {pts[3 * (size_t)p], pts[3 * (size_t)p + 1], pts[3 * (size_t)p + 2]};
        const int k0 = pcOff[p], k1 = pcOff[p + 1];
        // material zones: each sub-mesh interpolates from its own cells and
        // the point takes the mean over the materials around it
        // (mapSubMeshPointFields, solidSubMeshesTemplates.C:271-336), the
        // materials in the order of their first cell in the list
        int nz = 1;
        if (m.zone) {
            nz = 0;
            for (int k = k0; k < k1; ++k) {
                bool first = true;
                for (int k2 = k0; k2 < k; ++k2) first &= m.zone[pcCell[k2]] != m.zone[pcCell[k]];
                nz += first ? 1 : 0;
            }
        }
        double res[3] = {0, 0, 0};
        for (int kz = k0; kz < (m.zone ? k1 : k0 + 1); ++kz) {
            if (m.zone) {
                bool first = true;
                for (int k2 = k0; k2 < kz; ++k2) first &= m.zone[pcCell[k2]] != m.zone[pcCell[kz]];
                if (!first) continue;
            }
            const int z = (m.zone && kz < k1) ? m.zone[pcCell[kz]] : 0;
            double acc[3] = {0, 0, 0}, sw = 0.0;
            for (int k = k0; k < k1; ++k) {
                const int c = pcCell[k];
                if (m.zone && m.zone[c] != z) continue;
                double d[3];
#pragma unroll
                for (int q = 0; q < 3; ++q) d[q] = x[q] - m.cC[(size_t)q * C + c];
                const double w = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const double dG = d[0] * G[(size_t)j * C + c] + d[1] * G[(size_t)(3 + j) * C + c] +
                                      d[2] * G[(size_t)(6 + j) * C + c];
                    acc[j] += w * (D[(size_t)j * C + c] + dG);
                }
                sw += w;
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const double v = sw > 1e-300 ? acc[j] / sw : acc[j];
                res[j] = nz == 1 ? v : res[j] + v / nz;
            }
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) pD[3 * (size_t)p + j] = res[j];
    }
}

// von Mises equivalent stress sqrt(3/2 |dev sigma|^2) (solidModel.C:1957-1961)
__global__ __launch_bounds__(kBlock) void k_sigma_eq(const double* __restrict__ sig, double* __restrict__ eq, int n) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double xx = sig[i], xy = sig[(size_t)n + i], xz = sig[2 * (size_t)n + i];
        const double yy = sig[3 * (size_t)n + i], yz = sig[4 * (size_t)n + i], zz = sig[5 * (size_t)n + i];
        const double t3 = (xx + yy + zz) / 3.0;
        const double a = xx - t3, b = yy - t3, c = zz - t3;
        eq[i] = sqrt(1.5 * (a * a + b * b + c * c + 2.0 * (xy * xy + xz * xz + yz * yz)));
    }
}

// serving: (re)initialise one instance's state ranges; cold start zeroes the
// displacement, warm start keeps the resident field of buffer `dpar` and makes
// it the old-time field
__global__ __launch_bounds__(kBlock) void k_slot_init(double* __restrict__ D0, double* __restrict__ D1,
                                                      double* __restrict__ Dold, double* __restrict__ G,
                                                      double* __restrict__ g, int C, int B, int c0, int n, int b0,
                                                      int nb, int warm, int dpar) {
    double* D = dpar ? D1 : D0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const int P = c0 + i;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const size_t k = (size_t)q * C + P;
            if (!warm) D[k] = 0.0;
            Dold[k] = D[k];
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) G[(size_t)q * C + P] = 0.0;
    }
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < nb; i += gridDim.x * kBlock)
#pragma unroll
        for (int q = 0; q < 3; ++q) g[(size_t)q * B + b0 + i] = 0.0;
}

// PCG start from a given right-hand side b (for the standalone PCG entry):
// r = b, p = z = b/diag; normalised initial residual with the supplied normFactor
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_start(const DevMesh m, const BlockMap bm, const double* __restrict__ b,
                                                      T* __restrict__ r, T* __restrict__ p, Ctl* __restrict__ ctlAll,
                                                      double* __restrict__ part) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    const int C = m.nCells;
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        if (P >= cEnd) continue;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const size_t i = (size_t)j * C + P;
            const T rT = (T)b[i];
            r[i] = rT;
            const T z = invD[P] * rT;
            p[i] = z;
            acc[j] += fabs(b[i]);
            acc[3 + j] += (double)z * (double)rT;
        }
    }
    storeBlockPartials<6>(acc, 0u, part, lb, bm.nBlocks);
}

// lduMatrix::solver start for one component (component 0) of a caller's LDU
// matrix: r = source - A psi, and the partials of sum|r| and of OpenFOAM's
// normFactor, sum |A psi - rowSum avg(psi)| + |source - rowSum avg(psi)|,
// laid out as k_fin_residual reads them (components 1 and 2 zero)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ldu_start(const DevMesh m, const BlockMap bm, const double* __restrict__ psi,
                                                      const double* __restrict__ src, double avg, T* __restrict__ r,
                                                      double* __restrict__ part) {
    const int lb = logicalBlock(bm.nBlocks);
    const int inst = bm.inst[lb];
    const int C = m.nCells;
    const T* __restrict__ invD = MatSel<T>::invDiag(m);
    double acc[12];
#pragma unroll
    for (int i = 0; i < 12; ++i) acc[i] = 0.0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int s1 = bm.s1[lb];
    const int cEnd = bm.iCell0[inst] + bm.iNCell[inst];
    for (int s = bm.s0[lb] + wave; s < s1; s += kBlock / 64) {
        const int P = s * kSlice + lane;
        const int e0 = m.sliceStart[s];
        const int width = m.sliceWidth[s];
        if (P >= cEnd) continue;
        const double dg = m.diag[P];
        double Ap = dg * psi[P], asum = 0.0;
        for (int k = 0; k < width; ++k) {
            const int e = e0 + k * kSlice + lane;
            const int n = m.col[e];
            if (n >= 0) {
                const double a = m.eA[e];
                Ap -= a * psi[n];
                asum += a;
            }
        }
        const double t = (dg - asum) * avg;
        const double rv = src[P] - Ap;
        const T rT = (T)rv;
        r[P] = rT;
        r[(size_t)C + P] = (T)0;
        r[2 * (size_t)C + P] = (T)0;
        acc[0] += fabs(rv);
        acc[3] += fabs(Ap - t);
        acc[6] += fabs(src[P] - t);
        acc[9] += (double)(invD[P] * rT) * (double)rT;
    }
    storeBlockPartials<12>(acc, 0u, part, lb, bm.nBlocks);
}

// ===========================================================================
// finalize kernels: one workgroup per instance (grid = number of instances)
// ===========================================================================

// PCG::solve start: normFactor (lduMatrix::solver::normFactor), normalised
// initial residual, convergence check before the first iteration, z.r
__global__ __launch_bounds__(kFinBlock) void k_fin_residual(const BlockMap bm, Ctl* __restrict__ ctlAll,
                                                            const double* __restrict__ part) {
    const int inst = bm.finInst ? bm.finInst[blockIdx.x] : (int)blockIdx.x;
    Ctl* ctl = ctlAll + inst;
    if (ctl->outerDone) return;
    __shared__ double tot[12];
    finReduce<12>(part, bm.nBlocks, bm.iBlk0[inst], bm.iNBlk[inst], 0u, tot);
    if (threadIdx.x == 0) {
        double mag2 = 0.0;
        int any = 0;
        for (int j = 0; j < 3; ++j) {
            const double nf = tot[3 + j] + tot[6 + j] + 1e-20;
            const bool solved = (ctl->solveMask >> j) & 1;  // fvMatrix::solveSegregated valid components
            ctl->normFactor[j] = nf;
            ctl->initRes[j] = solved ? tot[j] / nf : 0.0;
            ctl->finalRes[j] = ctl->initRes[j];
            mag2 += ctl->initRes[j] * ctl->initRes[j];
            ctl->rho[j] = tot[9 + j];
            ctl->nIter[j] = 0;
            ctl->alpha[j] = 0.0;
            ctl->beta[j] = 0.0;
            // iterate unless already converged (finalRes = initRes < tolerance)
            const int act = solved && (ctl->pcgMaxIter > 0) && !(ctl->initRes[j] < ctl->pcgTol);
            ctl->active[j] = act;
            any |= act;
        }
        ctl->initResMag = sqrt(mag2);
        ctl->innerDone = !any;
        ctl->firstInner = 1;
    }
}

// PCG start from a caller-supplied right-hand side (normFactor given)
__global__ __launch_bounds__(kFinBlock) void k_fin_pcg_start(const BlockMap bm, Ctl* __restrict__ ctlAll,
                                                             const double* __restrict__ part) {
    const int inst = bm.finInst ? bm.finInst[blockIdx.x] : (int)blockIdx.x;
    Ctl* ctl = ctlAll + inst;
    __shared__ double tot[6];
    finReduce<6>(part, bm.nBlocks, bm.iBlk0[inst], bm.iNBlk[inst], 0u, tot);
    if (threadIdx.x == 0) {
        int any = 0;
        for (int j = 0; j < 3; ++j) {
            const bool solved = (ctl->solveMask >> j) & 1;
            ctl->initRes[j] = solved ? tot[j] / ctl->normFactor[j] : 0.0;
            ctl->finalRes[j] = ctl->initRes[j];
            ctl->rho[j] = tot[3 + j];
            ctl->nIter[j] = 0;
            const int act = solved && (ctl->pcgMaxIter > 0) && !(ctl->initRes[j] < ctl->pcgTol);
            ctl->active[j] = act;
            any |= act;
        }
        ctl->innerDone = !any;
        ctl->firstInner = 1;
    }
}

// alpha = rho / p.q, with PCG's singularity test |p.q|/normFactor < VSMALL
__global__ __launch_bounds__(kFinBlock) void k_fin_spmv(const BlockMap bm, Ctl* __restrict__ ctlAll,
                                                        const double* __restrict__ part) {
    const int inst = bm.finInst ? bm.finInst[blockIdx.x] : (int)blockIdx.x;
    Ctl* ctl = ctlAll + inst;
    if (ctl->outerDone || ctl->innerDone) return;
    __shared__ double tot[3];
    finReduce<3>(part, bm.nBlocks, bm.iBlk0[inst], bm.iNBlk[inst], 0u, tot);
    if (threadIdx.x == 0) {
        int any = 0;
        for (int j = 0; j < 3; ++j) {
            if (!ctl->active[j]) { ctl->alpha[j] = 0.0; continue; }
            if (fabs(tot[j]) / ctl->normFactor[j] < kVSmall) {
                ctl->active[j] = 0;
                ctl->alpha[j] = 0.0;
                continue;
            }
            ctl->alpha[j] = ctl->rho[j] / tot[j];
            any = 1;
        }
        if (!any) ctl->innerDone = 1;
    }
}

// PCG convergence per component (tolerance, relTol, maxIter) and beta
__global__ __launch_bounds__(kFinBlock) void k_fin_update(const BlockMap bm, Ctl* __restrict__ ctlAll,
                                                          const double* __restrict__ part) {
    const int inst = bm.finInst ? bm.finInst[blockIdx.x] : (int)blockIdx.x;
    Ctl* ctl = ctlAll + inst;
    if (ctl->outerDone || ctl->innerDone) return;
    __shared__ double tot[6];
    finReduce<6>(part, bm.nBlocks, bm.iBlk0[inst], bm.iNBlk[inst], 0u, tot);
    if (threadIdx.x == 0) {
        int any = 0;
        for (int j = 0; j < 3; ++j) {
            if (!ctl->active[j]) continue;
            ctl->nIter[j] += 1;
            ctl->innerTotal += 1;
            const double fr = tot[3 + j] / ctl->normFactor[j];
            ctl->finalRes[j] = fr;
            const bool conv = fr < ctl->pcgTol || (ctl->pcgRelTol > 1e-20 && fr < ctl->pcgRelTol * ctl->initRes[j]);
            const int act = !conv && ctl->nIter[j] < ctl->pcgMaxIter;
            if (ctl->pcMode == 0) {  // diagonal preconditioner: z.r came with the update
                ctl->beta[j] = tot[j] / ctl->rho[j];
                ctl->rho[j] = tot[j];
            }
            ctl->active[j] = act;
            any |= act;
        }
        ctl->innerDone = !any;
        ctl->firstInner = 0;
    }
}

// solidModel::converged (total-displacement form) and the mean field for the
// next normFactor; mode 1 (initialisation) only sets the mean
__global__ __launch_bounds__(kFinBlock) void k_fin_grad(const BlockMap bm, Ctl* __restrict__ ctlAll,
                                                        const double* __restrict__ part, double* __restrict__ hist,
                                                        int mode) {
    const int inst = bm.finInst ? bm.finInst[blockIdx.x] : (int)blockIdx.x;
    Ctl* ctl = ctlAll + inst;
    if (mode == 0 && ctl->outerDone) return;
    __shared__ double tot[6];
    finReduce<6>(part, bm.nBlocks, bm.iBlk0[inst], bm.iNBlk[inst], 7u, tot);
    if (threadIdx.x != 0) return;
    const double nc = (double)bm.iNCell[inst];
    for (int j = 0; j < 3; ++j) ctl->xbar[j] = tot[3 + j] / nc;
    if (mode != 0) return;
    ctl->dpar ^= 1;  // k_grad wrote D' into the other buffer
    double denom = tot[1];
    if (denom < kSmall) denom = tot[2] < kSmall ? kSmall : tot[2];
    const double rvf = tot[0] / denom;
    const double ir = ctl->initResMag;
    const int i = ctl->iCorr;
    ctl->rvf = rvf;
    const bool conv = i > 1 && ((ir < ctl->solTol && rvf < ctl->solTol) || rvf < ctl->altTol || ir < ctl->altTol);
    if (hist && i < ctl->histCap) {
        double* h = hist + (size_t)inst * 3 * ctl->histCap;
        h[i] = ir;
        h[ctl->histCap + i] = rvf;
        h[2 * ctl->histCap + i] = (double)(ctl->nIter[0] + ctl->nIter[1] + ctl->nIter[2]);
    }
    if (!isfinite(ir) || !isfinite(rvf)) {
        ctl->diverged = 1;
        ctl->outerDone = 1;
        ctl->nOuter = i + 1;
    } else if (conv) {
        ctl->converged = 1;
        ctl->outerDone = 1;
        ctl->nOuter = i + 1;
    } else if (i + 1 >= ctl->nCorr) {
        ctl->outerDone = 1;
        ctl->nOuter = i + 1;
    } else {
        ctl->iCorr = i + 1;
    }
}

// control block set from a host value passed by argument (no host->device copy)
__global__ void k_ctl_set(Ctl* __restrict__ c, const Ctl v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *c = v;
}

// p[0..n) = v (a kernel rather than a memset node, so callers may capture it)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T* __restrict__ p, size_t n, T v) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) p[i] = v;
}

// b = (double)a, exact
__global__ __launch_bounds__(kBlock) void k_widen(const float* __restrict__ a, double* __restrict__ b, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        b[i] = (double)a[i];
}

// plain streaming copy: the measured streaming ceiling and the known-size
// reference of the traffic counters.  One 32-byte load and store per thread
// over a grid covering the whole array (no grid-stride loop): of the copy
// forms measured on MI355X (grid-stride with one or four loads in flight per
// thread, 4-32 workgroups per CU, plain or non-temporal) this one streams
// fastest.  n is in 32-byte units.
typedef double dbl4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(kBlock) void k_copy(const dbl4* __restrict__ a, dbl4* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = a[i];
}

}  // namespace s4f
