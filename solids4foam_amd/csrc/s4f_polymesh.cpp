// s4f_polymesh.cpp -- host-side polyMesh generation and OpenFOAM geometry.
//
// The box generator reproduces blockMesh's single-block numbering (points and
// cells i-fastest, internal faces in upper-triangular owner/neighbour order,
// boundary faces per block side in blockMesh's createBoundary loop order),
// the mesh layout the reference cases are built with (e.g.
// tutorials/solids/linearElasticity/cantilever2d/segregatedCantilever2d/constant/polyMesh/blockMeshDict).
// Geometry follows OpenFOAM's primitiveMesh face-centre/area and
// cell-centre/volume decompositions, which the reference reads through
// mesh().C(), mesh().V(), mesh().Sf() and mesh().Cf().
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>
#include <algorithm>

#include "s4f_internal.h"

namespace s4f {

namespace {

inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// seeded uniform draw in [0,1) per (point, axis): splitmix64 finaliser of a
// Weyl sequence; counter-based so any point's jitter is independent of order
inline double jitterU01(uint64_t seed, uint64_t ctr) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * (ctr + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// OpenFOAM hex cell model: vertex v = (bit0 x, bit1 y, bit2 z) in the order
// 0:(000) 1:(100) 2:(110) 3:(010) 4:(001) 5:(101) 6:(111) 7:(011);
// outward faces x-min, x-max, y-min, y-max, z-min, z-max.
const int kHexFaces[6][4] = {{0, 4, 7, 3}, {1, 2, 6, 5}, {0, 1, 5, 4}, {3, 7, 6, 2}, {0, 3, 2, 1}, {4, 5, 6, 7}};
// OpenFOAM tet model faces, outward for a positively oriented tet.
const int kTetFaces[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
// Kuhn split: tet t follows the axis path kKuhnPath[t] from local vertex 0 to 6.
const int kKuhnPath[6][3] = {{0, 1, 2}, {1, 0, 2}, {1, 2, 0}, {2, 1, 0}, {2, 0, 1}, {0, 2, 1}};

inline int axisBit(int a) { return a == 0 ? 1 : (a == 1 ? 2 : 4); }
// hex-model vertex id of the corner with bits (x,y,z)
inline int hexVertexOfBits(int bits) {
    static const int m[8] = {0, 1, 3, 2, 4, 5, 7, 6};  // bits xyz (x=1,y=2,z=4) -> model id
    return m[bits];
}
int kuhnIndex(int a, int b, int c) {
    for (int t = 0; t < 6; ++t)
        if (kKuhnPath[t][0] == a && kKuhnPath[t][1] == b && kKuhnPath[t][2] == c) return t;
    return -1;
}
inline int permParity(int a, int b, int c) {
    // +1 even, -1 odd permutation of (0,1,2)
    int inv = (a > b) + (a > c) + (b > c);
    return (inv % 2) ? -1 : 1;
}

}  // namespace

void PolyMesh::computeGeometry() {
    const int nF = nFaces, nC = nCells;
    Cf.assign((size_t)nF * 3, 0.0);
    Sf.assign((size_t)nF * 3, 0.0);
    C.assign((size_t)nC * 3, 0.0);
    V.assign((size_t)nC, 0.0);
    const double* p = points.data();
    for (int f = 0; f < nF; ++f) {
        const int* fv = fVerts.data() + fOff[f];
        const int np = fOff[f + 1] - fOff[f];
        double* cf = &Cf[3 * (size_t)f];
        double* sf = &Sf[3 * (size_t)f];
        if (np == 3) {
            const double *a = p + 3 * fv[0], *b = p + 3 * fv[1], *c = p + 3 * fv[2];
            double e1[3], e2[3];
            for (int d = 0; d < 3; ++d) {
                cf[d] = (1.0 / 3.0) * (a[d] + b[d] + c[d]);
                e1[d] = b[d] - a[d];
                e2[d] = c[d] - a[d];
            }
            sf[0] = 0.5 * (e1[1] * e2[2] - e1[2] * e2[1]);
            sf[1] = 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]);
            sf[2] = 0.5 * (e1[0] * e2[1] - e1[1] * e2[0]);
        } else {
            double fc[3] = {p[3 * fv[0]], p[3 * fv[0] + 1], p[3 * fv[0] + 2]};
            for (int i = 1; i < np; ++i)
                for (int d = 0; d < 3; ++d) fc[d] += p[3 * fv[i] + d];
            for (int d = 0; d < 3; ++d) fc[d] /= np;
            double sumN[3] = {0, 0, 0}, sumA = 0.0, sumAc[3] = {0, 0, 0};
            for (int i = 0; i < np; ++i) {
                const double* pi = p + 3 * fv[i];
                const double* pn = p + 3 * fv[(i + 1) % np];
                double c[3], e1[3], e2[3], n[3];
                for (int d = 0; d < 3; ++d) {
                    c[d] = pi[d] + pn[d] + fc[d];
                    e1[d] = pn[d] - pi[d];
                    e2[d] = fc[d] - pi[d];
                }
                n[0] = e1[1] * e2[2] - e1[2] * e2[1];
                n[1] = e1[2] * e2[0] - e1[0] * e2[2];
                n[2] = e1[0] * e2[1] - e1[1] * e2[0];
                const double a = std::sqrt(dot3(n, n));
                for (int d = 0; d < 3; ++d) {
                    sumN[d] += n[d];
                    sumAc[d] += a * c[d];
                }
                sumA += a;
            }
            if (sumA < 1e-150) {
                for (int d = 0; d < 3; ++d) { cf[d] = fc[d]; sf[d] = 0.0; }
            } else {
                for (int d = 0; d < 3; ++d) { cf[d] = (1.0 / 3.0) * sumAc[d] / sumA; sf[d] = 0.5 * sumN[d]; }
            }
        }
    }
    // cell-centre estimate = mean face centre, accumulated over the owner
    // loop then the neighbour loop; then the face-pyramid decomposition
    std::vector<double> cEst((size_t)nC * 3, 0.0);
    std::vector<int> nCF(nC, 0);
    for (int f = 0; f < nF; ++f) {
        const int o = owner[f];
        for (int d = 0; d < 3; ++d) cEst[3 * (size_t)o + d] += Cf[3 * (size_t)f + d];
        nCF[o]++;
    }
    for (int f = 0; f < nInternalFaces; ++f) {
        const int n = neighbour[f];
        for (int d = 0; d < 3; ++d) cEst[3 * (size_t)n + d] += Cf[3 * (size_t)f + d];
        nCF[n]++;
    }
    for (int c = 0; c < nC; ++c)
        for (int d = 0; d < 3; ++d) cEst[3 * (size_t)c + d] /= nCF[c];
    for (int f = 0; f < nF; ++f) {
        const int o = owner[f];
        double dd[3];
        for (int d = 0; d < 3; ++d) dd[d] = Cf[3 * (size_t)f + d] - cEst[3 * (size_t)o + d];
        const double pyr3Vol = dot3(&Sf[3 * (size_t)f], dd);
        for (int d = 0; d < 3; ++d) {
            const double pc = (3.0 / 4.0) * Cf[3 * (size_t)f + d] + (1.0 / 4.0) * cEst[3 * (size_t)o + d];
            C[3 * (size_t)o + d] += pyr3Vol * pc;
        }
        V[o] += pyr3Vol;
    }
    for (int f = 0; f < nInternalFaces; ++f) {
        const int n = neighbour[f];
        double dd[3];
        for (int d = 0; d < 3; ++d) dd[d] = cEst[3 * (size_t)n + d] - Cf[3 * (size_t)f + d];
        const double pyr3Vol = dot3(&Sf[3 * (size_t)f], dd);
        for (int d = 0; d < 3; ++d) {
            const double pc = (3.0 / 4.0) * Cf[3 * (size_t)f + d] + (1.0 / 4.0) * cEst[3 * (size_t)n + d];
            C[3 * (size_t)n + d] += pyr3Vol * pc;
        }
        V[n] += pyr3Vol;
    }
    for (int c = 0; c < nC; ++c) {
        if (std::fabs(V[c]) > 1e-300) {
            for (int d = 0; d < 3; ++d) C[3 * (size_t)c + d] /= V[c];
        } else {
            for (int d = 0; d < 3; ++d) C[3 * (size_t)c + d] = cEst[3 * (size_t)c + d];
        }
        V[c] *= (1.0 / 3.0);
    }
    hasGeometry = true;
}

PolyMesh* makeBoxMesh(int nx, int ny, int nz, const double L[3], double jitter, uint64_t seed, bool tets) {
    if (nx < 1 || ny < 1 || nz < 1) throw std::invalid_argument("box mesh needs nx, ny, nz >= 1");
    const int64_t nHex64 = (int64_t)nx * ny * nz;
    if (nHex64 * (tets ? 6 : 1) > (int64_t)INT32_MAX / 8) throw std::invalid_argument("box mesh too large for int32 labels");
    auto pm = new PolyMesh();
    const int px = nx + 1, py = ny + 1;
    auto pid = [=](int i, int j, int k) { return i + px * (j + py * k); };
    pm->nPoints = px * py * (nz + 1);
    pm->points.resize((size_t)pm->nPoints * 3);
    const double hx = L[0] / nx, hy = L[1] / ny, hz = L[2] / nz;
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const int p = pid(i, j, k);
                double* x = &pm->points[3 * (size_t)p];
                x[0] = L[0] * i / nx;
                x[1] = L[1] * j / ny;
                x[2] = L[2] * k / nz;
                if (jitter != 0.0 && i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz) {
                    x[0] += jitter * hx * (2.0 * jitterU01(seed, 3ull * p + 0) - 1.0);
                    x[1] += jitter * hy * (2.0 * jitterU01(seed, 3ull * p + 1) - 1.0);
                    x[2] += jitter * hz * (2.0 * jitterU01(seed, 3ull * p + 2) - 1.0);
                }
            }
    const int nHex = (int)nHex64;
    pm->nPatches = 6;
    auto hexVerts = [&](int i, int j, int k, int* hv) {
        hv[0] = pid(i, j, k); hv[1] = pid(i + 1, j, k); hv[2] = pid(i + 1, j + 1, k); hv[3] = pid(i, j + 1, k);
        hv[4] = pid(i, j, k + 1); hv[5] = pid(i + 1, j, k + 1); hv[6] = pid(i + 1, j + 1, k + 1); hv[7] = pid(i, j + 1, k + 1);
    };

    if (!tets) {
        pm->nCells = nHex;
        const int nIF = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1);
        const int nBF = 2 * (ny * nz + nx * nz + nx * ny);
        pm->nInternalFaces = nIF;
        pm->nFaces = nIF + nBF;
        pm->fOff.resize((size_t)pm->nFaces + 1);
        pm->fVerts.resize((size_t)pm->nFaces * 4);
        pm->owner.resize(pm->nFaces);
        pm->neighbour.resize(nIF);
        int f = 0;
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int c = i + nx * (j + ny * k);
                    int hv[8];
                    hexVerts(i, j, k, hv);
                    const int nbr[3] = {i < nx - 1 ? c + 1 : -1, j < ny - 1 ? c + nx : -1, k < nz - 1 ? c + nx * ny : -1};
                    const int mf[3] = {1, 3, 5};
                    for (int q = 0; q < 3; ++q) {
                        if (nbr[q] < 0) continue;
                        pm->owner[f] = c;
                        pm->neighbour[f] = nbr[q];
                        pm->fOff[f] = 4 * f;
                        for (int v = 0; v < 4; ++v) pm->fVerts[4 * (size_t)f + v] = hv[kHexFaces[mf[q]][v]];
                        ++f;
                    }
                }
        auto bface = [&](int cell, int a, int b, int c2, int d) {
            pm->owner[f] = cell;
            pm->fOff[f] = 4 * f;
            int* v = &pm->fVerts[4 * (size_t)f];
            v[0] = a; v[1] = b; v[2] = c2; v[3] = d;
            ++f;
        };
        // blockMesh createBoundary: x-min, x-max (k outer, j inner), y-min,
        // y-max (i outer, k inner), z-min, z-max (i outer, j inner)
        int patch = 0;
        pm->patchStart.assign(6, 0);
        pm->patchSize.assign(6, 0);
        auto open = [&]() { pm->patchStart[patch] = f; };
        auto close = [&]() { pm->patchSize[patch] = f - pm->patchStart[patch]; ++patch; };
        open();
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j) bface(nx * (j + ny * k), pid(0, j, k), pid(0, j, k + 1), pid(0, j + 1, k + 1), pid(0, j + 1, k));
        close(); open();
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                bface(nx - 1 + nx * (j + ny * k), pid(nx, j, k), pid(nx, j + 1, k), pid(nx, j + 1, k + 1), pid(nx, j, k + 1));
        close(); open();
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k) bface(i + nx * (ny * k), pid(i, 0, k), pid(i + 1, 0, k), pid(i + 1, 0, k + 1), pid(i, 0, k + 1));
        close(); open();
        for (int i = 0; i < nx; ++i)
            for (int k = 0; k < nz; ++k)
                bface(i + nx * (ny - 1 + ny * k), pid(i, ny, k), pid(i, ny, k + 1), pid(i + 1, ny, k + 1), pid(i + 1, ny, k));
        close(); open();
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) bface(i + nx * j, pid(i, j, 0), pid(i, j + 1, 0), pid(i + 1, j + 1, 0), pid(i + 1, j, 0));
        close(); open();
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                bface(i + nx * (j + ny * (nz - 1)), pid(i, j, nz), pid(i + 1, j, nz), pid(i + 1, j + 1, nz), pid(i, j + 1, nz));
        close();
        pm->fOff[pm->nFaces] = 4 * pm->nFaces;
    } else {
        // Kuhn split.  The neighbour of every tet face is known analytically:
        // tet (a,b,c) = path v0 -> +e_a -> +e_b -> +e_c; its face opposite
        //   local vertex 3 (v111) lies on side x_c = 0 -> hex -c, tet (c,a,b)
        //   local vertex 0 (v000) lies on side x_a = 1 -> hex +a, tet (b,c,a)
        //   local vertex 1 -> same hex, tet (b,a,c);  local vertex 2 -> same hex, tet (a,c,b)
        const int nC = 6 * nHex;
        pm->nCells = nC;
        const int dims[3] = {nx, ny, nz};
        std::vector<int> cellPts((size_t)nC * 4);
        std::vector<int> nbrOf((size_t)nC * 4, -1);
        std::vector<signed char> sideOf((size_t)nC * 4, -1);
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int h = i + nx * (j + ny * k);
                    const int ijk[3] = {i, j, k};
                    int hv[8];
                    hexVerts(i, j, k, hv);
                    for (int t = 0; t < 6; ++t) {
                        const int a = kKuhnPath[t][0], b = kKuhnPath[t][1], c3 = kKuhnPath[t][2];
                        const int cell = 6 * h + t;
                        // local vertices in path order, then the positive orientation
                        int lv[4] = {0, axisBit(a), axisBit(a) | axisBit(b), 7};
                        int v[4];
                        for (int q = 0; q < 4; ++q) v[q] = hv[hexVertexOfBits(lv[q])];
                        // the path's local vertex slot for each OpenFOAM tet vertex
                        int slot[4] = {0, 1, 2, 3};
                        if (permParity(a, b, c3) < 0) { std::swap(v[1], v[2]); std::swap(slot[1], slot[2]); }
                        for (int q = 0; q < 4; ++q) cellPts[4 * (size_t)cell + q] = v[q];
                        // the OpenFOAM face lf is opposite tet vertex (tet model):
                        // face 0 (1 2 3) opp 0, face 1 (0 3 2) opp 1, face 2 (0 1 3) opp 2, face 3 (0 2 1) opp 3
                        for (int lf = 0; lf < 4; ++lf) {
                            const int oppPath = slot[lf];  // path position of the opposite vertex
                            int nb = -1, side = -1;
                            if (oppPath == 3) {          // side x_c = 0
                                if (ijk[c3] > 0) {
                                    const int hh = h - (c3 == 0 ? 1 : (c3 == 1 ? nx : nx * ny));
                                    nb = 6 * hh + kuhnIndex(c3, a, b);
                                } else side = 2 * c3;
                            } else if (oppPath == 0) {   // side x_a = 1
                                if (ijk[a] < dims[a] - 1) {
                                    const int hh = h + (a == 0 ? 1 : (a == 1 ? nx : nx * ny));
                                    nb = 6 * hh + kuhnIndex(b, c3, a);
                                } else side = 2 * a + 1;
                            } else if (oppPath == 1) {
                                nb = 6 * h + kuhnIndex(b, a, c3);
                            } else {
                                nb = 6 * h + kuhnIndex(a, c3, b);
                            }
                            nbrOf[4 * (size_t)cell + lf] = nb;
                            sideOf[4 * (size_t)cell + lf] = (signed char)side;
                        }
                    }
                }
        int nIF = 0;
        for (size_t q = 0; q < nbrOf.size(); ++q)
            if (nbrOf[q] >= 0) ++nIF;
        nIF /= 2;
        const int nBF = 4 * nC - 2 * nIF;
        pm->nInternalFaces = nIF;
        pm->nFaces = nIF + nBF;
        pm->fOff.resize((size_t)pm->nFaces + 1);
        pm->fVerts.resize((size_t)pm->nFaces * 3);
        pm->owner.resize(pm->nFaces);
        pm->neighbour.resize(nIF);
        int f = 0;
        for (int c = 0; c < nC; ++c) {
            int nbs[4], lfs[4], ne = 0;
            for (int lf = 0; lf < 4; ++lf) {
                const int nb = nbrOf[4 * (size_t)c + lf];
                if (nb > c) { nbs[ne] = nb; lfs[ne] = lf; ++ne; }
            }
            for (int x = 1; x < ne; ++x)  // insertion sort by neighbour
                for (int y = x; y > 0 && nbs[y - 1] > nbs[y]; --y) { std::swap(nbs[y], nbs[y - 1]); std::swap(lfs[y], lfs[y - 1]); }
            for (int q = 0; q < ne; ++q) {
                pm->owner[f] = c;
                pm->neighbour[f] = nbs[q];
                pm->fOff[f] = 3 * f;
                for (int v = 0; v < 3; ++v) pm->fVerts[3 * (size_t)f + v] = cellPts[4 * (size_t)c + kTetFaces[lfs[q]][v]];
                ++f;
            }
        }
        pm->patchStart.assign(6, 0);
        pm->patchSize.assign(6, 0);
        for (int side = 0; side < 6; ++side) {
            pm->patchStart[side] = f;
            for (int c = 0; c < nC; ++c)
                for (int lf = 0; lf < 4; ++lf) {
                    if (sideOf[4 * (size_t)c + lf] != side) continue;
                    pm->owner[f] = c;
                    pm->fOff[f] = 3 * f;
                    for (int v = 0; v < 3; ++v) pm->fVerts[3 * (size_t)f + v] = cellPts[4 * (size_t)c + kTetFaces[lf][v]];
                    ++f;
                }
            pm->patchSize[side] = f - pm->patchStart[side];
        }
        pm->fOff[pm->nFaces] = 3 * pm->nFaces;
    }
    pm->computeGeometry();
    return pm;
}

}  // namespace s4f
