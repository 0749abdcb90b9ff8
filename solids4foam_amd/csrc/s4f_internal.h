// s4f_internal.h -- shared host-side types of libs4f (not part of the ABI).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace s4f {

// An OpenFOAM-style polyMesh: points, faces (flat with offsets), owner for
// every face, neighbour for internal faces, contiguous boundary patches.
struct PolyMesh {
    int nPoints = 0, nCells = 0, nFaces = 0, nInternalFaces = 0, nPatches = 0;
    std::vector<double> points;   // nPoints*3
    std::vector<int> fOff;        // nFaces+1
    std::vector<int> fVerts;
    std::vector<int> owner;       // nFaces
    std::vector<int> neighbour;   // nInternalFaces
    std::vector<int> patchStart, patchSize;
    bool hasGeometry = false;
    std::vector<double> Cf, Sf, C, V;
    void computeGeometry();
};

PolyMesh* makeBoxMesh(int nx, int ny, int nz, const double L[3], double jitter, uint64_t seed, bool tets);

void setError(const std::string& msg);

}  // namespace s4f
