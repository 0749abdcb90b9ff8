// Host-code sanitizer driver: the product's polyMesh generator and geometry
// (solids4foam_amd/csrc/s4f_polymesh.cpp) and the FP64 CPU oracle
// (oracle/s4f_oracle.c), built with -fsanitize=address,undefined by
// tests/test_sanitize.py and run on small hex, tet and 2-D cases: mesh
// generation, geometry, case set-up, outer iterations, a converged solve,
// point interpolation, transient steps, and freeing everything.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "../../solids4foam_amd/csrc/s4f_internal.h"

extern "C" {
typedef struct or_mesh or_mesh;
typedef struct or_case or_case;
typedef struct {
    int nCorr; double solTol, altTol; int precond; double pcgTol, pcgRelTol; int pcgMaxIter;
    double rcScale, interfaceScale; int historyCap; double *histInitRes, *histRvf; int *histInner;
    double relaxD; int relaxAitken, d2dt2; double deltaT, deltaT0; double g[3];
} or_opts;
typedef struct { int nOuter, converged, innerTotal, maxIterReached; double initRes, rvf; } or_report;
or_mesh* or_box_mesh(int, int, int, double, double, double, double, uint64_t, int);
void or_mesh_free(or_mesh*);
void or_mesh_info(const or_mesh*, int*);
or_case* or_case_create(const or_mesh*, const double*, const double*, const int*, const double*, const double*, int,
                        double, double, int);
void or_case_free(or_case*);
void or_init(or_case*, const double*);
int or_iterate(or_case*, const or_opts*, double*, double*, int*);
int or_solve(or_case*, const or_opts*, or_report*);
void or_point_interpolate(const or_case*, const double*, const double*, double*);
long or_field_size(or_case*, const char*);
int or_get_field(or_case*, const char*, double*);
int or_set_field(or_case*, const char*, const double*);
void or_advance_time(or_case*);
void or_residual(or_case*);
void or_grad(or_case*);
}

namespace s4f {
void setError(const std::string&) {}
}

static or_opts opts() {
    or_opts o{};
    o.nCorr = 400; o.solTol = 1e-7; o.altTol = 1e-7; o.precond = 0; o.pcgTol = 1e-9; o.pcgRelTol = 0.1;
    o.pcgMaxIter = 1000; o.rcScale = 0.1; o.interfaceScale = 0.01;
    return o;
}

static int runCase(int nx, int ny, int nz, int tets, double jitter, int twoD, int transient) {
    or_mesh* m = or_box_mesh(nx, ny, nz, 1.0, 0.5, twoD ? 0.1 : 0.4, jitter, 7, tets);
    if (!m) return 1;
    int info[6];
    or_mesh_info(m, info);
    const int P = info[0], C = info[1], NF = info[2], F = info[3];
    const int B = NF - F;
    std::vector<double> E(C, 200e9), nu(C, 0.3), bv(3 * (size_t)B, 0.0), bp(B, 0.0);
    for (int c = 0; c < C; c += 5) E[c] = 20e9;
    int pt[6] = {0, 1, 1, 1, twoD ? 2 : 1, twoD ? 2 : 1};
    // patch order x-min, x-max, y-min, y-max, z-min, z-max: x-max loaded
    const int nx1 = ny * nz * (tets ? 2 : 1);  // faces of the x-min side
    for (int b = nx1; b < 2 * nx1 && b < B; ++b) bv[3 * (size_t)b + 1] = 1e6;
    or_case* c = or_case_create(m, E.data(), nu.data(), pt, bv.data(), bp.data(), tets, 0.1, 0.01, twoD);
    if (!c) return 2;
    or_init(c, nullptr);
    or_opts o = opts();
    if (transient) {
        o.d2dt2 = 1; o.deltaT = 1e-3; o.g[1] = -9.81;
        std::vector<double> rho(C, 7800.0);
        or_set_field(c, "rho", rho.data());
    }
    double ir, rvf;
    int inner;
    for (int k = 0; k < 3; ++k) or_iterate(c, &o, &ir, &rvf, &inner);
    or_report rep{};
    or_solve(c, &o, &rep);
    if (transient) {
        or_advance_time(c);
        or_solve(c, &o, &rep);
    }
    or_grad(c);
    or_residual(c);
    std::vector<double> D(or_field_size(c, "D")), G(or_field_size(c, "G")), pD(3 * (size_t)P);
    or_get_field(c, "D", D.data());
    or_get_field(c, "G", G.data());
    or_point_interpolate(c, D.data(), G.data(), pD.data());
    for (double v : pD)
        if (!std::isfinite(v)) return 3;
    or_case_free(c);
    or_mesh_free(m);
    // the product's generator and geometry on the same box
    const double L[3] = {1.0, 0.5, 0.4};
    std::unique_ptr<s4f::PolyMesh> pm(s4f::makeBoxMesh(nx, ny, nz, L, jitter, 7, tets != 0));
    pm->computeGeometry();
    if (pm->nCells != C || pm->nFaces != NF) return 4;
    return 0;
}

int main() {
    int rc = 0;
    rc |= runCase(6, 4, 3, 0, 0.15, 0, 0);
    rc |= runCase(3, 3, 2, 1, 0.1, 0, 0) << 4;
    rc |= runCase(8, 3, 1, 0, 0.0, 1, 0) << 8;
    rc |= runCase(4, 3, 3, 0, 0.1, 0, 1) << 12;
    std::printf("sanitize driver rc=%d\n", rc);
    return rc;
}
