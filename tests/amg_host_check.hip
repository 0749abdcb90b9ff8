// Host-side multigrid build of libs4f (greedy aggregation, Galerkin product,
// q = offdiag(A) P grouped per aggregate), checked on the CPU against a
// map-based restatement on a 3-D 7-point operator with varied coefficients.
#include "../solids4foam_amd/csrc/s4f_api.hip"
#include <map>
#include <chrono>
using namespace s4f;
int main(int argc, char** argv) {
    const int n = argc > 1 ? atoi(argv[1]) : 40;
    const int N = n * n * n;
    AmgHostLevel F;
    F.n = N;
    F.nSlices = (N + kSlice - 1) / kSlice;
    std::vector<std::vector<std::pair<int, double>>> rows(N);
    F.diag.assign(N, 0.0);
    for (int k = 0; k < n; ++k) for (int j = 0; j < n; ++j) for (int i = 0; i < n; ++i) {
        const int c = i + n * (j + n * k);
        const int nb[6][3] = {{-1,0,0},{1,0,0},{0,-1,0},{0,1,0},{0,0,-1},{0,0,1}};
        double d = 0.1;
        for (auto& o : nb) {
            const int ii = i + o[0], jj = j + o[1], kk = k + o[2];
            if (ii < 0 || jj < 0 || kk < 0 || ii >= n || jj >= n || kk >= n) continue;
            const double a = 1.0 + 0.01 * ((c * 7 + ii * 3 + jj * 5 + kk) % 13);
            rows[c].push_back({ii + n * (jj + n * kk), a});
        }
    }
    // symmetric coefficients: use the average of both directions
    for (int c = 0; c < N; ++c) for (auto& p : rows[c]) {
        for (auto& q : rows[p.first]) if (q.first == c) { double m = 0.5 * (p.second + q.second); p.second = m; q.second = m; }
    }
    for (int c = 0; c < N; ++c) { double s = 0.1; for (auto& p : rows[c]) s += p.second; F.diag[c] = s; }
    F.sliceStart.resize(F.nSlices + 1); F.sliceWidth.resize(F.nSlices);
    size_t E = 0;
    for (int s = 0; s < F.nSlices; ++s) {
        F.sliceStart[s] = (int)E; int w = 0;
        for (int l = 0; l < kSlice; ++l) { int c = s * kSlice + l; if (c < N) w = std::max(w, (int)rows[c].size()); }
        F.sliceWidth[s] = w; E += (size_t)w * kSlice;
    }
    F.sliceStart[F.nSlices] = (int)E;
    F.col.assign(E, kPad); F.a.assign(E, 0.0);
    for (int c = 0; c < N; ++c) { int s = c / kSlice, l = c % kSlice;
        for (size_t k = 0; k < rows[c].size(); ++k) { F.col[F.sliceStart[s] + k * kSlice + l] = rows[c][k].first; F.a[F.sliceStart[s] + k * kSlice + l] = rows[c][k].second; } }
    F.inst0 = {0}; F.instN = {N};
    std::vector<int> agg(N, -1);
    auto t0 = std::chrono::steady_clock::now();
    const int nc = aggregate(F, agg, 0, N, 0);
    auto t1 = std::chrono::steady_clock::now();
    AmgHostLevel G;
    galerkin(F, agg, nc, G);
    auto t2 = std::chrono::steady_clock::now();
    // reference: dense map per coarse row
    std::vector<std::map<int, double>> R(nc);
    std::vector<double> dg(nc, 0.0);
    for (int c = 0; c < N; ++c) {
        const int I = agg[c];
        dg[I] += F.diag[c];
        for (auto& p : rows[c]) { const int J = agg[p.first]; if (J == I) dg[I] -= p.second; else R[I][J] += p.second; }
    }
    double err = 0.0; long bad = 0;
    for (int I = 0; I < nc; ++I) {
        err = std::max(err, std::fabs(dg[I] - G.diag[I]) / std::fabs(dg[I]));
        int s = I / kSlice, l = I % kSlice, k = 0;
        for (auto& q : R[I]) {
            const size_t e = G.sliceStart[s] + (size_t)k * kSlice + l;
            if (G.col[e] != q.first) ++bad;
            else err = std::max(err, std::fabs(G.a[e] - q.second) / std::fabs(q.second));
            ++k;
        }
        if (k < G.sliceWidth[s] && G.col[G.sliceStart[s] + (size_t)k * kSlice + l] != kPad) ++bad;
    }
    std::vector<int> qc, qw; std::vector<double> q;
    AmgHostLevel F2 = F; F2.agg = agg;
    auto t3 = std::chrono::steady_clock::now();
    groupByAggregate(F2, qc, q, qw);
    auto t4 = std::chrono::steady_clock::now();
    // q check: per fine row, sum of off-diagonal coefficients per neighbour aggregate
    long qbad = 0; double qerr = 0.0;
    for (int c = 0; c < N; ++c) {
        std::map<int, double> m; for (auto& p : rows[c]) m[agg[p.first]] += p.second;
        int s = c / kSlice, l = c % kSlice, k = 0;
        for (auto& t : m) { size_t e = F.sliceStart[s] + (size_t)k * kSlice + l; if (qc[e] != t.first) ++qbad; else qerr = std::max(qerr, std::fabs(q[e] - t.second) / std::fabs(t.second)); ++k; }
    }
    printf("n %d: %d rows -> %d aggregates; aggregate %.3f s, galerkin %.3f s, group %.3f s; galerkin max rel err %.2e, misplaced %ld; q err %.2e misplaced %ld\n",
           n, N, nc, std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count(),
           std::chrono::duration<double>(t4 - t3).count(), err, bad, qerr, qbad);
    return (bad || qbad || err > 1e-12 || qerr > 1e-12) ? 1 : 0;
}
