// Host-side multigrid build of libs4f (greedy aggregation, Galerkin product,
// q = offdiag(A) P grouped per aggregate, and the smoothed first prolongation
// with its coarse operator, restriction lists and up-sweep operator), checked
// on the CPU against a map-based restatement on a 3-D 7-point operator with
// varied coefficients.
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
    // smoothed first prolongation P = (I - wp D^-1 A) P_agg, wp = 2/3: the
    // coarse operator P^T A P, the restriction lists P^T and the up-sweep
    // operator G = (D / w) P - A P against maps built from the rows directly
    const double wp = 2.0 / 3.0, om = 0.9;
    std::vector<std::map<int, double>> Pm(N), APm(N);
    for (int c = 0; c < N; ++c) {
        Pm[c][agg[c]] += 1.0 - wp;
        for (auto& p : rows[c]) Pm[c][agg[p.first]] += wp * p.second / F.diag[c];
    }
    for (int c = 0; c < N; ++c) {  // (A P)_c = D_c P_c - sum_j a_cj P_j
        for (auto& t : Pm[c]) APm[c][t.first] += F.diag[c] * t.second;
        for (auto& p : rows[c]) for (auto& t : Pm[p.first]) APm[c][t.first] -= p.second * t.second;
    }
    std::vector<std::map<int, double>> Acm(nc);
    for (int c = 0; c < N; ++c)
        for (auto& t : Pm[c]) for (auto& u : APm[c]) Acm[t.first][u.first] += t.second * u.second;
    AmgHostLevel Fs = F, Cs;
    Fs.agg = agg;
    auto t5 = std::chrono::steady_clock::now();
    galerkinSmoothed(Fs, nc, om, Cs, 0.0);
    auto t6 = std::chrono::steady_clock::now();
    double serr = 0.0, asym = 0.0; long sbad = 0;
    for (int I = 0; I < nc; ++I) {
        const double scale = std::fabs(Acm[I][I]);
        serr = std::max(serr, std::fabs(Cs.diag[I] - Acm[I][I]) / scale);
        std::map<int, double> got;
        int s = I / kSlice, l = I % kSlice;
        for (int k = 0; k < Cs.sliceWidth[s]; ++k) {
            const size_t e = Cs.sliceStart[s] + (size_t)k * kSlice + l;
            if (Cs.col[e] == kPad) continue;
            if (got.count(Cs.col[e]) || Cs.col[e] == I) ++sbad;
            got[Cs.col[e]] = -Cs.a[e];
        }
        for (auto& t : Acm[I]) if (t.first != I) serr = std::max(serr, std::fabs(got[t.first] - t.second) / scale);
        for (auto& t : got) if (!Acm[I].count(t.first)) ++sbad;
        for (auto& t : Acm[I]) asym = std::max(asym, std::fabs(t.second - Acm[t.first][I]) / scale);
    }
    // P^T lists: each coarse row's fine rows ascending with P's weights
    double werr = 0.0; long wbad = 0;
    std::vector<std::map<int, double>> PT(nc);
    for (int c = 0; c < N; ++c) for (auto& t : Pm[c]) PT[t.first][c] = t.second;
    for (int I = 0; I < nc; ++I) {
        if (Cs.memOff[I + 1] - Cs.memOff[I] != (int)PT[I].size()) { ++wbad; continue; }
        int k = Cs.memOff[I];
        for (auto& t : PT[I]) {
            if (Cs.mem[k] != t.first) ++wbad;
            else werr = std::max(werr, std::fabs(Cs.memW[k] - t.second));
            ++k;
        }
    }
    // the same lists as sliced ELL (coalesced restriction)
    for (int I = 0; I < nc; ++I) {
        const int sI = I / kSlice, l = I % kSlice, k0 = Cs.memOff[I], w = Cs.memOff[I + 1] - k0;
        if (Cs.rWidth[sI] < w) { ++wbad; continue; }
        for (int k = 0; k < Cs.rWidth[sI]; ++k) {
            const size_t e = Cs.rStart[sI] + (size_t)k * kSlice + l;
            if (k < w && (Cs.rMem[e] != Cs.mem[k0 + k] || Cs.rW[e] != Cs.memW[k0 + k])) ++wbad;
            if (k >= w && (Cs.rMem[e] != Cs.mem[k0] || Cs.rW[e] != 0.0)) ++wbad;
        }
    }
    // G rows
    double gerr = 0.0; long gbad = 0;
    for (int c = 0; c < N; ++c) {
        std::map<int, double> ref;
        for (auto& t : Pm[c]) ref[t.first] += F.diag[c] / om * t.second;
        for (auto& t : APm[c]) ref[t.first] -= t.second;
        std::map<int, double> got;
        const int s = c / kSlice, l = c % kSlice;
        for (int k = 0; k < Fs.gWidth[s]; ++k) {
            const size_t e = Fs.gStart[s] + (size_t)k * kSlice + l;
            got[Fs.gCol[e]] += Fs.g[e];
            if (Fs.gCol[e] < 0 || Fs.gCol[e] >= nc) ++gbad;
        }
        for (auto& t : ref) gerr = std::max(gerr, std::fabs(got[t.first] - t.second) / F.diag[c]);
        for (auto& t : got) if (!ref.count(t.first)) ++gbad;
    }
    // the filtered operator: entries kept are those at or above the threshold,
    // each row's sum (diagonal minus magnitudes) unchanged, still symmetric
    AmgHostLevel Fz = F, Cz;
    Fz.agg = agg;
    const double filt = 0.04;  // the filter's own check, whatever the library's default
    galerkinSmoothed(Fz, nc, om, Cz, filt);
    double ferr = 0.0; long fbad = 0, kept = 0, total = 0;
    for (int I = 0; I < nc; ++I) {
        const double scale = std::fabs(Acm[I][I]);
        double rowSum = 0.0;
        for (auto& t : Acm[I]) rowSum += t.second;
        std::map<int, double> got;
        int s = I / kSlice, l = I % kSlice;
        for (int k = 0; k < Cz.sliceWidth[s]; ++k) {
            const size_t e = Cz.sliceStart[s] + (size_t)k * kSlice + l;
            if (Cz.col[e] == kPad) continue;
            got[Cz.col[e]] = -Cz.a[e];
        }
        double zSum = Cz.diag[I];
        for (auto& t : got) zSum += t.second;
        ferr = std::max(ferr, std::fabs(zSum - rowSum) / scale);
        for (auto& t : Acm[I]) {
            if (t.first == I) continue;
            ++total;
            const bool big = std::fabs(t.second) >= filt * std::sqrt(std::fabs(Acm[I][I] * Acm[t.first][t.first]));
            if (big != (got.count(t.first) > 0)) {
                // entries within rounding of the threshold may fall either way
                const double thr = filt * std::sqrt(std::fabs(Acm[I][I] * Acm[t.first][t.first]));
                if (std::fabs(std::fabs(t.second) - thr) > 1e-12 * scale) ++fbad;
            }
            if (got.count(t.first)) {
                ++kept;
                ferr = std::max(ferr, std::fabs(got[t.first] - t.second) / scale);
            }
        }
    }
    printf("n %d smoothed: galerkinSmoothed %.3f s; coarse operator max err %.2e (asymmetry %.2e), misplaced %ld; "
           "restriction weights err %.2e, bad %ld; G err %.2e, bad %ld; filtered: %ld of %ld entries kept, "
           "row-sum / entry err %.2e, misfiltered %ld\n",
           n, std::chrono::duration<double>(t6 - t5).count(), serr, asym, sbad, werr, wbad, gerr, gbad, kept, total,
           ferr, fbad);
    return (bad || qbad || err > 1e-12 || qerr > 1e-12 || sbad || wbad || gbad || serr > 1e-12 || werr > 1e-14 ||
            gerr > 1e-12 || fbad || ferr > 1e-12 || kept == 0 || kept == total) ? 1 : 0;
}
