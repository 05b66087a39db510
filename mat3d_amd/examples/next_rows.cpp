// next_rows.cpp -- a C++ caller of the compat classes beside the 3D solvers:
// percolation, FFT_Analysis2D and recPlan (include/m3d_compat.hpp), written
// the way the reference's own drivers call its classes.
//
//   next_rows perc N ids.txt noOfPhases intrestPhase phase1 phase2 [folder]
//   next_rows plate2d N [print]
//   next_rows rec N tpcf.txt iterations beta seed [folder]
//
// Prints the public results one item per line for comparison.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <vector>

#include "m3d_compat.hpp"

namespace {

std::vector<double> read_values(const char* path, size_t n)
{
    std::ifstream f(path);
    std::vector<double> v(n);
    for (size_t i = 0; i < n; ++i)
        if (!(f >> v[i])) throw std::runtime_error("short file");
    return v;
}

int perc(int argc, char** argv)
{
    if (argc < 8) return 2;
    const int N = std::atoi(argv[2]);
    const std::vector<double> rve = read_values(argv[3], (size_t)N * N * N);
    m3d::percolation p(rve, N, N, N, std::atoi(argv[4]), std::atoi(argv[5]));
    p.runPercolationAnalysis(1);
    long long s1 = 0, s2 = 0;
    for (size_t i = 0; i < p.C1.size(); ++i) {
        s1 += p.C1[i];
        s2 = (s2 * 31 + p.C1[i]) % 1000000007LL;
    }
    std::printf("clusters %zu\nvolume %.17g\nc1 %lld %lld\n", p.all_clusters.size(), p.vol_intrstd_phs, s1, s2);
    std::printf("labels");
    for (int v : p.all_clusters) std::printf(" %d", v);
    std::printf("\nfractions");
    for (double v : p.clusters_vol_frac) std::printf(" %.17g", v);
    if (argc > 8) {
        // the reference's file output of the relabelled field
        p.printer(argv[8], "", "");
        p.vtiWriter(std::string(argv[8]) + "/clusters.vti");
    }
    p.SurfaceOfTwoPhase(rve, std::atoi(argv[6]), std::atoi(argv[7]));
    std::printf("\nsurface %lld\n", p.commonSurface);
    return 0;
}

int plate2d(int argc, char** argv)
{
    if (argc < 3) return 2;
    const int N = std::atoi(argv[2]);
    // one disc of radius N/4 in the middle: 1 = matrix, 0 = inclusion
    std::vector<double> rve((size_t)N * N, 1.0);
    for (int x = 0; x < N; ++x)
        for (int y = 0; y < N; ++y) {
            const double dx = x - N / 2, dy = y - N / 2;
            if (dx * dx + dy * dy <= (N / 4.0) * (N / 4.0)) rve[(size_t)y + (size_t)x * N] = 0.0;
        }
    double q[4] = {68.9e9, 0.35, 400e9, 0.23};
    m3d::FFT_Analysis2D a(N, N);
    a.material(q, rve);
    std::vector<std::complex<double>> E(3);
    a.setLoad(E);
    a.homogenization();
    if (argc > 3) a.printer(); // epsilon*.txt, sigma*.txt in the current directory
    for (int r = 0; r < 3; ++r)
        std::printf("%.17g %.17g %.17g\n", a.CHomo[r], a.CHomo[r + 3], a.CHomo[r + 6]);
    return 0;
}

int rec(int argc, char** argv)
{
    if (argc < 7) return 2;
    const int N = std::atoi(argv[2]);
    const std::vector<double> t = read_values(argv[3], (size_t)N * N * N);
    m3d::recPlan r(t.data(), N, N, N, std::strtoull(argv[6], nullptr, 10));
    r.runRec(std::atoi(argv[4]), std::atof(argv[5]));
    if (argc > 7) {
        r.saveRecImages(argv[7]);
        r.vtiWriter(std::string(argv[7]) + "/rec.vti", t.data());
    }
    double vf = 0;
    for (long long i = 0; i < r.xyz; ++i) vf += r.mTilda[2 * i];
    std::printf("iterations %zu\nerror %.17g\nfraction %.17g\nstatus %d\n", r.trackrealerrorA.size(),
                r.trackrealerrorA.empty() ? 0.0 : r.trackrealerrorA.back(), vf / (double)r.xyz, r.status);
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    try {
        if (argc > 1 && std::string(argv[1]) == "perc") return perc(argc, argv);
        if (argc > 1 && std::string(argv[1]) == "plate2d") return plate2d(argc, argv);
        if (argc > 1 && std::string(argv[1]) == "rec") return rec(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "next_rows: %s\n", e.what());
        return 1;
    }
    std::fprintf(stderr, "usage: next_rows perc|plate2d|rec ...\n");
    return 2;
}
