// mat3d_run.cpp -- the reference's command-line drivers on this library.
//
//   mat3d_run [--fp64] in_struct.txt            example/3d/main.cpp
//   mat3d_run [--fp64] --thermal in_therm.txt   example/3d_thermal/main.cpp
//
// Reads the same input decks and RVE files, runs the same sequence of calls
// (material, setLoad, runAnalysis or homogenization, vtiWriter) through the
// compat classes and writes the same output files (homogenization report,
// .vti).  Errors are reported and exit with status 1.
#include <cstdio>
#include <cstring>
#include <iostream>

#include "m3d_compat.hpp"

namespace {

int run_struct(const std::string& deck, int precision)
{
    const m3d::io::StructDeck d = m3d::io::read_struct_deck(deck);
    const int Nx = d.N, Ny = d.N, Nz = d.N;
    const std::vector<int32_t> rve = m3d::io::read_rve(d.rve, (long long)Nx * Ny * Nz);
    // material table of example/3d/main.cpp:104-111 (columns E, v, s_y, H, n, P, Y0)
    struct M {
        std::vector<double> v;
        int r;
        const double* data() const { return v.data(); }
        int rows() const { return r; }
    } mat{{d.E0, d.E1, d.v0, d.v1, 40e6, 250e6, 100e6, 8006, 1, .2, 8e-2, 8e-4, 8e2, 8e3}, 2};
    struct V {
        std::vector<double> v;
        const double* data() const { return v.data(); }
    } E{{d.E[0], d.E[1], d.E[2], d.E[3], d.E[4], d.E[5]}}, X{{0, 1, 0, 0, 0, 0}};
    struct I {
        const std::vector<int32_t>& v;
        const int32_t* data() const { return v.data(); }
    } ids{rve};
    m3d::FFT_Analysis3D test(Nx, Ny, Nz, "e", precision);
    test.material(mat, ids, 2);
    test.setLoad(E, X);
    if (!d.homogenize) test.runAnalysis(d.tol_field, d.tol_eq, 15);
    else {
        test.homogenization(d.tol_field, d.tol_eq, d.report);
        std::printf("E11=%.9g E22=%.9g E33=%.9g\n", test.E11, test.E22, test.E33);
    }
    if (d.write_vti) test.vtiWriter(d.vti);
    std::printf("iterations:");
    for (int w = 0; w < (d.homogenize ? 6 : 1); ++w) std::printf(" %d", test.info[w].iterations);
    std::printf("\n");
    return 0;
}

int run_therm(const std::string& deck, int precision)
{
    const m3d::io::ThermDeck d = m3d::io::read_therm_deck(deck);
    const int Nx = d.N, Ny = d.N, Nz = d.N;
    const std::vector<int32_t> rve = m3d::io::read_rve(d.rve, (long long)Nx * Ny * Nz);
    struct V {
        std::vector<double> v;
        const double* data() const { return v.data(); }
    } TG{{1, 0, 0}}; // example/3d_thermal/main.cpp:89-90 applies (1, 0, 0)
    struct I {
        const std::vector<int32_t>& v;
        const int32_t* data() const { return v.data(); }
    } ids{rve};
    m3d::FFT_analysis3d_thermal test(Nx, Ny, Nz, precision);
    test.setLoad(TG);
    test.material(d.km, d.ki, ids);
    if (!d.homogenize) test.runAnalysis(d.tol_field, d.tol_eq);
    else {
        test.homogenization(d.tol_field, d.tol_eq, d.report);
        std::printf("K11=%.9g K22=%.9g K33=%.9g\n", test.K11, test.K22, test.K33);
    }
    if (d.write_vti) test.vtiWriter(d.vti);
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    int precision = M3D_FP32_POLISH;
    bool thermal = false;
    std::string deck;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--fp64")) precision = M3D_FP64;
        else if (!std::strcmp(argv[i], "--thermal")) thermal = true;
        else deck = argv[i];
    }
    if (deck.empty()) {
        std::fprintf(stderr, "usage: %s [--fp64] [--thermal] deck.txt\n", argv[0]);
        return 2;
    }
    try {
        return thermal ? run_therm(deck, precision) : run_struct(deck, precision);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "mat3d_run: %s\n", e.what());
        return 1;
    }
}
