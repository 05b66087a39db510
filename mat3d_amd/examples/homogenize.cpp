// homogenize.cpp -- C++ caller of the MI355X solver through m3d_compat.hpp,
// following the call sequence of the reference CLI example/3d/main.cpp:117-128
// (construct, material, setLoad, runAnalysis or homogenization).  The
// reference's text-deck parser is replaced by positional arguments:
//
//   homogenize N rve.txt E_0 v_0 E_1 v_1 tol_field tol_eq S|H [fp64|fp32]
//
// rve.txt holds N^3 phase ids (0 or 1), one per line, z fastest, as the
// reference's rve65.txt.  Prints the run outcome and, for H, C_eff.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "m3d_compat.hpp"

// minimal column-major matrix with the data()/rows()/cols() interface
template <typename T>
struct Mat {
    int r, c;
    std::vector<T> v;
    Mat(int r_, int c_) : r(r_), c(c_), v((size_t)r_ * c_) {}
    T& operator()(int i, int j) { return v[i + (size_t)j * r]; }
    const T* data() const { return v.data(); }
    int rows() const { return r; }
    int cols() const { return c; }
};

int main(int argc, char** argv)
{
    if (argc < 10) {
        std::fprintf(stderr, "usage: %s N rve.txt E_0 v_0 E_1 v_1 tol_field tol_eq S|H [fp64|fp32]\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    const long xyz = (long)n * n * n;
    Mat<int> RVE(1, (int)xyz);
    {
        std::ifstream f(argv[2]);
        if (!f) { std::fprintf(stderr, "cannot open %s\n", argv[2]); return 2; }
        long id = 0;
        double a;
        while (f >> a) {
            if (id >= xyz) { std::fprintf(stderr, "Dimension is not correct\n"); return 2; }
            RVE.v[id++] = (int)a;
        }
        if (id != xyz) { std::fprintf(stderr, "Dimension is not correct\n"); return 2; }
    }
    // material table as in example/3d/main.cpp:103-111 (columns E, v, s_y, H, n, P, Y0)
    Mat<double> mat(2, 7);
    const double rows[2][7] = {{std::atof(argv[3]), std::atof(argv[4]), 40e6, 100e6, 1, 8e-2, 8e2},
                               {std::atof(argv[5]), std::atof(argv[6]), 250e6, 8006, .2, 8e-4, 8e3}};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 7; ++j) mat(i, j) = rows[i][j];
    const double tol_f = std::atof(argv[7]), tol_eq = std::atof(argv[8]);
    const std::string mode = argv[9];
    const int prec = (argc > 10 && std::string(argv[10]) == "fp64") ? M3D_FP64 : M3D_FP32_POLISH;
    try {
        m3d::FFT_Analysis3D test(n, n, n, "e", prec);
        test.material(mat, RVE, 2);
        Mat<double> E(6, 1), X(6, 1);
        E(1, 0) = 0.1; // E=.000,0.1,0,0,0,0 (example/3d/in_struct.txt)
        X(1, 0) = 1;
        test.setLoad(E, X);
        if (mode == "S") {
            test.runAnalysis(tol_f, tol_eq, 15);
            std::printf("iterations %d status %d mean_sigma22 %.9e\n", test.info[0].iterations, test.info[0].status,
                        [&] { double s = 0; for (long i = 0; i < xyz; ++i) s += test.sigma[6 * i + 1]; return s / xyz; }());
        } else {
            const double ret = test.homogenization(tol_f, tol_eq, "myOutput");
            std::printf("base %s ret %.9e\nC_eff=\n", test.strainBase ? "strain" : "stress", ret);
            for (int i = 0; i < 6; ++i) {
                for (int j = 0; j < 6; ++j) std::printf(" %.9e", test.CHomo[i + 6 * j]);
                std::printf("\n");
            }
        }
    } catch (const m3d::Error& e) {
        std::fprintf(stderr, "m3d error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
