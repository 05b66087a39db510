// io_probe.cpp -- drives include/m3d_io.hpp for tests/test_io.py (no GPU,
// no library: the format code only).
//   io_probe matrix                 a fixed 3x3 matrix in the Eigen stream format
//   io_probe deck <in_struct.txt>   the parsed fields of a structural deck
//   io_probe tdeck <in_therm.txt>   the parsed fields of a thermal deck
//   io_probe report_e < spec        elastic report (spec: Nx Ny Nz strainBase nph,
//                                   El[nph] v[nph] vol[nph], C S epHomo sigHomo)
//   io_probe report_t < spec        thermal report (spec: Nx Ny Nz km ki rve_sum, K R G Q)
//   io_probe bounds < spec          Hashin-Shtrikman bounds and conductivity estimates
//   io_probe vti <file>             a 2x3x4 image, components (i, -i, 0.5 i)
#include <algorithm>
#include <cstring>
#include <iostream>
#include <vector>

#include "m3d_io.hpp"

using namespace m3d::io;

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    if (mode == "matrix") {
        const double a[9] = {1.5, -2, 3e6, 0.25, 1, 0, 123456789, 0.001, -7};
        print_matrix(std::cout, a, 3, 3);
        std::cout << std::endl;
    } else if (mode == "deck") {
        const StructDeck d = read_struct_deck(argv[2]);
        std::cout << d.N << " " << d.E0 << " " << d.v0 << " " << d.E1 << " " << d.v1 << " " << d.tol_field << " "
                  << d.tol_eq << "\n" << d.rve << " " << d.write_vti << " " << d.vti << " " << d.homogenize << " "
                  << d.report << "\n";
        for (double e : d.E) std::cout << e << " ";
        std::cout << "\n";
    } else if (mode == "tdeck") {
        const ThermDeck d = read_therm_deck(argv[2]);
        std::cout << d.N << " " << d.km << " " << d.ki << " " << d.tol_field << " " << d.tol_eq << "\n"
                  << d.rve << " " << d.write_vti << " " << d.vti << " " << d.homogenize << " " << d.report << "\n";
    } else if (mode == "report_e") {
        ElasticReport r;
        int sb, n;
        std::cin >> r.Nx >> r.Ny >> r.Nz >> sb >> n;
        r.strainBase = sb != 0;
        r.El.resize(n); r.v.resize(n); r.vol.resize(n);
        for (auto& x : r.El) std::cin >> x;
        for (auto& x : r.v) std::cin >> x;
        for (auto& x : r.vol) std::cin >> x;
        std::vector<double> m(144);
        for (auto& x : m) std::cin >> x;
        r.CHomo = m.data(); r.SHomo = m.data() + 36; r.epHomo = m.data() + 72; r.sigHomo = m.data() + 108;
        write_report(std::cout, r);
    } else if (mode == "report_t") {
        ThermalReport r;
        std::cin >> r.Nx >> r.Ny >> r.Nz >> r.km >> r.ki >> r.rve_sum;
        std::vector<double> m(36);
        for (auto& x : m) std::cin >> x;
        r.KHomo = m.data(); r.RHomo = m.data() + 9; r.tempGradHomo = m.data() + 18; r.heatFluxHomo = m.data() + 27;
        write_report(std::cout, r);
    } else if (mode == "bounds") {
        // spec: n c[n] K[n] G[n] fm km fi ki -> HS lower K G E, upper K G E,
        // conductivity with the matrix, then the inclusion, as comparison medium
        int n;
        std::cin >> n;
        std::vector<double> c(n), K(n), G(n);
        for (auto& x : c) std::cin >> x;
        for (auto& x : K) std::cin >> x;
        for (auto& x : G) std::cin >> x;
        double fm, km, fi, ki;
        std::cin >> fm >> km >> fi >> ki;
        const auto Kr = std::minmax_element(K.begin(), K.end());
        const auto Gr = std::minmax_element(G.begin(), G.end());
        const IsoModuli lo = hashin_shtrikman(c, K, G, *Kr.first, *Gr.first);
        const IsoModuli hi = hashin_shtrikman(c, K, G, *Kr.second, *Gr.second);
        std::cout.precision(17);
        std::cout << lo.K << " " << lo.G << " " << lo.E << " " << hi.K << " " << hi.G << " " << hi.E << " "
                  << hs_conductivity(fm, km, fi, ki, km) << " " << hs_conductivity(fm, km, fi, ki, ki) << "\n";
    } else if (mode == "vti") {
        std::vector<double> a(24 * 2), b(24);
        for (int i = 0; i < 24; ++i) { a[2 * i] = i; a[2 * i + 1] = -i; b[i] = 0.5 * i; }
        write_vti(argv[2], 2, 3, 4, 4, {{a.data(), 2}, {b.data(), 1}});
    } else if (mode == "p2d") {
        // 2D printer in the current directory: 4 points, eps_c(i) = 10c + i + 0.5, sig = -eps
        std::vector<double> e(24), g(24);
        for (int i = 0; i < 4; ++i)
            for (int c = 0; c < 3; ++c) {
                e[2 * (c + 3 * i)] = 10 * c + i + 0.5;
                e[2 * (c + 3 * i) + 1] = 7;
                g[2 * (c + 3 * i)] = -(10 * c + i + 0.5);
            }
        printer_2d(4, e.data(), g.data());
    } else if (mode == "rec") {
        // mTilda real 0, 1, 0, 1 (mean 0.5), errors 3, 2, 1 of a 3-iteration run
        std::vector<double> m = {0, 9, 1, 9, 0, 9, 1, 9}, err = {3, 2, 1};
        save_rec_text(argv[2], 4, m.data(), err.data(), 3, 3);
    } else if (mode == "perc") {
        // percolation printer: C1 = 0..23, fractions 0.5, 0.25, 1/3 (phases 2, interest 1)
        std::vector<int32_t> c1(24);
        for (int i = 0; i < 24; ++i) c1[i] = i;
        printer_percolation(argv[2], 24, c1.data(), 1, 2, {0.5, 0.25, 1.0 / 3.0});
    } else {
        return 2;
    }
    return 0;
}
