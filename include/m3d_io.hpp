// m3d_io.hpp -- header-only on-disk formats around the solver (used by
// m3d_compat.hpp, usable on their own): the example drivers' input decks,
// the homogenization reports, the printer text dumps and a dependency-free
// VTK XML ImageData (.vti) writer.
//
//   read_struct_deck / read_therm_deck  example/3d/main.cpp:8-68,
//                                       example/3d_thermal/main.cpp:8-56
//   read_rve                            example/3d/main.cpp:72-93
//   write_report (elastic)              src/3d/FFT_analysis3D.cpp:905-996
//   write_report (thermal)              src/3dTherm/FFT_thermal3D.cpp:581-620
//   printer                             src/3d/FFT_analysis3D.cpp:1001-1053,
//                                       src/3dTherm/FFT_thermal3D.cpp:626-691
//   write_vti                           src/3d/FFT_analysis3D.cpp:1056-1104,
//                                       src/3dTherm/FFT_thermal3D.cpp:696-740,
//                                       src/percolation/percolation.cpp:443-473
//   printer_percolation                 src/percolation/percolation.cpp:417-441
//   printer_2d                          src/2d/FFT_analysis2D.cpp:328-360
//   save_rec_text                       src/rec/reconstruction.cpp:409-437 (text part)
//
// Text output goes through iostreams with the stream's default formatting,
// and matrices are printed the way Eigen's default stream operator prints
// them (columns right-aligned to the widest entry, one space apart), so a
// report reads as the reference's does.  Deviations: the .vti file is
// written by this header in ASCII (no VTK library), with the stress in
// components 6..11 (the reference reads sigma(i+6, .), past the matrix).
#pragma once

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include <stdexcept>

#include "m3d.h"

namespace m3d {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

namespace io {

// ------------------------------------------------------------------ decks
// The example drivers read "key value" pairs; a value whose atof() is
// nonzero is a number, anything else a string, in order of appearance; the
// last value is the load vector "a,b,c,...".
struct Deck {
    std::vector<double> num;
    std::vector<std::string> str;
    std::vector<double> load;
};

inline Deck read_deck(const std::string& path, size_t max_num, size_t nload)
{
    std::ifstream in(path);
    if (!in) throw Error(M3D_ERR_ARG, "deck: cannot open " + path);
    Deck d;
    std::string key, val;
    while (in >> key) {
        if (!(in >> val)) break;
        if (std::atof(val.c_str()) != 0.0 && d.num.size() < max_num) d.num.push_back(std::atof(val.c_str()));
        else d.str.push_back(val);
    }
    std::string rest = val;
    for (size_t i = 0; i < nload; ++i) {
        d.load.push_back(std::atof(rest.c_str()));
        const size_t c = rest.find(',');
        if (c != std::string::npos) rest = rest.substr(c + 1); // no comma: the value repeats, as there
    }
    return d;
}

// in_struct.txt: num = {N, E_0, v_0, E_1, v_1, tol_stress, tol_equilibrium},
// str = {rve file, Y|N, vti name, S|H, report name, load}
struct StructDeck {
    int N = 0;
    double E0 = 0, v0 = 0, E1 = 0, v1 = 0, tol_field = 0, tol_eq = 0;
    std::string rve, vti, report;
    bool write_vti = false, homogenize = false;
    double E[6] = {0, 0, 0, 0, 0, 0};
};

inline StructDeck read_struct_deck(const std::string& path)
{
    const Deck d = read_deck(path, 7, 6);
    if (d.num.size() < 7 || d.str.size() < 5) throw Error(M3D_ERR_ARG, "deck: the input data is incomplete or wrong");
    if (d.str[1] != "Y" && d.str[1] != "N") throw Error(M3D_ERR_ARG, "deck: vtiOutput_YorN must be Y or N");
    if (d.str[3] != "S" && d.str[3] != "H") throw Error(M3D_ERR_ARG, "deck: SingleAnalysisOrHomogenization must be S or H");
    StructDeck s;
    s.N = (int)d.num[0];
    s.E0 = d.num[1]; s.v0 = d.num[2]; s.E1 = d.num[3]; s.v1 = d.num[4];
    s.tol_field = d.num[5]; s.tol_eq = d.num[6];
    s.rve = d.str[0]; s.write_vti = d.str[1] == "Y"; s.vti = d.str[2];
    s.homogenize = d.str[3] == "H"; s.report = d.str[4];
    for (int i = 0; i < 6; ++i) s.E[i] = d.load[i];
    return s;
}

// in_therm.txt: num = {N, k_1 (matrix), k_0 (inclusion), tol_flux, tol_equilibrium},
// str = {rve file, Y|N, vti name, S|H, report name, gradient}
struct ThermDeck {
    int N = 0;
    double km = 0, ki = 0, tol_field = 0, tol_eq = 0;
    std::string rve, vti, report;
    bool write_vti = false, homogenize = false;
};

inline ThermDeck read_therm_deck(const std::string& path)
{
    const Deck d = read_deck(path, 5, 3);
    if (d.num.size() < 5 || d.str.size() < 5) throw Error(M3D_ERR_ARG, "deck: the input data is incomplete or wrong");
    if (d.str[1] != "Y" && d.str[1] != "N") throw Error(M3D_ERR_ARG, "deck: vtiOutput_YorN must be Y or N");
    if (d.str[3] != "S" && d.str[3] != "H") throw Error(M3D_ERR_ARG, "deck: SingleAnalysisOrHomogenization must be S or H");
    ThermDeck t;
    t.N = (int)d.num[0];
    t.km = d.num[1]; t.ki = d.num[2]; t.tol_field = d.num[3]; t.tol_eq = d.num[4];
    t.rve = d.str[0]; t.write_vti = d.str[1] == "Y"; t.vti = d.str[2];
    t.homogenize = d.str[3] == "H"; t.report = d.str[4];
    return t;
}

// one phase id per voxel, whitespace separated, voxel order z fastest
inline std::vector<int32_t> read_rve(const std::string& path, long long N)
{
    std::ifstream in(path);
    if (!in) throw Error(M3D_ERR_ARG, "rve: cannot open " + path);
    std::vector<int32_t> ids;
    ids.reserve((size_t)N);
    double a;
    while (in >> a) ids.push_back((int32_t)a);
    if ((long long)ids.size() != N) throw Error(M3D_ERR_ARG, "rve: dimension is not correct");
    return ids;
}

// ------------------------------------------------------------- formatting
// A column-major rows x cols matrix the way Eigen's operator<< prints it
// with the default format: each entry in the stream's current format,
// right-aligned to the widest entry, entries one space apart, rows on
// separate lines, no trailing newline.
inline void print_matrix(std::ostream& os, const double* a, int rows, int cols)
{
    std::streamsize width = 0;
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i) {
            std::stringstream ss;
            ss.copyfmt(os);
            ss << a[i + (size_t)j * rows];
            width = std::max<std::streamsize>(width, (std::streamsize)ss.str().size());
        }
    for (int i = 0; i < rows; ++i) {
        for (int j = 0; j < cols; ++j) {
            if (j) os << ' ';
            os.width(width);
            os << a[i + (size_t)j * rows];
        }
        if (i < rows - 1) os << '\n';
    }
}

inline std::vector<double> inverse(std::vector<double> A, int n)
{
    std::vector<double> X((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) X[i + (size_t)i * n] = 1.0;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[r + (size_t)c * n]) > std::fabs(A[p + (size_t)c * n])) p = r;
        for (int k = 0; k < n; ++k) {
            std::swap(A[c + (size_t)k * n], A[p + (size_t)k * n]);
            std::swap(X[c + (size_t)k * n], X[p + (size_t)k * n]);
        }
        const double d = A[c + (size_t)c * n];
        for (int k = 0; k < n; ++k) { A[c + (size_t)k * n] /= d; X[c + (size_t)k * n] /= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = A[r + (size_t)c * n];
            for (int k = 0; k < n; ++k) {
                A[r + (size_t)k * n] -= f * A[c + (size_t)k * n];
                X[r + (size_t)k * n] -= f * X[c + (size_t)k * n];
            }
        }
    }
    return X;
}

// ---------------------------------------------------------------- reports
// Inputs of an elastic report: grid, base, per-phase E and nu (material()
// columns 0, 1), volume fractions and the four 6x6 results (column-major).
struct ElasticReport {
    int Nx, Ny, Nz;
    bool strainBase;
    std::vector<double> El, v, vol;
    const double *CHomo, *SHomo, *epHomo, *sigHomo;
};

// ---------------------------------------------------------------- estimates
// Isotropic effective moduli: bulk K, shear G and Young's E = 9KG/(3K+G).
struct IsoModuli {
    double K, G, E;
};

// Hashin-Shtrikman bound of an isotropic n-phase composite (Hashin &
// Shtrikman, JMPS 11 (1963) 127-140) written as the comparison-medium average
// of Walpole (JMPS 14 (1966) 151-162): for comparison moduli (Kc, Gc),
//   K = [ sum_i c_i / (K_i + 4Gc/3) ]^-1 - 4Gc/3,
//   G = [ sum_i c_i / (G_i + Zc) ]^-1 - Zc,   Zc = Gc (9Kc + 8Gc) / (6 (Kc + 2Gc)).
// (Kc, Gc) = the smallest phase moduli gives the lower bound, the largest the
// upper bound.  For fractions c_i summing to one this is algebraically the
// 1963 paper's form (the homogenization report of FFT_analysis3D.cpp:905-942
// prints that bound).
inline IsoModuli hashin_shtrikman(const std::vector<double>& c, const std::vector<double>& K,
                                  const std::vector<double>& G, double Kc, double Gc)
{
    const double kshift = 4.0 * Gc / 3.0;
    const double gshift = Gc * (9.0 * Kc + 8.0 * Gc) / (6.0 * (Kc + 2.0 * Gc));
    double kinv = 0.0, ginv = 0.0;
    for (size_t i = 0; i < c.size(); ++i) {
        kinv += c[i] / (K[i] + kshift);
        ginv += c[i] / (G[i] + gshift);
    }
    IsoModuli m;
    m.K = 1.0 / kinv - kshift;
    m.G = 1.0 / ginv - gshift;
    m.E = 9.0 * m.K * m.G / (3.0 * m.K + m.G);
    return m;
}

// Conductivity of a matrix (km, fraction fm) with spherical inclusions (ki,
// fraction fi) by the Hashin-Shtrikman comparison medium kc (Hashin &
// Shtrikman, J. Appl. Phys. 33 (1962) 3125-3131):
//   k(kc) = [ fm / (km + 2kc) + fi / (ki + 2kc) ]^-1 - 2kc.
// kc = km is also the Mori-Tanaka (Maxwell Garnett) estimate of spheres in
// the matrix; kc = min / max of (km, ki) give the lower / upper bounds.
inline double hs_conductivity(double fm, double km, double fi, double ki, double kc)
{
    return 1.0 / (fm / (km + 2.0 * kc) + fi / (ki + 2.0 * kc)) - 2.0 * kc;
}

// homogenization report, FFT_analysis3D.cpp:905-996
inline void write_report(std::ostream& homo, const ElasticReport& ph)
{
    const ElasticReport& a = ph;
    const int n = (int)ph.El.size();
    std::vector<double> mu(n), la(n), K(n);
    for (int i = 0; i < n; ++i) {
        mu[i] = ph.El[i] / (1 + ph.v[i]) / 2.;
        la[i] = ph.v[i] * ph.El[i] / (1 + ph.v[i]) / (1 - 2 * ph.v[i]);
        K[i] = la[i] + (2. / 3.) * mu[i];
    }
    const auto Kr = std::minmax_element(K.begin(), K.end());
    const auto Gr = std::minmax_element(mu.begin(), mu.end());
    const IsoModuli lo = hashin_shtrikman(ph.vol, K, mu, *Kr.first, *Gr.first);
    const IsoModuli hi = hashin_shtrikman(ph.vol, K, mu, *Kr.second, *Gr.second);
    const char* rule = "      ----------------------------------------------------------------";

    homo << "*************************************************************************" << std::endl << std::endl;
    homo << "        F F T     H O M O G E N I Z A T I O N    B Y   M A T 3 D" << std::endl << std::endl;
    if (a.strainBase) homo << "                      Stran-Based Algorithm Is Used" << std::endl;
    else homo << "                      Stress-Based Algorithm Is Used" << std::endl;
    homo << "                           Nx=" << a.Nx << "   " << "Ny=" << a.Ny << "   " << "Nz=" << a.Nz << std::endl
         << std::endl;
    homo << rule << std::endl;
    for (int i = 0; i < n; ++i) {
        const std::string k = std::to_string(i) + " = ";
        homo << "                      Volume Fraction Phase " << k << ph.vol[i] << std::endl;
        homo << "     E" << k << ph.El[i] << "     " << "v" << k << ph.v[i] << "     " << "K" << k << K[i] << "     "
             << "G" << k << mu[i] << "     " << std::endl << std::endl;
    }
    homo << rule << std::endl << std::endl;
    homo << "                        Hashin-Shtrikman Bounds" << std::endl;
    homo << "         E_min=" << lo.E << "   " << "K_min=" << lo.K << "   " << "G_min=" << lo.G << std::endl;
    homo << "         E_max=" << hi.E << "   " << "K_max=" << hi.K << "   " << "G_max=" << hi.G << std::endl
         << std::endl << std::endl;
    homo << "************************************************************************" << std::endl << std::endl;
    homo << "Homogenized Elastic Modules:" << std::endl;
    homo << "E11=" << 1 / a.SHomo[0] << "     " << "E22=" << 1 / a.SHomo[7] << "     " << "E33=" << 1 / a.SHomo[14]
         << "     " << std::endl << std::endl;
    homo << std::endl << "C_Homogenized=" << std::endl << std::endl;
    print_matrix(homo, a.CHomo, 6, 6);
    homo << std::endl << rule << std::endl;
    homo << std::endl << "S_Homogenized=" << std::endl << std::endl;
    print_matrix(homo, a.SHomo, 6, 6);
    homo << std::endl << rule << std::endl;
    for (int i = 0; i < n; ++i) {
        std::vector<double> C(36, 0.0);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) C[r + 6 * c] = la[i] + (r == c ? 2 * mu[i] : 0.0);
        for (int r = 3; r < 6; ++r) C[r + 6 * r] = 2 * mu[i];
        const std::vector<double> S = inverse(C, 6);
        homo << std::endl << "C" << std::to_string(i) << "=" << std::endl;
        print_matrix(homo, C.data(), 6, 6);
        homo << std::endl << rule << std::endl;
        homo << std::endl << "S" << std::to_string(i) << "=" << std::endl;
        print_matrix(homo, S.data(), 6, 6);
        homo << std::endl << rule << std::endl;
    }
    homo << std::endl
         << "Homogenized Load Cases (Each row is one load case i.e (E11 E22 E33 E32 E31 E12) and Eij is summation of "
            "strain components in in  all voxels )"
         << std::endl << std::endl;
    print_matrix(homo, a.epHomo, 6, 6);
    homo << std::endl;
    homo << std::endl
         << "Homogenized stess (Each coulemn is (SS11 SS22 SS33 SS32 SS31 SS12) and SSij is summation of stress "
            "components in in all voxels )"
         << std::endl << std::endl;
    print_matrix(homo, a.sigHomo, 6, 6);
    homo << std::endl;
}

// thermal homogenization report, FFT_thermal3D.cpp:581-620.  rve_sum: the sum
// of the phase ids; the white-phase fraction RVE.sum()/xyz is an integer
// division there (:586) and is kept so.
struct ThermalReport {
    int Nx, Ny, Nz;
    double km, ki;
    long long rve_sum;
    const double *KHomo, *RHomo, *tempGradHomo, *heatFluxHomo;
};

inline void write_report(std::ostream& homo, const ThermalReport& a)
{
    const double km = a.km, ki = a.ki;
    const long long rve_sum = a.rve_sum;
    const long long xyz = (long long)a.Nx * a.Ny * a.Nz;
    const double fmat = (double)(rve_sum / xyz);
    // the report's three estimates (FFT_thermal3D.cpp:582-592): Mori-Tanaka
    // and the "lower" value are the matrix-referenced form, the "upper" the
    // inclusion-referenced one
    const double k_matrix = hs_conductivity(fmat, km, 1.0 - fmat, ki, km);
    const double k_incl = hs_conductivity(fmat, km, 1.0 - fmat, ki, ki);
    const double Km[9] = {km, 0, 0, 0, km, 0, 0, 0, km};
    const double Ki[9] = {ki, 0, 0, 0, ki, 0, 0, 0, ki};
    const double Rm[9] = {1 / km, 0, 0, 0, 1 / km, 0, 0, 0, 1 / km};
    homo << "Volume Fraction of the White Phase (or 1)=" << fmat << std::endl << std::endl;
    homo << "K phase matrix  ( or 1)=" << std::endl;
    print_matrix(homo, Km, 3, 3);
    homo << "     " << std::endl;
    homo << "K phase inclusion(or 0)=" << std::endl;
    print_matrix(homo, Ki, 3, 3);
    homo << "     " << std::endl;
    homo << "Nx=" << a.Nx << "   " << "Ny=" << a.Ny << "   " << a.Nz << std::endl << std::endl;
    homo << std::endl << "KHomogenized Mori-Tanaka=" << std::endl << std::endl;
    homo << k_matrix << std::endl;
    homo << std::endl << "Lower bound HS=" << std::endl << std::endl;
    homo << k_matrix << std::endl;
    homo << std::endl << "Upper bound LS=" << std::endl << std::endl;
    homo << k_incl << std::endl;
    homo << std::endl << "KHomogenized=" << std::endl << std::endl;
    print_matrix(homo, a.KHomo, 3, 3);
    homo << std::endl;
    homo << std::endl << "RHomogenized(Homogenized Resistivity tensor)=" << std::endl << std::endl;
    print_matrix(homo, a.RHomo, 3, 3);
    homo << std::endl;
    homo << std::endl << "Rm=" << std::endl << std::endl;
    print_matrix(homo, Rm, 3, 3);
    homo << std::endl;
    // the reference prints Km under the "Ri=" heading (:614-615)
    homo << std::endl << "Ri=" << std::endl << std::endl;
    print_matrix(homo, Km, 3, 3);
    homo << std::endl;
    homo << std::endl << std::endl << "--------------------------------------------------------------------" << std::endl
         << std::endl;
    homo << std::endl
         << "Homogenized Load Cases (Each row is one load case i.e (E11 E22 E33 ) and Eij is summation of temperatur "
            "gradient components in in  all voxels )"
         << std::endl << std::endl;
    print_matrix(homo, a.tempGradHomo, 3, 3);
    homo << std::endl;
    homo << std::endl
         << "Homogenized stess (Each coulemn is (SS11 SS22 SS33) and SSij is summation of heat flux in in all voxels )"
         << std::endl << std::endl;
    print_matrix(homo, a.heatFluxHomo, 3, 3);
    homo << std::endl;
}

// ---------------------------------------------------------------- printer
// FFT_analysis3D.cpp:1001-1053: one value per line per component in
// outFldr/{ep1..ep6,sig1..sig6}.txt and outFldr/ep_p ("ed": outFldr/d, the
// damage; "e": empty); for "ep"/"ed" also epsilon_sigma.txt (current
// directory) with the prop_homogenized rows after a zero row.
inline void printer(const std::string& outFldr, const std::string& behav, long long N, const double* eps,
                    const double* sig, const double* ep_or_d, const double* prop, int nInc)
{
    mkdir(outFldr.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    std::vector<std::string> names = {"ep1.txt",  "ep2.txt",  "ep3.txt",  "ep4.txt",  "ep5.txt",  "ep6.txt", "sig1.txt",
                                      "sig2.txt", "sig3.txt", "sig4.txt", "sig5.txt", "sig6.txt", "ep_p"};
    if (behav == "ed") names[12] = "d";
    for (int j = 0; j < 13; ++j) {
        std::ofstream out(outFldr + "/" + names[j]);
        for (long long i = 0; i < N; ++i) {
            if (j < 6) out << eps[j + 6 * i] << std::endl;
            else if (j < 12) out << sig[j - 6 + 6 * i] << std::endl;
            else if (behav != "e" && ep_or_d) out << ep_or_d[i] << std::endl;
        }
    }
    if (behav != "e" && prop) {
        std::ofstream op("epsilon_sigma.txt");
        op << "epsilon" << "  " << "sigma1" << "   " << "sigma2" << "   " << "sigma3" << "   " << "tau23" << "   "
           << "tau13" << "   " << "tau12" << "   " << (behav == "ep" ? "sigma_eff" : "d") << "   " << "epsilon_p_eff"
           << std::endl;
        op << 0 << "  " << 0 << "  " << 0 << "  " << 0 << "  " << 0 << "  " << 0 << "  " << 0 << "  " << 0 << "  " << 0
           << std::endl;
        for (int i = 0; i < nInc; ++i) {
            for (int c = 0; c < 9; ++c) op << (c ? "  " : "") << prop[i + (size_t)c * nInc];
            op << std::endl;
        }
    }
}

// FFT_thermal3D.cpp:626-691: folder/heatFlux{1,2,3}.txt, folder/tempGradient{1,2,3}.txt
inline void printer_thermal(const std::string& folder, long long N, const double* grad, const double* flux)
{
    mkdir(folder.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    for (int c = 0; c < 3; ++c) {
        std::ofstream q(folder + "/heatFlux" + std::to_string(c + 1) + ".txt");
        for (long long i = 0; i < N; ++i) q << flux[c + 3 * i] << std::endl;
    }
    for (int c = 0; c < 3; ++c) {
        std::ofstream g(folder + "/tempGradient" + std::to_string(c + 1) + ".txt");
        for (long long i = 0; i < N; ++i) g << grad[c + 3 * i] << std::endl;
    }
}

// percolation::printer, percolation.cpp:417-441: fold/clustersFile.txt (C1,
// one value per line) and fold/clustersdata.txt (a header with the interest
// phase and the number of phases, then one row per cluster: its label
// phasePlus + c and clusters_vol_frac[c], largest first).  The reference
// also echoes `fold` to stdout and takes two file names it never uses.
inline void printer_percolation(const std::string& fold, long long N, const int32_t* C1, int intrestPhase,
                                int noOfPhases, const std::vector<double>& clusters_vol_frac)
{
    mkdir(fold.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    {
        std::ofstream op0(fold + "/" + "clustersFile.txt");
        for (long long i = 0; i < N; ++i) op0 << C1[i] << std::endl;
    }
    std::ofstream op01(fold + "/" + "clustersdata.txt");
    op01 << "           intrested phase :  " << intrestPhase << "     Number of the phases: " << noOfPhases << std::endl;
    op01 << "-------------------------------------------------------------------------" << std::endl;
    op01 << "connected cluster name" << "        " << "volume fraction custer/volume fraction intrested phase"
         << std::endl;
    const int phasePlus = noOfPhases + 1;
    for (size_t c = 0; c < clusters_vol_frac.size(); ++c)
        op01 << "        " << (int)c + phasePlus << "                                   " << clusters_vol_frac[c]
             << std::endl;
}

// FFT_Analysis2D::printer, FFT_analysis2D.cpp:328-360: a "Results" folder
// (left empty, as there) and epsilon{1,2,3}.txt / sigma{1,2,3}.txt in the
// current directory, the real parts one per line.  eps_ri, sig_ri: 3 x xy
// complex values as (re, im) pairs.  The reference writes epsilon into the
// sigma files too; these hold sigma.
inline void printer_2d(long long xy, const double* eps_ri, const double* sig_ri)
{
    mkdir("Results", S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    for (int c = 0; c < 3; ++c) {
        std::ofstream e("epsilon" + std::to_string(c + 1) + ".txt");
        for (long long i = 0; i < xy; ++i) e << eps_ri[2 * (c + 3 * i)] << std::endl;
    }
    for (int c = 0; c < 3; ++c) {
        std::ofstream g("sigma" + std::to_string(c + 1) + ".txt");
        for (long long i = 0; i < xy; ++i) g << sig_ri[2 * (c + 3 * i)] << std::endl;
    }
}

// recPlan::saveRecImages, reconstruction.cpp:409-437, without the PNG slices
// (CImg): locFldr/rve_rec.txt (real part of mTilda, one per line) and
// locFldr/err.txt ("voloum fraction= " mean, then the first n - 1 errors of
// a run of n iterations; a run that stopped early has fewer, and the
// reference prints its unset entries).
inline void save_rec_text(const std::string& locFldr, long long xyz, const double* mTilda_ri, const double* err,
                          int nerr, int n)
{
    mkdir(locFldr.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH);
    double s = 0;
    {
        std::ofstream op2(locFldr + "/" + "rve_rec.txt");
        for (long long i = 0; i < xyz; ++i) {
            op2 << mTilda_ri[2 * i] << std::endl;
            s += mTilda_ri[2 * i];
        }
    }
    s = s / xyz;
    std::ofstream op1(locFldr + "/" + "err.txt");
    op1 << "voloum fraction= " << s << std::endl;
    for (int i = 0; i < n - 1 && i < nerr; ++i) op1 << err[i] << std::endl;
}

// --------------------------------------------------------------------- vti
// VTK XML ImageData with one point-data array "ImageScalars" of ncomp
// Float64 components; point (x, y, z) of the Nx x Ny x Nz image holds memory
// index x + Nx (y + Ny z), as the reference's writer fills it.
inline void write_vti(const std::string& fileName, int Nx, int Ny, int Nz, int ncomp,
                      const std::vector<std::pair<const double*, int>>& parts)
{
    std::ofstream f(fileName);
    if (!f) throw Error(M3D_ERR_ARG, "vti: cannot open " + fileName);
    const std::string ext = "0 " + std::to_string(Nx - 1) + " 0 " + std::to_string(Ny - 1) + " 0 " + std::to_string(Nz - 1);
    f << "<?xml version=\"1.0\"?>\n";
    f << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n";
    f << "  <ImageData WholeExtent=\"" << ext << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n";
    f << "    <Piece Extent=\"" << ext << "\">\n";
    f << "      <PointData Scalars=\"ImageScalars\">\n";
    f << "        <DataArray type=\"Float64\" Name=\"ImageScalars\" NumberOfComponents=\"" << ncomp
      << "\" format=\"ascii\">\n";
    f << std::setprecision(17);
    const long long N = (long long)Nx * Ny * Nz;
    for (long long i = 0; i < N; ++i) {
        int written = 0;
        for (const auto& pc : parts) {
            for (int c = 0; c < pc.second; ++c) {
                f << (written ? " " : "") << (pc.first ? pc.first[(size_t)pc.second * i + c] : 0.0);
                ++written;
            }
        }
        for (; written < ncomp; ++written) f << " 0";
        f << "\n";
    }
    f << "        </DataArray>\n      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n";
    f << "  </ImageData>\n</VTKFile>\n";
}

} // namespace io
} // namespace m3d
