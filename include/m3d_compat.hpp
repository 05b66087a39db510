// m3d_compat.hpp -- header-only C++ wrapper that re-exposes the reference's
// solver classes on top of the C-ABI in m3d.h.
//
//   FFT_Analysis3D          -- include/FFT_analysis3D.hpp:36-215
//   FFT_analysis3d_thermal  -- include/FFT_thermal3D.hpp:37-154
//   recPlan (TPCF form)     -- include/reconstruction.hpp:55-76
//   FFT_Analysis2D          -- include/FFT_analysis2D.hpp:20-99
//   percolation             -- include/percolation.hpp:36-71
//
// Method names, argument order and meaning are the reference's.  The matrix
// arguments are templates: anything with data()/rows()/cols() in the
// reference's column-major order works, Eigen's MatrixDdd / MatrixIdd
// included (no Eigen dependency here).  Public result members are plain
// std::vector<double> in column-major order (CHomo, SHomo, epHomo, sigHomo:
// 6x6; KHomo, RHomo, tempGradHomo, heatFluxHomo: 3x3; epsilon/sigma: 6 x N).
// Failures throw m3d::Error (the reference exits or terminates instead,
// e.g. src/3d/FFT_analysis3D.cpp:672-673).
#pragma once

#include <complex>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "m3d.h"
#include "m3d_io.hpp"

namespace m3d {

class FFT_Analysis3D {
  public:
    // FFT_Analysis3D(int Nx1, int Ny1, int Nz1, string mater_behav), FFT_analysis3D.hpp:196
    FFT_Analysis3D(int Nx1, int Ny1, int Nz1, const std::string& mater_behav, int precision = M3D_FP32_POLISH)
        : Nx(Nx1), Ny(Ny1), Nz(Nz1), xyz(Nx1 * Ny1 * Nz1), behav_(mater_behav)
    {
        int rc = m3d_elastic_create(Nx1, Ny1, Nz1, mater_behav.c_str(), precision, &h_);
        if (rc != M3D_OK) throw Error(rc, "m3d_elastic_create failed");
    }
    ~FFT_Analysis3D() { m3d_elastic_destroy(h_); }
    FFT_Analysis3D(const FFT_Analysis3D&) = delete;
    FFT_Analysis3D& operator=(const FFT_Analysis3D&) = delete;

    // void material(MatrixDdd& mat, MatrixIdd& RVE1, int noPhases), FFT_analysis3D.hpp:200
    template <class Mat, class Ids>
    void material(const Mat& mat, const Ids& RVE1, int noPhases)
    {
        std::vector<int32_t> ids(RVE1.data(), RVE1.data() + xyz);
        // mat is noPhases x 7 column-major; copy the first noPhases rows
        std::vector<double> m((size_t)noPhases * 7);
        for (int j = 0; j < 7; ++j)
            for (int i = 0; i < noPhases; ++i) m[i + (size_t)j * noPhases] = mat.data()[i + (size_t)j * mat.rows()];
        chk(m3d_elastic_material(h_, m.data(), ids.data(), noPhases));
        El_.assign(m.begin(), m.begin() + noPhases);
        v_.assign(m.begin() + noPhases, m.begin() + 2 * noPhases);
        vol_fractions.assign(noPhases, 0.0);
        int sb = 1;
        chk(m3d_elastic_info(h_, vol_fractions.data(), &sb, &la0, &mu0));
        strainBase = sb != 0;
    }

    // void setLoad(MatrixDdd& E1, MatrixDdd& X0), FFT_analysis3D.hpp:202
    template <class V>
    void setLoad(const V& E1, const V& X01)
    {
        for (int i = 0; i < 6; ++i) { E[i] = E1.data()[i]; X0[i] = X01.data()[i]; }
        chk(m3d_elastic_set_load(h_, E, X0));
    }

    // void runAnalysis(double tol_field, double tol_eq, int nInc1), FFT_analysis3D.hpp:206
    void runAnalysis(double tol_field, double tol_eq, int nInc1)
    {
        chk(m3d_elastic_run(h_, tol_field, tol_eq, nInc1, &info[0]));
        fetch_fields(0);
        last_run_basic_ = behav_ != "e";
        if (last_run_basic_) {
            // prop_homogenized (nInc x 9) and the per-voxel history
            int rows = 0;
            chk(m3d_elastic_get_prop(h_, nullptr, &rows));
            nInc = rows;
            prop_homogenized.assign((size_t)9 * rows, 0.0);
            chk(m3d_elastic_get_prop(h_, prop_homogenized.data(), &rows));
            std::vector<double> a((size_t)xyz), b((size_t)xyz);
            chk(m3d_elastic_get_state(h_, a.data(), b.data()));
            if (behav_ == "ep") { epsilon_p_eff = a; sigma_eff = b; }
            else d = a;
        }
    }

    // double homogenization(double, double, string), FFT_analysis3D.hpp:209;
    // a non-empty outputName writes the report outputName + ".txt" (:859)
    double homogenization(double tol_field, double tol_eq, const std::string& outputName = "")
    {
        CHomo.resize(36); SHomo.resize(36); epHomo.resize(36); sigHomo.resize(36);
        double ret = 0;
        chk(m3d_elastic_homogenization(h_, tol_field, tol_eq, CHomo.data(), SHomo.data(), epHomo.data(),
                                       sigHomo.data(), &ret, info));
        E11 = 1. / SHomo[0]; E22 = 1. / SHomo[7]; E33 = 1. / SHomo[14];
        fetch_fields(5); // the reference keeps the fields of the last load case
        last_run_basic_ = false;
        if (!outputName.empty()) {
            std::ofstream homo(outputName + ".txt");
            io::ElasticReport r{Nx, Ny, Nz, strainBase, El_, v_, vol_fractions,
                                CHomo.data(), SHomo.data(), epHomo.data(), sigHomo.data()};
            io::write_report(homo, r);
        }
        return ret;
    }

    // void printer(string outFldr), FFT_analysis3D.hpp:211
    void printer(const std::string& outFldr = "results")
    {
        const double* h = behav_ == "ep" ? epsilon_p_eff.data() : (behav_ == "ed" ? d.data() : nullptr);
        io::printer(outFldr, last_run_basic_ ? behav_ : std::string("e"), xyz, epsilon.data(), sigma.data(),
                    last_run_basic_ ? h : nullptr, last_run_basic_ ? prop_homogenized.data() : nullptr, nInc);
    }

    // void vtiWriter(string fileName), FFT_analysis3D.hpp:213
    void vtiWriter(const std::string& fileName)
    {
        const double* h = nullptr;
        const double* se = nullptr;
        if (last_run_basic_) {
            h = behav_ == "ep" ? epsilon_p_eff.data() : d.data();
            se = behav_ == "ep" ? sigma_eff.data() : nullptr;
        }
        io::write_vti(fileName, Nx, Ny, Nz, 14, {{epsilon.data(), 6}, {sigma.data(), 6}, {h, 1}, {se, 1}});
    }

    void set_option(const char* key, double v) { chk(m3d_elastic_set_option(h_, key, v)); }

    // public members of the reference class
    int Nx, Ny, Nz, xyz;
    double E[6] = {0, 0, 0, 0, 0, 0}, X0[6] = {0, 0, 0, 0, 0, 0};
    double la0 = 0, mu0 = 0, E11 = 0, E22 = 0, E33 = 0;
    bool strainBase = true;
    std::vector<double> vol_fractions, CHomo, SHomo, epHomo, sigHomo;
    std::vector<double> epsilon, sigma; // 6 x N column-major
    std::vector<double> prop_homogenized; // nInc x 9 column-major ("ep"/"ed")
    std::vector<double> epsilon_p_eff, sigma_eff, d; // N ("ep": the first two, "ed": d)
    int nInc = 0;
    m3d_run_info info[6] = {};

  private:
    void chk(int rc)
    {
        if (rc != M3D_OK) throw Error(rc, m3d_elastic_last_error(h_));
    }
    std::string behav_;
    bool last_run_basic_ = false;
    std::vector<double> El_, v_;
    void fetch_fields(int lc)
    {
        epsilon.resize((size_t)6 * xyz);
        sigma.resize((size_t)6 * xyz);
        chk(m3d_elastic_get_fields(h_, lc, epsilon.data(), sigma.data()));
    }
    m3d_elastic* h_ = nullptr;
};

class FFT_analysis3d_thermal {
  public:
    // FFT_analysis3d_thermal(int Nx1, int Ny1, int Nz1), FFT_thermal3D.hpp:133
    FFT_analysis3d_thermal(int Nx1, int Ny1, int Nz1, int precision = M3D_FP32_POLISH)
        : Nx(Nx1), Ny(Ny1), Nz(Nz1), xyz(Nx1 * Ny1 * Nz1)
    {
        int rc = m3d_thermal_create(Nx1, Ny1, Nz1, precision, &h_);
        if (rc != M3D_OK) throw Error(rc, "m3d_thermal_create failed");
    }
    ~FFT_analysis3d_thermal() { m3d_thermal_destroy(h_); }
    FFT_analysis3d_thermal(const FFT_analysis3d_thermal&) = delete;
    FFT_analysis3d_thermal& operator=(const FFT_analysis3d_thermal&) = delete;

    // void material(double km, double ki, MatrixIdd& RVE1), FFT_thermal3D.hpp:138
    template <class Ids>
    void material(double km, double ki, const Ids& RVE1)
    {
        std::vector<int32_t> ids(RVE1.data(), RVE1.data() + xyz);
        chk(m3d_thermal_material(h_, km, ki, ids.data()));
        if (have_load_) chk(m3d_thermal_set_load(h_, E));
        km_ = km;
        ki_ = ki;
        rve_sum_ = 0;
        for (int32_t q : ids) rve_sum_ += q;
    }

    // void setLoad(MatrixDdd& E1), FFT_thermal3D.hpp:140.  The reference example
    // calls it before material() (example/3d_thermal/main.cpp:94-95); the load
    // is kept and re-applied.
    template <class V>
    void setLoad(const V& E1)
    {
        for (int i = 0; i < 3; ++i) E[i] = E1.data()[i];
        have_load_ = true;
        chk(m3d_thermal_set_load(h_, E));
    }

    // void runAnalysis(double tol_field, double tol_eq), FFT_thermal3D.hpp:145
    void runAnalysis(double tol_field, double tol_eq)
    {
        chk(m3d_thermal_run(h_, tol_field, tol_eq, &info[0]));
        fetch_fields(0);
    }

    // void homogenization(double, double, string), FFT_thermal3D.hpp:148
    // a non-empty filename writes the report filename + ".txt" (:491)
    void homogenization(double tol_field, double tol_eq, const std::string& filename = "")
    {
        KHomo.resize(9); RHomo.resize(9); tempGradHomo.resize(9); heatFluxHomo.resize(9);
        chk(m3d_thermal_homogenization(h_, tol_field, tol_eq, KHomo.data(), RHomo.data(), tempGradHomo.data(),
                                       heatFluxHomo.data(), info));
        K11 = KHomo[0]; K22 = KHomo[4]; K33 = KHomo[8];
        fetch_fields(2);
        if (!filename.empty()) {
            std::ofstream homo(filename + ".txt");
            io::ThermalReport r{Nx, Ny, Nz, km_, ki_, rve_sum_,
                                KHomo.data(), RHomo.data(), tempGradHomo.data(), heatFluxHomo.data()};
            io::write_report(homo, r);
        }
    }

    // void printer(string folderName), FFT_thermal3D.hpp:150
    void printer(const std::string& folderName) { io::printer_thermal(folderName, xyz, tempGradient.data(), heatFlux.data()); }

    // void vtiWriter(string fileName), FFT_thermal3D.hpp:152
    void vtiWriter(const std::string& fileName)
    {
        io::write_vti(fileName, Nx, Ny, Nz, 6, {{tempGradient.data(), 3}, {heatFlux.data(), 3}});
    }

    void set_option(const char* key, double v) { chk(m3d_thermal_set_option(h_, key, v)); }

    int Nx, Ny, Nz, xyz;
    double E[3] = {0, 0, 0};
    double K11 = 0, K22 = 0, K33 = 0;
    std::vector<double> KHomo, RHomo, tempGradHomo, heatFluxHomo;
    std::vector<double> tempGradient, heatFlux; // 3 x N column-major
    m3d_run_info info[3] = {};

  private:
    void chk(int rc)
    {
        if (rc != M3D_OK) throw Error(rc, m3d_thermal_last_error(h_));
    }
    void fetch_fields(int lc)
    {
        tempGradient.resize((size_t)3 * xyz);
        heatFlux.resize((size_t)3 * xyz);
        chk(m3d_thermal_get_fields(h_, lc, tempGradient.data(), heatFlux.data()));
    }
    m3d_thermal* h_ = nullptr;
    bool have_load_ = false;
    double km_ = 0, ki_ = 0;
    long long rve_sum_ = 0;
};

// recPlan(double *TPCF, int Nx1, int Ny1, int Nz1), reconstruction.hpp:55, and
// runRec(int n, double b), :66.  mTilda holds Nx*Ny*Nz complex values as
// (re, im) pairs, fftw_complex's layout; trackrealerrorA one error per
// iteration.  The random phases come from a seeded stream (seed).
class recPlan {
  public:
    recPlan(const double* TPCF, int Nx1, int Ny1, int Nz1, uint64_t seed = 1)
        : xyz((long long)Nx1 * Ny1 * Nz1), Nx_(Nx1), Ny_(Ny1), Nz_(Nz1), seed_(seed)
    {
        int rc = m3d_rec_create(TPCF, Nx1, Ny1, Nz1, &h_);
        if (rc != M3D_OK) throw Error(rc, "m3d_rec_create failed");
    }
    ~recPlan() { m3d_rec_destroy(h_); }
    recPlan(const recPlan&) = delete;
    recPlan& operator=(const recPlan&) = delete;

    void runRec(int n, double b)
    {
        n_ = n;
        int it = 0;
        chk(m3d_rec_run(h_, n, b, seed_, &it, &status, &search_steps));
        mTilda.assign((size_t)(2 * xyz), 0.0);
        trackrealerrorA.assign((size_t)(it > 0 ? it : 1), 0.0);
        int k = 0;
        chk(m3d_rec_result(h_, mTilda.data(), trackrealerrorA.data(), (int)trackrealerrorA.size(), &k));
        trackrealerrorA.resize((size_t)k);
    }

    // void saveRecImages(string locFldr), :70 -- the text files (rve_rec.txt,
    // err.txt); the PNG slices need CImg and are not written
    void saveRecImages(const std::string& locFldr)
    {
        io::save_rec_text(locFldr, xyz, mTilda.data(), trackrealerrorA.data(), (int)trackrealerrorA.size(), n_);
    }
    // void vtiWriter(string fileName, double* tpcf), :72 -- real part of
    // mTilda and tpcf as two Float64 components
    void vtiWriter(const std::string& fileName, const double* tpcf)
    {
        std::vector<double> a((size_t)(2 * xyz));
        for (long long i = 0; i < xyz; ++i) {
            a[(size_t)(2 * i)] = mTilda[(size_t)(2 * i)];
            a[(size_t)(2 * i + 1)] = tpcf[i];
        }
        io::write_vti(fileName, Nx_, Ny_, Nz_, 2, {{a.data(), 2}});
    }

    std::vector<double> mTilda, trackrealerrorA;
    int status = 0, search_steps = 0;
    long long xyz;

  private:
    void chk(int rc)
    {
        if (rc != M3D_OK) throw Error(rc, m3d_rec_last_error(h_));
    }
    m3d_rec* h_ = nullptr;
    int Nx_, Ny_, Nz_, n_ = 0;
    uint64_t seed_;
};

// FFT_Analysis2D(int Nx1, int Ny1), FFT_analysis2D.hpp:82.  epsilon / sigma:
// 3 x Nx*Ny complex values, (re, im) pairs; CHomo, epHomo, sigHomo: 3x3
// column-major after homogenization().
class FFT_Analysis2D {
  public:
    FFT_Analysis2D(int Nx1, int Ny1) : xy((long long)Nx1 * Ny1)
    {
        int rc = m3d_2d_create(Nx1, Ny1, &h_);
        if (rc != M3D_OK) throw Error(rc, "m3d_2d_create failed");
    }
    ~FFT_Analysis2D() { m3d_2d_destroy(h_); }
    FFT_Analysis2D(const FFT_Analysis2D&) = delete;
    FFT_Analysis2D& operator=(const FFT_Analysis2D&) = delete;

    // void material(double *properties, MatrixDdd &RVE1), :89
    template <class Rve>
    void material(const double* properties, const Rve& RVE1)
    {
        chk(m3d_2d_material(h_, properties, RVE1.data()));
    }
    // void setLoad(MatrixCdd &E1), :92 (three complex values)
    template <class V>
    void setLoad(const V& E1)
    {
        const std::complex<double>* e = E1.data();
        double ri[6];
        for (int i = 0; i < 3; ++i) {
            ri[2 * i] = e[i].real();
            ri[2 * i + 1] = e[i].imag();
        }
        chk(m3d_2d_set_load(h_, ri));
    }
    // void runAnalysis(), :95
    void runAnalysis()
    {
        chk(m3d_2d_run(h_, &itNo));
        fetch();
    }
    // void homogenization(), :97
    void homogenization()
    {
        CHomo.assign(9, 0.0);
        epHomo.assign(9, 0.0);
        sigHomo.assign(9, 0.0);
        chk(m3d_2d_homogenization(h_, CHomo.data(), epHomo.data(), sigHomo.data()));
        fetch();
    }

    // void printer(), :103 -- epsilon{1,2,3}.txt and sigma{1,2,3}.txt
    void printer() { io::printer_2d(xy, epsilon.data(), sigma.data()); }

    std::vector<double> epsilon, sigma, CHomo, epHomo, sigHomo;
    int itNo = 0;
    long long xy;

  private:
    void fetch()
    {
        epsilon.resize((size_t)(6 * xy));
        sigma.resize((size_t)(6 * xy));
        chk(m3d_2d_get_fields(h_, epsilon.data(), sigma.data()));
    }
    void chk(int rc)
    {
        if (rc != M3D_OK) throw Error(rc, m3d_2d_last_error(h_));
    }
    m3d_2d* h_ = nullptr;
};

// percolation(MatrixDdd &RVE, int Nx1, int Ny1, int Nz1, int noOfPhases,
// int intrestPhase), percolation.hpp:39.  RVE: anything with data() over
// Nx*Ny*Nz values in the reference's order (MatrixDdd included); values are
// truncated to int as the reference's C1[i] = RVE(i).  C1 is the relabelled
// field after runPercolationAnalysis (the given field after
// SurfaceOfTwoPhase).
class percolation {
  public:
    template <class Rve>
    percolation(const Rve& RVE, int Nx1, int Ny1, int Nz1, int noOfPhases1, int intrestPhase1)
        : NX(Nx1), NY(Ny1), NZ(Nz1), noOfPhases(noOfPhases1), intrestPhase(intrestPhase1)
    {
        ids(RVE);
        int rc = m3d_perc_create(C1.data(), NX, NY, NZ, noOfPhases, intrestPhase, &h_);
        if (rc != M3D_OK) throw Error(rc, "m3d_perc_create failed");
    }
    ~percolation() { m3d_perc_destroy(h_); }
    percolation(const percolation&) = delete;
    percolation& operator=(const percolation&) = delete;

    // void runPercolationAnalysis(int total_iter), :42
    void runPercolationAnalysis(int total_iter)
    {
        int k = 0;
        chk(m3d_perc_run(h_, total_iter, &k, &vol_intrstd_phs));
        all_clusters.assign((size_t)k, 0);
        clusters_vol_frac.assign((size_t)k, 0.0);
        chk(m3d_perc_result(h_, C1.data(), all_clusters.data(), clusters_vol_frac.data(), k));
    }
    // void SurfaceOfTwoPhase(MatrixDdd &RVE, int phase1, int phase2), :47
    template <class Rve>
    void SurfaceOfTwoPhase(const Rve& RVE, int phase1, int phase2)
    {
        ids(RVE);
        chk(m3d_perc_surface(h_, C1.data(), phase1, phase2, &commonSurface));
    }

    // void printer(string fold, string clustresSizedata, string largestCluster), :53
    void printer(const std::string& fold, const std::string& clustresSizedata = "", const std::string& largestCluster = "")
    {
        (void)clustresSizedata;
        (void)largestCluster;
        io::printer_percolation(fold, (long long)C1.size(), C1.data(), intrestPhase, noOfPhases, clusters_vol_frac);
    }
    // void vtiWriter(string fileName), :58 -- C1 as one Float64 component
    void vtiWriter(const std::string& fileName)
    {
        std::vector<double> d(C1.begin(), C1.end());
        io::write_vti(fileName, NX, NY, NZ, 1, {{d.data(), 1}});
    }

    int NX, NY, NZ, noOfPhases, intrestPhase;
    long long commonSurface = 0;
    std::vector<int> all_clusters;
    std::vector<double> clusters_vol_frac;
    double vol_intrstd_phs = 0;
    std::vector<int32_t> C1;

  private:
    template <class Rve>
    void ids(const Rve& RVE)
    {
        const size_t n = (size_t)NX * NY * NZ;
        C1.resize(n);
        for (size_t i = 0; i < n; ++i) C1[i] = (int32_t)RVE.data()[i];
    }
    void chk(int rc)
    {
        if (rc != M3D_OK) throw Error(rc, m3d_perc_last_error(h_));
    }
    m3d_perc* h_ = nullptr;
};

} // namespace m3d
