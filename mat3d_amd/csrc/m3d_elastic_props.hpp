// m3d_elastic_props.hpp -- host-side FP64 material processing and effective
// tensor assembly of the elastic solver, shared by the single-RVE handle and
// the batch API.  Included after m3d_engine.hpp by m3d_elastic.hip only.
#pragma once

namespace {

// Per-RVE material state, FFT_analysis3D.cpp:420-549 (material()).
struct ElasticProps {
    int nph = 0;
    int strain_base = 1;           // 1: strain base (Gamma0, C_acc); 0: stress base
    double la0 = 0, mu0 = 0;       // reference medium
    std::vector<double> la, mu, K; // per phase
    std::vector<double> vol;       // volume fractions
    std::vector<double> Cm, Sm;    // per phase 6x6 column-major
    std::vector<double> mtab;      // per phase {la, mu, Sy, H, n, P, Y0, 0}
};

// material(): moduli (:441-466), volume fractions (:468-480), reference
// medium and base (:498-521).  `mat` is nph x 7 column-major
// [E, nu, sigma_y, H, n, P, Y0]; ids receives the N phase ids as bytes.
void elastic_props(const std::string& behav, const double* mat, const int32_t* rve, int nph, long long N, ElasticProps& P,
                   uint8_t* ids)
{
    if (!mat || !rve || nph < 1) throw ArgError{M3D_ERR_ARG, "material: null pointer or nph < 1"};
    if (nph > 256) throw ArgError{M3D_ERR_UNSUPPORTED, "material: more than 256 phases"};
    P.nph = nph;
    P.la.assign(nph, 0.0);
    P.mu.assign(nph, 0.0);
    P.K.assign(nph, 0.0);
    P.Cm.assign((size_t)36 * nph, 0.0);
    P.Sm.assign((size_t)36 * nph, 0.0);
    P.mtab.assign((size_t)8 * nph, 0.0);
    for (int i = 0; i < nph; ++i) {
        const double El = mat[i + 0 * nph], nu = mat[i + 1 * nph];
        const double mu = El / (1 + nu) / 2.;
        const double la = nu * El / (1 + nu) / (1 - 2 * nu);
        if (!std::isfinite(mu) || !std::isfinite(la) || mu <= 0)
            throw ArgError{M3D_ERR_NUMERIC, "material: non-finite or non-positive moduli in phase " + std::to_string(i)};
        P.mu[i] = mu;
        P.la[i] = la;
        P.K[i] = la + (2. / 3.) * mu;
        const double row[8] = {la, mu, mat[i + 2 * nph], mat[i + 3 * nph], mat[i + 4 * nph], mat[i + 5 * nph],
                               mat[i + 6 * nph], 0.0};
        std::memcpy(&P.mtab[(size_t)8 * i], row, sizeof row);
        iso6(la, mu, &P.Cm[36 * i]);
        if (!inv_n(&P.Cm[36 * i], &P.Sm[36 * i], 6)) throw ArgError{M3D_ERR_NUMERIC, "material: singular C"};
    }
    std::vector<long long> cnt(nph, 0);
    for (long long i = 0; i < N; ++i) {
        const int p = rve[i];
        if (p < 0 || p >= nph) throw ArgError{M3D_ERR_ARG, "material: phase id out of range at voxel " + std::to_string(i)};
        ids[i] = (uint8_t)p;
        cnt[p]++;
    }
    P.vol.assign(nph, 0.0);
    for (int k = 0; k < nph; ++k) P.vol[k] = (double)cnt[k] / (double)N;
    double mu_max, mu_min, la_max, la_min, K_max, K_min, vmax, dummy;
    int imu, idum;
    max_min(P.mu, mu_max, imu, mu_min);
    max_min(P.la, la_max, idum, la_min);
    max_min(P.K, K_max, idum, K_min);
    max_min(P.vol, vmax, idum, dummy);
    if (behav == "e") {
        const double K0 = std::sqrt(K_min * K_max);
        P.mu0 = std::sqrt(mu_min * mu_max);
        P.la0 = K0 - P.mu0 * 2. / 3.;
        P.strain_base = (P.vol[imu] == vmax) ? 1 : 0;
    } else {
        P.strain_base = 1;
        P.mu0 = (mu_min + mu_max) / 2;
        P.la0 = (la_min + la_max) / 2;
    }
}

// Per-phase coefficients of the isotropic maps used by the kernels
// (m3d_kernels.hpp, Elastic policy): stiffness or compliance, polarisation
// and the Eyre-Milton acceleration C_acc = 2(C+C0)^-1 C0 (:539-546).
// basic: the basic Moulinec-Suquet iteration with the load held fixed,
// u <- u_c (acceleration factors 1)
void elastic_tables(const ElasticProps& P, double* coef, bool basic = false)
{
    const double K0 = P.la0 + 2.0 * P.mu0 / 3.0, mu0 = P.mu0;
    std::fill(coef, coef + (size_t)256 * NCOEF, 0.0);
    for (int p = 0; p < P.nph; ++p) {
        double* c = coef + (size_t)p * NCOEF;
        const double K = P.K[p], mu = P.mu[p];
        if (P.strain_base) {
            c[0] = 3 * K;               c[1] = 2 * mu;
            c[2] = 3 * (K - K0);        c[3] = 2 * (mu - mu0);
            c[4] = 2 * K0 / (K + K0);   c[5] = 2 * mu0 / (mu + mu0);
        } else {
            c[0] = 1.0 / (3 * K);       c[1] = 1.0 / (2 * mu);
            c[2] = 1.0 / (3 * K) - 1.0 / (3 * K0);
            c[3] = 1.0 / (2 * mu) - 1.0 / (2 * mu0);
            c[4] = 2 * K / (K + K0);    c[5] = 2 * mu / (mu + mu0);
        }
        if (basic) c[4] = c[5] = 1.0;
    }
}

// N * (0.5 M[0](0,0) + 0.5 M[1](0,0)) * E.sum(), FFT_analysis3D.cpp:399,408
double elastic_norm(const ElasticProps& P, long long N, const double* E)
{
    const int p1 = P.nph > 1 ? 1 : 0;
    const std::vector<double>& M = P.strain_base ? P.Cm : P.Sm;
    double es = 0;
    for (int i = 0; i < 6; ++i) es += E[i];
    return (double)N * (0.5 * M[0] + 0.5 * M[36 * p1]) * es;
}

// The six unit loads of homogenization(), :872-875
void elastic_unit_loads(const ElasticProps& P, long long N, std::vector<std::array<double, 6>>& loads,
                        std::vector<double>& norms)
{
    for (int w = 0; w < 6; ++w) {
        std::array<double, 6> ld;
        ld.fill(0.0);
        ld[w] = P.strain_base ? 0.0005 : 1.e6;
        loads.push_back(ld);
        norms.push_back(elastic_norm(P, N, ld.data()));
    }
}

// Effective tensors from the six solves' means, :880-902 and :998.
// m: 6 x 12 means per load case, [<u> (6) | <v> (6)]; u is the field the
// scheme iterates on (strain in the strain base, stress in the stress base).
void elastic_assemble(int strain_base, const double* m, double* Cout, double* Sout, double* epOut, double* sgOut,
                      double* ret)
{
    double ep[36], sg[36];
    for (int w = 0; w < 6; ++w)
        for (int t = 0; t < 6; ++t) {
            const double mu_ = m[(size_t)w * 12 + t], mv = m[(size_t)w * 12 + 6 + t];
            const double eps_m = strain_base ? mu_ : mv;
            const double sig_m = strain_base ? mv : mu_;
            sg[w + 6 * t] = sig_m; // sigHomo(w,t)
            ep[t + 6 * w] = eps_m; // epHomo(t,w)
        }
    double Ch[36], Sh[36], inv[36];
    if (strain_base) {
        if (!inv_n(ep, inv, 6)) throw ArgError{M3D_ERR_NUMERIC, "homogenization: singular epHomo"};
        mul_n(inv, sg, Ch, 6);
        if (!inv_n(Ch, Sh, 6)) throw ArgError{M3D_ERR_NUMERIC, "homogenization: singular CHomo"};
    } else {
        if (!inv_n(sg, inv, 6)) throw ArgError{M3D_ERR_NUMERIC, "homogenization: singular sigHomo"};
        mul_n(inv, ep, Sh, 6);
        if (!inv_n(Sh, Ch, 6)) throw ArgError{M3D_ERR_NUMERIC, "homogenization: singular SHomo"};
    }
    if (Cout) std::memcpy(Cout, Ch, sizeof Ch);
    if (Sout) std::memcpy(Sout, Sh, sizeof Sh);
    if (epOut) std::memcpy(epOut, ep, sizeof ep);
    if (sgOut) std::memcpy(sgOut, sg, sizeof sg);
    if (ret) *ret = 1. / Sh[0] + 1. / Sh[7] + 1. / Sh[14];
}

} // namespace
