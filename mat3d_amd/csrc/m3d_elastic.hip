// m3d_elastic.hip -- C-ABI entry points of the elastic solver (include/m3d.h).
#include "m3d_engine.hpp"

// ====================================================================== ABI
struct m3d_elastic {
    Engine<Elastic>* eng = nullptr;
    std::string err;
    std::string behav;
    int nph = 0;
    bool have_mat = false, have_load = false, ran = false;
    int strain_base = 1;
    double la0 = 0, mu0 = 0;
    std::vector<double> vol, Cm, Sm; // per phase 6x6
    double E[6] = {0, 0, 0, 0, 0, 0}, X0[6] = {0, 0, 0, 0, 0, 0};
};

struct m3d_thermal {
    Engine<Thermal>* eng = nullptr;
    std::string err;
    int nph = 0;
    bool have_mat = false, have_load = false, ran = false;
    int grad_base = 1;
    double k0 = 0;
    std::vector<double> k;
    double E[3] = {0, 0, 0};
};

namespace {

template <typename Hd, typename F>
int guarded(Hd* h, F&& f)
{
    if (!h) return M3D_ERR_ARG;
    try {
        return f();
    } catch (const HipError& e) {
        h->err = std::string("HIP error in ") + e.what + ": " + hipGetErrorString(e.e);
        (void)hipGetLastError();
        return M3D_ERR_HIP;
    } catch (const ArgError& e) {
        h->err = e.msg;
        return e.code;
    } catch (const std::bad_alloc&) {
        h->err = "host allocation failed";
        return M3D_ERR_OOM;
    } catch (...) {
        h->err = "unexpected error";
        return M3D_ERR_HIP;
    }
}

// per-phase coefficients of the elastic maps (see m3d_kernels.hpp)
void elastic_tables(m3d_elastic* h, const std::vector<double>& la, const std::vector<double>& mu,
                    const std::vector<double>& K, std::vector<double>& coef)
{
    const int n = (int)la.size();
    const double K0 = h->la0 + 2.0 * h->mu0 / 3.0, mu0 = h->mu0;
    coef.assign((size_t)256 * NCOEF, 0.0);
    for (int p = 0; p < n; ++p) {
        double* c = coef.data() + (size_t)p * NCOEF;
        if (h->strain_base) {
            c[0] = 3 * K[p];            c[1] = 2 * mu[p];
            c[2] = 3 * (K[p] - K0);     c[3] = 2 * (mu[p] - mu0);
            c[4] = 2 * K0 / (K[p] + K0); c[5] = 2 * mu0 / (mu[p] + mu0);
        } else {
            c[0] = 1.0 / (3 * K[p]);    c[1] = 1.0 / (2 * mu[p]);
            c[2] = 1.0 / (3 * K[p]) - 1.0 / (3 * K0);
            c[3] = 1.0 / (2 * mu[p]) - 1.0 / (2 * mu0);
            c[4] = 2 * K[p] / (K[p] + K0); c[5] = 2 * mu[p] / (mu[p] + mu0);
        }
    }
}

double elastic_norm(const m3d_elastic* h, const double* E)
{
    // N * (0.5 M[0](0,0) + 0.5 M[1](0,0)) * E.sum(), FFT_analysis3D.cpp:399,408
    const int p1 = h->nph > 1 ? 1 : 0;
    const std::vector<double>& M = h->strain_base ? h->Cm : h->Sm;
    double es = 0;
    for (int i = 0; i < 6; ++i) es += E[i];
    return (double)h->eng->N * (0.5 * M[0] + 0.5 * M[36 * p1]) * es;
}

void fill_info(m3d_run_info* dst, const m3d_run_info& src)
{
    if (dst) *dst = src;
}

} // namespace

extern "C" {

const char* m3d_version(void) { return "mat3d_amd 0.1.0 (gfx950)"; }

int m3d_elastic_create(int nx, int ny, int nz, const char* behav, int precision, m3d_elastic** out)
{
    if (!out) return M3D_ERR_ARG;
    *out = nullptr;
    if (nx < 2 || ny < 2 || nz < 2 || (long long)nx * ny * nz > (1LL << 31)) return M3D_ERR_ARG;
    if (precision != M3D_FP64 && precision != M3D_FP32_POLISH) return M3D_ERR_ARG;
    std::string b = behav ? behav : "";
    if (b != "e" && b != "ep" && b != "ed") return M3D_ERR_ARG;
    m3d_elastic* h = new (std::nothrow) m3d_elastic;
    if (!h) return M3D_ERR_OOM;
    h->behav = b;
    int rc = guarded(h, [&]() {
        h->eng = new Engine<Elastic>(nx, ny, nz, precision);
        return M3D_OK;
    });
    if (rc != M3D_OK) {
        delete h->eng;
        delete h;
        return rc;
    }
    *out = h;
    return M3D_OK;
}

int m3d_elastic_destroy(m3d_elastic* h)
{
    if (!h) return M3D_ERR_ARG;
    delete h->eng;
    delete h;
    return M3D_OK;
}

const char* m3d_elastic_last_error(m3d_elastic* h) { return h ? h->err.c_str() : "null handle"; }

int m3d_elastic_set_option(m3d_elastic* h, const char* key, double value)
{
    return guarded(h, [&]() {
        std::string k = key ? key : "";
        if (k == "polish_factor" && value >= 1.0) h->eng->polish_factor = value;
        else if (k == "fp32_floor" && value >= 0.0) h->eng->fp32_floor = value;
        else if (k == "rollback" && value >= 0 && value <= 64) h->eng->rollback = (int)value;
        else if (k == "cap" && value >= 1) h->eng->cap = (int)value;
        else if (k == "chunk" && value >= 0) h->eng->chunk = (int)value;
        else throw ArgError{M3D_ERR_ARG, "unknown option or bad value: " + k};
        return M3D_OK;
    });
}

// material, FFT_analysis3D.cpp:420-549
int m3d_elastic_material(m3d_elastic* h, const double* mat, const int32_t* rve, int nph)
{
    return guarded(h, [&]() {
        if (!mat || !rve || nph < 1) throw ArgError{M3D_ERR_ARG, "material: null pointer or nph < 1"};
        if (nph > 256) throw ArgError{M3D_ERR_UNSUPPORTED, "material: more than 256 phases"};
        Engine<Elastic>& e = *h->eng;
        const long long N = e.N;
        std::vector<double> El(nph), nu(nph), mu(nph), la(nph), K(nph);
        h->Cm.assign((size_t)36 * nph, 0.0);
        h->Sm.assign((size_t)36 * nph, 0.0);
        for (int i = 0; i < nph; ++i) {
            El[i] = mat[i + 0 * nph];
            nu[i] = mat[i + 1 * nph];
            mu[i] = El[i] / (1 + nu[i]) / 2.;
            la[i] = nu[i] * El[i] / (1 + nu[i]) / (1 - 2 * nu[i]);
            K[i] = la[i] + (2. / 3.) * mu[i];
            if (!std::isfinite(mu[i]) || !std::isfinite(la[i]) || mu[i] <= 0)
                throw ArgError{M3D_ERR_NUMERIC, "material: non-finite or non-positive moduli"};
            iso6(la[i], mu[i], &h->Cm[36 * i]);
            if (!inv_n(&h->Cm[36 * i], &h->Sm[36 * i], 6)) throw ArgError{M3D_ERR_NUMERIC, "material: singular C"};
        }
        std::vector<uint8_t> ids((size_t)N);
        std::vector<long long> cnt(nph, 0);
        for (long long i = 0; i < N; ++i) {
            const int p = rve[i];
            if (p < 0 || p >= nph) throw ArgError{M3D_ERR_ARG, "material: phase id out of range at voxel " + std::to_string(i)};
            ids[i] = (uint8_t)p;
            cnt[p]++;
        }
        h->vol.assign(nph, 0.0);
        for (int k = 0; k < nph; ++k) h->vol[k] = (double)cnt[k] / (double)N;
        double mu_max, mu_min, la_max, la_min, K_max, K_min, vmax, dummy;
        int imu, idum;
        max_min(mu, mu_max, imu, mu_min);
        max_min(la, la_max, idum, la_min);
        max_min(K, K_max, idum, K_min);
        max_min(h->vol, vmax, idum, dummy);
        if (h->behav == "e") {
            const double K0 = std::sqrt(K_min * K_max);
            h->mu0 = std::sqrt(mu_min * mu_max);
            h->la0 = K0 - h->mu0 * 2. / 3.;
            h->strain_base = (h->vol[imu] == vmax) ? 1 : 0;
        } else {
            h->strain_base = 1;
            h->mu0 = (mu_min + mu_max) / 2;
            h->la0 = (la_min + la_max) / 2;
        }
        std::vector<double> coef;
        elastic_tables(h, la, mu, K, coef);
        std::vector<double> prm = {h->la0, h->mu0};
        e.upload_material(ids, coef, prm);
        e.g32.strain_base = e.g64.strain_base = h->strain_base;
        h->nph = nph;
        h->have_mat = true;
        return M3D_OK;
    });
}

int m3d_elastic_info(m3d_elastic* h, double* vol, int* strain_base, double* la0, double* mu0)
{
    return guarded(h, [&]() {
        if (!h->have_mat) throw ArgError{M3D_ERR_STATE, "info: material() not called"};
        if (vol) std::memcpy(vol, h->vol.data(), sizeof(double) * h->nph);
        if (strain_base) *strain_base = h->strain_base;
        if (la0) *la0 = h->la0;
        if (mu0) *mu0 = h->mu0;
        return M3D_OK;
    });
}

// setLoad, FFT_analysis3D.cpp:551-564
int m3d_elastic_set_load(m3d_elastic* h, const double* E, const double* X0)
{
    return guarded(h, [&]() {
        if (!E) throw ArgError{M3D_ERR_ARG, "set_load: null E"};
        for (int i = 0; i < 6; ++i) {
            h->E[i] = E[i];
            h->X0[i] = X0 ? X0[i] : E[i];
        }
        h->have_load = true;
        return M3D_OK;
    });
}

// runAnalysis "e" branch, FFT_analysis3D.cpp:751-794
int m3d_elastic_run(m3d_elastic* h, double tol_field, double tol_eq, int nInc, m3d_run_info* info)
{
    return guarded(h, [&]() {
        if (!h->have_mat || !h->have_load) throw ArgError{M3D_ERR_STATE, "run: material() and setLoad() first"};
        if (h->behav != "e") throw ArgError{M3D_ERR_UNSUPPORTED, "run: behaviour '" + h->behav + "' not provided yet"};
        (void)nInc; // the "e" branch ignores nInc (FFT_analysis3D.cpp:751)
        std::array<double, 6> ld;
        for (int i = 0; i < 6; ++i) ld[i] = h->E[i];
        h->eng->solve({ld}, tol_field, tol_eq, {elastic_norm(h, h->E)});
        fill_info(info, h->eng->info[0]);
        h->ran = true;
        return M3D_OK;
    });
}

// homogenization, FFT_analysis3D.cpp:856-999 (six load cases batched)
int m3d_elastic_homogenization(m3d_elastic* h, double tol_field, double tol_eq, double* Cout, double* Sout,
                               double* epOut, double* sgOut, double* ret, m3d_run_info* info6)
{
    return guarded(h, [&]() {
        if (!h->have_mat) throw ArgError{M3D_ERR_STATE, "homogenization: material() first"};
        Engine<Elastic>& e = *h->eng;
        std::vector<std::array<double, 6>> loads(6);
        std::vector<double> norms(6);
        for (int w = 0; w < 6; ++w) {
            loads[w].fill(0.0);
            loads[w][w] = h->strain_base ? 0.0005 : 1.e6; // :872-875
            norms[w] = elastic_norm(h, loads[w].data());
        }
        e.solve(loads, tol_field, tol_eq, norms);
        const std::vector<double> m = e.means();
        double ep[36], sg[36];
        for (int w = 0; w < 6; ++w)
            for (int t = 0; t < 6; ++t) {
                const double mu_ = m[(size_t)w * 12 + t], mv = m[(size_t)w * 12 + 6 + t];
                const double eps_m = h->strain_base ? mu_ : mv;
                const double sig_m = h->strain_base ? mv : mu_;
                sg[w + 6 * t] = sig_m; // sigHomo(w,t)
                ep[t + 6 * w] = eps_m; // epHomo(t,w)
            }
        double Ch[36], Sh[36], inv[36];
        if (h->strain_base) {
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
        if (info6)
            for (int w = 0; w < 6; ++w) info6[w] = e.info[w];
        h->ran = true;
        return M3D_OK;
    });
}

int m3d_elastic_get_fields(m3d_elastic* h, int lc, double* eps, double* sig)
{
    return guarded(h, [&]() {
        if (!h->ran) throw ArgError{M3D_ERR_STATE, "get_fields: nothing has run"};
        // u is epsilon in the strain base and sigma in the stress base
        if (h->strain_base) h->eng->fields(lc, eps, sig);
        else h->eng->fields(lc, sig, eps);
        return M3D_OK;
    });
}

int m3d_elastic_profile(m3d_elastic* h, int precision, int reps, double* ms6, int* geom9)
{
    return guarded(h, [&]() {
        if (!ms6 || reps < 1) throw ArgError{M3D_ERR_ARG, "profile: bad arguments"};
        h->eng->profile(precision, reps, ms6);
        if (geom9) h->eng->geometry(precision, geom9);
        h->ran = false; // the fields were overwritten
        return M3D_OK;
    });
}

} // extern "C"
