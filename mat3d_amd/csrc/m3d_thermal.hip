// m3d_thermal.hip -- C-ABI entry points of the thermal solver (include/m3d.h).
#include "m3d_engine.hpp"

// ====================================================================== ABI
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

void fill_info(m3d_run_info* dst, const m3d_run_info& src)
{
    if (dst) *dst = src;
}

} // namespace

extern "C" {
// --------------------------------------------------------------- thermal
int m3d_thermal_create(int nx, int ny, int nz, int precision, m3d_thermal** out)
{
    if (!out) return M3D_ERR_ARG;
    *out = nullptr;
    if (nx < 2 || ny < 2 || nz < 2 || (long long)nx * ny * nz > (1LL << 31)) return M3D_ERR_ARG;
    if (precision != M3D_FP64 && precision != M3D_FP32_POLISH) return M3D_ERR_ARG;
    m3d_thermal* h = new (std::nothrow) m3d_thermal;
    if (!h) return M3D_ERR_OOM;
    int rc = guarded(h, [&]() {
        h->eng = new Engine<Thermal>(nx, ny, nz, precision);
        h->eng->cap = 500; // FFT_thermal3D.cpp:476
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

int m3d_thermal_destroy(m3d_thermal* h)
{
    if (!h) return M3D_ERR_ARG;
    delete h->eng;
    delete h;
    return M3D_OK;
}

const char* m3d_thermal_last_error(m3d_thermal* h) { return h ? h->err.c_str() : "null handle"; }

int m3d_thermal_set_option(m3d_thermal* h, const char* key, double value)
{
    return guarded(h, [&]() {
        std::string k = key ? key : "";
        if (k == "polish_factor" && value >= 1.0) h->eng->polish_factor = value;
        else if (k == "fp32_floor" && value >= 0.0) h->eng->fp32_floor = value;
        else if (k == "rollback" && value >= 0 && value <= 64) h->eng->rollback = (int)value;
        else if (k == "cap" && value >= 1) h->eng->cap = (int)value;
        else if (k == "chunk" && value >= 0) h->eng->chunk = (int)value;
        else if (k == "refine" && (value == 0 || value == 1)) h->eng->refine = (int)value;
        else if (k == "autotune" && (value == 0 || value == 1)) h->eng->autotune = (int)value;
        else throw ArgError{M3D_ERR_ARG, "unknown option or bad value: " + k};
        return M3D_OK;
    });
}

namespace {
int thermal_material_impl(m3d_thermal* h, const std::vector<double>& k, const int32_t* rve, bool two_phase)
{
    Engine<Thermal>& e = *h->eng;
    const long long N = e.N;
    const int nph = (int)k.size();
    for (double v : k)
        if (!(v > 0) || !std::isfinite(v)) throw ArgError{M3D_ERR_NUMERIC, "material: conductivities must be > 0"};
    std::vector<uint8_t> ids((size_t)N);
    std::vector<long long> cnt(nph, 0);
    long long rsum = 0;
    for (long long i = 0; i < N; ++i) {
        int p = rve[i];
        if (two_phase) {
            rsum += p;
            p = (p != 0) ? 1 : 0; // RVE != 0 -> matrix (FFT_thermal3D.cpp:186-189)
        } else if (p < 0 || p >= nph) {
            throw ArgError{M3D_ERR_ARG, "material: phase id out of range at voxel " + std::to_string(i)};
        }
        ids[i] = (uint8_t)p;
        cnt[p]++;
    }
    double kmax, kmin;
    int imax;
    max_min(k, kmax, imax, kmin);
    h->k0 = std::sqrt(kmin * kmax);
    if (two_phase) {
        // vol_1 = RVE.sum()/xyz is an integer division in the reference (:399)
        const double vol_1 = (double)(rsum / N);
        const double ki = k[0], km = k[1];
        h->k0 = std::sqrt(ki * km);
        h->grad_base = ((vol_1 >= 0.5 && km > ki) || (vol_1 < 0.5 && km < ki)) ? 1 : 0;
    } else {
        std::vector<double> vol(nph);
        for (int p = 0; p < nph; ++p) vol[p] = (double)cnt[p] / (double)N;
        double vmax, d;
        int di;
        max_min(vol, vmax, di, d);
        h->grad_base = (vol[imax] == vmax) ? 1 : 0;
    }
    std::vector<double> coef((size_t)256 * NCOEF, 0.0);
    for (int p = 0; p < nph; ++p) {
        double* c = coef.data() + (size_t)p * NCOEF;
        if (h->grad_base) {
            c[0] = k[p];
            c[1] = k[p] - h->k0;
            c[2] = 2 * h->k0 / (k[p] + h->k0);
        } else {
            const double r = 1.0 / k[p], r0 = 1.0 / h->k0;
            c[0] = r;
            c[1] = r - r0;
            c[2] = 2 * r0 / (r + r0);
        }
    }
    std::vector<double> prm = {h->k0, 0.0};
    e.upload_material(ids, coef, prm);
    h->k = k;
    h->nph = nph;
    h->have_mat = true;
    return M3D_OK;
}

double thermal_norm(const m3d_thermal* h, const double* E)
{
    // N (0.5 Ki + 0.5 Km) E.sum() or the resistivity form, FFT_thermal3D.cpp:346,356
    const double k0 = h->k[0], k1 = h->k[h->nph > 1 ? 1 : 0];
    const double es = E[0] + E[1] + E[2];
    const double m = h->grad_base ? (0.5 * k0 + 0.5 * k1) : (0.5 / k0 + 0.5 / k1);
    return (double)h->eng->N * m * es;
}
} // namespace

int m3d_thermal_material(m3d_thermal* h, double km, double ki, const int32_t* rve)
{
    return guarded(h, [&]() {
        if (!rve) throw ArgError{M3D_ERR_ARG, "material: null RVE"};
        return thermal_material_impl(h, {ki, km}, rve, true);
    });
}

int m3d_thermal_material_n(m3d_thermal* h, const double* k, int nph, const int32_t* rve)
{
    return guarded(h, [&]() {
        if (!rve || !k || nph < 1) throw ArgError{M3D_ERR_ARG, "material_n: bad arguments"};
        if (nph > 256) throw ArgError{M3D_ERR_UNSUPPORTED, "material_n: more than 256 phases"};
        return thermal_material_impl(h, std::vector<double>(k, k + nph), rve, false);
    });
}

int m3d_thermal_set_load(m3d_thermal* h, const double* E)
{
    return guarded(h, [&]() {
        if (!E) throw ArgError{M3D_ERR_ARG, "set_load: null E"};
        for (int i = 0; i < 3; ++i) h->E[i] = E[i];
        h->have_load = true;
        return M3D_OK;
    });
}

int m3d_thermal_run(m3d_thermal* h, double tol_field, double tol_eq, m3d_run_info* info)
{
    return guarded(h, [&]() {
        if (!h->have_mat || !h->have_load) throw ArgError{M3D_ERR_STATE, "run: material() and setLoad() first"};
        std::array<double, 6> ld{};
        for (int i = 0; i < 3; ++i) ld[i] = h->E[i];
        h->eng->solve({ld}, tol_field, tol_eq, {thermal_norm(h, h->E)}, h->grad_base);
        fill_info(info, h->eng->info[0]);
        h->ran = true;
        return M3D_OK;
    });
}

int m3d_thermal_homogenization(m3d_thermal* h, double tol_field, double tol_eq, double* Kout, double* Rout,
                               double* Gout, double* Qout, m3d_run_info* info3)
{
    return guarded(h, [&]() {
        if (!h->have_mat) throw ArgError{M3D_ERR_STATE, "homogenization: material() first"};
        Engine<Thermal>& e = *h->eng;
        std::vector<std::array<double, 6>> loads(3);
        std::vector<double> norms(3);
        for (int w = 0; w < 3; ++w) {
            loads[w].fill(0.0);
            loads[w][w] = 1; // :506,548
            norms[w] = thermal_norm(h, loads[w].data());
        }
        e.solve(loads, tol_field, tol_eq, norms, h->grad_base);
        const std::vector<double> m = e.means();
        double G[9], Q[9], Kh[9], Rh[9], inv[9];
        for (int w = 0; w < 3; ++w)
            for (int c = 0; c < 3; ++c) {
                const double mu_ = m[(size_t)w * 6 + c], mv = m[(size_t)w * 6 + 3 + c];
                const double grad = h->grad_base ? mu_ : mv, flux = h->grad_base ? mv : mu_;
                if (h->grad_base) {
                    G[w + 3 * c] = grad; // tempGradHomo(w,c)
                    Q[c + 3 * w] = flux; // heatFluxHomo(c,w)
                } else {
                    Q[w + 3 * c] = flux; // heatFluxHomo(w,c)
                    G[c + 3 * w] = grad; // tempGradHomo(c,w)
                }
            }
        if (h->grad_base) {
            if (!inv_n(G, inv, 3)) throw ArgError{M3D_ERR_NUMERIC, "homogenization: singular tempGradHomo"};
            mul_n(inv, Q, Kh, 3);
            if (!inv_n(Kh, Rh, 3)) throw ArgError{M3D_ERR_NUMERIC, "homogenization: singular KHomo"};
        } else {
            if (!inv_n(Q, inv, 3)) throw ArgError{M3D_ERR_NUMERIC, "homogenization: singular heatFluxHomo"};
            mul_n(inv, G, Rh, 3);
            if (!inv_n(Rh, Kh, 3)) throw ArgError{M3D_ERR_NUMERIC, "homogenization: singular RHomo"};
        }
        if (Kout) std::memcpy(Kout, Kh, sizeof Kh);
        if (Rout) std::memcpy(Rout, Rh, sizeof Rh);
        if (Gout) std::memcpy(Gout, G, sizeof G);
        if (Qout) std::memcpy(Qout, Q, sizeof Q);
        if (info3)
            for (int w = 0; w < 3; ++w) info3[w] = e.info[w];
        h->ran = true;
        return M3D_OK;
    });
}

int m3d_thermal_get_fields(m3d_thermal* h, int lc, double* grad, double* flux)
{
    return guarded(h, [&]() {
        if (!h->ran) throw ArgError{M3D_ERR_STATE, "get_fields: nothing has run"};
        if (h->grad_base) h->eng->fields(lc, grad, flux);
        else h->eng->fields(lc, flux, grad);
        return M3D_OK;
    });
}

} // extern "C"
