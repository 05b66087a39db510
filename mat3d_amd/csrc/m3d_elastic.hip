// m3d_elastic.hip -- C-ABI entry points of the elastic solver (include/m3d.h).
#include "m3d_engine.hpp"
#include "m3d_basic_host.hpp"
#include "m3d_elastic_props.hpp"

#include <chrono>
#include <exception>
#include <thread>

// ====================================================================== ABI
struct m3d_elastic {
    Engine<Elastic>* eng = nullptr;
    BasicRunner* basic = nullptr; // runAnalysis "ep" / "ed"
    bool last_basic = false;
    std::string err;
    std::string behav;
    bool have_mat = false, have_load = false, ran = false;
    int scheme = 0; // 0 accelerated (reference "e" loop), 1 basic with the load held fixed
    int basic_fp32 = 0; // "ep" on an FP32+polish handle: FP32 phase per increment (opt-in)
    ElasticProps P;
    double E[6] = {0, 0, 0, 0, 0, 0}, X0[6] = {0, 0, 0, 0, 0, 0};
};

// one RVE of a batch and its results
struct BatchProblem {
    int nx, ny, nz;
    ElasticProps P;
    std::vector<uint8_t> ids;
    double C[36], S[36], ep[36], sg[36], ret = 0;
    m3d_run_info info[6];
};

// problems of one grid shape: one engine, kept across calls; `loaded`: the
// group fits one device batch and its materials are on the device
struct BatchGroup {
    std::array<int, 3> shape;
    std::vector<int> members;
    Engine<Elastic>* eng = nullptr;
    long long per = 0; // RVEs per device batch
    bool loaded = false;
};

struct m3d_batch {
    int precision = M3D_FP32_POLISH;
    std::string err;
    std::vector<BatchProblem> probs;
    std::vector<BatchGroup> groups; // rebuilt after m3d_batch_add
    bool grouped = false;
    bool ran = false;
    double polish_factor = 10.0, fp32_floor = 1e-6;
    int rollback = 5, cap = 2000, chunk = 0, max_solves = 0, refine = 1, scheme = 0, autotune = 1;
    double ms = 0, voxel_its = 0;
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
    } catch (const std::exception& e) {
        h->err = e.what();
        return M3D_ERR_UNSUPPORTED;
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
    delete h->basic;
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
        else if (k == "refine" && (value == 0 || value == 1)) h->eng->refine = (int)value;
        else if (k == "autotune" && (value == 0 || value == 1)) h->eng->autotune = (int)value;
        else if (k == "basic_fp32" && (value == 0 || value == 1)) {
            h->basic_fp32 = (int)value;
            delete h->basic; // buffers differ; results of an earlier run go with it
            h->basic = nullptr;
            if (h->last_basic) h->ran = false;
            h->last_basic = false;
        } else if (k == "scheme" && (value == 0 || value == 1)) {
            h->scheme = (int)value;
            if (h->have_mat) {
                std::vector<double> coef((size_t)256 * NCOEF);
                elastic_tables(h->P, coef.data(), h->scheme == 1);
                h->eng->upload_coef(coef, {h->P.la0, h->P.mu0});
            }
        } else throw ArgError{M3D_ERR_ARG, "unknown option or bad value: " + k};
        return M3D_OK;
    });
}

// material, FFT_analysis3D.cpp:420-549
int m3d_elastic_material(m3d_elastic* h, const double* mat, const int32_t* rve, int nph)
{
    return guarded(h, [&]() {
        Engine<Elastic>& e = *h->eng;
        std::vector<uint8_t> ids((size_t)e.N);
        ElasticProps P;
        elastic_props(h->behav, mat, rve, nph, e.N, P, ids.data());
        std::vector<double> coef((size_t)256 * NCOEF);
        elastic_tables(P, coef.data(), h->scheme == 1);
        e.upload_material(ids, coef, {P.la0, P.mu0});
        h->P = std::move(P);
        h->have_mat = true;
        return M3D_OK;
    });
}

// public member prop_homogenized (nInc x 9), FFT_analysis3D.hpp:156, filled by
// export_homogen_var (:722-746)
int m3d_elastic_get_prop(m3d_elastic* h, double* prop, int* rows)
{
    return guarded(h, [&]() {
        if (!h->basic || !h->last_basic) throw ArgError{M3D_ERR_STATE, "get_prop: no \"ep\"/\"ed\" run"};
        if (rows) *rows = h->basic->rows;
        if (prop) {
            const int R = h->basic->rows, Rall = (int)(h->basic->prop.size() / 9);
            for (int j = 0; j < 9; ++j)
                for (int i = 0; i < R; ++i) prop[i + (size_t)j * R] = h->basic->prop[i + (size_t)j * Rall];
        }
        return M3D_OK;
    });
}

int m3d_elastic_info(m3d_elastic* h, double* vol, int* strain_base, double* la0, double* mu0)
{
    return guarded(h, [&]() {
        if (!h->have_mat) throw ArgError{M3D_ERR_STATE, "info: material() not called"};
        if (vol) std::memcpy(vol, h->P.vol.data(), sizeof(double) * h->P.nph);
        if (strain_base) *strain_base = h->P.strain_base;
        if (la0) *la0 = h->P.la0;
        if (mu0) *mu0 = h->P.mu0;
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
        if (h->behav != "e") {
            // basic scheme with mixed control, J2 ("ep") or damage ("ed"), :796-853
            if (nInc < 1) throw ArgError{M3D_ERR_ARG, "run: nInc must be >= 1"};
            if (!h->basic)
                h->basic = new BasicRunner(*h->eng, h->behav == "ep" ? BEH_EP : BEH_ED,
                                           h->eng->precision == M3D_FP32_POLISH && h->basic_fp32 && h->behav == "ep");
            const int p1 = h->P.nph > 1 ? 1 : 0;
            const double nc = (double)h->eng->N * (0.5 * h->P.Cm[0] + 0.5 * h->P.Cm[36 * p1]);
            h->basic->run(h->P.mtab, h->P.la0, h->P.mu0, nc, h->E, h->X0, tol_field, tol_eq, nInc, h->eng->cap);
            fill_info(info, h->basic->info);
            h->last_basic = true;
            h->ran = true;
            return M3D_OK;
        }
        (void)nInc; // the "e" branch ignores nInc (FFT_analysis3D.cpp:751)
        std::array<double, 6> ld;
        for (int i = 0; i < 6; ++i) ld[i] = h->E[i];
        h->eng->solve({ld}, tol_field, tol_eq, {elastic_norm(h->P, h->eng->N, h->E)}, h->P.strain_base);
        fill_info(info, h->eng->info[0]);
        h->last_basic = false;
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
        std::vector<std::array<double, 6>> loads;
        std::vector<double> norms;
        elastic_unit_loads(h->P, e.N, loads, norms);
        e.solve(loads, tol_field, tol_eq, norms, h->P.strain_base);
        const std::vector<double> m = e.means();
        elastic_assemble(h->P.strain_base, m.data(), Cout, Sout, epOut, sgOut, ret);
        if (info6)
            for (int w = 0; w < 6; ++w) info6[w] = e.info[w];
        h->last_basic = false;
        h->ran = true;
        return M3D_OK;
    });
}

int m3d_elastic_get_fields(m3d_elastic* h, int lc, double* eps, double* sig)
{
    return guarded(h, [&]() {
        if (!h->ran) throw ArgError{M3D_ERR_STATE, "get_fields: nothing has run"};
        if (h->last_basic) {
            if (lc != 0) throw ArgError{M3D_ERR_ARG, "get_fields: runAnalysis holds one load case"};
            h->basic->fields(eps, sig);
            return M3D_OK;
        }
        // u is epsilon in the strain base and sigma in the stress base
        if (h->P.strain_base) h->eng->fields(lc, eps, sig);
        else h->eng->fields(lc, sig, eps);
        return M3D_OK;
    });
}

// public members epsilon_p_eff / sigma_eff ("ep") or d ("ed") of the last
// runAnalysis, FFT_analysis3D.hpp:56-98 (read by printer and vtiWriter,
// FFT_analysis3D.cpp:1001-1104)
int m3d_elastic_get_state(m3d_elastic* h, double* ep_eff_or_d, double* sig_eff)
{
    return guarded(h, [&]() {
        if (!h->basic || !h->last_basic) throw ArgError{M3D_ERR_STATE, "get_state: no \"ep\"/\"ed\" run"};
        h->basic->state(ep_eff_or_d, sig_eff);
        return M3D_OK;
    });
}

int m3d_elastic_get_tuning(m3d_elastic* h, int* xg_threads, int* xg_wide, int* y_threads, int* xg_radix, int* y_radix,
                           int* from_file)
{
    return guarded(h, [&]() {
        const XyTune& t = h->eng->tune32;
        if (xg_threads) *xg_threads = t.xg_nt;
        if (xg_wide) *xg_wide = t.xg_wide;
        if (y_threads) *y_threads = t.y_nt;
        if (xg_radix) *xg_radix = t.xg_rp;
        if (y_radix) *y_radix = t.y_rp;
        if (from_file) *from_file = t.from_file;
        return M3D_OK;
    });
}

int m3d_tuning_report(char* buf, int len, int* needed)
{
    std::string out;
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        for (const auto& kv : tune_cache()) {
            const TuneKey& k = kv.first;
            const XyTune& t = kv.second;
            char line[256];
            std::snprintf(line, sizeof line, "%s %dx%dx%d c%d solves %d: xgreen %d threads, %s kz tile, radix plan %d; "
                                             "ypass %d threads, radix plan %d (%s)\n",
                          k.dev.c_str(), k.nx, k.ny, k.nz, k.c, k.ns, t.xg_nt, t.xg_wide ? "wide" : "default", t.xg_rp,
                          t.y_nt, t.y_rp, t.from_file ? "tuning file" : "timed");
            out += line;
        }
    }
    if (needed) *needed = (int)out.size() + 1;
    if (buf && len > 0) {
        const size_t n = std::min((size_t)len - 1, out.size());
        std::memcpy(buf, out.data(), n);
        buf[n] = 0;
    }
    return M3D_OK;
}

const char* m3d_tuning_file(void)
{
    static std::string p;
    p = tune_file_path();
    return p.c_str();
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

// device copy rate for measurement tools (no reference counterpart)
int m3d_device_copy(double gib, int reps, double* ms_per_copy)
{
    if (!(gib > 0) || reps < 1 || !ms_per_copy) return M3D_ERR_ARG;
    const size_t bytes = (size_t)(gib * 1073741824.0) / 256 * 256;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = M3D_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) rc = M3D_ERR_OOM;
    if (rc == M3D_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = M3D_ERR_HIP;
    if (rc == M3D_OK) {
        float ms = 0;
        bool ok = hipMemset(a, 1, bytes) == hipSuccess && hipMemcpy(b, a, bytes, hipMemcpyDeviceToDevice) == hipSuccess &&
                  hipEventRecord(e0, nullptr) == hipSuccess;
        for (int r = 0; ok && r < reps; ++r) ok = hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, nullptr) == hipSuccess;
        ok = ok && hipEventRecord(e1, nullptr) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (ok) *ms_per_copy = (double)ms / reps;
        else rc = M3D_ERR_HIP;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    (void)hipGetLastError();
    return rc;
}

// ------------------------------------------------------------------ batch
int m3d_batch_create(int precision, m3d_batch** out)
{
    if (!out) return M3D_ERR_ARG;
    *out = nullptr;
    if (precision != M3D_FP64 && precision != M3D_FP32_POLISH) return M3D_ERR_ARG;
    m3d_batch* b = new (std::nothrow) m3d_batch;
    if (!b) return M3D_ERR_OOM;
    b->precision = precision;
    *out = b;
    return M3D_OK;
}

namespace {
void batch_free_groups(m3d_batch* b)
{
    for (BatchGroup& G : b->groups) delete G.eng;
    b->groups.clear();
    b->grouped = false;
}
} // namespace

int m3d_batch_destroy(m3d_batch* b)
{
    if (!b) return M3D_ERR_ARG;
    batch_free_groups(b);
    delete b;
    return M3D_OK;
}

const char* m3d_batch_last_error(m3d_batch* b) { return b ? b->err.c_str() : "null handle"; }

int m3d_batch_set_option(m3d_batch* b, const char* key, double value)
{
    return guarded(b, [&]() {
        std::string k = key ? key : "";
        if (k == "polish_factor" && value >= 1.0) b->polish_factor = value;
        else if (k == "fp32_floor" && value >= 0.0) b->fp32_floor = value;
        else if (k == "rollback" && value >= 0 && value <= 64) b->rollback = (int)value;
        else if (k == "cap" && value >= 1) b->cap = (int)value;
        else if (k == "chunk" && value >= 0) b->chunk = (int)value;
        else if (k == "max_solves" && value >= 0) {
            b->max_solves = (int)value;
            batch_free_groups(b); // device batch sizes change
        }
        else if (k == "refine" && (value == 0 || value == 1)) b->refine = (int)value;
        else if (k == "autotune" && (value == 0 || value == 1)) b->autotune = (int)value;
        else if (k == "scheme" && (value == 0 || value == 1)) {
            b->scheme = (int)value;
            for (BatchGroup& G : b->groups) G.loaded = false; // coefficient tables change
        }
        else throw ArgError{M3D_ERR_ARG, "unknown option or bad value: " + k};
        return M3D_OK;
    });
}

// material() of one more RVE, FFT_analysis3D.cpp:420-549
int m3d_batch_add(m3d_batch* b, int nx, int ny, int nz, const double* mat, const int32_t* rve, int nph, int* index)
{
    return guarded(b, [&]() {
        if (nx < 2 || ny < 2 || nz < 2 || (long long)nx * ny * nz > (1LL << 31))
            throw ArgError{M3D_ERR_ARG, "batch_add: grid sizes"};
        (void)make_geo<double>(nx, ny, nz, 6); // throws for unsupported sizes
        (void)make_geo<float>(nx, ny, nz, 6);
        BatchProblem q;
        q.nx = nx;
        q.ny = ny;
        q.nz = nz;
        const long long N = (long long)nx * ny * nz;
        q.ids.resize((size_t)N);
        elastic_props("e", mat, rve, nph, N, q.P, q.ids.data());
        b->probs.push_back(std::move(q));
        b->ran = false;
        batch_free_groups(b);
        if (index) *index = (int)b->probs.size() - 1;
        return M3D_OK;
    });
}

// homogenization() of every RVE, FFT_analysis3D.cpp:856-999: problems of one
// grid shape share an engine; their 6 load cases each run as one batch of
// independent solves.  When every shape group fits the device memory in one
// device batch, the groups' engines run at the same time, one host thread
// and one stream each (allocation and autotuning done first, one engine
// after the other), so small grids and the host's polling between graph
// replays of one group overlap the others' kernels; otherwise the groups run
// one after another, sub-batched to fit.
namespace {

struct GroupRun {
    std::vector<std::array<double, 6>> loads;
    std::vector<double> norms;
    std::vector<int> sbv, rvv;
    std::vector<double> means;
    std::exception_ptr err;
};

void batch_group_setup(m3d_batch* b, BatchGroup& G)
{
    Engine<Elastic>& e = *G.eng;
    e.polish_factor = b->polish_factor;
    e.fp32_floor = b->fp32_floor;
    e.rollback = b->rollback;
    e.cap = b->cap;
    e.chunk = b->chunk;
    e.refine = b->refine;
    e.autotune = b->autotune;
}

// loads, norms, base flags and RVE indices of members [k0, k1) of G; uploads
// their phase ids and coefficient tables unless resident
void batch_group_load(m3d_batch* b, BatchGroup& G, size_t k0, size_t k1, GroupRun& run)
{
    Engine<Elastic>& e = *G.eng;
    const std::vector<int>& members = G.members;
    const int R = (int)(k1 - k0);
    run.loads.clear();
    run.norms.clear();
    run.sbv.clear();
    run.rvv.clear();
    for (int r = 0; r < R; ++r) {
        const BatchProblem& q = b->probs[members[k0 + r]];
        elastic_unit_loads(q.P, e.N, run.loads, run.norms);
        for (int w = 0; w < 6; ++w) {
            run.sbv.push_back(q.P.strain_base);
            run.rvv.push_back(r);
        }
    }
    if (!G.loaded) {
        std::vector<uint8_t> ids((size_t)R * e.N);
        std::vector<double> coef((size_t)R * 256 * NCOEF), prm((size_t)2 * R);
        for (int r = 0; r < R; ++r) {
            const BatchProblem& q = b->probs[members[k0 + r]];
            std::memcpy(ids.data() + (size_t)r * e.N, q.ids.data(), (size_t)e.N);
            elastic_tables(q.P, coef.data() + (size_t)r * 256 * NCOEF, b->scheme == 1);
            prm[2 * r] = q.P.la0;
            prm[2 * r + 1] = q.P.mu0;
        }
        e.upload_material(ids, coef, prm);
        G.loaded = (R == (int)members.size()); // one device batch: keep it resident
    }
}

void batch_group_store(m3d_batch* b, BatchGroup& G, size_t k0, size_t k1, const std::vector<double>& m)
{
    Engine<Elastic>& e = *G.eng;
    for (int r = 0; r < (int)(k1 - k0); ++r) {
        BatchProblem& q = b->probs[G.members[k0 + r]];
        elastic_assemble(q.P.strain_base, m.data() + (size_t)r * 72, q.C, q.S, q.ep, q.sg, &q.ret);
        for (int w = 0; w < 6; ++w) {
            q.info[w] = e.info[(size_t)6 * r + w];
            b->voxel_its += (double)e.N * q.info[w].iterations;
        }
    }
}

} // namespace

int m3d_batch_homogenization(m3d_batch* b, double tol_field, double tol_eq)
{
    return guarded(b, [&]() {
        if (b->probs.empty()) throw ArgError{M3D_ERR_STATE, "batch_homogenization: no problems added"};
        b->ran = false;
        b->ms = 0;
        b->voxel_its = 0;
        if (!b->grouped) {
            batch_free_groups(b);
            for (int i = 0; i < (int)b->probs.size(); ++i) {
                const std::array<int, 3> sh{b->probs[i].nx, b->probs[i].ny, b->probs[i].nz};
                auto it = std::find_if(b->groups.begin(), b->groups.end(), [&](const BatchGroup& G) { return G.shape == sh; });
                if (it == b->groups.end()) {
                    b->groups.push_back(BatchGroup{});
                    b->groups.back().shape = sh;
                    it = b->groups.end() - 1;
                }
                it->members.push_back(i);
            }
            b->grouped = true;
        }
        // engines; device bytes the groups still need to run as one device batch each
        size_t extra = 0;
        for (BatchGroup& G : b->groups) {
            if (!G.eng) G.eng = new Engine<Elastic>(G.shape[0], G.shape[1], G.shape[2], b->precision);
            batch_group_setup(b, G);
            const int ns = 6 * (int)G.members.size();
            if (G.eng->cap_solves < ns || (b->precision == M3D_FP32_POLISH && G.eng->ring32 != G.eng->ring_needed()))
                extra += (size_t)ns * G.eng->bytes_per_solve() + (size_t)G.members.size() * G.eng->N;
        }
        size_t fr = 0, tot = 0;
        HCHK(hipMemGetInfo(&fr, &tot));
        const bool together = b->max_solves == 0 && b->groups.size() > 1 && (double)extra <= 0.7 * (double)fr;
        if (together) {
            std::vector<GroupRun> runs(b->groups.size());
            for (size_t gi = 0; gi < b->groups.size(); ++gi) {
                BatchGroup& G = b->groups[gi];
                G.per = (long long)G.members.size();
                batch_group_load(b, G, 0, G.members.size(), runs[gi]);
                G.eng->prepare(6 * (int)G.members.size());
            }
            int dev = 0;
            HCHK(hipGetDevice(&dev));
            const auto t0 = std::chrono::steady_clock::now();
            std::vector<std::thread> th;
            th.reserve(b->groups.size());
            for (size_t gi = 0; gi < b->groups.size(); ++gi) {
                th.emplace_back([&, gi]() {
                    GroupRun& run = runs[gi];
                    try {
                        HCHK(hipSetDevice(dev));
                        Engine<Elastic>& e = *b->groups[gi].eng;
                        e.solve(run.loads, tol_field, tol_eq, run.norms, 1, run.sbv, run.rvv);
                        run.means = e.means();
                    } catch (...) {
                        run.err = std::current_exception();
                    }
                });
            }
            for (std::thread& t : th) t.join();
            b->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            for (GroupRun& run : runs)
                if (run.err) std::rethrow_exception(run.err);
            for (size_t gi = 0; gi < b->groups.size(); ++gi)
                batch_group_store(b, b->groups[gi], 0, b->groups[gi].members.size(), runs[gi].means);
        } else {
            for (BatchGroup& G : b->groups) {
                Engine<Elastic>& e = *G.eng;
                if (G.per < 1) {
                    // RVEs per device batch: 70% of free memory, or the max_solves option
                    HCHK(hipMemGetInfo(&fr, &tot));
                    G.per = (long long)(0.7 * (double)fr / (double)(6 * e.bytes_per_solve() + (size_t)e.N));
                    if (b->max_solves > 0) G.per = std::min<long long>(G.per, b->max_solves / 6);
                    if (G.per < 1) throw ArgError{M3D_ERR_OOM, "batch_homogenization: one RVE's 6 solves do not fit"};
                }
                for (size_t k0 = 0; k0 < G.members.size(); k0 += (size_t)G.per) {
                    const size_t k1 = std::min(G.members.size(), k0 + (size_t)G.per);
                    GroupRun run;
                    batch_group_load(b, G, k0, k1, run);
                    e.solve(run.loads, tol_field, tol_eq, run.norms, 1, run.sbv, run.rvv);
                    batch_group_store(b, G, k0, k1, e.means());
                    b->ms += e.info.empty() ? 0.0 : e.info[0].time_ms;
                }
            }
        }
        b->ran = true;
        return M3D_OK;
    });
}

int m3d_batch_results(m3d_batch* b, int index, double* C, double* S, double* epHomo, double* sigHomo, double* ret,
                      m3d_run_info* info6)
{
    return guarded(b, [&]() {
        if (!b->ran) throw ArgError{M3D_ERR_STATE, "batch_results: run m3d_batch_homogenization first"};
        if (index < 0 || index >= (int)b->probs.size()) throw ArgError{M3D_ERR_ARG, "batch_results: index out of range"};
        const BatchProblem& q = b->probs[index];
        if (C) std::memcpy(C, q.C, sizeof q.C);
        if (S) std::memcpy(S, q.S, sizeof q.S);
        if (epHomo) std::memcpy(epHomo, q.ep, sizeof q.ep);
        if (sigHomo) std::memcpy(sigHomo, q.sg, sizeof q.sg);
        if (ret) *ret = q.ret;
        if (info6)
            for (int w = 0; w < 6; ++w) info6[w] = q.info[w];
        return M3D_OK;
    });
}

int m3d_batch_stats(m3d_batch* b, int* count, double* ms, double* voxel_iterations)
{
    return guarded(b, [&]() {
        if (count) *count = (int)b->probs.size();
        if (ms) *ms = b->ms;
        if (voxel_iterations) *voxel_iterations = b->voxel_its;
        return M3D_OK;
    });
}

} // extern "C"
