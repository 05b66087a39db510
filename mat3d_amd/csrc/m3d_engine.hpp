#pragma once
// m3d_engine.hpp -- host runtime of libmat3d.so (gfx950), shared by the ABI units.
//
// Owns the device state of one solver handle, builds the FFT plans and tile
// shapes, drives the iteration as captured hipGraphs of a chunk of
// iterations and assembles the homogenized tensors in FP64 on the host.
// Control decisions (stop rule, equilibrium-check scheduling) are taken on
// the device; the host only polls the per-solve done flags between graph
// replays.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <sys/stat.h>
#include <unistd.h>

#include "../../include/m3d.h"
#include "m3d_kernels.hpp"
#include "m3d_fast.hpp"

using namespace m3d;

namespace {

constexpr size_t LDS_MAX = 65536;

// spectrum layout of the specialised kernels: x planes per block (power of
// two; 1 = plain [x][y][kz] rows) and, for plain rows, extra complex values
// between x planes
#ifndef M3D_XBLOCK
#define M3D_XBLOCK 1
#endif
#ifndef M3D_XPAD
#define M3D_XPAD 128
#endif

// byte alignment of the spectrum rows of the specialised kernels (0: 16
// bytes, whole kz slots of padding; rows of even C are then packed at Nzh*C
// complex values).  64-byte rows were measured slower in an A/B on one box at
// 256^3 (z pass +6%, y passes +7%, x pass with the Green step -1%).
#ifndef M3D_ROW_ALIGN
#define M3D_ROW_ALIGN 0
#endif

// workgroups of a check-pass launch (the check kernels loop over their
// tiles; 0: the physics' own choice, PH::chk_grid)
#ifndef M3D_CHK_WG
#define M3D_CHK_WG 0
#endif

struct HipError {
    hipError_t e;
    const char* what;
};

#define HCHK(x)                                                                                                        \
    do {                                                                                                               \
        hipError_t _e = (x);                                                                                           \
        if (_e != hipSuccess) throw HipError{_e, #x};                                                                  \
    } while (0)

struct ArgError {
    int code;
    std::string msg;
};

// ------------------------------------------------------------ FFT planning
bool make_axis(int L, int table_len, FftAxis& a)
{
    std::memset(&a, 0, sizeof a);
    a.L = L;
    a.tws = table_len / L;
    int m = L, e2 = 0;
    while (m % 2 == 0) { m /= 2; ++e2; }
    std::vector<int> r;
    // powers of two: radix 8 where possible, 4*4 instead of 8*2
    while (e2 >= 3) {
        if (e2 == 4) { r.push_back(4); r.push_back(4); e2 = 0; break; }
        r.push_back(8);
        e2 -= 3;
    }
    if (e2 == 2) r.push_back(4);
    if (e2 == 1) r.push_back(2);
    for (int p : {3, 5, 7, 11, 13})
        while (m % p == 0) { r.push_back(p); m /= p; }
    if (m != 1 || (int)r.size() > 16) return false;
    a.np = (int)r.size();
    a.rad = 0;
    for (int i = 0; i < a.np; ++i) a.rad |= (unsigned long long)r[i] << (4 * i);
    return true;
}

template <typename T>
std::vector<cx<T>> twiddles(int n)
{
    std::vector<cx<T>> t(n);
    for (int m = 0; m < n; ++m) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        t[m] = cx<T>{(T)cosl(a), (T)(-sinl(a))};
    }
    return t;
}

// tile shapes for precision of complex size csz
void plan_tiles(Geo& g, size_t csz)
{
    const int C = g.C;
    const int Lz = g.zeven ? g.Nz / 2 : g.Nz;
    g.lsz = g.zeven ? g.Nzh : g.Nz;
    g.TY = 0;
    for (int ty = g.Ny; ty >= 1; --ty) {
        if (g.Ny % ty) continue;
        if ((long)ty * C * Lz > (long)EPT * ZTHREADS) continue;
        if ((size_t)ty * C * g.lsz * csz > LDS_MAX) continue;
        g.TY = ty;
        break;
    }
    if (!g.TY) throw ArgError{M3D_ERR_UNSUPPORTED, "z line too long for one workgroup"};
    g.nty = g.Ny / g.TY;
    auto tk = [&](int L) {
        long a = (long)EPT * XYTHREADS / ((long)C * L);
        long b = (long)(LDS_MAX / ((size_t)C * L * csz));
        long t = std::min(a, b);
        t = std::min<long>(t, g.Nzh);
        return (int)t;
    };
    g.TKy = tk(g.Ny);
    g.TKx = tk(g.Nx);
    if (g.TKy < 1 || g.TKx < 1) throw ArgError{M3D_ERR_UNSUPPORTED, "x/y line too long for one workgroup"};
    g.nky = (g.Nzh + g.TKy - 1) / g.TKy;
    g.nkx = (g.Nzh + g.TKx - 1) / g.TKx;
    g.nbz = g.Nx * g.nty;
    g.nbx = g.Ny * g.nkx;
    g.TKc = g.TKx;
    g.nbc = g.nbx;
}

// line lengths with compile-time specialised kernels (per physics)
bool fast_xy(int C, int L)
{
    if (C == 6) return L == 64 || L == 128 || L == 256 || L == 192 || L == 96 || L == 32 || L == 48;
    return L == 64 || L == 128 || L == 256 || L == 512;
}
bool fast_z(int C, int LZ)
{
    if (C == 6) return LZ == 32 || LZ == 64 || LZ == 128 || LZ == 96 || LZ == 48 || LZ == 16 || LZ == 24;
    return LZ == 32 || LZ == 64 || LZ == 128 || LZ == 256;
}

template <typename T>
Geo make_geo(int Nx, int Ny, int Nz, int C)
{
    const size_t csz = sizeof(cx<T>);
    Geo g;
    std::memset(&g, 0, sizeof g);
    g.Nx = Nx; g.Ny = Ny; g.Nz = Nz; g.Nzh = Nz / 2 + 1; g.C = C;
    g.N = (long long)Nx * Ny * Nz;
    g.H = (long long)Nx * Ny * g.Nzh;
    g.zeven = (Nz % 2 == 0);
    g.invN = 1.0 / (double)g.N;
    if (!make_axis(g.zeven ? Nz / 2 : Nz, Nz, g.az) || !make_axis(Ny, Ny, g.ay) || !make_axis(Nx, Nx, g.ax))
        throw ArgError{M3D_ERR_UNSUPPORTED, "grid size has a prime factor larger than 13"};
    plan_tiles(g, csz);
    const FftAxis* a[3] = {&g.az, &g.ay, &g.ax};
    for (const FftAxis* x : a)
        for (int i = 0; i < x->np; ++i) {
            const int r = x->radix(i);
            if (r != 2 && r != 3 && r != 4 && r != 8) g.gen = 1;
        }
    // compile-time specialised passes where available
    // the z kernels are instantiated for the tile a cubic grid would use
    if (g.zeven && fast_z(C, Nz / 2) && fast::zlines<T>(Nz / 2, C, Ny) == fast::zlines<T>(Nz / 2, C, Nz)) {
        g.fz = 1;
        g.TY = fast::zlines<T>(Nz / 2, C, Ny);
        g.nty = Ny / g.TY;
        g.nbz = Nx * g.nty;
    }
    if (fast_xy(C, Ny)) {
        g.fy = 1;
        g.TKy = fast::xytile<T>(Ny, C);
        g.nky = (g.Nzh + g.TKy - 1) / g.TKy;
    }
    if (fast_xy(C, Nx)) {
        g.fx = 1;
        g.TKx = fast::xtile<T>(Nx, C);
        g.nkx = (g.Nzh + g.TKx - 1) / g.TKx;
        g.nbx = Ny * g.nkx;
        g.TKc = fast::xctile<T>(Nx, C);
        g.nbc = Ny * ((g.Nzh + g.TKc - 1) / g.TKc);
    }
    // spectrum row layout (the runtime-planned kernels address rows (x, y)
    // at (x * Ny + y) * Nzh * C: plain rows without padding)
    const long long rowc = (long long)g.Nzh * C;
    const bool all_fast = g.fz && g.fy && g.fx;
    // the specialised kernels may start every row on an M3D_ROW_ALIGN-byte
    // boundary (whole kz slots of padding): their y and x passes move a row's
    // kz tile as one contiguous chunk, and a chunk that straddles 64-byte
    // sectors is written as partial sectors (measured: 64-byte rows made the
    // z and y passes slower).  By default rows are padded to 16 bytes only,
    // which changes nothing for even C and gives odd C in FP32 (thermal, rows
    // of Nzh*3 values) an even row pitch, so that its y and x passes move
    // 16-byte pairs too
    long long rowp = rowc;
    constexpr long long row_align = M3D_ROW_ALIGN > 0 ? M3D_ROW_ALIGN : 16;
    if (all_fast)
        while ((rowp * (long long)csz) % row_align) rowp += C;
    g.lxb = 0;
    g.xbm = 0;
    g.sxb = rowp;
    g.sy = rowp;
    g.sxa = (long long)Ny * rowp;
    if (all_fast && M3D_XBLOCK > 1 && Nx % M3D_XBLOCK == 0) {
        while ((1 << g.lxb) < M3D_XBLOCK) ++g.lxb;
        g.xbm = M3D_XBLOCK - 1;
        g.sxa = (long long)Ny * M3D_XBLOCK * rowp;
        g.sy = (long long)M3D_XBLOCK * rowp;
    } else if (all_fast) {
        g.sxa += M3D_XPAD; // x planes M3D_XPAD complex values apart beyond their rows
    }
    g.spec = (long long)((Nx - 1) >> g.lxb) * g.sxa + (long long)((Nx - 1) & g.xbm) * g.sxb + (long long)(Ny - 1) * g.sy + rowc;
    return g;
}

// ---------------------------------------------------------- small algebra
bool inv_n(const double* A, double* X, int n)
{
    double M[36];
    for (int i = 0; i < n * n; ++i) { M[i] = A[i]; X[i] = 0; }
    for (int i = 0; i < n; ++i) X[i + i * n] = 1;
    for (int c = 0; c < n; ++c) {
        int p = c;
        double best = std::fabs(M[c + c * n]);
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(M[r + c * n]) > best) { best = std::fabs(M[r + c * n]); p = r; }
        if (best == 0.0 || !std::isfinite(best)) return false;
        if (p != c)
            for (int k = 0; k < n; ++k) { std::swap(M[c + k * n], M[p + k * n]); std::swap(X[c + k * n], X[p + k * n]); }
        double d = M[c + c * n];
        for (int k = 0; k < n; ++k) { M[c + k * n] /= d; X[c + k * n] /= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = M[r + c * n];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) { M[r + k * n] -= f * M[c + k * n]; X[r + k * n] -= f * X[c + k * n]; }
        }
    }
    return true;
}

void mul_n(const double* A, const double* B, double* Cm, int n)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double s = 0;
            for (int k = 0; k < n; ++k) s += A[i + k * n] * B[k + j * n];
            Cm[i + j * n] = s;
        }
}

void iso6(double la, double mu, double* C6)
{
    std::memset(C6, 0, sizeof(double) * 36);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C6[i + 6 * j] = la + (i == j ? 2 * mu : 0.0);
    for (int i = 3; i < 6; ++i) C6[i + 6 * i] = 2 * mu;
}

// max_min_Array semantics (first index of the extreme), FFT_analysis3D.cpp:5-23
void max_min(const std::vector<double>& a, double& mx, int& imx, double& mn)
{
    mx = a[0]; imx = 0; mn = a[0];
    for (size_t i = 1; i < a.size(); ++i) {
        if (a[i] > mx) { mx = a[i]; imx = (int)i; }
        if (a[i] < mn) mn = a[i];
    }
}

// --------------------------------------------------------------- engine
// ------------------------------------------------------------------ autotuner
// Variants of the FP32 y pass and x pass with the Green step: threads per
// workgroup (256 / 512), the radix plan of the line FFTs (radix 16 first or
// radix 8 first) and, for the x pass, the kz tile (xtile / 1.5x wider).
// Each is timed on the live buffers of the first FP32 solve of a (device,
// grid, components, solves) shape.  The choices are kept for the process and
// in a text file (M3D_TUNE_FILE, default $XDG_CACHE_HOME or
// $HOME/.cache/mat3d_amd/tune.txt; "none" keeps them in the process only),
// so later processes on the same device take them without timing.
struct XyTune {
    int xg_nt = 512, xg_wide = 0, xg_rp = 0, y_nt = 512, y_rp = 0;
    int from_file = 0; // taken from the tuning file (1) or timed here (0)
};

struct TuneKey {
    std::string dev; // device architecture and CU count
    int nx, ny, nz, c, ns;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(dev, nx, ny, nz, c, ns) < std::tie(o.dev, o.nx, o.ny, o.nz, o.c, o.ns);
    }
};

inline std::map<TuneKey, XyTune>& tune_cache()
{
    static std::map<TuneKey, XyTune> m;
    return m;
}
inline std::mutex& tune_mutex()
{
    static std::mutex mu;
    return mu;
}

// entries of the tuning file are lines "dev nx ny nz c ns xg_nt xg_wide
// xg_rp y_nt y_rp" (precision: FP32 passes)
inline std::string tune_file_path()
{
    const char* e = std::getenv("M3D_TUNE_FILE");
    if (e && *e) return (std::string(e) == "none") ? std::string() : std::string(e);
    const char* x = std::getenv("XDG_CACHE_HOME");
    const char* h = std::getenv("HOME");
    if (x && *x) return std::string(x) + "/mat3d_amd/tune.txt";
    if (h && *h) return std::string(h) + "/.cache/mat3d_amd/tune.txt";
    return std::string();
}

// caller holds tune_mutex
inline void tune_file_load()
{
    static bool loaded = false;
    if (loaded) return;
    loaded = true;
    const std::string path = tune_file_path();
    if (path.empty()) return;
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return;
    char dev[128];
    TuneKey k;
    XyTune t;
    while (std::fscanf(f, "%127s %d %d %d %d %d %d %d %d %d %d", dev, &k.nx, &k.ny, &k.nz, &k.c, &k.ns, &t.xg_nt,
                       &t.xg_wide, &t.xg_rp, &t.y_nt, &t.y_rp) == 11) {
        const bool ok = (t.xg_nt == 256 || t.xg_nt == 512) && (t.y_nt == 256 || t.y_nt == 512) &&
                        (t.xg_wide == 0 || t.xg_wide == 1) && (t.xg_rp == 0 || t.xg_rp == 1) &&
                        (t.y_rp == 0 || t.y_rp == 1);
        if (!ok) continue; // a line this build does not know: timed again
        k.dev = dev;
        t.from_file = 1;
        tune_cache()[k] = t;
    }
    std::fclose(f);
}

// caller holds tune_mutex; rewrites the file from the cache (tmp + rename)
inline void tune_file_store()
{
    const std::string path = tune_file_path();
    if (path.empty()) return;
    const size_t slash = path.rfind('/');
    if (slash != std::string::npos) {
        // create the directory chain (best effort: no file, no persistence)
        std::string d;
        for (size_t i = 0; i <= slash; ++i) {
            if (path[i] == '/' && !d.empty()) (void)mkdir(d.c_str(), 0755);
            d.push_back(path[i]);
        }
    }
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return;
    for (const auto& kv : tune_cache()) {
        const TuneKey& k = kv.first;
        const XyTune& t = kv.second;
        std::fprintf(f, "%s %d %d %d %d %d %d %d %d %d %d\n", k.dev.c_str(), k.nx, k.ny, k.nz, k.c, k.ns, t.xg_nt,
                     t.xg_wide, t.xg_rp, t.y_nt, t.y_rp);
    }
    const bool ok = std::fclose(f) == 0;
    if (ok) (void)std::rename(tmp.c_str(), path.c_str());
    else (void)std::remove(tmp.c_str());
}

inline std::string device_tag()
{
    int dev = 0;
    hipDeviceProp_t pr;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&pr, dev) != hipSuccess) return "unknown";
    std::string a = pr.gcnArchName;
    const size_t c = a.find(':');
    if (c != std::string::npos) a = a.substr(0, c);
    return a + "x" + std::to_string(pr.multiProcessorCount);
}

template <typename PH>
struct Engine {
    static constexpr int C = PH::C;
    int Nx, Ny, Nz;
    int precision;
    Geo g32{}, g64{};
    int cap_solves = 0;
    long long N = 0, H = 0;
    // device buffers
    float* U32 = nullptr;  // ring of rollback+1 FP32 iterates x solves
    float* R32 = nullptr;  // anchored polish: r = T(A) - A per solve (FP32)
    double* U64 = nullptr; // two FP64 iterates x solves
    int ring32 = 0;        // ring size the FP32 buffer was allocated with
    unsigned char* Sbuf = nullptr;
    unsigned char* Kbuf = nullptr;
    double* psum = nullptr;
    double* pchk = nullptr;
    Ctl* ctl = nullptr;
    double* mpart = nullptr;
    int* rve_of = nullptr;
    uint8_t* ph = nullptr;
    size_t ph_cap = 0;
    float* coef32 = nullptr;
    double* coef64 = nullptr;
    float* prm32 = nullptr;
    double* prm64 = nullptr;
    cx<float>* tw32[3] = {nullptr, nullptr, nullptr};
    cx<double>* tw64[3] = {nullptr, nullptr, nullptr};
    // twiddle tables {z, y, x} of one precision
    template <typename T>
    const cx<T>* const* twiddles_of() const
    {
        if constexpr (sizeof(T) == 8) return tw64;
        else return tw32;
    }
    double* fbuf = nullptr; // field output staging (2*C*N doubles)
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // host state
    int nsolve = 0;
    std::vector<int> final_prec; // per solve: 0 fp32, 1 fp64
    std::vector<m3d_run_info> info;
    double polish_factor = 10.0;
    int refine = 1; // anchored FP32 polish (1) or plain FP64 polish iterations (0)
    int autotune = 1; // time the x/y pass variants at the first FP32 solve of a shape
    XyTune tune32;    // variants of the FP32 x/y passes in use
    bool tuned32 = false;
    int tuned_ns = 0; // batch size the variants were timed at
    int rollback = 5;
    double rollback_margin = 2.0;
    double fp32_floor = 1e-6;
    int cap = 2000;
    int chunk = 0;
    int mnblk = 256;

    Engine(int nx, int ny, int nz, int prec) : Nx(nx), Ny(ny), Nz(nz), precision(prec)
    {
        g32 = make_geo<float>(nx, ny, nz, C);
        g64 = make_geo<double>(nx, ny, nz, C);
        N = g64.N;
        H = g64.H;
        HCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HCHK(hipEventCreate(&ev0));
        HCHK(hipEventCreate(&ev1));
        // twiddle tables: z table has length Nz, y: Ny, x: Nx
        const int lens[3] = {Nz, Ny, Nx};
        for (int a = 0; a < 3; ++a) {
            auto t32 = twiddles<float>(lens[a]);
            auto t64 = twiddles<double>(lens[a]);
            HCHK(hipMalloc(&tw32[a], sizeof(cx<float>) * lens[a]));
            HCHK(hipMalloc(&tw64[a], sizeof(cx<double>) * lens[a]));
            HCHK(hipMemcpy(tw32[a], t32.data(), sizeof(cx<float>) * lens[a], hipMemcpyHostToDevice));
            HCHK(hipMemcpy(tw64[a], t64.data(), sizeof(cx<double>) * lens[a], hipMemcpyHostToDevice));
        }
        HCHK(hipMalloc(&ph, (size_t)N));
        ph_cap = (size_t)N;
    }

    void free_solves()
    {
        if (U32) hipFree(U32);
        if (R32) hipFree(R32);
        if (U64) hipFree(U64);
        U32 = nullptr;
        R32 = nullptr;
        U64 = nullptr;
        if (Sbuf) hipFree(Sbuf);
        if (Kbuf) hipFree(Kbuf);
        if (psum) hipFree(psum);
        if (pchk) hipFree(pchk);
        if (ctl) hipFree(ctl);
        if (mpart) hipFree(mpart);
        if (rve_of) hipFree(rve_of);
        Sbuf = Kbuf = nullptr;
        psum = pchk = mpart = nullptr;
        ctl = nullptr;
        rve_of = nullptr;
        cap_solves = 0;
    }

    ~Engine()
    {
        if (st) hipStreamSynchronize(st);
        free_solves();
        if (ph) hipFree(ph);
        if (coef32) hipFree(coef32);
        if (coef64) hipFree(coef64);
        if (prm32) hipFree(prm32);
        if (prm64) hipFree(prm64);
        if (fbuf) hipFree(fbuf);
        for (int a = 0; a < 3; ++a) {
            if (tw32[a]) hipFree(tw32[a]);
            if (tw64[a]) hipFree(tw64[a]);
        }
        if (ev0) hipEventDestroy(ev0);
        if (ev1) hipEventDestroy(ev1);
        if (st) hipStreamDestroy(st);
    }

    template <typename P>
    void alloc(P** p, size_t bytes)
    {
        hipError_t e = hipMalloc((void**)p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *p = nullptr;
            throw ArgError{M3D_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed"};
        }
    }

    int ring_needed() const { return std::max(2, rollback + 1); }

    // bytes of one solve's spectrum buffer: room for either precision's layout
    size_t spec_bytes() const
    {
        const size_t b32 = (size_t)g32.spec * sizeof(cx<float>), b64 = (size_t)g64.spec * sizeof(cx<double>);
        const size_t b = std::max(b32, b64);
        return (b + 255) / 256 * 256;
    }

    // device bytes one solve adds (fields, spectra), for batch sizing
    size_t bytes_per_solve() const
    {
        const size_t nc = (size_t)N * C, hc = spec_bytes() / sizeof(cx<double>);
        size_t b = sizeof(double) * nc * 2 + 2 * sizeof(cx<double>) * hc;
        if (precision == M3D_FP32_POLISH) b += sizeof(float) * nc * (ring_needed() + 1);
        return b;
    }

    void ensure_solves(int n)
    {
        if (n <= cap_solves && (precision != M3D_FP32_POLISH || ring32 == ring_needed())) return;
        free_solves();
        const size_t nc = (size_t)N * C;
        const size_t hc = spec_bytes() / sizeof(cx<double>);
        if (precision == M3D_FP32_POLISH) {
            ring32 = ring_needed();
            alloc(&U32, sizeof(float) * nc * n * ring32);
            alloc(&R32, sizeof(float) * nc * n);
        }
        alloc(&U64, sizeof(double) * nc * n * 2);
        alloc(&Sbuf, sizeof(cx<double>) * hc * n);
        alloc(&Kbuf, sizeof(cx<double>) * hc * n);
        const int nbz = std::max(g32.nbz, g64.nbz), nbx = std::max(g32.nbc, g64.nbc);
        alloc(&psum, sizeof(double) * (size_t)nbz * n);
        alloc(&pchk, sizeof(double) * (size_t)2 * (nbx + 1) * n);
        alloc(&ctl, sizeof(Ctl) * n);
        alloc(&mpart, sizeof(double) * (size_t)2 * C * mnblk * n);
        alloc(&rve_of, sizeof(int) * n);
        std::vector<int> z(n, 0);
        HCHK(hipMemcpy(rve_of, z.data(), sizeof(int) * n, hipMemcpyHostToDevice));
        cap_solves = n;
    }

    template <typename T>
    Ptr<T> ptrs(const Geo& g)
    {
        Ptr<T> p;
        std::memset(&p, 0, sizeof p);
        const size_t nc = (size_t)N * C;
        if constexpr (sizeof(T) == 4) {
            p.U = U32;
            p.ring = ring32;
            p.coef = coef32;
            p.twz = tw32[0];
            p.twy = tw32[1];
            p.twx = tw32[2];
        } else {
            p.U = U64;
            p.ring = 2;
            p.coef = coef64;
            p.twz = tw64[0];
            p.twy = tw64[1];
            p.twx = tw64[2];
        }
        p.sU = (long long)nc;
        p.sB = (long long)nc * cap_solves;
        p.S = reinterpret_cast<cx<T>*>(Sbuf);
        p.K = reinterpret_cast<cx<T>*>(Kbuf);
        p.sS = (long long)(spec_bytes() / sizeof(cx<T>));
        p.ph = ph;
        p.sPh = N;
        p.rve_of = rve_of;
        p.sCoef = 256 * NCOEF;
        p.psum = psum;
        p.pchk = pchk;
        p.ctl = ctl;
        p.R = R32;
        p.eU = U32;
        p.eB = (long long)nc * cap_solves;
        p.ering = ring32 > 0 ? ring32 : 1;
        (void)g;
        return p;
    }

    template <typename T>
    const T* prm()
    {
        if constexpr (sizeof(T) == 4) return prm32;
        else return prm64;
    }

    // ---------------------------------------------------------- launches
    // ------------------------------------------------- per-pass dispatch
    template <typename T, int MODE, int LZ>
    bool zf(const Geo& g, const Ptr<T>& p)
    {
        if (g.Nz != 2 * LZ) return false;
        constexpr int TYc = fast::zlines<T>(LZ, C, 2 * LZ);
        if (g.TY != TYc) throw ArgError{M3D_ERR_UNSUPPORTED, "z tile mismatch"};
        const size_t sm = ((size_t)(LZ + 1) * TYc * C + 2 * LZ) * sizeof(cx<T>);
        hipLaunchKernelGGL((fast::kz<T, PH, MODE, LZ, TYc>), dim3(grid_of(g.nsolve * g.nbz, MODE == MODE_CHK_CUR ||
                                                                                            MODE == MODE_CHK_PREV)),
                           dim3(fast::ZNT), sm, st, g, p);
        return true;
    }

    template <typename T, int MODE>
    void zpass(const Geo& g, const Ptr<T>& p)
    {
        if (g.fz) {
            bool ok;
            if constexpr (C == 6)
                ok = zf<T, MODE, 128>(g, p) || zf<T, MODE, 64>(g, p) || zf<T, MODE, 32>(g, p) || zf<T, MODE, 96>(g, p) ||
                     zf<T, MODE, 48>(g, p) || zf<T, MODE, 16>(g, p) || zf<T, MODE, 24>(g, p);
            else
                ok = zf<T, MODE, 256>(g, p) || zf<T, MODE, 128>(g, p) || zf<T, MODE, 64>(g, p) || zf<T, MODE, 32>(g, p);
            if (!ok) throw ArgError{M3D_ERR_UNSUPPORTED, "no z kernel"};
            return;
        }
        const size_t zsm = (size_t)g.TY * C * g.lsz * sizeof(cx<T>);
        const dim3 grid(g.nsolve * g.nbz);
        // the generic kernel adds r of anchored solves at run time; the anchor
        // mode is the specialised kernels' only (callers check g.fz)
        static_assert((MODE != MODE_IUA && MODE != MODE_FWDP) || sizeof(T) == 8, "polish modes are FP64");
        if (MODE == MODE_IUA || MODE == MODE_FWDP)
            throw ArgError{M3D_ERR_STATE, "polish z pass without the specialised kernels"};
        constexpr int GM = (MODE == MODE_IUFR || MODE == MODE_IUA) ? MODE_IUF : (MODE == MODE_FWDP ? MODE_FWD : MODE);
        if (g.gen) hipLaunchKernelGGL((k_zpass<T, PH, GM, true>), grid, dim3(ZTHREADS), zsm, st, g, p);
        else hipLaunchKernelGGL((k_zpass<T, PH, GM, false>), grid, dim3(ZTHREADS), zsm, st, g, p);
    }

    // threads of a finaliser workgroup over n partial sums per solve: the
    // wide form where the loads dominate (large grids); 256 threads
    // otherwise, since a 1024-thread workgroup waits for a whole free CU
    // while the kernels of other engines' streams share the GPU
    static unsigned fin_threads(long long n) { return n >= 8192 ? ATHREADS : FTHREADS; }

    // workgroups of a launch: one per tile, or for the check passes (which
    // loop over their tiles) at most M3D_CHK_WG (PH::chk_grid by default)
    static unsigned grid_of(long long tiles, bool check)
    {
        constexpr long long cap = M3D_CHK_WG > 0 ? M3D_CHK_WG : PH::chk_grid;
        return (unsigned)(check ? std::min<long long>(tiles, cap) : tiles);
    }

    template <typename T, int KIND, int DIR, int WHICH, int L, int TK, int NTH, int RP = 0>
    void xy_launch(const Geo& g, const Ptr<T>& p)
    {
        const int nkt = (g.Nzh + TK - 1) / TK;
        const int outer = KIND == 0 ? g.Nx : g.Ny;
        const size_t sm = ((size_t)L * TK * C + L) * sizeof(cx<T>);
        const dim3 grid(grid_of(g.nsolve * outer * nkt, WHICH != BUF_MAIN));
        constexpr bool LOOP = WHICH != BUF_MAIN;
        if (sm > 65536) {
            // gfx950: up to 160 KB of LDS per workgroup
            HCHK(hipFuncSetAttribute(reinterpret_cast<const void*>(&fast::kxy<T, PH, KIND, DIR, WHICH, L, TK, NTH, LOOP, RP>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm));
        }
        hipLaunchKernelGGL((fast::kxy<T, PH, KIND, DIR, WHICH, L, TK, NTH, LOOP, RP>), grid, dim3(NTH), sm, st, g, p,
                           prm<T>());
    }

    // KIND 0: y pass, 1: x pass + Green, 2: x pass of the check FFT.  The FP32
    // y pass and x pass with the Green step run the variant the autotuner
    // picked (threads per workgroup, kz tile width); the rest one shape.
    template <typename T, int KIND, int DIR, int WHICH, int L>
    bool xyf(const Geo& g, const Ptr<T>& p)
    {
        const int Lax = KIND == 0 ? g.Ny : g.Nx;
        if (Lax != L) return false;
        constexpr int TKc = KIND == 0 ? fast::xytile<T>(L, C) : (KIND == 1 ? fast::xtile<T>(L, C) : fast::xctile<T>(L, C));
        const int TK = KIND == 0 ? g.TKy : (KIND == 1 ? g.TKx : g.TKc);
        if (TK != TKc) throw ArgError{M3D_ERR_UNSUPPORTED, "xy tile mismatch"};
        if constexpr (sizeof(T) == 4 && KIND == 1) {
            constexpr int TKw = fast::xtile_wide<T>(L, C);
            constexpr bool RPS = fast::plans_differ(L);
            const XyTune& tu = tune32;
            const int rp = RPS ? tu.xg_rp : 0;
            if (tu.xg_wide) {
                if (tu.xg_nt == 512) {
                    if (RPS && rp) xy_launch<T, KIND, DIR, WHICH, L, TKw, 512, RPS>(g, p);
                    else xy_launch<T, KIND, DIR, WHICH, L, TKw, 512>(g, p);
                } else {
                    if (RPS && rp) xy_launch<T, KIND, DIR, WHICH, L, TKw, 256, RPS>(g, p);
                    else xy_launch<T, KIND, DIR, WHICH, L, TKw, 256>(g, p);
                }
            } else {
                if (tu.xg_nt == 512) {
                    if (RPS && rp) xy_launch<T, KIND, DIR, WHICH, L, TKc, 512, RPS>(g, p);
                    else xy_launch<T, KIND, DIR, WHICH, L, TKc, 512>(g, p);
                } else {
                    if (RPS && rp) xy_launch<T, KIND, DIR, WHICH, L, TKc, 256, RPS>(g, p);
                    else xy_launch<T, KIND, DIR, WHICH, L, TKc, 256>(g, p);
                }
            }
        } else if constexpr (sizeof(T) == 4 && KIND == 0 && WHICH == BUF_MAIN) {
            // the main y passes (forward and inverse) take the tuned variant;
            // the check's y pass keeps the default radix plan
            constexpr bool RPS = fast::plans_differ(L);
            const int rp = RPS ? tune32.y_rp : 0;
            if (tune32.y_nt == 512) {
                if (RPS && rp) xy_launch<T, KIND, DIR, WHICH, L, TKc, 512, RPS>(g, p);
                else xy_launch<T, KIND, DIR, WHICH, L, TKc, 512>(g, p);
            } else {
                if (RPS && rp) xy_launch<T, KIND, DIR, WHICH, L, TKc, 256, RPS>(g, p);
                else xy_launch<T, KIND, DIR, WHICH, L, TKc, 256>(g, p);
            }
        } else if constexpr (sizeof(T) == 4 && KIND == 0) {
            if (tune32.y_nt == 512) xy_launch<T, KIND, DIR, WHICH, L, TKc, 512>(g, p);
            else xy_launch<T, KIND, DIR, WHICH, L, TKc, 256>(g, p);
        } else {
            xy_launch<T, KIND, DIR, WHICH, L, TKc, fast::xy_threads<KIND, T>()>(g, p);
        }
        return true;
    }

    template <typename T, int KIND, int DIR, int WHICH>
    bool xyfast(const Geo& g, const Ptr<T>& p)
    {
        if constexpr (C == 6)
            return xyf<T, KIND, DIR, WHICH, 256>(g, p) || xyf<T, KIND, DIR, WHICH, 128>(g, p) ||
                   xyf<T, KIND, DIR, WHICH, 64>(g, p) || xyf<T, KIND, DIR, WHICH, 192>(g, p) ||
                   xyf<T, KIND, DIR, WHICH, 96>(g, p) || xyf<T, KIND, DIR, WHICH, 32>(g, p) ||
                   xyf<T, KIND, DIR, WHICH, 48>(g, p);
        else
            return xyf<T, KIND, DIR, WHICH, 512>(g, p) || xyf<T, KIND, DIR, WHICH, 256>(g, p) ||
                   xyf<T, KIND, DIR, WHICH, 128>(g, p) || xyf<T, KIND, DIR, WHICH, 64>(g, p);
    }

    template <typename T, int DIR, int WHICH>
    void ypass(const Geo& g, const Ptr<T>& p)
    {
        if (g.fy) {
            if (!xyfast<T, 0, DIR, WHICH>(g, p)) throw ArgError{M3D_ERR_UNSUPPORTED, "no y kernel"};
            return;
        }
        const size_t ysm = (size_t)g.Ny * g.TKy * C * sizeof(cx<T>);
        const dim3 grid(g.nsolve * g.Nx * g.nky);
        if (g.gen) hipLaunchKernelGGL((k_ypass<T, C, DIR, WHICH, true>), grid, dim3(XYTHREADS), ysm, st, g, p);
        else hipLaunchKernelGGL((k_ypass<T, C, DIR, WHICH, false>), grid, dim3(XYTHREADS), ysm, st, g, p);
    }

    template <typename T>
    void xgreen(const Geo& g, const Ptr<T>& p)
    {
        if (g.fx) {
            if (!xyfast<T, 1, -1, BUF_MAIN>(g, p)) throw ArgError{M3D_ERR_UNSUPPORTED, "no x kernel"};
            return;
        }
        const size_t xsm = (size_t)g.Nx * g.TKx * C * sizeof(cx<T>);
        const dim3 grid(g.nsolve * g.nbx);
        if (g.gen) hipLaunchKernelGGL((k_xgreen<T, PH, true>), grid, dim3(XYTHREADS), xsm, st, g, p, prm<T>());
        else hipLaunchKernelGGL((k_xgreen<T, PH, false>), grid, dim3(XYTHREADS), xsm, st, g, p, prm<T>());
    }

    template <typename T, int WHICH>
    void xcheck(const Geo& g, const Ptr<T>& p)
    {
        if (g.fx) {
            if (!xyfast<T, 2, -1, WHICH>(g, p)) throw ArgError{M3D_ERR_UNSUPPORTED, "no x check kernel"};
            return;
        }
        const size_t xsm = (size_t)g.Nx * g.TKx * C * sizeof(cx<T>);
        const dim3 grid(g.nsolve * g.nbx);
        if (g.gen) hipLaunchKernelGGL((k_xcheck<T, PH, WHICH, true>), grid, dim3(XYTHREADS), xsm, st, g, p);
        else hipLaunchKernelGGL((k_xcheck<T, PH, WHICH, false>), grid, dim3(XYTHREADS), xsm, st, g, p);
    }

    template <typename T>
    // promote: the FP64 start of the polish reads the FP32 iterate and writes
    // its FP64 copy itself (specialised kernels; k_promote otherwise)
    void launch_start(const Geo& g, const Ptr<T>& p, bool promote = false)
    {
        if constexpr (sizeof(T) == 8) {
            if (promote) zpass<T, MODE_FWDP>(g, p);
            else zpass<T, MODE_FWD>(g, p);
        } else {
            zpass<T, MODE_FWD>(g, p);
        }
        hipLaunchKernelGGL((k_fin0<C>), dim3(g.nsolve), dim3(fin_threads(g.nbz)), 0, st, g, p.ctl, p.psum);
        ypass<T, -1, BUF_MAIN>(g, p);
        xgreen<T>(g, p);
        ypass<T, 1, BUF_MAIN>(g, p);
    }

    // green = false: the iteration's update and stop check only, without the
    // transforms that prepare the next iteration's Green output (the FP64
    // anchor iteration of the polish, which the FP32 corrections follow)
    template <typename T>
    void launch_iter(const Geo& g, const Ptr<T>& p, bool green = true)
    {
        // the FP64 anchor iteration on the specialised kernels also writes r
        // and e (MODE_IUA; k_anchor otherwise)
        if constexpr (sizeof(T) == 8) {
            if (!green && g.fz) zpass<T, MODE_IUA>(g, p);
            else zpass<T, MODE_IUF>(g, p);
        } else {
            zpass<T, MODE_IUF>(g, p);
        }
        hipLaunchKernelGGL((k_finA<C>), dim3(g.nsolve), dim3(fin_threads(g.nbz)), 0, st, g, p.ctl, p.psum);
        zpass<T, MODE_CHK_CUR>(g, p);
        ypass<T, -1, BUF_CHK_CUR>(g, p);
        xcheck<T, BUF_CHK_CUR>(g, p);
        zpass<T, MODE_CHK_PREV>(g, p);
        ypass<T, -1, BUF_CHK_PREV>(g, p);
        xcheck<T, BUF_CHK_PREV>(g, p);
        hipLaunchKernelGGL((k_finB<C>), dim3(g.nsolve), dim3(fin_threads(g.nbc)), 0, st, g, p.ctl, p.pchk);
        if (!green) return;
        ypass<T, -1, BUF_MAIN>(g, p);
        xgreen<T>(g, p);
        ypass<T, 1, BUF_MAIN>(g, p);
    }

    int chunk_for(const Geo& g) const
    {
        if (chunk > 0) return chunk;
        const double work = (double)g.N * g.nsolve;
        if (work >= 3.0e7) return 2;
        if (work >= 4.0e6) return 4;
        return 8;
    }

    // anchored polish: FP32 passes on the corrections e (r added in the
    // update), equilibrium checks in FP64 on anchor + e
    void launch_start_corr(const Geo& gf, const Ptr<float>& pf)
    {
        zpass<float, MODE_FWD>(gf, pf);
        hipLaunchKernelGGL((k_fin0<C>), dim3(gf.nsolve), dim3(fin_threads(gf.nbz)), 0, st, gf, pf.ctl, pf.psum);
        ypass<float, -1, BUF_MAIN>(gf, pf);
        xgreen<float>(gf, pf);
        ypass<float, 1, BUF_MAIN>(gf, pf);
    }

    void launch_iter_corr(const Geo& gf, const Ptr<float>& pf, const Geo& gd, const Ptr<double>& pd)
    {
        zpass<float, MODE_IUFR>(gf, pf);
        hipLaunchKernelGGL((k_finA<C>), dim3(gf.nsolve), dim3(fin_threads(gf.nbz)), 0, st, gf, pf.ctl, pf.psum);
        zpass<double, MODE_CHK_CUR>(gd, pd);
        ypass<double, -1, BUF_CHK_CUR>(gd, pd);
        xcheck<double, BUF_CHK_CUR>(gd, pd);
        zpass<double, MODE_CHK_PREV>(gd, pd);
        ypass<double, -1, BUF_CHK_PREV>(gd, pd);
        xcheck<double, BUF_CHK_PREV>(gd, pd);
        hipLaunchKernelGGL((k_finB<C>), dim3(gd.nsolve), dim3(fin_threads(gd.nbc)), 0, st, gd, pd.ctl, pd.pchk);
        ypass<float, -1, BUF_MAIN>(gf, pf);
        xgreen<float>(gf, pf);
        ypass<float, 1, BUF_MAIN>(gf, pf);
    }

    // capture `ck` iterations (iter()) in a graph and replay it until every
    // solve of the batch is done; the host polls the done flags in between
    template <typename Iter>
    void run_graph(int ck, Iter&& iter, std::vector<Ctl>& hctl)
    {
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < ck; ++i) iter();
        hipError_t ce = hipStreamEndCapture(st, &graph);
        if (ce != hipSuccess) throw HipError{ce, "hipStreamEndCapture"};
        ce = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ce != hipSuccess) {
            hipGraphDestroy(graph);
            throw HipError{ce, "hipGraphInstantiate"};
        }
        try {
            for (int guard = 0; guard < 100000; ++guard) {
                HCHK(hipGraphLaunch(exec, st));
                HCHK(hipMemcpyAsync(hctl.data(), ctl, sizeof(Ctl) * nsolve, hipMemcpyDeviceToHost, st));
                HCHK(hipStreamSynchronize(st));
                bool all = true;
                for (int s = 0; s < nsolve; ++s) all = all && hctl[s].done;
                if (all) break;
            }
        } catch (...) {
            hipGraphExecDestroy(exec);
            hipGraphDestroy(graph);
            throw;
        }
        hipGraphExecDestroy(exec);
        hipGraphDestroy(graph);
    }

    // run the iteration in precision T until every solve of the batch is done
    template <typename T>
    void run_phase(Geo g, std::vector<Ctl>& hctl)
    {
        g.nsolve = nsolve;
        Ptr<T> p = ptrs<T>(g);
        launch_start<T>(g, p);
        HCHK(hipGetLastError());
        run_graph(chunk_for(g), [&]() { launch_iter<T>(g, p); }, hctl);
    }

    // anchored polish of the solves flagged corr (r and e from k_anchor, or
    // from the anchor iteration's z pass on the specialised kernels)
    void run_corr(Geo gf, Geo gd, std::vector<Ctl>& hctl)
    {
        gf.nsolve = gd.nsolve = nsolve;
        Ptr<float> pf = ptrs<float>(gf);
        Ptr<double> pd = ptrs<double>(gd);
        if (!gd.fz) hipLaunchKernelGGL((k_anchor<C>), dim3(1024), dim3(256), 0, st, gd, pd, pf, ctl);
        launch_start_corr(gf, pf);
        HCHK(hipGetLastError());
        run_graph(chunk_for(gf), [&]() { launch_iter_corr(gf, pf, gd, pd); }, hctl);
    }

    // time the FP32 x/y pass variants on this shape (first FP32 solve); the
    // kernels run on the live buffers, whose contents the solve then rewrites
    void tune_xy(Geo g)
    {
        g.nsolve = nsolve;
        const TuneKey key{device_tag(), Nx, Ny, Nz, C, nsolve};
        {
            std::lock_guard<std::mutex> lk(tune_mutex());
            tune_file_load();
            auto it = tune_cache().find(key);
            if (it != tune_cache().end()) {
                tune32 = it->second;
                tuned32 = true;
                return;
            }
        }
        Ptr<float> p = ptrs<float>(g);
        // each variant's time is the best of `rounds` timings of `reps`
        // launches, the variants interleaved within every round, so a drift
        // of the clocks during the tuning does not decide the choice
        constexpr int rounds = 4, reps = 2;
        auto time_it = [&](auto&& launch) {
            HCHK(hipEventRecord(ev0, st));
            for (int r = 0; r < reps; ++r) launch();
            HCHK(hipEventRecord(ev1, st));
            HCHK(hipEventSynchronize(ev1));
            float ms = 0;
            HCHK(hipEventElapsedTime(&ms, ev0, ev1));
            return ms;
        };
        XyTune best = tune32;
        best.from_file = 0;
        if (g.fx) {
            std::vector<std::array<int, 3>> vars; // {threads, wide, radix plan}
            for (int rp : {0, 1})
                for (int nt : {256, 512})
                    for (int wide : {0, 1}) vars.push_back({nt, wide, rp});
            std::vector<float> bt(vars.size(), 1e30f);
            for (int r = 0; r <= rounds; ++r)
                for (size_t v = 0; v < vars.size(); ++v) {
                    tune32.xg_nt = vars[v][0];
                    tune32.xg_wide = vars[v][1];
                    tune32.xg_rp = vars[v][2];
                    const float t = time_it([&]() { xgreen<float>(g, p); });
                    if (r > 0) bt[v] = std::min(bt[v], t); // the first timing pass warms every variant up
                }
            const size_t k = std::min_element(bt.begin(), bt.end()) - bt.begin();
            best.xg_nt = vars[k][0];
            best.xg_wide = vars[k][1];
            best.xg_rp = vars[k][2];
        }
        tune32 = best;
        if (g.fy) {
            std::vector<std::array<int, 2>> vars; // {threads, radix plan}
            for (int rp : {0, 1})
                for (int nt : {256, 512}) vars.push_back({nt, rp});
            std::vector<float> bt(vars.size(), 1e30f);
            for (int r = 0; r <= rounds; ++r)
                for (size_t v = 0; v < vars.size(); ++v) {
                    tune32.y_nt = vars[v][0];
                    tune32.y_rp = vars[v][1];
                    const float t = time_it([&]() { ypass<float, -1, BUF_MAIN>(g, p); });
                    if (r > 0) bt[v] = std::min(bt[v], t);
                }
            const size_t k = std::min_element(bt.begin(), bt.end()) - bt.begin();
            best.y_nt = vars[k][0];
            best.y_rp = vars[k][1];
        }
        HCHK(hipGetLastError());
        tune32 = best;
        tuned32 = true;
        std::lock_guard<std::mutex> lk(tune_mutex());
        tune_cache()[key] = best;
        tune_file_store();
    }

    // device buffers for n solves and, for an FP32 + polish engine, the
    // autotuner's choice for this shape, ahead of solve(): the batch driver
    // prepares every engine on one thread before the engines run together,
    // so no allocation or timing overlaps another engine's kernels.  The
    // timing runs the passes on the live (not yet loaded) buffers with every
    // solve marked running.
    void prepare(int n)
    {
        nsolve = n;
        ensure_solves(n);
        if (precision == M3D_FP32_POLISH && autotune && (!tuned32 || tuned_ns != n)) {
            std::vector<Ctl> hctl(n);
            for (Ctl& c : hctl) {
                std::memset(&c, 0, sizeof c);
                c.eq_it = -1;
                c.cap = cap;
            }
            HCHK(hipMemcpyAsync(ctl, hctl.data(), sizeof(Ctl) * n, hipMemcpyHostToDevice, st));
            std::vector<int> rv(n, 0);
            HCHK(hipMemcpyAsync(rve_of, rv.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
            tune_xy(g32);
            tuned_ns = n;
            HCHK(hipStreamSynchronize(st));
        }
    }

    // loads: nsolve x 6 (only the first C entries used); tol per phase
    // loads, norms, base flags and RVE indices per solve (sb/rve empty: the
    // handle's single RVE with base flag sb0)
    void solve(const std::vector<std::array<double, 6>>& loads, double tol_f, double tol_eq, const std::vector<double>& norm,
               int sb0, const std::vector<int>& sbv = {}, const std::vector<int>& rve = {})
    {
        nsolve = (int)loads.size();
        ensure_solves(nsolve);
        std::vector<int> rv(nsolve, 0);
        for (int s = 0; s < nsolve && s < (int)rve.size(); ++s) rv[s] = rve[s];
        HCHK(hipMemcpyAsync(rve_of, rv.data(), sizeof(int) * nsolve, hipMemcpyHostToDevice, st));
        std::vector<Ctl> hctl(nsolve);
        for (int s = 0; s < nsolve; ++s) {
            Ctl& c = hctl[s];
            std::memset(&c, 0, sizeof c);
            c.eq_it = -1;
            c.cap = cap;
            c.norm = norm[s];
            c.sb = s < (int)sbv.size() ? sbv[s] : sb0;
            for (int k = 0; k < 6; ++k) c.E[k] = loads[s][k];
            // the FP32 phase stops at polish_factor x tol, but never below what
            // FP32 fields resolve (fp32_floor); the FP64 polish does the rest
            const bool two = (precision == M3D_FP32_POLISH);
            c.tol_f = two ? std::max(tol_f * polish_factor, fp32_floor) : tol_f;
            c.tol_eq = two ? std::max(tol_eq * polish_factor, fp32_floor) : tol_eq;
        }
        HCHK(hipMemcpyAsync(ctl, hctl.data(), sizeof(Ctl) * nsolve, hipMemcpyHostToDevice, st));
        if (precision == M3D_FP32_POLISH && autotune && (!tuned32 || tuned_ns != nsolve)) {
            tune_xy(g32);
            tuned_ns = nsolve;
        }
        final_prec.assign(nsolve, 1);
        info.assign(nsolve, m3d_run_info{});
        HCHK(hipEventRecord(ev0, st));
        Geo gf = g32, gd = g64;
        gf.nsolve = gd.nsolve = nsolve;
        if (precision == M3D_FP32_POLISH) {
            Ptr<float> pf = ptrs<float>(gf);
            hipLaunchKernelGGL((k_fill_load<float, C>), dim3(1024), dim3(256), 0, st, gf, pf, 0);
            run_phase<float>(gf, hctl);
            // roll back `rollback` iterations and redo them in FP64 under the
            // exact rule: the FP32 rule fires at or before the reference's stop
            // iteration, so every candidate stop iterate is an FP64 iterate
            std::vector<int> it32(nsolve), jr(nsolve, 0);
            for (int s = 0; s < nsolve; ++s) {
                it32[s] = hctl[s].it;
                info[s].iters_fp32 = hctl[s].it;
                // borderline switch iterate: both criteria already within
                // rollback_margin x tol in FP32 -> the reference may stop
                // right here, so the last iterations are redone in FP64.
                // Iteration cap reached in FP32: the reference returns its
                // FP64 iterate at the cap (FFT_analysis3D.cpp:785-792,
                // FFT_thermal3D.cpp:474-480), so the last iterations are
                // redone in FP64 too, up to the cap or to the exact stop rule
                // if the FP32 phase only stalled on its floored tolerance
                const bool capped = hctl[s].done != 1;
                const bool border = hctl[s].dsum < rollback_margin * tol_f && hctl[s].deq < rollback_margin * tol_eq;
                const int it = hctl[s].it;
                if (capped) {
                    jr[s] = std::max(0, it - rollback);
                } else if (border) {
                    // the reference rule needs Delta-sum < tol at its stop iteration,
                    // and the FP32 Delta-sum differs from the FP64 one far less than
                    // rollback_margin: so it can only stop within the trailing run of
                    // iterations whose FP32 Delta-sum is below rollback_margin x tol.
                    // Redo that run (at least the switch iterate), at most `rollback`
                    int k = it;
                    if (rollback <= 15) {
                        while (k - 1 > std::max(1, it - rollback) &&
                               hctl[s].dsum_hist[(k - 1) & 15] < rollback_margin * tol_f)
                            --k;
                        jr[s] = std::max(0, k - 1);
                    } else {
                        jr[s] = std::max(0, it - rollback);
                    }
                } else {
                    jr[s] = it;
                }
                hctl[s].it = jr[s];
                hctl[s].done = 0;
                hctl[s].need_cur = hctl[s].need_prev = 0;
                hctl[s].eq_it = -1;
                hctl[s].tol_f = tol_f;
                hctl[s].tol_eq = tol_eq;
            }
            HCHK(hipMemcpyAsync(ctl, hctl.data(), sizeof(Ctl) * nsolve, hipMemcpyHostToDevice, st));
            // promote the FP32 iterates to FP64: inside the polish's first z pass
            // on the specialised kernels, else by k_promote
            const bool fused_promote = refine && gd.fz;
            if (!fused_promote)
                hipLaunchKernelGGL((k_promote<C>), dim3(1024), dim3(256), 0, st, gd, ptrs<float>(gf), ptrs<double>(gd), ctl);
            bool any = false;
            for (int s = 0; s < nsolve; ++s) any = any || (hctl[s].done == 0);
            if (any && refine) {
                // one FP64 iteration from the switch iterate A gives r = T(A) - A;
                // the rest of the polish iterates the FP32 correction e <- r + M e
                // (the load term drops out of M) with every stop-rule quantity
                // taken in FP64 from A + e.  The corrections transform e only,
                // so that iteration stops after its update and stop check.
                Ptr<double> pd = ptrs<double>(gd);
                launch_start<double>(gd, pd, fused_promote);
                launch_iter<double>(gd, pd, false);
                HCHK(hipGetLastError());
                HCHK(hipMemcpyAsync(hctl.data(), ctl, sizeof(Ctl) * nsolve, hipMemcpyDeviceToHost, st));
                HCHK(hipStreamSynchronize(st));
                bool left = false;
                for (int s = 0; s < nsolve; ++s) {
                    if (hctl[s].done) continue;
                    hctl[s].corr = 1;
                    hctl[s].anc_it = hctl[s].it - 1;
                    for (int k = 0; k < 6; ++k) hctl[s].E[k] = 0.0;
                    left = true;
                }
                if (left) {
                    HCHK(hipMemcpyAsync(ctl, hctl.data(), sizeof(Ctl) * nsolve, hipMemcpyHostToDevice, st));
                    run_corr(gf, gd, hctl);
                }
            } else if (any) {
                run_phase<double>(gd, hctl);
            }
            for (int s = 0; s < nsolve; ++s) info[s].iters_fp64 = hctl[s].it - jr[s];
        } else {
            Ptr<double> pd = ptrs<double>(gd);
            hipLaunchKernelGGL((k_fill_load<double, C>), dim3(1024), dim3(256), 0, st, gd, pd, 0);
            run_phase<double>(gd, hctl);
            for (int s = 0; s < nsolve; ++s) info[s].iters_fp64 = hctl[s].it;
        }
        HCHK(hipEventRecord(ev1, st));
        HCHK(hipStreamSynchronize(st));
        float ms = 0;
        HCHK(hipEventElapsedTime(&ms, ev0, ev1));
        for (int s = 0; s < nsolve; ++s) {
            info[s].iterations = hctl[s].it;
            info[s].status = hctl[s].done;
            info[s].eqVar = hctl[s].eqvar;
            info[s].diff_sum = hctl[s].dsum;
            info[s].diff_eq = hctl[s].deq;
            info[s].time_ms = ms;
        }
    }

    // means of u and v per solve: out[s][0..C) = <u>, out[s][C..2C) = <v>
    std::vector<double> means()
    {
        std::vector<double> out((size_t)nsolve * 2 * C, 0.0);
        std::vector<double> part((size_t)nsolve * 2 * C * mnblk);
        for (int pr = 0; pr < 2; ++pr) {
            bool any = false;
            for (int s = 0; s < nsolve; ++s) any = any || (final_prec[s] == pr);
            if (!any) continue;
            if (pr == 0) {
                Geo g = g32;
                g.nsolve = nsolve;
                hipLaunchKernelGGL((k_means<float, PH>), dim3(mnblk, nsolve), dim3(256), 0, st, g, ptrs<float>(g), mpart, mnblk);
            } else {
                Geo g = g64;
                g.nsolve = nsolve;
                hipLaunchKernelGGL((k_means<double, PH>), dim3(mnblk, nsolve), dim3(256), 0, st, g, ptrs<double>(g), mpart, mnblk);
            }
            HCHK(hipGetLastError());
            HCHK(hipMemcpyAsync(part.data(), mpart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
            HCHK(hipStreamSynchronize(st));
            for (int s = 0; s < nsolve; ++s) {
                if (final_prec[s] != pr) continue;
                for (int q = 0; q < 2 * C; ++q) {
                    double a = 0;
                    for (int b = 0; b < mnblk; ++b) a += part[((size_t)s * 2 * C + q) * mnblk + b];
                    out[(size_t)s * 2 * C + q] = a / (double)N;
                }
            }
        }
        return out;
    }

    void fields(int s, double* uo, double* vo)
    {
        if (s < 0 || s >= nsolve) throw ArgError{M3D_ERR_ARG, "load case index out of range"};
        if (!fbuf) alloc(&fbuf, sizeof(double) * 2 * C * (size_t)N);
        double* du = fbuf;
        double* dv = fbuf + (size_t)C * N;
        if (final_prec[s] == 0) {
            Geo g = g32;
            g.nsolve = nsolve;
            hipLaunchKernelGGL((k_fields<float, PH>), dim3(2048), dim3(256), 0, st, g, ptrs<float>(g), s, du, dv);
        } else {
            Geo g = g64;
            g.nsolve = nsolve;
            hipLaunchKernelGGL((k_fields<double, PH>), dim3(2048), dim3(256), 0, st, g, ptrs<double>(g), s, du, dv);
        }
        HCHK(hipGetLastError());
        if (uo) HCHK(hipMemcpyAsync(uo, du, sizeof(double) * C * (size_t)N, hipMemcpyDeviceToHost, st));
        if (vo) HCHK(hipMemcpyAsync(vo, dv, sizeof(double) * C * (size_t)N, hipMemcpyDeviceToHost, st));
        HCHK(hipStreamSynchronize(st));
    }

    // Average duration (ms) of standalone launches of the per-iteration
    // kernels over the current batch: {zpass(update), ypass fwd, xgreen,
    // ypass inv, zpass(check), xcheck}.  Uses the handle's stream; the field
    // state is overwritten (benchmark use only).
    template <typename T>
    void profile_t(int reps, double* ms)
    {
        Geo g = sizeof(T) == 4 ? g32 : g64;
        g.nsolve = nsolve;
        Ptr<T> p = ptrs<T>(g);
        std::vector<Ctl> hctl(nsolve), save(nsolve);
        HCHK(hipMemcpy(save.data(), ctl, sizeof(Ctl) * nsolve, hipMemcpyDeviceToHost));
        hctl = save;
        for (auto& c : hctl) {
            c.done = 0;
            c.need_cur = 1;
            c.need_prev = 1;
            c.corr = 0;
        }
        HCHK(hipMemcpy(ctl, hctl.data(), sizeof(Ctl) * nsolve, hipMemcpyHostToDevice));
        for (int k = 0; k < 6; ++k) {
            auto launch = [&]() {
                switch (k) {
                case 0: zpass<T, MODE_IUF>(g, p); break;
                case 1: ypass<T, -1, BUF_MAIN>(g, p); break;
                case 2: xgreen<T>(g, p); break;
                case 3: ypass<T, 1, BUF_MAIN>(g, p); break;
                case 4: zpass<T, MODE_CHK_CUR>(g, p); break;
                default: xcheck<T, BUF_CHK_CUR>(g, p); break;
                }
            };
            launch(); // warm
            HCHK(hipEventRecord(ev0, st));
            for (int r = 0; r < reps; ++r) launch();
            HCHK(hipEventRecord(ev1, st));
            HCHK(hipEventSynchronize(ev1));
            float t = 0;
            HCHK(hipEventElapsedTime(&t, ev0, ev1));
            ms[k] = (double)t / reps;
        }
        HCHK(hipGetLastError());
        HCHK(hipMemcpy(ctl, save.data(), sizeof(Ctl) * nsolve, hipMemcpyHostToDevice));
    }

    void profile(int prec, int reps, double* ms)
    {
        if (nsolve < 1) throw ArgError{M3D_ERR_STATE, "profile: run or homogenize first"};
        if (prec == M3D_FP32_POLISH && precision != M3D_FP32_POLISH)
            throw ArgError{M3D_ERR_ARG, "profile: FP32 kernels need an FP32+polish handle"};
        if (prec == M3D_FP32_POLISH) profile_t<float>(reps, ms);
        else profile_t<double>(reps, ms);
    }

    // geometry summary for byte accounting: {Nx,Ny,Nz,Nzh,C,nsolve,TY,TKy,TKx}
    void geometry(int prec, int* out)
    {
        const Geo& g = prec == M3D_FP32_POLISH ? g32 : g64;
        int v[9] = {g.Nx, g.Ny, g.Nz, g.Nzh, g.C, nsolve, g.TY, g.TKy, g.TKx};
        std::memcpy(out, v, sizeof v);
    }

    // replace the coefficient tables and reference media (same layout as
    // upload_material), keeping the phase ids
    void upload_coef(const std::vector<double>& coef, const std::vector<double>& prmv)
    {
        std::vector<float> c32(coef.begin(), coef.end());
        std::vector<float> p32(prmv.begin(), prmv.end());
        HCHK(hipMemcpy(coef32, c32.data(), sizeof(float) * c32.size(), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(coef64, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(prm32, p32.data(), sizeof(float) * p32.size(), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(prm64, prmv.data(), sizeof(double) * prmv.size(), hipMemcpyHostToDevice));
    }

    // upload phase ids (validated, < 256 phases) and coefficient tables
    // ids: N phase ids per RVE; coef: 256*NCOEF per RVE; prmv: 2 per RVE
    void upload_material(const std::vector<uint8_t>& ids, const std::vector<double>& coef, const std::vector<double>& prmv)
    {
        if (ids.size() > ph_cap) {
            if (ph) (void)hipFree(ph);
            ph = nullptr;
            ph_cap = 0;
            alloc(&ph, ids.size());
            ph_cap = ids.size();
        }
        HCHK(hipMemcpy(ph, ids.data(), ids.size(), hipMemcpyHostToDevice));
        std::vector<float> c32(coef.begin(), coef.end());
        std::vector<float> p32(prmv.begin(), prmv.end());
        if (coef32) hipFree(coef32);
        if (coef64) hipFree(coef64);
        if (prm32) hipFree(prm32);
        if (prm64) hipFree(prm64);
        coef32 = nullptr; coef64 = nullptr; prm32 = nullptr; prm64 = nullptr;
        alloc(&coef32, sizeof(float) * c32.size());
        alloc(&coef64, sizeof(double) * coef.size());
        alloc(&prm32, sizeof(float) * p32.size());
        alloc(&prm64, sizeof(double) * prmv.size());
        HCHK(hipMemcpy(coef32, c32.data(), sizeof(float) * c32.size(), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(coef64, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(prm32, p32.data(), sizeof(float) * p32.size(), hipMemcpyHostToDevice));
        HCHK(hipMemcpy(prm64, prmv.data(), sizeof(double) * prmv.size(), hipMemcpyHostToDevice));
    }
};

} // namespace

