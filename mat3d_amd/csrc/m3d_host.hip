// m3d_host.hip -- host runtime and C-ABI of libmat3d.so (gfx950).
//
// Owns the device state of one solver handle, builds the FFT plans and tile
// shapes, drives the iteration as captured hipGraphs of a chunk of
// iterations and assembles the homogenized tensors in FP64 on the host.
// Control decisions (stop rule, equilibrium-check scheduling) are taken on
// the device; the host only polls the per-solve done flags between graph
// replays.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/m3d.h"
#include "m3d_kernels.hpp"

using namespace m3d;

namespace {

constexpr size_t LDS_MAX = 65536;

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
}

Geo make_geo(int Nx, int Ny, int Nz, int C, size_t csz)
{
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
template <typename PH>
struct Engine {
    static constexpr int C = PH::C;
    int Nx, Ny, Nz;
    int precision;
    Geo g32{}, g64{};
    int cap_solves = 0;
    long long N = 0, H = 0;
    // device buffers
    float* U32[2] = {nullptr, nullptr};
    double* U64[2] = {nullptr, nullptr};
    unsigned char* Sbuf = nullptr;
    unsigned char* Kbuf = nullptr;
    double* psum = nullptr;
    double* pchk = nullptr;
    Ctl* ctl = nullptr;
    double* mpart = nullptr;
    int* rve_of = nullptr;
    uint8_t* ph = nullptr;
    float* coef32 = nullptr;
    double* coef64 = nullptr;
    float* prm32 = nullptr;
    double* prm64 = nullptr;
    cx<float>* tw32[3] = {nullptr, nullptr, nullptr};
    cx<double>* tw64[3] = {nullptr, nullptr, nullptr};
    double* fbuf = nullptr; // field output staging (2*C*N doubles)
    hipStream_t st = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // host state
    int nsolve = 0;
    std::vector<int> final_prec; // per solve: 0 fp32, 1 fp64
    std::vector<m3d_run_info> info;
    double polish_factor = 10.0;
    int cap = 2000;
    int chunk = 0;
    int mnblk = 256;

    Engine(int nx, int ny, int nz, int prec) : Nx(nx), Ny(ny), Nz(nz), precision(prec)
    {
        g32 = make_geo(nx, ny, nz, C, sizeof(cx<float>));
        g64 = make_geo(nx, ny, nz, C, sizeof(cx<double>));
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
    }

    void free_solves()
    {
        for (int b = 0; b < 2; ++b) {
            if (U32[b]) hipFree(U32[b]);
            if (U64[b]) hipFree(U64[b]);
            U32[b] = nullptr;
            U64[b] = nullptr;
        }
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

    void ensure_solves(int n)
    {
        if (n <= cap_solves) return;
        free_solves();
        const size_t nc = (size_t)N * C;
        const size_t hc = (size_t)H * C;
        if (precision == M3D_FP32_POLISH) {
            alloc(&U32[0], sizeof(float) * nc * n);
            alloc(&U32[1], sizeof(float) * nc * n);
        }
        alloc(&U64[0], sizeof(double) * nc * n);
        alloc(&U64[1], sizeof(double) * nc * n);
        alloc(&Sbuf, sizeof(cx<double>) * hc * n);
        alloc(&Kbuf, sizeof(cx<double>) * hc * n);
        const int nbz = std::max(g32.nbz, g64.nbz), nbx = std::max(g32.nbx, g64.nbx);
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
            p.U0 = U32[0];
            p.U1 = U32[1];
            p.coef = coef32;
            p.twz = tw32[0];
            p.twy = tw32[1];
            p.twx = tw32[2];
        } else {
            p.U0 = U64[0];
            p.U1 = U64[1];
            p.coef = coef64;
            p.twz = tw64[0];
            p.twy = tw64[1];
            p.twx = tw64[2];
        }
        p.sU = (long long)nc;
        p.S = reinterpret_cast<cx<T>*>(Sbuf);
        p.K = reinterpret_cast<cx<T>*>(Kbuf);
        p.sS = (long long)H * C * (sizeof(cx<double>) / sizeof(cx<T>));
        p.ph = ph;
        p.sPh = N;
        p.rve_of = rve_of;
        p.sCoef = 0;
        p.psum = psum;
        p.pchk = pchk;
        p.ctl = ctl;
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
    template <typename T, bool GEN>
    void launch_start_g(const Geo& g, const Ptr<T>& p)
    {
        const int ns = g.nsolve;
        const size_t zsm = (size_t)g.TY * C * g.lsz * sizeof(cx<T>);
        const size_t ysm = (size_t)g.Ny * g.TKy * C * sizeof(cx<T>);
        const size_t xsm = (size_t)g.Nx * g.TKx * C * sizeof(cx<T>);
        hipLaunchKernelGGL((k_zpass<T, PH, MODE_FWD, GEN>), dim3(ns * g.nbz), dim3(ZTHREADS), zsm, st, g, p);
        hipLaunchKernelGGL(k_fin0, dim3(ns), dim3(FTHREADS), 0, st, g, p.ctl, p.psum);
        hipLaunchKernelGGL((k_ypass<T, C, -1, BUF_MAIN, GEN>), dim3(ns * g.Nx * g.nky), dim3(XYTHREADS), ysm, st, g, p);
        hipLaunchKernelGGL((k_xgreen<T, PH, GEN>), dim3(ns * g.nbx), dim3(XYTHREADS), xsm, st, g, p, prm<T>());
        hipLaunchKernelGGL((k_ypass<T, C, 1, BUF_MAIN, GEN>), dim3(ns * g.Nx * g.nky), dim3(XYTHREADS), ysm, st, g, p);
    }

    template <typename T, bool GEN>
    void launch_iter_g(const Geo& g, const Ptr<T>& p)
    {
        const int ns = g.nsolve;
        const size_t zsm = (size_t)g.TY * C * g.lsz * sizeof(cx<T>);
        const size_t ysm = (size_t)g.Ny * g.TKy * C * sizeof(cx<T>);
        const size_t xsm = (size_t)g.Nx * g.TKx * C * sizeof(cx<T>);
        const dim3 gz(ns * g.nbz), gy(ns * g.Nx * g.nky), gx(ns * g.nbx);
        hipLaunchKernelGGL((k_zpass<T, PH, MODE_IUF, GEN>), gz, dim3(ZTHREADS), zsm, st, g, p);
        hipLaunchKernelGGL(k_finA, dim3(ns), dim3(FTHREADS), 0, st, g, p.ctl, p.psum);
        hipLaunchKernelGGL((k_zpass<T, PH, MODE_CHK_CUR, GEN>), gz, dim3(ZTHREADS), zsm, st, g, p);
        hipLaunchKernelGGL((k_ypass<T, C, -1, BUF_CHK_CUR, GEN>), gy, dim3(XYTHREADS), ysm, st, g, p);
        hipLaunchKernelGGL((k_xcheck<T, PH, BUF_CHK_CUR, GEN>), gx, dim3(XYTHREADS), xsm, st, g, p);
        hipLaunchKernelGGL((k_zpass<T, PH, MODE_CHK_PREV, GEN>), gz, dim3(ZTHREADS), zsm, st, g, p);
        hipLaunchKernelGGL((k_ypass<T, C, -1, BUF_CHK_PREV, GEN>), gy, dim3(XYTHREADS), ysm, st, g, p);
        hipLaunchKernelGGL((k_xcheck<T, PH, BUF_CHK_PREV, GEN>), gx, dim3(XYTHREADS), xsm, st, g, p);
        hipLaunchKernelGGL(k_finB, dim3(ns), dim3(FTHREADS), 0, st, g, p.ctl, p.pchk);
        hipLaunchKernelGGL((k_ypass<T, C, -1, BUF_MAIN, GEN>), gy, dim3(XYTHREADS), ysm, st, g, p);
        hipLaunchKernelGGL((k_xgreen<T, PH, GEN>), gx, dim3(XYTHREADS), xsm, st, g, p, prm<T>());
        hipLaunchKernelGGL((k_ypass<T, C, 1, BUF_MAIN, GEN>), gy, dim3(XYTHREADS), ysm, st, g, p);
    }

    static bool needs_gen(const Geo& g)
    {
        const FftAxis* a[3] = {&g.az, &g.ay, &g.ax};
        for (const FftAxis* x : a)
            for (int i = 0; i < x->np; ++i) {
                const int r = x->radix(i);
                if (r != 2 && r != 3 && r != 4 && r != 8) return true;
            }
        return false;
    }

    template <typename T>
    void launch_start(const Geo& g, const Ptr<T>& p)
    {
        if (needs_gen(g)) launch_start_g<T, true>(g, p);
        else launch_start_g<T, false>(g, p);
    }

    template <typename T>
    void launch_iter(const Geo& g, const Ptr<T>& p)
    {
        if (needs_gen(g)) launch_iter_g<T, true>(g, p);
        else launch_iter_g<T, false>(g, p);
    }

    int chunk_for(const Geo& g) const
    {
        if (chunk > 0) return chunk;
        const double work = (double)g.N * g.nsolve;
        if (work >= 3.0e7) return 2;
        if (work >= 4.0e6) return 4;
        return 8;
    }

    // run the iteration in precision T until every solve of the batch is done
    template <typename T>
    void run_phase(Geo g, std::vector<Ctl>& hctl)
    {
        g.nsolve = nsolve;
        Ptr<T> p = ptrs<T>(g);
        launch_start<T>(g, p);
        HCHK(hipGetLastError());
        const int ck = chunk_for(g);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < ck; ++i) launch_iter<T>(g, p);
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

    // loads: nsolve x 6 (only the first C entries used); tol per phase
    void solve(const std::vector<std::array<double, 6>>& loads, double tol_f, double tol_eq, const std::vector<double>& norm)
    {
        nsolve = (int)loads.size();
        ensure_solves(nsolve);
        std::vector<Ctl> hctl(nsolve);
        for (int s = 0; s < nsolve; ++s) {
            Ctl& c = hctl[s];
            std::memset(&c, 0, sizeof c);
            c.eq_it = -1;
            c.cap = cap;
            c.norm = norm[s];
            for (int k = 0; k < 6; ++k) c.E[k] = loads[s][k];
            const bool two = (precision == M3D_FP32_POLISH);
            c.tol_f = two ? tol_f * polish_factor : tol_f;
            c.tol_eq = two ? tol_eq * polish_factor : tol_eq;
        }
        HCHK(hipMemcpyAsync(ctl, hctl.data(), sizeof(Ctl) * nsolve, hipMemcpyHostToDevice, st));
        final_prec.assign(nsolve, 1);
        info.assign(nsolve, m3d_run_info{});
        HCHK(hipEventRecord(ev0, st));
        Geo gf = g32, gd = g64;
        gf.nsolve = gd.nsolve = nsolve;
        if (precision == M3D_FP32_POLISH) {
            Ptr<float> pf = ptrs<float>(gf);
            hipLaunchKernelGGL((k_fill_load<float, C>), dim3(1024), dim3(256), 0, st, gf, pf, 0);
            run_phase<float>(gf, hctl);
            std::vector<int> it32(nsolve);
            for (int s = 0; s < nsolve; ++s) {
                it32[s] = hctl[s].it;
                info[s].iters_fp32 = hctl[s].it;
                if (hctl[s].done == 1) {
                    hctl[s].done = 0;
                    hctl[s].tol_f = tol_f;
                    hctl[s].tol_eq = tol_eq;
                } else {
                    final_prec[s] = 0; // iteration cap reached in FP32: keep that state
                }
            }
            hipLaunchKernelGGL((k_promote<C>), dim3(1024), dim3(256), 0, st, gd, U32[0], U32[1], (long long)N * C, U64[0],
                               U64[1], (long long)N * C, ctl);
            HCHK(hipMemcpyAsync(ctl, hctl.data(), sizeof(Ctl) * nsolve, hipMemcpyHostToDevice, st));
            bool any = false;
            for (int s = 0; s < nsolve; ++s) any = any || (hctl[s].done == 0);
            if (any) run_phase<double>(gd, hctl);
            for (int s = 0; s < nsolve; ++s) info[s].iters_fp64 = hctl[s].it - it32[s];
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
    template <typename T, bool GEN>
    void profile_g(int reps, double* ms)
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
        }
        HCHK(hipMemcpy(ctl, hctl.data(), sizeof(Ctl) * nsolve, hipMemcpyHostToDevice));
        const int ns = g.nsolve;
        const size_t zsm = (size_t)g.TY * C * g.lsz * sizeof(cx<T>);
        const size_t ysm = (size_t)g.Ny * g.TKy * C * sizeof(cx<T>);
        const size_t xsm = (size_t)g.Nx * g.TKx * C * sizeof(cx<T>);
        const dim3 gz(ns * g.nbz), gy(ns * g.Nx * g.nky), gx(ns * g.nbx);
        for (int k = 0; k < 6; ++k) {
            auto launch = [&]() {
                switch (k) {
                case 0: hipLaunchKernelGGL((k_zpass<T, PH, MODE_IUF, GEN>), gz, dim3(ZTHREADS), zsm, st, g, p); break;
                case 1: hipLaunchKernelGGL((k_ypass<T, C, -1, BUF_MAIN, GEN>), gy, dim3(XYTHREADS), ysm, st, g, p); break;
                case 2: hipLaunchKernelGGL((k_xgreen<T, PH, GEN>), gx, dim3(XYTHREADS), xsm, st, g, p, prm<T>()); break;
                case 3: hipLaunchKernelGGL((k_ypass<T, C, 1, BUF_MAIN, GEN>), gy, dim3(XYTHREADS), ysm, st, g, p); break;
                case 4: hipLaunchKernelGGL((k_zpass<T, PH, MODE_CHK_CUR, GEN>), gz, dim3(ZTHREADS), zsm, st, g, p); break;
                default: hipLaunchKernelGGL((k_xcheck<T, PH, BUF_CHK_CUR, GEN>), gx, dim3(XYTHREADS), xsm, st, g, p); break;
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
        if (prec == M3D_FP32_POLISH) {
            if (needs_gen(g32)) profile_g<float, true>(reps, ms);
            else profile_g<float, false>(reps, ms);
        } else {
            if (needs_gen(g64)) profile_g<double, true>(reps, ms);
            else profile_g<double, false>(reps, ms);
        }
    }

    // geometry summary for byte accounting: {Nx,Ny,Nz,Nzh,C,nsolve,TY,TKy,TKx}
    void geometry(int prec, int* out)
    {
        const Geo& g = prec == M3D_FP32_POLISH ? g32 : g64;
        int v[9] = {g.Nx, g.Ny, g.Nz, g.Nzh, g.C, nsolve, g.TY, g.TKy, g.TKx};
        std::memcpy(out, v, sizeof v);
    }

    // upload phase ids (validated, < 256 phases) and coefficient tables
    void upload_material(const std::vector<uint8_t>& ids, const std::vector<double>& coef, const std::vector<double>& prmv)
    {
        HCHK(hipMemcpy(ph, ids.data(), (size_t)N, hipMemcpyHostToDevice));
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
        else if (k == "cap" && value >= 1) h->eng->cap = (int)value;
        else if (k == "chunk" && value >= 0) h->eng->chunk = (int)value;
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
    e.g32.strain_base = e.g64.strain_base = h->grad_base;
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
        h->eng->solve({ld}, tol_field, tol_eq, {thermal_norm(h, h->E)});
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
        e.solve(loads, tol_field, tol_eq, norms);
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
