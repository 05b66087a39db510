// m3d_2d.hip -- FFT_Analysis2D (src/2d/FFT_analysis2D.cpp) on the GPU, FP64:
// the basic scheme on complex 2D fields [E11, E22, E12] (tensor shear) with
// a fixed 25 iterations per runAnalysis.
//
// Per iteration (four launches):
//   polar : eps = t / xy of the previous iteration (FFTBCKWRD_epsilon
//           :143-148); sigma = C(phase) eps (multipCx, :96-106); tau =
//           sigma - C0 eps (:240)
//   y     : forward FFT along y of the three interleaved components (p3)
//   x     : forward FFT along x, then t = -G(xi) t with the reference's
//           6-entry Green operator per frequency (setGF :53-93, recomputed
//           here; x > (Nx-1)/2 maps to -(Ny - x) as in the reference), DC =
//           E xy (multiplyGF :114-141), then the backward FFT along x
//   y     : backward FFT along y (p4)
// convergenceCheck (:150-187) only prints, and its sum reads past the two-entry
// frequency vector; its iteration counter is kept, its transform is not.
// The 1D transforms are the y/x passes of m3d_fast.hpp on a three-component
// view: y lines are rows of one x plane (3 values apart), x lines are rows
// Ny*3 values apart with the y index in the role of the contiguous kz tile.
#include "m3d_engine.hpp"

#include <cstring>
#include <string>
#include <vector>

namespace m3d {
namespace p2 {

struct Comp3 {
    static constexpr int C = 3;
    // the x pass with the Green step on the view of an Nx x Ny grid as
    // g.Nx x 1 x g.Nzh (kx: x frequency, kz: y frequency); prm = {la0, mu0},
    // ctl.E = the load as (re, im) pairs
    template <typename T>
    static __device__ void green(const Geo& g, const Ctl& ctl, const T* prm, int kx, int, int kz, cx<T>* v);
};

constexpr int NT = 256;

struct Mat {
    double Cm[9], Ci[9], C0[9];
    double la0, mu0;
};

// from_t: the previous iteration's strain is still the unscaled inverse
// transform in t; eps = t / xy is stored first (the scaling step of the
// previous iteration, done here to save a launch)
__global__ void k_polar(const uint8_t* ph, Mat m, cx<double>* eps, cx<double>* sig, cx<double>* t, long long xy,
                        int from_t)
{
#pragma clang fp contract(off)
    const long long s0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, st = (long long)gridDim.x * blockDim.x;
    for (long long i = s0; i < xy; i += st) {
        const double* C = ph[i] ? m.Cm : m.Ci;
        if (from_t) {
            const double n = (double)xy;
            for (int c = 0; c < 3; ++c) eps[3 * i + c] = cx<double>{t[3 * i + c].re / n, t[3 * i + c].im / n};
        }
        const cx<double> e0 = eps[3 * i], e1 = eps[3 * i + 1], e2 = eps[3 * i + 2];
        for (int r = 0; r < 3; ++r) {
            const cx<double> sg{(C[3 * r] * e0.re + C[3 * r + 1] * e1.re) + C[3 * r + 2] * e2.re,
                                (C[3 * r] * e0.im + C[3 * r + 1] * e1.im) + C[3 * r + 2] * e2.im};
            const cx<double> c0{(m.C0[3 * r] * e0.re + m.C0[3 * r + 1] * e1.re) + m.C0[3 * r + 2] * e2.re,
                                (m.C0[3 * r] * e0.im + m.C0[3 * r + 1] * e1.im) + m.C0[3 * r + 2] * e2.im};
            sig[3 * i + r] = sg;
            t[3 * i + r] = cx<double>{sg.re - c0.re, sg.im - c0.im};
        }
    }
}

__device__ __forceinline__ double dl(int i, int j) { return i == j ? 1.0 : 0.0; }

// the Green step at frequency (x, y) of the Nx x Ny grid on its three values
// v (multiplyGF, :114-141): DC = E xy, else v = -G(xi) v with the shear
// weighted twice; E as (re, im) pairs
__device__ __forceinline__ void green2d(int x, int y, int Nx, int Ny, double la0, double mu0, const double* E,
                                        cx<double>* v)
{
#pragma clang fp contract(off)
    if (x == 0 && y == 0) {
        const double n = (double)((long long)Nx * Ny);
        for (int c = 0; c < 3; ++c) v[c] = cx<double>{E[2 * c] * n, E[2 * c + 1] * n};
        return;
    }
    const double d1 = 1 / (4 * mu0);
    const double d2 = (la0 + mu0) / (mu0 * (la0 + 2 * mu0));
    const int q[6][4] = {{0, 0, 0, 0}, {0, 0, 1, 1}, {0, 0, 0, 1}, {1, 1, 1, 1}, {1, 1, 0, 1}, {0, 1, 0, 1}};
    const int w[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    const int x1 = (x <= (Nx - 1) / 2) ? x : -(Ny - x);
    const int y1 = (y <= (Ny - 1) / 2) ? y : -(Ny - y);
    const double zet[2] = {x1 * 1.0, y1 * 1.0};
    const double n2 = zet[0] * zet[0] + zet[1] * zet[1];
    double G[6];
    for (int r = 0; r < 6; ++r) {
        const int a = q[r][0], b = q[r][1], k = q[r][2], h = q[r][3];
        G[r] = (d1 / n2) * (((dl(k, a) * zet[h] * zet[b] + dl(h, a) * zet[k] * zet[b]) + dl(k, b) * zet[h] * zet[a]) +
                            dl(h, b) * zet[k] * zet[a]) -
               (d2 * zet[a] * zet[b] * zet[k] * zet[h]) / (n2 * n2);
    }
    const cx<double> a0 = v[0], a1 = v[1], a2 = v[2];
    for (int j = 0; j < 3; ++j) {
        const double re = -((G[w[j][0]] * a0.re + G[w[j][1]] * a1.re) + G[w[j][2]] * a2.re * 2);
        const double im = -((G[w[j][0]] * a0.im + G[w[j][1]] * a1.im) + G[w[j][2]] * a2.im * 2);
        v[j] = cx<double>{re, im};
    }
}

template <typename T>
__device__ void Comp3::green(const Geo& g, const Ctl& ctl, const T* prm, int kx, int, int kz, cx<T>* v)
{
    green2d(kx, kz, g.Nx, g.Nzh, prm[0], prm[1], ctl.E, v);
}

static_assert(fast::has_green_hook<Comp3>::value, "the 2D Green step must replace the 3D one in the x pass");

__global__ void k_scale(const cx<double>* t, cx<double>* eps, long long n3, double xy)
{
    const long long s0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, st = (long long)gridDim.x * blockDim.x;
    for (long long k = s0; k < n3; k += st) eps[k] = cx<double>{t[k].re / xy, t[k].im / xy};
}

inline bool fast_len(int L)
{
    return L == 32 || L == 48 || L == 64 || L == 96 || L == 128 || L == 192 || L == 256 || L == 512;
}

struct Solver {
    int Nx, Ny;
    long long xy;
    hipStream_t st = nullptr;
    cx<double>*eps = nullptr, *sig = nullptr, *t = nullptr;
    uint8_t* ph = nullptr;
    Ctl* ctl = nullptr; // never done: gates nothing in the passes; E holds the load
    cx<double>*twx = nullptr, *twy = nullptr;
    double* dprm = nullptr; // {la0, mu0} for the Green step
    int* drve = nullptr;    // {0}: the one solve's parameter row
    Geo gy{}, gx1{};
    Mat m{};
    cx<double> E[3] = {{0, 0}, {0, 0}, {0, 0}};
    bool have_mat = false;
    int itNo = 0;
    int nb = 1024;

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

    Solver(int nx, int ny) : Nx(nx), Ny(ny)
    {
        xy = (long long)nx * ny;
        if (!fast_len(nx) || !fast_len(ny))
            throw ArgError{M3D_ERR_UNSUPPORTED, "FFT_Analysis2D: Nx and Ny must be 32, 48, 64, 96, 128, 192, 256 or 512"};
        HCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        alloc(&eps, sizeof(cx<double>) * 3 * xy);
        alloc(&sig, sizeof(cx<double>) * 3 * xy);
        alloc(&t, sizeof(cx<double>) * 3 * xy);
        alloc(&ph, (size_t)xy);
        alloc(&ctl, sizeof(Ctl));
        alloc(&dprm, sizeof(double) * 2);
        alloc(&drve, sizeof(int));
        HCHK(hipMemset(drve, 0, sizeof(int)));
        HCHK(hipMemset(eps, 0, sizeof(cx<double>) * 3 * xy));
        HCHK(hipMemset(sig, 0, sizeof(cx<double>) * 3 * xy));
        HCHK(hipMemset(ctl, 0, sizeof(Ctl)));
        auto tx = twiddles<double>(nx), ty = twiddles<double>(ny);
        alloc(&twx, sizeof(cx<double>) * nx);
        alloc(&twy, sizeof(cx<double>) * ny);
        HCHK(hipMemcpy(twx, tx.data(), sizeof(cx<double>) * nx, hipMemcpyHostToDevice));
        HCHK(hipMemcpy(twy, ty.data(), sizeof(cx<double>) * ny, hipMemcpyHostToDevice));
        // y lines: one tile per x plane (rows y at stride 3, one "kz" column of
        // 3 components); x lines: rows x at stride Ny*3, the Ny points of a row
        // in the role of the kz tile
        std::memset(&gy, 0, sizeof gy);
        gy.Nx = nx; gy.Ny = ny; gy.Nz = 1; gy.Nzh = 1; gy.C = 3; gy.nsolve = 1;
        gy.sy = 3; gy.sxa = (long long)ny * 3; gy.sxb = 0; gy.lxb = 0; gy.xbm = 0;
        // x lines with the Green step: the grid as nx x 1 x ny, rows x at
        // stride ny*3, the y points in the role of the kz tile
        gx1 = gy;
        gx1.Nx = nx; gx1.Ny = 1; gx1.Nzh = ny;
        gx1.sy = (long long)ny * 3; gx1.sxa = (long long)ny * 3; gx1.sxb = 0; gx1.lxb = 0; gx1.xbm = 0;
    }

    ~Solver()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : {(void*)eps, (void*)sig, (void*)t, (void*)ph, (void*)ctl, (void*)dprm, (void*)drve, (void*)twx,
                        (void*)twy})
            if (p) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }

    template <int L, int TK, int DIR>
    bool pass_at(const Geo& g, const cx<double>* tw)
    {
        if (g.Ny != L) return false;
        const int nkt = (g.Nzh + TK - 1) / TK;
        const size_t sm = ((size_t)L * TK * 3 + L) * sizeof(cx<double>);
        Ptr<double> p;
        std::memset(&p, 0, sizeof p);
        p.S = t;
        p.K = t;
        p.twy = tw;
        p.ctl = ctl;
        hipLaunchKernelGGL((fast::kxy<double, Comp3, 0, DIR, BUF_MAIN, L, TK, NT>), dim3(g.Nx * nkt), dim3(NT), sm, st, g,
                           p, (const double*)nullptr);
        return true;
    }
    // y pass: one kz column (TK = 1); x pass: kz tiles of the library's width
    template <int DIR>
    void ypass()
    {
        if (!(pass_at<512, 1, DIR>(gy, twy) || pass_at<256, 1, DIR>(gy, twy) || pass_at<192, 1, DIR>(gy, twy) ||
              pass_at<128, 1, DIR>(gy, twy) || pass_at<96, 1, DIR>(gy, twy) || pass_at<64, 1, DIR>(gy, twy) ||
              pass_at<48, 1, DIR>(gy, twy) || pass_at<32, 1, DIR>(gy, twy)))
            throw ArgError{M3D_ERR_UNSUPPORTED, "FFT_Analysis2D: no y kernel"};
    }
    // x lines: forward FFT, the Green step (Comp3::green), inverse FFT in one
    // pass over each tile
    template <int L>
    bool xg_at()
    {
        if (gx1.Nx != L) return false;
        constexpr int TK = fast::xtile<double>(L, 3);
        const int nkt = (gx1.Nzh + TK - 1) / TK;
        const size_t sm = ((size_t)L * TK * 3 + L) * sizeof(cx<double>);
        if (sm > 65536) throw ArgError{M3D_ERR_UNSUPPORTED, "FFT_Analysis2D: x tile too large"};
        Ptr<double> p;
        std::memset(&p, 0, sizeof p);
        p.S = t;
        p.K = t;
        p.twx = twx;
        p.ctl = ctl;
        p.rve_of = drve;
        hipLaunchKernelGGL((fast::kxy<double, Comp3, 1, -1, BUF_MAIN, L, TK>), dim3(nkt), dim3(NT), sm, st, gx1, p,
                           (const double*)dprm);
        return true;
    }
    void xgreen()
    {
        if (!(xg_at<512>() || xg_at<256>() || xg_at<192>() || xg_at<128>() || xg_at<96>() || xg_at<64>() ||
              xg_at<48>() || xg_at<32>()))
            throw ArgError{M3D_ERR_UNSUPPORTED, "FFT_Analysis2D: no x kernel"};
    }

    // material(properties, RVE1), :189-221
    void material(const double* q, const double* rve)
    {
        const double Em = q[0], vm = q[1], Ei = q[2], vi = q[3];
        const double mum = Em / (1 + vm) / 2, lam = vm * Em / (1 + vm) / (1 - 2 * vm);
        const double mui = Ei / (1 + vi) / 2, lai = vi * Ei / (1 + vi) / (1 - 2 * vi);
        const double cm[9] = {lam + 2 * mum, lam, 0, lam, lam + 2 * mum, 0, 0, 0, 2 * mum};
        const double ci[9] = {lai + 2 * mui, lai, 0, lai, lai + 2 * mui, 0, 0, 0, 2 * mui};
        const double alf = 1;
        for (int k = 0; k < 9; ++k) {
            m.Cm[k] = cm[k];
            m.Ci[k] = ci[k];
            m.C0[k] = alf * (cm[k] + ci[k]) / 2.;
        }
        m.la0 = alf * (lai + lam) / 2.;
        m.mu0 = alf * (mui + mum) / 2.;
        const double lm[2] = {m.la0, m.mu0};
        HCHK(hipMemcpy(dprm, lm, sizeof lm, hipMemcpyHostToDevice));
        std::vector<uint8_t> p((size_t)xy);
        for (long long i = 0; i < xy; ++i) p[(size_t)i] = rve[i] == 1 ? 1 : 0; // RVE(i)==1: matrix (:99-103)
        HCHK(hipMemcpy(ph, p.data(), (size_t)xy, hipMemcpyHostToDevice));
        have_mat = true;
    }

    // runAnalysis, :232-246
    void run()
    {
        if (!have_mat) throw ArgError{M3D_ERR_STATE, "runAnalysis: material() first"};
        double e_ri[6];
        for (int c = 0; c < 3; ++c) {
            e_ri[2 * c] = E[c].re;
            e_ri[2 * c + 1] = E[c].im;
        }
        HCHK(hipMemcpy(reinterpret_cast<char*>(ctl) + offsetof(Ctl, E), e_ri, sizeof e_ri, hipMemcpyHostToDevice));
        for (int it = 0; it < 25; ++it) {
            hipLaunchKernelGGL(k_polar, dim3(nb), dim3(NT), 0, st, ph, m, eps, sig, t, xy, it > 0 ? 1 : 0);
            ypass<-1>();
            xgreen();
            ypass<1>();
            itNo += 1;
        }
        hipLaunchKernelGGL(k_scale, dim3(nb), dim3(NT), 0, st, t, eps, 3 * xy, (double)xy);
        HCHK(hipGetLastError());
        HCHK(hipStreamSynchronize(st));
    }
};

} // namespace p2
} // namespace m3d

using namespace m3d;

struct m3d_2d {
    p2::Solver* s = nullptr;
    std::string err;
};

namespace {

template <typename F>
int guarded(m3d_2d* h, F&& f)
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
    }
}

// 3x3 column-major inverse by Gauss-Jordan with partial pivoting (Eigen's inverse())
bool inv3(const double* A, double* X)
{
    double M[9];
    std::memcpy(M, A, sizeof M);
    for (int k = 0; k < 9; ++k) X[k] = (k % 4 == 0) ? 1.0 : 0.0;
    for (int c = 0; c < 3; ++c) {
        int piv = c;
        for (int r = c + 1; r < 3; ++r)
            if (std::fabs(M[r + 3 * c]) > std::fabs(M[piv + 3 * c])) piv = r;
        if (M[piv + 3 * c] == 0) return false;
        if (piv != c)
            for (int k = 0; k < 3; ++k) {
                std::swap(M[c + 3 * k], M[piv + 3 * k]);
                std::swap(X[c + 3 * k], X[piv + 3 * k]);
            }
        const double d = M[c + 3 * c];
        for (int k = 0; k < 3; ++k) {
            M[c + 3 * k] /= d;
            X[c + 3 * k] /= d;
        }
        for (int r = 0; r < 3; ++r) {
            if (r == c) continue;
            const double f = M[r + 3 * c];
            for (int k = 0; k < 3; ++k) {
                M[r + 3 * k] -= f * M[c + 3 * k];
                X[r + 3 * k] -= f * X[c + 3 * k];
            }
        }
    }
    return true;
}

} // namespace

int m3d_2d_create(int nx, int ny, m3d_2d** out)
{
    if (!out || nx < 1 || ny < 1) return M3D_ERR_ARG;
    *out = nullptr;
    m3d_2d* h = new (std::nothrow) m3d_2d;
    if (!h) return M3D_ERR_OOM;
    const int rc = guarded(h, [&]() {
        h->s = new p2::Solver(nx, ny);
        return M3D_OK;
    });
    if (rc != M3D_OK) {
        delete h->s;
        delete h;
        return rc;
    }
    *out = h;
    return M3D_OK;
}

int m3d_2d_destroy(m3d_2d* h)
{
    if (!h) return M3D_ERR_ARG;
    delete h->s;
    delete h;
    return M3D_OK;
}

const char* m3d_2d_last_error(m3d_2d* h) { return h ? h->err.c_str() : "null handle"; }

int m3d_2d_material(m3d_2d* h, const double* properties, const double* rve)
{
    return guarded(h, [&]() {
        if (!properties || !rve) throw ArgError{M3D_ERR_ARG, "material: null argument"};
        h->s->material(properties, rve);
        return M3D_OK;
    });
}

int m3d_2d_set_load(m3d_2d* h, const double* E_ri)
{
    return guarded(h, [&]() {
        if (!E_ri) throw ArgError{M3D_ERR_ARG, "setLoad: null argument"};
        p2::Solver& s = *h->s;
        for (int c = 0; c < 3; ++c) s.E[c] = cx<double>{E_ri[2 * c], E_ri[2 * c + 1]};
        std::vector<cx<double>> e((size_t)(3 * s.xy));
        for (long long i = 0; i < s.xy; ++i)
            for (int c = 0; c < 3; ++c) e[(size_t)(3 * i + c)] = s.E[c];
        HCHK(hipMemcpy(s.eps, e.data(), sizeof(cx<double>) * e.size(), hipMemcpyHostToDevice));
        return M3D_OK;
    });
}

int m3d_2d_run(m3d_2d* h, int* itNo)
{
    return guarded(h, [&]() {
        h->s->run();
        if (itNo) *itNo = h->s->itNo;
        return M3D_OK;
    });
}

int m3d_2d_homogenization(m3d_2d* h, double* CHomo, double* epHomo, double* sigHomo)
{
    return guarded(h, [&]() {
        p2::Solver& s = *h->s;
        double ep[9], sg[9];
        std::vector<cx<double>> e((size_t)(3 * s.xy)), g((size_t)(3 * s.xy));
        for (int w = 0; w < 3; ++w) {
            for (int c = 0; c < 3; ++c) s.E[c] = cx<double>{c == w ? 0.005 : 0.0, 0.0};
            s.run();
            HCHK(hipMemcpy(e.data(), s.eps, sizeof(cx<double>) * e.size(), hipMemcpyDeviceToHost));
            HCHK(hipMemcpy(g.data(), s.sig, sizeof(cx<double>) * g.size(), hipMemcpyDeviceToHost));
            double se[3] = {0, 0, 0}, ss[3] = {0, 0, 0};
            for (long long i = 0; i < s.xy; ++i)
                for (int c = 0; c < 3; ++c) {
                    se[c] += e[(size_t)(3 * i + c)].re;
                    ss[c] += g[(size_t)(3 * i + c)].re;
                }
            for (int c = 0; c < 3; ++c) {
                ep[w + 3 * c] = se[c] / (double)s.xy;
                sg[c + 3 * w] = ss[c] / (double)s.xy;
            }
        }
        double inv[9];
        if (!inv3(ep, inv)) throw ArgError{M3D_ERR_NUMERIC, "homogenization: singular epHomo"};
        double C[9];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                double a = 0;
                for (int k = 0; k < 3; ++k) a += inv[r + 3 * k] * sg[k + 3 * c];
                C[r + 3 * c] = a;
            }
        if (CHomo) std::memcpy(CHomo, C, sizeof C);
        if (epHomo) std::memcpy(epHomo, ep, sizeof ep);
        if (sigHomo) std::memcpy(sigHomo, sg, sizeof sg);
        return M3D_OK;
    });
}

int m3d_2d_get_fields(m3d_2d* h, double* eps_ri, double* sig_ri)
{
    return guarded(h, [&]() {
        p2::Solver& s = *h->s;
        if (eps_ri) HCHK(hipMemcpy(eps_ri, s.eps, sizeof(cx<double>) * 3 * s.xy, hipMemcpyDeviceToHost));
        if (sig_ri) HCHK(hipMemcpy(sig_ri, s.sig, sizeof(cx<double>) * 3 * s.xy, hipMemcpyDeviceToHost));
        return M3D_OK;
    });
}
