// m3d_rec.hip -- phase-retrieval reconstruction of a two-phase microstructure
// from its two-point correlation function: recPlan(double *TPCF, Nx, Ny, Nz)
// and runRec(n, beta) of src/rec/reconstruction.cpp:211-243, 285-406, FP64.
//
// Per iteration (one captured sequence, replayed in chunks; every kernel
// exits once the device-side done flag is set):
//   begin   : the previous iteration's end (below); j += 1; itNo >= n ->
//             done = 2 (":Iteration Number Exceeded"); every 150th iteration
//             selects a re-randomisation stream
//   fwd FFT : MT = FFT3(mT)  (z pass out of place, then y and x passes)
//   [rand]  : X = FFT3(uniform[-1,1) field)   (re-randomisation iterations only)
//   phase   : err partials sum |(|MT|) - normC| (funcNormC); ang = arg(MT), or
//             0.3 arg(X) + 0.7 arg(MT); MH = normC e^{i ang}
//   err     : trackrealerrorA(i) = err / xyz; err < 0.1 -> done = 1
//   inv FFT : IFFT3(MH) (unnormalised; x, y, then z), the z pass storing
//             the projection: mH = MH / xyz; j < 100: the reference's relaxed
//             projection with beta, else clamp to [0, 1]; mT = mH
//   end     : j > 200 -> j = 0 (after a clamp iteration); itNo += 1 --
//             marked by the projection, applied by the next begin
// The 3D complex FFT is z-lines in LDS (contiguous, tiles of TY lines), then
// the y and x passes of m3d_fast.hpp on a one-component view of the grid.
// The final volume-fraction search (:374-402) counts voxels above a moving
// threshold on the device, one launch per step.
#include "m3d_engine.hpp"

#include <cstring>
#include <string>
#include <vector>

namespace m3d {
namespace rec {

// one complex component per grid point
struct Scalar1 {
    static constexpr int C = 1;
};

constexpr int NT = 256;

// counter-based uniform stream in [-1, 1): value i of stream k is
// splitmix64(seed + (k << 40) + i) (published mixing constants), 53 bits
__host__ __device__ inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline double uniform(uint64_t seed, uint64_t stream, uint64_t i)
{
    return 2.0 * ((double)(mix64(seed + (stream << 40) + i) >> 11) * 0x1.0p-53) - 1.0;
}

struct State {
    Ctl ctl;  // ctl.done: 1 converged, 2 iteration number exceeded
    Ctl xctl; // done unless this iteration re-randomises (gates the FFT of X)
    int i, j, n, rr;
    int pend; // the last iteration's end (j reset, i += 1) is still to apply
    double beta;
    uint64_t seed;
};

// z lines of length L, TY lines per block (line l at l*L), LDS [z][line]
// mH = MH / xyz; j < 100: the reference's relaxed projection with beta
// (t: the current mT), else clamp to [0, 1]
__device__ __forceinline__ cx<double> project(cx<double> MH, double t, int j, double beta, double xyz)
{
    cx<double> h{MH.re / xyz, MH.im / xyz};
    if (j < 100) {
        if (h.re < 0) h = cx<double>{t - beta * h.re, 0.0};
        if (h.re > 1) h = cx<double>{t - beta * (h.re - 1), 0.0};
        if (h.re > 1) h = cx<double>{1.0, 0.0};
    } else {
        const double re = h.re;
        if (re < 0) h = cx<double>{0.0, 0.0};
        if (re > 1) h = cx<double>{1.0, 0.0};
    }
    return h;
}

// PROJ: the last pass of the inverse transform, stored through the
// projection (project()) into dst = mT (each element read and rewritten by
// one thread)
template <int L, int TY, int DIR, bool PROJ>
__global__ void __launch_bounds__(NT) kz(const Ctl* ctl, const cx<double>* src, cx<double>* dst, const cx<double>* twz,
                                         State* stt, double xyz)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<double>* sl = reinterpret_cast<cx<double>*>(smem_raw);
    cx<double>* tw = sl + L * TY;
    if (ctl->done) return;
    const long long l0 = (long long)blockIdx.x * TY;
    for (int i = threadIdx.x; i < L; i += NT) tw[i] = twz[i];
    const cx<double>* s = src + l0 * L;
    for (int e = threadIdx.x; e < TY * L; e += NT) {
        const int line = e / L, z = e - line * L;
        sl[z * TY + line] = s[e];
    }
    __syncthreads();
    fast::fft<double, L, TY, NT, DIR, 1>(sl, tw);
    cx<double>* d = dst + l0 * L;
    if constexpr (PROJ) {
        const int j = stt->j;
        const double beta = stt->beta;
        for (int e = threadIdx.x; e < TY * L; e += NT) {
            const int line = e / L, z = e - line * L;
            d[e] = project(sl[z * TY + line], d[e].re, j, beta, xyz);
        }
        // the iteration's end is applied by the next k_begin (no block here
        // reads pend; j was read above)
        if (blockIdx.x == 0 && threadIdx.x == 0) stt->pend = 1;
    } else {
        for (int e = threadIdx.x; e < TY * L; e += NT) {
            const int line = e / L, z = e - line * L;
            d[e] = sl[z * TY + line];
        }
    }
}

// the previous iteration's end (:393-399 of the reference's loop), then
// this iteration's start
__global__ void k_begin(State* st)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (st->pend) {
        st->pend = 0;
        if (st->j >= 100 && st->j > 200) st->j = 0;
        st->i += 1;
    }
    st->xctl.done = 1;
    if (st->ctl.done) return;
    st->j += 1;
    if (st->i >= st->n) {
        st->ctl.done = 2;
        return;
    }
    st->rr = (remainder((double)(st->i + 1), 150.0) == 0) ? (st->i + 1) / 150 : 0;
    st->xctl.done = st->rr ? 0 : 1;
}

__global__ void k_init(State* st, cx<double>* mT, const double* m0, long long N)
{
    const long long s0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, step = (long long)gridDim.x * blockDim.x;
    for (long long s = s0; s < N; s += step) {
        const double a = m0 ? m0[s] : (uniform(st->seed, 0, (uint64_t)s) + 1.0) * 0.5;
        mT[s] = cx<double>{a, 0.0};
    }
}

__global__ void k_rand(const State* st, cx<double>* X, long long N)
{
    if (st->xctl.done) return;
    const long long s0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, step = (long long)gridDim.x * blockDim.x;
    for (long long s = s0; s < N; s += step) X[s] = cx<double>{uniform(st->seed, (uint64_t)st->rr, (uint64_t)s), 0.0};
}

__global__ void __launch_bounds__(NT) k_phase(const State* st, const cx<double>* MT, const cx<double>* X,
                                              const double* normC, cx<double>* MH, double* part, long long N)
{
    __shared__ double red[NT / 64];
    if (st->ctl.done) return;
    const bool rr = st->rr != 0;
    double acc = 0;
    const long long s0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, step = (long long)gridDim.x * blockDim.x;
    for (long long s = s0; s < N; s += step) {
        const cx<double> m = MT[s];
        const double c = normC[s];
        acc += fabs(sqrt(m.re * m.re + m.im * m.im) - c);
        double ang = atan2(m.im, m.re);
        if (rr) ang = atan2(X[s].im, X[s].re) * .3 + ang * .7;
        MH[s] = cx<double>{c * cos(ang), c * sin(ang)};
    }
    const double b = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ void __launch_bounds__(NT) k_err(State* st, const double* part, int nb, double* err, double invN)
{
    __shared__ double red[NT / 64];
    if (st->ctl.done) return;
    double a = 0;
    for (int b = threadIdx.x; b < nb; b += NT) a += part[b];
    const double e = block_sum(a, red);
    if (threadIdx.x == 0) {
        const double v = e * invN;
        err[st->i] = v;
        if (v < 0.1) st->ctl.done = 1;
    }
}


__global__ void __launch_bounds__(NT) k_count(const cx<double>* mT, double thr, double* part, long long N)
{
    __shared__ double red[NT / 64];
    double a = 0;
    const long long s0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, step = (long long)gridDim.x * blockDim.x;
    for (long long s = s0; s < N; s += step) a += (mT[s].re < thr) ? 0.0 : 1.0;
    const double b = block_sum(a, red);
    if (threadIdx.x == 0) part[blockIdx.x] = b;
}

__global__ void k_binarize(cx<double>* mT, double thr, long long N)
{
    const long long s0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, step = (long long)gridDim.x * blockDim.x;
    for (long long s = s0; s < N; s += step) mT[s].re = (mT[s].re < thr) ? 0.0 : 1.0;
}

inline bool fast_len(int L) { return L == 32 || L == 48 || L == 64 || L == 96 || L == 128 || L == 192 || L == 256; }

// z-line tile: TY lines with TY * L <= 2048 complex values (32 KB of LDS),
// TY a power of two
constexpr int zlines(int L)
{
    int ty = 1;
    while (ty * 2 * L <= 2048 && ty < 64) ty *= 2;
    return ty;
}

struct Plan {
    int Nx, Ny, Nz;
    long long N;
    hipStream_t st = nullptr;
    cx<double>*mT = nullptr, *MT = nullptr, *MH = nullptr, *X = nullptr;
    double *normC = nullptr, *part = nullptr, *errd = nullptr;
    const double* m0 = nullptr; // device copy of a start field, if set
    double* m0d = nullptr;
    State* state = nullptr;
    cx<double>* tw[3] = {nullptr, nullptr, nullptr}; // z, y, x
    Geo gy{}, gx{};
    int nblk = 1024;
    int cap_err = 0;
    std::vector<double> err;
    double vol0 = 0; // normC[0] / xyz
    int status = 0, search = 0;

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

    Plan(const double* tpcf, int nx, int ny, int nz) : Nx(nx), Ny(ny), Nz(nz)
    {
        N = (long long)nx * ny * nz;
        if (!fast_len(nx) || !fast_len(ny) || !fast_len(nz))
            throw ArgError{M3D_ERR_UNSUPPORTED, "recPlan: each grid length must be one of 32, 48, 64, 96, 128, 192, 256"};
        HCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        alloc(&mT, sizeof(cx<double>) * N);
        alloc(&MT, sizeof(cx<double>) * N);
        alloc(&MH, sizeof(cx<double>) * N);
        alloc(&X, sizeof(cx<double>) * N);
        alloc(&normC, sizeof(double) * N);
        alloc(&part, sizeof(double) * nblk);
        alloc(&state, sizeof(State));
        const int lens[3] = {nz, ny, nx};
        for (int a = 0; a < 3; ++a) {
            auto t = twiddles<double>(lens[a]);
            alloc(&tw[a], sizeof(cx<double>) * lens[a]);
            HCHK(hipMemcpy(tw[a], t.data(), sizeof(cx<double>) * lens[a], hipMemcpyHostToDevice));
        }
        // normC = sqrt(|TPCF| xyz) (:240-241), on the host in FP64
        std::vector<double> nc((size_t)N);
        for (long long i = 0; i < N; ++i) nc[(size_t)i] = std::sqrt(std::fabs(tpcf[i]) * (double)N);
        vol0 = nc[0] / (1. * (double)N);
        HCHK(hipMemcpy(normC, nc.data(), sizeof(double) * N, hipMemcpyHostToDevice));
        // one-component views for the y pass (lines along y, rows of Nz values)
        // and the x pass (the same with x and y exchanged)
        std::memset(&gy, 0, sizeof gy);
        gy.Nx = nx; gy.Ny = ny; gy.Nz = nz; gy.Nzh = nz; gy.C = 1; gy.nsolve = 1;
        gy.sy = nz; gy.sxa = (long long)ny * nz; gy.sxb = 0; gy.lxb = 0; gy.xbm = 0;
        gx = gy;
        gx.Nx = ny; gx.Ny = nx;
        gx.sy = (long long)ny * nz; gx.sxa = nz;
    }

    ~Plan()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (void* p : {(void*)mT, (void*)MT, (void*)MH, (void*)X, (void*)normC, (void*)part, (void*)errd, (void*)m0d,
                        (void*)state, (void*)tw[0], (void*)tw[1], (void*)tw[2]})
            if (p) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }

    template <int L>
    bool z_at(const Ctl* c, const cx<double>* src, cx<double>* dst, int dir, bool proj)
    {
        if (Nz != L) return false;
        constexpr int TY = zlines(L);
        const long long lines = (long long)Nx * Ny;
        if (lines % TY) throw ArgError{M3D_ERR_UNSUPPORTED, "recPlan: Nx*Ny must be a multiple of the z tile"};
        const size_t sm = ((size_t)L * TY + L) * sizeof(cx<double>);
        const dim3 grid(lines / TY);
        const double n = (double)N;
        if (dir < 0) hipLaunchKernelGGL((kz<L, TY, -1, false>), grid, dim3(NT), sm, st, c, src, dst, tw[0], state, n);
        else if (!proj) hipLaunchKernelGGL((kz<L, TY, 1, false>), grid, dim3(NT), sm, st, c, src, dst, tw[0], state, n);
        else hipLaunchKernelGGL((kz<L, TY, 1, true>), grid, dim3(NT), sm, st, c, src, dst, tw[0], state, n);
        return true;
    }
    void zpass(const Ctl* c, const cx<double>* src, cx<double>* dst, int dir, bool proj = false)
    {
        if (!(z_at<256>(c, src, dst, dir, proj) || z_at<192>(c, src, dst, dir, proj) ||
              z_at<128>(c, src, dst, dir, proj) || z_at<96>(c, src, dst, dir, proj) ||
              z_at<64>(c, src, dst, dir, proj) || z_at<48>(c, src, dst, dir, proj) || z_at<32>(c, src, dst, dir, proj)))
            throw ArgError{M3D_ERR_UNSUPPORTED, "recPlan: no z kernel"};
    }

    // y pass (g = gy, twiddles of y) or x pass (g = gx, twiddles of x)
    template <int L, int DIR>
    bool xy_at(const Geo& g, const Ctl* c, cx<double>* data, const cx<double>* twl)
    {
        if (g.Ny != L) return false;
        constexpr int TK = fast::xytile<double>(L, 1);
        const int nkt = (g.Nzh + TK - 1) / TK;
        const size_t sm = ((size_t)L * TK + L) * sizeof(cx<double>);
        Ptr<double> p;
        std::memset(&p, 0, sizeof p);
        p.S = data;
        p.K = data;
        p.sS = 0;
        p.twy = twl;
        p.ctl = const_cast<Ctl*>(c);
        hipLaunchKernelGGL((fast::kxy<double, Scalar1, 0, DIR, BUF_MAIN, L, TK, NT>), dim3(g.Nx * nkt), dim3(NT), sm, st,
                           g, p, (const double*)nullptr);
        return true;
    }
    template <int DIR>
    void xypass(const Geo& g, const Ctl* c, cx<double>* data, const cx<double>* twl)
    {
        if (!(xy_at<256, DIR>(g, c, data, twl) || xy_at<192, DIR>(g, c, data, twl) || xy_at<128, DIR>(g, c, data, twl) ||
              xy_at<96, DIR>(g, c, data, twl) || xy_at<64, DIR>(g, c, data, twl) || xy_at<48, DIR>(g, c, data, twl) ||
              xy_at<32, DIR>(g, c, data, twl)))
            throw ArgError{M3D_ERR_UNSUPPORTED, "recPlan: no y/x kernel"};
    }

    // unnormalised 3D FFT src -> dst (dst may equal src), gated by c->done
    void fft3(const Ctl* c, const cx<double>* src, cx<double>* dst, int dir)
    {
        zpass(c, src, dst, dir);
        if (dir < 0) {
            xypass<-1>(gy, c, dst, tw[1]);
            xypass<-1>(gx, c, dst, tw[2]);
        } else {
            xypass<1>(gy, c, dst, tw[1]);
            xypass<1>(gx, c, dst, tw[2]);
        }
    }

    // with_x: the re-randomisation path (field and its FFT, gated on the
    // device); chunks without a re-randomisation iteration leave it out
    void iteration(bool with_x)
    {
        const int nb = nblk;
        hipLaunchKernelGGL(k_begin, dim3(1), dim3(64), 0, st, state);
        fft3(&state->ctl, mT, MT, -1);
        if (with_x) {
            hipLaunchKernelGGL(k_rand, dim3(nb), dim3(NT), 0, st, state, X, N);
            fft3(&state->xctl, X, X, -1);
        }
        hipLaunchKernelGGL(k_phase, dim3(nb), dim3(NT), 0, st, state, MT, X, normC, MH, part, N);
        hipLaunchKernelGGL(k_err, dim3(1), dim3(NT), 0, st, state, part, nb, errd, 1.0 / (double)N);
        // inverse transform as x, y, then z with the projection into mT
        xypass<1>(gx, &state->ctl, MH, tw[2]);
        xypass<1>(gy, &state->ctl, MH, tw[1]);
        zpass(&state->ctl, MH, mT, 1, true);
    }

    void set_initial(const double* m)
    {
        if (!m0d) alloc(&m0d, sizeof(double) * N);
        HCHK(hipMemcpy(m0d, m, sizeof(double) * N, hipMemcpyHostToDevice));
        m0 = m0d;
    }

    // runRec(n, beta), :285-406
    int run(int n, double beta, uint64_t seed)
    {
        if (n < 0) throw ArgError{M3D_ERR_ARG, "runRec: n must be >= 0"};
        if (cap_err < std::max(n, 1)) {
            if (errd) (void)hipFree(errd);
            errd = nullptr;
            alloc(&errd, sizeof(double) * std::max(n, 1));
            cap_err = std::max(n, 1);
        }
        State h;
        std::memset(&h, 0, sizeof h);
        h.n = n;
        h.beta = beta;
        h.seed = seed;
        h.xctl.done = 1;
        HCHK(hipMemcpyAsync(state, &h, sizeof h, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_init, dim3(nblk), dim3(NT), 0, st, state, mT, m0, N);
        HCHK(hipGetLastError());
        // chunks of iterations in captured graphs, replayed until done: one
        // with the re-randomisation path, one without, chosen per chunk from
        // the iteration counter (it advances by one per iteration until done)
        const int ck = 16;
        struct Guard {
            hipGraph_t g[2] = {nullptr, nullptr};
            hipGraphExec_t x[2] = {nullptr, nullptr};
            ~Guard()
            {
                for (int k = 0; k < 2; ++k) {
                    if (x[k]) (void)hipGraphExecDestroy(x[k]);
                    if (g[k]) (void)hipGraphDestroy(g[k]);
                }
            }
        } gr;
        for (int v = 0; v < 2; ++v) {
            HCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < ck; ++k) iteration(v == 1);
            HCHK(hipStreamEndCapture(st, &gr.g[v]));
            HCHK(hipGraphInstantiate(&gr.x[v], gr.g[v], nullptr, nullptr, 0));
        }
        const long long max_replays = (long long)n / ck + 2;
        for (long long r = 0; r <= max_replays; ++r) {
            // iteration i re-randomises when (i + 1) % 150 == 0 (k_begin);
            // a pending end still moves the counter by one
            const int i0 = h.i + (h.pend ? 1 : 0);
            bool rr = false;
            for (int k = 0; k < ck; ++k) rr = rr || (i0 + k + 1) % 150 == 0;
            HCHK(hipGraphLaunch(gr.x[rr ? 1 : 0], st));
            HCHK(hipMemcpyAsync(&h, state, sizeof h, hipMemcpyDeviceToHost, st));
            HCHK(hipStreamSynchronize(st));
            if (h.ctl.done) break;
        }
        if (!h.ctl.done) throw ArgError{M3D_ERR_STATE, "runRec: the iteration did not end"};
        status = h.ctl.done;
        const int nerr = (status == 1) ? h.i + 1 : h.i;
        err.assign((size_t)nerr, 0.0);
        if (nerr > 0) HCHK(hipMemcpy(err.data(), errd, sizeof(double) * nerr, hipMemcpyDeviceToHost));
        // volume fraction search (:374-402): threshold .5 - x, x moved by .001
        std::vector<double> pc((size_t)nblk);
        int ctr = 0;
        double x = 0, thr = .5;
        while (true) {
            thr = .5 - x;
            hipLaunchKernelGGL(k_count, dim3(nblk), dim3(NT), 0, st, mT, thr, part, N);
            HCHK(hipGetLastError());
            HCHK(hipMemcpyAsync(pc.data(), part, sizeof(double) * nblk, hipMemcpyDeviceToHost, st));
            HCHK(hipStreamSynchronize(st));
            double sum = 0;
            for (double v : pc) sum += v;
            const double vol_dif = (sum / (1. * (double)N) - vol0);
            if (vol_dif > 0) x -= .001;
            else x += .001;
            if (std::fabs(vol_dif) < .0001 || ctr > 200) break;
            ctr += 1;
        }
        hipLaunchKernelGGL(k_binarize, dim3(nblk), dim3(NT), 0, st, mT, thr, N);
        HCHK(hipGetLastError());
        HCHK(hipStreamSynchronize(st));
        search = ctr;
        return nerr;
    }
};

} // namespace rec
} // namespace m3d

using namespace m3d;

struct m3d_rec {
    rec::Plan* plan = nullptr;
    std::string err;
    bool ran = false;
};

namespace {

template <typename F>
int guarded(m3d_rec* h, F&& f)
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

} // namespace

int m3d_rec_create(const double* tpcf, int nx, int ny, int nz, m3d_rec** out)
{
    if (!out || !tpcf || nx < 1 || ny < 1 || nz < 1) return M3D_ERR_ARG;
    *out = nullptr;
    m3d_rec* h = new (std::nothrow) m3d_rec;
    if (!h) return M3D_ERR_OOM;
    const int rc = guarded(h, [&]() {
        h->plan = new rec::Plan(tpcf, nx, ny, nz);
        return M3D_OK;
    });
    if (rc != M3D_OK) {
        delete h->plan;
        delete h;
        return rc;
    }
    *out = h;
    return M3D_OK;
}

int m3d_rec_destroy(m3d_rec* h)
{
    if (!h) return M3D_ERR_ARG;
    delete h->plan;
    delete h;
    return M3D_OK;
}

const char* m3d_rec_last_error(m3d_rec* h) { return h ? h->err.c_str() : "null handle"; }

int m3d_rec_set_initial(m3d_rec* h, const double* m0)
{
    return guarded(h, [&]() {
        if (!m0) throw ArgError{M3D_ERR_ARG, "set_initial: null field"};
        h->plan->set_initial(m0);
        return M3D_OK;
    });
}

int m3d_rec_run(m3d_rec* h, int n, double beta, uint64_t seed, int* iterations, int* status, int* search_steps)
{
    return guarded(h, [&]() {
        const int it = h->plan->run(n, beta, seed);
        if (iterations) *iterations = it;
        if (status) *status = h->plan->status;
        if (search_steps) *search_steps = h->plan->search;
        h->ran = true;
        return M3D_OK;
    });
}

int m3d_rec_result(m3d_rec* h, double* mtilda_ri, double* err, int maxn, int* nerr)
{
    return guarded(h, [&]() {
        if (!h->ran) throw ArgError{M3D_ERR_STATE, "result: runRec first"};
        rec::Plan& p = *h->plan;
        if (mtilda_ri) HCHK(hipMemcpy(mtilda_ri, p.mT, sizeof(cx<double>) * p.N, hipMemcpyDeviceToHost));
        const int k = (int)p.err.size();
        if (err) std::memcpy(err, p.err.data(), sizeof(double) * (size_t)std::min(k, std::max(maxn, 0)));
        if (nerr) *nerr = k;
        return M3D_OK;
    });
}
