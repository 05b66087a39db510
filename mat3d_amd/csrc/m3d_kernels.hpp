// m3d_kernels.hpp -- gfx950 kernels of the FFT homogenization loop.
//
// One accelerated (Eyre-Milton) iteration of src/3d/FFT_analysis3D.cpp:751-794
// (and src/3dTherm/FFT_thermal3D.cpp:446-485) is realised as four passes over
// HBM, all working in place on one half-spectrum buffer per solve:
//
//   zpass  (MODE_IUF): c2r along z of the Green output  -> u_c
//                      u <- u + A(x)(u_c - u)            (update,  :283-295)
//                      v  = M(x) u, FP64 partial sums    (multipCx :236-247, sum_field)
//                      tau = (M(x) - M0) u               (polarisation :757-767)
//                      r2c along z of tau
//   ypass  (fwd)     : complex FFT along y
//   xgreen           : complex FFT along x, u_c^ = -G0(xi) tau^ / N, DC = E
//                      (multiplyGF :297-348 with the Nyquist projection that
//                      FFTW's c2r applies), inverse FFT along x
//   ypass  (inv)     : inverse complex FFT along y
//
// The equilibrium check (convergenceCheck :351-416) needs the spectrum of
// sigma.  It is only needed when the stop rule can fire, i.e. when
// |diff_sum_field| < tol_field, so it runs on demand (zpass MODE_CHK, ypass,
// xcheck) and is skipped otherwise; the stop decision is the reference's.
//
// Layouts (AoS, the reference's own, FFT_analysis3D.cpp:110-150):
//   real field  U[((x*Ny + y)*Nz + z)*C + c]
//   spectrum    S[((x*Ny + y)*Nzh + kz)*C + c],  Nzh = Nz/2 + 1
// Tensor index 0 is the z axis, 1 is y, 2 is x (FFT_analysis3D.cpp:211).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "m3d_fft.hpp"

namespace m3d {

constexpr int NCOEF = 8; // per-phase coefficients
constexpr int ZTHREADS = 256;
constexpr int XYTHREADS = 512;
constexpr int EPT = 16;  // complex values per thread per pass

// z-pass modes: forward only, update (inverse, update, forward), check FFT of
// the current / previous iterate, update of an anchored FP32 correction (adds
// r), the FP64 anchor iteration of the polish (inverse and update only;
// writes r = u_new - u_old and the first correction e = r in FP32), and the
// FP64 start of the polish from the FP32 iterate (forward only; reads the
// FP32 ring, writes the promoted FP64 iterate).  The last two: the
// specialised z pass only.
enum { MODE_FWD = 0, MODE_IUF = 1, MODE_CHK_CUR = 2, MODE_CHK_PREV = 3, MODE_IUFR = 4, MODE_IUA = 5, MODE_FWDP = 6 };
enum { BUF_MAIN = 0, BUF_CHK_CUR = 2, BUF_CHK_PREV = 3, BUF_K = 4 };

// per-solve control block, read and written on the device only
struct Ctl {
    int it;        // itNo
    int done;      // 0 running, 1 converged, 2 iteration cap
    int need_cur;  // equilibrium check of the current iterate needed
    int need_prev; // ... and of the previous iterate
    int eq_it;     // iterate index eq_val belongs to (-1: none)
    int cap;       // iteration cap (reference: 2000 / 500)
    int first;     // basic scheme: 1 while the stage after the increment start runs
    int err;       // basic scheme: constitutive update failed (Newton NaN / no convergence)
    int sb;        // elastic: 1 strain base, 0 stress base; thermal: 1 gradient base
    // anchored polish: the iterate is A + e with the FP64 anchor A (iterate
    // anc_it of the FP64 ring) and an FP32 correction e (FP32 ring); the
    // FP32 passes iterate e <- r + M e (r = T(A) - A, load term dropped)
    int corr;
    int anc_it;
    int pad;
    double sum_anchor; // sum of the dual field of the anchor (FP64)
    double sum_prev;  // sum of the dual field at iterate `it`
    double eq_val;    // eqVar at iterate eq_it
    double dsum, deq, eqvar;
    double tol_f, tol_eq;
    double norm;      // N * (M[0]_00/2 + M[1]_00/2) * sum(E)
    double E[6];      // macroscopic load (DC of the Green output)
    // basic scheme with mixed control (updateE, FFT_analysis3D.cpp:583-594)
    double X0[6];     // stress direction
    double dEsum;     // dE.sum()
    double incNo;     // increment number (1-based)
    double nc;        // N * (C[0]_00/2 + C[1]_00/2): Delta-sum normalisation without sum(E)
    double la0, mu0;  // reference medium (S0 = C0^-1 applied in closed form)
    double dsum_hist[16]; // dsum of iterate i at [i % 16] (how far an FP32 switch rolls back)
};

struct Geo {
    int Nx, Ny, Nz, Nzh, C;
    long long N, H;
    FftAxis az, ay, ax; // az: Nz/2 (even Nz) or Nz (odd Nz)
    int zeven;
    int TY, nty;        // z tiles: TY lines (y) per tile, Ny/TY tiles per x plane
    int TKy, nky;       // y tiles: TKy kz values, ceil(Nzh/TKy) tiles
    int TKx, nkx;       // x tiles
    int lsz;            // complex stride of one z line in LDS
    int nsolve;
    int nbz, nbx;       // blocks per solve of the z pass and the x pass
    int TKc, nbc;       // x pass of the equilibrium check: kz tile, partial sums per solve
    double invN;
    // spectrum layout of the specialised kernels: the half-spectrum row
    // (x, y) (Nzh*C complex values) starts at
    //   (x >> lxb) * sxa + (x & xbm) * sxb + y * sy
    // plain rows: lxb = 0, xbm = 0, sxa = Ny*Nzh*C (+ padding), sy = Nzh*C;
    // x planes in blocks of 2^lxb: sxa = Ny*2^lxb*Nzh*C, sxb = Nzh*C, sy = 2^lxb*Nzh*C
    int lxb, xbm;
    long long sxa, sxb, sy;
    long long spec;     // complex values of one solve's spectrum (rows + padding)
    int fz, fy, fx;     // host only: compile-time specialised kernel used for this axis
    int gen;            // host only: generic kernels need the odd radices
};

template <typename T>
struct Ptr {
    T* U;                   // ring of `ring` iterate buffers, each holding every solve
    long long sU;           // per-solve stride of U (reals)
    long long sB;           // stride between ring buffers (reals)
    int ring;               // iterate j lives in buffer j mod ring
    cx<T>* S;               // main spectrum
    cx<T>* K;               // check spectrum
    long long sS;           // per-solve stride of S and K (complex)
    const uint8_t* ph;      // phase ids, per RVE
    long long sPh;
    const int* rve_of;      // solve -> RVE index
    const T* coef;          // per RVE: [maxph][NCOEF]
    int sCoef;
    const cx<T>* twz;
    const cx<T>* twy;
    const cx<T>* twx;
    double* psum;           // z-pass partial sums, [nsolve][nbz]
    double* pchk;           // check partials, [nsolve][2][nbc + 1]
    Ctl* ctl;
    // anchored polish (solves with ctl.corr): FP32 passes add R to the
    // update; FP64 passes read the iterate as U[anc_it] + e, e from the FP32
    // ring eU (stride eB between ring buffers, `ering` buffers)
    const float* R;
    const float* eU;
    long long eB;
    int ering;
};

// ----------------------------------------------------------------- physics
// Elastic: isotropic maps M u = a tr(u)/3 I + b dev(u) on Voigt vectors with
// tensor shears ([11,22,33,23,13,12], FFT_analysis3D.cpp:457-462).
struct Elastic {
    static constexpr int C = 6;
    // workgroups of a check-pass launch: the check is rarely needed (Δsum
    // reaches its tolerance near the stop), so the launches that leave at
    // once dominate and a smaller grid wins (measured: 8192 against 32768,
    // 256^3 headline -1.5%, 64-RVE batch -2.3%)
    static constexpr int chk_grid = 8192;
    template <typename T>
    __device__ static __forceinline__ void iso(T a, T b, const T* u, T* o)
    {
        const T t = (u[0] + u[1] + u[2]) * T(1.0 / 3.0);
        const T q = (a - b) * t;
        o[0] = b * u[0] + q;
        o[1] = b * u[1] + q;
        o[2] = b * u[2] + q;
        o[3] = b * u[3];
        o[4] = b * u[4];
        o[5] = b * u[5];
    }
    // coef: [a_v,b_v, a_t,b_t, a_a,b_a]
    template <typename T>
    __device__ static __forceinline__ void dual(const T* k, const T* u, T* o) { iso(k[0], k[1], u, o); }
    template <typename T>
    __device__ static __forceinline__ void polar(const T* k, const T* u, T* o) { iso(k[2], k[3], u, o); }
    template <typename T>
    __device__ static __forceinline__ void accel(const T* k, const T* u, T* o) { iso(k[4], k[5], u, o); }
    // frequency map, FFT_analysis3D.cpp:202-209
    __device__ static __forceinline__ int fmap(int k, int N) { return k <= (N - 1) / 2 ? k : -(N - k); }
    __device__ static __forceinline__ int fmapx(int k, int Nx, int Ny) { return fmap(k, Nx); }
};

// Thermal: scalar conductivities, 3 components (FFT_thermal3D.cpp:180-268).
struct Thermal {
    static constexpr int C = 3;
    // at high conductivity contrast the check runs in most iterations (the
    // 512^3 three-phase workload: every one), where 32768 workgroups beat 8192
    // by 2.5%
    static constexpr int chk_grid = 32768;
    // coef: [k_v, k_t, k_a]
    template <typename T>
    __device__ static __forceinline__ void dual(const T* k, const T* u, T* o)
    {
        o[0] = k[0] * u[0]; o[1] = k[0] * u[1]; o[2] = k[0] * u[2];
    }
    template <typename T>
    __device__ static __forceinline__ void polar(const T* k, const T* u, T* o)
    {
        o[0] = k[1] * u[0]; o[1] = k[1] * u[1]; o[2] = k[1] * u[2];
    }
    template <typename T>
    __device__ static __forceinline__ void accel(const T* k, const T* u, T* o)
    {
        o[0] = k[2] * u[0]; o[1] = k[2] * u[1]; o[2] = k[2] * u[2];
    }
};

// ------------------------------------------------------- Green operators
// per-solve constants of the isotropic Green operators, computed once per
// tile instead of per frequency (same values and operation order as the
// per-frequency expressions they replace)
template <typename T>
struct GreenK {
    T la0, mu0;
    T d2;   // (la0 + mu0) / (mu0 (la0 + 2 mu0))
    T c3;   // 2 la0 mu0 / (la0 + 2 mu0)
    T m2;   // 2 mu0
    T h2mu; // 1 / (2 mu0): FP32 multiplies by it where FP64 divides
    T k0;   // thermal reference conductivity
};

template <typename T>
__device__ __forceinline__ GreenK<T> green_k(T la0, T mu0)
{
    GreenK<T> K;
    K.la0 = la0;
    K.mu0 = mu0;
    K.d2 = (la0 + mu0) / (mu0 * (la0 + T(2) * mu0));
    K.c3 = T(2) * la0 * mu0 / (la0 + T(2) * mu0);
    K.m2 = T(2) * mu0;
    K.h2mu = T(1) / (T(2) * mu0);
    K.k0 = T(0);
    return K;
}

// the constants of a solve's medium: elastic prm = {la0, mu0}, thermal {k0}
template <typename T, int C>
__device__ __forceinline__ GreenK<T> green_k_of(const T* prm)
{
    if constexpr (C == 6) return green_k<T>(prm[0], prm[1]);
    GreenK<T> K{};
    K.k0 = prm[0];
    return K;
}

// Strain base, Gamma0 (FFT_analysis3D.cpp:223-224) contracted with a
// symmetric tau (all shear pairs summed): closed form
//   (G:t)_ij = (xi_i (t xi)_j + xi_j (t xi)_i) / (2 mu0 |xi|^2)
//              - (la0+mu0)/(mu0(la0+2mu0)) xi_i xi_j (xi.t.xi) / |xi|^4
// Stress base, Delta0 (:226-228): with P = I - n n, n = xi/|xi|,
//   (D:t)_ij = c3 P_ij tr(P t) + 2 mu0 (P t P)_ij,  c3 = 2 la0 mu0/(la0+2mu0)
// Returns G:t (complex, per Voigt component).
template <typename T>
__device__ __forceinline__ void green_elastic_k(const T xi[3], const cx<T> t[6], cx<T> o[6], int strain_base,
                                                const GreenK<T>& K)
{
    const T n2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
    // FP32: the hardware reciprocal (1 ulp) in place of the IEEE division
    // sequence; FP32 iterates are refined in FP64, whose path divides exactly
    T inv;
    if constexpr (sizeof(T) == 4) inv = __builtin_amdgcn_rcpf(n2);
    else inv = T(1) / n2;
    // w = t xi, tensor from Voigt: 0:(0,0) 1:(1,1) 2:(2,2) 3:(1,2) 4:(0,2) 5:(0,1)
    cx<T> w0 = cscale(t[0], xi[0]) + cscale(t[5], xi[1]) + cscale(t[4], xi[2]);
    cx<T> w1 = cscale(t[5], xi[0]) + cscale(t[1], xi[1]) + cscale(t[3], xi[2]);
    cx<T> w2 = cscale(t[4], xi[0]) + cscale(t[3], xi[1]) + cscale(t[2], xi[2]);
    cx<T> s = cscale(w0, xi[0]) + cscale(w1, xi[1]) + cscale(w2, xi[2]);
    const cx<T> w[3] = {w0, w1, w2};
    const int pi[6] = {0, 1, 2, 1, 0, 0};
    const int pj[6] = {0, 1, 2, 2, 2, 1};
    if (strain_base) {
        T a;
        if constexpr (sizeof(T) == 4) a = inv * K.h2mu;
        else a = inv / (T(2) * K.mu0);
        const T b = K.d2 * inv * inv;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int i = pi[q], j = pj[q];
            cx<T> r = cscale(cscale(w[j], xi[i]) + cscale(w[i], xi[j]), a);
            o[q] = r - cscale(s, b * xi[i] * xi[j]);
        }
    } else {
        // n n = xi xi / n2 ; s/n2 = n.t.n ; w/|xi| = t n ; with
        // (P t P)_ij = t_ij - n_i w'_j - w'_i n_j + sn n_i n_j (w' = t n) and
        // P_ij = delta_ij - xi_i xi_j / n2 the output collects into
        //   o_ij = 2mu0 t_ij - (2mu0/n2)(xi_i w_j + xi_j w_i) + xi_i xi_j G + delta_ij D
        // with D = c3 tr(P t) and G = (2mu0 sn - D) / n2, formed once per frequency
        const cx<T> trt = t[0] + t[1] + t[2];
        const cx<T> sn = cscale(s, inv); // n.t.n
        const cx<T> D = cscale(trt - sn, K.c3);
        const cx<T> G = cscale(cscale(sn, K.m2) - D, inv);
        const T a = K.m2 * inv;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const int i = pi[q], j = pj[q];
            const cx<T> x = cscale(w[j], xi[i]) + cscale(w[i], xi[j]);
            cx<T> r = cscale(t[q], K.m2) - cscale(x, a) + cscale(G, xi[i] * xi[j]);
            if (i == j) r = r + D;
            o[q] = r;
        }
    }
}

template <typename T>
__device__ __forceinline__ void green_elastic(const T xi[3], const cx<T> t[6], cx<T> o[6], int strain_base,
                                              T la0, T mu0)
{
    green_elastic_k<T>(xi, t, o, strain_base, green_k<T>(la0, mu0));
}

// Thermal flux base:     D_ij = k0 (delta_ij - xi_i xi_j/|xi|^2) (:161-169)
template <typename T>
__device__ __forceinline__ void green_thermal(const T xi[3], const cx<T> t[3], cx<T> o[3], int grad_base, T k0)
{
    const T n2 = xi[0] * xi[0] + xi[1] * xi[1] + xi[2] * xi[2];
    cx<T> s = cscale(t[0], xi[0]) + cscale(t[1], xi[1]) + cscale(t[2], xi[2]);
    // FP32: hardware reciprocal, as in green_elastic
    auto rcp = [](T x) {
        if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
        else return T(1) / x;
    };
    if (grad_base) {
        const T a = rcp(k0 * n2);
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = cscale(s, a * xi[c]);
    } else {
        const T a = rcp(n2);
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = cscale(t[c] - cscale(s, a * xi[c]), k0);
    }
}

// ------------------------------------------------------------ reductions
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// deterministic block sum (fixed tree); result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) r += red[i];
    }
    return r;
}

template <typename T>
__device__ __forceinline__ T* u_buf(const Ptr<T>& p, int it, int s)
{
    int b = it % p.ring;
    if (b < 0) b += p.ring;
    return p.U + (long long)b * p.sB + (long long)s * p.sU;
}

// FP32 correction e of iterate `it` of an anchored solve
template <typename T>
__device__ __forceinline__ const float* e_buf(const Ptr<T>& p, int it, int s)
{
    int b = it % p.ering;
    if (b < 0) b += p.ering;
    return p.eU + (long long)b * p.eB + (long long)s * p.sU;
}

// --------------------------------------------------------------- z pass
// One block: solve s, plane x, lines y0..y0+TY-1, all components.
// LDS: TY*C lines of lsz complex each; line l = yl*C + c.
template <typename T, typename PH, int MODE, bool GEN>
__global__ void __launch_bounds__(ZTHREADS) k_zpass(Geo g, Ptr<T> p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[ZTHREADS / 64];
    constexpr int C = PH::C;
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    T* sr = reinterpret_cast<T*>(smem_raw);

    const int per = g.Nx * g.nty;
    const int s = blockIdx.x / per;
    const int t = blockIdx.x % per;
    const int x = t / g.nty;
    const int y0 = (t % g.nty) * g.TY;
    const Ctl& ctl = p.ctl[s];
    if (ctl.done) return;
    if (MODE == MODE_CHK_CUR && !ctl.need_cur) return;
    if (MODE == MODE_CHK_PREV && !ctl.need_prev) return;
    const int it = ctl.it;

    const int Nz = g.Nz, Nzh = g.Nzh, lsz = g.lsz, TY = g.TY;
    const int nl = TY * C;
    const long long vox0 = ((long long)x * g.Ny + y0) * Nz;         // first voxel of tile
    const long long spec0 = ((long long)x * g.Ny + y0) * Nzh * C;  // first spectrum entry
    const int rv = p.rve_of[s];
    const uint8_t* ph = p.ph + (long long)rv * p.sPh + vox0;
    const T* coef = p.coef + (long long)rv * p.sCoef;

    const T* Uin;
    T* Uout = nullptr;
    if (MODE == MODE_FWD) Uin = u_buf(p, it, s);
    else if (MODE == MODE_IUF) { Uin = u_buf(p, it, s); Uout = u_buf(p, it + 1, s); }
    else if (MODE == MODE_CHK_CUR) Uin = u_buf(p, it, s);
    else Uin = u_buf(p, it - 1, s);
    // anchored polish: FP32 updates add r, FP64 checks read anchor + e
    const bool corr = ctl.corr != 0;
    const float* Rin = nullptr;
    const float* Ein = nullptr;
    if (corr) {
        if (sizeof(T) == 4 && MODE == MODE_IUF) Rin = p.R + (long long)s * p.sU + vox0 * C;
        if (sizeof(T) == 8 && (MODE == MODE_CHK_CUR || MODE == MODE_CHK_PREV)) {
            Uin = u_buf(p, ctl.anc_it, s);
            Ein = e_buf(p, MODE == MODE_CHK_CUR ? it : it - 1, s) + vox0 * C;
        }
    }
    Uin += vox0 * C;
    if (Uout) Uout += vox0 * C;

    const LineGeo lg{nl, nl, lsz, 0, 1};
    const int tid = threadIdx.x, nthr = blockDim.x;

    if (MODE == MODE_IUF) {
        // 1. load the Green output tile (contiguous in global memory)
        const cx<T>* src = p.S + (long long)s * p.sS + spec0;
        const int tot = TY * Nzh * C;
        for (int e = tid; e < tot; e += nthr) {
            const int yl = e / (Nzh * C);
            const int r = e - yl * Nzh * C;
            const int kz = r / C, c = r - kz * C;
            sl[(yl * C + c) * lsz + kz] = src[e];
        }
        __syncthreads();
        // 2. c2r pre-processing (real parts of the kz=0 and Nyquist bins only)
        if (g.zeven) {
            const int L = Nz / 2;
            const int half = L / 2 + 1;
            for (int e = tid; e < nl * half; e += nthr) {
                const int l = e / half, k = e - l * half;
                cx<T>* ln = sl + l * lsz;
                if (k == 0) {
                    const T x0 = ln[0].re, xl = ln[L].re;
                    ln[0] = cx<T>{x0 + xl, x0 - xl};
                } else {
                    const int kk = L - k;
                    const cx<T> Xk = ln[k], Xc = ln[kk];
                    // Z[k]  = (X[k]  + X[kk]*) + i W^-k  (X[k]  - X[kk]*)
                    // Z[kk] = (X[kk] + X[k]*)  + i W^-kk (X[kk] - X[k]*)
                    cx<T> w = p.twz[k];
                    w.im = -w.im;
                    const cx<T> sm = Xk + cconj(Xc), df = Xk - cconj(Xc);
                    const cx<T> zk = sm + rot90<1>(cmul(w, df));
                    ln[k] = zk;
                    if (kk != k) {
                        cx<T> w2 = p.twz[kk];
                        w2.im = -w2.im;
                        const cx<T> sm2 = Xc + cconj(Xk), df2 = Xc - cconj(Xk);
                        ln[kk] = sm2 + rot90<1>(cmul(w2, df2));
                    }
                }
            }
            __syncthreads();
            fft_lines<T, 1, EPT, false, GEN>(sl, lg, g.az, p.twz);
        } else {
            // odd Nz: Hermitian extension, then a full complex inverse FFT
            for (int e = tid; e < nl * Nzh; e += nthr) {
                const int l = e / Nzh, k = e - l * Nzh;
                cx<T>* ln = sl + l * lsz;
                if (k == 0) ln[0].im = T(0);
                else ln[Nz - k] = cconj(ln[k]);
            }
            __syncthreads();
            fft_lines<T, 1, EPT, false, GEN>(sl, lg, g.az, p.twz);
            // pack real parts as reals at the front of the line
            // (in-place compaction needs a barrier between read and write)
            T vals[EPT];
#pragma unroll
            for (int i = 0; i < EPT; ++i) {
                const int e = tid + i * nthr;
                if (e < nl * Nz) vals[i] = sl[(e / Nz) * lsz + (e % Nz)].re;
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < EPT; ++i) {
                const int e = tid + i * nthr;
                if (e < nl * Nz) sr[(e / Nz) * 2 * lsz + (e % Nz)] = vals[i];
            }
            __syncthreads();
        }
    }

    // 3. per-voxel work
    double acc = 0.0;
    const int nv = TY * Nz;
    for (int v = tid; v < nv; v += nthr) {
        const int yl = v / Nz, z = v - yl * Nz;
        const T* kp = coef + (int)ph[v] * NCOEF;
        T u[C], o[C];
#pragma unroll
        for (int c = 0; c < C; ++c) u[c] = Uin[(long long)v * C + c];
        if (Ein) {
#pragma unroll
            for (int c = 0; c < C; ++c) u[c] += (T)Ein[(long long)v * C + c];
        }
        if (MODE == MODE_IUF) {
            T d[C], a[C];
#pragma unroll
            for (int c = 0; c < C; ++c) d[c] = sr[(yl * C + c) * 2 * lsz + z] - u[c];
            PH::accel(kp, d, a);
#pragma unroll
            for (int c = 0; c < C; ++c) u[c] = u[c] + a[c];
            if (Rin) {
#pragma unroll
                for (int c = 0; c < C; ++c) u[c] += (T)Rin[(long long)v * C + c];
            }
#pragma unroll
            for (int c = 0; c < C; ++c) Uout[(long long)v * C + c] = u[c];
        }
        if (MODE == MODE_FWD || MODE == MODE_IUF) {
            PH::dual(kp, u, o);
#pragma unroll
            for (int c = 0; c < C; ++c) acc += (double)o[c];
            PH::polar(kp, u, o);
        } else {
            // sigma: the dual field in the strain (gradient) base, u itself otherwise
            if (ctl.sb) PH::dual(kp, u, o);
            else {
#pragma unroll
                for (int c = 0; c < C; ++c) o[c] = u[c];
            }
        }
#pragma unroll
        for (int c = 0; c < C; ++c) sr[(yl * C + c) * 2 * lsz + z] = o[c];
    }
    if (MODE == MODE_FWD || MODE == MODE_IUF) {
        const double bs = block_sum(acc, red);
        if (tid == 0) p.psum[(long long)s * g.nbz + t] = bs;
    }
    __syncthreads();

    // 4. forward r2c along z
    if (g.zeven) {
        fft_lines<T, -1, EPT, false, GEN>(sl, lg, g.az, p.twz);
        const int L = Nz / 2;
        const int half = L / 2 + 1;
        for (int e = tid; e < nl * half; e += nthr) {
            const int l = e / half, k = e - l * half;
            cx<T>* ln = sl + l * lsz;
            if (k == 0) {
                const cx<T> z0 = ln[0];
                ln[0] = cx<T>{z0.re + z0.im, T(0)};
                ln[L] = cx<T>{z0.re - z0.im, T(0)};
            } else {
                const int kk = L - k;
                const cx<T> zk = ln[k], zc = cconj(ln[kk]);
                const cx<T> E = cscale(zk + zc, T(0.5));
                const cx<T> O = cscale(rot90<-1>(zk - zc), T(0.5)); // -i/2 (zk - zc)
                ln[k] = E + cmul(p.twz[k], O);
                if (kk != k) ln[kk] = cconj(E) + cmul(p.twz[kk], cconj(O));
            }
        }
        __syncthreads();
    } else {
        // odd Nz: spread reals into complex slots (backwards-safe via registers)
        T vals[EPT];
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * nthr;
            if (e < nl * Nz) vals[i] = sr[(e / Nz) * 2 * lsz + (e % Nz)];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < EPT; ++i) {
            const int e = tid + i * nthr;
            if (e < nl * Nz) sl[(e / Nz) * lsz + (e % Nz)] = cx<T>{vals[i], T(0)};
        }
        __syncthreads();
        fft_lines<T, -1, EPT, false, GEN>(sl, lg, g.az, p.twz);
    }

    // 5. store the tile (contiguous)
    {
        cx<T>* dst = ((MODE == MODE_FWD || MODE == MODE_IUF) ? p.S : p.K) + (long long)s * p.sS + spec0;
        const int tot = TY * Nzh * C;
        for (int e = tid; e < tot; e += nthr) {
            const int yl = e / (Nzh * C);
            const int r = e - yl * Nzh * C;
            const int kz = r / C, c = r - kz * C;
            dst[e] = sl[(yl * C + c) * lsz + kz];
        }
    }
}

// --------------------------------------------------------------- y pass
// Block: solve s, plane x, kz tile [kz0, kz0+nk); LDS [y][kzl*C + c].
template <typename T, int C, int DIR, int WHICH, bool GEN>
__global__ void __launch_bounds__(XYTHREADS) k_ypass(Geo g, Ptr<T> p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    const int per = g.Nx * g.nky;
    const int s = blockIdx.x / per;
    const int t = blockIdx.x % per;
    const int x = t / g.nky;
    const int kz0 = (t % g.nky) * g.TKy;
    const Ctl& ctl = p.ctl[s];
    if (ctl.done) return;
    if (WHICH == BUF_CHK_CUR && !ctl.need_cur) return;
    if (WHICH == BUF_CHK_PREV && !ctl.need_prev) return;
    const int nk = min(g.TKy, g.Nzh - kz0);
    const int w = nk * C;
    cx<T>* base = (WHICH == BUF_MAIN ? p.S : p.K) + (long long)s * p.sS + ((long long)x * g.Ny * g.Nzh + kz0) * C;
    const long long ystride = (long long)g.Nzh * C;
    const int tot = g.Ny * w;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int y = e / w, r = e - y * w;
        sl[e] = base[y * ystride + r];
    }
    __syncthreads();
    fft_lines<T, DIR, EPT, true, GEN>(sl, LineGeo{w, w, 1, 0, w}, g.ay, p.twy);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int y = e / w, r = e - y * w;
        base[y * ystride + r] = sl[e];
    }
}

// ------------------------------------------------ x pass with Green step
// frequency vector (tensor order z, y, x) of spectrum index (kx, ky, kz)
__device__ __forceinline__ void freq_el(const Geo& g, int kx, int ky, int kz, int f[3])
{
    f[0] = Elastic::fmap(kz, g.Nz);
    f[1] = Elastic::fmap(ky, g.Ny);
    f[2] = Elastic::fmap(kx, g.Nx);
}

// thermal maps: gradient base and check use the N/2-positive map
// (FFT_thermal3D.cpp:113-120); flux base the (N-1)/2 map with -(Ny-x) for x (:150-151)
__device__ __forceinline__ int fmap_t(int k, int N) { return k <= N / 2 ? k : -(N - k); }
__device__ __forceinline__ void freq_th(const Geo& g, int sb, int kx, int ky, int kz, int f[3])
{
    if (sb) {
        f[0] = fmap_t(kz, g.Nz);
        f[1] = fmap_t(ky, g.Ny);
        f[2] = fmap_t(kx, g.Nx);
    } else {
        f[0] = Elastic::fmap(kz, g.Nz);
        f[1] = Elastic::fmap(ky, g.Ny);
        f[2] = kx <= (g.Nx - 1) / 2 ? kx : -(g.Ny - kx);
    }
}

template <typename T, typename PH>
__device__ __forceinline__ void green_apply_k(const Geo& g, int sb, const GreenK<T>& K, int kx, int ky, int kz, cx<T>* v)
{
    constexpr int C = PH::C;
    int f[3];
    if constexpr (C == 6) freq_el(g, kx, ky, kz, f);
    else freq_th(g, sb, kx, ky, kz, f);
    T xi[3] = {T(f[0]), T(f[1]), T(f[2])};
    cx<T> in[C], o[C];
#pragma unroll
    for (int c = 0; c < C; ++c) in[c] = v[c];
    if constexpr (C == 6) green_elastic_k<T>(xi, in, o, sb, K);
    else green_thermal<T>(xi, in, o, sb, K.k0);
    // c2r of a multi-dimensional half spectrum keeps only the Hermitian part
    // of the kz=0 and kz=Nz/2 planes: average with the operator of the
    // conjugate-partner index when its frequency vector is not -xi.
    const bool plane = (kz == 0) || (g.zeven && kz == g.Nz / 2);
    if (plane) {
        const int px = (g.Nx - kx) % g.Nx, py = (g.Ny - ky) % g.Ny;
        int fp[3];
        if constexpr (C == 6) freq_el(g, px, py, kz, fp);
        else freq_th(g, sb, px, py, kz, fp);
        if (fp[0] != -f[0] || fp[1] != -f[1] || fp[2] != -f[2]) {
            T xp[3] = {T(fp[0]), T(fp[1]), T(fp[2])};
            cx<T> o2[C];
            if constexpr (C == 6) green_elastic_k<T>(xp, in, o2, sb, K);
            else green_thermal<T>(xp, in, o2, sb, K.k0);
#pragma unroll
            for (int c = 0; c < C; ++c) o[c] = cscale(o[c] + o2[c], T(0.5));
        }
    }
    const T sc = T(-g.invN);
#pragma unroll
    for (int c = 0; c < C; ++c) v[c] = cscale(o[c], sc);
}

template <typename T, typename PH>
__device__ __forceinline__ void green_apply(const Geo& g, int sb, const T* prm, int kx, int ky, int kz, cx<T>* v)
{
    green_apply_k<T, PH>(g, sb, green_k_of<T, PH::C>(prm), kx, ky, kz, v);
}

// prm: elastic {la0, mu0}; thermal {k0}
template <typename T, typename PH, bool GEN>
__global__ void __launch_bounds__(XYTHREADS) k_xgreen(Geo g, Ptr<T> p, const T* __restrict__ prm_all)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    constexpr int C = PH::C;
    const int per = g.Ny * g.nkx;
    const int s = blockIdx.x / per;
    const int t = blockIdx.x % per;
    const int y = t / g.nkx;
    const int kz0 = (t % g.nkx) * g.TKx;
    const Ctl& ctl = p.ctl[s];
    if (ctl.done) return;
    const T* prm = prm_all + 2 * p.rve_of[s];
    const int nk = min(g.TKx, g.Nzh - kz0);
    const int w = nk * C;
    cx<T>* base = p.S + (long long)s * p.sS + ((long long)y * g.Nzh + kz0) * C;
    const long long xstride = (long long)g.Ny * g.Nzh * C;
    const int tot = g.Nx * w;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int x = e / w, r = e - x * w;
        sl[e] = base[x * xstride + r];
    }
    __syncthreads();
    const LineGeo lg{w, w, 1, 0, w};
    fft_lines<T, -1, EPT, true, GEN>(sl, lg, g.ax, p.twx);
    for (int f = threadIdx.x; f < g.Nx * nk; f += blockDim.x) {
        const int kx = f / nk, kzl = f - kx * nk;
        const int kz = kz0 + kzl;
        cx<T>* v = sl + kx * w + kzl * C;
        if (kx == 0 && y == 0 && kz == 0) {
#pragma unroll
            for (int c = 0; c < C; ++c) v[c] = cx<T>{T(ctl.E[c]), T(0)};
        } else {
            green_apply<T, PH>(g, ctl.sb, prm, kx, y, kz, v);
        }
    }
    __syncthreads();
    fft_lines<T, 1, EPT, true, GEN>(sl, lg, g.ax, p.twx);
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int x = e / w, r = e - x * w;
        base[x * xstride + r] = sl[e];
    }
}

// ---------------------------------------------- x pass of the check FFT
// eqVar partials: mechanical sums kz >= 1 of sqrt|sum_c |s_c|^2 W_c| with
// W = (xi0^2, xi1^2, xi2^2, 2xi1+2xi2, 2xi0+2xi2, 2xi0+2xi1) (FFT_analysis3D.cpp:366-389);
// thermal sums kz >= 0 of sqrt(sum_c |q_c|^2 xi_c^2) (FFT_thermal3D.cpp:315-335).
template <typename T, typename PH, int WHICH, bool GEN>
__global__ void __launch_bounds__(XYTHREADS) k_xcheck(Geo g, Ptr<T> p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[XYTHREADS / 64];
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    constexpr int C = PH::C;
    const int per = g.Ny * g.nkx;
    const int s = blockIdx.x / per;
    const int t = blockIdx.x % per;
    const int y = t / g.nkx;
    const int kz0 = (t % g.nkx) * g.TKx;
    const Ctl& ctl = p.ctl[s];
    if (ctl.done) return;
    if (WHICH == BUF_CHK_CUR && !ctl.need_cur) return;
    if (WHICH == BUF_CHK_PREV && !ctl.need_prev) return;
    const int nk = min(g.TKx, g.Nzh - kz0);
    const int w = nk * C;
    const cx<T>* base = p.K + (long long)s * p.sS + ((long long)y * g.Nzh + kz0) * C;
    const long long xstride = (long long)g.Ny * g.Nzh * C;
    const int tot = g.Nx * w;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int x = e / w, r = e - x * w;
        sl[e] = base[x * xstride + r];
    }
    __syncthreads();
    fft_lines<T, -1, EPT, true, GEN>(sl, LineGeo{w, w, 1, 0, w}, g.ax, p.twx);
    double acc = 0.0;
    for (int f = threadIdx.x; f < g.Nx * nk; f += blockDim.x) {
        const int kx = f / nk, kzl = f - kx * nk;
        const int kz = kz0 + kzl;
        const cx<T>* v = sl + kx * w + kzl * C;
        double a2[C];
#pragma unroll
        for (int c = 0; c < C; ++c) a2[c] = (double)v[c].re * (double)v[c].re + (double)v[c].im * (double)v[c].im;
        if constexpr (C == 6) {
            if (kz >= 1) {
                const double z0 = Elastic::fmap(kz, g.Nz), z1 = Elastic::fmap(y, g.Ny), z2 = Elastic::fmap(kx, g.Nx);
                const double q = a2[0] * z0 * z0 + a2[1] * z1 * z1 + a2[2] * z2 * z2 + a2[3] * ((z1 + z1) + (z2 + z2)) +
                                 a2[4] * ((z0 + z0) + (z2 + z2)) + a2[5] * ((z0 + z0) + (z1 + z1));
                acc += sqrt(fabs(q));
            }
        } else {
            const double z0 = fmap_t(kz, g.Nz), z1 = fmap_t(y, g.Ny), z2 = fmap_t(kx, g.Nx);
            acc += sqrt(a2[0] * z0 * z0 + a2[1] * z1 * z1 + a2[2] * z2 * z2);
        }
        if (kx == 0 && y == 0 && kz == 0) {
            double dd;
            if constexpr (C == 6) dd = sqrt(a2[0] + a2[1] + a2[2] + 2 * a2[3] + 2 * a2[4] + 2 * a2[5]);
            else dd = sqrt(a2[0] + a2[1] + a2[2]);
            p.pchk[((long long)s * 2 + (WHICH == BUF_CHK_CUR ? 0 : 1)) * (g.nbc + 1) + g.nbc] = dd;
        }
    }
    const double bs = block_sum(acc, red);
    if (threadIdx.x == 0) p.pchk[((long long)s * 2 + (WHICH == BUF_CHK_CUR ? 0 : 1)) * (g.nbc + 1) + t] = bs;
}

// ---------------------------------------------------------- finalizers
constexpr int FTHREADS = 256;
// the accelerated scheme's finalisers: one workgroup per solve over up to
// Nx*Ny/TY (z pass) or Ny*Nzh/TK (check x pass) partial sums, 65536 at 512^3
constexpr int ATHREADS = 1024;

// sum of a[0..n) by one workgroup in a fixed order: every thread keeps 8
// partial sums of its strided elements (8 loads in flight, not one chain of
// dependent adds), combined pairwise, then the workgroup's tree
__device__ __forceinline__ double ordered_sum(const double* a, int n, double* red)
{
    double p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0, p6 = 0, p7 = 0;
    const int st = blockDim.x;
    int i = threadIdx.x;
    for (; i + 7 * st < n; i += 8 * st) {
        p0 += a[i];
        p1 += a[i + st];
        p2 += a[i + 2 * st];
        p3 += a[i + 3 * st];
        p4 += a[i + 4 * st];
        p5 += a[i + 5 * st];
        p6 += a[i + 6 * st];
        p7 += a[i + 7 * st];
    }
    for (; i < n; i += st) p0 += a[i];
    return block_sum(((p0 + p1) + (p2 + p3)) + ((p4 + p5) + (p6 + p7)), red);
}

// after the forward-only z pass: sum of the dual field at the start iterate
template <int C>
__global__ void __launch_bounds__(ATHREADS) k_fin0(Geo g, Ctl* ctl, const double* psum)
{
    __shared__ double red[ATHREADS / 64];
    const int s = blockIdx.x;
    Ctl& c = ctl[s];
    const int done = c.done;
    __syncthreads();
    if (done) return;
    const double S = ordered_sum(psum + (long long)s * g.nbz, g.nbz, red);
    if (threadIdx.x == 0) {
        // anchored solves sum e; the anchor's sum completes the iterate's
        if (c.corr) c.sum_prev = S + c.sum_anchor;
        else {
            c.sum_prev = S;
            c.sum_anchor = S;
        }
        c.eq_it = -1;
        c.need_cur = 0;
        c.need_prev = 0;
    }
}

// after the update pass: itNo += 1, diff_sum_field, decide whether the
// equilibrium check is needed (FFT_analysis3D.cpp:392-416, stop rule :785-792)
template <int C>
__global__ void __launch_bounds__(ATHREADS) k_finA(Geo g, Ctl* ctl, const double* psum)
{
    __shared__ double red[ATHREADS / 64];
    const int s = blockIdx.x;
    Ctl& c = ctl[s];
    const int done = c.done;
    __syncthreads();
    if (done) return;
    const double S0 = ordered_sum(psum + (long long)s * g.nbz, g.nbz, red);
    if (threadIdx.x == 0) {
        const double S = c.corr ? S0 + c.sum_anchor : S0;
        c.it += 1;
        c.dsum = fabs((S - c.sum_prev) / c.norm);
        c.dsum_hist[c.it & 15] = c.dsum;
        c.sum_prev = S;
        const int need = (c.it > 1) && (fabs(c.dsum) < c.tol_f);
        c.need_cur = need;
        c.need_prev = need && (c.eq_it != c.it - 1);
        if (!need && c.it > c.cap) c.done = 2;
    }
}

// after the check FFTs: eqVar, diff_avr_eqVar and the stop decision
template <int C>
__global__ void __launch_bounds__(ATHREADS) k_finB(Geo g, Ctl* ctl, const double* pchk)
{
    __shared__ double red[ATHREADS / 64];
    const int s = blockIdx.x;
    Ctl& c = ctl[s];
    const int done = c.done, need_cur = c.need_cur, need_prev = c.need_prev;
    __syncthreads();
    if (done || !need_cur) return;
    const double* pc = pchk + (long long)s * 2 * (g.nbc + 1);
    const double cc = ordered_sum(pc, g.nbc, red);
    double ccp = 0;
    if (need_prev) ccp = ordered_sum(pc + g.nbc + 1, g.nbc, red);
    if (threadIdx.x == 0) {
        const double eq = cc / pc[g.nbc] * g.invN;
        const double eqp = need_prev ? (ccp / pc[2 * g.nbc + 1] * g.invN) : c.eq_val;
        c.eqvar = eq;
        c.deq = fabs(eq - eqp);
        c.eq_val = eq;
        c.eq_it = c.it;
        if (fabs(c.deq) < c.tol_eq) c.done = 1;
        else if (c.it > c.cap) c.done = 2;
        c.need_cur = 0;
        c.need_prev = 0;
    }
}

// --------------------------------------------------------------- utility
template <typename T, int C>
__global__ void k_fill_load(Geo g, Ptr<T> p, int buf)
{
    const long long n = g.N * C;
    for (int s = 0; s < g.nsolve; ++s) {
        T* U = u_buf(p, buf, s);
        const Ctl& c = p.ctl[s];
        for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
            U[i] = T(c.E[i % C]);
    }
}

// FP32 -> FP64 state transfer at the precision switch: the FP64 iterate
// `it` <- the FP32 iterate `it` (solves that are done are skipped)
template <int C>
__global__ void k_promote(Geo g, Ptr<float> p32, Ptr<double> p64, const Ctl* ctl)
{
    const long long n = g.N * C;
    for (int s = 0; s < g.nsolve; ++s) {
        if (ctl[s].done) continue;
        const int it = ctl[s].it;
        const float* src = u_buf(p32, it, s);
        double* dst = u_buf(p64, it, s);
        for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
            dst[i] = (double)src[i];
    }
}

// start of the anchored polish, after one FP64 iteration from the anchor
// A = iterate it-1: r = T(A) - A = u[it] - u[it-1] (FP64 difference, kept
// in FP32), and the correction of iterate `it` is e = r.  Skips solves that
// are done or not anchored.
template <int C>
__global__ void k_anchor(Geo g, Ptr<double> p64, Ptr<float> p32, const Ctl* ctl)
{
    const long long n = g.N * C;
    for (int s = 0; s < g.nsolve; ++s) {
        if (ctl[s].done || !ctl[s].corr) continue;
        const int it = ctl[s].it;
        const double* a = u_buf(p64, it - 1, s);
        const double* b = u_buf(p64, it, s);
        float* e = u_buf(p32, it, s);
        float* r = const_cast<float*>(p32.R) + (long long)s * p32.sU;
        for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
            const float d = (float)(b[i] - a[i]);
            e[i] = d;
            r[i] = d;
        }
    }
}

// means of u and of the dual field v = M(x) u at the final iterate; output
// per solve [2*C] partial sums per block
template <typename T, typename PH>
__global__ void __launch_bounds__(256) k_means(Geo g, Ptr<T> p, double* out, int nblk)
{
    __shared__ double red[256 / 64];
    constexpr int C = PH::C;
    const int s = blockIdx.y;
    const Ctl& ctl = p.ctl[s];
    const bool corr = sizeof(T) == 8 && ctl.corr;
    const T* U = u_buf(p, corr ? ctl.anc_it : ctl.it, s);
    const float* E = corr ? e_buf(p, ctl.it, s) : nullptr;
    const int rv = p.rve_of[s];
    const uint8_t* ph = p.ph + (long long)rv * p.sPh;
    const T* coef = p.coef + (long long)rv * p.sCoef;
    double su[C], sv[C];
#pragma unroll
    for (int c = 0; c < C; ++c) { su[c] = 0; sv[c] = 0; }
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < g.N; v += (long long)gridDim.x * blockDim.x) {
        T u[C], o[C];
#pragma unroll
        for (int c = 0; c < C; ++c) u[c] = U[v * C + c];
        if (E) {
#pragma unroll
            for (int c = 0; c < C; ++c) u[c] += (T)E[v * C + c];
        }
        PH::dual(coef + (int)ph[v] * NCOEF, u, o);
#pragma unroll
        for (int c = 0; c < C; ++c) { su[c] += (double)u[c]; sv[c] += (double)o[c]; }
    }
    for (int c = 0; c < C; ++c) {
        double a = block_sum(su[c], red);
        if (threadIdx.x == 0) out[((long long)s * 2 * C + c) * nblk + blockIdx.x] = a;
        double bsum = block_sum(sv[c], red);
        if (threadIdx.x == 0) out[((long long)s * 2 * C + C + c) * nblk + blockIdx.x] = bsum;
    }
}

// final fields in the reference order: eps/sigma (or gradient/flux) FP64 AoS
template <typename T, typename PH>
__global__ void k_fields(Geo g, Ptr<T> p, int s, double* uo, double* vo)
{
    constexpr int C = PH::C;
    const Ctl& ctl = p.ctl[s];
    const bool corr = sizeof(T) == 8 && ctl.corr;
    const T* U = u_buf(p, corr ? ctl.anc_it : ctl.it, s);
    const float* E = corr ? e_buf(p, ctl.it, s) : nullptr;
    const int rv = p.rve_of[s];
    const uint8_t* ph = p.ph + (long long)rv * p.sPh;
    const T* coef = p.coef + (long long)rv * p.sCoef;
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < g.N; v += (long long)gridDim.x * blockDim.x) {
        T u[C], o[C];
#pragma unroll
        for (int c = 0; c < C; ++c) u[c] = U[v * C + c];
        if (E) {
#pragma unroll
            for (int c = 0; c < C; ++c) u[c] += (T)E[v * C + c];
        }
        PH::dual(coef + (int)ph[v] * NCOEF, u, o);
#pragma unroll
        for (int c = 0; c < C; ++c) {
            uo[v * C + c] = (double)u[c];
            vo[v * C + c] = (double)o[c];
        }
    }
}

} // namespace m3d
