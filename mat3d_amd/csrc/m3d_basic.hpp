// m3d_basic.hpp -- basic (Moulinec-Suquet) scheme with mixed control and the
// J2 / damage constitutive updates: runAnalysis "ep" / "ed",
// src/3d/FFT_analysis3D.cpp:796-853.
//
// Per iteration (one stage = four kernels + two finalisers):
//   zb   : c2r along z of eps^ (6 comps) -> eps;  constitutive() -> sigma
//          (elastoPlastic :640-691 or elastoDamage :619-638, per voxel,
//          history in HBM); FP64 partial sums of every component of sigma and
//          eps (updateE :583-594 needs the means); r2c along z of sigma and
//          eps -> a 12-component half spectrum [sigma^ | eps^]
//   fin1 : itNo, diff_sum_field (:399, with the E of the Green step that made
//          this iterate), updateE -> E for the next Green step
//   y    : forward FFT along y (12 components)
//   xb   : forward FFT along x; eqVar partials from sigma^ (convergenceCheck
//          :351-389 -- the reference transforms sigma once more for it, here it
//          is the spectrum the next Green step uses anyway); Green step
//          t3 = eps^ - Gamma0:sigma^ (multiplyGF :323-338), DC = E N; inverse
//          FFT along x of the 6 output components
//   fin2 : eqVar, diff_avr_eqVar, stop rule (:820-826)
//   y    : inverse FFT along y (6 components)
// The reference medium is always the strain-based Gamma0 with arithmetic
// means of lambda and mu (:517-521).
#pragma once
#include "m3d_kernels.hpp"

namespace m3d {

enum { BEH_EP = 1, BEH_ED = 2 };
// FINAL (specialised kernels only): the fields and state of the last iterate
// of an increment, recomputed from the Green output its z pass read (the
// FP64 iterations do not write them)
enum { BMODE_START = 0, BMODE_ITER = 1, BMODE_RESTART = 2, BMODE_FINAL = 3 };
constexpr int NBSUM = 12; // sums per solve in the z pass: sigma_c, eps_c

template <typename T>
struct BPtr {
    T* eps;          // 6N per solve (sU stride)
    T* sig;          // 6N
    T* sig_old;      // 6N  ("ep")
    T* eps_old;      // 6N
    T* ep;           // N   equivalent plastic strain ("ep") / damage d ("ed")
    T* phi;          // N   yield function of the last constitutive call ("ep")
    T* dep;          // N   plastic multiplier increment ("ep")
    T* seff;         // N   von Mises stress after the return map ("ep")
    T* sig_star;     // 6N  sigma_old - C eps_old, once per increment ("ep", FP64 specialised kernels)
    double* pfin;    // [14][nbz] block sums of the fused FINAL pass ("ep", FP64 specialised kernels)
    long long sU;    // per-solve stride of the 6N fields
    long long sN;    // per-solve stride of the N fields
    cx<T>* S6;       // Green output, 6 components (read by the z pass)
    cx<T>* S12;      // [sigma^ | eps^], 12 components (generic kernels, START)
                     // or sigma^, 6 components (specialised ITER)
    cx<T>* S6w;      // Green output written by the specialised x pass (the other S6 buffer)
    cx<T>* Eh;       // eps^ after the Green step, tile order of the specialised x pass
    long long sS6, sS12;
    const uint8_t* ph;
    const double* mat; // per phase {la, mu, Sy, H, n, P, Y0, 0}
    const cx<T>* twz;
    const cx<T>* twy;
    const cx<T>* twx;
    double* psum;    // [nsolve][NBSUM][nbz]
    double* pchk;    // [nsolve][nbx + 1]
    Ctl* ctl;
};

// von Mises stress and deviator, sigma_eff_calculator :597-606
template <typename T>
__device__ __forceinline__ T vm_eff(const T* s, T* sd)
{
    const T sm = (s[0] + s[1] + s[2]) / T(3);
#pragma unroll
    for (int c = 0; c < 6; ++c) sd[c] = s[c] - (c < 3 ? sm : T(0));
    return sqrt(T(1.5)) * sqrt(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2] + T(2) * sd[3] * sd[3] +
                               T(2) * sd[4] * sd[4] + T(2) * sd[5] * sd[5]);
}

// isotropic C d with C = lambda tr I + 2 mu (tensor shears)
template <typename T>
__device__ __forceinline__ void iso_C(T la, T mu, const T* d, T* o)
{
    const T tr = d[0] + d[1] + d[2];
#pragma unroll
    for (int c = 0; c < 6; ++c) o[c] = T(2) * mu * d[c] + (c < 3 ? la * tr : T(0));
}

// elastoPlastic for one voxel.  Returns false when the Newton iteration for
// the plastic multiplier produces NaN (the reference terminates) or does not
// reach |res| < 1e-5 within 200 steps.
// LIN: every phase has hardening exponent n = 1, where pow(a, 1) = a and
// pow(a, 0) = 1 exactly, so the same arithmetic runs without the pow code
// (and its registers)
// STAR: the trial stress sigma_old + C (eps - eps_old) is formed as
// sig_star + C eps with sig_star = sigma_old - C eps_old kept per increment
// (passed in sig_old; eps_old unused): one 6-vector read instead of two
template <typename T, bool LIN = false, bool STAR = false>
__device__ __forceinline__ bool j2_update(const double* m, const T* eps, const T* eps_old, const T* sig_old, T ep_eff,
                                          T* sig, T& phi, T& dep, T& seff)
{
    const T la = T(m[0]), mu = T(m[1]), Sy = T(m[2]), H = T(m[3]), n = T(m[4]);
    T d[6], c[6], sd[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) d[k] = STAR ? eps[k] : eps[k] - eps_old[k];
    iso_C(la, mu, d, c);
#pragma unroll
    for (int k = 0; k < 6; ++k) sig[k] = sig_old[k] + c[k];
    seff = vm_eff(sig, sd);
    auto pw = [&](T a, T e) {
        if constexpr (LIN) return e == T(0) ? T(1) : a; // e is n or n - 1 with n = 1
        else return pow(a, e);
    };
    phi = seff - Sy - H * pw(ep_eff, n);
    if (phi > T(0)) {
        T x = 0, x_o = 0, res;
        bool ok = false;
        if (sizeof(T) == 4 && n == T(1)) {
            // FP32 phase, linear hardening: the root of
            // seff - 3 mu x - Sy - H (ep + x) = 0 (the reference's first Newton
            // step from x = 0 is this root)
            x = phi / (T(3) * mu + H);
            ok = true;
        } else {
            // FP32: |res| < 1e-5 is below the resolution of stresses in FP32,
            // so the FP32 phase stops at a residual relative to the yield
            // stress; the FP64 polish applies the reference's absolute rule
            const T rtol = sizeof(T) == 4 ? T(1e-6) * (Sy + seff) : T(1e-5);
            for (int iter = 0; iter < 200; ++iter) {
                x = x_o - (seff - T(3) * mu * x_o - Sy - H * pw(ep_eff + x_o, n)) /
                              (-T(3) * mu - H * n * pw(ep_eff + x_o, n - T(1)));
                x_o = x;
                res = seff - T(3) * mu * x - Sy - H * pw(ep_eff + x, n);
                if (isnan(x)) break;
                if (fabs(res) < rtol) { ok = true; break; }
            }
        }
        if (!ok) return false;
        dep = x;
        // sigma -= (1.5 x / seff) C sd; C sd = 2 mu sd for the deviator
        const T f = T(1.5) * x / seff;
        iso_C(la, mu, sd, c);
#pragma unroll
        for (int k = 0; k < 6; ++k) sig[k] -= f * c[k];
        seff = vm_eff(sig, sd);
    } else {
        dep = T(0);
    }
    return true;
}

// elastoDamage for one voxel; d is read (previous iterate) and updated
template <typename T>
__device__ __forceinline__ void damage_update(const double* m, const T* eps, T& d, T* sig)
{
    const T la = T(m[0]), mu = T(m[1]), P = T(m[5]), Y0 = T(m[6]);
    T c[6];
    iso_C(la, mu, eps, c);
    const T Y = sqrt(c[0] * eps[0] + c[1] * eps[1] + c[2] * eps[2] + T(2) * c[3] * eps[3] + T(2) * c[4] * eps[4] +
                     T(2) * c[5] * eps[5]);
    const T d_old = d;
    if (Y > Y0) {
        const T dn = T(1) - exp(-P * (Y - Y0));
        d = dn > d_old ? dn : d_old;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) sig[k] = (T(1) - d) * c[k];
}

// ------------------------------------------------------------------- z pass
// Block: solve s, plane x, lines y0..y0+TY-1; LDS lines l = yl*12 + c
// (c < 6: sigma, c >= 6: eps), each lsz complex.
template <typename T, int BEH, int MODE, bool GEN>
__global__ void __launch_bounds__(ZTHREADS) k_zbasic(Geo g, BPtr<T> p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[ZTHREADS / 64];
    constexpr int C = 12;
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    T* sr = reinterpret_cast<T*>(smem_raw);
    const int per = g.Nx * g.nty;
    const int s = blockIdx.x / per;
    const int t = blockIdx.x % per;
    const int x = t / g.nty;
    const int y0 = (t % g.nty) * g.TY;
    Ctl& ctl = p.ctl[s];
    if (ctl.done) return;
    const int Nz = g.Nz, Nzh = g.Nzh, lsz = g.lsz, TY = g.TY;
    const long long vox0 = ((long long)x * g.Ny + y0) * Nz;
    const int tid = threadIdx.x, nthr = blockDim.x;

    T* eps = p.eps + (long long)s * p.sU + vox0 * 6;
    T* sig = p.sig + (long long)s * p.sU + vox0 * 6;
    const uint8_t* ph = p.ph + vox0;

    if (MODE == BMODE_ITER) {
        // Green output (6 comps) -> lines yl*12 + 6 + c
        const cx<T>* src = p.S6 + (long long)s * p.sS6 + ((long long)x * g.Ny + y0) * Nzh * 6;
        const int tot = TY * Nzh * 6;
        for (int e = tid; e < tot; e += nthr) {
            const int yl = e / (Nzh * 6);
            const int r = e - yl * Nzh * 6;
            const int kz = r / 6, c = r - kz * 6;
            sl[(yl * C + 6 + c) * lsz + kz] = src[e];
        }
        __syncthreads();
        const LineGeo lg6{TY * 6, 6, lsz, C * lsz, 1};
        cx<T>* s6 = sl + 6 * lsz;
        if (g.zeven) {
            const int L = Nz / 2;
            const int half = L / 2 + 1;
            for (int e = tid; e < TY * 6 * half; e += nthr) {
                const int l = e / half, k = e - l * half;
                cx<T>* ln = s6 + lg6.off(l);
                if (k == 0) {
                    const T x0 = ln[0].re, xl = ln[L].re;
                    ln[0] = cx<T>{x0 + xl, x0 - xl};
                } else {
                    const int kk = L - k;
                    const cx<T> Xk = ln[k], Xc = ln[kk];
                    cx<T> w = p.twz[k];
                    w.im = -w.im;
                    ln[k] = (Xk + cconj(Xc)) + rot90<1>(cmul(w, Xk - cconj(Xc)));
                    if (kk != k) {
                        cx<T> w2 = p.twz[kk];
                        w2.im = -w2.im;
                        ln[kk] = (Xc + cconj(Xk)) + rot90<1>(cmul(w2, Xc - cconj(Xk)));
                    }
                }
            }
            __syncthreads();
            fft_lines<T, 1, EPT, false, GEN>(s6, lg6, g.az, p.twz);
        } else {
            for (int e = tid; e < TY * 6 * Nzh; e += nthr) {
                const int l = e / Nzh, k = e - l * Nzh;
                cx<T>* ln = s6 + lg6.off(l);
                if (k == 0) ln[0].im = T(0);
                else ln[Nz - k] = cconj(ln[k]);
            }
            __syncthreads();
            fft_lines<T, 1, EPT, false, GEN>(s6, lg6, g.az, p.twz);
            // compact real parts into the real view of each line
            for (int l = 0; l < TY * 6; ++l) {
                cx<T>* ln = s6 + lg6.off(l);
                T* rl = reinterpret_cast<T*>(ln);
                T v[EPT];
#pragma unroll
                for (int i = 0; i < EPT; ++i) {
                    const int z = tid + i * nthr;
                    if (z < Nz) v[i] = ln[z].re;
                }
                __syncthreads();
#pragma unroll
                for (int i = 0; i < EPT; ++i) {
                    const int z = tid + i * nthr;
                    if (z < Nz) rl[z] = v[i];
                }
                __syncthreads();
            }
        }
    }

    // constitutive update per voxel
    double acc[NBSUM];
#pragma unroll
    for (int q = 0; q < NBSUM; ++q) acc[q] = 0.0;
    bool bad = false;
    const int nv = TY * Nz;
    for (int v = tid; v < nv; v += nthr) {
        const int yl = v / Nz, z = v - yl * Nz;
        const int k = ph[v];
        const double* m = p.mat + k * 8;
        T e[6], sg[6];
        if (MODE == BMODE_ITER) {
            const T invN = T(1); // the Green step already divided by N
#pragma unroll
            for (int c = 0; c < 6; ++c) e[c] = sr[(yl * C + 6 + c) * 2 * lsz + z] * invN;
#pragma unroll
            for (int c = 0; c < 6; ++c) eps[(long long)v * 6 + c] = e[c];
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) e[c] = eps[(long long)v * 6 + c];
        }
        const long long vg = vox0 + v;
        if (BEH == BEH_EP) {
            T so[6], eo[6], phi, dep, seff;
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                so[c] = p.sig_old[(long long)s * p.sU + vg * 6 + c];
                eo[c] = p.eps_old[(long long)s * p.sU + vg * 6 + c];
            }
            const T epe = p.ep[(long long)s * p.sN + vg];
            if (!j2_update<T>(m, e, eo, so, epe, sg, phi, dep, seff)) bad = true;
            p.phi[(long long)s * p.sN + vg] = phi;
            p.dep[(long long)s * p.sN + vg] = dep;
            p.seff[(long long)s * p.sN + vg] = seff;
        } else {
            T d = p.ep[(long long)s * p.sN + vg];
            damage_update<T>(m, e, d, sg);
            p.ep[(long long)s * p.sN + vg] = d;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            sig[(long long)v * 6 + c] = sg[c];
            acc[c] += (double)sg[c];
            acc[6 + c] += (double)e[c];
            sr[(yl * C + c) * 2 * lsz + z] = sg[c];
            sr[(yl * C + 6 + c) * 2 * lsz + z] = e[c];
        }
    }
    if (bad) ctl.err = 1;
#pragma unroll
    for (int q = 0; q < NBSUM; ++q) {
        const double bs = block_sum(acc[q], red);
        if (tid == 0) p.psum[((long long)s * NBSUM + q) * g.nbz + t] = bs;
    }
    __syncthreads();

    // forward r2c of all 12 components
    const LineGeo lg{TY * C, TY * C, lsz, 0, 1};
    if (g.zeven) {
        fft_lines<T, -1, EPT, false, GEN>(sl, lg, g.az, p.twz);
        const int L = Nz / 2;
        const int half = L / 2 + 1;
        for (int e = tid; e < TY * C * half; e += nthr) {
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
                const cx<T> O = cscale(rot90<-1>(zk - zc), T(0.5));
                ln[k] = E + cmul(p.twz[k], O);
                if (kk != k) ln[kk] = cconj(E) + cmul(p.twz[kk], cconj(O));
            }
        }
        __syncthreads();
    } else {
        for (int l = 0; l < TY * C; ++l) {
            cx<T>* ln = sl + l * lsz;
            const T* rl = reinterpret_cast<const T*>(ln);
            T v[EPT];
#pragma unroll
            for (int i = 0; i < EPT; ++i) {
                const int z = tid + i * nthr;
                if (z < Nz) v[i] = rl[z];
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < EPT; ++i) {
                const int z = tid + i * nthr;
                if (z < Nz) ln[z] = cx<T>{v[i], T(0)};
            }
            __syncthreads();
        }
        fft_lines<T, -1, EPT, false, GEN>(sl, lg, g.az, p.twz);
    }
    cx<T>* dst = p.S12 + (long long)s * p.sS12 + ((long long)x * g.Ny + y0) * Nzh * C;
    const int tot = TY * Nzh * C;
    for (int e = tid; e < tot; e += nthr) {
        const int yl = e / (Nzh * C);
        const int r = e - yl * Nzh * C;
        const int kz = r / C, c = r - kz * C;
        dst[e] = sl[(yl * C + c) * lsz + kz];
    }
}

// --------------------------------------------------------------- x pass
// Block: solve s, row y, kz tile; LDS [x][kzl*12 + c] (12 comps in, 6 out).
template <typename T, bool GEN>
__global__ void __launch_bounds__(XYTHREADS) k_xbasic(Geo g, BPtr<T> p)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[XYTHREADS / 64];
    constexpr int C = 12;
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    const int per = g.Ny * g.nkx;
    const int s = blockIdx.x / per;
    const int t = blockIdx.x % per;
    const int y = t / g.nkx;
    const int kz0 = (t % g.nkx) * g.TKx;
    const Ctl& ctl = p.ctl[s];
    if (ctl.done) return;
    const int nk = min(g.TKx, g.Nzh - kz0);
    const int w = nk * C;
    const cx<T>* base = p.S12 + (long long)s * p.sS12 + ((long long)y * g.Nzh + kz0) * C;
    const long long xstride = (long long)g.Ny * g.Nzh * C;
    const int tot = g.Nx * w;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int xx = e / w, r = e - xx * w;
        sl[e] = base[xx * xstride + r];
    }
    __syncthreads();
    fft_lines<T, -1, EPT, true, GEN>(sl, LineGeo{w, w, 1, 0, w}, g.ax, p.twx);
    const T la0 = T(ctl.la0), mu0 = T(ctl.mu0);
    double acc = 0.0;
    for (int f = threadIdx.x; f < g.Nx * nk; f += blockDim.x) {
        const int kx = f / nk, kzl = f - kx * nk;
        const int kz = kz0 + kzl;
        cx<T>* v = sl + kx * w + kzl * C;
        // eqVar partial from sigma^ (kz >= 1), dd at DC
        double a2[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) a2[c] = (double)v[c].re * (double)v[c].re + (double)v[c].im * (double)v[c].im;
        if (kz >= 1) {
            const double z0 = Elastic::fmap(kz, g.Nz), z1 = Elastic::fmap(y, g.Ny), z2 = Elastic::fmap(kx, g.Nx);
            const double q = a2[0] * z0 * z0 + a2[1] * z1 * z1 + a2[2] * z2 * z2 + a2[3] * ((z1 + z1) + (z2 + z2)) +
                             a2[4] * ((z0 + z0) + (z2 + z2)) + a2[5] * ((z0 + z0) + (z1 + z1));
            acc += sqrt(fabs(q));
        }
        if (kx == 0 && y == 0 && kz == 0) {
            p.pchk[(long long)s * (g.nbx + 1) + g.nbx] = sqrt(a2[0] + a2[1] + a2[2] + 2 * a2[3] + 2 * a2[4] + 2 * a2[5]);
#pragma unroll
            for (int c = 0; c < 6; ++c) v[c] = cx<T>{T(ctl.E[c]), T(0)};
        } else {
            // t3 = eps^ - Gamma0:sigma^ (Nyquist planes: Hermitian part of the Gamma0 term)
            int fz = Elastic::fmap(kz, g.Nz), fy = Elastic::fmap(y, g.Ny), fxx = Elastic::fmap(kx, g.Nx);
            T xi[3] = {T(fz), T(fy), T(fxx)};
            cx<T> sg[6], o[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) sg[c] = v[c];
            green_elastic<T>(xi, sg, o, 1, la0, mu0);
            const bool plane = (kz == 0) || (g.zeven && kz == g.Nz / 2);
            if (plane) {
                const int px = (g.Nx - kx) % g.Nx, py = (g.Ny - y) % g.Ny;
                const int gz = Elastic::fmap(kz, g.Nz), gy = Elastic::fmap(py, g.Ny), gx = Elastic::fmap(px, g.Nx);
                if (gz != -fz || gy != -fy || gx != -fxx) {
                    T xp[3] = {T(gz), T(gy), T(gx)};
                    cx<T> o2[6];
                    green_elastic<T>(xp, sg, o2, 1, la0, mu0);
#pragma unroll
                    for (int c = 0; c < 6; ++c) o[c] = cscale(o[c] + o2[c], T(0.5));
                }
            }
            const T sc = T(g.invN);
#pragma unroll
            for (int c = 0; c < 6; ++c) v[c] = cscale(v[6 + c] - o[c], sc);
        }
    }
    const double bs = block_sum(acc, red);
    if (threadIdx.x == 0) p.pchk[(long long)s * (g.nbx + 1) + t] = bs;
    __syncthreads();
    fft_lines<T, 1, EPT, true, GEN>(sl, LineGeo{nk * 6, 6, 1, C, w}, g.ax, p.twx);
    cx<T>* ob = p.S6 + (long long)s * p.sS6 + ((long long)y * g.Nzh + kz0) * 6;
    const long long ostride = (long long)g.Ny * g.Nzh * 6;
    const int tot6 = g.Nx * nk * 6;
    for (int e = threadIdx.x; e < tot6; e += blockDim.x) {
        const int xx = e / (nk * 6), r = e - xx * nk * 6;
        const int kzl = r / 6, c = r - kzl * 6;
        ob[xx * ostride + r] = sl[xx * w + kzl * C + c];
    }
}

// ----------------------------------------------------------- finalisers
// S0 u for the isotropic reference medium: S0 = (1/3K0) J + (1/2mu0) Kdev
__device__ __forceinline__ void s0_apply(double la0, double mu0, const double* u, double* o)
{
    const double K0 = la0 + 2.0 * mu0 / 3.0;
    const double a = 1.0 / (3.0 * K0), b = 1.0 / (2.0 * mu0);
    const double t = (u[0] + u[1] + u[2]) / 3.0;
    for (int c = 0; c < 6; ++c) o[c] = b * u[c] + (c < 3 ? (a - b) * t : 0.0);
}

__device__ __forceinline__ double prod6(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2] + 2 * a[3] * b[3] + 2 * a[4] * b[4] + 2 * a[5] * b[5];
}

// updateE, FFT_analysis3D.cpp:583-594, from the means of sigma and eps
__device__ __forceinline__ void update_E(Ctl& c, const double* sbar, const double* ebar)
{
    double S0s[6], y[6], x[6];
    s0_apply(c.la0, c.mu0, sbar, S0s);
    for (int i = 0; i < 6; ++i) x[i] = S0s[i] - ebar[i];
    s0_apply(c.la0, c.mu0, c.X0, y);
    const double k = (c.dEsum * c.incNo + prod6(x, c.X0)) / prod6(y, c.X0);
    for (int i = 0; i < 6; ++i) c.E[i] = k * y[i] - S0s[i] + ebar[i];
}

// MODE_START: after the increment's first constitutive call (:809):
//   sum_field_old = sum(sigma^0); E <- updateE(sigma^0, eps^0); itNo = 0
// MODE_ITER : itNo += 1, diff_sum_field with the E that produced this
//   iterate (:399), sum_field_old for the next check, E <- updateE
// MODE_RESTART: the iterate itNo recomputed in FP64 from its FP32 strain
//   (end of an FP32 phase): sum_field_old and E as in MODE_ITER, itNo kept;
//   the check after it records eqVar without a stop decision
//
// The z pass leaves NBSUM partial sums per workgroup (nbz of them per sum,
// 18 432 at 192^3).  k_bsum_parts adds them in BSPLIT contiguous shares, one
// workgroup per (share, sum), and k_bfin1 adds the BSPLIT shares: a fixed
// summation order, so the result does not depend on timing.
constexpr int BSPLIT = 64;

template <int U = 0> // a template so that every unit including this header may instantiate it
__global__ void __launch_bounds__(256) k_bsum_parts(Geo g, const Ctl* ctl, const double* psum, double* part)
{
    __shared__ double red[256 / 64];
    const int s = blockIdx.z, q = blockIdx.y, k = blockIdx.x;
    if (ctl[s].done) return;
    const int per = (g.nbz + BSPLIT - 1) / BSPLIT;
    const int b = k * per, e = min(g.nbz, b + per);
    const double* a = psum + ((long long)s * NBSUM + q) * g.nbz;
    double v = 0;
    for (int i = b + threadIdx.x; i < e; i += blockDim.x) v += a[i];
    v = block_sum(v, red);
    if (threadIdx.x == 0) part[((long long)s * NBSUM + q) * BSPLIT + k] = v;
}

template <int MODE>
__global__ void __launch_bounds__(FTHREADS) k_bfin1(Geo g, Ctl* ctl, const double* part)
{
    __shared__ double red[FTHREADS / 64];
    const int s = blockIdx.x;
    Ctl& c = ctl[s];
    const int done = c.done;
    __syncthreads();
    if (done) return;
    double m[NBSUM];
    for (int q = 0; q < NBSUM; ++q) m[q] = ordered_sum(part + ((long long)s * NBSUM + q) * BSPLIT, BSPLIT, red);
    if (threadIdx.x == 0) {
        double sbar[6], ebar[6], S = 0;
        for (int q = 0; q < 6; ++q) {
            sbar[q] = m[q] * g.invN;
            ebar[q] = m[6 + q] * g.invN;
            S += m[q];
        }
        if (MODE == BMODE_ITER) {
            c.it += 1;
            double es = 0;
            for (int q = 0; q < 6; ++q) es += c.E[q];
            c.dsum = fabs((S - c.sum_prev) / (c.nc * es));
            c.first = 0;
        } else if (MODE == BMODE_RESTART) {
            c.first = 2;
        } else {
            c.it = 0;
            c.first = 1;
        }
        c.sum_prev = S;
        update_E(c, sbar, ebar);
    }
}

// after the x pass: eqVar of this iterate, diff_avr_eqVar, stop rule (:820-826)
template <int U = 0> // a template so that every unit including this header may instantiate it
__global__ void __launch_bounds__(FTHREADS) k_bfin2(Geo g, Ctl* ctl, const double* pchk)
{
    __shared__ double red[FTHREADS / 64];
    const int s = blockIdx.x;
    Ctl& c = ctl[s];
    const int done = c.done;
    __syncthreads();
    if (done) return;
    const double* pc = pchk + (long long)s * (g.nbx + 1);
    const double cc = ordered_sum(pc, g.nbx, red);
    if (threadIdx.x == 0) {
        if (c.first == 1) return; // the stage after the increment start checks nothing
        const double eq = cc / pc[g.nbx] * g.invN;
        if (c.first == 2) {
            // restart stage: eqVar of the recomputed iterate, no stop decision
            c.eq_val = eq;
            c.eqvar = eq;
            c.first = 0;
            return;
        }
        c.deq = fabs(eq - c.eq_val);
        c.eq_val = eq;
        c.eqvar = eq;
        if (c.err) c.done = 3;
        else if (c.it > 1 && fabs(c.deq) < c.tol_eq && fabs(c.dsum) < c.tol_f) c.done = 1;
        else if (c.it > c.cap) c.done = 2;
    }
}

// sums for export_homogen_var (:722-746): eps_c, sigma_c, seff|d, ep (14 per solve)
template <typename T, int BEH>
__global__ void __launch_bounds__(256) k_bsums(Geo g, BPtr<T> p, double* out, int nblk)
{
    __shared__ double red[256 / 64];
    const int s = blockIdx.y;
    double a[14];
    for (int q = 0; q < 14; ++q) a[q] = 0;
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < g.N; v += (long long)gridDim.x * blockDim.x) {
        for (int c = 0; c < 6; ++c) {
            a[c] += (double)p.eps[(long long)s * p.sU + v * 6 + c];
            a[6 + c] += (double)p.sig[(long long)s * p.sU + v * 6 + c];
        }
        if (BEH == BEH_EP) {
            a[12] += (double)p.seff[(long long)s * p.sN + v];
            a[13] += (double)p.ep[(long long)s * p.sN + v];
        } else {
            a[12] += (double)p.ep[(long long)s * p.sN + v];
        }
    }
    for (int q = 0; q < 14; ++q) {
        const double b = block_sum(a[q], red);
        if (threadIdx.x == 0) out[((long long)s * 14 + q) * nblk + blockIdx.x] = b;
    }
}

// update_old_Variable, FFT_analysis3D.cpp:693-719
template <typename T, int BEH>
__global__ void k_bupdate(Geo g, BPtr<T> p, int nsolve)
{
    // element-wise over the 6N tensors and the N scalars (consecutive
    // threads, consecutive values)
    const long long st = (long long)gridDim.x * blockDim.x;
    const long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    for (int s = 0; s < nsolve; ++s) {
        const long long b6 = (long long)s * p.sU, b1 = (long long)s * p.sN;
        for (long long i = i0; i < 6 * g.N; i += st) {
            if (BEH == BEH_EP) p.sig_old[b6 + i] = p.sig[b6 + i];
            const T e = p.eps[b6 + i], tmp = p.eps_old[b6 + i];
            p.eps_old[b6 + i] = e;
            p.eps[b6 + i] = e + (e - tmp);
        }
        if (BEH == BEH_EP)
            for (long long v = i0; v < g.N; v += st)
                if (p.phi[b1 + v] > T(0)) p.ep[b1 + v] += p.dep[b1 + v];
    }
}

// the 14 export sums of the fused FINAL pass: its block sums pfin[q][nbz],
// added in nblk contiguous shares (fixed order), out[q][b] as k_bsums writes
template <int U = 0> // a template so that every unit including this header may instantiate it
__global__ void __launch_bounds__(256) k_bfsums(Geo g, const double* pfin, double* out, int nblk)
{
    __shared__ double red[256 / 64];
    const int q = blockIdx.y, b = blockIdx.x;
    const int per = (g.nbz + nblk - 1) / nblk;
    const int i0 = b * per, i1 = min(g.nbz, i0 + per);
    double v = 0;
    for (int i = i0 + threadIdx.x; i < i1; i += blockDim.x) v += pfin[(long long)q * g.nbz + i];
    v = block_sum(v, red);
    if (threadIdx.x == 0) out[(long long)q * nblk + b] = v;
}

// sigma_star = sigma_old - C eps_old for the next increment's start stage on
// the specialised FP64 "ep" kernels, after k_bupdate (one solve)
template <typename T>
__global__ void k_bstar(Geo g, BPtr<T> p)
{
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < g.N; v += (long long)gridDim.x * blockDim.x) {
        const double* mt = p.mat + p.ph[v] * 8;
        T eo[6], ce[6];
        for (int c = 0; c < 6; ++c) eo[c] = p.eps_old[v * 6 + c];
        iso_C(T(mt[0]), T(mt[1]), eo, ce);
        for (int c = 0; c < 6; ++c) p.sig_star[v * 6 + c] = p.sig_old[v * 6 + c] - ce[c];
    }
}

// copy fields between the FP64 and FP32 states of a solve; what: 1 eps,
// 2 sig, 4 sig_old, 8 eps_old, 16 ep|d, 32 phi, dep and seff, 64 the Green
// output spectrum S6 (6 x H complex)
enum { BC_EPS = 1, BC_SIG = 2, BC_SIGOLD = 4, BC_EPSOLD = 8, BC_EP = 16, BC_PHI = 32, BC_S6 = 64 };
template <typename TS, typename TD>
__global__ void k_bcopy(Geo g, BPtr<TS> a, BPtr<TD> b, int what)
{
    const long long n6 = 6 * g.N, n1 = g.N, h6 = 6 * g.H;
    const long long i0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, st = (long long)gridDim.x * blockDim.x;
    for (long long i = i0; i < n6; i += st) {
        if (what & BC_EPS) b.eps[i] = TD(a.eps[i]);
        if (what & BC_SIG) b.sig[i] = TD(a.sig[i]);
        if (what & BC_SIGOLD) b.sig_old[i] = TD(a.sig_old[i]);
        if (what & BC_EPSOLD) b.eps_old[i] = TD(a.eps_old[i]);
    }
    for (long long i = i0; i < n1; i += st) {
        if (what & BC_EP) b.ep[i] = TD(a.ep[i]);
        if (what & BC_PHI) {
            b.phi[i] = TD(a.phi[i]);
            b.dep[i] = TD(a.dep[i]);
            b.seff[i] = TD(a.seff[i]);
        }
    }
    if (what & BC_S6)
        for (long long i = i0; i < h6; i += st) b.S6[i] = cx<TD>{TD(a.S6[i].re), TD(a.S6[i].im)};
}

template <typename T>
__global__ void k_binit(Geo g, BPtr<T> p, int nsolve, int beh)
{
    for (int s = 0; s < nsolve; ++s) {
        const Ctl& c = p.ctl[s];
        for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < g.N; v += (long long)gridDim.x * blockDim.x) {
            const long long i6 = (long long)s * p.sU + v * 6, i1 = (long long)s * p.sN + v;
            for (int k = 0; k < 6; ++k) {
                p.eps[i6 + k] = T(c.E[k]); // setLoad(dE, X0) (:799)
                p.eps_old[i6 + k] = T(0);
                if (beh == BEH_EP) p.sig_old[i6 + k] = T(0);
                if (beh == BEH_EP && p.sig_star) p.sig_star[i6 + k] = T(0); // sigma_old - C eps_old
            }
            p.ep[i1] = beh == BEH_EP ? T(.000000001236547) : T(0); // :67-69 / :89-90
        }
    }
}

} // namespace m3d
