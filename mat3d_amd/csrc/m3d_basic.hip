// m3d_basic.hip -- compile-time specialised kernels of the basic scheme with
// mixed control and the J2 / damage updates (runAnalysis "ep" / "ed",
// src/3d/FFT_analysis3D.cpp:796-853), in FP64 and FP32.  The host driver is BasicRunner
// (m3d_basic_host.hpp); it calls the launchers at the end of this file for
// grids whose line lengths have specialised kernels and the runtime-planned
// kernels of m3d_basic.hpp otherwise.
#include "m3d_basic_fast.hpp"

#include <stdexcept>

namespace m3d {
namespace bfast {

// component-count tags for the y pass
struct Comp12 {
    static constexpr int C = 12;
};
struct Comp6 {
    static constexpr int C = 6;
};

// ------------------------------------------------------------------ z pass
// Block: plane x, lines y0..y0+TY-1.  Slot m of the half-length complex FFT
// holds voxels (2m, 2m+1) of its line in real space.
//   START: eps from HBM, the constitutive update (elastoPlastic :640-691 or
//     elastoDamage :619-638) -> sigma, r2c of the 12 lines -> [sigma^ | eps^]
//     in S12 (12 components per kz); LDS row k holds W = 12 TY columns,
//     column yl*12 + c: c < 6 sigma, c >= 6 eps.
//   ITER: eps^ (Green output, 6 comps, S6) -> c2r -> eps -> sigma, written
//     over eps in LDS (W = 6 TY columns, column yl*6 + c), r2c of the 6
//     lines -> sigma^ in S12 with 6 components per kz.  eps^ itself is not
//     transformed again: the x pass keeps the spectrum of the strain it
//     produced (the forward transform of the c2r of a Hermitian spectrum is
//     that spectrum), so FFT_epsilon (:813) costs nothing here.
//   FINAL: the ITER pass of the increment's last iterate without the r2c,
//     storing eps, sigma and the return-map state: the FP64 ITER pass keeps
//     them out of HBM (only that iterate needs them, update_old_Variable
//     :693-719, export_homogen_var :722-746), and its Green output is still
//     in the S6 buffer it was read from (the x pass writes the other one).
// All modes: FP64 partial sums of sigma and eps (updateE :583-594).  The
// damage variable is written every iteration: the next one compares with it.
// The J2 trial stress sigma_old + C (eps - eps_old) is formed as
// sigma_star + C eps with sigma_star = sigma_old - C eps_old (FP64): every
// stage reads 48 bytes of history per voxel instead of 96.  FINAL (FP64 "ep")
// also ends the increment: it writes the export sums (export_homogen_var
// :722-746) and applies update_old_Variable (:693-719) -- eps_p += d_eps_p
// where phi > 0, sigma_old = sigma, eps_old = eps, eps += eps - eps_old --
// and the next sigma_star, so no other pass reads the fields for them.
template <typename T, int BEH, int MODE, int LZ, int TY, bool LIN = false>
__global__ void __launch_bounds__(ZNT) M3D_BZ_ATTR kzb(Geo g, BPtr<T> p)
{
    constexpr bool FIN = MODE == BMODE_FINAL;
    constexpr bool START = MODE == BMODE_START;
    constexpr bool STORE = !(MODE == BMODE_ITER && sizeof(T) == 8);
    // FP64 "ep": every stage reads sigma_star = sigma_old - C eps_old in place
    // of sigma_old and eps_old; FINAL ends the increment (FUSE)
    constexpr bool STAR = sizeof(T) == 8 && BEH == BEH_EP;
    constexpr bool FUSE = FIN && STAR;
    constexpr int C = START ? 12 : 6; // LDS columns per line
    constexpr int EO = START ? 6 : 0; // column of eps_0 within a line
    constexpr int W = TY * C;
    constexpr int WS = W + M3D_BZ_PAD; // LDS row stride (complex values)
    constexpr int NZ = 2 * LZ;
    constexpr int NZH = LZ + 1;
    constexpr int NT = ZNT;
    constexpr int NW = NT / 64;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[NBSUM + 2][NW];
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* tw = sl + NZH * WS;

    const int per = g.Nx * (g.Ny / TY);
    const int t = fast::xcd_remap(blockIdx.x, gridDim.x);
    const int x = t / (g.Ny / TY);
    const int y0 = (t - x * (g.Ny / TY)) * TY;
    Ctl& ctl = p.ctl[0];
    if (ctl.done || t >= per) return;
    const int tid = threadIdx.x;
    const long long vox0 = ((long long)x * g.Ny + y0) * NZ;
    T* eps = p.eps + vox0 * 6;
    T* sig = p.sig + vox0 * 6;
    const uint8_t* ph = p.ph + vox0;
    for (int i = tid; i < NZ; i += NT) tw[i] = p.twz[i];
    // one voxel pair per thread; the iterations (PRE) load the pair's J2
    // history and phases before the transforms, so that they arrive during them
    static_assert(TY * LZ <= NT, "one voxel pair per thread");
    constexpr bool PRE = STAR;
    T hs[12], hep[2];
    int hph[2] = {0, 0};
    if (PRE && tid < TY * LZ) {
        const int yl = tid / LZ, m = tid - yl * LZ;
        const int v0 = yl * NZ + 2 * m;
        fast::vload<T, 12>(p.sig_star + (vox0 + v0) * 6, hs);
        hep[0] = p.ep[vox0 + v0];
        hep[1] = p.ep[vox0 + v0 + 1];
        hph[0] = ph[v0];
        hph[1] = ph[v0 + 1];
    }

    if constexpr (!START) {
        // the TY lines of eps^ are contiguous in S6: [yl][kz][c]
        const cx<T>* src = p.S6 + ((long long)x * g.Ny + y0) * NZH * 6;
        for (int e = tid; e < TY * NZH * 6; e += NT) {
            const int yl = e / (NZH * 6);
            const int r = e - yl * (NZH * 6);
            const int kz = r / 6, c = r - kz * 6;
            sl[kz * WS + yl * 6 + c] = src[e];
        }
        __syncthreads();
        // c2r pre-processing of the eps lines; kz = 0 and Nz/2 keep real parts
        constexpr int half = LZ / 2 + 1;
        for (int e = tid; e < W * half; e += NT) {
            const int l = e % W, k = e / W;
            if (k == 0) {
                const T x0 = sl[l].re, xl = sl[LZ * WS + l].re;
                sl[l] = cx<T>{x0 + xl, x0 - xl};
            } else {
                const int kk = LZ - k;
                const cx<T> Xk = sl[k * WS + l], Xc = sl[kk * WS + l];
                cx<T> w = tw[k];
                w.im = -w.im;
                sl[k * WS + l] = (Xk + cconj(Xc)) + rot90<1>(cmul(w, Xk - cconj(Xc)));
                if (kk != k) {
                    cx<T> w2 = tw[kk];
                    w2.im = -w2.im;
                    sl[kk * WS + l] = (Xc + cconj(Xk)) + rot90<1>(cmul(w2, Xc - cconj(Xk)));
                }
            }
        }
        __syncthreads();
        fast::fft<T, LZ, W, NT, 1, 2, 0, 1, WS>(sl, tw);
    } else {
        __syncthreads();
    }

    // constitutive update per voxel pair
    double acc[NBSUM + 2]; // sigma, eps; FUSE: sigma_eff, eps_p
#pragma unroll
    for (int q = 0; q < NBSUM + 2; ++q) acc[q] = 0.0;
    bool bad = false;
    for (int q = tid; q < TY * LZ; q += NT) {
        const int yl = q / LZ, m = q - yl * LZ;
        const int v0 = yl * NZ + 2 * m;
        T e2[12], s2[12];
        if constexpr (!START) {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const cx<T> u = sl[m * WS + yl * C + c];
                e2[c] = u.re;
                e2[6 + c] = u.im;
            }
            if (STORE && !FUSE) fast::vstore<T, 12>(eps + (long long)v0 * 6, e2);
        } else {
            fast::vload<T, 12>(eps + (long long)v0 * 6, e2);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const long long vg = vox0 + v0 + h;
            const double* mt = p.mat + (PRE ? hph[h] : ph[v0 + h]) * 8;
            T* e = e2 + 6 * h;
            T* sg = s2 + 6 * h;
            if (BEH == BEH_EP) {
                T so[6], eo[6], phi, dep, seff;
                if constexpr (PRE) {
                    // sigma_star = sigma_old - C eps_old of this increment
#pragma unroll
                    for (int c = 0; c < 6; ++c) so[c] = hs[6 * h + c];
                    if (!j2_update<T, LIN, true>(mt, e, eo, so, hep[h], sg, phi, dep, seff)) bad = true;
                } else {
                    fast::vload<T, 6>(p.sig_old + vg * 6, so);
                    fast::vload<T, 6>(p.eps_old + vg * 6, eo);
                    if (!j2_update<T, LIN>(mt, e, eo, so, p.ep[vg], sg, phi, dep, seff)) bad = true;
                }
                if constexpr (FUSE) {
                    // export sums of this iterate, then update_old_Variable
                    acc[NBSUM] += (double)seff;
                    acc[NBSUM + 1] += (double)hep[h];
                    p.seff[vg] = seff;
                    p.ep[vg] = phi > T(0) ? hep[h] + dep : hep[h];
                    T eprev[6], en[6], ek[6], sk[6], st[6], ce[6];
                    fast::vload<T, 6>(p.eps_old + vg * 6, eprev);
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        ek[c] = e[c];
                        sk[c] = sg[c];
                        en[c] = e[c] + (e[c] - eprev[c]);
                    }
                    iso_C(T(mt[0]), T(mt[1]), ek, ce);
#pragma unroll
                    for (int c = 0; c < 6; ++c) st[c] = sk[c] - ce[c];
                    fast::vstore<T, 6>(p.eps + vg * 6, en);
                    fast::vstore<T, 6>(p.eps_old + vg * 6, ek);
                    fast::vstore<T, 6>(p.sig_old + vg * 6, sk);
                    fast::vstore<T, 6>(p.sig_star + vg * 6, st);
                } else if (STORE) {
                    p.phi[vg] = phi;
                    p.dep[vg] = dep;
                    p.seff[vg] = seff;
                }
            } else {
                T d = p.ep[vg];
                damage_update<T>(mt, e, d, sg);
                p.ep[vg] = d;
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                acc[c] += (double)sg[c];
                acc[6 + c] += (double)e[c];
            }
        }
        if (STORE) fast::vstore<T, 12>(sig + (long long)v0 * 6, s2);
        if (FIN) continue;
        // sigma over the slots this thread read eps from (ITER), or beside eps (START)
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            sl[m * WS + yl * C + c] = cx<T>{s2[c], s2[6 + c]};
            if (START) sl[m * WS + yl * C + EO + c] = cx<T>{e2[c], e2[6 + c]};
        }
    }
    if (bad) ctl.err = 1;
    if (FIN && !FUSE) return;
    // fixed-order block sums of the 12 means (FUSE: the 14 export sums, in
    // the order eps, sigma, sigma_eff, eps_p)
    {
        constexpr int NQ = FUSE ? NBSUM + 2 : NBSUM;
        const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            const double v = wave_sum(acc[q]);
            if (lane == 0) red[q][wid] = v;
        }
        __syncthreads();
        if (tid < NQ) {
            double r = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) r += red[tid][w];
            if constexpr (FUSE) {
                const int o = tid < 6 ? tid + 6 : (tid < 12 ? tid - 6 : tid); // sigma_c, eps_c -> eps_c, sigma_c
                p.pfin[(long long)o * g.nbz + t] = r;
            } else {
                p.psum[(long long)tid * g.nbz + t] = r;
            }
        }
    }
    if constexpr (FUSE) return;
    __syncthreads();

    // r2c of the W lines
    fast::fft<T, LZ, W, NT, -1, 2, 0, 1, WS>(sl, tw);
    {
        constexpr int half = LZ / 2 + 1;
        for (int e = tid; e < W * half; e += NT) {
            const int l = e % W, k = e / W;
            if (k == 0) {
                const cx<T> z0 = sl[l];
                sl[l] = cx<T>{z0.re + z0.im, T(0)};
                sl[LZ * WS + l] = cx<T>{z0.re - z0.im, T(0)};
            } else {
                const int kk = LZ - k;
                const cx<T> zk = sl[k * WS + l], zc = cconj(sl[kk * WS + l]);
                const cx<T> E = cscale(zk + zc, T(0.5));
                const cx<T> O = cscale(rot90<-1>(zk - zc), T(0.5));
                sl[k * WS + l] = E + cmul(tw[k], O);
                if (kk != k) sl[kk * WS + l] = cconj(E) + cmul(tw[kk], cconj(O));
            }
        }
    }
    __syncthreads();
    cx<T>* dst = p.S12 + ((long long)x * g.Ny + y0) * NZH * C;
    for (int e = tid; e < TY * NZH * C; e += NT) {
        const int yl = e / (NZH * C);
        const int r = e - yl * (NZH * C);
        const int kz = r / C, c = r - kz * C;
        dst[e] = sl[kz * WS + yl * C + c];
    }
}

// ------------------------------------------------------------------ x pass
// Block: row y, kz tile [kz0, kz0+nk).  Forward x FFT of the sigma^ lines,
// eqVar partials from sigma^ (convergenceCheck :351-389), Green step
// eps^ - Gamma0:sigma^ (multiplyGF :323-338, DC = E), inverse x FFT of its
// 6 lines per kz into the S6 buffer p.S6w.
//   E6 = false (after a START z pass): sigma^ and eps^ both come from S12
//     (12 components per kz) and both are x-transformed here; LDS row kx
//     holds W = 12 TK columns, column kzl*12 + c (c >= 6 eps).
//   E6 = true (after an ITER z pass): sigma^ comes from S12 with 6
//     components per kz (LDS row kx: W = 6 TK columns, kzl*6 + c); eps^ is
//     read from Eh, where the previous x pass left it already x-transformed,
//     by the thread that applies the Green step to it (L*TK <= XNT: one
//     frequency per thread), issued with the sigma^ loads so that it arrives
//     during the forward FFT.
// Both write the new eps^ (unnormalised, before the inverse x FFT) to Eh,
// one contiguous [kx][kzl][c] block per tile, for the next iteration.
template <typename T, int L, int TK, bool E6>
__global__ void __launch_bounds__(XNT) M3D_BX_ATTR kxb(Geo g, BPtr<T> p)
{
    constexpr int C = E6 ? 6 : 12; // LDS columns per kz
    constexpr int W = TK * C;
    constexpr int NT = XNT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[NT / 64];
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    cx<T>* tw = sl + L * W;

    const int nkt = (g.Nzh + TK - 1) / TK;
    const int t = fast::xcd_remap(blockIdx.x, gridDim.x);
    const int y = t / nkt;
    const int kz0 = (t - y * nkt) * TK;
    const Ctl& ctl = p.ctl[0];
    if (ctl.done || y >= g.Ny) return;
    const int nk = min(TK, g.Nzh - kz0);
    const int tid = threadIdx.x;
    cx<T>* eh = p.Eh + (long long)t * (L * TK * 6);
    static_assert(L * TK <= NT, "one frequency per thread");
    for (int i = tid; i < L; i += NT) tw[i] = p.twx[i];
    const cx<T>* base = p.S12 + ((long long)y * g.Nzh + kz0) * C;
    const long long xstride = (long long)g.Ny * g.Nzh * C;
    cx<T> ev[6];
    {
        using RT = fast::RowTile<T, L, W, NT, sizeof(T) == 4>;
        typename RT::V r[RT::PER];
        fast::rows_load<T, L, W, NT, sizeof(T) == 4>(r, base, fast::uniform_rows(xstride), nk * C);
        if (E6 && tid < L * nk) {
#pragma unroll
            for (int c = 0; c < 6; ++c) ev[c] = eh[(long long)tid * 6 + c];
        }
        fast::rows_commit<T, L, W, NT, sizeof(T) == 4>(sl, r);
    }
    __syncthreads();
    fast::fft<T, L, W, NT, -1, 1>(sl, tw);

    const T la0 = T(ctl.la0), mu0 = T(ctl.mu0);
    const T sc = T(g.invN);
    double acc = 0.0;
    for (int f = tid; f < L * nk; f += NT) {
        const int kx = f / nk, kzl = f - kx * nk;
        const int kz = kz0 + kzl;
        cx<T>* v = sl + kx * W + kzl * C;
        cx<T>* eo = eh + (long long)f * 6;
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
            p.pchk[g.nbx] = sqrt(a2[0] + a2[1] + a2[2] + 2 * a2[3] + 2 * a2[4] + 2 * a2[5]);
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                v[c] = cx<T>{T(ctl.E[c]), T(0)};
                eo[c] = cx<T>{T(ctl.E[c] / g.invN), T(0)};
            }
        } else {
            if constexpr (!E6) {
#pragma unroll
                for (int c = 0; c < 6; ++c) ev[c] = v[6 + c];
            }
            const int fz = Elastic::fmap(kz, g.Nz), fy = Elastic::fmap(y, g.Ny), fxx = Elastic::fmap(kx, g.Nx);
            T xi[3] = {T(fz), T(fy), T(fxx)};
            cx<T> sg[6], o[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) sg[c] = v[c];
            green_elastic<T>(xi, sg, o, 1, la0, mu0);
            // kz = 0 and Nz/2 planes: Hermitian part of the Gamma0 term
            if (kz == 0 || kz == g.Nz / 2) {
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
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const cx<T> n = ev[c] - o[c];
                eo[c] = n;
                v[c] = cscale(n, sc);
            }
        }
    }
    const double bs = block_sum(acc, red);
    if (tid == 0) p.pchk[t] = bs;
    __syncthreads();
    cx<T>* ob = p.S6w + ((long long)y * g.Nzh + kz0) * 6;
    const long long ostride = (long long)g.Ny * g.Nzh * 6;
    if constexpr (E6) {
        fast::fft<T, L, W, NT, 1, 1>(sl, tw);
        for (int e = tid; e < L * nk * 6; e += NT) {
            const int xx = e / (nk * 6), r = e - xx * (nk * 6);
            ob[xx * ostride + r] = sl[xx * W + r];
        }
    } else {
        fast::fft_sub<T, L, TK * 6, NT, 1, 1, W, 6, C, 0>(sl, tw);
        for (int e = tid; e < L * nk * 6; e += NT) {
            const int xx = e / (nk * 6), r = e - xx * (nk * 6);
            const int kzl = r / 6, c = r - kzl * 6;
            ob[xx * ostride + r] = sl[xx * W + kzl * C + c];
        }
    }
}

// ---------------------------------------------------------------- launchers
namespace {

template <typename T, int BEH, int MODE, int LZ, bool LIN>
bool z_at(const Geo& g, const BPtr<T>& p, hipStream_t st)
{
    if (g.Nz != 2 * LZ) return false;
    constexpr int TY = zlines(LZ, 2 * LZ);
    if (g.TY != TY) throw std::runtime_error("basic z tile mismatch");
    constexpr int C = MODE == BMODE_START ? 12 : 6;
    const size_t sm = ((size_t)(LZ + 1) * (TY * C + M3D_BZ_PAD) + 2 * LZ) * sizeof(cx<T>);
    hipLaunchKernelGGL((kzb<T, BEH, MODE, LZ, TY, LIN>), dim3(g.nbz), dim3(ZNT), sm, st, g, p);
    return true;
}

template <typename T, int BEH, int MODE, bool LIN = false>
void z_mode(const Geo& g, const BPtr<T>& p, hipStream_t st)
{
    const bool ok = z_at<T, BEH, MODE, 128, LIN>(g, p, st) || z_at<T, BEH, MODE, 96, LIN>(g, p, st) ||
                    z_at<T, BEH, MODE, 64, LIN>(g, p, st) || z_at<T, BEH, MODE, 48, LIN>(g, p, st) ||
                    z_at<T, BEH, MODE, 32, LIN>(g, p, st) || z_at<T, BEH, MODE, 24, LIN>(g, p, st) ||
                    z_at<T, BEH, MODE, 16, LIN>(g, p, st);
    if (!ok) throw std::runtime_error("no basic z kernel for this grid");
}

template <typename T, typename PH, int DIR, int WHICH, int L>
bool y_at(const Geo& g, const Ptr<T>& p, hipStream_t st)
{
    if (g.Ny != L) return false;
    constexpr int TK = fast::xytile<T>(L, PH::C);
    const int nkt = (g.Nzh + TK - 1) / TK;
    const size_t sm = ((size_t)L * TK * PH::C + L) * sizeof(cx<T>);
    hipLaunchKernelGGL((fast::kxy<T, PH, 0, DIR, WHICH, L, TK, 256>), dim3(g.Nx * nkt), dim3(256), sm, st, g, p,
                       (const T*)nullptr);
    return true;
}

template <typename T, typename PH, int DIR, int WHICH>
void y_run(const Geo& g, const Ptr<T>& p, hipStream_t st)
{
    const bool ok = y_at<T, PH, DIR, WHICH, 256>(g, p, st) || y_at<T, PH, DIR, WHICH, 192>(g, p, st) ||
                    y_at<T, PH, DIR, WHICH, 128>(g, p, st) || y_at<T, PH, DIR, WHICH, 96>(g, p, st) ||
                    y_at<T, PH, DIR, WHICH, 64>(g, p, st) || y_at<T, PH, DIR, WHICH, 48>(g, p, st) ||
                    y_at<T, PH, DIR, WHICH, 32>(g, p, st);
    if (!ok) throw std::runtime_error("no basic y kernel for this grid");
}

template <typename T, int L, bool E6>
bool x_at(const Geo& g, const BPtr<T>& p, hipStream_t st)
{
    if (g.Nx != L) return false;
    constexpr int TK = xtile(L);
    if (g.TKx != TK) throw std::runtime_error("basic x tile mismatch");
    const size_t sm = ((size_t)L * TK * (E6 ? 6 : 12) + L) * sizeof(cx<T>);
    if (sm > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&kxb<T, L, TK, E6>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
    hipLaunchKernelGGL((kxb<T, L, TK, E6>), dim3(g.nbx), dim3(XNT), sm, st, g, p);
    return true;
}

template <typename T, bool E6>
void x_run(const Geo& g, const BPtr<T>& p, hipStream_t st)
{
    const bool ok = x_at<T, 256, E6>(g, p, st) || x_at<T, 192, E6>(g, p, st) || x_at<T, 128, E6>(g, p, st) ||
                    x_at<T, 96, E6>(g, p, st) || x_at<T, 64, E6>(g, p, st) || x_at<T, 48, E6>(g, p, st) ||
                    x_at<T, 32, E6>(g, p, st);
    if (!ok) throw std::runtime_error("no basic x kernel for this grid");
}

} // namespace

template <typename T>
void launch_z(int beh, int mode, const Geo& g, const BPtr<T>& p, hipStream_t st, bool lin)
{
    if (beh == BEH_EP && lin && sizeof(T) == 8) {
        // FP64, every phase with n = 1: the instantiation without pow
        if (mode == BMODE_START) z_mode<T, BEH_EP, BMODE_START, true>(g, p, st);
        else if (mode == BMODE_FINAL) z_mode<T, BEH_EP, BMODE_FINAL, true>(g, p, st);
        else z_mode<T, BEH_EP, BMODE_ITER, true>(g, p, st);
    } else if (beh == BEH_EP) {
        if (mode == BMODE_START) z_mode<T, BEH_EP, BMODE_START>(g, p, st);
        else if (mode == BMODE_FINAL) z_mode<T, BEH_EP, BMODE_FINAL>(g, p, st);
        else z_mode<T, BEH_EP, BMODE_ITER>(g, p, st);
    } else {
        if (mode == BMODE_START) z_mode<T, BEH_ED, BMODE_START>(g, p, st);
        else if (mode == BMODE_FINAL) z_mode<T, BEH_ED, BMODE_FINAL>(g, p, st);
        else z_mode<T, BEH_ED, BMODE_ITER>(g, p, st);
    }
}

template <typename T>
void launch_y12(const Geo& g, const Ptr<T>& p, hipStream_t st)
{
    y_run<T, Comp12, -1, BUF_K>(g, p, st);
}

template <typename T>
void launch_y6_fwd(const Geo& g, const Ptr<T>& p, hipStream_t st)
{
    y_run<T, Comp6, -1, BUF_K>(g, p, st);
}

template <typename T>
void launch_y6(const Geo& g, const Ptr<T>& p, hipStream_t st)
{
    y_run<T, Comp6, 1, BUF_MAIN>(g, p, st);
}

template <typename T>
void launch_x(const Geo& g, const BPtr<T>& p, hipStream_t st, bool e6)
{
    if (e6) x_run<T, true>(g, p, st);
    else x_run<T, false>(g, p, st);
}

template void launch_z<double>(int, int, const Geo&, const BPtr<double>&, hipStream_t, bool);
template void launch_z<float>(int, int, const Geo&, const BPtr<float>&, hipStream_t, bool);
template void launch_y12<double>(const Geo&, const Ptr<double>&, hipStream_t);
template void launch_y12<float>(const Geo&, const Ptr<float>&, hipStream_t);
template void launch_y6_fwd<double>(const Geo&, const Ptr<double>&, hipStream_t);
template void launch_y6_fwd<float>(const Geo&, const Ptr<float>&, hipStream_t);
template void launch_y6<double>(const Geo&, const Ptr<double>&, hipStream_t);
template void launch_y6<float>(const Geo&, const Ptr<float>&, hipStream_t);
template void launch_x<double>(const Geo&, const BPtr<double>&, hipStream_t, bool);
template void launch_x<float>(const Geo&, const BPtr<float>&, hipStream_t, bool);

} // namespace bfast
} // namespace m3d
