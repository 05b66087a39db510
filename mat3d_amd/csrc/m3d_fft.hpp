// m3d_fft.hpp -- LDS-resident Stockham FFT building blocks for gfx950.
//
// A workgroup owns a tile of lines in LDS and transforms all of them with a
// sequence of in-place Stockham passes: every thread loads the R inputs of
// each of its butterflies into registers, the workgroup synchronises, the
// butterflies are computed and stored to their autosorted positions.  No
// bit reversal pass is needed and the output is in natural order.
//
// Lines are described by a two-level geometry so one routine covers both
// layouts used by the solver:
//   * "interleaved" lines (y/x passes): element e of line l lives at
//     l + e*es (es = number of lines), consecutive threads take consecutive
//     lines -> conflict-free LDS access;
//   * "contiguous" lines (z pass): element e of line l at l*ls + e,
//     consecutive threads take consecutive elements.
//
// Twiddles come from a per-axis table tw[m] = exp(-2 pi i m / Ltab); a pass
// over lines of length L uses stride Ltab/L into it.
#pragma once
#include <hip/hip_runtime.h>

namespace m3d {

template <typename T>
struct cx {
    T re, im;
};

template <typename T>
__device__ __forceinline__ cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return cx<T>{a.re + b.re, a.im + b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return cx<T>{a.re - b.re, a.im - b.im}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b)
{
    return cx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return cx<T>{a.re * s, a.im * s}; }
template <typename T>
__device__ __forceinline__ cx<T> cconj(cx<T> a) { return cx<T>{a.re, -a.im}; }
// multiply by -i (forward, DIR<0) or +i (inverse, DIR>0)
template <int DIR, typename T>
__device__ __forceinline__ cx<T> rot90(cx<T> a)
{
    return DIR < 0 ? cx<T>{a.im, -a.re} : cx<T>{-a.im, a.re};
}

struct FftAxis {
    int L;                   // transform length
    int np;                  // number of passes (<= 16)
    int tws;                 // stride into the twiddle table (table length / L)
    int pad;
    unsigned long long rad;  // radix of pass p in bits [4p, 4p+4) (kept packed so the
                             // kernel never indexes its by-value arguments dynamically)
    __host__ __device__ int radix(int p) const { return (int)((rad >> (4 * p)) & 15ull); }
};

template <typename T, int R, int DIR>
__device__ __forceinline__ void butterfly(cx<T> (&v)[R], const cx<T>* __restrict__ tw, int wstep)
{
    if constexpr (R == 2) {
        cx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    } else if constexpr (R == 3) {
        const T c = T(-0.5);
        const T s = T(0.86602540378443864676372317075294); // sin(2 pi / 3)
        cx<T> t1 = v[1] + v[2];
        cx<T> t2 = v[1] - v[2];
        cx<T> m = cx<T>{v[0].re + c * t1.re, v[0].im + c * t1.im};
        // forward: X1 = m - i s t2, X2 = m + i s t2
        cx<T> r = rot90<DIR>(t2);
        v[0] = v[0] + t1;
        v[1] = cx<T>{m.re + s * r.re, m.im + s * r.im};
        v[2] = cx<T>{m.re - s * r.re, m.im - s * r.im};
    } else if constexpr (R == 4) {
        cx<T> s02 = v[0] + v[2], d02 = v[0] - v[2];
        cx<T> s13 = v[1] + v[3], d13 = rot90<DIR>(v[1] - v[3]);
        v[0] = s02 + s13;
        v[2] = s02 - s13;
        v[1] = d02 + d13;
        v[3] = d02 - d13;
    } else if constexpr (R == 8) {
        const T h = T(0.70710678118654752440084436210485);
        cx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cx<T> o[4] = {v[1], v[3], v[5], v[7]};
        // 4-point DFTs
        {
            cx<T> s02 = e[0] + e[2], d02 = e[0] - e[2];
            cx<T> s13 = e[1] + e[3], d13 = rot90<DIR>(e[1] - e[3]);
            e[0] = s02 + s13; e[2] = s02 - s13; e[1] = d02 + d13; e[3] = d02 - d13;
        }
        {
            cx<T> s02 = o[0] + o[2], d02 = o[0] - o[2];
            cx<T> s13 = o[1] + o[3], d13 = rot90<DIR>(o[1] - o[3]);
            o[0] = s02 + s13; o[2] = s02 - s13; o[1] = d02 + d13; o[3] = d02 - d13;
        }
        // twiddles W8^k (forward: exp(-i pi k / 4))
        cx<T> o1 = o[1], o2 = o[2], o3 = o[3];
        cx<T> r1 = rot90<DIR>(o1);
        o1 = cx<T>{h * (o1.re + r1.re), h * (o1.im + r1.im)}; // (1 -/+ i)/sqrt2
        o2 = rot90<DIR>(o2);
        cx<T> r3 = rot90<DIR>(o3);
        o3 = cx<T>{h * (r3.re - o3.re), h * (r3.im - o3.im)}; // (-1 -/+ i)/sqrt2
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[5] = e[1] - o1;
        v[2] = e[2] + o2;
        v[6] = e[2] - o2;
        v[3] = e[3] + o3;
        v[7] = e[3] - o3;
    } else if constexpr (R == 16) {
        // 16 = 4 x 4: 4-point DFTs over n1 of x[4 n1 + n2], twiddles
        // W16^(n2 k1), 4-point DFTs over n2; X[k1 + 4 k2] lands in v[k1 + 4 k2]
        const T c1 = T(0.92387953251128675612818318939679); // cos(pi/8)
        const T s1 = T(0.38268343236508977172845998403040); // sin(pi/8)
        const T h = T(0.70710678118654752440084436210485);
        cx<T> y[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            const cx<T> a = v[n2], b = v[n2 + 4], c = v[n2 + 8], d = v[n2 + 12];
            const cx<T> s02 = a + c, d02 = a - c, s13 = b + d, d13 = rot90<DIR>(b - d);
            y[n2][0] = s02 + s13;
            y[n2][1] = d02 + d13;
            y[n2][2] = s02 - s13;
            y[n2][3] = d02 - d13;
        }
        // W16^m = (cos, DIR*sin)(2 pi m / 16): forward exp(-i ...), inverse exp(+i ...)
        auto tmul = [](cx<T> z, T wr, T wi) { return cx<T>{z.re * wr - z.im * wi, z.re * wi + z.im * wr}; };
        const T sd = DIR > 0 ? T(1) : T(-1);
        y[1][1] = tmul(y[1][1], c1, sd * s1);   // m = 1
        y[1][2] = tmul(y[1][2], h, sd * h);     // m = 2
        y[1][3] = tmul(y[1][3], s1, sd * c1);   // m = 3
        y[2][1] = tmul(y[2][1], h, sd * h);     // m = 2
        y[2][2] = rot90<DIR>(y[2][2]);          // m = 4: -i forward, +i inverse
        y[2][3] = tmul(y[2][3], -h, sd * h);    // m = 6
        y[3][1] = tmul(y[3][1], s1, sd * c1);   // m = 3
        y[3][2] = tmul(y[3][2], -h, sd * h);    // m = 6
        y[3][3] = tmul(y[3][3], -c1, -sd * s1); // m = 9
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            const cx<T> a = y[0][k1], b = y[1][k1], c = y[2][k1], d = y[3][k1];
            const cx<T> s02 = a + c, d02 = a - c, s13 = b + d, d13 = rot90<DIR>(b - d);
            v[k1] = s02 + s13;
            v[k1 + 4] = d02 + d13;
            v[k1 + 8] = s02 - s13;
            v[k1 + 12] = d02 - d13;
        }
    } else {
        // generic odd radix: direct DFT, W_R^m = tw[m * wstep]
        cx<T> out[R];
#pragma unroll
        for (int q = 0; q < R; ++q) {
            cx<T> acc = v[0];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cx<T> w = tw[((r * q) % R) * wstep];
                if (DIR > 0) w.im = -w.im;
                acc = acc + cmul(v[r], w);
            }
            out[q] = acc;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) v[q] = out[q];
    }
}

// Line geometry inside LDS: offset(l) = (l % nin)*lin + (l / nin)*lout,
// element stride es.
struct LineGeo {
    int nl;   // number of lines
    int nin;  // lines in the inner group
    int lin;  // stride between inner lines
    int lout; // stride between inner groups
    int es;   // element stride
    __device__ __forceinline__ int off(int l) const { return (l % nin) * lin + (l / nin) * lout; }
};

// One in-place Stockham pass of radix R over all lines of the tile.
// EPT bounds the number of complex values a thread holds (items*R).
template <typename T, int R, int DIR, int EPT, bool LINES_FAST>
__device__ __forceinline__ void fft_pass(cx<T>* __restrict__ s, const LineGeo g, int L, int Ns,
                                         const cx<T>* __restrict__ tw, int tws)
{
    constexpr int IPT = (EPT + R - 1) / R;
    const int nR = L / R;
    const int items = g.nl * nR;
    const int tstep = (L / (Ns * R)) * tws;
    const int nthr = blockDim.x;
    cx<T> v[IPT][R];
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int it = threadIdx.x + i * nthr;
        if (it < items) {
            int l, j;
            if (LINES_FAST) { l = it % g.nl; j = it / g.nl; }
            else { j = it % nR; l = it / nR; }
            const int lo = g.off(l);
            const int k = j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> a = s[lo + (j + r * nR) * g.es];
                if (r > 0 && Ns > 1) {
                    cx<T> w = tw[k * r * tstep];
                    if (DIR > 0) w.im = -w.im;
                    a = cmul(a, w);
                }
                v[i][r] = a;
            }
            butterfly<T, R, DIR>(v[i], tw, nR * tws);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int it = threadIdx.x + i * nthr;
        if (it < items) {
            int l, j;
            if (LINES_FAST) { l = it % g.nl; j = it / g.nl; }
            else { j = it % nR; l = it / nR; }
            const int lo = g.off(l);
            const int k = j % Ns;
            const int ob = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) s[lo + (ob + r * Ns) * g.es] = v[i][r];
        }
    }
    __syncthreads();
}

// Full transform of all tile lines (in place).  Radices 2,3,4,8 always;
// GEN adds the odd radices 5,7,11,13 (kept out of the common instantiation
// because their register footprint would cap the occupancy of every kernel).
template <typename T, int DIR, int EPT, bool LINES_FAST, bool GEN>
__device__ void fft_lines(cx<T>* __restrict__ s, const LineGeo g, const FftAxis ax, const cx<T>* __restrict__ tw)
{
    int Ns = 1;
    for (int p = 0; p < ax.np; ++p) {
        const int R = ax.radix(p);
        if (R == 8) fft_pass<T, 8, DIR, EPT, LINES_FAST>(s, g, ax.L, Ns, tw, ax.tws);
        else if (R == 4) fft_pass<T, 4, DIR, EPT, LINES_FAST>(s, g, ax.L, Ns, tw, ax.tws);
        else if (R == 3) fft_pass<T, 3, DIR, EPT, LINES_FAST>(s, g, ax.L, Ns, tw, ax.tws);
        else if (R == 2) fft_pass<T, 2, DIR, EPT, LINES_FAST>(s, g, ax.L, Ns, tw, ax.tws);
        else if constexpr (GEN) {
            if (R == 5) fft_pass<T, 5, DIR, EPT, LINES_FAST>(s, g, ax.L, Ns, tw, ax.tws);
            else if (R == 7) fft_pass<T, 7, DIR, EPT, LINES_FAST>(s, g, ax.L, Ns, tw, ax.tws);
            else if (R == 11) fft_pass<T, 11, DIR, EPT, LINES_FAST>(s, g, ax.L, Ns, tw, ax.tws);
            else if (R == 13) fft_pass<T, 13, DIR, EPT, LINES_FAST>(s, g, ax.L, Ns, tw, ax.tws);
        }
        Ns *= R;
    }
}

} // namespace m3d
