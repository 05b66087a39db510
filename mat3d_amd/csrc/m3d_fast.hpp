// m3d_fast.hpp -- compile-time specialised passes for the common grid sizes.
//
// Same dataflow and layouts as m3d_kernels.hpp, but the line length, the
// tile shape and the workgroup size are template constants, so every index
// computation of the Stockham passes folds to shifts and constant
// multiplies, every loop unrolls and the per-thread register footprint is
// known.  Twiddle tables are staged in LDS once per workgroup.  Lines are
// interleaved in LDS (element e of line l at e*W + l) so a wavefront touches
// consecutive LDS words in every pass.
//
// The z pass works on voxel pairs: the half-length complex FFT of a real
// z-line packs samples (2m, 2m+1) into complex slot m, so the c2r output of
// slot m of the C component lines *is* the C-vector of voxels 2m and 2m+1.
// A thread owning slot m reads the 2C unknowns of its voxel pair with vector
// loads, applies update / dual / polarisation, and writes the 2C polarisation
// values back into the same slots for the r2c.
#pragma once
#include <type_traits>
#include "m3d_kernels.hpp"

namespace m3d {
namespace fast {

// native vector types for register staging (HIP's float4 is a struct and
// arrays of it are not promoted to registers when copied)
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef double vd2 __attribute__((ext_vector_type(2)));
// one complex value (8 or 16 bytes) or a pair of FP32 complex values (16 bytes)
template <typename T, bool PAIR>
using stage_t = typename std::conditional<PAIR, vf4, typename std::conditional<sizeof(T) == 4, vf2, vd2>::type>::type;

// ---------------------------------------------------------------- planning
constexpr int n_twos(int L) { return (L % 2 == 0) ? 1 + n_twos(L / 2) : 0; }
constexpr int odd_part(int L) { return (L % 2 == 0) ? odd_part(L / 2) : L; }
constexpr int n_threes(int L) { return (L % 3 == 0) ? 1 + n_threes(L / 3) : 0; }
// radix sequence of the compile-time passes: 16s for the powers of two (a
// leftover factor 8 or 4 as one more pass; 2 as 16*2 -> 8*4), then 3s.  Fewer
// passes mean fewer LDS round trips per line.
#ifndef M3D_RADIX16
#define M3D_RADIX16 1
#endif
constexpr int n8(int e) { return e >= 3 ? ((e == 4) ? 0 : 1 + n8(e - 3)) : 0; }
constexpr int rest2(int e) { return e >= 3 ? ((e == 4) ? 4 : rest2(e - 3)) : e; }
// powers-of-two radices of 2^e, pass p (0 if p is past them).  Radix plan
// rp = 0 is the build's default (radix 16 first unless M3D_RADIX16 is 0),
// rp = 1 the other one (radix 8 first, or 16 first when the default is 8);
// the autotuner times both for the y pass and the x pass with the Green step.
constexpr int pow2_radix(int e, int p, int rp = 0)
{
    if ((M3D_RADIX16 != 0) == (rp != 0)) {
        const int a = n8(e), r = rest2(e);
        if (p < a) return 8;
        p -= a;
        if (r == 4) return p < 2 ? 4 : 0;
        if (r == 2) return p < 1 ? 4 : 0;
        if (r == 1) return p < 1 ? 2 : 0;
        return 0;
    }
    const int n16 = e / 4, rem = e % 4;
    if (rem == 1 && n16 >= 1) {
        // 16^(n16-1) * 8 * 4
        if (p < n16 - 1) return 16;
        p -= n16 - 1;
        return p == 0 ? 8 : (p == 1 ? 4 : 0);
    }
    if (p < n16) return 16;
    p -= n16;
    if (p == 0 && rem > 0) return 1 << rem;
    return 0;
}
constexpr int pow2_passes(int e, int rp = 0)
{
    int n = 0;
    while (pow2_radix(e, n, rp) != 0) ++n;
    return n;
}
constexpr int npass(int L, int rp = 0) { return pow2_passes(n_twos(L), rp) + n_threes(L); }
constexpr int radix(int L, int p, int rp = 0)
{
    const int e = n_twos(L), a = pow2_passes(e, rp);
    return p < a ? pow2_radix(e, p, rp) : 3;
}
// the two radix plans of length L differ (else plan 1 is not worth timing)
constexpr bool plans_differ(int L)
{
    if (npass(L, 0) != npass(L, 1)) return true;
    for (int p = 0; p < npass(L, 0); ++p)
        if (radix(L, p, 0) != radix(L, p, 1)) return true;
    return false;
}
constexpr bool supported(int L) { return odd_part(L) == 1 || odd_part(L) == 3 || odd_part(L) == 9; }

// tile shapes (shared with the host planner)
#ifndef M3D_ZBUDGET
#define M3D_ZBUDGET 3072 // complex values per z tile (FP32; FP64 half)
#endif
template <typename T>
constexpr int zlines(int LZ, int C, int Ny)
{
    // TY: largest power of two with TY*C*LZ <= budget that divides Ny
    int budget = sizeof(T) == 4 ? M3D_ZBUDGET : M3D_ZBUDGET / 2;
    int ty = 1;
    while (ty < 16 && ty * 2 * C * LZ <= budget && Ny % (ty * 2) == 0) ty *= 2;
    return ty;
}
#ifndef M3D_YTILE
#define M3D_YTILE 6144
#endif
template <typename T>
constexpr int xytile(int L, int C)
{
    int budget = sizeof(T) == 4 ? M3D_YTILE : M3D_YTILE / 2;
    int tk = budget / (C * L);
    return tk < 1 ? 1 : tk;
}
// x passes: complex values per tile (FP32; FP64 half); the x lines are one
// (y, kz) row apart, Ny*Nzh*C values, so wider kz tiles amortise each row visit
#ifndef M3D_XTILE
#define M3D_XTILE 6144
#endif
template <typename T>
constexpr int xtile(int L, int C)
{
    int budget = sizeof(T) == 4 ? M3D_XTILE : M3D_XTILE / 2;
    int tk = budget / (C * L);
    return tk < 1 ? 1 : tk;
}
// x pass of the equilibrium check (FP32): narrower kz tiles, so more
// workgroups share a CU's LDS while the tile is read (the pass only reads)
#ifndef M3D_XCTILE
#define M3D_XCTILE 3072
#endif
template <typename T>
constexpr int xctile(int L, int C)
{
    int budget = sizeof(T) == 4 ? M3D_XCTILE : M3D_XTILE / 2;
    int tk = budget / (C * L);
    return tk < 1 ? 1 : tk;
}
// the autotuner's wide alternative for the x pass with the Green step (1.5x)
template <typename T>
constexpr int xtile_wide(int L, int C)
{
    int budget = (sizeof(T) == 4 ? M3D_XTILE : M3D_XTILE / 2) * 3 / 2;
    int tk = budget / (C * L);
    return tk < 1 ? 1 : tk;
}
constexpr int ZNT = 256;

// threads per workgroup of the x/y passes: y pass, x pass with the Green
// step, x pass of the check FFT
#ifndef M3D_XY_NT
#define M3D_XY_NT 512
#endif
#ifndef M3D_XG_NT
#define M3D_XG_NT 512
#endif
#ifndef M3D_XC_NT
#define M3D_XC_NT 256
#endif
// (FP64 tiles hold half the lines: they keep 256 threads)
template <int KIND, typename T>
constexpr int xy_threads()
{
    if (sizeof(T) == 8) return 256;
    return KIND == 1 ? M3D_XG_NT : (KIND == 2 ? M3D_XC_NT : M3D_XY_NT);
}

// minimum waves per SIMD the register allocator must leave room for (0: no
// constraint); these passes are HBM-bound, so occupancy buys latency hiding
#ifndef M3D_KZ_WAVES
#define M3D_KZ_WAVES 4
#endif
#ifndef M3D_KXY_WAVES
#define M3D_KXY_WAVES 0
#endif
#if M3D_KZ_WAVES > 0
#define M3D_KZ_ATTR __attribute__((amdgpu_waves_per_eu(M3D_KZ_WAVES)))
#else
#define M3D_KZ_ATTR
#endif
#if M3D_KXY_WAVES > 0
#define M3D_KXY_ATTR __attribute__((amdgpu_waves_per_eu(M3D_KXY_WAVES)))
#else
#define M3D_KXY_ATTR
#endif

// XCD-aware block order: the dispatcher hands block b to XCD b % 8; give
// each XCD a contiguous run of logical tiles so tiles sharing cache lines
// run under one L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    const int per = nb >> 3;
    if (b < (per << 3)) return (b & 7) * per + (b >> 3);
    return b;
}

// The same per solve: nb = nsolve * per tiles, solve-major; XCD k takes the
// k-th contiguous eighth of every solve's tiles, so a batch in which only
// some solves still run keeps all 8 XCDs busy.  Falls back to xcd_remap
// when per is not a multiple of 8.
__device__ __forceinline__ int xcd_split(int b, int nb, int per)
{
    if (per < 8 || (per & 7) || nb % per) return xcd_remap(b, nb);
    const int ps = per >> 3;
    const int j = b >> 3, s = j / ps;
    return s * per + (b & 7) * ps + (j - s * ps);
}

// thread index as an opaque value: address arithmetic derived from it is
// redone where it is used instead of being hoisted out of a tile loop and
// held in registers across it
__device__ __forceinline__ int lane_tid()
{
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// Index arithmetic at full rate: on CDNA the 32-bit integer multiply
// (v_mul_lo_u32, v_mul_hi_u32) and the 64-bit multiply-add issue at a quarter
// of the rate of v_mul_u32_u24, and the x pass with the Green step is bound by
// VALU issue.  cdiv<D, XMAX>(x) is x / D for 0 <= x < XMAX: a shift for powers
// of two, else one 24-bit multiply by M = ceil(2^S / D) and a shift, with S the
// smallest shift for which x * M fits 32 bits and XMAX * (M D - 2^S) <= 2^S
// (then floor(x M / 2^S) = floor(x / D) for every x < XMAX).
template <int D, int XMAX>
struct FastDiv {
    static constexpr int pick()
    {
        for (int S = 0; S < 32; ++S) {
            const unsigned long long M = ((1ull << S) + D - 1) / D;
            const unsigned long long e = M * D - (1ull << S);
            if (M < (1ull << 24) && (unsigned long long)XMAX * M < (1ull << 32) && (unsigned long long)XMAX * e <= (1ull << S))
                return S;
        }
        return -1;
    }
    static constexpr int S = pick();
    static constexpr unsigned M = S < 0 ? 0u : (unsigned)(((1ull << (S < 0 ? 0 : S)) + D - 1) / D);
};

template <int D, int XMAX>
__device__ __forceinline__ int cdiv(int x)
{
    static_assert(D > 0 && XMAX > 0 && XMAX <= (1 << 24), "cdiv range");
    if constexpr ((D & (D - 1)) == 0) {
        return (int)((unsigned)x / (unsigned)D);
    } else {
        static_assert(FastDiv<D, XMAX>::S >= 0, "no 24-bit reciprocal for this divisor and range");
        return (int)(__umul24((unsigned)x, FastDiv<D, XMAX>::M) >> FastDiv<D, XMAX>::S);
    }
}

// --------------------------------------------------------------- passes
template <int DIR, typename T>
__device__ __forceinline__ cx<T> twmul(cx<T> a, cx<T> w)
{
    if (DIR > 0) w.im = -w.im;
    return cmul(a, w);
}

// one Stockham pass over W interleaved lines of length L (in place).  The
// LDS row of sample j holds WS >= W columns; line l sits in column
// (l / G) * GS + OFF + l % G (defaults: columns 0..W-1), so a subset of the
// lines (e.g. 6 of every 12) can be transformed in place.
template <typename T, int L, int W, int NT, int R, int Ns, int DIR, int TWS, int WS = W, int G = W, int GS = 0,
          int OFF = 0>
__device__ __forceinline__ void pass(cx<T>* __restrict__ s, const cx<T>* __restrict__ tw)
{
    constexpr int nR = L / R;
    constexpr int items = W * nR;
    constexpr int IPT = (items + NT - 1) / NT;
    constexpr int tstep = (L / (Ns * R)) * TWS;
    const int tid = lane_tid();
    cx<T> v[IPT][R];
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int it = tid + i * NT;
        if (items % NT == 0 || it < items) {
            const int j = cdiv<W, IPT * NT>(it), l0 = it - j * W, k = j % Ns;
            int l;
            if constexpr (G == W) l = OFF + l0;
            else {
                const int gq = cdiv<G, W>(l0);
                l = gq * GS + OFF + (l0 - gq * G);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                cx<T> a = s[(j + r * nR) * WS + l];
                if (Ns > 1 && r > 0) a = twmul<DIR>(a, tw[k * r * tstep]);
                v[i][r] = a;
            }
            butterfly<T, R, DIR>(v[i], tw, 0);
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int it = tid + i * NT;
        if (items % NT == 0 || it < items) {
            const int j = cdiv<W, IPT * NT>(it), l0 = it - j * W, k = j % Ns;
            int l;
            if constexpr (G == W) l = OFF + l0;
            else {
                const int gq = cdiv<G, W>(l0);
                l = gq * GS + OFF + (l0 - gq * G);
            }
            const int ob = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) s[(ob + r * Ns) * WS + l] = v[i][r];
        }
    }
    __syncthreads();
}

template <typename T, int L, int W, int NT, int DIR, int TWS, int P = 0, int Ns = 1, int WS = W, int G = W, int GS = 0,
          int OFF = 0, int RP = 0>
__device__ __forceinline__ void fft(cx<T>* __restrict__ s, const cx<T>* __restrict__ tw)
{
    if constexpr (P < npass(L, RP)) {
        constexpr int R = radix(L, P, RP);
        pass<T, L, W, NT, R, Ns, DIR, TWS, WS, G, GS, OFF>(s, tw);
        fft<T, L, W, NT, DIR, TWS, P + 1, Ns * R, WS, G, GS, OFF, RP>(s, tw);
    }
}

// the lines OFF + (0..G-1) of every group of GS columns (WS columns in all)
template <typename T, int L, int W, int NT, int DIR, int TWS, int WS, int G, int GS, int OFF>
__device__ __forceinline__ void fft_sub(cx<T>* __restrict__ s, const cx<T>* __restrict__ tw)
{
    fft<T, L, W, NT, DIR, TWS, 0, 1, WS, G, GS, OFF>(s, tw);
}

// vector load/store of n contiguous values into a register array (16-byte
// chunks when n*sizeof(T) allows it, else 8); components are moved one by one
// so the register array is never addressed as memory
template <typename T, int n>
__device__ __forceinline__ void vload(const T* __restrict__ g, T (&r)[n])
{
    if constexpr (sizeof(T) == 4 && n % 4 == 0) {
#pragma unroll
        for (int i = 0; i < n / 4; ++i) {
            const float4 q = reinterpret_cast<const float4*>(g)[i];
            r[4 * i] = q.x; r[4 * i + 1] = q.y; r[4 * i + 2] = q.z; r[4 * i + 3] = q.w;
        }
    } else if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int i = 0; i < n / 2; ++i) {
            const float2 q = reinterpret_cast<const float2*>(g)[i];
            r[2 * i] = q.x; r[2 * i + 1] = q.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < n / 2; ++i) {
            const double2 q = reinterpret_cast<const double2*>(g)[i];
            r[2 * i] = q.x; r[2 * i + 1] = q.y;
        }
    }
}
template <typename T, int n>
__device__ __forceinline__ void vstore(T* __restrict__ g, const T (&r)[n])
{
    if constexpr (sizeof(T) == 4 && n % 4 == 0) {
#pragma unroll
        for (int i = 0; i < n / 4; ++i)
            reinterpret_cast<float4*>(g)[i] = make_float4(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]);
    } else if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int i = 0; i < n / 2; ++i) reinterpret_cast<float2*>(g)[i] = make_float2(r[2 * i], r[2 * i + 1]);
    } else {
#pragma unroll
        for (int i = 0; i < n / 2; ++i) reinterpret_cast<double2*>(g)[i] = make_double2(r[2 * i], r[2 * i + 1]);
    }
}

// ---------------------------------------------------------- tile movers
// Rows of W contiguous complex values in global memory (row r at
// base + r*stride) <-> LDS rows sl[r*W + j].  All of a thread's loads are
// issued before any LDS store so the workgroup keeps ROWS*W*sizeof values
// in flight; pairs of complex values move as one 16-byte access when the
// layout allows it (VEC).  Entries j >= wv are zero-filled / not stored.
// offset (complex values) of row r of a tile: (r >> sh) * a + (r & mask) * b
struct RowMap {
    long long a, b;
    int sh, mask;
    __device__ __forceinline__ long long operator()(int r) const
    {
        return (long long)(r >> sh) * a + (long long)(r & mask) * b;
    }
    // offset of row r < ROWS: plain rows (one stride) whose offsets stay
    // below 2^32 take one full-rate 24-bit multiply (a uniform branch)
    template <int ROWS>
    __device__ __forceinline__ long long off(int r) const
    {
        if (sh == 0 && mask == 0 && a < (long long)(0xffffffffull / ROWS) && a < (1ll << 24))
            return (long long)__umul24((unsigned)r, (unsigned)a);
        return (*this)(r);
    }
};
__device__ __forceinline__ RowMap uniform_rows(long long stride) { return RowMap{stride, 0, 0, 0}; }

// a thread's share of a ROWS x W tile, split into an issue step (global ->
// registers) and a commit step (registers -> LDS)
template <typename T, int ROWS, int W, int NT, bool VEC>
struct RowTile {
    using V = stage_t<T, VEC>;
    static constexpr int G = VEC ? 2 : 1;
    static constexpr int WG = W / G;
    static constexpr int N = ROWS * WG;
    static constexpr int PER = (N + NT - 1) / NT;
};

// a * b for a, b < 2^24 as one full-rate v_mul_u32_u24 (low 32 bits of the
// product).  Written out because the compiler selects the quarter-rate
// 32-bit multiply when the operand that bounds b sits in another basic block.
__device__ __forceinline__ unsigned mul_u24(unsigned a, unsigned b)
{
    unsigned r;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// plain rows whose byte stride s satisfies s < 2^24 and ROWS * s < 2^32: a
// row's byte offset is one full-rate 24-bit multiply from the tile's
// (uniform) base, and no 64-bit address arithmetic is done per access
template <typename T, int ROWS>
__device__ __forceinline__ bool rows_fit32(const RowMap& rm)
{
    constexpr long long lim = (1ll << 24) < (1ll << 32) / ROWS ? (1ll << 24) : (1ll << 32) / ROWS;
    return rm.sh == 0 && rm.mask == 0 && rm.a >= 0 && rm.a * (long long)sizeof(cx<T>) < lim;
}

template <typename T, int ROWS, int W, int NT, bool VEC, typename V, int PER>
__device__ __forceinline__ void rows_load(V (&r)[PER], const cx<T>* __restrict__ base, RowMap rm, int wv)
{
    using RT = RowTile<T, ROWS, W, NT, VEC>;
    static_assert(PER == RT::PER, "register tile size");
    const int t0 = lane_tid();
    if (rows_fit32<T, ROWS>(rm)) {
        const unsigned sb = (unsigned)(rm.a * (long long)sizeof(cx<T>)); // < 2^24
        const char* bb = reinterpret_cast<const char*>(base);
        auto addr = [&](int e, int& jj) {
            const int row = cdiv<RT::WG, RT::PER * NT>(e);
            jj = e - row * RT::WG;
            return reinterpret_cast<const V*>(bb + (mul_u24((unsigned)row, sb) + (unsigned)(jj * (int)sizeof(V))));
        };
        if (wv == W) {
            // every column valid (all kz tiles but a ragged last one)
#pragma unroll
            for (int i = 0; i < RT::PER; ++i) {
                const int e = t0 + i * NT;
                int jj;
                if (RT::N % NT == 0 || e < RT::N) r[i] = *addr(e, jj);
            }
        } else {
#pragma unroll
            for (int i = 0; i < RT::PER; ++i) {
                const int e = t0 + i * NT;
                if (RT::N % NT == 0 || e < RT::N) {
                    int jj;
                    const V* a = addr(e, jj);
                    r[i] = jj * RT::G < wv ? *a : V(0);
                }
            }
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < RT::PER; ++i) {
        const int e = t0 + i * NT;
        if (RT::N % NT == 0 || e < RT::N) {
            const int row = cdiv<RT::WG, RT::PER * NT>(e), j = (e - row * RT::WG) * RT::G;
            r[i] = j < wv ? *reinterpret_cast<const V*>(base + rm.template off<ROWS>(row) + j) : V(0);
        }
    }
}

template <typename T, int ROWS, int W, int NT, bool VEC, typename V, int PER>
__device__ __forceinline__ void rows_commit(cx<T>* __restrict__ sl, const V (&r)[PER])
{
    using RT = RowTile<T, ROWS, W, NT, VEC>;
    static_assert(PER == RT::PER, "register tile size");
    const int t0 = lane_tid();
#pragma unroll
    for (int i = 0; i < RT::PER; ++i) {
        const int e = t0 + i * NT;
        if (RT::N % NT == 0 || e < RT::N) {
            const int row = cdiv<RT::WG, RT::PER * NT>(e), j = (e - row * RT::WG) * RT::G;
            *reinterpret_cast<V*>(sl + row * W + j) = r[i];
        }
    }
}

template <typename T, int ROWS, int W, int NT, bool VEC>
__device__ __forceinline__ void rows_to_lds(cx<T>* __restrict__ sl, const cx<T>* __restrict__ base, RowMap rm, int wv)
{
    typename RowTile<T, ROWS, W, NT, VEC>::V r[RowTile<T, ROWS, W, NT, VEC>::PER];
    rows_load<T, ROWS, W, NT, VEC>(r, base, rm, wv);
    rows_commit<T, ROWS, W, NT, VEC>(sl, r);
}

template <typename T, int ROWS, int W, int NT, bool VEC>
__device__ __forceinline__ void lds_to_rows(const cx<T>* __restrict__ sl, cx<T>* __restrict__ base, RowMap rm, int wv)
{
    using V = stage_t<T, VEC>;
    constexpr int G = VEC ? 2 : 1;
    constexpr int WG = W / G;
    constexpr int N = ROWS * WG;
    constexpr int PER = (N + NT - 1) / NT;
    const int t0 = lane_tid();
    if (rows_fit32<T, ROWS>(rm)) {
        const unsigned sb = (unsigned)(rm.a * (long long)sizeof(cx<T>)); // < 2^24
        char* bb = reinterpret_cast<char*>(base);
        const bool full = wv == W;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int e = t0 + i * NT;
            if (N % NT == 0 || e < N) {
                const int row = cdiv<WG, PER * NT>(e), jj = e - row * WG;
                if (full || jj * G < wv)
                    *reinterpret_cast<V*>(bb + (mul_u24((unsigned)row, sb) + (unsigned)(jj * (int)sizeof(V)))) =
                        *reinterpret_cast<const V*>(sl + row * W + jj * G);
            }
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = t0 + i * NT;
        if (N % NT == 0 || e < N) {
            const int row = cdiv<WG, PER * NT>(e), j = (e - row * WG) * G;
            if (j < wv)
                *reinterpret_cast<V*>(base + rm.template off<ROWS>(row) + j) = *reinterpret_cast<const V*>(sl + row * W + j);
        }
    }
}

// spectrum chunk of TY rows [yl][kz][c] (row yl at yl * RS) <-> LDS
// sl[kz*W + yl*C + c]; pairs (c, c+1) move as one 16-byte access for even C
// in FP32.  The load is split in an issue step (global -> registers) and a
// commit step (registers -> LDS) so other loads can be issued in between.
template <typename T, int TY, int NZH, int C, int NT>
struct ZChunk {
    static constexpr int W = TY * C;
    static constexpr bool VEC = (sizeof(T) == 4) && (C % 2 == 0);
    static constexpr int G = VEC ? 2 : 1;
    static constexpr int N = TY * NZH * C / G;
    static constexpr int PER = (N + NT - 1) / NT;
    using V = stage_t<T, VEC>;
};

template <typename Z, int NZH, int C, typename T, typename V>
__device__ __forceinline__ void zchunk_issue(V (&r)[Z::PER], const cx<T>* __restrict__ src, long long RS)
{
    constexpr int NT = ZNT;
#pragma unroll
    for (int i = 0; i < Z::PER; ++i) {
        const int e = threadIdx.x + i * NT;
        if (Z::N % NT == 0 || e < Z::N) {
            const int f = e * Z::G;
            const int yl = f / (NZH * C);
            const int rr = f - yl * (NZH * C);
            r[i] = *reinterpret_cast<const V*>(src + yl * RS + rr);
        }
    }
}

template <typename Z, int NZH, int C, typename T, typename V>
__device__ __forceinline__ void zchunk_commit(const V (&r)[Z::PER], cx<T>* __restrict__ sl)
{
    constexpr int NT = ZNT;
#pragma unroll
    for (int i = 0; i < Z::PER; ++i) {
        const int e = threadIdx.x + i * NT;
        if (Z::N % NT == 0 || e < Z::N) {
            const int f = e * Z::G;
            const int yl = f / (NZH * C);
            const int rr = f - yl * (NZH * C);
            const int kz = rr / C, c = rr - kz * C;
            *reinterpret_cast<V*>(sl + kz * Z::W + yl * C + c) = r[i];
        }
    }
}

template <typename T, int TY, int NZH, int C, int NT>
__device__ __forceinline__ void lds_to_zchunk(const cx<T>* __restrict__ sl, cx<T>* __restrict__ dst, long long RS)
{
    constexpr int W = TY * C;
    constexpr bool VEC = (sizeof(T) == 4) && (C % 2 == 0);
    constexpr int G = VEC ? 2 : 1;
    constexpr int N = TY * NZH * C / G;
    constexpr int PER = (N + NT - 1) / NT;
    using V = typename std::conditional<VEC, float4, cx<T>>::type;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int e = threadIdx.x + i * NT;
        if (N % NT == 0 || e < N) {
            const int f = e * G;
            const int yl = f / (NZH * C);
            const int rr = f - yl * (NZH * C);
            const int kz = rr / C, c = rr - kz * C;
            *reinterpret_cast<V*>(dst + yl * RS + rr) = *reinterpret_cast<const V*>(sl + kz * W + yl * C + c);
        }
    }
}

// ------------------------------------------------------------------ z pass
// Block: solve s, plane x, lines y0..y0+TY-1.  LDS: (LZ+1)*W complex lines
// (W = TY*C, line l = yl*C + c) + the Nz twiddle table.
template <typename T, typename PH, int MODE, int LZ, int TY>
__device__ __forceinline__ void kz_tile(const Geo& g, const Ptr<T>& p, int b, cx<T>* sl, double* red)
{
    constexpr int C = PH::C;
    constexpr int W = TY * C;
    constexpr int NZ = 2 * LZ;
    constexpr int NZH = LZ + 1;
    constexpr int NT = ZNT;
    cx<T>* tw = sl + NZH * W;

    const int per = g.Nx * (g.Ny / TY);
    const int s = b / per;
    const int t = b - s * per;
    const int x = t / (g.Ny / TY);
    const int y0 = (t - x * (g.Ny / TY)) * TY;
    constexpr bool IUF = MODE == MODE_IUF || MODE == MODE_IUFR || MODE == MODE_IUA;
    constexpr bool ANC = MODE == MODE_IUA; // anchor iteration: no forward transform
    constexpr bool CHK = MODE == MODE_CHK_CUR || MODE == MODE_CHK_PREV;
    const Ctl& ctl = p.ctl[s];
    if (ctl.done) return;
    if (MODE == MODE_CHK_CUR && !ctl.need_cur) return;
    if (MODE == MODE_CHK_PREV && !ctl.need_prev) return;
    const int it = ctl.it;
    const int tid = threadIdx.x;

    const long long vox0 = ((long long)x * g.Ny + y0) * NZ;
    // first spectrum row (x, y0) of the tile and the row stride
    const long long RS = g.sy;
    const long long spec0 = (long long)(x >> g.lxb) * g.sxa + (long long)(x & g.xbm) * g.sxb + y0 * RS;
    const int rv = p.rve_of[s];
    const uint8_t* ph = p.ph + (long long)rv * p.sPh + vox0;
    const T* coef = p.coef + (long long)rv * p.sCoef;
    constexpr bool FWD = MODE == MODE_FWD || MODE == MODE_FWDP;
    const T* Uin;
    T* Uout = nullptr;
    // FWDP: the FP32 iterate `it` (FP32 ring), promoted to the FP64 iterate `it`
    const float* Uin32 = nullptr;
    T* Uprom = nullptr;
    if constexpr (MODE == MODE_FWDP) {
        Uin32 = e_buf(p, it, s) + vox0 * C;
        Uprom = u_buf(p, it, s) + vox0 * C;
    }
    if (FWD || MODE == MODE_CHK_CUR) Uin = u_buf(p, it, s);
    else if (IUF) { Uin = u_buf(p, it, s); Uout = u_buf(p, it + 1, s); }
    else Uin = u_buf(p, it - 1, s);
    // anchored polish: the FP64 check reads anchor + FP32 correction
    const float* Ein = nullptr;
    if constexpr (sizeof(T) == 8 && CHK) {
        if (ctl.corr) {
            Uin = u_buf(p, ctl.anc_it, s);
            Ein = e_buf(p, MODE == MODE_CHK_CUR ? it : it - 1, s) + vox0 * C;
        }
    }
    const float* Rin = MODE == MODE_IUFR ? p.R + (long long)s * p.sU + vox0 * C : nullptr;
    // anchor iteration: r and the correction e of iterate it + 1 (FP32 ring)
    float* Ranc = nullptr;
    float* Eanc = nullptr;
    if constexpr (ANC) {
        Ranc = const_cast<float*>(p.R) + (long long)s * p.sU + vox0 * C;
        Eanc = const_cast<float*>(e_buf(p, it + 1, s)) + vox0 * C;
    }
    Uin += vox0 * C;
    if (Uout) Uout += vox0 * C;

    // issue every global load of the tile first: spectrum chunk, then the
    // unknowns and phase ids of this thread's voxel pairs
    constexpr int NPAIR = TY * LZ;
    constexpr int QI = (NPAIR + NT - 1) / NT;
    for (int i = tid; i < NZ; i += NT) tw[i] = p.twz[i];
    using ZC = ZChunk<T, TY, NZH, C, NT>;
    typename ZC::V chunk[ZC::PER];
    if (IUF) zchunk_issue<ZC, NZH, C>(chunk, p.S + (long long)s * p.sS + spec0, RS);
    T u[QI][2 * C];
    T rr[MODE == MODE_IUFR ? QI : 1][2 * C];
    uint16_t pp[QI];
#pragma unroll
    for (int i = 0; i < QI; ++i) {
        const int q = tid + i * NT;
        if (NPAIR % NT == 0 || q < NPAIR) {
            const int yl = q / LZ, m = q - yl * LZ;
            const int v0 = yl * NZ + 2 * m;
            if constexpr (MODE == MODE_FWDP) {
                float f[2 * C];
                vload<float, 2 * C>(Uin32 + (long long)v0 * C, f);
#pragma unroll
                for (int c = 0; c < 2 * C; ++c) u[i][c] = (T)f[c];
                vstore<T, 2 * C>(Uprom + (long long)v0 * C, u[i]);
            } else {
                vload<T, 2 * C>(Uin + (long long)v0 * C, u[i]);
            }
            if constexpr (MODE == MODE_IUFR) vload<T, 2 * C>(Rin + (long long)v0 * C, rr[i]);
            pp[i] = *reinterpret_cast<const uint16_t*>(ph + v0);
            if constexpr (sizeof(T) == 8 && CHK) {
                if (Ein) {
                    float e2[2 * C];
                    vload<float, 2 * C>(Ein + (long long)v0 * C, e2);
#pragma unroll
                    for (int c = 0; c < 2 * C; ++c) u[i][c] += (T)e2[c];
                }
            }
        }
    }

    if (IUF) {
        zchunk_commit<ZC, NZH, C>(chunk, sl);
        __syncthreads();
        // c2r pre-processing; kz=0 and kz=Nz/2 keep their real parts only
        constexpr int half = LZ / 2 + 1;
        for (int e = tid; e < W * half; e += NT) {
            const int l = e % W, k = e / W;
            if (k == 0) {
                const T x0 = sl[l].re, xl = sl[LZ * W + l].re;
                sl[l] = cx<T>{x0 + xl, x0 - xl};
            } else {
                const int kk = LZ - k;
                const cx<T> Xk = sl[k * W + l], Xc = sl[kk * W + l];
                cx<T> w = tw[k];
                w.im = -w.im;
                sl[k * W + l] = (Xk + cconj(Xc)) + rot90<1>(cmul(w, Xk - cconj(Xc)));
                if (kk != k) {
                    cx<T> w2 = tw[kk];
                    w2.im = -w2.im;
                    sl[kk * W + l] = (Xc + cconj(Xk)) + rot90<1>(cmul(w2, Xc - cconj(Xk)));
                }
            }
        }
        __syncthreads();
        fft<T, LZ, W, NT, 1, 2>(sl, tw);
    } else {
        __syncthreads();
    }

    // per voxel pair (2m, 2m+1) of line yl
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < QI; ++i) {
        const int q = tid + i * NT;
        if (NPAIR % NT == 0 || q < NPAIR) {
            const int yl = q / LZ, m = q - yl * LZ;
            const int v0 = yl * NZ + 2 * m;
            T u0[C], u1[C], o0[C], o1[C];
#pragma unroll
            for (int c = 0; c < C; ++c) { u0[c] = u[i][c]; u1[c] = u[i][C + c]; }
            const T* k0 = coef + (int)(pp[i] & 0xff) * NCOEF;
            const T* k1 = coef + (int)(pp[i] >> 8) * NCOEF;
            cx<T>* slot = sl + m * W + yl * C;
            if (IUF) {
                T d0[C], d1[C], a0[C], a1[C];
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    const cx<T> uc = slot[c];
                    d0[c] = uc.re - u0[c];
                    d1[c] = uc.im - u1[c];
                }
                PH::accel(k0, d0, a0);
                PH::accel(k1, d1, a1);
                T un[2 * C];
#pragma unroll
                for (int c = 0; c < C; ++c) {
                    u0[c] += a0[c];
                    u1[c] += a1[c];
                    if constexpr (MODE == MODE_IUFR) {
                        u0[c] += rr[i][c];
                        u1[c] += rr[i][C + c];
                    }
                    un[c] = u0[c];
                    un[C + c] = u1[c];
                }
                vstore<T, 2 * C>(Uout + (long long)v0 * C, un);
                if constexpr (ANC) {
                    // r = T(A) - A in FP64, kept in FP32 (the difference of the stored iterates)
                    float d[2 * C];
#pragma unroll
                    for (int c = 0; c < 2 * C; ++c) d[c] = (float)(un[c] - u[i][c]);
                    vstore<float, 2 * C>(Ranc + (long long)v0 * C, d);
                    vstore<float, 2 * C>(Eanc + (long long)v0 * C, d);
                }
            }
            if (FWD || IUF) {
                PH::dual(k0, u0, o0);
                PH::dual(k1, u1, o1);
#pragma unroll
                for (int c = 0; c < C; ++c) acc += (double)o0[c] + (double)o1[c];
                if constexpr (!ANC) {
                    PH::polar(k0, u0, o0);
                    PH::polar(k1, u1, o1);
                }
            } else if (ctl.sb) {
                PH::dual(k0, u0, o0);
                PH::dual(k1, u1, o1);
            } else {
#pragma unroll
                for (int c = 0; c < C; ++c) { o0[c] = u0[c]; o1[c] = u1[c]; }
            }
            if constexpr (!ANC) {
#pragma unroll
                for (int c = 0; c < C; ++c) slot[c] = cx<T>{o0[c], o1[c]};
            }
        }
    }
    if (FWD || IUF) {
        const double bs = block_sum(acc, red);
        if (tid == 0) p.psum[(long long)s * g.nbz + t] = bs;
    }
    if constexpr (ANC) return;
    __syncthreads();

    fft<T, LZ, W, NT, -1, 2>(sl, tw);
    {
        constexpr int half = LZ / 2 + 1;
        for (int e = tid; e < W * half; e += NT) {
            const int l = e % W, k = e / W;
            if (k == 0) {
                const cx<T> z0 = sl[l];
                sl[l] = cx<T>{z0.re + z0.im, T(0)};
                sl[LZ * W + l] = cx<T>{z0.re - z0.im, T(0)};
            } else {
                const int kk = LZ - k;
                const cx<T> zk = sl[k * W + l], zc = cconj(sl[kk * W + l]);
                const cx<T> E = cscale(zk + zc, T(0.5));
                const cx<T> O = cscale(rot90<-1>(zk - zc), T(0.5));
                sl[k * W + l] = E + cmul(tw[k], O);
                if (kk != k) sl[kk * W + l] = cconj(E) + cmul(tw[kk], cconj(O));
            }
        }
    }
    __syncthreads();
    {
        cx<T>* dst = ((FWD || IUF) ? p.S : p.K) + (long long)s * p.sS + spec0;
        lds_to_zchunk<T, TY, NZH, C, NT>(sl, dst, RS);
    }
}

// one solve's check is needed (the check passes are captured in every
// iteration and gated on these flags)
template <int WHICH>
__device__ __forceinline__ bool chk_needed(const Ctl& c)
{
    return !c.done && (WHICH == BUF_CHK_CUR ? c.need_cur : c.need_prev);
}

// The main passes run one workgroup per tile.  The check passes run a capped
// grid that loops over the tiles: when no solve needs the check, the whole
// launch ends after a few thousand workgroups instead of dispatching one per
// tile only to exit.
template <typename T, typename PH, int MODE, int LZ, int TY>
__global__ void __launch_bounds__(ZNT) M3D_KZ_ATTR kz(Geo g, Ptr<T> p)
{
    constexpr int NT = ZNT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[NT / 64];
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    constexpr bool CHK = MODE == MODE_CHK_CUR || MODE == MODE_CHK_PREV;
    if constexpr (CHK) {
        constexpr int WHICH = MODE == MODE_CHK_CUR ? BUF_CHK_CUR : BUF_CHK_PREV;
        int mine = 0;
        for (int s = threadIdx.x; s < g.nsolve; s += NT) mine |= chk_needed<WHICH>(p.ctl[s]);
        if (!__syncthreads_or(mine)) return;
        const int total = g.nsolve * g.Nx * (g.Ny / TY);
        // a workgroup stays on XCD blockIdx.x % 8 and the grid is a multiple
        // of 8 (or covers every tile once): the remap hands each XCD its
        // contiguous eighth of the tiles, as for the one-tile launches
        for (int b = blockIdx.x; b < total; b += gridDim.x) {
            kz_tile<T, PH, MODE, LZ, TY>(g, p, xcd_split(b, total, g.Nx * (g.Ny / TY)), sl, red);
            __syncthreads();
        }
    } else {
        kz_tile<T, PH, MODE, LZ, TY>(g, p, xcd_split(blockIdx.x, gridDim.x, g.Nx * (g.Ny / TY)), sl, red);
    }
}

// --------------------------------------------------------- y / x passes
// KIND 0: y pass (DIR), KIND 1: x pass with Green step, KIND 2: x pass of the check FFT

// tile (s, o, kz0) of a y or x pass: its first spectrum row and row map
template <typename T, int C, int KIND, int WHICH>
__device__ __forceinline__ const cx<T>* xy_rows(const Geo& g, const Ptr<T>& p, int s, int o, int kz0, RowMap& rm)
{
    if (KIND == 0) {
        // tile: plane x = o, lines along y (row y at y * sy)
        rm = RowMap{g.sy, 0, 0, 0};
        return (WHICH == BUF_MAIN ? p.S : p.K) + (long long)s * p.sS + (long long)(o >> g.lxb) * g.sxa +
               (long long)(o & g.xbm) * g.sxb + (long long)kz0 * C;
    }
    // tile: row y = o, lines along x; row (x, y) at (x >> lxb) * sxa + (x & xbm) * sxb + y * sy
    rm = (g.lxb == 0 && g.xbm == 0) ? RowMap{g.sxa, 0, 0, 0} : RowMap{g.sxa, g.sxb, g.lxb, g.xbm};
    return (KIND == 1 ? p.S : p.K) + (long long)s * p.sS + (long long)o * g.sy + (long long)kz0 * C;
}

// a physics type may bring its own Green step for the x pass:
// PH::green<T>(g, ctl, prm, kx, y, kz, v) on one frequency's values in LDS,
// DC included
template <typename P, typename = void>
struct has_green_hook : std::false_type {};
template <typename P>
struct has_green_hook<P, std::void_t<decltype(&P::template green<double>)>> : std::true_type {};

// a tile already in LDS: forward FFT along the lines, then the Green step and
// the inverse FFT (KIND 1) or the equilibrium partial sums (KIND 2; t is the
// tile's index within its solve).  Returns with the tile in LDS (KIND 0, 1).
template <typename T, typename PH, int KIND, int DIR, int WHICH, int L, int TK, int NT, int RP = 0>
__device__ __forceinline__ void xy_tile(const Geo& g, const Ptr<T>& p, const Ctl& ctl, const T* __restrict__ prm, cx<T>* sl,
                                        const cx<T>* tw, double* red, int s, int o, int kz0, int nk, int t)
{
    constexpr int C = PH::C;
    constexpr int W = TK * C;
    const int tid = lane_tid();
    fft<T, L, W, NT, (KIND == 0 ? DIR : -1), 1, 0, 1, W, W, 0, 0, RP>(sl, tw);

    if constexpr (KIND == 1) {
        const int y = o;
        GreenK<T> K{};
        if constexpr (!has_green_hook<PH>::value) K = green_k_of<T, C>(prm);
        auto step = [&](int kx, int kzl) {
            const int kz = kz0 + kzl;
            cx<T>* v = sl + kx * W + kzl * C;
            if constexpr (has_green_hook<PH>::value) {
                PH::template green<T>(g, ctl, prm, kx, y, kz, v);
            } else if (kx == 0 && y == 0 && kz == 0) {
#pragma unroll
                for (int c = 0; c < C; ++c) v[c] = cx<T>{T(ctl.E[c]), T(0)};
            } else {
                green_apply_k<T, PH>(g, ctl.sb, K, kx, y, kz, v);
            }
        };
        // full kz tiles (all but the last) divide by the constant TK
        if (nk == TK) {
            for (int f = tid; f < L * TK; f += NT) {
                const int kx = cdiv<TK, L * TK + NT>(f);
                step(kx, f - kx * TK);
            }
        } else {
            for (int f = tid; f < L * nk; f += NT) {
                const int kx = f / nk;
                step(kx, f - kx * nk);
            }
        }
        __syncthreads();
        fft<T, L, W, NT, 1, 1, 0, 1, W, W, 0, 0, RP>(sl, tw);
    }
    if constexpr (KIND == 2) {
        const int y = o;
        double acc = 0.0;
        for (int f = tid; f < L * nk; f += NT) {
            const int kx = nk == TK ? cdiv<TK, L * TK + NT>(f) : f / nk, kzl = f - kx * nk;
            const int kz = kz0 + kzl;
            const cx<T>* v = sl + kx * W + kzl * C;
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
        if (tid == 0) p.pchk[((long long)s * 2 + (WHICH == BUF_CHK_CUR ? 0 : 1)) * (g.nbc + 1) + t] = bs;
    }
}

// solve s takes part in this pass
template <int WHICH>
__device__ __forceinline__ bool xy_active(const Ctl& c)
{
    if (c.done) return false;
    if (WHICH == BUF_CHK_CUR && !c.need_cur) return false;
    if (WHICH == BUF_CHK_PREV && !c.need_prev) return false;
    return true;
}

// one workgroup per tile
template <typename T, typename PH, int KIND, int DIR, int WHICH, int L, int TK, int NTH = xy_threads<KIND, T>(), int RP = 0>
__device__ __forceinline__ void xy_body(const Geo& g, const Ptr<T>& p, const T* __restrict__ prm_all, int b, cx<T>* sl,
                                        double* red)
{
    constexpr int C = PH::C;
    constexpr int W = TK * C;
    constexpr int NT = NTH;
    cx<T>* tw = sl + L * W;

    const int nkt = (g.Nzh + TK - 1) / TK;
    const int lines_out = KIND == 0 ? g.Nx : g.Ny; // outer index of the tile
    const int per = lines_out * nkt;
    const int s = b / per;
    const int t = b - s * per;
    const int o = t / nkt;
    const int kz0 = (t - o * nkt) * TK;
    if (!xy_active<WHICH>(p.ctl[s])) return;
    const int nk = min(TK, g.Nzh - kz0);
    const int wv = nk * C;

    RowMap rm;
    cx<T>* base = const_cast<cx<T>*>(xy_rows<T, C, KIND, WHICH>(g, p, s, o, kz0, rm));
    const cx<T>* twg = KIND == 0 ? p.twy : p.twx;
    // FP32 tiles move pairs of complex values (16 bytes) per access: always for
    // even C; for odd C (thermal) when the tile's row chunks are pairs of
    // 16-byte-aligned values (even row pitch, even chunk width), a uniform
    // test per tile, else single values
    constexpr bool VEC_EVEN = (sizeof(T) == 4) && (C % 2 == 0);
    constexpr bool VEC_ODD = (sizeof(T) == 4) && (C % 2 != 0) && ((TK * C) % 2 == 0);
    bool vec = VEC_EVEN;
    if constexpr (VEC_ODD)
        vec = (reinterpret_cast<unsigned long long>(base) & 15) == 0 && rm.sh == 0 && rm.mask == 0 && (rm.a & 1) == 0 &&
              (wv & 1) == 0;
    auto load = [&](auto vc) {
        constexpr bool VEC = decltype(vc)::value;
        // the tile's loads are issued before the twiddle table is copied to LDS
        using RT = RowTile<T, L, W, NT, VEC>;
        typename RT::V r[RT::PER];
        rows_load<T, L, W, NT, VEC>(r, base, rm, wv);
        for (int i = threadIdx.x; i < L; i += NT) tw[i] = twg[i];
        rows_commit<T, L, W, NT, VEC>(sl, r);
    };
    if constexpr (VEC_ODD) {
        if (vec) load(std::true_type{});
        else load(std::false_type{});
    } else {
        load(std::integral_constant<bool, VEC_EVEN>{});
    }
    __syncthreads();
    // the Green step's reference medium (KIND 1 only: other users leave rve_of unset)
    const T* prm = nullptr;
    if constexpr (KIND == 1) prm = prm_all + 2 * p.rve_of[s];
    xy_tile<T, PH, KIND, DIR, WHICH, L, TK, NT, RP>(g, p, p.ctl[s], prm, sl, tw, red, s, o, kz0, nk, t);
    if constexpr (KIND != 2) {
        if constexpr (VEC_ODD) {
            if (vec) lds_to_rows<T, L, W, NT, true>(sl, base, rm, wv);
            else lds_to_rows<T, L, W, NT, false>(sl, base, rm, wv);
        } else {
            lds_to_rows<T, L, W, NT, VEC_EVEN>(sl, base, rm, wv);
        }
    }
}

// one workgroup per tile; LOOP (the solver's check passes only): a capped
// grid loops over the tiles of every solve and leaves at once when no solve
// needs the check (see kz)
template <typename T, typename PH, int KIND, int DIR, int WHICH, int L, int TK, int NTH = xy_threads<KIND, T>(),
          bool LOOP = false, int RP = 0>
__global__ void __launch_bounds__(NTH) M3D_KXY_ATTR kxy(Geo g, Ptr<T> p, const T* __restrict__ prm_all)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ double red[NTH / 64];
    cx<T>* sl = reinterpret_cast<cx<T>*>(smem_raw);
    static_assert(!LOOP || WHICH != BUF_MAIN, "only the check passes loop over their tiles");
    if constexpr (LOOP) {
        int mine = 0;
        for (int s = threadIdx.x; s < g.nsolve; s += NTH) mine |= chk_needed<WHICH>(p.ctl[s]);
        if (!__syncthreads_or(mine)) return;
        const int total = g.nsolve * (KIND == 0 ? g.Nx : g.Ny) * ((g.Nzh + TK - 1) / TK);
        for (int b = blockIdx.x; b < total; b += gridDim.x) {
            xy_body<T, PH, KIND, DIR, WHICH, L, TK, NTH, RP>(g, p, prm_all,
                                                              xcd_split(b, total, total / max(1, g.nsolve)), sl, red);
            __syncthreads();
        }
    } else {
        xy_body<T, PH, KIND, DIR, WHICH, L, TK, NTH, RP>(g, p, prm_all,
                                                              xcd_split(blockIdx.x, gridDim.x, (int)gridDim.x / max(1, g.nsolve)), sl,
                                                              red);
    }
}

} // namespace fast
} // namespace m3d
