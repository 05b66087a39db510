// m3d_perc.hip -- cluster statistics of one phase (src/percolation/percolation.cpp)
// on the GPU.
//
// The reference labels the voxels of the interest phase 1, 2, ... in raster
// order and then, in 5 rounds of 8 sweep orders times 6 trigger tests, gives
// every interest voxel of the 2x2x2 cube whose largest corner is (i, j, k)
// the smallest label among them whenever one of the six index permutations of
// (i, j, k) is an interest voxel (percolation.cpp:75-114, :117-286).  The
// trigger does not depend on the labels, so the set of cubes that merge is
// fixed by the phase field alone; once the sweeps stop changing anything,
// every connected set of the graph "interest voxels that share an active
// cube" carries the label of its first voxel in raster order.  This file
// computes that fixed point directly:
//
//   k_init     parent = first voxel of the voxel's run along z, flags, counts
//   k_link     union-find over the active cubes, skipping joins that other
//              joins imply; the roots are lowered by atomic min, so every
//              root is the smallest index of its set
//   k_count / k_scan_top / k_enum
//              roots numbered in raster order (the clusters' first-appearance
//              order, :318-336) with the raster rank of each root among the
//              interest voxels (the reference's label)
//   k_sizes    every interest voxel points at its root; voxels per cluster
//   k_hist / scan / k_rank
//              counting sort of the sizes: rank = number of clusters that are
//              strictly larger (std::find of the cluster's fraction in the
//              descending list, :355-376), and the descending fractions
//   k_relabel  interest voxels = phasePlus + rank, other voxels keep their id
//
// On a phase field where the reference's 240 sweeps reach their fixed point
// the outputs are the reference's bit for bit; where they stop short of it
// (long winding clusters), the reference leaves one connected set split over
// several labels and this file reports the set as one cluster.
//
// SurfaceOfTwoPhase (:384-414) counts, for every voxel of phase1, its -x,
// -y and -z face neighbours of phase2, with -1 outside the box.
#include "m3d_engine.hpp"

#include <climits>
#include <cstring>
#include <string>
#include <vector>

namespace m3d {
namespace perc {

constexpr int NT = 256;          // threads per workgroup
constexpr int TPB = 16;          // rounds of NT elements per counting block
constexpr int BT = NT * TPB;     // elements per counting block
constexpr int NTOP = 1024;       // threads of the one-workgroup scan

__device__ __forceinline__ int ld(const int* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// parent chains decrease strictly, so the walk ends at a self-parent
__device__ __forceinline__ int find_root(const int* lab, int x)
{
    int p = ld(lab + x);
    while (p != x) {
        x = p;
        p = ld(lab + x);
    }
    return x;
}

// joins the sets of a and b by hooking the larger root under the smaller
// (atomic min); a lost race leaves the hooked root's new parent in `old`, and
// the loop goes on with it, so no link is dropped.  Every pass either ends or
// lowers one of the two values, so the loop ends.
__device__ void unite(int* lab, int a, int b)
{
    bool done;
    do {
        a = find_root(lab, a);
        b = find_root(lab, b);
        if (a < b) {
            const int old = atomicMin(lab + b, a);
            done = (old == b);
            b = old;
        } else if (b < a) {
            const int old = atomicMin(lab + a, b);
            done = (old == a);
            a = old;
        } else
            done = true;
    } while (!done);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// number of set lanes below this one in a ballot
__device__ __forceinline__ int below(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// adds the workgroup's sum of cnt to red[0] and its maximum of mx to the int
// at red + 1, one atomic each per workgroup
__device__ void block_total(unsigned long long cnt, int mx, unsigned long long* red)
{
    __shared__ unsigned long long sc[NT / 64];
    __shared__ int sm[NT / 64];
    cnt = wave_sum(cnt);
    mx = wave_max(mx);
    if (lane_id() == 0) {
        sc[threadIdx.x >> 6] = cnt;
        sm[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / 64; ++w) {
            cnt += sc[w];
            mx = max(mx, sm[w]);
        }
        if (cnt) atomicAdd(red, cnt);
        atomicMax(reinterpret_cast<int*>(red + 1), mx);
    }
}

__global__ void k_setup(unsigned long long* red)
{
    if (threadIdx.x == 0) {
        red[0] = 0;                                   // interest voxels
        reinterpret_cast<int*>(red + 1)[0] = INT_MIN; // largest other id
        red[2] = 0;                                   // surface count
    }
}

// one wave per z row (x, y): interest voxels start with the first voxel of
// their run along z as parent (z neighbours always share an active cube:
// the one whose largest corner is the upper of the two), other voxels -1
__global__ void __launch_bounds__(NT) k_init(const int32_t* ids, int nx, int ny, int nz, int32_t phase, int32_t* lab,
                                             uint8_t* flag, unsigned long long* red)
{
    const int lane = lane_id();
    const long long rows = (long long)nx * ny, wst = (long long)gridDim.x * (NT / 64);
    unsigned long long cnt = 0;
    int mx = INT_MIN;
    for (long long row = blockIdx.x * (long long)(NT / 64) + (threadIdx.x >> 6); row < rows; row += wst) {
        const long long r0 = row * nz;
        int carry = -1; // start of the run that reaches the end of the previous chunk
        for (int z0 = 0; z0 < nz; z0 += 64) {
            const int z = z0 + lane;
            const bool ok = z < nz;
            const int id = ok ? ids[r0 + z] : 0;
            const bool in = ok && id == phase;
            const unsigned long long m = __ballot(in);
            const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
            const unsigned long long zm = ~m & upto;
            int start;
            if (zm)
                start = z0 + (63 - __clzll((long long)zm)) + 1;
            else
                start = carry >= 0 ? carry : z0;
            if (ok) {
                lab[r0 + z] = in ? (int)(r0 + start) : -1;
                flag[r0 + z] = in;
            }
            cnt += in;
            if (ok && !in) mx = max(mx, id);
            const int s63 = __shfl(start, 63, 64);
            carry = (m >> 63) & 1 ? s63 : -1;
        }
    }
    block_total(cnt, mx, red);
}

// offsets e in 1..7 (bits x, y, z) that contain d or lie in d
__device__ __forceinline__ unsigned nested(int d)
{
    unsigned r = 0;
    for (int e = 1; e < 8; ++e)
        if (e != d && ((d & e) == d || (d & e) == e)) r |= 1u << e;
    return r;
}

// one thread per cube; the cube (x, y, z) covers the voxels (x-1..x,
// y-1..y, z-1..z) of the box (coordinates -1 are the margin) and is active
// when one of the six permutations of (x, y, z) is an interest voxel.
//
// When the cube's largest corner v is an interest voxel, v is joined only to
// the corners that no other union already reaches (every dependency below
// is on a voxel with smaller coordinates, so the argument is well founded):
//   - the -z face neighbour: its parent chain already joins v (z runs);
//   - a corner c = v - d is reached through the face neighbour n_i = v - e_i
//     when d_i = 1, n_i is an interest voxel and n_i != c: c lies in the cube
//     of n_i, which is active because n_i is an interest voxel;
//   - a corner with d_z = 0 is reached when v - e_z and c - e_z are both
//     interest voxels: v - e_z joins c - e_z (directly or by this same rule
//     one step lower), and z runs join v and c to them.
// Otherwise (v not an interest voxel) see below.
__global__ void __launch_bounds__(NT) k_link(const uint8_t* f, int32_t* lab, int N)
{
    const long long n = (long long)N * N * N;
    const long long v = blockIdx.x * (long long)NT + threadIdx.x;
    if (v >= n) return;
    const int z = (int)(v % N);
    const long long t = v / N;
    const int y = (int)(t % N), x = (int)(t / N);
    auto I = [N](int a, int b, int c) { return ((long long)a * N + b) * N + c; };
    const long long sx = (long long)N * N, sy = N;
    if (f[v]) {
        const bool hx = x > 0 && f[v - sx], hy = y > 0 && f[v - sy], hz = z > 0 && f[v - 1];
        for (int d = 1; d < 8; ++d) {
            const int dx = (d >> 2) & 1, dy = (d >> 1) & 1, dz = d & 1;
            if (d == 1) continue; // -z neighbour: same run
            if ((dx && x == 0) || (dy && y == 0) || (dz && z == 0)) continue;
            const long long c = v - dx * sx - dy * sy - dz;
            if (!f[c]) continue;
            const bool face = dx + dy + dz == 1;
            if (!face && ((dx && hx) || (dy && hy) || (dz && hz))) continue;
            if (!dz && hz && f[c - 1]) continue;
            unite(lab, (int)v, (int)c);
        }
        return;
    }
    // v is not an interest voxel: corner v - d1 already reaches v - d2 when
    // d1 is a subset of d2 (the second lies in the cube of the first, which
    // is active); only one corner of each group so formed is joined, and
    // only when the cube is active and there are two groups or more
    unsigned m = 0;
    for (int d = 1; d < 8; ++d) {
        const int dx = (d >> 2) & 1, dy = (d >> 1) & 1, dz = d & 1;
        if ((dx && x == 0) || (dy && y == 0) || (dz && z == 0)) continue;
        if (f[v - dx * sx - dy * sy - dz]) m |= 1u << d;
    }
    if (__popc(m) < 2) return;
    unsigned rem = m, reps = 0;
    while (rem) {
        const int d0 = __ffs(rem) - 1;
        unsigned comp = 1u << d0;
        for (;;) {
            unsigned g = comp;
            for (int d = 1; d < 8; ++d)
                if ((comp >> d) & 1) g |= nested(d) & m;
            if (g == comp) break;
            comp = g;
        }
        reps |= 1u << d0;
        rem &= ~comp;
    }
    if (__popc(reps) < 2) return;
    const bool act = f[I(x, z, y)] | f[I(y, x, z)] | f[I(y, z, x)] | f[I(z, y, x)] | f[I(z, x, y)];
    if (!act) return;
    int a = -1;
    for (int d = 1; d < 8; ++d) {
        if (!((reps >> d) & 1)) continue;
        const long long c = v - ((d >> 2) & 1) * sx - ((d >> 1) & 1) * sy - (d & 1);
        if (a < 0)
            a = (int)c;
        else
            unite(lab, a, (int)c);
    }
}

// interest voxels (ci) and roots (cr) per block of BT elements
__global__ void __launch_bounds__(NT) k_count(const uint8_t* f, const int32_t* lab, long long n, int* ci, int* cr)
{
    __shared__ int s[2][NT / 64];
    int a = 0, b = 0;
    for (int r = 0; r < TPB; ++r) {
        const long long e = blockIdx.x * (long long)BT + r * NT + threadIdx.x;
        if (e < n && f[e]) {
            a += 1;
            b += lab[e] == (int)e;
        }
    }
    a = (int)wave_sum((unsigned long long)a);
    b = (int)wave_sum((unsigned long long)b);
    if (lane_id() == 0) {
        s[0][threadIdx.x >> 6] = a;
        s[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int ta = 0, tb = 0;
        for (int w = 0; w < NT / 64; ++w) {
            ta += s[0][w];
            tb += s[1][w];
        }
        ci[blockIdx.x] = ta;
        cr[blockIdx.x] = tb;
    }
}

// sum of v over the values of a block of NT threads, and this thread's
// exclusive prefix; `s` holds NT/64 + 1 ints
template <int W>
__device__ __forceinline__ int block_scan(int v, int* s, int& total)
{
    const int lane = lane_id(), w = threadIdx.x >> 6;
    int inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
    }
    if (lane == 63) s[w] = inc;
    __syncthreads();
    int base = 0, tot = 0;
    for (int k = 0; k < W; ++k) {
        const int sk = s[k];
        base += k < w ? sk : 0;
        tot += sk;
    }
    __syncthreads();
    total = tot;
    return base + inc - v;
}

// exclusive scan in place of a[0..m) (and b, when given), one workgroup;
// the totals go to tot[0] (and tot[1])
__global__ void __launch_bounds__(NTOP) k_scan_top(int* a, int* b, int m, long long* tot)
{
    __shared__ int s[NTOP / 64 + 1];
    long long ra = 0, rb = 0;
    for (int c0 = 0; c0 < m; c0 += NTOP) {
        const int e = c0 + threadIdx.x;
        const int va = e < m ? a[e] : 0;
        int ta;
        const int pa = block_scan<NTOP / 64>(va, s, ta);
        if (e < m) a[e] = (int)(ra + pa);
        ra += ta;
        if (b) {
            const int vb = e < m ? b[e] : 0;
            int tb;
            const int pb = block_scan<NTOP / 64>(vb, s, tb);
            if (e < m) b[e] = (int)(rb + pb);
            rb += tb;
        }
    }
    if (threadIdx.x == 0) {
        tot[0] = ra;
        if (b) tot[1] = rb;
    }
}

// roots numbered in raster order: qm[root] = cluster index q, clab[q] =
// plus + raster rank of the root among the interest voxels
__global__ void __launch_bounds__(NT) k_enum(const uint8_t* f, const int32_t* lab, long long n, const int* ci,
                                             const int* cr, int plus, int32_t* qm, int32_t* clab)
{
    const int lane = lane_id(), w = threadIdx.x >> 6;
    __shared__ int s[2][NT / 64];
    int oi = ci[blockIdx.x], orr = cr[blockIdx.x];
    for (int r = 0; r < TPB; ++r) {
        const long long e = blockIdx.x * (long long)BT + r * NT + threadIdx.x;
        const bool in = e < n && f[e];
        const bool root = in && lab[e] == (int)e;
        const unsigned long long mi = __ballot(in), mr = __ballot(root);
        if (lane == 0) {
            s[0][w] = __popcll(mi);
            s[1][w] = __popcll(mr);
        }
        __syncthreads();
        int bi = 0, br = 0, ti = 0, tr = 0;
        for (int k = 0; k < NT / 64; ++k) {
            bi += k < w ? s[0][k] : 0;
            br += k < w ? s[1][k] : 0;
            ti += s[0][k];
            tr += s[1][k];
        }
        __syncthreads();
        if (root) {
            const int q = orr + br + below(mr);
            qm[e] = q;
            clab[q] = plus + oi + bi + below(mi);
        }
        oi += ti;
        orr += tr;
    }
}

// adds `inc` at idx[lane] for every set lane, one atomic per distinct index
// in the wave
__device__ __forceinline__ void wave_add(int* a, bool has, int idx)
{
    unsigned long long act = __ballot(has);
    while (act) {
        const int leader = __ffsll((long long)act) - 1;
        const int li = __shfl(idx, leader, 64);
        const unsigned long long same = __ballot(has && idx == li);
        if (lane_id() == leader) atomicAdd(a + li, __popcll(same));
        if (has && idx == li) has = false;
        act &= ~same;
    }
}

// voxels per cluster (and every voxel's parent set to its root): each
// workgroup counts BT voxels into a small table in
// LDS (open addressing on the cluster index) and adds the table to csize;
// a cluster that finds no free slot goes to csize directly
constexpr int HSLOTS = 256;

__global__ void __launch_bounds__(NT) k_sizes(const uint8_t* f, int32_t* lab, const int32_t* qm, long long n,
                                              int* csize)
{
    __shared__ int key[HSLOTS], val[HSLOTS];
    for (int i = threadIdx.x; i < HSLOTS; i += NT) {
        key[i] = -1;
        val[i] = 0;
    }
    __syncthreads();
    const int lane = lane_id();
    for (int r = 0; r < TPB; ++r) {
        const long long e = blockIdx.x * (long long)BT + r * NT + threadIdx.x;
        bool has = e < n && f[e];
        int q = 0;
        if (has) {
            // the voxel's root; stored as its parent for k_relabel
            const int r = find_root(lab, (int)e);
            lab[e] = r;
            q = qm[r];
        }
        unsigned long long act = __ballot(has);
        while (act) {
            const int leader = __ffsll((long long)act) - 1;
            const int lq = __shfl(q, leader, 64);
            const unsigned long long same = __ballot(has && q == lq);
            if (lane == leader) {
                const int c = __popcll(same);
                int h = (int)(((unsigned)lq * 2654435761u) >> 24) & (HSLOTS - 1);
                bool put = false;
                for (int p = 0; p < 16 && !put; ++p) {
                    const int k = atomicCAS(key + h, -1, lq);
                    if (k == -1 || k == lq) {
                        atomicAdd(val + h, c);
                        put = true;
                    }
                    h = (h + 1) & (HSLOTS - 1);
                }
                if (!put) atomicAdd(csize + lq, c);
            }
            if (has && q == lq) has = false;
            act &= ~same;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < HSLOTS; i += NT)
        if (key[i] >= 0) atomicAdd(csize + key[i], val[i]);
}

__global__ void __launch_bounds__(NT) k_hist(const int* csize, int K, int* hist)
{
    const int q = blockIdx.x * NT + threadIdx.x;
    const bool has = q < K;
    wave_add(hist, has, has ? csize[q] : 0);
}

__global__ void __launch_bounds__(NT) k_part(const int* a, long long L, int* part)
{
    __shared__ int s[NT / 64];
    int t = 0;
    for (int r = 0; r < TPB; ++r) {
        const long long e = blockIdx.x * (long long)BT + r * NT + threadIdx.x;
        if (e < L) t += a[e];
    }
    t = (int)wave_sum((unsigned long long)(unsigned)t);
    if (lane_id() == 0) s[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        int u = 0;
        for (int w = 0; w < NT / 64; ++w) u += s[w];
        part[blockIdx.x] = u;
    }
}

// out = inclusive prefix sums of a, from the block offsets in part
__global__ void __launch_bounds__(NT) k_apply(const int* a, long long L, const int* part, int* out)
{
    __shared__ int s[NT / 64 + 1];
    int run = part[blockIdx.x];
    for (int r = 0; r < TPB; ++r) {
        const long long e = blockIdx.x * (long long)BT + r * NT + threadIdx.x;
        const int v = e < L ? a[e] : 0;
        int tot;
        const int p = block_scan<NT / 64>(v, s, tot);
        if (e < L) out[e] = run + p + v;
        run += tot;
    }
}

// rank of cluster q = clusters strictly larger; the descending fractions
// fill the positions rank .. rank + (clusters of that size) - 1
__global__ void __launch_bounds__(NT) k_rank(const int* csize, int K, const int* cum, int* hist, double vol, int* rnk,
                                             double* frac)
{
    const int q = blockIdx.x * NT + threadIdx.x;
    if (q >= K) return;
    const int sz = csize[q];
    const int base = K - cum[sz];
    rnk[q] = base;
    const int ord = atomicSub(hist + sz, 1) - 1;
    frac[base + ord] = (double)sz / vol;
}

__global__ void __launch_bounds__(NT) k_relabel(const int32_t* ids, const uint8_t* f, const int32_t* lab,
                                                const int32_t* qm, const int* rnk, int plus, long long n, int32_t* c1)
{
    const long long v = blockIdx.x * (long long)NT + threadIdx.x;
    if (v >= n) return;
    c1[v] = f[v] ? plus + rnk[qm[lab[v]]] : ids[v];
}

__global__ void __launch_bounds__(NT) k_surface(const int32_t* ids, int nx, int ny, int nz, int p1, int p2,
                                                unsigned long long* red)
{
    const long long n = (long long)nx * ny * nz;
    unsigned long long cnt = 0;
    const long long st = (long long)gridDim.x * NT;
    for (long long v = blockIdx.x * (long long)NT + threadIdx.x; v < n; v += st) {
        if (ids[v] != p1) continue;
        const int z = (int)(v % nz);
        const long long t = v / nz;
        const int y = (int)(t % ny), x = (int)(t / ny);
        cnt += (x > 0 ? ids[v - (long long)ny * nz] : -1) == p2;
        cnt += (y > 0 ? ids[v - nz] : -1) == p2;
        cnt += (z > 0 ? ids[v - 1] : -1) == p2;
    }
    block_total(cnt, INT_MIN, red + 2);
}

__global__ void __launch_bounds__(NT) k_fill_i(int* a, long long n, int v)
{
    const long long st = (long long)gridDim.x * NT;
    for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += st) a[i] = v;
}

__global__ void __launch_bounds__(NT) k_copy_i(const int* a, long long n, int* b)
{
    const long long st = (long long)gridDim.x * NT;
    for (long long i = blockIdx.x * (long long)NT + threadIdx.x; i < n; i += st) b[i] = a[i];
}

template <typename P>
void dgrow(P*& p, size_t& cap, size_t count)
{
    if (count <= cap && p) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t bytes = sizeof(P) * (count ? count : 1);
    if (hipMalloc((void**)&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        throw ArgError{M3D_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed"};
    }
    cap = count;
}

// a buffer of fixed size for the handle's lifetime
template <typename P>
void dalloc(P*& p, size_t count)
{
    size_t cap = 0;
    dgrow(p, cap, count);
}

// stage boundaries timed on the handle's stream
enum Stage { ST_INIT, ST_LINK, ST_ENUM, ST_SIZES, ST_SORT, ST_RELABEL, ST_N };

} // namespace perc
} // namespace m3d

using namespace m3d;

struct m3d_perc {
    int nx = 0, ny = 0, nz = 0, nph = 0, phase = 0, plus = 0;
    long long n = 0;
    hipStream_t st = nullptr;
    int32_t *ids = nullptr, *lab = nullptr, *qm = nullptr, *c1 = nullptr;
    uint8_t* flag = nullptr;
    unsigned long long* red = nullptr;
    long long* tot = nullptr;
    int *ci = nullptr, *cr = nullptr;
    size_t cap_ci = 0, cap_cr = 0;
    int *clab = nullptr, *csize = nullptr, *rnk = nullptr, *hist = nullptr, *cum = nullptr, *part = nullptr;
    size_t cap_clab = 0, cap_csize = 0, cap_rnk = 0, cap_hist = 0, cap_cum = 0, cap_part = 0;
    double* frac = nullptr;
    size_t cap_frac = 0;
    int K = 0;
    double vol = 0;
    bool ran = false;   // clusters of the last analysis are there
    bool c1_ok = false; // the phase field (relabelled, or as given to the surface count) is there
    bool timing = false;
    hipEvent_t ev[perc::ST_N + 1] = {};
    double ms[perc::ST_N] = {};
    std::string err;
};

namespace {

template <typename F>
int guarded(m3d_perc* h, F&& f)
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

void release(m3d_perc* h)
{
    void* bufs[] = {h->ids, h->lab, h->qm, h->c1, h->flag, h->red, h->tot, h->ci, h->cr, h->clab,
                    h->csize, h->rnk, h->hist, h->cum, h->part, h->frac};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t& e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->st) (void)hipStreamDestroy(h->st);
}

unsigned blocks(long long n, int per) { return (unsigned)((n + per - 1) / per); }

void mark(m3d_perc* h, int k)
{
    if (h->timing) HCHK(hipEventRecord(h->ev[k], h->st));
}

// the pipeline up to the cluster count; returns K
void label(m3d_perc* h)
{
    using namespace perc;
    const long long n = h->n;
    const unsigned gb = blocks(n, NT), cb = blocks(n, BT);
    hipStream_t s = h->st;
    mark(h, ST_INIT);
    hipLaunchKernelGGL(k_setup, dim3(1), dim3(64), 0, s, h->red);
    hipLaunchKernelGGL(k_init, dim3(std::min(blocks((long long)h->nx * h->ny, NT / 64), 1024u)), dim3(NT), 0, s,
                       (const int32_t*)h->ids,
                       h->nx, h->ny, h->nz, h->phase, h->lab, h->flag, h->red);
    HCHK(hipGetLastError());
    mark(h, ST_LINK);
    hipLaunchKernelGGL(k_link, dim3(gb), dim3(NT), 0, s, (const uint8_t*)h->flag, h->lab, h->nx);
    HCHK(hipGetLastError());
    mark(h, ST_ENUM);
    hipLaunchKernelGGL(k_count, dim3(cb), dim3(NT), 0, s, (const uint8_t*)h->flag, (const int32_t*)h->lab, n, h->ci, h->cr);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(NTOP), 0, s, h->ci, h->cr, (int)cb, h->tot);
    HCHK(hipGetLastError());
    unsigned long long red[2];
    long long tot[2];
    HCHK(hipMemcpyAsync(red, h->red, sizeof red, hipMemcpyDeviceToHost, s));
    HCHK(hipMemcpyAsync(tot, h->tot, sizeof tot, hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    int mx;
    std::memcpy(&mx, &red[1], sizeof mx);
    if (mx >= h->plus)
        throw ArgError{M3D_ERR_ARG, "phase id " + std::to_string(mx) + " is not below noOfPhases + 1 = " +
                                        std::to_string(h->plus) + " (it would read as a cluster label)"};
    if ((long long)red[0] != tot[0]) throw ArgError{M3D_ERR_STATE, "interest voxel counts disagree"};
    h->vol = (double)red[0];
    h->K = (int)tot[1];
    dgrow(h->clab, h->cap_clab, (size_t)h->K);
    hipLaunchKernelGGL(k_enum, dim3(cb), dim3(NT), 0, s, (const uint8_t*)h->flag, (const int32_t*)h->lab, n,
                       (const int*)h->ci, (const int*)h->cr, h->plus, h->qm, h->clab);
    HCHK(hipGetLastError());
}

void stats(m3d_perc* h)
{
    using namespace perc;
    const long long n = h->n;
    const int K = h->K;
    hipStream_t s = h->st;
    dgrow(h->csize, h->cap_csize, (size_t)K);
    dgrow(h->rnk, h->cap_rnk, (size_t)K);
    dgrow(h->frac, h->cap_frac, (size_t)K);
    const long long L = (long long)h->vol + 1;
    const unsigned lb = blocks(L, BT);
    dgrow(h->hist, h->cap_hist, (size_t)L);
    dgrow(h->cum, h->cap_cum, (size_t)L);
    dgrow(h->part, h->cap_part, (size_t)lb);
    mark(h, ST_SIZES);
    if (K > 0) {
        hipLaunchKernelGGL(k_fill_i, dim3(std::min(blocks(K, NT), 1024u)), dim3(NT), 0, s, h->csize, (long long)K, 0);
        hipLaunchKernelGGL(k_sizes, dim3(blocks(n, BT)), dim3(NT), 0, s, (const uint8_t*)h->flag, h->lab,
                           (const int32_t*)h->qm, n, h->csize);
    }
    HCHK(hipGetLastError());
    mark(h, ST_SORT);
    if (K > 0) {
        hipLaunchKernelGGL(k_fill_i, dim3(std::min(blocks(L, NT), 1024u)), dim3(NT), 0, s, h->hist, L, 0);
        hipLaunchKernelGGL(k_hist, dim3(blocks(K, NT)), dim3(NT), 0, s, (const int*)h->csize, K, h->hist);
        hipLaunchKernelGGL(k_part, dim3(lb), dim3(NT), 0, s, (const int*)h->hist, L, h->part);
        hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(NTOP), 0, s, h->part, (int*)nullptr, (int)lb, h->tot + 2);
        hipLaunchKernelGGL(k_apply, dim3(lb), dim3(NT), 0, s, (const int*)h->hist, L, (const int*)h->part, h->cum);
        hipLaunchKernelGGL(k_rank, dim3(blocks(K, NT)), dim3(NT), 0, s, (const int*)h->csize, K, (const int*)h->cum,
                           h->hist, h->vol, h->rnk, h->frac);
    }
    HCHK(hipGetLastError());
    mark(h, ST_RELABEL);
    hipLaunchKernelGGL(k_relabel, dim3(blocks(n, NT)), dim3(NT), 0, s, (const int32_t*)h->ids, (const uint8_t*)h->flag,
                       (const int32_t*)h->lab, (const int32_t*)h->qm, (const int*)h->rnk, h->plus, n, h->c1);
    HCHK(hipGetLastError());
    mark(h, ST_N);
}

} // namespace

extern "C" {

const char* m3d_perc_last_error(m3d_perc* h) { return h ? h->err.c_str() : "null handle"; }

int m3d_perc_create(const int32_t* rve, int nx, int ny, int nz, int noOfPhases, int intrestPhase, m3d_perc** out)
{
    if (!out) return M3D_ERR_ARG;
    *out = nullptr;
    if (!rve || nx < 1 || ny < 1 || nz < 1 || noOfPhases < 0) return M3D_ERR_ARG;
    if ((long long)nx * ny * nz > (long long)INT_MAX) return M3D_ERR_UNSUPPORTED;
    m3d_perc* h = new (std::nothrow) m3d_perc;
    if (!h) return M3D_ERR_OOM;
    const int rc = guarded(h, [&]() {
        h->nx = nx;
        h->ny = ny;
        h->nz = nz;
        h->nph = noOfPhases;
        h->phase = intrestPhase;
        h->plus = noOfPhases + 1;
        h->n = (long long)nx * ny * nz;
        HCHK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
        for (hipEvent_t& e : h->ev) HCHK(hipEventCreate(&e));
        perc::dalloc(h->ids, (size_t)h->n);
        perc::dalloc(h->lab, (size_t)h->n);
        perc::dalloc(h->qm, (size_t)h->n);
        perc::dalloc(h->c1, (size_t)h->n);
        perc::dalloc(h->flag, (size_t)h->n);
        perc::dalloc(h->red, 4);
        perc::dalloc(h->tot, 4);
        const size_t cb = (size_t)blocks(h->n, perc::BT);
        perc::dgrow(h->ci, h->cap_ci, cb);
        perc::dgrow(h->cr, h->cap_cr, cb);
        HCHK(hipMemcpyAsync(h->ids, rve, sizeof(int32_t) * (size_t)h->n, hipMemcpyHostToDevice, h->st));
        HCHK(hipStreamSynchronize(h->st));
        return M3D_OK;
    });
    if (rc != M3D_OK) {
        release(h);
        delete h;
        return rc;
    }
    *out = h;
    return M3D_OK;
}

int m3d_perc_destroy(m3d_perc* h)
{
    if (!h) return M3D_ERR_ARG;
    (void)hipStreamSynchronize(h->st);
    release(h);
    delete h;
    return M3D_OK;
}

int m3d_perc_run(m3d_perc* h, int total_iter, int* nclusters, double* vol_intrstd)
{
    (void)total_iter;
    return guarded(h, [&]() {
        if (h->nx != h->ny || h->ny != h->nz)
            throw ArgError{M3D_ERR_UNSUPPORTED, "runPercolationAnalysis needs Nx = Ny = Nz (its trigger tests permute "
                                                "the cube's indices)"};
        if (h->phase >= h->plus)
            throw ArgError{M3D_ERR_ARG, "intrestPhase must be below noOfPhases + 1"};
        h->ran = false;
        h->c1_ok = false;
        label(h);
        stats(h);
        HCHK(hipStreamSynchronize(h->st));
        if (h->timing)
            for (int k = 0; k < perc::ST_N; ++k) {
                float t = 0;
                HCHK(hipEventElapsedTime(&t, h->ev[k], h->ev[k + 1]));
                h->ms[k] = t;
            }
        h->ran = true;
        h->c1_ok = true;
        if (nclusters) *nclusters = h->K;
        if (vol_intrstd) *vol_intrstd = h->vol;
        return M3D_OK;
    });
}

int m3d_perc_result(m3d_perc* h, int32_t* c1, int32_t* all_clusters, double* clusters_vol_frac, int maxn)
{
    return guarded(h, [&]() {
        if (c1 && !h->c1_ok) throw ArgError{M3D_ERR_STATE, "run the analysis or the surface count first"};
        if ((all_clusters || clusters_vol_frac) && !h->ran) throw ArgError{M3D_ERR_STATE, "run the analysis first"};
        const int m = h->ran ? std::max(0, std::min(maxn, h->K)) : 0;
        hipStream_t s = h->st;
        if (c1) HCHK(hipMemcpyAsync(c1, h->c1, sizeof(int32_t) * (size_t)h->n, hipMemcpyDeviceToHost, s));
        if (all_clusters && m)
            HCHK(hipMemcpyAsync(all_clusters, h->clab, sizeof(int32_t) * (size_t)m, hipMemcpyDeviceToHost, s));
        if (clusters_vol_frac && m)
            HCHK(hipMemcpyAsync(clusters_vol_frac, h->frac, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        return M3D_OK;
    });
}

int m3d_perc_surface(m3d_perc* h, const int32_t* rve, int phase1, int phase2, long long* common)
{
    return guarded(h, [&]() {
        if (!rve || !common) throw ArgError{M3D_ERR_ARG, "null argument"};
        hipStream_t s = h->st;
        HCHK(hipMemcpyAsync(h->ids, rve, sizeof(int32_t) * (size_t)h->n, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(perc::k_setup, dim3(1), dim3(64), 0, s, h->red);
        hipLaunchKernelGGL(perc::k_surface, dim3(2048), dim3(perc::NT), 0, s, (const int32_t*)h->ids, h->nx, h->ny, h->nz,
                           phase1, phase2, h->red);
        // the field now holds the new phases, as the reference's C1 does
        hipLaunchKernelGGL(perc::k_copy_i, dim3(2048), dim3(perc::NT), 0, s, (const int*)h->ids, h->n, h->c1);
        HCHK(hipGetLastError());
        unsigned long long red[3];
        HCHK(hipMemcpyAsync(red, h->red, sizeof red, hipMemcpyDeviceToHost, s));
        HCHK(hipStreamSynchronize(s));
        *common = (long long)red[2];
        h->c1_ok = true;
        return M3D_OK;
    });
}

int m3d_perc_set_timing(m3d_perc* h, int on)
{
    if (!h) return M3D_ERR_ARG;
    h->timing = on != 0;
    return M3D_OK;
}

int m3d_perc_timings(m3d_perc* h, double* ms, int n)
{
    return guarded(h, [&]() {
        if (!ms) throw ArgError{M3D_ERR_ARG, "null argument"};
        if (!h->timing || !h->ran) throw ArgError{M3D_ERR_STATE, "enable timing and run first"};
        for (int k = 0; k < n && k < perc::ST_N; ++k) ms[k] = h->ms[k];
        return M3D_OK;
    });
}

} // extern "C"
