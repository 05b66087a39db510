// m3d_gen.hip -- voxel work of the seeded microstructure generators of
// mat3d_amd/gen.py on the GPU: balls (random sequential addition, centres
// chosen on the host), straight finite cylinders (fibres, stamped in batches
// of 32 with the inclusion count after each batch) and periodic Voronoi
// cells.  The random draws and the placement decisions stay on the host, in
// the generators' own order; the kernels only rasterise, with the same
// integer tests and the same floating-point operations in the same order
// (no contraction), so the phase ids are the generators' bit for bit.
#include "m3d_engine.hpp"

#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

namespace m3d {
namespace gen {

constexpr int NT = 256;

__global__ void k_fill(int32_t* ids, long long n, int32_t v)
{
    const long long s0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, st = (long long)gridDim.x * blockDim.x;
    for (long long i = s0; i < n; i += st) ids[i] = v;
}

// one workgroup per ball {cx, cy, cz, r, phase}: voxels c + o with
// |o|^2 <= r^2, periodic (gen._stamp_ball)
__global__ void __launch_bounds__(NT) k_balls(int32_t* ids, int nx, int ny, int nz, const int32_t* balls)
{
    const int32_t* b = balls + 5 * (long long)blockIdx.x;
    const int cx = b[0], cy = b[1], cz = b[2], r = b[3], ph = b[4];
    const int w = 2 * r + 1;
    const long long cube = (long long)w * w * w;
    for (long long e = threadIdx.x; e < cube; e += NT) {
        const int ox = (int)(e / ((long long)w * w)) - r;
        const int oy = (int)((e / w) % w) - r;
        const int oz = (int)(e % w) - r;
        if (ox * ox + oy * oy + oz * oz > r * r) continue;
        const int ix = ((cx + ox) % nx + nx) % nx, iy = ((cy + oy) % ny + ny) % ny, iz = ((cz + oz) % nz + nz) % nz;
        ids[((long long)ix * ny + iy) * nz + iz] = ph;
    }
}

// one workgroup per cylinder {v0, v1, v2, c0, c1, c2, lo0, lo1, lo2, hi0,
// hi1, hi2}: voxels (a0, a1, a2) of the box lo..hi with X = a - c,
// t = X.v, d2 = |X|^2 - t^2, |t| <= half and d2 <= r^2 (gen.fibres), in
// numpy's operation order without fused multiply-adds
__global__ void __launch_bounds__(NT) k_cylinders(int32_t* ids, int nx, int ny, int nz, const double* cyl, double half,
                                                  double r2, int32_t incl)
{
#pragma clang fp contract(off)
    const double* f = cyl + 12 * (long long)blockIdx.x;
    const double v0 = f[0], v1 = f[1], v2 = f[2], c0 = f[3], c1 = f[4], c2 = f[5];
    const long long lo0 = (long long)f[6], lo1 = (long long)f[7], lo2 = (long long)f[8];
    const long long n0 = (long long)f[9] - lo0 + 1, n1 = (long long)f[10] - lo1 + 1, n2 = (long long)f[11] - lo2 + 1;
    const long long tot = n0 * n1 * n2;
    for (long long e = threadIdx.x; e < tot; e += NT) {
        const long long a0 = lo0 + e / (n1 * n2), a1 = lo1 + (e / n2) % n1, a2 = lo2 + e % n2;
        const double X = (double)a0 - c0, Y = (double)a1 - c1, Z = (double)a2 - c2;
        const double t = (X * v0 + Y * v1) + Z * v2;
        const double d2 = ((X * X + Y * Y) + Z * Z) - t * t;
        if (fabs(t) <= half && d2 <= r2) {
            const long long ix = ((a0 % nx) + nx) % nx, iy = ((a1 % ny) + ny) % ny, iz = ((a2 % nz) + nz) % nz;
            ids[(ix * ny + iy) * nz + iz] = incl;
        }
    }
}

__global__ void __launch_bounds__(NT) k_count(const int32_t* ids, long long n, int32_t v, unsigned long long* out)
{
    __shared__ unsigned long long red[NT / 64];
    unsigned long long a = 0;
    const long long s0 = blockIdx.x * (long long)blockDim.x + threadIdx.x, st = (long long)gridDim.x * blockDim.x;
    for (long long i = s0; i < n; i += st) a += ids[i] == v ? 1ull : 0ull;
    for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s = 0;
        for (int w = 0; w < NT / 64; ++w) s += red[w];
        out[blockIdx.x] = s;
    }
}

// nearest seed under the minimum-image distance (per axis min(|d|, L - |d|)),
// the first seed on exact ties; seeds staged in LDS in chunks
__global__ void __launch_bounds__(NT) k_voronoi(int32_t* ids, int nx, int ny, int nz, const double* pts, int g)
{
#pragma clang fp contract(off)
    __shared__ double sp[3 * 512];
    const long long n = (long long)nx * ny * nz;
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const double bx = nx, by = ny, bz = nz;
    double qx = 0, qy = 0, qz = 0;
    if (i < n) {
        qx = (double)(i / ((long long)ny * nz));
        qy = (double)((i / nz) % ny);
        qz = (double)(i % nz);
    }
    double best = 1e300;
    int arg = 0;
    for (int k0 = 0; k0 < g; k0 += 512) {
        const int m = min(512, g - k0);
        __syncthreads();
        for (int e = threadIdx.x; e < 3 * m; e += NT) sp[e] = pts[3 * (long long)k0 + e];
        __syncthreads();
        for (int k = 0; k < m; ++k) {
            double dx = fabs(qx - sp[3 * k]), dy = fabs(qy - sp[3 * k + 1]), dz = fabs(qz - sp[3 * k + 2]);
            dx = fmin(dx, bx - dx);
            dy = fmin(dy, by - dy);
            dz = fmin(dz, bz - dz);
            const double d2 = (dx * dx + dy * dy) + dz * dz;
            if (d2 < best) {
                best = d2;
                arg = k0 + k;
            }
        }
    }
    if (i < n) ids[i] = arg;
}

struct Buf {
    int32_t* ids = nullptr;
    void* aux = nullptr;
    unsigned long long* cnt = nullptr;
    ~Buf()
    {
        if (ids) (void)hipFree(ids);
        if (aux) (void)hipFree(aux);
        if (cnt) (void)hipFree(cnt);
    }
};

template <typename P>
void dalloc(P** p, size_t bytes)
{
    hipError_t e = hipMalloc((void**)p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        throw ArgError{M3D_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed"};
    }
}

int run(const char* what, std::string& err, const std::function<void()>& f)
{
    try {
        f();
        return M3D_OK;
    } catch (const HipError& e) {
        err = std::string(what) + ": HIP error in " + e.what + ": " + hipGetErrorString(e.e);
        (void)hipGetLastError();
        return M3D_ERR_HIP;
    } catch (const ArgError& e) {
        err = std::string(what) + ": " + e.msg;
        return e.code;
    } catch (const std::exception& e) {
        err = std::string(what) + ": " + e.what();
        return M3D_ERR_UNSUPPORTED;
    }
}

thread_local std::string last_err;

// Fixed device buffer of m3d_gen_cylinders_noalloc: the staged cylinder list
// (12 doubles each) followed by the per-workgroup inclusion counts.  Every
// GPU the library runs on holds its own copy until the process ends.  Host
// code never passes `scratch` itself to a kernel or a copy (on the host the
// name is only a stand-in): it uses the device address that
// hipGetSymbolAddress returns for the current device.
constexpr int NOALLOC_MAX_CYL = 8192;
constexpr int NOALLOC_COUNT_BLOCKS = 1024;
static __device__ __align__(256) double scratch[12 * NOALLOC_MAX_CYL + NOALLOC_COUNT_BLOCKS];

double* scratch_ptr()
{
    static std::mutex mtx;
    static std::vector<double*> per_device;
    int dev = 0;
    HCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mtx);
    if ((int)per_device.size() <= dev) per_device.resize(dev + 1, nullptr);
    if (!per_device[dev]) {
        void* p = nullptr;
        HCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        per_device[dev] = static_cast<double*>(p);
    }
    return per_device[dev];
}

} // namespace gen
} // namespace m3d

using namespace m3d;

const char* m3d_gen_last_error(void) { return gen::last_err.c_str(); }

int m3d_gen_balls(int nx, int ny, int nz, const int32_t* balls, int nballs, int32_t fill, int32_t* ids)
{
    if (nx < 1 || ny < 1 || nz < 1 || nballs < 0 || !ids || (nballs > 0 && !balls)) return M3D_ERR_ARG;
    for (int k = 0; k < nballs; ++k)
        if (balls[5 * (size_t)k + 3] < 0) return M3D_ERR_ARG;
    return gen::run("gen_balls", gen::last_err, [&]() {
        const long long n = (long long)nx * ny * nz;
        gen::Buf b;
        gen::dalloc(&b.ids, sizeof(int32_t) * n);
        gen::dalloc(&b.aux, sizeof(int32_t) * 5 * (size_t)nballs);
        if (nballs) HCHK(hipMemcpy(b.aux, balls, sizeof(int32_t) * 5 * (size_t)nballs, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(gen::k_fill, dim3(1024), dim3(gen::NT), 0, nullptr, b.ids, n, fill);
        if (nballs)
            hipLaunchKernelGGL(gen::k_balls, dim3(nballs), dim3(gen::NT), 0, nullptr, b.ids, nx, ny, nz, (const int32_t*)b.aux);
        HCHK(hipGetLastError());
        HCHK(hipMemcpy(ids, b.ids, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    });
}

int m3d_gen_cylinders(int nx, int ny, int nz, const double* cyl, int ncyl, double half, double radius, int32_t incl,
                      int batch, long long target, int32_t* ids, int* used, long long* count)
{
    if (nx < 1 || ny < 1 || nz < 1 || ncyl < 0 || batch < 1 || !ids || (ncyl > 0 && !cyl)) return M3D_ERR_ARG;
    for (int k = 0; k < ncyl; ++k)
        for (int a = 0; a < 3; ++a)
            if (cyl[12 * (size_t)k + 9 + a] < cyl[12 * (size_t)k + 6 + a]) return M3D_ERR_ARG;
    return gen::run("gen_cylinders", gen::last_err, [&]() {
        const long long n = (long long)nx * ny * nz;
        const int nb = 1024;
        gen::Buf b;
        gen::dalloc(&b.ids, sizeof(int32_t) * n);
        gen::dalloc(&b.aux, sizeof(double) * 12 * (size_t)ncyl);
        gen::dalloc(&b.cnt, sizeof(unsigned long long) * nb);
        HCHK(hipMemcpy(b.ids, ids, sizeof(int32_t) * n, hipMemcpyHostToDevice));
        if (ncyl) HCHK(hipMemcpy(b.aux, cyl, sizeof(double) * 12 * (size_t)ncyl, hipMemcpyHostToDevice));
        std::vector<unsigned long long> part(nb);
        // stamp batch by batch until the inclusion count reaches the target
        long long c = 0;
        int k = 0;
        auto counted = [&]() {
            hipLaunchKernelGGL(gen::k_count, dim3(nb), dim3(gen::NT), 0, nullptr, b.ids, n, incl, b.cnt);
            HCHK(hipGetLastError());
            HCHK(hipMemcpy(part.data(), b.cnt, sizeof(unsigned long long) * nb, hipMemcpyDeviceToHost));
            long long s = 0;
            for (unsigned long long v : part) s += (long long)v;
            return s;
        };
        c = counted();
        while (c < target && k + batch <= ncyl) {
            hipLaunchKernelGGL(gen::k_cylinders, dim3(batch), dim3(gen::NT), 0, nullptr, b.ids, nx, ny, nz,
                               (const double*)b.aux + 12 * (size_t)k, half, radius * radius, incl);
            k += batch;
            c = counted();
        }
        HCHK(hipMemcpy(ids, b.ids, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        if (used) *used = k;
        if (count) *count = c;
    });
}

int m3d_gen_voronoi(int nx, int ny, int nz, const double* pts, int g, int32_t* ids)
{
    if (nx < 1 || ny < 1 || nz < 1 || g < 1 || !pts || !ids) return M3D_ERR_ARG;
    return gen::run("gen_voronoi", gen::last_err, [&]() {
        const long long n = (long long)nx * ny * nz;
        gen::Buf b;
        gen::dalloc(&b.ids, sizeof(int32_t) * n);
        gen::dalloc(&b.aux, sizeof(double) * 3 * (size_t)g);
        HCHK(hipMemcpy(b.aux, pts, sizeof(double) * 3 * (size_t)g, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(gen::k_voronoi, dim3((unsigned)((n + gen::NT - 1) / gen::NT)), dim3(gen::NT), 0, nullptr, b.ids,
                           nx, ny, nz, (const double*)b.aux, g);
        HCHK(hipGetLastError());
        HCHK(hipMemcpy(ids, b.ids, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    });
}

int m3d_gen_cylinders_noalloc(int nx, int ny, int nz, const double* cyl, int ncyl, double half, double radius,
                              int32_t incl, int batch, long long target, int32_t* ids_dev, int* used, long long* count)
{
    if (nx < 1 || ny < 1 || nz < 1 || ncyl < 0 || batch < 1 || !ids_dev || (ncyl > 0 && !cyl)) return M3D_ERR_ARG;
    if (ncyl > gen::NOALLOC_MAX_CYL) {
        gen::last_err = "gen_cylinders_noalloc: " + std::to_string(ncyl) + " cylinders exceed the fixed device buffer (" +
                        std::to_string(gen::NOALLOC_MAX_CYL) + " at most per call)";
        return M3D_ERR_ARG;
    }
    for (int k = 0; k < ncyl; ++k)
        for (int a = 0; a < 3; ++a)
            if (cyl[12 * (size_t)k + 9 + a] < cyl[12 * (size_t)k + 6 + a]) return M3D_ERR_ARG;
    return gen::run("gen_cylinders_noalloc", gen::last_err, [&]() {
        const long long n = (long long)nx * ny * nz;
        const int nb = gen::NOALLOC_COUNT_BLOCKS;
        double* sp = gen::scratch_ptr();
        double* cyl_d = sp;
        unsigned long long* cnt = reinterpret_cast<unsigned long long*>(sp + 12 * (size_t)gen::NOALLOC_MAX_CYL);
        if (ncyl) HCHK(hipMemcpy(cyl_d, cyl, sizeof(double) * 12 * (size_t)ncyl, hipMemcpyHostToDevice));
        unsigned long long part[gen::NOALLOC_COUNT_BLOCKS];
        // every count overwrites all nb entries, so nothing of an earlier
        // call's contents is read
        auto counted = [&]() {
            hipLaunchKernelGGL(gen::k_count, dim3(nb), dim3(gen::NT), 0, nullptr, ids_dev, n, incl, cnt);
            HCHK(hipGetLastError());
            HCHK(hipMemcpy(part, cnt, sizeof(part), hipMemcpyDeviceToHost));
            long long s = 0;
            for (unsigned long long v : part) s += (long long)v;
            return s;
        };
        long long c = counted();
        int k = 0;
        while (c < target && k + batch <= ncyl) {
            hipLaunchKernelGGL(gen::k_cylinders, dim3(batch), dim3(gen::NT), 0, nullptr, ids_dev, nx, ny, nz,
                               (const double*)cyl_d + 12 * (size_t)k, half, radius * radius, incl);
            k += batch;
            c = counted();
        }
        HCHK(hipGetLastError());
        if (used) *used = k;
        if (count) *count = c;
    });
}
