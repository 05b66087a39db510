// m3d_basic_fast.hpp -- interface of the compile-time specialised kernels of
// the basic scheme ("ep" / "ed", FFT_analysis3D.cpp:796-853), defined in
// m3d_basic.hip.  Same passes as the runtime-planned kernels of
// m3d_basic.hpp (z pass with the constitutive update, y pass on 12
// components, x pass with eqVar partials and the Green step, y pass on 6),
// with every FFT length and tile a template constant.
#pragma once
#include "m3d_basic.hpp"
#include "m3d_fast.hpp"

namespace m3d {
namespace bfast {

#ifndef M3D_BZ_THREADS
#define M3D_BZ_THREADS 256
#endif
#ifndef M3D_BZ_TILE
#define M3D_BZ_TILE 2304
#endif
#ifndef M3D_BZ_PAD
#define M3D_BZ_PAD 0 // complex values of padding per LDS row of the z pass
#endif
#ifndef M3D_BX_THREADS
#define M3D_BX_THREADS 512
#endif
#ifndef M3D_BX_TILE
#define M3D_BX_TILE 6144
#endif
// minimum waves per SIMD the register allocator must leave room for in the
// z and x passes (0: no constraint)
#ifndef M3D_BZ_WAVES
#define M3D_BZ_WAVES 0
#endif
#ifndef M3D_BX_WAVES
#define M3D_BX_WAVES 0
#endif
#if M3D_BZ_WAVES > 0
#define M3D_BZ_ATTR __attribute__((amdgpu_waves_per_eu(M3D_BZ_WAVES)))
#else
#define M3D_BZ_ATTR
#endif
#if M3D_BX_WAVES > 0
#define M3D_BX_ATTR __attribute__((amdgpu_waves_per_eu(M3D_BX_WAVES)))
#else
#define M3D_BX_ATTR
#endif
constexpr int ZNT = M3D_BZ_THREADS; // threads of the z pass
constexpr int XNT = M3D_BX_THREADS; // threads of the x pass

// z tile: TY lines (y) of 12 components each, TY*12*LZ <= M3D_BZ_TILE complex
constexpr int zlines(int LZ, int Ny)
{
    int ty = 1;
    while (ty < 8 && (ty * 2) * 12 * LZ <= M3D_BZ_TILE && Ny % (ty * 2) == 0) ty *= 2;
    return ty;
}
// x tile: TK kz values of 12 components (TK * L <= M3D_BX_TILE / 12)
constexpr int xtile(int L)
{
    const int tk = M3D_BX_TILE / (12 * L);
    return tk < 1 ? 1 : tk;
}
// y tiles (the generic fast y pass with 12 / 6 components)
constexpr int ytile12(int L) { return fast::xytile<double>(L, 12); }
constexpr int ytile6(int L) { return fast::xytile<double>(L, 6); }

// line lengths with specialised basic kernels: z half-lengths and x/y lengths
inline bool z_supported(int LZ) { return LZ == 16 || LZ == 24 || LZ == 32 || LZ == 48 || LZ == 64 || LZ == 96 || LZ == 128; }
inline bool xy_supported(int L) { return L == 32 || L == 48 || L == 64 || L == 96 || L == 128 || L == 192 || L == 256; }
inline bool supported(int Nx, int Ny, int Nz)
{
    return Nz % 2 == 0 && z_supported(Nz / 2) && xy_supported(Nx) && xy_supported(Ny) && Ny % zlines(Nz / 2, Ny) == 0 &&
           zlines(Nz / 2, Ny) == zlines(Nz / 2, Nz);
}

// host launchers (m3d_basic.hip); g carries Nx..Nzh, N, H, invN, nbz, nbx
// of the tiles above, TY, TKx; the y launch takes the 12- or 6-component
// view of the spectra in p (S: 6-component Green output, K: 12-component)
// (T = double or float)
// lin: every phase has hardening exponent n = 1 (FP64 "ep" then runs the
// return map without pow, bit for bit the same arithmetic)
template <typename T>
void launch_z(int beh, int mode, const Geo& g, const BPtr<T>& p, hipStream_t st, bool lin = false);
template <typename T>
void launch_y12(const Geo& g, const Ptr<T>& p, hipStream_t st);
// forward y pass of the 6-component sigma^ spectrum that an ITER z pass
// leaves in the K view
template <typename T>
void launch_y6_fwd(const Geo& g, const Ptr<T>& p, hipStream_t st);
// e6: the x pass after an ITER z pass (sigma^ with 6 components, eps^ from Eh)
template <typename T>
void launch_x(const Geo& g, const BPtr<T>& p, hipStream_t st, bool e6);
template <typename T>
void launch_y6(const Geo& g, const Ptr<T>& p, hipStream_t st);

} // namespace bfast
} // namespace m3d
