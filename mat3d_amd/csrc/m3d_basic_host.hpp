// m3d_basic_host.hpp -- host driver of the basic scheme ("ep" / "ed").
#pragma once
#include "m3d_basic.hpp"
#include "m3d_basic_fast.hpp"
#include "m3d_engine.hpp"

namespace {

// ------------------------------------------------------ basic scheme runner
// runAnalysis "ep" / "ed" (FFT_analysis3D.cpp:796-853) on one solve.  FP64
// handles, and grids without specialised kernels, iterate in FP64.  On an
// FP32+polish handle each increment iterates in FP32 (fields, FFTs, return
// map) until the stop rule holds at polish_factor x tol, then continues in
// FP64 from the promoted Green output under the reference's rule; the
// history (sigma_old, eps_old, ep) is kept and updated in FP64.
struct BasicRunner {
    Engine<Elastic>& e;
    Geo g12{}, g6{};
    Geo gf{};          // tiles of the compile-time specialised kernels
    Geo gf6{};         // the same with the row strides of the 6-component spectrum
    bool fast = false; // grid has specialised basic kernels (m3d_basic.hip)
    int beh;
    bool two = false;       // FP32 phase + FP64 polish per increment
    double* f6 = nullptr;   // eps, sig, sig_old, eps_old, sig_star (5 x 6N)
    double* f1 = nullptr;   // ep|d, phi, dep, seff (4 x N)
    unsigned char* S6 = nullptr;  // two 6-component spectra: the specialised passes alternate them
    unsigned char* S12 = nullptr;
    unsigned char* Eh = nullptr;  // eps^ of the specialised x pass
    float* f6s = nullptr;   // FP32 phase: the same fields in FP32
    float* f1s = nullptr;
    unsigned char* S6s = nullptr;
    unsigned char* S12s = nullptr;
    unsigned char* Ehs = nullptr;
    size_t neh = 0;         // complex values of Eh
    double* psum = nullptr;
    double* psum2 = nullptr; // [NBSUM][BSPLIT] shares of psum
    double* pfin = nullptr;  // [14][nbz] block sums of the fused FINAL pass
    bool lin = false;        // every phase has hardening exponent n = 1
    double* pchk = nullptr;
    double* bs = nullptr;
    double* mat = nullptr;
    Ctl* ctl = nullptr;
    int nblk = 256;
    std::vector<double> prop; // rows x 9, column-major like prop_homogenized
    int rows = 0;
    int exit_inc = -1; // increment of the "ed" softening exit, -1 if none
    m3d_run_info info{};

    BasicRunner(Engine<Elastic>& eng, int b, bool fp32) : e(eng), beh(b)
    {
        g12 = make_geo<double>(e.Nx, e.Ny, e.Nz, 12);
        g6 = make_geo<double>(e.Nx, e.Ny, e.Nz, 6);
        g12.fz = g12.fy = g12.fx = 0;
        g6.fz = g6.fy = g6.fx = 0;
        // make_geo may have switched tiles to the specialised kernels' shapes
        plan_tiles(g12, sizeof(cx<double>));
        plan_tiles(g6, sizeof(cx<double>));
        g12.nsolve = g6.nsolve = 1;
        // the basic scheme's spectra keep plain [x][y][kz] rows without padding
        for (Geo* q : {&g12, &g6}) {
            q->lxb = q->xbm = 0;
            q->sy = q->sxb = (long long)q->Nzh * q->C;
            q->sxa = (long long)q->Ny * q->sy;
        }
        fast = bfast::supported(e.Nx, e.Ny, e.Nz);
        if (fast) {
            gf = g12;
            gf.TY = bfast::zlines(e.Nz / 2, e.Ny);
            gf.nty = e.Ny / gf.TY;
            gf.nbz = e.Nx * gf.nty;
            gf.TKx = bfast::xtile(e.Nx);
            gf.nkx = (gf.Nzh + gf.TKx - 1) / gf.TKx;
            gf.nbx = e.Ny * gf.nkx;
            gf6 = gf;
            gf6.sy = gf6.sxb = (long long)gf.Nzh * 6;
            gf6.sxa = (long long)gf.Ny * gf6.sy;
        }
        const int nbz = fast ? std::max(g12.nbz, gf.nbz) : g12.nbz;
        const int nbx = fast ? std::max(g12.nbx, gf.nbx) : g12.nbx;
        const size_t N = (size_t)e.N, H = (size_t)e.H;
        try {
            e.alloc(&f6, sizeof(double) * 30 * N);
            e.alloc(&f1, sizeof(double) * 4 * N);
            e.alloc(&S6, sizeof(cx<double>) * 12 * H);
            e.alloc(&S12, sizeof(cx<double>) * 12 * H);
            if (fast) {
                neh = (size_t)gf.nbx * e.Nx * gf.TKx * 6;
                e.alloc(&Eh, sizeof(cx<double>) * neh);
            }
            e.alloc(&psum, sizeof(double) * NBSUM * nbz);
            e.alloc(&psum2, sizeof(double) * NBSUM * BSPLIT);
            if (fast) e.alloc(&pfin, sizeof(double) * 14 * (size_t)gf.nbz);
            e.alloc(&pchk, sizeof(double) * (nbx + 1));
            e.alloc(&bs, sizeof(double) * 14 * nblk);
            e.alloc(&mat, sizeof(double) * 256 * 8);
            e.alloc(&ctl, sizeof(Ctl));
            two = fp32 && fast;
            if (two) {
                e.alloc(&f6s, sizeof(float) * 24 * N);
                e.alloc(&f1s, sizeof(float) * 4 * N);
                e.alloc(&S6s, sizeof(cx<float>) * 12 * H);
                e.alloc(&S12s, sizeof(cx<float>) * 12 * H);
                e.alloc(&Ehs, sizeof(cx<float>) * neh);
            }
        } catch (...) {
            release();
            throw;
        }
    }
    void release()
    {
        for (void* q : {(void*)f6, (void*)f1, (void*)S6, (void*)S12, (void*)Eh, (void*)psum, (void*)psum2, (void*)pfin, (void*)pchk,
                        (void*)bs, (void*)mat, (void*)ctl, (void*)f6s, (void*)f1s, (void*)S6s, (void*)S12s, (void*)Ehs})
            if (q) (void)hipFree(q);
        f6 = f1 = psum = psum2 = pfin = pchk = bs = mat = nullptr;
        S6 = S12 = Eh = nullptr;
        f6s = f1s = nullptr;
        S6s = S12s = Ehs = nullptr;
        ctl = nullptr;
    }
    ~BasicRunner() { release(); }

    // the S6 buffer `half` (0 or 1) of precision T
    template <typename T>
    cx<T>* s6_of(int half)
    {
        cx<T>* b = reinterpret_cast<cx<T>*>(sizeof(T) == 8 ? S6 : S6s);
        return b + (half ? 6 * e.H : 0);
    }
    // par < 0: the generic kernels (one S6 buffer, read and written in place);
    // par = 0 / 1: the specialised z pass reads S6 buffer par and the x pass
    // writes buffer 1 - par
    template <typename T>
    BPtr<T> bptr_t(int par = -1)
    {
        BPtr<T> p;
        std::memset(&p, 0, sizeof p);
        const long long N = e.N;
        T* a6 = sizeof(T) == 8 ? (T*)f6 : (T*)f6s;
        T* a1 = sizeof(T) == 8 ? (T*)f1 : (T*)f1s;
        p.eps = a6;
        p.sig = a6 + 6 * N;
        p.sig_old = a6 + 12 * N;
        p.eps_old = a6 + 18 * N;
        p.sig_star = sizeof(T) == 8 ? a6 + 24 * N : nullptr;
        p.pfin = sizeof(T) == 8 ? pfin : nullptr;
        p.ep = a1;
        p.phi = a1 + N;
        p.dep = a1 + 2 * N;
        p.seff = a1 + 3 * N;
        p.sU = 6 * N;
        p.sN = N;
        p.S6 = reinterpret_cast<cx<T>*>(sizeof(T) == 8 ? S6 : S6s);
        p.S12 = reinterpret_cast<cx<T>*>(sizeof(T) == 8 ? S12 : S12s);
        p.sS6 = p.sS12 = 12 * e.H;
        if (par >= 0) {
            p.S6 = s6_of<T>(par);
            p.S6w = s6_of<T>(1 - par);
            p.Eh = reinterpret_cast<cx<T>*>(sizeof(T) == 8 ? Eh : Ehs);
        }
        p.ph = e.ph;
        p.mat = mat;
        const cx<T>* const* tw = e.template twiddles_of<T>();
        p.twz = tw[0];
        p.twy = tw[1];
        p.twx = tw[2];
        p.psum = psum;
        p.pchk = pchk;
        p.ctl = ctl;
        return p;
    }
    template <typename T>
    Ptr<T> yptr_t(int par = -1)
    {
        Ptr<T> p;
        std::memset(&p, 0, sizeof p);
        p.S = par >= 0 ? s6_of<T>(1 - par) : reinterpret_cast<cx<T>*>(sizeof(T) == 8 ? S6 : S6s);
        p.K = reinterpret_cast<cx<T>*>(sizeof(T) == 8 ? S12 : S12s);
        p.sS = 12 * e.H;
        const cx<T>* const* tw = e.template twiddles_of<T>();
        p.twz = tw[0];
        p.twy = tw[1];
        p.twx = tw[2];
        p.ctl = ctl;
        return p;
    }
    BPtr<double> bptr() { return bptr_t<double>(); }
    Ptr<double> yptr() { return yptr_t<double>(); }

    // the per-solve sums of the z pass (two launches, fixed order) and the
    // increment/iteration bookkeeping that follows them
    template <int MODE>
    void bfin1(const Geo& g)
    {
        hipLaunchKernelGGL(k_bsum_parts<0>, dim3(BSPLIT, NBSUM, 1), dim3(256), 0, e.st, g, ctl, psum, psum2);
        hipLaunchKernelGGL((k_bfin1<MODE>), dim3(1), dim3(FTHREADS), 0, e.st, g, ctl, psum2);
    }

    template <bool GEN>
    void stage(int mode)
    {
        hipStream_t st = e.st;
        BPtr<double> bp = bptr();
        Ptr<double> yp = yptr();
        const size_t zsm = (size_t)g12.TY * 12 * g12.lsz * sizeof(cx<double>);
        const size_t ysm12 = (size_t)g12.Ny * g12.TKy * 12 * sizeof(cx<double>);
        const size_t ysm6 = (size_t)g6.Ny * g6.TKy * 6 * sizeof(cx<double>);
        const size_t xsm = (size_t)g12.Nx * g12.TKx * 12 * sizeof(cx<double>);
        if (mode == BMODE_START) {
            if (beh == BEH_EP)
                hipLaunchKernelGGL((k_zbasic<double, BEH_EP, BMODE_START, GEN>), dim3(g12.nbz), dim3(ZTHREADS), zsm, st, g12, bp);
            else
                hipLaunchKernelGGL((k_zbasic<double, BEH_ED, BMODE_START, GEN>), dim3(g12.nbz), dim3(ZTHREADS), zsm, st, g12, bp);
            bfin1<BMODE_START>(g12);
        } else {
            if (beh == BEH_EP)
                hipLaunchKernelGGL((k_zbasic<double, BEH_EP, BMODE_ITER, GEN>), dim3(g12.nbz), dim3(ZTHREADS), zsm, st, g12, bp);
            else
                hipLaunchKernelGGL((k_zbasic<double, BEH_ED, BMODE_ITER, GEN>), dim3(g12.nbz), dim3(ZTHREADS), zsm, st, g12, bp);
            bfin1<BMODE_ITER>(g12);
        }
        hipLaunchKernelGGL((k_ypass<double, 12, -1, BUF_K, GEN>), dim3(g12.Nx * g12.nky), dim3(XYTHREADS), ysm12, st, g12, yp);
        hipLaunchKernelGGL((k_xbasic<double, GEN>), dim3(g12.nbx), dim3(XYTHREADS), xsm, st, g12, bp);
        hipLaunchKernelGGL((k_bfin2<0>), dim3(1), dim3(FTHREADS), 0, st, g12, ctl, pchk);
        hipLaunchKernelGGL((k_ypass<double, 6, 1, BUF_MAIN, GEN>), dim3(g6.Nx * g6.nky), dim3(XYTHREADS), ysm6, st, g6, yp);
    }

    // one stage with the specialised kernels (the sequence of stage()); the
    // restart stage is a start stage (strain from HBM) that keeps itNo.  A
    // start stage transforms [sigma | eps] and writes S6 buffer 0 (par = 1);
    // an iteration reads S6 buffer par, transforms sigma only (the x pass
    // keeps eps^ in Eh) and writes buffer 1 - par.
    template <typename T = double>
    void stage_fast(int mode, int par)
    {
        hipStream_t st = e.st;
        const bool iter = mode == BMODE_ITER;
        BPtr<T> bp = bptr_t<T>(par);
        Ptr<T> yp = yptr_t<T>(par);
        bfast::launch_z<T>(beh, iter ? BMODE_ITER : BMODE_START, gf, bp, st, lin);
        if (mode == BMODE_START) bfin1<BMODE_START>(gf);
        else if (mode == BMODE_RESTART)
            bfin1<BMODE_RESTART>(gf);
        else bfin1<BMODE_ITER>(gf);
        if (iter) bfast::launch_y6_fwd<T>(gf6, yp, st);
        else bfast::launch_y12<T>(gf, yp, st);
        bfast::launch_x<T>(gf, bp, st, iter);
        hipLaunchKernelGGL((k_bfin2<0>), dim3(1), dim3(FTHREADS), 0, st, gf, ctl, pchk);
        bfast::launch_y6<T>(gf6, yp, st);
    }

    // the fields and state of the increment's last iterate on the
    // specialised FP64 path, whose iterations do not store them: the z pass
    // in FINAL mode on the S6 buffer that iterate was computed from (for
    // "ep" it also leaves the export sums in pfin and applies
    // update_old_Variable).  The
    // iterations after the start stage (itNo it0) alternate the buffers from
    // buffer 0, so iterate itNo c.it read buffer (c.it - it0 - 1) & 1.  The
    // z pass skips a solve marked done, so the flag is cleared around it; c
    // is the host copy of the control block after the iterations.
    void final_fast(Ctl& c, int it0)
    {
        hipStream_t st = e.st;
        Ctl run = c;
        run.done = 0;
        HCHK(hipMemcpyAsync(ctl, &run, sizeof run, hipMemcpyHostToDevice, st));
        bfast::launch_z<double>(beh, BMODE_FINAL, gf, bptr_t<double>((c.it - it0 - 1) & 1), st, lin);
        HCHK(hipGetLastError());
        Ctl after;
        HCHK(hipMemcpyAsync(&after, ctl, sizeof after, hipMemcpyDeviceToHost, st));
        HCHK(hipStreamSynchronize(st));
        if (after.err) c.done = 3;
        HCHK(hipMemcpyAsync(ctl, &c, sizeof c, hipMemcpyHostToDevice, st));
        HCHK(hipStreamSynchronize(st));
    }

    void run_stage(int mode, int par)
    {
        if (fast) stage_fast(mode, par);
        else if (g12.gen) stage<true>(mode);
        else stage<false>(mode);
    }

    // mtab: per phase {la, mu, Sy, H, n, P, Y0, 0}; la0/mu0: reference medium;
    // nc = N (C[0]_00/2 + C[1]_00/2)
    void run(const std::vector<double>& mtab, double la0, double mu0, double nc, const double* E, const double* X0,
             double tol_f, double tol_eq, int nInc, int cap)
    {
        hipStream_t st = e.st;
        HCHK(hipMemcpyAsync(mat, mtab.data(), sizeof(double) * mtab.size(), hipMemcpyHostToDevice, st));
        lin = true;
        for (size_t k = 4; k < mtab.size(); k += 8) lin = lin && mtab[k] == 1.0;
        Ctl c;
        std::memset(&c, 0, sizeof c);
        double dEs = 0;
        for (int i = 0; i < 6; ++i) {
            c.E[i] = E[i] / (double)nInc; // dE = E/nInc; E = dE (:799)
            c.X0[i] = X0[i];
            dEs += c.E[i];
        }
        c.dEsum = dEs;
        c.la0 = la0;
        c.mu0 = mu0;
        c.nc = nc;
        c.cap = cap;
        c.tol_f = tol_f;
        c.tol_eq = tol_eq;
        c.eq_it = -1;
        HCHK(hipMemcpyAsync(ctl, &c, sizeof c, hipMemcpyHostToDevice, st));
        BPtr<double> bp = bptr();
        hipLaunchKernelGGL((k_binit<double>), dim3(1024), dim3(256), 0, st, g12, bp, 1, beh);
        int t = 0;
        for (int i = 0; i < 6; ++i)
            if (X0[i] != 0) { t = i; break; }
        prop.assign((size_t)9 * nInc, 0.0);
        rows = nInc;
        exit_inc = -1;
        int nx = 0, total_it = 0, total32 = 0;
        std::vector<double> part((size_t)14 * nblk);
        const int ck = 4;
        // captured graphs of `ck` iterations (FP64, and FP32 for the FP32
        // phase), replayed in every increment
        struct GraphGuard {
            hipGraph_t g = nullptr;
            hipGraphExec_t x = nullptr;
            ~GraphGuard()
            {
                if (x) (void)hipGraphExecDestroy(x);
                if (g) (void)hipGraphDestroy(g);
            }
        } g64, g32;
        auto capture = [&](GraphGuard& gg, bool fp32) {
            HCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < ck; ++k) {
                if (fp32) stage_fast<float>(BMODE_ITER, k & 1);
                else run_stage(BMODE_ITER, k & 1);
            }
            HCHK(hipStreamEndCapture(st, &gg.g));
            hipError_t ge = hipGraphInstantiate(&gg.x, gg.g, nullptr, nullptr, 0);
            if (ge != hipSuccess) throw HipError{ge, "hipGraphInstantiate"};
        };
        capture(g64, false);
        if (two) capture(g32, true);
        auto replay = [&](hipGraphExec_t x) {
            for (int guard = 0; guard < 100000; ++guard) {
                HCHK(hipGraphLaunch(x, st));
                HCHK(hipMemcpyAsync(&c, ctl, sizeof c, hipMemcpyDeviceToHost, st));
                HCHK(hipStreamSynchronize(st));
                if (c.done) break;
            }
        };
        BPtr<float> bp32{};
        if (two) bp32 = bptr_t<float>();
        const int nb = 1024;
        HCHK(hipEventRecord(e.ev0, st));
        for (int inc = 0; inc < nInc; ++inc) {
            // host copy of the device state, then reset the per-increment fields
            HCHK(hipMemcpyAsync(&c, ctl, sizeof c, hipMemcpyDeviceToHost, st));
            HCHK(hipStreamSynchronize(st));
            c.incNo = inc + 1;
            c.done = 0;
            c.it = 0;
            c.err = 0;
            bool fp64_run = true;
            bool fused = false; // the FINAL pass (FP64 "ep") ended the increment
            if (two) {
                // FP32 phase from the increment's FP64 start state
                c.tol_f = std::max(tol_f * e.polish_factor, e.fp32_floor);
                c.tol_eq = std::max(tol_eq * e.polish_factor, e.fp32_floor);
                HCHK(hipMemcpyAsync(ctl, &c, sizeof c, hipMemcpyHostToDevice, st));
                hipLaunchKernelGGL((k_bcopy<double, float>), dim3(nb), dim3(256), 0, st, g12, bp, bp32,
                                   BC_EPS | BC_SIGOLD | BC_EPSOLD | BC_EP);
                stage_fast<float>(BMODE_START, 1);
                HCHK(hipGetLastError());
                replay(g32.x);
                if (c.done == 2) {
                    // iteration cap reached in FP32: that state is the result
                    hipLaunchKernelGGL((k_bcopy<float, double>), dim3(nb), dim3(256), 0, st, g12, bp32, bp,
                                       BC_EPS | BC_SIG | BC_EP | BC_PHI);
                    total32 += c.it;
                    fp64_run = false;
                } else if (c.done == 1) {
                    // the FP32 rule holds at iterate itNo: recompute that iterate
                    // in FP64 from its strain (and the damage of its constitutive
                    // call), then iterate in FP64 under the reference rule
                    hipLaunchKernelGGL((k_bcopy<float, double>), dim3(nb), dim3(256), 0, st, g12, bp32, bp,
                                       BC_EPS | (beh == BEH_ED ? BC_EP : 0));
                    total32 += c.it;
                    c.done = 0;
                    c.tol_f = tol_f;
                    c.tol_eq = tol_eq;
                    const int it0 = c.it;
                    HCHK(hipMemcpyAsync(ctl, &c, sizeof c, hipMemcpyHostToDevice, st));
                    stage_fast<double>(BMODE_RESTART, 1);
                    HCHK(hipGetLastError());
                    replay(g64.x);
                    if (c.done != 3) {
                        final_fast(c, it0);
                        fused = beh == BEH_EP;
                    }
                    fp64_run = false;
                } else {
                    // the FP32 return map failed: redo the increment in FP64 (its
                    // start state is untouched)
                    c.done = 0;
                    c.it = 0;
                    c.err = 0;
                }
            }
            if (fp64_run) {
                c.tol_f = tol_f;
                c.tol_eq = tol_eq;
                HCHK(hipMemcpyAsync(ctl, &c, sizeof c, hipMemcpyHostToDevice, st));
                run_stage(BMODE_START, 1);
                HCHK(hipGetLastError());
                replay(g64.x);
                if (fast && c.done != 3) {
                    final_fast(c, 0);
                    fused = beh == BEH_EP;
                }
            }
            if (c.done == 3)
                throw ArgError{M3D_ERR_NUMERIC, "constitutive update: Newton iteration failed (NaN or no convergence)"};
            total_it += c.it;
            info.iterations = c.it;
            info.status = c.done;
            info.eqVar = c.eqvar;
            info.diff_sum = c.dsum;
            info.diff_eq = c.deq;
            // export_homogen_var(inc) (:722-746): from the fused FINAL pass's
            // block sums, or from the fields
            if (fused)
                hipLaunchKernelGGL((k_bfsums<0>), dim3(nblk, 14), dim3(256), 0, st, gf, (const double*)pfin, bs, nblk);
            else if (beh == BEH_EP)
                hipLaunchKernelGGL((k_bsums<double, BEH_EP>), dim3(nblk, 1), dim3(256), 0, st, g12, bp, bs, nblk);
            else
                hipLaunchKernelGGL((k_bsums<double, BEH_ED>), dim3(nblk, 1), dim3(256), 0, st, g12, bp, bs, nblk);
            HCHK(hipGetLastError());
            HCHK(hipMemcpyAsync(part.data(), bs, sizeof(double) * part.size(), hipMemcpyDeviceToHost, st));
            HCHK(hipStreamSynchronize(st));
            double m[14];
            for (int q = 0; q < 14; ++q) {
                double a = 0;
                for (int b = 0; b < nblk; ++b) a += part[(size_t)q * nblk + b];
                m[q] = a / (double)e.N;
            }
            const size_t R = (size_t)nInc;
            prop[inc + 0 * R] = m[t];
            for (int j = 1; j < 7; ++j) prop[inc + (size_t)j * R] = m[6 + j - 1];
            prop[inc + 7 * R] = m[12];
            prop[inc + 8 * R] = beh == BEH_EP ? m[13] : 0.0;
            // "ed" softening exit (:829-847), evaluated on the row just computed
            if (beh == BEH_ED && inc != 0)
                for (int j = 0; j < 6; ++j)
                    if (X0[j] != 0 && prop[inc + (size_t)(j + 1) * R] < prop[inc - 1 + (size_t)(j + 1) * R]) nx += 1;
            // softening exit: nInc = _i (:849); prop_homogenized keeps its
            // nInc rows, the ones past the exit increment stay zero
            if (nx == 3) {
                exit_inc = inc;
                break;
            }
            // update_old_Variable (:693-719); E += dE does not feed back (updateE recomputes E).
            // The fused FINAL pass has applied it; else here, and on the specialised
            // kernels sigma_star for the next start stage
            if (!fused) {
                if (beh == BEH_EP)
                    hipLaunchKernelGGL((k_bupdate<double, BEH_EP>), dim3(8192), dim3(256), 0, st, g12, bp, 1);
                else
                    hipLaunchKernelGGL((k_bupdate<double, BEH_ED>), dim3(8192), dim3(256), 0, st, g12, bp, 1);
                if (beh == BEH_EP && fast) hipLaunchKernelGGL((k_bstar<double>), dim3(8192), dim3(256), 0, st, g12, bp);
            }
        }
        HCHK(hipEventRecord(e.ev1, st));
        HCHK(hipStreamSynchronize(st));
        float ms = 0;
        HCHK(hipEventElapsedTime(&ms, e.ev0, e.ev1));
        info.time_ms = ms;
        info.iters_fp64 = total_it - total32;
        info.iters_fp32 = total32;
    }

    void fields(double* eps, double* sig)
    {
        const size_t n6 = (size_t)6 * e.N;
        if (eps) HCHK(hipMemcpy(eps, f6, sizeof(double) * n6, hipMemcpyDeviceToHost));
        if (sig) HCHK(hipMemcpy(sig, f6 + n6, sizeof(double) * n6, hipMemcpyDeviceToHost));
    }

    // per-voxel history: equivalent plastic strain ("ep") or damage ("ed"),
    // and the von Mises stress of the last return map ("ep"; zeros for "ed")
    void state(double* ep_or_d, double* seff)
    {
        const size_t n = (size_t)e.N;
        if (ep_or_d) HCHK(hipMemcpy(ep_or_d, f1, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (seff) {
            if (beh == BEH_EP) HCHK(hipMemcpy(seff, f1 + 3 * n, sizeof(double) * n, hipMemcpyDeviceToHost));
            else std::fill(seff, seff + n, 0.0);
        }
    }
};

} // namespace
