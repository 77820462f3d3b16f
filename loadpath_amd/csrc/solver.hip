// The grid handle, the V-cycle and the batched PCG, and the C-ABI around them.
//
// Reference: mg.GridElasticity (mg.py:411-988): _smooth (Chebyshev,
// mg.py:796-835), _vcycle (mg.py:837-860), precondition (mg.py:862-864),
// _pcg (mg.py:872-926), solve_many / _solve with the STALL -> strict fallback
// (mg.py:928-965), _estimate_lmax (mg.py:771-794), strain_energy
// (mg.py:975-988). The reference's GPU path (gpu.TorchMG, gpu.py:280-521)
// syncs to the host for every dot product and solves the coarsest level on
// the host; here alpha, beta and |r| live on the device and the host reads
// one small status block per CG iteration.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>

#include "ke_pattern.h"
#include "lp_internal.h"

namespace lp {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

namespace {

// fixed hash start vector in [-0.5, 0.5) for lambda_max when none is given
template <typename T>
__global__ void k_hash_start(T* v, const uint8_t* __restrict__ mask, size_t nn, unsigned seed) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nn) return;
    unsigned long long h = (unsigned long long)i * 0x9E3779B97F4A7C15ull + seed * 0xD1B54A32D192ED03ull;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    const double r = (double)(h >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    const int c = (int)(i / nn);
    v[i] = ((mask[i % nn] >> c) & 1) ? (T)r : T(0);
}

template <typename T>
__global__ void k_mask_vec(T* v, const uint8_t* __restrict__ mask, size_t nn) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nn) return;
    const int c = (int)(i / nn);
    if (!((mask[i % nn] >> c) & 1)) v[i] = T(0);
}

template <typename T>
__global__ void k_from_double_masked(const double* __restrict__ src, T* dst, const uint8_t* __restrict__ mask, size_t nn) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * nn) return;
    const int c = (int)(i / nn);
    const T v = (T)src[i];
    dst[i] = ((mask[i % nn] >> c) & 1) ? v * T(1) : T(0);
}

// se[rhs][cell] = u_c^T Ke u_c (FP64), one thread per cell
__global__ void k_strain_energy(const double* __restrict__ u, const double* __restrict__ ke, Geom g, double* se,
                                size_t vstride) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (cell >= g.ncell) return;
    const int Z = (int)(cell % g.nz), Y = (int)((cell / g.nz) % g.ny), X = (int)(cell / ((size_t)g.nz * g.ny));
    const double* uu = u + rhs * vstride;
    double ue[24];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const size_t node = ((size_t)(X + (b & 1)) * g.Ny + (Y + ((b >> 1) & 1))) * g.Nz + (Z + ((b >> 2) & 1));
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[3 * b + c] = uu[c * g.nn + node];
    }
    double acc = 0;
#pragma unroll
    for (int i = 0; i < 24; ++i) {
        double t = 0;
#pragma unroll
        for (int j = 0; j < 24; ++j) t += ke[i * 24 + j] * ue[j];
        acc += ue[i] * t;
    }
    se[rhs * g.ncell + cell] = acc;
}

inline unsigned nblk(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void swap_buf(DBuf& a, DBuf& b) {
    std::swap(a.p, b.p);
    std::swap(a.bytes, b.bytes);
}

int api_guard_code(int code) { return code; }

template <typename F> int guarded(F&& f) {
    try {
        return f();
    } catch (const HipFailure&) {
        return LP_ERR_HIP;
    } catch (const ArgFailure&) {
        return LP_ERR_ARG;
    } catch (const std::exception& e) {
        set_error(e.what());
        return LP_ERR_STATE;
    }
}

}  // namespace

// ---- lambda_max ------------------------------------------------------------------
static double estimate_lmax_level(Grid& G, int i, const double* start_host) {
    Level& L = G.lv[i];
    const int eb = G.vbytes();
    const size_t n = 3 * L.g.nn;
    hipStream_t s = G.stream;
    L.lmax_vec.alloc(n * eb);
    if (start_host) {
        DBuf& tmp = G.work.lmax_tmp;
        tmp.alloc(n * 8);
        LP_HIP(hipMemcpyAsync(tmp.p, start_host, n * 8, hipMemcpyHostToDevice, s));
        if (eb == 8)
            hipLaunchKernelGGL(k_from_double_masked<double>, dim3(nblk(n)), dim3(kBlock), 0, s, tmp.as<double>(),
                               L.lmax_vec.as<double>(), L.mask.as<uint8_t>(), L.g.nn);
        else
            hipLaunchKernelGGL(k_from_double_masked<float>, dim3(nblk(n)), dim3(kBlock), 0, s, tmp.as<double>(),
                               L.lmax_vec.as<float>(), L.mask.as<uint8_t>(), L.g.nn);
        LP_HIP(hipGetLastError());
        LP_HIP(hipStreamSynchronize(s));
    } else if (L.lmax_warm) {
        if (eb == 8)
            hipLaunchKernelGGL(k_mask_vec<double>, dim3(nblk(n)), dim3(kBlock), 0, s, L.lmax_vec.as<double>(),
                               L.mask.as<uint8_t>(), L.g.nn);
        else
            hipLaunchKernelGGL(k_mask_vec<float>, dim3(nblk(n)), dim3(kBlock), 0, s, L.lmax_vec.as<float>(),
                               L.mask.as<uint8_t>(), L.g.nn);
        LP_HIP(hipGetLastError());
    } else {
        if (eb == 8)
            hipLaunchKernelGGL(k_hash_start<double>, dim3(nblk(n)), dim3(kBlock), 0, s, L.lmax_vec.as<double>(),
                               L.mask.as<uint8_t>(), L.g.nn, (unsigned)i);
        else
            hipLaunchKernelGGL(k_hash_start<float>, dim3(nblk(n)), dim3(kBlock), 0, s, L.lmax_vec.as<float>(),
                               L.mask.as<uint8_t>(), L.g.nn, (unsigned)i);
        LP_HIP(hipGetLastError());
    }
    const int its = G.opt.lmax_iters > 0 ? G.opt.lmax_iters : 12;
    DBuf& sc = G.work.lmax_sc;
    DBuf& av = G.work.lmax_av;
    sc.alloc(2 * sizeof(double) * its);
    av.alloc(n * eb);
    const void* dinv = eb == 8 ? L.dinv64.p : L.dinv32.p;
    const int nbd = dot_blocks(n);
    const int nba = apply_blocks(G, i, eb);
    ensure_partials(G, (size_t)nbd + nba);
    double* S = sc.as<double>();
    // per step: v . v (partials Pd), then one launch that stores its sum,
    // normalizes v and sums the previous step's v . (dinv A v) partials (Pa),
    // then the operator; the last step's partials are summed after the loop
    double* Pd = G.partials.as<double>();
    double* Pa = Pd + nbd;
    for (int it = 0; it < its; ++it) {
        launch_dot(L.lmax_vec.p, L.lmax_vec.p, eb, n, n, 1, Pd, nullptr, s);
        launch_norm_scale(L.lmax_vec.p, eb, n, Pd, nbd, S + 2 * it, it > 0 ? Pa : nullptr, nba,
                          it > 0 ? S + 2 * it - 1 : nullptr, s);
        EpiArgs E;
        E.kind = EPI_POWER;
        E.out = av.p;
        E.dinv = dinv;
        E.partials = Pa;
        launch_apply(G, i, eb, L.lmax_vec.p, n, 1, E, s);
        swap_buf(L.lmax_vec, av);
    }
    launch_reduce(Pa, nba, 1, S + 2 * its - 1, 0, s);
    std::vector<double> h(2 * its);
    LP_HIP(hipMemcpyAsync(h.data(), S, h.size() * 8, hipMemcpyDeviceToHost, s));
    LP_HIP(hipStreamSynchronize(s));
    L.lmax_warm = true;
    ++G.epoch;
    double lam = 1.0;
    for (int it = 0; it < its; ++it) {
        if (h[2 * it] == 0.0) return 2.0;   // the reference returns 2.0 when the vector vanishes
        lam = h[2 * it + 1];
    }
    return 1.3 * std::max(lam, 1e-12);
}

// ---- smoothers and V-cycle ----------------------------------------------------------
static void smooth(Grid& G, int i, int nrhs, bool from_zero) {
    Level& L = G.lv[i];
    const int eb = G.vbytes();
    const size_t n = 3 * L.g.nn;
    hipStream_t s = G.stream;
    const void* dinv = eb == 8 ? L.dinv64.p : L.dinv32.p;
    if (G.opt.smoother == LP_SMOOTH_JACOBI) {
        int sweeps = G.opt.degree;
        if (from_zero) {
            launch_jacobi_start(L.vb.p, dinv, G.opt.omega, L.vx.p, eb, n, n, nrhs, s);
            --sweeps;
        }
        for (int k = 0; k < sweeps; ++k) {
            EpiArgs E;
            E.kind = EPI_JACOBI;
            E.out = L.vt.p;
            E.b = L.vb.p;
            E.dinv = dinv;
            E.c1 = G.opt.omega;
            launch_apply(G, i, eb, L.vx.p, n, nrhs, E, s);
            swap_buf(L.vx, L.vt);
        }
        return;
    }
    const double hi = L.lmax, lo = hi / 30.0;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    if (G.tune_fused_cheb && mfma_level(G, i, eb) && G.opt.degree >= 2) {
        // the same recurrence with its vector passes fused into the operator
        // launches: the first step stages d = (dinv b) / theta itself, the
        // last adds the next direction to x (no start or update passes)
        const int napply = G.opt.degree - 1;
        // from a nonzero x with one step to take (degree 2): the residual
        // step stores r only and the step stages d = (dinv r) / theta from
        // it, so d is neither written nor read back
        const bool resid_only = !from_zero && napply == 1 && G.tune_cheb_r != 0;
        if (!from_zero) {
            EpiArgs E;
            E.b = L.vb.p;
            E.dinv = dinv;
            if (resid_only) {
                E.kind = EPI_RESID;
                E.out = L.vr.p;
            } else {
                E.kind = EPI_RESID_SCALED;
                E.r = L.vr.p;
                E.d2 = L.vd.p;
                E.c1 = 1.0 / theta;
            }
            launch_apply(G, i, eb, L.vx.p, n, nrhs, E, s);
        }
        for (int it = 0; it < napply; ++it) {
            const bool first = from_zero && it == 0, last = it + 1 == napply;
            const double rho_new = 1.0 / (2.0 * sigma - rho);
            EpiArgs E;
            E.kind = first ? (last ? EPI_CHEB_ONLY : EPI_CHEB_FIRST) : (last ? EPI_CHEB_LAST : EPI_CHEB);
            if (resid_only) E.kind = EPI_CHEB_LAST_R;
            E.x = L.vx.p;
            E.r = L.vr.p;
            E.d2 = L.vd2.p;
            E.dinv = dinv;
            E.c1 = rho_new * rho;
            E.c2 = 2.0 * rho_new / delta;
            E.c3 = 1.0 / theta;
            launch_apply(G, i, eb, first ? L.vb.p : (resid_only ? L.vr.p : L.vd.p), n, nrhs, E, s);
            if (!last) swap_buf(L.vd, L.vd2);
            rho = rho_new;
        }
        return;
    }
    if (from_zero) {
        launch_cheb_start(L.vb.p, dinv, 1.0 / theta, L.vx.p, L.vr.p, L.vd.p, eb, n, n, nrhs, s);
    } else {
        EpiArgs E;
        E.kind = EPI_RESID_SCALED;
        E.b = L.vb.p;
        E.r = L.vr.p;
        E.d2 = L.vd.p;
        E.dinv = dinv;
        E.c1 = 1.0 / theta;
        launch_apply(G, i, eb, L.vx.p, n, nrhs, E, s);
    }
    for (int it = 0; it < G.opt.degree; ++it) {
        const bool last = it + 1 == G.opt.degree;
        const double rho_new = 1.0 / (2.0 * sigma - rho);
        if (!last) {
            EpiArgs E;
            E.kind = EPI_CHEB;
            E.x = L.vx.p;
            E.r = L.vr.p;
            E.d = L.vd.p;
            E.d2 = L.vd2.p;
            E.dinv = dinv;
            E.c1 = rho_new * rho;
            E.c2 = 2.0 * rho_new / delta;
            launch_apply(G, i, eb, L.vd.p, n, nrhs, E, s);
            swap_buf(L.vd, L.vd2);
        } else {
            launch_axpy_inplace(L.vx.p, L.vd.p, eb, n, n, nrhs, s);
        }
        rho = rho_new;
    }
}

void vcycle(Grid& G, int i, int nrhs) {
    Level& L = G.lv[i];
    const int eb = G.vbytes();
    const size_t n = 3 * L.g.nn;
    hipStream_t s = G.stream;
    if (i == G.nlevels - 1) {
        if (G.ncoarse > 0) {
            launch_dense_matvec(G.cinv.as<double>(), G.ncoarse, L.vb.p, L.vx.p, eb, nrhs, n, s);
        } else {
            smooth(G, i, nrhs, true);
            smooth(G, i, nrhs, false);
        }
        return;
    }
    smooth(G, i, nrhs, true);
    Level& C = G.lv[i + 1];
    const size_t nc = 3 * C.g.nn;
    {
        EpiArgs E;
        E.kind = EPI_RESID_MASK;
        E.out = L.vt.p;
        E.b = L.vb.p;
        launch_apply(G, i, eb, L.vx.p, n, nrhs, E, s);
    }
    launch_restrict_masked(L.vt.p, L.g, C.mask.as<uint8_t>(), C.vb.p, C.g, eb, nrhs, n, nc, s);
    vcycle(G, i + 1, nrhs);
    launch_prolong_add(C.vx.p, C.mask.as<uint8_t>(), C.g, L.mask.as<uint8_t>(), L.vx.p, L.g, eb, nrhs, nc, n, s);
    smooth(G, i, nrhs, false);
}

// z = M r for the FP64 CG vectors r -> z
// one V-cycle on the level-0 vectors: replayed from a captured hipGraph
// when that is valid (Chebyshev smoothing, whose buffer swaps cancel over a
// V-cycle; no profiling events), captured again when the epoch changed
static void vcycle_run(Grid& G, int nrhs) {
    const bool use_graph = G.tune_graph && G.opt.smoother == LP_SMOOTH_CHEBYSHEV && !G.prof.on;
    if (!use_graph) {
        vcycle(G, 0, nrhs);
        return;
    }
    if (!G.vgraph || G.vgraph_nrhs != nrhs || G.vgraph_epoch != G.epoch) {
        if (G.vgraph) {
            LP_HIP(hipGraphExecDestroy(G.vgraph));
            G.vgraph = nullptr;
        }
        hipGraph_t graph = nullptr;
        LP_HIP(hipStreamBeginCapture(G.stream, hipStreamCaptureModeThreadLocal));
        try {
            vcycle(G, 0, nrhs);
        } catch (...) {
            hipStreamEndCapture(G.stream, &graph);
            if (graph) hipGraphDestroy(graph);
            throw;
        }
        LP_HIP(hipStreamEndCapture(G.stream, &graph));
        const hipError_t e = hipGraphInstantiate(&G.vgraph, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        LP_HIP(e);
        G.vgraph_nrhs = nrhs;
        G.vgraph_epoch = G.epoch;
    }
    LP_HIP(hipGraphLaunch(G.vgraph, G.stream));
}

static void precondition(Grid& G, const double* r, double* z, int nrhs) {
    const int eb = G.vbytes();
    Level& L = G.lv[0];
    const size_t n = 3 * L.g.nn;
    // (the FP16-MFMA operators of the mixed path scale each cell's operand
    // themselves, so the V-cycle input needs no range scaling)
    launch_cast(r, 8, L.vb.p, eb, n, n, nrhs, G.stream);
    vcycle(G, 0, nrhs);
    launch_cast(L.vx.p, eb, z, 8, n, n, nrhs, G.stream);
}

static void read_scal(Grid& G, int nrhs, std::vector<CgScal>& out) {
    out.resize(nrhs);
    LP_HIP(hipMemcpyAsync(G.host_pinned, G.scal.p, nrhs * sizeof(CgScal), hipMemcpyDeviceToHost, G.stream));
    LP_HIP(hipStreamSynchronize(G.stream));
    std::memcpy(out.data(), G.host_pinned, nrhs * sizeof(CgScal));
}

// Batched PCG on G.b -> G.x (FP64). Returns per-rhs iterations and flags.
static void pcg_run(Grid& G, int nrhs, bool warm, int* iters, int* converged) {
    hipStream_t s = G.stream;
    const size_t n = 3 * G.g0.nn;
    const int nbd = dot_blocks(n);
    const int nba = apply_blocks(G, 0, 8);
    ensure_partials(G, (size_t)nrhs * std::max(2 * nbd, nba));
    double* P = G.partials.as<double>();
    CgScal* S = G.scal.as<CgScal>();
    const bool flex = G.flex();
    launch_dot(G.b.p, G.b.p, 8, n, n, nrhs, P, nullptr, s);
    launch_cg_finalize(G, nrhs, 3, nbd, s);
    if (warm) {
        EpiArgs E;
        E.kind = EPI_STORE_DOT;
        E.out = G.ap.p;
        E.partials = P;
        launch_apply(G, 0, 8, G.x.p, n, nrhs, E, s);
        launch_cg_finalize(G, nrhs, 5, nba, s);
        launch_dot(G.b.p, G.x.p, 8, n, n, nrhs, P, nullptr, s);
        launch_cg_finalize(G, nrhs, 6, nbd, s);
        launch_warm(G, nrhs, 0, s);
        launch_cg_finalize(G, nrhs, 7, nbd, s);
        launch_warm(G, nrhs, 1, s);
    } else {
        launch_fill(G.x.p, 8, n * nrhs, 0.0, s);
        launch_cast(G.b.p, 8, G.r.p, 8, n, n, nrhs, s);
    }
    // z = M r: the V-cycle's result is only summed against r (and Ap for
    // the flexible beta) and folded into p, never stored in FP64; later
    // iterations get the V-cycle input from the CG step itself
    Level& L0 = G.lv[0];
    const int eb = G.vbytes();
    launch_cast(G.r.p, 8, L0.vb.p, eb, n, n, nrhs, s);
    vcycle_run(G, nrhs);
    launch_cast_dot(L0.vx.p, eb, G.r.as<double>(), nullptr, n, n, nrhs, P, S, s);
    launch_cg_finalize(G, nrhs, 4, nbd, s);
    launch_cast(L0.vx.p, eb, G.p.p, 8, n, n, nrhs, s);
    std::vector<CgScal> h;
    bool pending = false;   // alpha of the last iteration not yet added to x
    // from the second iteration on, the operator launch also forms the new
    // direction p = z + beta p_old (into p2, then swapped) and adds the last
    // step alpha p_old to x: the separate update pass is gone
    const bool fused = G.tune_cg_fused != 0;
    bool deferred = false;  // the last iteration's p and x updates left to the next operator launch
    for (int it = 1; it <= G.opt.maxiter; ++it) {
        EpiArgs E;
        E.kind = EPI_STORE_DOT;
        E.out = G.ap.p;
        E.partials = P;
        E.scal = S;
        if (deferred) {
            E.kind = EPI_CG_APPLY;
            E.d = L0.vx.p;
            E.zbytes = eb;
            E.d2 = G.p2.p;
            E.x = G.x.p;
        }
        launch_apply(G, 0, 8, G.p.p, n, nrhs, E, s);
        if (deferred) {
            swap_buf(G.p, G.p2);
            deferred = false;
        }
        launch_cg_finalize(G, nrhs, 0, nba, s);
        launch_cg_step(G, nrhs, L0.vb.p, eb, s);
        launch_cg_finalize(G, nrhs, 1, nbd, s);
        pending = true;
        read_scal(G, nrhs, h);
        bool any = false;
        for (int k = 0; k < nrhs; ++k) any |= h[k].active != 0;
        if (!any) break;
        vcycle_run(G, nrhs);
        if (flex) {
            launch_cast_dot(L0.vx.p, eb, G.r.as<double>(), nullptr, n, n, nrhs, P, S, s, G.ap.as<double>());
            launch_cg_beta_flex(G, nrhs, nbd, s);
        } else {
            launch_cast_dot(L0.vx.p, eb, G.r.as<double>(), nullptr, n, n, nrhs, P, S, s);
            launch_cg_finalize(G, nrhs, 2, nbd, s);
        }
        // x += alpha p with this iteration's p, then p = z + beta p (in the
        // next operator launch when fused)
        if (fused)
            deferred = true;
        else
            launch_cg_update(G, nrhs, L0.vx.p, eb, false, s);
        pending = false;
    }
    if (pending) launch_cg_update(G, nrhs, L0.vx.p, eb, true, s);
    // the cap reached with the updates still deferred: apply them (the
    // direction too, as the separate pass would have)
    if (deferred) launch_cg_update(G, nrhs, L0.vx.p, eb, false, s);
    read_scal(G, nrhs, h);
    for (int k = 0; k < nrhs; ++k) {
        iters[k] = h[k].iters + (h[k].failed ? 1 : 0);
        converged[k] = h[k].converged;
    }
}

// The mixed path's CG in FP32 with reliable FP64 updates (iterative
// refinement inside one Krylov run): the direction p, its product A p (the
// FP32 vector operator), the residual r (the V-cycle's input) and the
// solution since the last update xlo are FP32; the FP64 solution x is
// updated, and the residual recomputed in FP64 as b - A x, whenever a
// right-hand side's FP32 residual falls below delta times its largest since
// the last update or below the tolerance. The recomputed residual replaces
// the FP32 one and the iteration goes on with the same direction (a
// Fletcher-Reeves beta at that step); convergence is decided on the FP64
// residual only, so the answer meets the tolerance as the FP64 CG's does.
// The warm start and the stopping rule are the FP64 CG's (mg.py:872-926).
static void pcg32_run(Grid& G, int nrhs, bool warm, int* iters, int* converged) {
    hipStream_t s = G.stream;
    const size_t n = 3 * G.g0.nn;
    const int nbd = dot_blocks(n);
    const int nba64 = apply_blocks(G, 0, 8);
    const int nba32 = fine_vector_blocks(G);   // the FP32 vector operator's partials per rhs
    ensure_partials(G, (size_t)nrhs * std::max({2 * nbd, nba64, nba32}));
    double* P = G.partials.as<double>();
    CgScal* S = G.scal.as<CgScal>();
    Level& L0 = G.lv[0];
    float* r32 = L0.vb.as<float>();
    const float* z32 = L0.vx.as<float>();
    const bool flex = G.flex();
    launch_dot(G.b.p, G.b.p, 8, n, n, nrhs, P, nullptr, s);
    launch_cg_finalize(G, nrhs, 3, nbd, s);
    if (warm) {
        EpiArgs E;
        E.kind = EPI_STORE_DOT;
        E.out = G.ap.p;
        E.partials = P;
        launch_apply(G, 0, 8, G.x.p, n, nrhs, E, s);
        launch_cg_finalize(G, nrhs, 5, nba64, s);
        launch_dot(G.b.p, G.x.p, 8, n, n, nrhs, P, nullptr, s);
        launch_cg_finalize(G, nrhs, 6, nbd, s);
        launch_warm(G, nrhs, 0, s);
        launch_cg_finalize(G, nrhs, 7, nbd, s);
        launch_warm(G, nrhs, 1, s);
        launch_cast(G.r.p, 8, r32, 4, n, n, nrhs, s);
    } else {
        launch_fill(G.x.p, 8, n * nrhs, 0.0, s);
        launch_cast(G.b.p, 8, r32, 4, n, n, nrhs, s);
    }
    launch_fill(G.xlo32.p, 4, n * nrhs, 0.0, s);
    std::vector<CgScal> h;
    read_scal(G, nrhs, h);
    // per rhs: the largest FP32 residual norm since the last update
    std::vector<double> rmax(nrhs);
    for (int k = 0; k < nrhs; ++k) rmax[k] = h[k].bnorm;
    if (warm) {
        // the warm start's residual norm: |b - s A x| when kept, else |b|
        launch_dot(G.r.p, G.r.p, 8, n, n, nrhs, P, nullptr, s);
        launch_cg_finalize(G, nrhs, 5, nbd, s);
        read_scal(G, nrhs, h);
        for (int k = 0; k < nrhs; ++k) rmax[k] = std::sqrt(h[k].tmp);
    }
    vcycle_run(G, nrhs);
    launch_dot2_f32(G, nrhs, r32, z32, nullptr, s);
    launch_cg_finalize(G, nrhs, 4, nbd, s);
    launch_cg32_update(G, nrhs, z32, true, s);
    for (int it = 1; it <= G.opt.maxiter; ++it) {
        EpiArgs E;
        E.kind = EPI_STORE_DOT;
        E.out = G.ap32.p;
        E.partials = P;
        E.scal = S;
        launch_apply_fine(G, 0, 4, G.p32.p, n, nrhs, E, s);
        launch_cg_finalize(G, nrhs, 0, nba32, s);
        launch_cg32_step(G, nrhs, r32, s);
        launch_cg_finalize(G, nrhs, 8, nbd, s);
        read_scal(G, nrhs, h);
        bool any = false, check = false;
        for (int k = 0; k < nrhs; ++k) {
            if (!h[k].active) continue;
            any = true;
            const double rn = std::sqrt(h[k].rr);
            if (!(rn > h[k].tol) || rn < G.cg32_delta * rmax[k] || it == G.opt.maxiter) check = true;
            rmax[k] = std::max(rmax[k], rn);
        }
        if (!any) break;
        bool replaced = false;
        if (check) {
            // the reliable update: x += xlo; r = b - A x in FP64; r32 = r
            launch_cg32_fold(G, nrhs, s);
            EpiArgs R;
            R.kind = EPI_RESID;
            R.b = G.b.p;
            R.out = G.r.p;
            R.partials = P;
            R.scal = S;
            launch_apply(G, 0, 8, G.x.p, n, nrhs, R, s);
            launch_cg_finalize(G, nrhs, 9, nba64, s);
            launch_cast_active(G, nrhs, G.r.as<double>(), r32, s);
            read_scal(G, nrhs, h);
            ++G.cg32_updates;
            replaced = true;
            any = false;
            for (int k = 0; k < nrhs; ++k) {
                if (!h[k].active) continue;
                any = true;
                rmax[k] = std::sqrt(h[k].rr);
            }
            if (!any) break;
        }
        vcycle_run(G, nrhs);
        if (replaced && G.tune_cg32_restart) {
            // iterative refinement: the FP32 PCG restarts from the FP64
            // residual (gpu.py:429-481's outer step)
            launch_dot2_f32(G, nrhs, r32, z32, nullptr, s);
            launch_cg_finalize(G, nrhs, 10, nbd, s);
        } else if (flex && (!replaced || G.cg32_flex_replaced)) {
            launch_dot2_f32(G, nrhs, r32, z32, G.ap32.as<float>(), s);
            launch_cg_beta_flex(G, nrhs, nbd, s);
        } else {
            launch_dot2_f32(G, nrhs, r32, z32, nullptr, s);
            launch_cg_finalize(G, nrhs, 2, nbd, s);
        }
        launch_cg32_update(G, nrhs, z32, false, s);
    }
    // rhs still active at the cap: their last steps are in xlo
    launch_cg32_fold(G, nrhs, s);
    read_scal(G, nrhs, h);
    for (int k = 0; k < nrhs; ++k) {
        iters[k] = h[k].iters + (h[k].failed ? 1 : 0);
        converged[k] = h[k].converged;
    }
}

// PCG with the mixed path's safety net: when a right-hand side breaks down
// or reaches the cap with the FP16 levels in the V-cycle, the solve is
// repeated with FP32 operators on every level, warm from the iterates (the
// warm start drops a non-finite or useless iterate to a cold start); the
// switch holds until the hierarchy is rebuilt
void pcg(Grid& G, int nrhs, bool warm, int* iters, int* converged) {
    const bool net = G.fp16() && G.tune_fallback != 0;
    const int full = G.opt.maxiter;
    if (net && G.tune_fallback == 2) G.opt.maxiter = 1;
    // the first attempt of the mixed path may run the CG in FP32 with
    // reliable FP64 updates; the repeat is the FP64 CG
    if (G.fp16() && G.tune_cg32 != 0) pcg32_run(G, nrhs, warm, iters, converged);
    else pcg_run(G, nrhs, warm, iters, converged);
    G.opt.maxiter = full;
    if (!net) return;
    bool ok = true;
    for (int k = 0; k < nrhs; ++k) ok &= converged[k] != 0;
    if (ok) return;
    G.fp16_off = true;
    ++G.epoch;   // a captured V-cycle holds the FP16 kernels
    ++G.fallbacks;
    std::vector<int> it2(nrhs), ok2(nrhs);
    pcg_run(G, nrhs, true, it2.data(), ok2.data());
    for (int k = 0; k < nrhs; ++k) {
        iters[k] += it2[k];
        converged[k] = ok2[k];
    }
}

static void estimate_all_lmax(Grid& G) {
    for (int i = 0; i < G.nlevels; ++i) G.lv[i].lmax = estimate_lmax_level(G, i, nullptr);
}

static int set_moduli_impl(Grid& G, bool defer) {
    int rc = build_hierarchy(G);
    if (rc == LP_SINGULAR) {
        if (G.strict) return LP_SINGULAR;
        G.strict = true;
        rc = build_hierarchy(G);
        if (rc != LP_OK) return rc;
    }
    G.built = true;
    if (!defer && G.opt.smoother == LP_SMOOTH_CHEBYSHEV) estimate_all_lmax(G);
    return LP_OK;
}

int rebuild_hierarchy(Grid& G) { return set_moduli_impl(G, false); }

// GridElasticity._solve (mg.py:950-965) on the batch in G.b / G.x: the plain
// hierarchy gets at most kStall CG iterations; if a right-hand side has not
// converged by then, the strict hierarchy is built from the same moduli (and
// stays: G.strict is sticky, as the reference's attribute is) and every
// right-hand side is solved again from its original start. The iteration
// counts add up. Returns LP_OK or the rebuild's status; converged[] tells
// whether each right-hand side met the tolerance.
int solve_stall(Grid& G, int nrhs, bool warm, int* iters, int* converged) {
    const bool stall = G.stall != 0 && !G.strict && G.nlevels >= 2;
    if (!stall) {
        pcg(G, nrhs, warm, iters, converged);
        return LP_OK;
    }
    const size_t bytes = 3 * G.g0.nn * (size_t)nrhs * 8;
    if (warm) LP_HIP(hipMemcpyAsync(G.x0.p, G.x.p, bytes, hipMemcpyDeviceToDevice, G.stream));
    const int full = G.opt.maxiter;
    G.opt.maxiter = std::min(full, G.stall == 2 ? 1 : kStall);
    try {
        pcg(G, nrhs, warm, iters, converged);
    } catch (...) {
        G.opt.maxiter = full;
        throw;
    }
    G.opt.maxiter = full;
    bool ok = true;
    for (int k = 0; k < nrhs; ++k) ok &= converged[k] != 0;
    if (ok) return LP_OK;
    G.strict = true;
    ++G.strict_switches;
    const int rc = set_moduli_impl(G, false);
    if (rc != LP_OK) return rc;
    if (warm) LP_HIP(hipMemcpyAsync(G.x.p, G.x0.p, bytes, hipMemcpyDeviceToDevice, G.stream));
    std::vector<int> it2(nrhs), ok2(nrhs);
    pcg(G, nrhs, warm, it2.data(), ok2.data());
    for (int k = 0; k < nrhs; ++k) {
        iters[k] += it2[k];
        converged[k] = ok2[k];
    }
    return LP_OK;
}

}  // namespace lp

using namespace lp;

struct lp_grid : public Grid {};

namespace {

void* to_device(Grid& G, DBuf& stage, const void* p, size_t bytes, bool device_ptr) {
    if (device_ptr) return const_cast<void*>(p);
    stage.alloc(bytes);
    LP_HIP(hipMemcpyAsync(stage.p, p, bytes, hipMemcpyHostToDevice, G.stream));
    return stage.p;
}

void from_device(Grid& G, void* host, const void* dev, size_t bytes) {
    LP_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, G.stream));
    LP_HIP(hipStreamSynchronize(G.stream));
}

}  // namespace

extern "C" {

const char* lp_last_error(void) { return g_err.c_str(); }
int lp_version(void) { return 1; }

int lp_grid_create(lp_grid** out, int nx, int ny, int nz, const uint8_t* fixed_grid, const double* ke,
                   const lp_options* opt) {
    return guarded([&]() -> int {
        if (!out || nx < 1 || ny < 1 || nz < 1 || !fixed_grid || !ke || !opt) {
            set_error("lp_grid_create: bad argument");
            return LP_ERR_ARG;
        }
        if ((nx > 1 && nx % 2) || (ny > 1 && ny % 2) || (nz > 1 && nz % 2)) {
            set_error("lp_grid_create: cell counts must be even (pad as mg.py:438-441)");
            return LP_ERR_ARG;
        }
        // the level kernels address one right-hand side's FP64 node vector
        // with 32-bit byte offsets and use offsets from 2^31 up as "outside"
        if ((size_t)(nx + 1) * (ny + 1) * (nz + 1) * 3 * sizeof(double) >= (1ull << 31)) {
            set_error("lp_grid_create: grid too large (at most 89,478,485 nodes)");
            return LP_ERR_ARG;
        }
        lp_grid* G = new lp_grid();
        try {
            G->opt = *opt;
            if (G->opt.max_rhs < 1) G->opt.max_rhs = 1;
            if (G->opt.degree < 1) G->opt.degree = 2;
            if (G->opt.maxiter < 1) G->opt.maxiter = 500;
            G->max_rhs = G->opt.max_rhs;
            G->g0.set(nx, ny, nz);
            LP_HIP(hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking));
            std::memcpy(G->ke, ke, sizeof(G->ke));
            // Ke from its symmetry orbits (ke_pattern.h) when it has the cube's
            // symmetry to rounding; otherwise the fine kernel reads all 576
            {
                double cnt[KE_NV] = {}, kmax = 0;
                for (int e = 0; e < 576; ++e) {
                    kmax = std::max(kmax, std::fabs(ke[e]));
                    if (KE_SGN[e]) {
                        G->ke_orbit[KE_IDX[e]] += KE_SGN[e] * ke[e];
                        cnt[KE_IDX[e]] += 1;
                    }
                }
                for (int v = 0; v < KE_NV; ++v) G->ke_orbit[v] /= std::max(cnt[v], 1.0);
                double dev = 0;
                for (int e = 0; e < 576; ++e) {
                    const double r = KE_SGN[e] ? KE_SGN[e] * G->ke_orbit[KE_IDX[e]] : 0.0;
                    dev = std::max(dev, std::fabs(r - ke[e]));
                }
                G->ke_sym = dev <= 1e-13 * kmax;
            }
            const size_t nn = G->g0.nn;
            // fixed bits per node
            std::vector<uint8_t> fb(nn, 0);
            for (int c = 0; c < 3; ++c)
                for (size_t i = 0; i < nn; ++i)
                    if (fixed_grid[c * nn + i]) fb[i] |= (uint8_t)(1 << c);
            G->fixed.alloc(nn);
            LP_HIP(hipMemcpy(G->fixed.p, fb.data(), nn, hipMemcpyHostToDevice));
            std::vector<float> kf(576);
            for (int i = 0; i < 576; ++i) kf[i] = (float)ke[i];
            G->ke64.alloc(576 * 8);
            G->ke32.alloc(576 * 4);
            LP_HIP(hipMemcpy(G->ke64.p, ke, 576 * 8, hipMemcpyHostToDevice));
            LP_HIP(hipMemcpy(G->ke32.p, kf.data(), 576 * 4, hipMemcpyHostToDevice));
            // M_q = P_q^T Ke P_q (mg.py:626)
            std::vector<double> pk(8 * 576), mq(8 * 576, 0.0);
            child_prolongations_host(pk.data());
            for (int q = 0; q < 8; ++q) {
                const double* P = pk.data() + q * 576;
                double t[576];
                for (int a = 0; a < 24; ++a)
                    for (int j = 0; j < 24; ++j) {
                        double v = 0;
                        for (int b = 0; b < 24; ++b) v += ke[a * 24 + b] * P[b * 24 + j];
                        t[a * 24 + j] = v;
                    }
                for (int i = 0; i < 24; ++i)
                    for (int j = 0; j < 24; ++j) {
                        double v = 0;
                        for (int a = 0; a < 24; ++a) v += P[a * 24 + i] * t[a * 24 + j];
                        mq[q * 576 + i * 24 + j] = v;
                    }
            }
            std::vector<float> mqf(8 * 576);
            for (int i = 0; i < 8 * 576; ++i) mqf[i] = (float)mq[i];
            G->mq64.alloc(8 * 576 * 8);
            G->mq32.alloc(8 * 576 * 4);
            LP_HIP(hipMemcpy(G->mq64.p, mq.data(), 8 * 576 * 8, hipMemcpyHostToDevice));
            LP_HIP(hipMemcpy(G->mq32.p, mqf.data(), 8 * 576 * 4, hipMemcpyHostToDevice));
            const size_t vb = 3 * nn * G->max_rhs * 8;
            G->x.alloc(vb);
            G->r.alloc(vb);
            G->p.alloc(vb);
            G->p2.alloc(vb);
            G->ap.alloc(vb);
            G->b.alloc(vb);
            G->x0.alloc(vb);
            if (opt->precision == LP_PREC_MIXED) {
                G->p32.alloc(vb / 2);
                G->ap32.alloc(vb / 2);
                G->xlo32.alloc(vb / 2);
            }
            G->scal.alloc(G->max_rhs * sizeof(CgScal));
            LP_HIP(hipMemset(G->scal.p, 0, G->max_rhs * sizeof(CgScal)));
            G->host_pinned_bytes = G->max_rhs * sizeof(CgScal) + 64;
            LP_HIP(hipHostMalloc(&G->host_pinned, G->host_pinned_bytes, hipHostMallocDefault));
            G->lv[0].e64.alloc(G->g0.ncell * 8);
        } catch (...) {
            delete G;
            throw;
        }
        *out = G;
        return LP_OK;
    });
}

void lp_grid_destroy(lp_grid* g) {
    if (!g) return;
    hipStreamSynchronize(g->stream);
    if (g->host_pinned) hipHostFree(g->host_pinned);
    hipStream_t s = g->stream;
    delete g;
    if (s) hipStreamDestroy(s);
}

int lp_set_moduli(lp_grid* g, const double* e, int flags) {
    return guarded([&]() -> int {
        if (!g || !e) return LP_ERR_ARG;
        const size_t bytes = g->g0.ncell * 8;
        if (flags & 1)
            LP_HIP(hipMemcpyAsync(g->lv[0].e64.p, e, bytes, hipMemcpyDeviceToDevice, g->stream));
        else
            LP_HIP(hipMemcpyAsync(g->lv[0].e64.p, e, bytes, hipMemcpyHostToDevice, g->stream));
        return set_moduli_impl(*g, (flags & 2) != 0);
    });
}

int lp_estimate_lmax(lp_grid* g, int level, const double* start) {
    return guarded([&]() -> int {
        if (!g || !g->built || level < 0 || level >= g->nlevels) return LP_ERR_ARG;
        g->lv[level].lmax = estimate_lmax_level(*g, level, start);
        return LP_OK;
    });
}

int lp_grid_info(lp_grid* g, int* nlevels, int* dims, int* strict, double* lmax) {
    return guarded([&]() -> int {
        if (!g) return LP_ERR_ARG;
        if (nlevels) *nlevels = g->nlevels;
        for (int i = 0; i < g->nlevels; ++i) {
            if (dims) {
                dims[3 * i] = g->lv[i].g.nx;
                dims[3 * i + 1] = g->lv[i].g.ny;
                dims[3 * i + 2] = g->lv[i].g.nz;
            }
            if (lmax) lmax[i] = g->lv[i].lmax;
        }
        if (strict) *strict = g->strict ? 1 : 0;
        return LP_OK;
    });
}

int lp_apply(lp_grid* g, int level, int dtype_bytes, const void* u, void* out, int nrhs, int device_ptr) {
    return guarded([&]() -> int {
        if (!g || !g->built || level < 0 || level >= g->nlevels || !u || !out || nrhs < 1) return LP_ERR_ARG;
        if (dtype_bytes != 4 && dtype_bytes != 8) return LP_ERR_ARG;
        if (dtype_bytes == 8 && level > 0 && g->opt.precision != LP_PREC_FP64) {
            set_error("lp_apply: FP64 coarse operators exist only in the FP64 hierarchy");
            return LP_ERR_ARG;
        }
        const Level& L = g->lv[level];
        const size_t n = 3 * L.g.nn;
        const size_t bytes = n * nrhs * dtype_bytes;
        const void* du = to_device(*g, g->stage_a, u, bytes, device_ptr);
        void* dout = device_ptr ? out : (g->stage_b.alloc(bytes), g->stage_b.p);
        EpiArgs E;
        E.kind = EPI_STORE;
        E.out = dout;
        launch_apply(*g, level, dtype_bytes, du, n, nrhs, E, g->stream);
        if (!device_ptr) from_device(*g, out, dout, bytes);
        else LP_HIP(hipStreamSynchronize(g->stream));
        return LP_OK;
    });
}

int lp_precondition(lp_grid* g, const double* r, double* z, int nrhs, int device_ptr) {
    return guarded([&]() -> int {
        if (!g || !g->built || nrhs < 1 || nrhs > g->max_rhs) return LP_ERR_ARG;
        const size_t n = 3 * g->g0.nn;
        const size_t bytes = n * nrhs * 8;
        const double* dr = static_cast<const double*>(to_device(*g, g->stage_a, r, bytes, device_ptr));
        double* dz = device_ptr ? z : (g->stage_b.alloc(bytes), g->stage_b.as<double>());
        precondition(*g, dr, dz, nrhs);
        if (!device_ptr) from_device(*g, z, dz, bytes);
        else LP_HIP(hipStreamSynchronize(g->stream));
        return LP_OK;
    });
}

int lp_solve(lp_grid* g, int nrhs, const double* b, double* x, int warm, int device_ptr, int* iters, int* converged) {
    return guarded([&]() -> int {
        if (!g || !g->built || nrhs < 1 || nrhs > g->max_rhs || !b || !x || !iters || !converged) return LP_ERR_ARG;
        Grid& G = *g;
        const size_t n = 3 * G.g0.nn;
        const size_t bytes = n * nrhs * 8;
        auto load = [&]() {
            if (device_ptr) {
                LP_HIP(hipMemcpyAsync(G.b.p, b, bytes, hipMemcpyDeviceToDevice, G.stream));
                if (warm) LP_HIP(hipMemcpyAsync(G.x.p, x, bytes, hipMemcpyDeviceToDevice, G.stream));
            } else {
                LP_HIP(hipMemcpyAsync(G.b.p, b, bytes, hipMemcpyHostToDevice, G.stream));
                if (warm) LP_HIP(hipMemcpyAsync(G.x.p, x, bytes, hipMemcpyHostToDevice, G.stream));
            }
        };
        load();
        std::vector<int> it1(nrhs), ok1(nrhs);
        const int rc = solve_stall(G, nrhs, warm != 0, it1.data(), ok1.data());
        if (rc != LP_OK) return rc;
        bool all_ok = true;
        for (int k = 0; k < nrhs; ++k) {
            iters[k] = it1[k];
            converged[k] = ok1[k];
            all_ok &= ok1[k] != 0;
        }
        if (device_ptr) {
            LP_HIP(hipMemcpyAsync(x, G.x.p, bytes, hipMemcpyDeviceToDevice, G.stream));
            LP_HIP(hipStreamSynchronize(G.stream));
        } else {
            from_device(G, x, G.x.p, bytes);
        }
        if (!all_ok) {
            set_error("linear solve did not converge; the structure may be a mechanism (check supports)");
            return LP_NOT_CONVERGED;
        }
        return LP_OK;
    });
}

__global__ void k_mask_rhs(const double* __restrict__ f, const uint8_t* __restrict__ mask, double* b, size_t nn) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= 3 * nn) return;
    const int c = (int)(i / nn);
    b[rhs * 3 * nn + i] = ((mask[i % nn] >> c) & 1) ? f[rhs * 3 * nn + i] : 0.0;
}

int lp_mask_rhs(lp_grid* g, int nrhs, const double* f, double* b, int device_ptr) {
    return guarded([&]() -> int {
        if (!g || !g->built || nrhs < 1 || !f || !b) return LP_ERR_ARG;
        Grid& G = *g;
        const size_t n = 3 * G.g0.nn;
        const double* df = static_cast<const double*>(to_device(G, G.stage_a, f, n * nrhs * 8, device_ptr));
        double* db = device_ptr ? b : (G.stage_b.alloc(n * nrhs * 8), G.stage_b.as<double>());
        hipLaunchKernelGGL(k_mask_rhs, dim3(nblk(n), nrhs), dim3(kBlock), 0, G.stream, df, G.lv[0].mask.as<uint8_t>(),
                           db, G.g0.nn);
        LP_HIP(hipGetLastError());
        if (!device_ptr) from_device(G, b, db, n * nrhs * 8);
        else LP_HIP(hipStreamSynchronize(G.stream));
        return LP_OK;
    });
}

int lp_strain_energy(lp_grid* g, int nrhs, const double* u, double* se, int device_ptr) {
    return guarded([&]() -> int {
        if (!g || nrhs < 1 || !u || !se) return LP_ERR_ARG;
        Grid& G = *g;
        const size_t n = 3 * G.g0.nn;
        const double* du = static_cast<const double*>(to_device(G, G.stage_a, u, n * nrhs * 8, device_ptr));
        const size_t sb = G.g0.ncell * nrhs * 8;
        double* dse = device_ptr ? se : (G.stage_b.alloc(sb), G.stage_b.as<double>());
        hipLaunchKernelGGL(k_strain_energy, dim3(nblk(G.g0.ncell), nrhs), dim3(kBlock), 0, G.stream, du,
                           G.ke64.as<double>(), G.g0, dse, n);
        LP_HIP(hipGetLastError());
        if (!device_ptr) from_device(G, se, dse, sb);
        else LP_HIP(hipStreamSynchronize(G.stream));
        return LP_OK;
    });
}

int lp_level_data(lp_grid* g, int level, int what, void* out) {
    return guarded([&]() -> int {
        if (!g || !g->built || level < 0 || level >= g->nlevels || !out) return LP_ERR_ARG;
        Grid& G = *g;
        Level& L = G.lv[level];
        if (what == 0) {
            from_device(G, out, L.mask.p, L.g.nn);
        } else if (what == 1) {
            from_device(G, out, L.dinv64.p, 3 * L.g.nn * 8);
        } else if (what == 2) {
            DBuf tmp;
            tmp.alloc(L.g.ncell * 24 * 8);
            level_cell_dofs(G, level, tmp.as<int64_t>());
            from_device(G, out, tmp.p, L.g.ncell * 24 * 8);
        } else if (what == 3) {
            from_device(G, out, L.held.p, L.g.nn);
        } else {
            return LP_ERR_ARG;
        }
        return LP_OK;
    });
}

int lp_prolong(const double* coarse, double* fine, int cx, int cy, int cz, int device_ptr) {
    return guarded([&]() -> int {
        if (!coarse || !fine || cx < 1 || cy < 1 || cz < 1) return LP_ERR_ARG;
        const size_t nc = (size_t)cx * cy * cz, nf = (size_t)(2 * cx - 1) * (2 * cy - 1) * (2 * cz - 1);
        DBuf a, b;
        const double* dc = coarse;
        double* df = fine;
        if (!device_ptr) {
            a.alloc(3 * nc * 8);
            b.alloc(3 * nf * 8);
            LP_HIP(hipMemcpy(a.p, coarse, 3 * nc * 8, hipMemcpyHostToDevice));
            dc = a.as<double>();
            df = b.as<double>();
        }
        for (int c = 0; c < 3; ++c) launch_prolong_plain(dc + c * nc, df + c * nf, cx, cy, cz, nullptr);
        if (!device_ptr) LP_HIP(hipMemcpy(fine, df, 3 * nf * 8, hipMemcpyDeviceToHost));
        else LP_HIP(hipDeviceSynchronize());
        return LP_OK;
    });
}

int lp_restrict(const double* fine, double* coarse, int cx, int cy, int cz, int device_ptr) {
    return guarded([&]() -> int {
        if (!coarse || !fine || cx < 1 || cy < 1 || cz < 1) return LP_ERR_ARG;
        const size_t nc = (size_t)cx * cy * cz, nf = (size_t)(2 * cx - 1) * (2 * cy - 1) * (2 * cz - 1);
        DBuf a, b;
        const double* df = fine;
        double* dc = coarse;
        if (!device_ptr) {
            a.alloc(3 * nf * 8);
            b.alloc(3 * nc * 8);
            LP_HIP(hipMemcpy(a.p, fine, 3 * nf * 8, hipMemcpyHostToDevice));
            df = a.as<double>();
            dc = b.as<double>();
        }
        for (int c = 0; c < 3; ++c) launch_restrict_plain(df + c * nf, dc + c * nc, cx, cy, cz, nullptr);
        if (!device_ptr) LP_HIP(hipMemcpy(coarse, dc, 3 * nc * 8, hipMemcpyDeviceToHost));
        else LP_HIP(hipDeviceSynchronize());
        return LP_OK;
    });
}

int lp_profile(lp_grid* g, int level, int dtype_bytes, int enable) {
    return guarded([&]() -> int {
        if (!g) return LP_ERR_ARG;
        Grid::Prof& P = g->prof;
        LP_HIP(hipStreamSynchronize(g->stream));
        P.on = enable != 0;
        P.level = level;
        P.eb = dtype_bytes;
        P.used = 0;
        P.dropped = 0;
        P.bytes.clear();
        if (P.on && P.ev.empty()) {
            P.ev.resize(2 * 8192);
            for (hipEvent_t& e : P.ev) LP_HIP(hipEventCreate(&e));
        }
        return LP_OK;
    });
}

int lp_profile_stats(lp_grid* g, long long* launches, double* total_ms, double* total_bytes) {
    return guarded([&]() -> int {
        if (!g || !launches || !total_ms || !total_bytes) return LP_ERR_ARG;
        Grid::Prof& P = g->prof;
        LP_HIP(hipStreamSynchronize(g->stream));
        double ms = 0, by = 0;
        for (size_t i = 0; i < P.used; ++i) {
            float t = 0;
            LP_HIP(hipEventElapsedTime(&t, P.ev[2 * i], P.ev[2 * i + 1]));
            ms += t;
            by += P.bytes[i];
        }
        *launches = (long long)P.used;
        *total_ms = ms;
        *total_bytes = by;
        return P.dropped ? LP_NOT_CONVERGED : LP_OK;
    });
}

int lp_set_solve_tol(lp_grid* g, double rtol, int maxiter) {
    return guarded([&]() -> int {
        if (!g || !(rtol > 0.0) || maxiter < 1) {
            set_error("lp_set_solve_tol: bad argument");
            return LP_ERR_ARG;
        }
        g->opt.rtol = rtol;
        g->opt.maxiter = maxiter;
        return LP_OK;
    });
}

int lp_tune(lp_grid* g, int param, int value) {
    if (!g) return LP_ERR_ARG;
    if (param == 0) {
        if (value != 1 && value != 2) return LP_ERR_ARG;
        const int old = g->tune_fine_ny;
        g->tune_fine_ny = value;
        ++g->epoch;
        return old;
    }
    if (param == 1) {
        if (value < 1 || value > 64) return LP_ERR_ARG;
        const int old = g->tune_fine_xchunk;
        g->tune_fine_xchunk = value;
        ++g->epoch;
        return old;
    }
    if (param == 2) {
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->tune_fused_cheb;
        g->tune_fused_cheb = value;
        ++g->epoch;
        return old;
    }
    if (param == 3) {
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->tune_graph;
        g->tune_graph = value;
        return old;
    }
    if (param == 4) {
        if (value < -1 || value > 1) return LP_ERR_ARG;
        const int old = g->tune_flex;
        g->tune_flex = value;
        return old;
    }
    if (param == 5) {
        if (value < 0 || value > 2) return LP_ERR_ARG;
        const int old = g->tune_fallback;
        g->tune_fallback = value;
        return old;
    }
    if (param == 6) return g->fallbacks;
    if (param == 7) {
        if (value < 0 || value > 2) return LP_ERR_ARG;
        const int old = g->tune_wave;
        g->tune_wave = value;
        ++g->epoch;
        return old;
    }
    if (param == 8) {
        if (value < 0 || value > 64) return LP_ERR_ARG;
        const int old = g->tune_weights_xchunk;
        if (value == 0) return old;
        g->tune_weights_xchunk = value;
        ++g->epoch;
        return old;
    }
    if (param == 9) {
        if (value < 0 || value > 2) return LP_ERR_ARG;
        const int old = g->stall;
        g->stall = value;
        return old;
    }
    if (param == 10) {
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->strict ? 1 : 0;
        g->strict = value != 0;
        return old;
    }
    if (param == 11) return g->strict_switches;
    if (param == 13) {
        if (value != 0 && value != 1) return LP_ERR_ARG;
        if (value && !g->xlo32.p) return LP_ERR_ARG;   // mixed-precision grids only
        const int old = g->tune_cg32;
        g->tune_cg32 = value;
        return old;
    }
    if (param == 14) return (int)std::min<long long>(g->cg32_updates, 0x7fffffff);
    if (param == 15) {
        // reliable-update threshold delta in thousandths (1..999); the sign
        // selects the beta at a replaced step: > 0 Fletcher-Reeves, < 0
        // the flexible one from the FP32 recursion's A p
        if (value == 0 || value > 999 || value < -999) return LP_ERR_ARG;
        const int old = (int)std::lround(g->cg32_delta * 1000.0) * (g->cg32_flex_replaced ? -1 : 1);
        g->cg32_delta = std::abs(value) / 1000.0;
        g->cg32_flex_replaced = value < 0;
        return old;
    }
    if (param == 17) {
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->tune_rhs_inner;
        g->tune_rhs_inner = value;
        ++g->epoch;
        return old;
    }
    if (param == 16) {
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->tune_bf16;
        g->tune_bf16 = value;
        ++g->epoch;
        return old;
    }
    if (param == 12) {
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->tune_l1_explicit;
        g->tune_l1_explicit = value;
        return old;
    }
    if (param == 19) {
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->tune_cg32_restart;
        g->tune_cg32_restart = value;
        return old;
    }
    if (param == 20) {
        // mixed V-cycle: every coarse level (2) or the Galerkin levels (1) as
        // FP16 cell matrices (takes effect at the next hierarchy build)
        if (value < 0 || value > 2) return LP_ERR_ARG;
        const int old = g->tune_half_explicit;
        g->tune_half_explicit = value;
        return old;
    }
    if (param == 21) {
        // FP64 CG: the direction and x updates fused into the next operator
        // launch (1, the default) or a separate pass (0)
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->tune_cg_fused;
        g->tune_cg_fused = value;
        return old;
    }
    if (param == 22) {
        // FP16 cell-matrix levels: LDS-staged node bricks of ty x tz nodes
        // marching x-chunks of planes (value ty * 10000 + tz * 100 +
        // x-chunk; ty * tz <= 32, (ty + 2)(tz + 2) <= 64), or 0 for the
        // node-per-thread gathers
        if (!half_brick_shape_ok(value)) return LP_ERR_ARG;
        const int old = g->tune_half_brick;
        g->tune_half_brick = value;
        ++g->epoch;
        return old;
    }
    if (param == 23) {
        // the brick form only on levels of at least this many bricks
        if (value < 0) return LP_ERR_ARG;
        const int old = g->tune_half_brick_min;
        g->tune_half_brick_min = value;
        ++g->epoch;
        return old;
    }
    if (param == 24) {
        // fused Chebyshev smoothers of degree 2 from a nonzero x: the
        // residual step stores r only and the last step stages d from it
        // (1, the default), or the residual step stores r and d (0)
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->tune_cheb_r;
        g->tune_cheb_r = value;
        ++g->epoch;
        return old;
    }
    if (param == 18) {
        if (value != 0 && value != 1) return LP_ERR_ARG;
        const int old = g->tune_l1_explicit32;
        g->tune_l1_explicit32 = value;
        return old;
    }
    return LP_ERR_ARG;
}

int lp_autotune(lp_grid* g, int nrhs, int reps, int* xchunk, double* ms_best) {
    return guarded([&]() -> int {
        if (!g || !g->built || nrhs < 1 || nrhs > g->max_rhs || reps < 1) return LP_ERR_ARG;
        Grid& G = *g;
        const int eb = G.vbytes();
        const int cand[] = {4, 8, 12, 16, 24, 32};
        int best = G.tune_fine_xchunk;
        double tbest = 1e300;
        for (int c : cand) {
            if (c > G.g0.Nx) continue;
            G.tune_fine_xchunk = c;
            double ms = 0.0;
            const int rc = lp_time_apply(g, 0, eb, nrhs, reps, &ms);
            if (rc != LP_OK) return rc;
            if (ms < tbest) {
                tbest = ms;
                best = c;
            }
        }
        G.tune_fine_xchunk = best;
        ++G.epoch;
        if (xchunk) *xchunk = best;
        if (ms_best) *ms_best = tbest;
        // the weights level's FP16-MFMA kernel: node planes per block (a
        // coarse grid wants short chunks, or it runs a handful of blocks)
        if (G.nlevels > 1 && weights_mfma_path(G, 1, eb)) {
            const int wcand[] = {2, 4, 8, 16};
            int wbest = G.tune_weights_xchunk;
            double wt = 1e300;
            for (int c : wcand) {
                if (c > G.lv[1].g.Nx) continue;
                G.tune_weights_xchunk = c;
                ++G.epoch;
                double ms = 0.0;
                const int rc = lp_time_apply(g, 1, eb, nrhs, reps, &ms);
                if (rc != LP_OK) return rc;
                if (ms < wt) {
                    wt = ms;
                    wbest = c;
                }
            }
            G.tune_weights_xchunk = wbest;
            ++G.epoch;
        }
        return LP_OK;
    });
}

int lp_time_apply(lp_grid* g, int level, int dtype_bytes, int nrhs, int reps, double* ms_per_apply) {
    return guarded([&]() -> int {
        if (!g || !g->built || level < 0 || level >= g->nlevels || reps < 1 || !ms_per_apply) return LP_ERR_ARG;
        if (nrhs < 1 || nrhs > g->max_rhs) return LP_ERR_ARG;
        Grid& G = *g;
        Level& L = G.lv[level];
        const size_t n = 3 * L.g.nn;
        DBuf u, o;
        u.alloc(n * nrhs * dtype_bytes);
        o.alloc(n * nrhs * dtype_bytes);
        launch_fill(u.p, dtype_bytes, n * nrhs, 1.0, G.stream);
        EpiArgs E;
        E.kind = EPI_STORE;
        E.out = o.p;
        launch_apply(G, level, dtype_bytes, u.p, n, nrhs, E, G.stream);
        hipEvent_t a, b;
        LP_HIP(hipEventCreate(&a));
        LP_HIP(hipEventCreate(&b));
        LP_HIP(hipEventRecord(a, G.stream));
        for (int k = 0; k < reps; ++k) launch_apply(G, level, dtype_bytes, u.p, n, nrhs, E, G.stream);
        LP_HIP(hipEventRecord(b, G.stream));
        LP_HIP(hipEventSynchronize(b));
        float ms = 0;
        LP_HIP(hipEventElapsedTime(&ms, a, b));
        hipEventDestroy(a);
        hipEventDestroy(b);
        *ms_per_apply = ms / reps;
        return LP_OK;
    });
}

}  // extern "C"
