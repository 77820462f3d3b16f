// The SIMP loop on the device: moduli, sensitivities, density filter and the
// OC update with its volume bisection, around the batched MG-PCG solve.
//
// Reference: the top3d loop benchmarks/reference.py:214-257 (SIMP p, Emin,
// density filter H/Hs, OC with bisection l2 = 1e9 / rel. tol 1e-3, move);
// the cone filter simp.GridFilter (simp.py:1143-1189) and top3d's _filter
// (reference.py:188-211); element energy femops.element_energy
// (femops.py:236-242) and the compliance sensitivity (reference.py:230-233,
// simp.py:1897-1904). Load cases are summed (sum of compliances).
//
// Every step stays on the device. The bisection's l1 / l2 live in device
// memory: each evaluation launches an OC+filter pass and a finalize that
// updates the bracket and raises a done flag; later evaluations exit at
// once, so the host queues a fixed number and syncs once per iteration.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "lp_internal.h"

namespace lp {
void pcg(Grid& G, int nrhs, bool warm, int* iters, int* converged);
}

using namespace lp;

struct lp_grid : public Grid {};

namespace {

struct Tap {
    int dx, dy, dz, pad;
    double w;
};

struct Bisect {
    double l1, l2, lmid, target;
    int done, evals, pad0, pad1;
};

inline unsigned nblk(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ double bsum(double v, double* sred) {
    v = wsum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sred[wv] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sred[i];
    __syncthreads();
    return s;
}

// E = emin + xphys^p (1 - emin) (reference.py:229)
__global__ void k_moduli(const double* __restrict__ xphys, double* e, size_t n, double penal, double emin) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    e[i] = emin + pow(xphys[i], penal) * (1.0 - emin);
}

// b = loads * mask
__global__ void k_mask_loads(const double* __restrict__ f, const uint8_t* __restrict__ mask, double* b, size_t nn,
                             size_t stride) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= 3 * nn) return;
    const int c = (int)(i / nn);
    const double v = f[rhs * stride + i];
    b[rhs * stride + i] = ((mask[i % nn] >> c) & 1) ? v : 0.0;
}

// ce = sum_rhs u_c^T Ke u_c; partial of sum E ce; dc = -p (1-emin) x^(p-1) ce
__global__ void k_sens(const double* __restrict__ u, const double* __restrict__ ke, Geom g, int nrhs, size_t vstride,
                       const double* __restrict__ xphys, const double* __restrict__ e, double penal, double emin,
                       double* ce_out, double* dc, double* partials) {
    __shared__ double sred[4];
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0;
    if (cell < g.ncell) {
        const unsigned ci = (unsigned)cell, nyz = (unsigned)g.ny * (unsigned)g.nz;
        const int X = (int)(ci / nyz), Y = (int)((ci - (unsigned)X * nyz) / (unsigned)g.nz);
        const int Z = (int)(ci - (unsigned)X * nyz - (unsigned)Y * (unsigned)g.nz);
        double ce = 0;
        for (int r = 0; r < nrhs; ++r) {
            const double* uu = u + r * vstride;
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
            ce += acc;
        }
        ce_out[cell] = ce;
        const double xp = xphys[cell];
        dc[cell] = -penal * (1.0 - emin) * pow(xp, penal - 1.0) * ce;
        part = e[cell] * ce;
    }
    const double s = bsum(part, sred);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// out[c] = post[c] * sum_t w_t v[c+d_t] pre[c+d_t]  (zero outside the grid)
__global__ void k_filter(const double* __restrict__ v, const double* __restrict__ pre, const double* __restrict__ post,
                         const uint8_t* __restrict__ passive, const Tap* __restrict__ taps, int ntaps, int nx, int ny,
                         int nz, double* out, double* partials, const Bisect* bis) {
    __shared__ double sred[4];
    if (bis && bis->done) return;
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)nx * ny * nz;
    double part = 0;
    if (cell < n) {
        const unsigned ci = (unsigned)cell, nyz = (unsigned)ny * (unsigned)nz;
        const int x = (int)(ci / nyz), y = (int)((ci - (unsigned)x * nyz) / (unsigned)nz);
        const int z = (int)(ci - (unsigned)x * nyz - (unsigned)y * (unsigned)nz);
        double acc = 0;
        for (int t = 0; t < ntaps; ++t) {
            const Tap tp = taps[t];
            const int X = x + tp.dx, Y = y + tp.dy, Z = z + tp.dz;
            if (X < 0 || Y < 0 || Z < 0 || X >= nx || Y >= ny || Z >= nz) continue;
            const size_t k = ((size_t)X * ny + Y) * nz + Z;
            acc += tp.w * (pre ? v[k] * pre[k] : v[k]);
        }
        if (post) acc *= post[cell];
        if (passive && passive[cell]) acc = 0.0;
        out[cell] = acc;
        part = acc;
    }
    if (partials) {
        const double s = bsum(part, sred);
        if (threadIdx.x == 0) partials[blockIdx.x] = s;
    }
}

// Heaviside projection and its derivative (simp.py:1443-1450)
__global__ void k_project(const double* __restrict__ xt, size_t n, double beta, double eta, double* out,
                          double* dout) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double den = tanh(beta * eta) + tanh(beta * (1.0 - eta));
    const double t = tanh(beta * (xt[i] - eta));
    if (out) out[i] = (tanh(beta * eta) + t) / den;
    if (dout) dout[i] = beta * (1.0 - t * t) / den;
}

// MMA inputs: df0 = s dc, dcon = t dv, and block partials of sum xphys
__global__ void k_mma_inputs(const double* __restrict__ dc, const double* __restrict__ dv,
                             const double* __restrict__ xphys, size_t n, double s, double t, double* df0,
                             double* dcon, double* partials) {
    __shared__ double sred[4];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        df0[i] = s * dc[i];
        dcon[i] = t * dv[i];
        part = xphys[i];
    }
    const double r = bsum(part, sred);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// xnew = max(0, max(x - move, min(1, min(x + move, x sqrt(-dc/dv/lmid)))))
__device__ __forceinline__ double oc_value(const double* __restrict__ x, const double* __restrict__ dc,
                                          const double* __restrict__ dv, const uint8_t* __restrict__ passive,
                                          size_t i, double move, double lmid) {
    const double xi = x[i];
    double v = xi * sqrt(-dc[i] / dv[i] / lmid);
    v = fmin(xi + move, v);
    v = fmin(1.0, v);
    v = fmax(xi - move, v);
    v = fmax(0.0, v);
    if (passive && passive[i]) v = 0.0;
    return v;
}

// the design of the bisection's last evaluated multiplier
__global__ void k_oc(const double* __restrict__ x, const double* __restrict__ dc, const double* __restrict__ dv,
                     const uint8_t* __restrict__ passive, size_t n, double move, const Bisect* bis, double* xnew) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    xnew[i] = oc_value(x, dc, dv, passive, i, move, bis->lmid);
}

// one bisection evaluation at lmid = (l1 + l2) / 2: block partials of the
// physical volume sum xphys = sum_c xnew_c W_c, W = H (1/hs) over the
// non-passive cells (H is symmetric), so no filter pass is needed per
// evaluation; nothing is stored
__global__ void k_oc_volume(const double* __restrict__ x, const double* __restrict__ dc,
                            const double* __restrict__ dv, const uint8_t* __restrict__ passive,
                            const double* __restrict__ w, size_t n, double move, const Bisect* bis,
                            double* partials) {
    __shared__ double sred[4];
    if (bis->done) return;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double lmid = 0.5 * (bis->l2 + bis->l1);
    const double part = i < n ? oc_value(x, dc, dv, passive, i, move, lmid) * w[i] : 0.0;
    const double r = bsum(part, sred);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// t = 1/hs on the non-passive cells, 0 on the passive ones
__global__ void k_mask_passive(const double* __restrict__ hs_inv, const uint8_t* __restrict__ passive, size_t n,
                               double* t) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) t[i] = passive[i] ? 0.0 : hs_inv[i];
}

__global__ void k_bisect_finalize(const double* __restrict__ partials, int nb, Bisect* bis, double tol) {
    __shared__ double sred[4];
    if (bis->done) return;
    double acc = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partials[i];
    const double s = bsum(acc, sred);
    if (threadIdx.x != 0) return;
    const double lmid = 0.5 * (bis->l2 + bis->l1);
    if (s > bis->target) bis->l1 = lmid;
    else bis->l2 = lmid;
    bis->lmid = lmid;
    bis->evals += 1;
    if (!((bis->l2 - bis->l1) / (bis->l1 + bis->l2) > tol)) bis->done = 1;
}

__global__ void k_reduce_sum(const double* __restrict__ partials, int nb, double* out) {
    __shared__ double sred[4];
    double acc = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partials[i];
    const double s = bsum(acc, sred);
    if (threadIdx.x == 0) *out = s;
}

__global__ void k_change(const double* __restrict__ a, const double* __restrict__ b, size_t n, unsigned long long* out) {
    double m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, fabs(a[i] - b[i]));
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

__global__ void k_inv(const double* __restrict__ a, double* b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = 1.0 / a[i];
}

std::vector<Tap> make_taps(double radius, int reach) {
    std::vector<Tap> taps;
    for (int dx = -reach; dx <= reach; ++dx)
        for (int dy = -reach; dy <= reach; ++dy)
            for (int dz = -reach; dz <= reach; ++dz) {
                const double w = radius - std::sqrt((double)(dx * dx + dy * dy + dz * dz));
                if (w > 0) taps.push_back(Tap{dx, dy, dz, 0, w});
            }
    return taps;
}

template <typename F> int guarded_s(F&& f) {
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

struct lp_simp {
    lp_grid* g = nullptr;
    int nrhs = 1;
    double volfrac = 0.3, rmin = 1.5, penal = 3.0, emin = 1e-9, move = 0.2;
    int filter_kind = LP_FILTER_TOP3D;
    bool warm = true;
    bool have_u = false;
    size_t ncell = 0;
    int ntaps = 0;
    double target = 0;
    DBuf taps, x, xnew, xphys, dc, dv, ce, hs_inv, ones, passive, loads, bis, partials, scalar, chg;
    DBuf dv0;    // H (1/hs): the filtered volume sensitivity (constant)
    DBuf wvol;   // H (1/hs on the non-passive cells): weights of the volume sum
    bool has_passive = false;
    // optimizer: 0 optimality criteria with the volume bisection, 1 MMA
    int optimizer = 0;
    lp_mma* mma = nullptr;
    double c_ref = 0.0;
    DBuf df0, dcon;
    ~lp_simp() {
        if (mma) lp_mma_destroy(mma);
    }
};

extern "C" {

int lp_simp_create(lp_simp** out, lp_grid* g, int nrhs, const double* loads, const uint8_t* active, double volfrac,
                   double rmin, int filter_kind, double penal, double emin, double move, int warm_start) {
    return guarded_s([&]() -> int {
        if (!out || !g || nrhs < 1 || nrhs > g->max_rhs || !loads) return LP_ERR_ARG;
        lp_simp* S = new lp_simp();
        try {
            S->g = g;
            S->nrhs = nrhs;
            S->volfrac = volfrac;
            S->rmin = rmin;
            S->filter_kind = filter_kind;
            S->penal = penal;
            S->emin = emin;
            S->move = move;
            S->warm = warm_start != 0;
            const Geom& g0 = g->g0;
            const size_t n = g0.ncell;
            S->ncell = n;
            hipStream_t s = g->stream;
            const int reach = filter_kind == LP_FILTER_TOP3D ? (int)std::ceil(rmin) - 1 : (int)std::floor(rmin);
            std::vector<Tap> taps = make_taps(rmin, std::max(reach, 0));
            // a unit axis carries no offsets along it (reference.py:196)
            std::vector<Tap> kept;
            for (const Tap& t : taps)
                if (!((t.dx && g0.nx == 1) || (t.dy && g0.ny == 1) || (t.dz && g0.nz == 1))) kept.push_back(t);
            S->ntaps = (int)kept.size();
            S->taps.alloc(kept.size() * sizeof(Tap));
            LP_HIP(hipMemcpy(S->taps.p, kept.data(), kept.size() * sizeof(Tap), hipMemcpyHostToDevice));
            for (DBuf* b : {&S->x, &S->xnew, &S->xphys, &S->dc, &S->dv, &S->ce, &S->hs_inv, &S->ones}) b->alloc(n * 8);
            S->passive.alloc(n);
            std::vector<uint8_t> pas(n, 0);
            size_t n_active = n;
            if (active) {
                n_active = 0;
                for (size_t i = 0; i < n; ++i) {
                    pas[i] = active[i] ? 0 : 1;
                    n_active += active[i] ? 1 : 0;
                }
                S->has_passive = n_active != n;
            }
            LP_HIP(hipMemcpy(S->passive.p, pas.data(), n, hipMemcpyHostToDevice));
            std::vector<double> x0(n);
            for (size_t i = 0; i < n; ++i) x0[i] = pas[i] ? 0.0 : volfrac;
            LP_HIP(hipMemcpy(S->x.p, x0.data(), n * 8, hipMemcpyHostToDevice));
            LP_HIP(hipMemcpy(S->xphys.p, x0.data(), n * 8, hipMemcpyHostToDevice));
            launch_fill(S->ones.p, 8, n, 1.0, s);
            // hs: row sums (top3d) or the full kernel weight (padded)
            if (filter_kind == LP_FILTER_TOP3D) {
                hipLaunchKernelGGL(k_filter, dim3(nblk(n)), dim3(kBlock), 0, s, S->ones.as<double>(), (const double*)nullptr,
                                   (const double*)nullptr, (const uint8_t*)nullptr, S->taps.as<Tap>(), S->ntaps, g0.nx,
                                   g0.ny, g0.nz, S->hs_inv.as<double>(), (double*)nullptr, (const Bisect*)nullptr);
                LP_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_inv, dim3(nblk(n)), dim3(kBlock), 0, s, S->hs_inv.as<double>(), S->hs_inv.as<double>(), n);
                LP_HIP(hipGetLastError());
            } else {
                double full = 0;
                for (const Tap& t : kept) full += t.w;
                launch_fill(S->hs_inv.p, 8, n, 1.0 / full, s);
            }
            // constant filtered fields: dv = H (1/hs) and the volume weights
            S->dv0.alloc(n * 8);
            S->wvol.alloc(n * 8);
            hipLaunchKernelGGL(k_filter, dim3(nblk(n)), dim3(kBlock), 0, s, S->ones.as<double>(), S->hs_inv.as<double>(),
                               (const double*)nullptr, (const uint8_t*)nullptr, S->taps.as<Tap>(), S->ntaps, g0.nx, g0.ny,
                               g0.nz, S->dv0.as<double>(), (double*)nullptr, (const Bisect*)nullptr);
            LP_HIP(hipGetLastError());
            if (S->has_passive) {
                hipLaunchKernelGGL(k_mask_passive, dim3(nblk(n)), dim3(kBlock), 0, s, S->hs_inv.as<double>(),
                                   S->passive.as<uint8_t>(), n, S->xnew.as<double>());
                LP_HIP(hipGetLastError());
                hipLaunchKernelGGL(k_filter, dim3(nblk(n)), dim3(kBlock), 0, s, S->xnew.as<double>(),
                                   (const double*)nullptr, (const double*)nullptr, (const uint8_t*)nullptr,
                                   S->taps.as<Tap>(), S->ntaps, g0.nx, g0.ny, g0.nz, S->wvol.as<double>(),
                                   (double*)nullptr, (const Bisect*)nullptr);
                LP_HIP(hipGetLastError());
            } else {
                LP_HIP(hipMemcpyAsync(S->wvol.p, S->dv0.p, n * 8, hipMemcpyDeviceToDevice, s));
            }
            const size_t vn = 3 * g0.nn;
            S->loads.alloc(vn * nrhs * 8);
            LP_HIP(hipMemcpy(S->loads.p, loads, vn * nrhs * 8, hipMemcpyHostToDevice));
            S->bis.alloc(sizeof(Bisect));
            S->partials.alloc(nblk(n) * 8);
            S->scalar.alloc(64);
            S->chg.alloc(8);
            // the volume target counts every cell of the box (top3d: volfrac * n)
            S->target = volfrac * (double)n;
            LP_HIP(hipStreamSynchronize(s));
        } catch (...) {
            delete S;
            throw;
        }
        *out = S;
        return LP_OK;
    });
}

void lp_simp_destroy(lp_simp* s) { delete s; }

int lp_simp_iterate(lp_simp* S, int niter, lp_simp_stats* stats) {
    return guarded_s([&]() -> int {
        if (!S || niter < 0 || (niter > 0 && !stats)) return LP_ERR_ARG;
        Grid& G = *S->g;
        hipStream_t s = G.stream;
        const size_t n = S->ncell;
        const Geom& g0 = G.g0;
        const size_t vn = 3 * g0.nn;
        const unsigned nb = nblk(n);
        for (int it = 0; it < niter; ++it) {
            const auto t0 = std::chrono::steady_clock::now();
            // moduli and hierarchy
            hipLaunchKernelGGL(k_moduli, dim3(nb), dim3(kBlock), 0, s, S->xphys.as<double>(), G.lv[0].e64.as<double>(), n,
                               S->penal, S->emin);
            LP_HIP(hipGetLastError());
            int rc = build_hierarchy(G);
            if (rc == LP_SINGULAR && !G.strict) {
                G.strict = true;
                rc = build_hierarchy(G);
            }
            if (rc != LP_OK) return rc;
            G.built = true;
            if (G.opt.smoother == LP_SMOOTH_CHEBYSHEV)
                for (int i = 0; i < G.nlevels; ++i) lp_estimate_lmax(S->g, i, nullptr);
            // solve all load cases together
            hipLaunchKernelGGL(k_mask_loads, dim3(nblk(vn), S->nrhs), dim3(kBlock), 0, s, S->loads.as<double>(),
                               G.lv[0].mask.as<uint8_t>(), G.b.as<double>(), g0.nn, vn);
            LP_HIP(hipGetLastError());
            std::vector<int> iters(S->nrhs), ok(S->nrhs);
            const bool warm = S->warm && S->have_u;
            pcg(G, S->nrhs, warm, iters.data(), ok.data());
            bool all = true;
            for (int k = 0; k < S->nrhs; ++k) all &= ok[k] != 0;
            if (!all) {
                set_error("linear solve did not converge; the structure may be a mechanism (check supports)");
                return LP_NOT_CONVERGED;
            }
            S->have_u = true;
            // sensitivities and compliance
            hipLaunchKernelGGL(k_sens, dim3(nb), dim3(kBlock), 0, s, G.x.as<double>(), G.ke64.as<double>(), g0, S->nrhs,
                               vn, S->xphys.as<double>(), G.lv[0].e64.as<double>(), S->penal, S->emin,
                               S->ce.as<double>(), S->dc.as<double>(), S->partials.as<double>());
            LP_HIP(hipGetLastError());
            double* scal = S->scalar.as<double>();
            hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kBlock), 0, s, S->partials.as<double>(), (int)nb, scal + 0);
            LP_HIP(hipGetLastError());
            // filter: dc = H(dc / hs); dv = H(1 / hs)
            hipLaunchKernelGGL(k_filter, dim3(nb), dim3(kBlock), 0, s, S->dc.as<double>(), S->hs_inv.as<double>(),
                               (const double*)nullptr, (const uint8_t*)nullptr, S->taps.as<Tap>(), S->ntaps, g0.nx, g0.ny,
                               g0.nz, S->xnew.as<double>(), (double*)nullptr, (const Bisect*)nullptr);
            LP_HIP(hipGetLastError());
            std::swap(S->dc.p, S->xnew.p);
            LP_HIP(hipMemcpyAsync(S->dv.p, S->dv0.p, n * 8, hipMemcpyDeviceToDevice, s));
            if (S->optimizer == 1) {
                // MMA on f0 = 10 c / c_ref with the volume constraint
                // sum xphys / (vf n) - 1 <= 0, both gradients through the
                // filter (the normalization of loadpath's MMA loops on
                // compliance-type objectives, thermal.py:1035-1043)
                double c = 0.0;
                LP_HIP(hipMemcpyAsync(&c, scal, 8, hipMemcpyDeviceToHost, s));
                LP_HIP(hipStreamSynchronize(s));
                if (S->c_ref == 0.0) S->c_ref = std::fabs(c) > 0 ? std::fabs(c) : 1.0;
                hipLaunchKernelGGL(k_mma_inputs, dim3(nb), dim3(kBlock), 0, s, S->dc.as<double>(), S->dv.as<double>(),
                                   S->xphys.as<double>(), n, 10.0 / S->c_ref, 1.0 / (S->volfrac * (double)n),
                                   S->df0.as<double>(), S->dcon.as<double>(), S->partials.as<double>());
                hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kBlock), 0, s, S->partials.as<double>(), (int)nb, scal + 1);
                LP_HIP(hipGetLastError());
                double vol = 0.0;
                LP_HIP(hipMemcpyAsync(&vol, scal + 1, 8, hipMemcpyDeviceToHost, s));
                LP_HIP(hipStreamSynchronize(s));
                const double con = vol / (S->volfrac * (double)n) - 1.0;
                const int rc = lp_mma_update(S->mma, S->x.as<double>(), 10.0 * c / S->c_ref, S->df0.as<double>(), &con,
                                             S->dcon.as<double>(), S->xnew.as<double>(), 1);
                if (rc != LP_OK) return rc;
                hipLaunchKernelGGL(k_filter, dim3(nb), dim3(kBlock), 0, s, S->xnew.as<double>(), (const double*)nullptr,
                                   S->hs_inv.as<double>(), (const uint8_t*)nullptr, S->taps.as<Tap>(), S->ntaps, g0.nx,
                                   g0.ny, g0.nz, S->xphys.as<double>(), (double*)nullptr, (const Bisect*)nullptr);
                LP_HIP(hipGetLastError());
                LP_HIP(hipMemsetAsync(S->chg.p, 0, 8, s));
                hipLaunchKernelGGL(k_change, dim3(std::min(nb, 1024u)), dim3(kBlock), 0, s, S->xnew.as<double>(),
                                   S->x.as<double>(), n, S->chg.as<unsigned long long>());
                LP_HIP(hipGetLastError());
                std::swap(S->x.p, S->xnew.p);
                double chg = 0.0;
                LP_HIP(hipMemcpyAsync(&chg, S->chg.p, 8, hipMemcpyDeviceToHost, s));
                LP_HIP(hipStreamSynchronize(s));
                lp_simp_stats& st = stats[it];
                st.compliance = c;
                st.change = chg;
                st.volume = vol / (double)n;
                st.cg_iters = *std::max_element(iters.begin(), iters.end());
                st.cg_iters_total = 0;
                for (int k = 0; k < S->nrhs; ++k) st.cg_iters_total += iters[k];
                st.oc_evals = 0;
                st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                continue;
            }
            // bisection on the device
            Bisect b0{0.0, 1e9, 0.0, S->target, 0, 0, 0, 0};
            LP_HIP(hipMemcpyAsync(S->bis.p, &b0, sizeof(b0), hipMemcpyHostToDevice, s));
            const uint8_t* pas = S->has_passive ? S->passive.as<uint8_t>() : nullptr;
            // evaluations queued in rounds (they exit once the bracket has
            // closed; the bracket 1e9 at relative 1e-3 takes ~50)
            Bisect bq;
            for (int e = 0; e < 100;) {
                const int round = e == 0 ? 56 : 16;
                for (int k = 0; k < round && e < 100; ++k, ++e) {
                    hipLaunchKernelGGL(k_oc_volume, dim3(nb), dim3(kBlock), 0, s, S->x.as<double>(),
                                       S->dc.as<double>(), S->dv.as<double>(), pas, S->wvol.as<double>(), n, S->move,
                                       S->bis.as<Bisect>(), S->partials.as<double>());
                    hipLaunchKernelGGL(k_bisect_finalize, dim3(1), dim3(kBlock), 0, s, S->partials.as<double>(),
                                       (int)nb, S->bis.as<Bisect>(), 1e-3);
                }
                LP_HIP(hipGetLastError());
                LP_HIP(hipMemcpyAsync(&bq, S->bis.p, sizeof(bq), hipMemcpyDeviceToHost, s));
                LP_HIP(hipStreamSynchronize(s));
                if (bq.done) break;
            }
            // the design and its physical densities at the last evaluated multiplier
            hipLaunchKernelGGL(k_oc, dim3(nb), dim3(kBlock), 0, s, S->x.as<double>(), S->dc.as<double>(),
                               S->dv.as<double>(), pas, n, S->move, S->bis.as<Bisect>(), S->xnew.as<double>());
            hipLaunchKernelGGL(k_filter, dim3(nb), dim3(kBlock), 0, s, S->xnew.as<double>(), (const double*)nullptr,
                               S->hs_inv.as<double>(), pas, S->taps.as<Tap>(), S->ntaps, g0.nx, g0.ny, g0.nz,
                               S->xphys.as<double>(), (double*)nullptr, (const Bisect*)nullptr);
            LP_HIP(hipGetLastError());
            // change, volume, x = xnew
            LP_HIP(hipMemsetAsync(S->chg.p, 0, 8, s));
            hipLaunchKernelGGL(k_change, dim3(std::min(nb, 1024u)), dim3(kBlock), 0, s, S->xnew.as<double>(),
                               S->x.as<double>(), n, S->chg.as<unsigned long long>());
            LP_HIP(hipGetLastError());
            std::swap(S->x.p, S->xnew.p);
            double h[2];
            Bisect bh;
            LP_HIP(hipMemcpyAsync(h, scal, 8, hipMemcpyDeviceToHost, s));
            LP_HIP(hipMemcpyAsync(h + 1, S->chg.p, 8, hipMemcpyDeviceToHost, s));
            LP_HIP(hipMemcpyAsync(&bh, S->bis.p, sizeof(bh), hipMemcpyDeviceToHost, s));
            LP_HIP(hipStreamSynchronize(s));
            lp_simp_stats& st = stats[it];
            st.compliance = h[0];
            st.change = h[1];
            st.volume = 0;
            st.cg_iters = *std::max_element(iters.begin(), iters.end());
            st.cg_iters_total = 0;
            for (int k = 0; k < S->nrhs; ++k) st.cg_iters_total += iters[k];
            st.oc_evals = bh.evals;
            st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (!bh.done) {
                set_error("OC bisection did not close within 100 evaluations");
                return LP_ERR_STATE;
            }
        }
        return LP_OK;
    });
}

int lp_project(long long n, const double* xt, double beta, double eta, double* out, double* dout, int device_ptr) {
    return guarded_s([&]() -> int {
        if (n < 0 || !xt || (!out && !dout) || !(beta > 0)) return LP_ERR_ARG;
        if (n == 0) return LP_OK;
        const size_t nb = (size_t)n * 8;
        DBuf a, b, c;
        const double* dx = xt;
        double* o = out;
        double* d = dout;
        if (!device_ptr) {
            a.alloc(nb);
            LP_HIP(hipMemcpy(a.p, xt, nb, hipMemcpyHostToDevice));
            dx = a.as<double>();
            if (out) { b.alloc(nb); o = b.as<double>(); }
            if (dout) { c.alloc(nb); d = c.as<double>(); }
        }
        hipLaunchKernelGGL(k_project, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, 0, dx, (size_t)n, beta,
                           eta, o, d);
        LP_HIP(hipGetLastError());
        if (!device_ptr) {
            if (out) LP_HIP(hipMemcpy(out, o, nb, hipMemcpyDeviceToHost));
            if (dout) LP_HIP(hipMemcpy(dout, d, nb, hipMemcpyDeviceToHost));
        }
        LP_HIP(hipDeviceSynchronize());
        return LP_OK;
    });
}

int lp_simp_set_optimizer(lp_simp* S, int optimizer) {
    return guarded_s([&]() -> int {
        if (!S || (optimizer != 0 && optimizer != 1)) return LP_ERR_ARG;
        if (optimizer == 1 && S->has_passive) {
            set_error("lp_simp_set_optimizer: MMA runs on boxes without passive cells");
            return LP_ERR_ARG;
        }
        S->optimizer = optimizer;
        if (optimizer == 1 && !S->mma) {
            const size_t n = S->ncell;
            std::vector<double> lo(n, 0.0), hi(n, 1.0);
            const int rc = lp_mma_create(&S->mma, (long long)n, 1, lo.data(), hi.data(), S->move, 1000.0, 0.5, 1.2, 0.7);
            if (rc != LP_OK) return rc;
            S->df0.alloc(n * 8);
            S->dcon.alloc(n * 8);
            S->c_ref = 0.0;
        }
        return LP_OK;
    });
}

int lp_simp_get(lp_simp* S, int what, double* out) {
    return guarded_s([&]() -> int {
        if (!S || !out) return LP_ERR_ARG;
        DBuf* src = what == 0 ? &S->x : what == 1 ? &S->xphys : what == 2 ? &S->dc : what == 3 ? &S->ce : nullptr;
        if (!src) return LP_ERR_ARG;
        LP_HIP(hipMemcpyAsync(out, src->p, S->ncell * 8, hipMemcpyDeviceToHost, S->g->stream));
        LP_HIP(hipStreamSynchronize(S->g->stream));
        return LP_OK;
    });
}

int lp_simp_set_x(lp_simp* S, const double* x) {
    return guarded_s([&]() -> int {
        if (!S || !x) return LP_ERR_ARG;
        LP_HIP(hipMemcpyAsync(S->x.p, x, S->ncell * 8, hipMemcpyHostToDevice, S->g->stream));
        LP_HIP(hipMemcpyAsync(S->xphys.p, x, S->ncell * 8, hipMemcpyHostToDevice, S->g->stream));
        LP_HIP(hipStreamSynchronize(S->g->stream));
        return LP_OK;
    });
}

int lp_simp_displacements(lp_simp* S, double* out) {
    return guarded_s([&]() -> int {
        if (!S || !out) return LP_ERR_ARG;
        const size_t bytes = 3 * S->g->g0.nn * S->nrhs * 8;
        LP_HIP(hipMemcpyAsync(out, S->g->x.p, bytes, hipMemcpyDeviceToHost, S->g->stream));
        LP_HIP(hipStreamSynchronize(S->g->stream));
        return LP_OK;
    });
}

int lp_filter_apply(int nx, int ny, int nz, double radius, int reach, const double* v, double* out, int device_ptr) {
    return guarded_s([&]() -> int {
        if (nx < 1 || ny < 1 || nz < 1 || !v || !out || reach < 0) return LP_ERR_ARG;
        std::vector<Tap> taps = make_taps(radius, reach);
        const size_t n = (size_t)nx * ny * nz;
        DBuf t, a, b;
        t.alloc(std::max<size_t>(1, taps.size()) * sizeof(Tap));
        if (!taps.empty()) LP_HIP(hipMemcpy(t.p, taps.data(), taps.size() * sizeof(Tap), hipMemcpyHostToDevice));
        const double* dv = v;
        double* dout = out;
        if (!device_ptr) {
            a.alloc(n * 8);
            b.alloc(n * 8);
            LP_HIP(hipMemcpy(a.p, v, n * 8, hipMemcpyHostToDevice));
            dv = a.as<double>();
            dout = b.as<double>();
        }
        hipLaunchKernelGGL(k_filter, dim3(nblk(n)), dim3(kBlock), 0, nullptr, dv, (const double*)nullptr,
                           (const double*)nullptr, (const uint8_t*)nullptr, t.as<Tap>(), (int)taps.size(), nx, ny, nz,
                           dout, (double*)nullptr, (const Bisect*)nullptr);
        LP_HIP(hipGetLastError());
        if (!device_ptr) LP_HIP(hipMemcpy(out, dout, n * 8, hipMemcpyDeviceToHost));
        else LP_HIP(hipDeviceSynchronize());
        return LP_OK;
    });
}

}  // extern "C"
