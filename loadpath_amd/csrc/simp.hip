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

// e = emin + x^p (1 - emin) and de = de/dx = p x^(p-1) (1 - emin)
__global__ void k_moduli_de(const double* __restrict__ x, double* e, double* de, size_t n, double penal,
                            double emin) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i];
    e[i] = emin + pow(xi, penal) * (1.0 - emin);
    if (de) de[i] = penal * pow(xi, penal - 1.0) * (1.0 - emin);
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

// ---- the production loop's ties and live-cell reduction -----------------

constexpr double kAliveMin = 1e-3;      // simp.ALIVE (simp.py:59)
constexpr double kReduceBelow = 0.8;    // simp.REDUCE_BELOW (simp.py:61)

__global__ void k_fill_u8(uint8_t* v, size_t n, uint8_t x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = x;
}

// mirror plane through the box centre across axis a (domain.mirror,
// simp._symmetrize, simp.py:1476-1482): a cell and its image both take
// 0.5 (v + v_image); one thread per pair, the centre plane of an odd axis
// unchanged (0.5 (v + v) = v)
__global__ void k_tie_mirror(double* v, int nx, int ny, int nz, int axis) {
    const int d[3] = {nx, ny, nz};
    const size_t n = (size_t)nx * ny * nz;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c[3];
    c[2] = (int)(i % nz);
    c[1] = (int)((i / nz) % ny);
    c[0] = (int)(i / ((size_t)nz * ny));
    const int m = d[axis] - 1 - c[axis];
    if (m <= c[axis]) return;               // the lower cell of each pair acts
    int cm[3] = {c[0], c[1], c[2]};
    cm[axis] = m;
    const size_t j = ((size_t)cm[0] * ny + cm[1]) * nz + cm[2];
    const double t = 0.5 * (v[i] + v[j]);
    v[i] = t;
    v[j] = t;
}

// extrusion columns along axis a (manufacture.tie_groups, simp._extrude,
// simp.py:1453-1463): every design cell of a column takes the column's mean
// over its design cells, summed in increasing order along the axis; one
// thread per column
__global__ void k_tie_extrude(double* v, const uint8_t* __restrict__ pres, int nx, int ny, int nz, int axis) {
    const int d[3] = {nx, ny, nz};
    const int o0 = axis == 0 ? 1 : 0, o1 = axis == 2 ? 1 : 2;
    const size_t ncol = (size_t)d[o0] * d[o1];
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol) return;
    int c[3];
    c[o0] = (int)(t / d[o1]);
    c[o1] = (int)(t % d[o1]);
    double sum = 0.0;
    int cnt = 0;
    for (int k = 0; k < d[axis]; ++k) {
        c[axis] = k;
        const size_t i = ((size_t)c[0] * ny + c[1]) * nz + c[2];
        if (!pres[i]) {
            sum += v[i];
            ++cnt;
        }
    }
    if (cnt == 0) return;
    const double mean = sum / (double)cnt;
    for (int k = 0; k < d[axis]; ++k) {
        c[axis] = k;
        const size_t i = ((size_t)c[0] * ny + c[1]) * nz + c[2];
        if (!pres[i]) v[i] = mean;
    }
}

__global__ void k_zero_passive(double* v, const uint8_t* __restrict__ passive, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && passive[i]) v[i] = 0.0;
}

// alive = xe > ALIVE or an anchor cell (simp.py:726); partial count
__global__ void k_alive(const double* __restrict__ xe, const uint8_t* __restrict__ anchor, size_t n, double alive_min,
                        uint8_t* alive, double* partials) {
    __shared__ double sred[4];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        const uint8_t a = (xe[i] > alive_min || anchor[i]) ? 1 : 0;
        alive[i] = a;
        part = a;
    }
    const double r = bsum(part, sred);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// the components holding a live cell that touches a support (_connected,
// simp.py:264-282): their roots are flagged
__global__ void k_keep_roots(const uint8_t* __restrict__ alive, const uint8_t* __restrict__ support,
                             const int* __restrict__ labels, size_t n, uint8_t* rootflag) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && alive[i] && support[i]) rootflag[labels[i]] = 1;
}

// moduli of the reduced system: E where the cell is live and in a kept
// component, 0 elsewhere (_solve_mg, simp.py:806-807); partial kept count
__global__ void k_keep_moduli(double* e, const uint8_t* __restrict__ alive, const int* __restrict__ labels,
                              const uint8_t* __restrict__ rootflag, size_t n, double* partials) {
    __shared__ double sred[4];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        const bool keep = alive[i] && labels[i] >= 0 && rootflag[labels[i]];
        if (!keep) e[i] = 0.0;
        part = keep ? 1.0 : 0.0;
    }
    const double r = bsum(part, sred);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// a load on a dof no live cell holds (simp.py:808): flag set
__global__ void k_dead_loads(const double* __restrict__ f, const uint8_t* __restrict__ mask,
                             const uint8_t* __restrict__ held, size_t nn, size_t stride, int* flag) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= 3 * nn) return;
    const int c = (int)(i / nn);
    const size_t node = i % nn;
    const bool live = (((mask[node] | held[node]) >> c) & 1) != 0;
    if (!live && f[rhs * stride + i] != 0.0) *flag = 1;
}

// ce = sum_rhs u_c^T Ke u_c; partial of sum E ce; dc = -p (1-emin) x^(p-1) ce
__global__ void k_sens(const double* __restrict__ u, const double* __restrict__ ke, Geom g, int nrhs, size_t vstride,
                       const double* __restrict__ xphys, const double* __restrict__ e, double penal, double emin,
                       const double* __restrict__ wcase, double* ce_out, double* dc, double* partials) {
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
            ce += wcase ? wcase[r] * acc : acc;
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

// ---- the production loop (simp._optimize_run) ----------------------------

__device__ __forceinline__ double proj_value(double xt, double beta, double eta) {
    const double den = tanh(beta * eta) + tanh(beta * (1.0 - eta));
    return (tanh(beta * eta) + tanh(beta * (xt - eta))) / den;
}
__device__ __forceinline__ double proj_deriv(double xt, double beta, double eta) {
    const double den = tanh(beta * eta) + tanh(beta * (1.0 - eta));
    const double t = tanh(beta * (xt - eta));
    return beta * (1.0 - t * t) / den;
}

// blueprint xp = P(xt, eta) and analysed design xe = P(xt, eta_e), both 1 on
// preserved cells and 0 on passive ones (simp.py:1775-1812); the moduli
// E = emin + xe^p (1 - emin) (simp.py:1861); partials of sum xp over the
// design (the recorded volume)
__global__ void k_loop_physical(const double* __restrict__ xt, const uint8_t* __restrict__ pres,
                                const uint8_t* __restrict__ passive, const double* __restrict__ wdes, size_t n,
                                double beta, double eta, double eta_e, double penal, double emin, double* xp,
                                double* xe, double* e, double* partials) {
    __shared__ double sred[4];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        double p = proj_value(xt[i], beta, eta), q = proj_value(xt[i], beta, eta_e);
        if (pres[i]) p = q = 1.0;
        if (passive && passive[i]) p = q = 0.0;
        xp[i] = p;
        xe[i] = q;
        e[i] = emin + pow(q, penal) * (1.0 - emin);
        part = p * wdes[i];
    }
    const double r = bsum(part, sred);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// the chain rule's cell factors (simp.py:1994-2001, 2077):
// g = dc_xp P'(xt, eta_e) / hs, gv = w P'(xt, eta) / hs, 0 on preserved
// and passive cells
__global__ void k_loop_chain(const double* __restrict__ dcx, const double* __restrict__ xt,
                             const uint8_t* __restrict__ pres, const uint8_t* __restrict__ passive,
                             const double* __restrict__ wdes, const double* __restrict__ hs_inv, size_t n, double beta,
                             double eta, double eta_e, double* g, double* gv) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool off = pres[i] || (passive && passive[i]);
    const double dpe = off ? 0.0 : proj_deriv(xt[i], beta, eta_e);
    const double dp = off ? 0.0 : proj_deriv(xt[i], beta, eta);
    g[i] = dcx[i] * dpe * hs_inv[i];
    gv[i] = wdes[i] * dp * hs_inv[i];
}

// dc <= 0, dv >= 1e-12 in place (simp.py:2191, 2204), ratio = sqrt(-dc / dv);
// with a previous step, partials of a = gp . dx and b = (dc + lp dv) . dx
// (simp.py:2237-2238) in rows 0 and 1
__global__ void k_clamp_only(double* __restrict__ dc, double* __restrict__ dv, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dc[i] = fmin(dc[i], 0.0);
    dv[i] = fmax(dv[i], 1e-12);
}

__global__ void k_loop_clamp(double* __restrict__ dc, double* __restrict__ dv, double* __restrict__ ratio, size_t n,
                             const double* __restrict__ dxp, const double* __restrict__ gp, double lam_p,
                             double* partials) {
    __shared__ double sred[4];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double pa = 0.0, pb = 0.0;
    if (i < n) {
        const double c = fmin(dc[i], 0.0), v = fmax(dv[i], 1e-12);
        dc[i] = c;
        dv[i] = v;
        ratio[i] = sqrt(-c / v);
        if (dxp) {
            pa = gp[i] * dxp[i];
            pb = (c + lam_p * v) * dxp[i];
        }
    }
    if (!dxp) return;
    const double ra = bsum(pa, sred);
    const double rb = bsum(pb, sred);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = ra;
        partials[gridDim.x + blockIdx.x] = rb;
    }
}

// one step at multiplier lm: xn = clip(x ratio / sqrt(lm), x -/+ move in
// [0, 1]), preserved 1, passive 0 (simp.py:2288-2302)
__global__ void k_loop_step(const double* __restrict__ x, const double* __restrict__ ratio,
                            const uint8_t* __restrict__ pres, const uint8_t* __restrict__ passive, size_t n,
                            double move, double sqrt_lm, double* xn) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[i];
    const double lo = fmax(0.0, xi - move), hi = fmin(1.0, xi + move);
    double v = fmin(fmax(xi * ratio[i] / sqrt_lm, lo), hi);
    if (pres[i]) v = 1.0;
    if (passive && passive[i]) v = 0.0;
    xn[i] = v;
}

// partials of the stepped blueprint's volume sum_design P(xt, eta)
// (simp.py:2299-2300; preserved cells carry no design weight)
__global__ void k_loop_volume(const double* __restrict__ xt, const uint8_t* __restrict__ pres,
                              const uint8_t* __restrict__ passive, const double* __restrict__ wdes, size_t n,
                              double beta, double eta, double* partials) {
    __shared__ double sred[4];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double part = 0.0;
    if (i < n) {
        double p = proj_value(xt[i], beta, eta);
        if (pres[i]) p = 1.0;
        if (passive && passive[i]) p = 0.0;
        part = p * wdes[i];
    }
    const double r = bsum(part, sred);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// the step just taken, for the next iteration's measurement: dx = xn - x,
// gp = dc + lm dv (simp.py:2381)
__global__ void k_loop_stepprev(const double* __restrict__ xn, const double* __restrict__ x,
                                const double* __restrict__ dc, const double* __restrict__ dv, size_t n, double lm,
                                double* dxp, double* gp) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    dxp[i] = xn[i] - x[i];
    gp[i] = dc[i] + lm * dv[i];
}

__global__ void k_reduce_sum(const double* __restrict__ partials, int nb, double* out) {
    __shared__ double sred[4];
    double acc = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partials[i];
    const double s = bsum(acc, sred);
    if (threadIdx.x == 0) *out = s;
}

// block partial sums of a cell field (one per block; k_reduce_sum adds them
// in block order)
__global__ void k_sum_partials(const double* __restrict__ a, size_t n, double* __restrict__ partials) {
    __shared__ double sred[4];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double s = bsum(i < n ? a[i] : 0.0, sred);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
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

// cone weights w = radius - |d| pitch > 0 for offsets |d_i| <= reach, in
// (dx, dy, dz) order (simp.GridFilter, simp.py:1152-1159)
std::vector<Tap> make_taps(double radius, int reach, double pitch = 1.0) {
    std::vector<Tap> taps;
    for (int dx = -reach; dx <= reach; ++dx)
        for (int dy = -reach; dy <= reach; ++dy)
            for (int dz = -reach; dz <= reach; ++dz) {
                const double dist = std::sqrt((double)(dx * dx + dy * dy + dz * dz));
                const double w = radius - (pitch == 1.0 ? dist : dist * pitch);
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
    // the production loop (lp_simp_set_loop)
    struct Loop {
        bool on = false;
        lp_loop_opts o{};
        std::vector<int> schedule;
        double beta = 1.0, move_k = 0.2, lam_last = 0.0, lam_p = 0.0, c0 = 0.0, target = 0.0, n_design = 0.0;
        int since_beta = 0, it = 0;
        bool have_prev = false, converged = false, weighted = false;
        std::vector<double> comp_hist;
        DBuf pres, wdes, xt, xe, g, gv, ratio, dxp, gp, wcase;
        // ties: mirror axes in order, extrusion axis (-1 none)
        std::vector<int> mirror;
        int extrusion = -1;
        // the multigrid path's live-cell reduction (simp.py:795-826)
        bool reduce = false;
        DBuf anchor, support, alive, labels, rootflag, e_full, flag;
        int reduced_iters = 0, full_retries = 0, full_for_loads = 0;
    } L;
    ~lp_simp() {
        if (mma) lp_mma_destroy(mma);
    }
};

namespace {

// iterations at projection sharpness beta (simp.py:1436-1440)
int loop_hold(const lp_simp::Loop& L, double beta) {
    const int k = (int)std::lround(std::log2(std::max(beta, 1.0)));
    return k < (int)L.schedule.size() ? L.schedule[k] : L.o.beta_every;
}

double reduce_host(lp_simp* S, const double* partials, int nb, int slot) {
    hipStream_t s = S->g->stream;
    double* scal = S->scalar.as<double>();
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kBlock), 0, s, partials, nb, scal + slot);
    LP_HIP(hipGetLastError());
    double v = 0.0;
    LP_HIP(hipMemcpyAsync(&v, scal + slot, 8, hipMemcpyDeviceToHost, s));
    LP_HIP(hipStreamSynchronize(s));
    return v;
}

// the design ties of the loop on a cell field (mirror planes in order, then
// the extrusion columns)
void loop_ties(lp_simp* S, double* v) {
    const lp_simp::Loop& L = S->L;
    const Geom& g0 = S->g->g0;
    hipStream_t s = S->g->stream;
    for (int a : L.mirror)
        hipLaunchKernelGGL(k_tie_mirror, dim3(nblk(S->ncell)), dim3(kBlock), 0, s, v, g0.nx, g0.ny, g0.nz, a);
    if (L.extrusion >= 0) {
        const int d[3] = {g0.nx, g0.ny, g0.nz};
        const size_t ncol = S->ncell / (size_t)d[L.extrusion];
        hipLaunchKernelGGL(k_tie_extrude, dim3(nblk(ncol)), dim3(kBlock), 0, s, v, L.pres.as<uint8_t>(), g0.nx, g0.ny,
                           g0.nz, L.extrusion);
    }
    LP_HIP(hipGetLastError());
}

bool loop_tied(const lp_simp* S) { return !S->L.mirror.empty() || S->L.extrusion >= 0; }

void loop_filter(lp_simp* S, const double* v, const double* post, double* out) {
    const Geom& g0 = S->g->g0;
    hipLaunchKernelGGL(k_filter, dim3(nblk(S->ncell)), dim3(kBlock), 0, S->g->stream, v, (const double*)nullptr, post,
                       (const uint8_t*)nullptr, S->taps.as<Tap>(), S->ntaps, g0.nx, g0.ny, g0.nz, out, (double*)nullptr,
                       (const Bisect*)nullptr);
    LP_HIP(hipGetLastError());
}

// one iteration of simp._optimize_run's loop (simp.py:1840-2449)
int loop_iterate(lp_simp* S, lp_simp_stats& st) {
    lp_simp::Loop& L = S->L;
    Grid& G = *S->g;
    hipStream_t s = G.stream;
    const size_t n = S->ncell;
    const Geom& g0 = G.g0;
    const size_t vn = 3 * g0.nn;
    const unsigned nb = nblk(n);
    const auto t0 = std::chrono::steady_clock::now();
    const lp_loop_opts& o = L.o;
    const double beta = L.beta, eta_e = o.eta + o.erode;
    const uint8_t* pas = S->has_passive ? S->passive.as<uint8_t>() : nullptr;
    const uint8_t* pres = L.pres.as<uint8_t>();
    L.it += 1;
    // designs and moduli (simp.py:1857-1861)
    loop_filter(S, S->x.as<double>(), S->hs_inv.as<double>(), L.xt.as<double>());
    hipLaunchKernelGGL(k_loop_physical, dim3(nb), dim3(kBlock), 0, s, L.xt.as<double>(), pres, pas,
                       L.wdes.as<double>(), n, beta, o.eta, eta_e, o.penal, o.emin, S->xphys.as<double>(),
                       L.xe.as<double>(), G.lv[0].e64.as<double>(), S->partials.as<double>());
    LP_HIP(hipGetLastError());
    const double vol_now = reduce_host(S, S->partials.as<double>(), (int)nb, 2);
    double* e64 = G.lv[0].e64.as<double>();
    // the multigrid path's reduction (_solve_mg, simp.py:795-826): under
    // REDUCE_BELOW of the cells live, the cells that are not live or not
    // joined to a support get modulus 0
    bool reduced = false;
    if (L.reduce) {
        hipLaunchKernelGGL(k_alive, dim3(nb), dim3(kBlock), 0, s, L.xe.as<double>(), L.anchor.as<uint8_t>(), n,
                           kAliveMin, L.alive.as<uint8_t>(), S->partials.as<double>());
        LP_HIP(hipGetLastError());
        const double n_alive = reduce_host(S, S->partials.as<double>(), (int)nb, 5);
        if (n_alive < kReduceBelow * (double)n) {
            LP_HIP(hipMemcpyAsync(L.e_full.p, e64, n * 8, hipMemcpyDeviceToDevice, s));
            label26_roots(g0.nx, g0.ny, g0.nz, L.alive.as<uint8_t>(), L.labels.as<int>(), L.flag.as<int>(), s);
            hipLaunchKernelGGL(k_fill_u8, dim3(nb), dim3(kBlock), 0, s, L.rootflag.as<uint8_t>(), n, (uint8_t)0);
            hipLaunchKernelGGL(k_keep_roots, dim3(nb), dim3(kBlock), 0, s, L.alive.as<uint8_t>(),
                               L.support.as<uint8_t>(), L.labels.as<int>(), n, L.rootflag.as<uint8_t>());
            hipLaunchKernelGGL(k_keep_moduli, dim3(nb), dim3(kBlock), 0, s, e64, L.alive.as<uint8_t>(),
                               L.labels.as<int>(), L.rootflag.as<uint8_t>(), n, S->partials.as<double>());
            LP_HIP(hipGetLastError());
            const double n_keep = reduce_host(S, S->partials.as<double>(), (int)nb, 6);
            reduced = n_keep < (double)n;
            if (!reduced) LP_HIP(hipMemcpyAsync(e64, L.e_full.p, n * 8, hipMemcpyDeviceToDevice, s));
        }
    }
    auto rebuild = [&]() -> int {
        int r = build_hierarchy(G);
        if (r == LP_SINGULAR && !G.strict) {
            G.strict = true;
            r = build_hierarchy(G);
        }
        if (r != LP_OK) return r;
        G.built = true;
        if (G.opt.smoother == LP_SMOOTH_CHEBYSHEV)
            for (int i = 0; i < G.nlevels; ++i) lp_estimate_lmax(S->g, i, nullptr);
        return LP_OK;
    };
    auto full_grid = [&]() -> int {
        LP_HIP(hipMemcpyAsync(e64, L.e_full.p, n * 8, hipMemcpyDeviceToDevice, s));
        reduced = false;
        return rebuild();
    };
    int rc = rebuild();
    if (reduced && rc == LP_OK) {
        // a load on a dof the kept cells do not hold: the whole grid
        LP_HIP(hipMemsetAsync(L.flag.p, 0, 4, s));
        hipLaunchKernelGGL(k_dead_loads, dim3(nblk(vn), S->nrhs), dim3(kBlock), 0, s, S->loads.as<double>(),
                           G.lv[0].mask.as<uint8_t>(), G.lv[0].held.as<uint8_t>(), g0.nn, vn, L.flag.as<int>());
        LP_HIP(hipGetLastError());
        int dead = 0;
        LP_HIP(hipMemcpyAsync(&dead, L.flag.p, 4, hipMemcpyDeviceToHost, s));
        LP_HIP(hipStreamSynchronize(s));
        if (dead) {
            ++L.full_for_loads;
            rc = full_grid();
        }
    }
    if (rc != LP_OK && reduced) {
        ++L.full_retries;
        rc = full_grid();
    }
    if (rc != LP_OK) return rc;
    std::vector<int> iters(S->nrhs), ok(S->nrhs);
    auto solve_now = [&](bool warm) -> int {
        hipLaunchKernelGGL(k_mask_loads, dim3(nblk(vn), S->nrhs), dim3(kBlock), 0, s, S->loads.as<double>(),
                           G.lv[0].mask.as<uint8_t>(), G.b.as<double>(), g0.nn, vn);
        LP_HIP(hipGetLastError());
        return solve_stall(G, S->nrhs, warm, iters.data(), ok.data());
    };
    auto all_ok = [&]() {
        for (int k = 0; k < S->nrhs; ++k)
            if (!ok[k]) return false;
        return true;
    };
    const bool warm = S->warm && S->have_u;
    if (warm && L.reduce) {
        // the last displacements, masked to this system's live dofs, start
        // the solve (mg.py:941: x0 * mask)
        hipLaunchKernelGGL(k_mask_loads, dim3(nblk(vn), S->nrhs), dim3(kBlock), 0, s, G.x.as<double>(),
                           G.lv[0].mask.as<uint8_t>(), G.x.as<double>(), g0.nn, vn);
        LP_HIP(hipGetLastError());
    }
    rc = solve_now(warm);
    if (reduced && (rc != LP_OK || !all_ok())) {
        // the kept cells hang on the rest by an edge or a corner, or a
        // coarse level lost a thin support: the whole grid, from zero
        // (simp.py:817-822)
        ++L.full_retries;
        rc = full_grid();
        if (rc != LP_OK) return rc;
        rc = solve_now(false);
    }
    if (rc != LP_OK) return rc;
    if (!all_ok()) {
        set_error("linear solve did not converge; the structure may be a mechanism (check supports)");
        return LP_NOT_CONVERGED;
    }
    if (reduced) ++L.reduced_iters;
    S->have_u = true;
    // compliance and dc_xp on the analysed design (simp.py:1897-1904)
    hipLaunchKernelGGL(k_sens, dim3(nb), dim3(kBlock), 0, s, G.x.as<double>(), G.ke64.as<double>(), g0, S->nrhs, vn,
                       L.xe.as<double>(), G.lv[0].e64.as<double>(), o.penal, o.emin,
                       L.weighted ? L.wcase.as<double>() : (const double*)nullptr, S->ce.as<double>(),
                       S->dc.as<double>(), S->partials.as<double>());
    LP_HIP(hipGetLastError());
    const double comp = reduce_host(S, S->partials.as<double>(), (int)nb, 0);
    if (L.c0 == 0.0) L.c0 = comp;
    // dc = H (dc_xp P'_e / hs), dv = H (w P' / hs) (H symmetric)
    hipLaunchKernelGGL(k_loop_chain, dim3(nb), dim3(kBlock), 0, s, S->dc.as<double>(), L.xt.as<double>(), pres, pas,
                       L.wdes.as<double>(), S->hs_inv.as<double>(), n, beta, o.eta, eta_e, L.g.as<double>(),
                       L.gv.as<double>());
    LP_HIP(hipGetLastError());
    loop_filter(S, L.g.as<double>(), nullptr, S->dc.as<double>());
    loop_filter(S, L.gv.as<double>(), nullptr, S->dv.as<double>());
    if (loop_tied(S)) {
        // the clamped sensitivities of tied cells are averaged like the
        // cells (simp.py:2191-2226); the clamp below then leaves them
        hipLaunchKernelGGL(k_clamp_only, dim3(nb), dim3(kBlock), 0, s, S->dc.as<double>(), S->dv.as<double>(), n);
        LP_HIP(hipGetLastError());
        loop_ties(S, S->dc.as<double>());
        loop_ties(S, S->dv.as<double>());
    }
    // clamps, ratio and the measured step (simp.py:2228-2245)
    double sec = -1.0;
    double move;
    hipLaunchKernelGGL(k_loop_clamp, dim3(nb), dim3(kBlock), 0, s, S->dc.as<double>(), S->dv.as<double>(),
                       L.ratio.as<double>(), n, L.have_prev && o.adaptive_move ? L.dxp.as<double>() : (const double*)nullptr,
                       L.gp.as<double>(), L.lam_p, S->partials.as<double>());
    LP_HIP(hipGetLastError());
    if (o.adaptive_move) {
        if (L.have_prev) {
            const double a_s = reduce_host(S, S->partials.as<double>(), (int)nb, 3);
            const double b_s = reduce_host(S, S->partials.as<double>() + nb, (int)nb, 4);
            if (a_s < 0) {
                const double t_s = b_s > a_s ? a_s / (a_s - b_s) : INFINITY;
                sec = std::min(t_s, 99.0);
                L.move_k = std::min(L.move_k * std::min(std::max(t_s, 0.5), 1.2), o.move);
            }
        }
        move = L.move_k;
    } else {
        move = beta <= 4 ? o.move : o.move / 2;
    }
    // the volume multiplier (simp.py:2303-2373)
    int n_eval = 0;
    auto stepped = [&](double lm) -> double {
        ++n_eval;
        hipLaunchKernelGGL(k_loop_step, dim3(nb), dim3(kBlock), 0, s, S->x.as<double>(), L.ratio.as<double>(), pres,
                           pas, n, move, std::sqrt(lm), S->xnew.as<double>());
        LP_HIP(hipGetLastError());
        if (loop_tied(S)) {
            // each trial design is tied, then its void cells emptied again
            // (simp.py:2296-2304)
            loop_ties(S, S->xnew.as<double>());
            if (pas) hipLaunchKernelGGL(k_zero_passive, dim3(nb), dim3(kBlock), 0, s, S->xnew.as<double>(), pas, n);
            LP_HIP(hipGetLastError());
        }
        loop_filter(S, S->xnew.as<double>(), S->hs_inv.as<double>(), L.g.as<double>());
        hipLaunchKernelGGL(k_loop_volume, dim3(nb), dim3(kBlock), 0, s, L.g.as<double>(), pres, pas,
                           L.wdes.as<double>(), n, beta, o.eta, S->partials.as<double>());
        LP_HIP(hipGetLastError());
        return reduce_host(S, S->partials.as<double>(), (int)nb, 1) - L.target;
    };
    const double s_lo = std::log(1e-12), s_hi = std::log(1e12);
    double sm = L.lam_last > 0 ? std::min(std::max(std::log(L.lam_last), s_lo), s_hi) : 0.0;
    double lm = std::exp(sm);
    double f = stepped(lm);
    double s1 = sm, s2 = sm, f1 = 0, f2 = 0;
    bool h1 = false, h2 = false;
    if (f > 0) { f1 = f; h1 = true; } else { f2 = f; h2 = true; }
    double reach = 0.25;
    while ((!h1 || !h2) && f != 0.0) {
        if (!h1) {
            if (s1 <= s_lo) break;
            s1 = std::max(s1 - reach, s_lo);
            sm = s1;
        } else {
            if (s2 >= s_hi) break;
            s2 = std::min(s2 + reach, s_hi);
            sm = s2;
        }
        reach *= 4.0;
        lm = std::exp(sm);
        f = stepped(lm);
        if (f > 0) { f1 = f; s1 = sm; h1 = true; } else { f2 = f; s2 = sm; h2 = true; }
    }
    if (h1 && h2) {
        int side = 0;
        for (int k = 0; k < 60; ++k) {
            if (std::fabs(f) <= 1e-13 * L.target || f2 == f1 ||
                s2 - s1 <= 4e-16 * std::max(std::max(std::fabs(s1), std::fabs(s2)), 1.0))
                break;
            sm = (s1 * f2 - s2 * f1) / (f2 - f1);
            if (!(s1 < sm && sm < s2)) sm = 0.5 * (s1 + s2);
            lm = std::exp(sm);
            f = stepped(lm);
            if (f > 0) {
                s1 = sm; f1 = f;
                if (side == 1) f2 *= 0.5;
                side = 1;
            } else {
                s2 = sm; f2 = f;
                if (side == -1) f1 *= 0.5;
                side = -1;
            }
        }
    }
    L.lam_last = lm;
    const bool vol_held = h1 && h2;
    // change, the step for the next measurement, x = xnew
    LP_HIP(hipMemsetAsync(S->chg.p, 0, 8, s));
    hipLaunchKernelGGL(k_change, dim3(std::min(nb, 1024u)), dim3(kBlock), 0, s, S->xnew.as<double>(), S->x.as<double>(),
                       n, S->chg.as<unsigned long long>());
    LP_HIP(hipGetLastError());
    if (o.adaptive_move) {
        hipLaunchKernelGGL(k_loop_stepprev, dim3(nb), dim3(kBlock), 0, s, S->xnew.as<double>(), S->x.as<double>(),
                           S->dc.as<double>(), S->dv.as<double>(), n, lm, L.dxp.as<double>(), L.gp.as<double>());
        LP_HIP(hipGetLastError());
        L.have_prev = true;
        L.lam_p = lm;
    }
    std::swap(S->x.p, S->xnew.p);
    double change = 0.0;
    LP_HIP(hipMemcpyAsync(&change, S->chg.p, 8, hipMemcpyDeviceToHost, s));
    LP_HIP(hipStreamSynchronize(s));
    L.since_beta += 1;
    L.comp_hist.push_back(comp);
    st.compliance = comp;
    st.change = change;
    st.volume = L.n_design > 0 ? vol_now / L.n_design : 0.0;
    st.cg_iters = *std::max_element(iters.begin(), iters.end());
    st.cg_iters_total = 0;
    for (int k = 0; k < S->nrhs; ++k) st.cg_iters_total += iters[k];
    st.oc_evals = n_eval;
    st.beta = beta;
    st.move = move;
    st.step_fit = sec;
    st.lambda = lm;
    st.vol_held = vol_held ? 1 : 0;
    st.converged = 0;
    // continuation and the stop tests (simp.py:2413-2449)
    const bool measured = o.adaptive_move != 0;
    const bool free_step = !measured || change < 0.999 * move;
    if (beta < o.beta_max && ((change < o.tol * 2 && free_step) || L.since_beta >= loop_hold(L, beta))) {
        L.beta = std::min(beta * 2, o.beta_max);
        L.since_beta = 0;
        L.have_prev = false;
        L.move_k = L.move_k / 2;
    } else if (beta >= o.beta_max && L.it >= 3) {
        if (change < o.tol && free_step) {
            L.converged = true;
        } else if (measured && vol_held) {
            if (L.since_beta >= loop_hold(L, beta)) L.converged = true;
        } else if (L.since_beta >= o.stall_window && (int)L.comp_hist.size() > o.stall_window) {
            const double old = L.comp_hist[L.comp_hist.size() - 1 - o.stall_window];
            if (std::fabs(comp - old) <= o.stall_rtol * std::fabs(old)) L.converged = true;
        }
    }
    st.converged = L.converged ? 1 : 0;
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return LP_OK;
}

}  // namespace

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
            S->partials.alloc(2 * nblk(n) * 8);
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
        if (S->L.on) {
            for (int it = 0; it < niter; ++it) {
                stats[it] = lp_simp_stats{};
                const int rc = loop_iterate(S, stats[it]);
                if (rc != LP_OK) return rc;
            }
            return LP_OK;
        }
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
            rc = solve_stall(G, S->nrhs, warm, iters.data(), ok.data());
            if (rc != LP_OK) return rc;
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
                               (const double*)nullptr, S->ce.as<double>(), S->dc.as<double>(), S->partials.as<double>());
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
            hipLaunchKernelGGL(k_sum_partials, dim3(nb), dim3(kBlock), 0, s, S->xphys.as<double>(), n,
                               S->partials.as<double>());
            hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(kBlock), 0, s, S->partials.as<double>(), (int)nb, scal + 1);
            LP_HIP(hipGetLastError());
            std::swap(S->x.p, S->xnew.p);
            double h[3];
            Bisect bh;
            LP_HIP(hipMemcpyAsync(h, scal, 8, hipMemcpyDeviceToHost, s));
            LP_HIP(hipMemcpyAsync(h + 1, S->chg.p, 8, hipMemcpyDeviceToHost, s));
            LP_HIP(hipMemcpyAsync(h + 2, scal + 1, 8, hipMemcpyDeviceToHost, s));
            LP_HIP(hipMemcpyAsync(&bh, S->bis.p, sizeof(bh), hipMemcpyDeviceToHost, s));
            LP_HIP(hipStreamSynchronize(s));
            lp_simp_stats& st = stats[it];
            st.compliance = h[0];
            st.change = h[1];
            st.volume = h[2] / (double)n;
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

int lp_simp_set_loop(lp_simp* S, const lp_loop_opts* o, const uint8_t* preserve, const double* case_weights) {
    return guarded_s([&]() -> int {
        if (!S || !o || o->beta_start <= 0 || o->beta_max < o->beta_start || o->beta_every < 1 || o->move <= 0 ||
            o->penal <= 0 || o->emin <= 0 || o->emin >= 1 || o->stall_window < 1 || o->eta <= 0 || o->eta >= 1 ||
            o->eta + o->erode <= 0 || o->eta + o->erode >= 1 || (o->schedule_len > 0 && !o->schedule)) {
            set_error("lp_simp_set_loop: bad options");
            return LP_ERR_ARG;
        }
        lp_simp::Loop& L = S->L;
        const size_t n = S->ncell;
        hipStream_t s = S->g->stream;
        L.o = *o;
        L.schedule.assign(o->schedule, o->schedule + std::max(o->schedule_len, 0));
        for (DBuf* b : {&L.wdes, &L.xt, &L.xe, &L.g, &L.gv, &L.ratio, &L.dxp, &L.gp})
            if (!b->p) b->alloc(n * 8);
        if (!L.pres.p) L.pres.alloc(n);
        std::vector<uint8_t> pr(n, 0), pas(n, 0);
        if (preserve)
            for (size_t i = 0; i < n; ++i) pr[i] = preserve[i] ? 1 : 0;
        if (S->has_passive) LP_HIP(hipMemcpy(pas.data(), S->passive.p, n, hipMemcpyDeviceToHost));
        std::vector<double> w(n), x0(n);
        double nd = 0;
        for (size_t i = 0; i < n; ++i) {
            w[i] = pr[i] ? 0.0 : 1.0;
            nd += w[i];
            x0[i] = pr[i] ? 1.0 : (pas[i] ? 0.0 : S->volfrac);
        }
        if (nd == 0) {
            set_error("lp_simp_set_loop: nothing to optimize: every cell is preserved");
            return LP_ERR_ARG;
        }
        LP_HIP(hipMemcpy(L.pres.p, pr.data(), n, hipMemcpyHostToDevice));
        LP_HIP(hipMemcpy(L.wdes.p, w.data(), n * 8, hipMemcpyHostToDevice));
        LP_HIP(hipMemcpy(S->x.p, x0.data(), n * 8, hipMemcpyHostToDevice));
        L.weighted = case_weights != nullptr;
        if (L.weighted) {
            if (!L.wcase.p) L.wcase.alloc(S->nrhs * 8);
            LP_HIP(hipMemcpy(L.wcase.p, case_weights, S->nrhs * 8, hipMemcpyHostToDevice));
        }
        L.n_design = nd;
        L.target = S->volfrac * nd;
        L.beta = o->beta_start;
        L.move_k = o->move;
        L.lam_last = 0.0;
        L.lam_p = 0.0;
        L.c0 = 0.0;
        L.since_beta = 0;
        L.it = 0;
        L.have_prev = false;
        L.converged = false;
        L.comp_hist.clear();
        L.on = true;
        S->have_u = false;
        LP_HIP(hipStreamSynchronize(s));
        return LP_OK;
    });
}

int lp_simp_moduli(long long n, const double* x, double penal, double emin, double* e, double* de, int device_ptr) {
    return guarded_s([&]() -> int {
        if (n < 0 || !x || !e) return LP_ERR_ARG;
        if (n == 0) return LP_OK;
        DBuf dx, dee, dd;
        const double* px = x;
        double* pe = e;
        double* pd = de;
        if (!device_ptr) {
            dx.alloc(n * 8);
            dee.alloc(n * 8);
            LP_HIP(hipMemcpy(dx.p, x, n * 8, hipMemcpyHostToDevice));
            px = dx.as<double>();
            pe = dee.as<double>();
            if (de) {
                dd.alloc(n * 8);
                pd = dd.as<double>();
            }
        }
        hipLaunchKernelGGL(k_moduli_de, dim3(nblk((size_t)n)), dim3(kBlock), 0, nullptr, px, pe, pd, (size_t)n, penal,
                           emin);
        LP_HIP(hipGetLastError());
        if (!device_ptr) {
            LP_HIP(hipMemcpy(e, pe, n * 8, hipMemcpyDeviceToHost));
            if (de) LP_HIP(hipMemcpy(de, pd, n * 8, hipMemcpyDeviceToHost));
        } else {
            LP_HIP(hipDeviceSynchronize());
        }
        return LP_OK;
    });
}

int lp_simp_set_ties(lp_simp* S, const int* mirror_axes, int n_mirror, int extrusion_axis) {
    return guarded_s([&]() -> int {
        if (!S || n_mirror < 0 || n_mirror > 3 || (n_mirror > 0 && !mirror_axes) || extrusion_axis < -1 ||
            extrusion_axis > 2) {
            set_error("lp_simp_set_ties: bad argument");
            return LP_ERR_ARG;
        }
        if (!S->L.on) {
            set_error("lp_simp_set_ties: the production loop is not set (lp_simp_set_loop first)");
            return LP_ERR_STATE;
        }
        std::vector<int> m;
        for (int k = 0; k < n_mirror; ++k) {
            if (mirror_axes[k] < 0 || mirror_axes[k] > 2) {
                set_error("lp_simp_set_ties: mirror axis out of range");
                return LP_ERR_ARG;
            }
            m.push_back(mirror_axes[k]);
        }
        S->L.mirror = m;
        S->L.extrusion = extrusion_axis;
        return LP_OK;
    });
}

int lp_simp_set_reduction(lp_simp* S, const uint8_t* support_touch, const uint8_t* anchor) {
    return guarded_s([&]() -> int {
        if (!S) return LP_ERR_ARG;
        lp_simp::Loop& L = S->L;
        if (!support_touch || !anchor) {
            L.reduce = false;
            return LP_OK;
        }
        const size_t n = S->ncell;
        for (DBuf* b : {&L.anchor, &L.support, &L.alive, &L.rootflag})
            if (!b->p) b->alloc(n);
        if (!L.labels.p) L.labels.alloc(n * 4);
        if (!L.e_full.p) L.e_full.alloc(n * 8);
        if (!L.flag.p) L.flag.alloc(4);
        LP_HIP(hipMemcpy(L.support.p, support_touch, n, hipMemcpyHostToDevice));
        LP_HIP(hipMemcpy(L.anchor.p, anchor, n, hipMemcpyHostToDevice));
        L.reduce = true;
        L.reduced_iters = L.full_retries = L.full_for_loads = 0;
        return LP_OK;
    });
}

int lp_simp_loop_info(lp_simp* S, int* reduced_iters, int* full_retries, int* full_for_loads) {
    if (!S) return LP_ERR_ARG;
    if (reduced_iters) *reduced_iters = S->L.reduced_iters;
    if (full_retries) *full_retries = S->L.full_retries;
    if (full_for_loads) *full_for_loads = S->L.full_for_loads;
    return LP_OK;
}

int lp_simp_set_optimizer(lp_simp* S, int optimizer) {
    return guarded_s([&]() -> int {
        if (!S || (optimizer != 0 && optimizer != 1)) return LP_ERR_ARG;
        if (optimizer == 1 && S->has_passive) {
            set_error("lp_simp_set_optimizer: MMA runs on boxes without passive cells");
            return LP_ERR_ARG;
        }
        S->optimizer = optimizer;
        S->L.on = false;   // back from the production loop, if it was on
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

static int filter_apply(int nx, int ny, int nz, const std::vector<Tap>& taps, const double* v, double* out,
                        int device_ptr);

int lp_filter_apply(int nx, int ny, int nz, double radius, int reach, const double* v, double* out, int device_ptr) {
    return guarded_s([&]() -> int {
        if (nx < 1 || ny < 1 || nz < 1 || !v || !out || reach < 0) return LP_ERR_ARG;
        return filter_apply(nx, ny, nz, make_taps(radius, reach), v, out, device_ptr);
    });
}

int lp_grid_filter(int nx, int ny, int nz, double pitch, double radius, const double* v, double* out,
                   int device_ptr) {
    return guarded_s([&]() -> int {
        if (nx < 1 || ny < 1 || nz < 1 || !v || !out || !(pitch > 0) || !(radius >= 0)) return LP_ERR_ARG;
        const int reach = (int)std::floor(radius / pitch);
        return filter_apply(nx, ny, nz, make_taps(radius, reach, pitch), v, out, device_ptr);
    });
}

static int filter_apply(int nx, int ny, int nz, const std::vector<Tap>& taps, const double* v, double* out,
                        int device_ptr) {
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
}

}  // extern "C"
