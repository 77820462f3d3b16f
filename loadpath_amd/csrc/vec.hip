// Vector passes of the PCG and the smoothers, with deterministic two-stage
// reductions (per-block partials in a fixed order, then one block per rhs).
//
// Reference: mgops._cg_step / _xpby / _dot / _cast_into / _cheb_start
// (mgops.py:320-360, 557-585) and mg._pcg (mg.py:872-926). Scalars (alpha,
// beta, |r|) stay on the device in CgScal; the host reads only the per-rhs
// active / converged flags.
#include "lp_internal.h"

namespace lp {

namespace {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__device__ double block_sum(double v, double* sred) {
    v = wsum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sred[wv] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sred[i];
    __syncthreads();
    return s;  // valid on thread 0
}

template <typename T>
__global__ void k_dot(const T* __restrict__ a, const T* __restrict__ b, size_t n, size_t stride,
                      double* partials, const CgScal* scal) {
    __shared__ double sred[4];
    const int rhs = blockIdx.y;
    double acc = 0;
    if (!(scal && !scal[rhs].active)) {
        const T* pa = a + rhs * stride;
        const T* pb = b + rhs * stride;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
            acc += (double)(pa[i] * pb[i]);
    }
    const double s = block_sum(acc, sred);
    if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = s;
}

__global__ void k_reduce(const double* __restrict__ partials, int nb, double* dst, size_t dstride) {
    __shared__ double sred[4];
    const int rhs = blockIdx.x;
    double acc = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partials[(size_t)rhs * nb + i];
    const double s = block_sum(acc, sred);
    if (threadIdx.x == 0) *(double*)((char*)dst + rhs * dstride) = s;
}

template <typename S, typename D>
__global__ void k_cast(const S* __restrict__ src, D* __restrict__ dst, size_t n, size_t stride) {
    const int rhs = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[rhs * stride + i] = (D)src[rhs * stride + i];
}

template <typename T> __global__ void k_fill(T* dst, size_t n, T v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = v;
}

// mode 0: v *= num[rhs]; mode 1: v /= (T)num[rhs] (num = sqrt of stored value when mode 2)
template <typename T>
__global__ void k_scale_dev(T* v, size_t n, size_t stride, const double* num, size_t nstride, int mode) {
    const int rhs = blockIdx.y;
    const double s = *(const double*)((const char*)num + rhs * nstride);
    T f;
    if (mode == 0) f = (T)s;
    else f = (T)sqrt(s);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (mode == 0) v[rhs * stride + i] = v[rhs * stride + i] * f;
        else v[rhs * stride + i] = f != T(0) ? v[rhs * stride + i] / f : v[rhs * stride + i];
    }
}

// one power step's normalization in one launch: every block sums the dot
// partials of v . v in k_reduce's order (the same bits in every block and as
// k_reduce's), block 0 stores that sum in *sv and, when `prev` is given, the
// sum of the previous step's partials in *sp (k_reduce's order too), and
// each block divides its share of v by the root of the sum (k_scale_dev's
// mode 1: the reduce and scale launches of a power step, and the reduce of
// the step before, as one)
template <typename T>
__global__ void k_norm_scale(T* v, size_t n, const double* __restrict__ dotp, int nbd, double* sv,
                             const double* __restrict__ prev, int nbp, double* sp) {
    __shared__ double sred[4];
    double acc = 0;
    for (int i = threadIdx.x; i < nbd; i += blockDim.x) acc += dotp[i];
    const double s = block_sum(acc, sred);
    __shared__ double ss;
    if (threadIdx.x == 0) ss = s;
    __syncthreads();
    const double sum = ss;
    if (blockIdx.x == 0) {
        if (threadIdx.x == 0) *sv = sum;
        if (prev != nullptr) {
            double a = 0;
            for (int i = threadIdx.x; i < nbp; i += blockDim.x) a += prev[i];
            const double t = block_sum(a, sred);
            if (threadIdx.x == 0) *sp = t;
        }
    }
    const T f = (T)sqrt(sum);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = f != T(0) ? v[i] / f : v[i];
}

// x = 0; r = b; d = (dinv b) s   (mgops._cheb_start, mgops.py:320-327)
template <typename T>
__global__ void k_cheb_start(const T* __restrict__ b, const T* __restrict__ dinv, T s, T* __restrict__ x,
                             T* __restrict__ r, T* __restrict__ d, size_t n, size_t stride) {
    const int rhs = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t k = rhs * stride + i;
        const T bi = b[k];
        x[k] = T(0);
        r[k] = bi;
        d[k] = (dinv[i] * bi) * s;
    }
}

template <typename T>
__global__ void k_axpy1(T* __restrict__ x, const T* __restrict__ d, size_t n, size_t stride) {
    const int rhs = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[rhs * stride + i] += d[rhs * stride + i];
}

// first damped-Jacobi sweep from zero: x = w dinv b
template <typename T>
__global__ void k_jacobi_start(const T* __restrict__ b, const T* __restrict__ dinv, double w, T* __restrict__ x,
                               size_t n, size_t stride) {
    const int rhs = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[rhs * stride + i] = (T)(w * (double)dinv[i] * (double)b[rhs * stride + i]);
}

// x += alpha p; r -= alpha ap; partial |r|^2   (mgops._cg_step, mgops.py:346-354)
template <typename V>
__global__ void k_cg_step(double* __restrict__ r, const double* __restrict__ ap, size_t n, size_t stride,
                          const CgScal* scal, double* partials, V* __restrict__ rv) {
    __shared__ double sred[4];
    const int rhs = blockIdx.y;
    double acc = 0;
    if (scal[rhs].active) {
        const double a = scal[rhs].alpha;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const size_t k = rhs * stride + i;
            const double ri = r[k] - a * ap[k];
            r[k] = ri;
            rv[k] = (V)ri;
            acc += ri * ri;
        }
    }
    const double s = block_sum(acc, sred);
    if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = s;
}

// x += alpha p with the direction of this iteration (rhs whose alpha is
// pending), then p = z + beta p (mgops._xpby, mgops.py:357-360) for the
// active ones, z = the V-cycle's result read in its own precision (never
// stored in FP64); xonly: the x update alone (the last iteration)
template <typename S>
__global__ void k_cg_update(double* __restrict__ x, double* __restrict__ p, const S* __restrict__ z, size_t n,
                            size_t stride, const CgScal* scal, int xonly) {
    const int rhs = blockIdx.y;
    const bool upd = scal[rhs].xpend != 0;
    const bool dir = !xonly && scal[rhs].active;
    if (!upd && !dir) return;
    const double a = scal[rhs].alpha, beta = scal[rhs].beta;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t k = rhs * stride + i;
        const double pk = p[k];
        // explicit FMAs: the fused operator (EPI_CG_APPLY) rounds the same way
        if (upd) x[k] = fma(a, pk, x[k]);
        if (dir) p[k] = fma(beta, pk, (double)z[k]);
    }
}

// Finalize stages (one block per rhs): reduce the partials, update CgScal.
__global__ void k_cg_finalize(const double* __restrict__ partials, int nb, CgScal* scal, int stage,
                              double rtol) {
    __shared__ double sred[4];
    const int rhs = blockIdx.x;
    double acc = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partials[(size_t)rhs * nb + i];
    const double s = block_sum(acc, sred);
    if (threadIdx.x != 0) return;
    CgScal& c = scal[rhs];
    switch (stage) {
        case 0:  // pap = p.Ap -> alpha
            c.xpend = 0;
            if (!c.active) break;
            c.pap = s;
            if (!(s > 0)) {
                c.failed = 1;
                c.active = 0;
            } else {
                c.alpha = c.rz / s;
                c.xpend = 1;
            }
            break;
        case 1:  // rr -> convergence
            if (!c.active) break;
            c.rr = s;
            c.iters += 1;
            if (sqrt(s) <= c.tol) {
                c.converged = 1;
                c.active = 0;
            }
            break;
        case 2:  // rz_new -> beta
            if (!c.active) break;
            c.beta = s / c.rz;
            c.rz = s;
            break;
        case 3:  // |b|^2 -> bnorm, tol; reset counters
            c.bnorm = sqrt(s);
            c.tol = rtol * c.bnorm;
            c.iters = 0;
            c.failed = 0;
            c.converged = c.bnorm == 0 ? 1 : 0;
            c.active = c.bnorm == 0 ? 0 : 1;
            c.rz = 0;
            c.alpha = 0;
            c.beta = 0;
            c.xpend = 0;
            break;
        case 4:  // rz = r.z
            if (!c.active) break;
            c.rz = s;
            break;
        case 5:  // xax (warm start): stored in tmp
            c.tmp = s;
            break;
        case 6:  // b.x -> warm-start scale s = bx / xax (or 0), stored in alpha
            c.alpha = (c.tmp > 0) ? s / c.tmp : 0.0;
            break;
        case 7:  // |b - s Ax|^2 -> keep the scaled start (beta = 1) or drop it (beta = 0)
            c.beta = (c.alpha > 0 && sqrt(s) < c.bnorm) ? 1.0 : 0.0;
            break;
        case 8:  // the FP32 recurrence's |r|^2: counted, convergence left to the true residual
            if (!c.active) break;
            c.rr = s;
            c.iters += 1;
            break;
        case 9:  // the true residual |b - A x|^2 (FP64) of a reliable update -> convergence
            if (!c.active) break;
            c.rr = s;
            if (sqrt(s) <= c.tol) {
                c.converged = 1;
                c.active = 0;
            }
            break;
        case 10:  // rz of a restart: the next direction is z itself
            if (!c.active) break;
            c.beta = 0.0;
            c.rz = s;
            break;
        default: break;
    }
}

// warm start: if beta==1: x *= s; r = b - s ax; else x = 0; r = b
__global__ void k_warm_apply(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ b,
                             const double* __restrict__ ax, size_t n, size_t stride, const CgScal* scal) {
    const int rhs = blockIdx.y;
    const bool keep = scal[rhs].beta == 1.0;
    const double s = scal[rhs].alpha;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t k = rhs * stride + i;
        if (keep) {
            x[k] *= s;
            r[k] = b[k] - s * ax[k];
        } else {
            x[k] = 0;
            r[k] = b[k];
        }
    }
}

// |b - s ax|^2 partials for the warm-start test
__global__ void k_warm_norm(const double* __restrict__ b, const double* __restrict__ ax, size_t n,
                            size_t stride, const CgScal* scal, double* partials) {
    __shared__ double sred[4];
    const int rhs = blockIdx.y;
    const double s = scal[rhs].alpha;
    double acc = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t k = rhs * stride + i;
        const double v = b[k] - s * ax[k];
        acc += v * v;
    }
    const double t = block_sum(acc, sred);
    if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = t;
}

// z = (double) v (the V-cycle's result); partials of r . z for active rhs
// and, when ap is set, of z . ap in the rows after the nrhs rows of r . z
// (the flexible beta's second sum)
template <typename S>
__global__ void k_cast_dot(const S* __restrict__ v, const double* __restrict__ r, double* __restrict__ z, size_t n,
                           size_t stride, double* partials, const CgScal* scal, const double* __restrict__ ap) {
    __shared__ double sred[4];
    const int rhs = blockIdx.y;
    const bool on = !(scal && !scal[rhs].active);
    double acc = 0, acc2 = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t k = rhs * stride + i;
        const double zi = (double)v[k];
        if (z) z[k] = zi;
        if (on) {
            acc += r[k] * zi;
            if (ap) acc2 += ap[k] * zi;
        }
    }
    const double t = block_sum(acc, sred);
    if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = t;
    if (ap) {
        __syncthreads();
        const double t2 = block_sum(acc2, sred);
        if (threadIdx.x == 0) partials[((size_t)gridDim.y + rhs) * gridDim.x + blockIdx.x] = t2;
    }
}

// flexible (Polak-Ribiere) beta for an inexact preconditioner:
// beta = z1 . (r1 - r0) / (z0 . r0) = -alpha (z1 . A p0) / rz0, since
// r1 - r0 = -alpha A p0; rz = z1 . r1
__global__ void k_cg_beta_flex(const double* __restrict__ partials, int nb, int nrhs, CgScal* scal) {
    __shared__ double sred[4];
    const int rhs = blockIdx.x;
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        a += partials[(size_t)rhs * nb + i];
        b += partials[((size_t)nrhs + rhs) * nb + i];
    }
    const double rz1 = block_sum(a, sred);
    __syncthreads();
    const double zap = block_sum(b, sred);
    if (threadIdx.x != 0) return;
    CgScal& c = scal[rhs];
    if (!c.active) return;
    c.beta = -c.alpha * zap / c.rz;
    c.rz = rz1;
}

// ---- FP32 CG with reliable FP64 updates (the mixed path) ----------------

// xlo += alpha p (rhs whose alpha is pending); r -= alpha ap; partial |r|^2.
// r is the V-cycle's input vector itself (FP32)
__global__ void k_cg32_step(float* __restrict__ r, const float* __restrict__ ap, float* __restrict__ xlo,
                            const float* __restrict__ p, size_t n, size_t stride, const CgScal* scal,
                            double* partials) {
    __shared__ double sred[4];
    const int rhs = blockIdx.y;
    double acc = 0;
    if (scal[rhs].active) {
        const float a = (float)scal[rhs].alpha;
        const bool upd = scal[rhs].xpend != 0;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const size_t k = rhs * stride + i;
            if (upd) xlo[k] += a * p[k];
            const float ri = r[k] - a * ap[k];
            r[k] = ri;
            acc += (double)ri * (double)ri;
        }
    }
    const double t = block_sum(acc, sred);
    if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = t;
}

// partials of r . z (rows 0..nrhs-1) and z . ap (rows nrhs..2 nrhs-1), FP32
// fields summed in double, active rhs only
__global__ void k_dot2_f32(const float* __restrict__ r, const float* __restrict__ z, const float* __restrict__ ap,
                           size_t n, size_t stride, double* partials, const CgScal* scal) {
    __shared__ double sred[4];
    const int rhs = blockIdx.y;
    double acc = 0, acc2 = 0;
    if (scal[rhs].active) {
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const size_t k = rhs * stride + i;
            const double zi = (double)z[k];
            acc += (double)r[k] * zi;
            if (ap) acc2 += (double)ap[k] * zi;
        }
    }
    const double t = block_sum(acc, sred);
    if (threadIdx.x == 0) partials[(size_t)rhs * gridDim.x + blockIdx.x] = t;
    if (ap) {
        __syncthreads();
        const double t2 = block_sum(acc2, sred);
        if (threadIdx.x == 0) partials[((size_t)gridDim.y + rhs) * gridDim.x + blockIdx.x] = t2;
    }
}

// p = z + beta p for the active rhs (first: p = z)
__global__ void k_cg32_update(float* __restrict__ p, const float* __restrict__ z, size_t n, size_t stride,
                              const CgScal* scal, int first) {
    const int rhs = blockIdx.y;
    if (!scal[rhs].active) return;
    const float beta = first ? 0.0f : (float)scal[rhs].beta;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t k = rhs * stride + i;
        p[k] = first ? z[k] : z[k] + beta * p[k];
    }
}

// the reliable update: x += (double) xlo; xlo = 0 (every rhs: a converged
// one's xlo is already 0, a broken-down one's holds its last steps)
__global__ void k_cg32_fold(double* __restrict__ x, float* __restrict__ xlo, size_t n, size_t stride) {
    const int rhs = blockIdx.y;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t k = rhs * stride + i;
        x[k] += (double)xlo[k];
        xlo[k] = 0.0f;
    }
}

// rv = (float) r for the active rhs
__global__ void k_cast_active(const double* __restrict__ r, float* __restrict__ rv, size_t n, size_t stride,
                              const CgScal* scal) {
    const int rhs = blockIdx.y;
    if (!scal[rhs].active) return;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        rv[rhs * stride + i] = (float)r[rhs * stride + i];
}

inline int grid_for(size_t n) {
    size_t b = (n + kBlock - 1) / kBlock;
    if (b > 2048) b = 2048;
    if (b < 1) b = 1;
    return (int)b;
}

}  // namespace

void ensure_partials(Grid& G, size_t n) {
    if (n > G.partials_cap) {
        G.partials.alloc(n * sizeof(double));
        G.partials_cap = n;
    }
}

int dot_blocks(size_t n) {
    size_t b = (n + 4 * kBlock - 1) / (4 * kBlock);
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    return (int)b;
}

void launch_dot(const void* a, const void* b, int eb, size_t n, size_t stride, int nrhs, double* partials,
                const CgScal* scal, hipStream_t s) {
    dim3 grid(dot_blocks(n), nrhs);
    if (eb == 8)
        hipLaunchKernelGGL(k_dot<double>, grid, dim3(kBlock), 0, s, (const double*)a, (const double*)b, n, stride,
                           partials, scal);
    else
        hipLaunchKernelGGL(k_dot<float>, grid, dim3(kBlock), 0, s, (const float*)a, (const float*)b, n, stride,
                           partials, scal);
    LP_HIP(hipGetLastError());
}

void launch_reduce(const double* partials, int nb, int nrhs, double* dst, size_t dstride, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce, dim3(nrhs), dim3(kBlock), 0, s, partials, nb, dst, dstride);
    LP_HIP(hipGetLastError());
}

void launch_cast(const void* src, int sb, void* dst, int db, size_t n, size_t stride, int nrhs, hipStream_t s) {
    dim3 grid(grid_for(n), nrhs);
    if (sb == 8 && db == 4)
        hipLaunchKernelGGL((k_cast<double, float>), grid, dim3(kBlock), 0, s, (const double*)src, (float*)dst, n, stride);
    else if (sb == 4 && db == 8)
        hipLaunchKernelGGL((k_cast<float, double>), grid, dim3(kBlock), 0, s, (const float*)src, (double*)dst, n, stride);
    else if (sb == 8 && db == 8)
        hipLaunchKernelGGL((k_cast<double, double>), grid, dim3(kBlock), 0, s, (const double*)src, (double*)dst, n, stride);
    else
        hipLaunchKernelGGL((k_cast<float, float>), grid, dim3(kBlock), 0, s, (const float*)src, (float*)dst, n, stride);
    LP_HIP(hipGetLastError());
}

void launch_cast_dot(const void* v, int vb, const double* r, double* z, size_t n, size_t stride, int nrhs,
                     double* partials, const CgScal* scal, hipStream_t s, const double* ap) {
    dim3 grid(dot_blocks(n), nrhs);
    if (vb == 4)
        hipLaunchKernelGGL(k_cast_dot<float>, grid, dim3(kBlock), 0, s, (const float*)v, r, z, n, stride, partials,
                           scal, ap);
    else
        hipLaunchKernelGGL(k_cast_dot<double>, grid, dim3(kBlock), 0, s, (const double*)v, r, z, n, stride, partials,
                           scal, ap);
    LP_HIP(hipGetLastError());
}

void launch_cg_beta_flex(Grid& G, int nrhs, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_beta_flex, dim3(nrhs), dim3(kBlock), 0, s, G.partials.as<double>(), nb, nrhs,
                       G.scal.as<CgScal>());
    LP_HIP(hipGetLastError());
}

void launch_fill(void* dst, int eb, size_t n, double v, hipStream_t s) {
    if (eb == 8) hipLaunchKernelGGL(k_fill<double>, dim3(grid_for(n)), dim3(kBlock), 0, s, (double*)dst, n, v);
    else if (eb == 4) hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n)), dim3(kBlock), 0, s, (float*)dst, n, (float)v);
    else hipLaunchKernelGGL(k_fill<uint8_t>, dim3(grid_for(n)), dim3(kBlock), 0, s, (uint8_t*)dst, n, (uint8_t)v);
    LP_HIP(hipGetLastError());
}

void launch_norm_scale(void* v, int eb, size_t n, const double* dotp, int nbd, double* sv, const double* prev,
                       int nbp, double* sp, hipStream_t s) {
    dim3 grid(grid_for(n));
    if (eb == 8)
        hipLaunchKernelGGL(k_norm_scale<double>, grid, dim3(kBlock), 0, s, (double*)v, n, dotp, nbd, sv, prev, nbp, sp);
    else
        hipLaunchKernelGGL(k_norm_scale<float>, grid, dim3(kBlock), 0, s, (float*)v, n, dotp, nbd, sv, prev, nbp, sp);
    LP_HIP(hipGetLastError());
}

void launch_scale_by_dev(void* v, int eb, size_t n, size_t stride, int nrhs, const double* num, size_t nstride,
                         int mode, hipStream_t s) {
    dim3 grid(grid_for(n), nrhs);
    if (eb == 8)
        hipLaunchKernelGGL(k_scale_dev<double>, grid, dim3(kBlock), 0, s, (double*)v, n, stride, num, nstride, mode);
    else
        hipLaunchKernelGGL(k_scale_dev<float>, grid, dim3(kBlock), 0, s, (float*)v, n, stride, num, nstride, mode);
    LP_HIP(hipGetLastError());
}

void launch_cheb_start(const void* b, const void* dinv, double s_, void* x, void* r, void* d, int eb, size_t n,
                       size_t stride, int nrhs, hipStream_t st) {
    dim3 grid(grid_for(n), nrhs);
    if (eb == 8)
        hipLaunchKernelGGL(k_cheb_start<double>, grid, dim3(kBlock), 0, st, (const double*)b, (const double*)dinv, s_,
                           (double*)x, (double*)r, (double*)d, n, stride);
    else
        hipLaunchKernelGGL(k_cheb_start<float>, grid, dim3(kBlock), 0, st, (const float*)b, (const float*)dinv,
                           (float)s_, (float*)x, (float*)r, (float*)d, n, stride);
    LP_HIP(hipGetLastError());
}

void launch_axpy_inplace(void* x, const void* d, int eb, size_t n, size_t stride, int nrhs, hipStream_t st) {
    dim3 grid(grid_for(n), nrhs);
    if (eb == 8) hipLaunchKernelGGL(k_axpy1<double>, grid, dim3(kBlock), 0, st, (double*)x, (const double*)d, n, stride);
    else hipLaunchKernelGGL(k_axpy1<float>, grid, dim3(kBlock), 0, st, (float*)x, (const float*)d, n, stride);
    LP_HIP(hipGetLastError());
}

void launch_jacobi_start(const void* b, const void* dinv, double w, void* x, int eb, size_t n, size_t stride,
                         int nrhs, hipStream_t st) {
    dim3 grid(grid_for(n), nrhs);
    if (eb == 8)
        hipLaunchKernelGGL(k_jacobi_start<double>, grid, dim3(kBlock), 0, st, (const double*)b, (const double*)dinv, w,
                           (double*)x, n, stride);
    else
        hipLaunchKernelGGL(k_jacobi_start<float>, grid, dim3(kBlock), 0, st, (const float*)b, (const float*)dinv, w,
                           (float*)x, n, stride);
    LP_HIP(hipGetLastError());
}

void launch_cg_step(Grid& G, int nrhs, void* rv, int vb, hipStream_t s) {
    const size_t n = 3 * G.g0.nn;
    const int nb = dot_blocks(n);
    if (vb == 4)
        hipLaunchKernelGGL(k_cg_step<float>, dim3(nb, nrhs), dim3(kBlock), 0, s, G.r.as<double>(), G.ap.as<double>(),
                           n, n, G.scal.as<CgScal>(), G.partials.as<double>(), static_cast<float*>(rv));
    else
        hipLaunchKernelGGL(k_cg_step<double>, dim3(nb, nrhs), dim3(kBlock), 0, s, G.r.as<double>(), G.ap.as<double>(),
                           n, n, G.scal.as<CgScal>(), G.partials.as<double>(), static_cast<double*>(rv));
    LP_HIP(hipGetLastError());
}

void launch_cg_update(Grid& G, int nrhs, const void* z, int zb, bool xonly, hipStream_t s) {
    const size_t n = 3 * G.g0.nn;
    if (zb == 4)
        hipLaunchKernelGGL(k_cg_update<float>, dim3(grid_for(n), nrhs), dim3(kBlock), 0, s, G.x.as<double>(),
                           G.p.as<double>(), static_cast<const float*>(z), n, n, G.scal.as<CgScal>(), xonly ? 1 : 0);
    else
        hipLaunchKernelGGL(k_cg_update<double>, dim3(grid_for(n), nrhs), dim3(kBlock), 0, s, G.x.as<double>(),
                           G.p.as<double>(), static_cast<const double*>(z), n, n, G.scal.as<CgScal>(), xonly ? 1 : 0);
    LP_HIP(hipGetLastError());
}

void launch_cg32_step(Grid& G, int nrhs, float* r, hipStream_t s) {
    const size_t n = 3 * G.g0.nn;
    hipLaunchKernelGGL(k_cg32_step, dim3(dot_blocks(n), nrhs), dim3(kBlock), 0, s, r, G.ap32.as<float>(),
                       G.xlo32.as<float>(), G.p32.as<float>(), n, n, G.scal.as<CgScal>(), G.partials.as<double>());
    LP_HIP(hipGetLastError());
}

void launch_dot2_f32(Grid& G, int nrhs, const float* r, const float* z, const float* ap, hipStream_t s) {
    const size_t n = 3 * G.g0.nn;
    hipLaunchKernelGGL(k_dot2_f32, dim3(dot_blocks(n), nrhs), dim3(kBlock), 0, s, r, z, ap, n, n,
                       G.partials.as<double>(), G.scal.as<CgScal>());
    LP_HIP(hipGetLastError());
}

void launch_cg32_update(Grid& G, int nrhs, const float* z, bool first, hipStream_t s) {
    const size_t n = 3 * G.g0.nn;
    hipLaunchKernelGGL(k_cg32_update, dim3(grid_for(n), nrhs), dim3(kBlock), 0, s, G.p32.as<float>(), z, n, n,
                       G.scal.as<CgScal>(), first ? 1 : 0);
    LP_HIP(hipGetLastError());
}

void launch_cg32_fold(Grid& G, int nrhs, hipStream_t s) {
    const size_t n = 3 * G.g0.nn;
    hipLaunchKernelGGL(k_cg32_fold, dim3(grid_for(n), nrhs), dim3(kBlock), 0, s, G.x.as<double>(),
                       G.xlo32.as<float>(), n, n);
    LP_HIP(hipGetLastError());
}

void launch_cast_active(Grid& G, int nrhs, const double* r, float* rv, hipStream_t s) {
    const size_t n = 3 * G.g0.nn;
    hipLaunchKernelGGL(k_cast_active, dim3(grid_for(n), nrhs), dim3(kBlock), 0, s, r, rv, n, n, G.scal.as<CgScal>());
    LP_HIP(hipGetLastError());
}

void launch_cg_finalize(Grid& G, int nrhs, int stage, int nb, hipStream_t s) {
    hipLaunchKernelGGL(k_cg_finalize, dim3(nrhs), dim3(kBlock), 0, s, G.partials.as<double>(), nb,
                       G.scal.as<CgScal>(), stage, G.opt.rtol);
    LP_HIP(hipGetLastError());
}

void launch_warm(Grid& G, int nrhs, int which, hipStream_t s) {
    const size_t n = 3 * G.g0.nn;
    if (which == 0) {
        hipLaunchKernelGGL(k_warm_norm, dim3(dot_blocks(n), nrhs), dim3(kBlock), 0, s, G.b.as<double>(),
                           G.ap.as<double>(), n, n, G.scal.as<CgScal>(), G.partials.as<double>());
    } else {
        hipLaunchKernelGGL(k_warm_apply, dim3(grid_for(n), nrhs), dim3(kBlock), 0, s, G.x.as<double>(),
                           G.r.as<double>(), G.b.as<double>(), G.ap.as<double>(), n, n, G.scal.as<CgScal>());
    }
    LP_HIP(hipGetLastError());
}

}  // namespace lp
