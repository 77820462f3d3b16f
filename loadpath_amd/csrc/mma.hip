// The method of moving asymptotes on the device (Svanberg 1987; the 2007
// formulation), for n design variables and m <= kMmaMaxM constraints.
//
// Reference: loadpath/mma.py:31-214 (class MMA: update, tighten, _subsolv).
//
// The n-length work runs on the GPU; the m-sized pieces (the (m+1) x (m+1)
// Newton system of the subproblem, the dual variables y, z, lambda, mu,
// zeta, s and the interior-point control) run on the host. Every n-length
// pass is one kernel that also produces the reductions the host needs next
// (residual norm and max, g = sum P/(u - x) + Q/(x - l), the reduced Newton
// matrix, the step-to-boundary maxima) as per-block partials, summed on the
// host in a fixed order (deterministic).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "lp_internal.h"

namespace lp {

namespace {

constexpr int kMmaMaxM = 8;
constexpr int kMmaBlocks = 1024;
constexpr double kAlbefa = 0.1;
constexpr double kRaa0 = 1e-5;

struct SmallVec {
    double v[kMmaMaxM];
};

struct MmaPtr {
    double *x, *low, *upp, *xmin, *xmax, *alfa, *beta, *p0, *q0, *P, *Q;
    double *xs, *xsi, *eta, *bx, *bxsi, *beta_, *dx, *dxsi, *deta;
    long long n;
    int m;
};

__device__ __forceinline__ double sanitize_grad(double g) {
    if (isnan(g)) return 0.0;
    if (isinf(g)) return g > 0 ? 1e6 : -1e6;
    return g;
}

__device__ __forceinline__ void block_reduce(double* vals, int nv, bool is_max_from, int max_from, double* out) {
    // vals[k] for k < max_from are summed, k >= max_from maxed; one slot per
    // block in out[k * gridDim.x + blockIdx.x]
    __shared__ double sm[4][2 + 3 * kMmaMaxM + kMmaMaxM * (kMmaMaxM + 1) / 2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = 0; k < nv; ++k) {
        double v = vals[k];
        for (int o = 32; o > 0; o >>= 1) {
            const double t = __shfl_xor(v, o);
            v = (is_max_from && k >= max_from) ? fmax(v, t) : v + t;
        }
        if (lane == 0) sm[w][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < nv) {
        const int k = threadIdx.x;
        double v = sm[0][k];
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
            v = (is_max_from && k >= max_from) ? fmax(v, sm[i][k]) : v + sm[i][k];
        out[(size_t)k * gridDim.x + blockIdx.x] = v;
    }
}

// asymptotes (mma.py:76-87) and the sanitized, clipped design
__global__ void k_mma_asym(const double* __restrict__ xin, const double* __restrict__ xold1,
                           const double* __restrict__ xold2, MmaPtr S, int first, double asyinit, double incr,
                           double decr) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < S.n;
         j += (long long)gridDim.x * blockDim.x) {
        double x = xin[j];
        if (isnan(x)) x = 0.5;
        x = fmin(fmax(x, S.xmin[j]), S.xmax[j]);
        S.x[j] = x;
        const double span = S.xmax[j] - S.xmin[j];
        if (first) {
            S.low[j] = x - asyinit * span;
            S.upp[j] = x + asyinit * span;
        } else {
            const double t = (x - xold1[j]) * (xold1[j] - xold2[j]);
            const double f = t > 0 ? incr : (t < 0 ? decr : 1.0);
            const double lo = x - f * (xold1[j] - S.low[j]);
            const double up = x + f * (S.upp[j] - xold1[j]);
            S.low[j] = fmin(fmax(lo, x - 10.0 * span), x - 0.01 * span);
            S.upp[j] = fmin(fmax(up, x + 0.01 * span), x + 10.0 * span);
        }
    }
}

// move limits and the convex approximation (mma.py:88-107); partials of
// sum_j P_ij / (u - x) + Q_ij / (x - l)
__global__ void k_mma_approx(MmaPtr S, const double* __restrict__ df0, const double* __restrict__ df, double move,
                             double* partials) {
    double acc[kMmaMaxM];
    for (int i = 0; i < S.m; ++i) acc[i] = 0.0;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < S.n;
         j += (long long)gridDim.x * blockDim.x) {
        const double x = S.x[j], lo = S.low[j], up = S.upp[j];
        const double span = S.xmax[j] - S.xmin[j];
        S.alfa[j] = fmax(fmax(lo + kAlbefa * (x - lo), x - move * span), S.xmin[j]);
        S.beta[j] = fmin(fmin(up - kAlbefa * (up - x), x + move * span), S.xmax[j]);
        const double reg = kRaa0 / fmax(span, 1e-5);
        const double ux = up - x, xl = x - lo;
        const double u2 = ux * ux, l2 = xl * xl;
        const double g0 = sanitize_grad(df0[j]);
        const double gp = fmax(g0, 0.0), gm = fmax(-g0, 0.0);
        const double r0 = 0.001 * (gp + gm) + reg;
        S.p0[j] = (gp + r0) * u2;
        S.q0[j] = (gm + r0) * l2;
        for (int i = 0; i < S.m; ++i) {
            const double g = sanitize_grad(df[(size_t)i * S.n + j]);
            const double p = fmax(g, 0.0), q = fmax(-g, 0.0);
            const double r = 0.001 * (p + q) + reg;
            const double Pi = (p + r) * u2, Qi = (q + r) * l2;
            S.P[(size_t)i * S.n + j] = Pi;
            S.Q[(size_t)i * S.n + j] = Qi;
            acc[i] += Pi / ux + Qi / xl;
        }
    }
    block_reduce(acc, S.m, false, 0, partials);
}

__global__ void k_mma_start(MmaPtr S) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < S.n;
         j += (long long)gridDim.x * blockDim.x) {
        const double x = 0.5 * (S.alfa[j] + S.beta[j]);
        S.xs[j] = x;
        S.xsi[j] = fmax(1.0 / (x - S.alfa[j]), 1.0);
        S.eta[j] = fmax(1.0 / (S.beta[j] - x), 1.0);
    }
}

// trial state (x, xsi, eta) = base + step * direction (step < 0: the state as
// is), then the n-part of the perturbed KKT residual at lambda: partials of
// [sum r^2, g_1..g_m, max |r|]
__global__ void k_mma_resid(MmaPtr S, double step, SmallVec lam, double epsi, double* partials) {
    double acc[2 + kMmaMaxM];
    for (int k = 0; k < 2 + S.m; ++k) acc[k] = 0.0;
    double rsq = 0.0, rmax = 0.0;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < S.n;
         j += (long long)gridDim.x * blockDim.x) {
        double x, xsi, eta;
        if (step >= 0) {
            x = S.bx[j] + step * S.dx[j];
            xsi = S.bxsi[j] + step * S.dxsi[j];
            eta = S.beta_[j] + step * S.deta[j];
            S.xs[j] = x;
            S.xsi[j] = xsi;
            S.eta[j] = eta;
        } else {
            x = S.xs[j];
            xsi = S.xsi[j];
            eta = S.eta[j];
        }
        const double ux = S.upp[j] - x, xl = x - S.low[j];
        double plam = S.p0[j], qlam = S.q0[j];
        for (int i = 0; i < S.m; ++i) {
            const double Pi = S.P[(size_t)i * S.n + j], Qi = S.Q[(size_t)i * S.n + j];
            plam += Pi * lam.v[i];
            qlam += Qi * lam.v[i];
            acc[1 + i] += Pi / ux + Qi / xl;
        }
        const double r1 = plam / (ux * ux) - qlam / (xl * xl) - xsi + eta;
        const double r5 = xsi * (x - S.alfa[j]) - epsi;
        const double r6 = eta * (S.beta[j] - x) - epsi;
        rsq += r1 * r1 + r5 * r5 + r6 * r6;
        rmax = fmax(rmax, fmax(fabs(r1), fmax(fabs(r5), fabs(r6))));
    }
    acc[0] = rsq;
    acc[1 + S.m] = rmax;
    block_reduce(acc, 2 + S.m, true, 1 + S.m, partials);
}

// Newton step, n part (mma.py:158-172): partials of [g (m), GG (delx/diagx)
// (m), (GG / diagx) GG^T (upper triangle)]
__global__ void k_mma_newton(MmaPtr S, SmallVec lam, double epsi, double* partials) {
    constexpr int NV = 2 * kMmaMaxM + kMmaMaxM * (kMmaMaxM + 1) / 2;
    double acc[NV];
    const int m = S.m;
    const int nv = 2 * m + m * (m + 1) / 2;
    for (int k = 0; k < nv; ++k) acc[k] = 0.0;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < S.n;
         j += (long long)gridDim.x * blockDim.x) {
        const double x = S.xs[j], xsi = S.xsi[j], eta = S.eta[j];
        const double ux = S.upp[j] - x, xl = x - S.low[j];
        const double u2 = ux * ux, l2 = xl * xl;
        double plam = S.p0[j], qlam = S.q0[j];
        double gg[kMmaMaxM];
        for (int i = 0; i < m; ++i) {
            const double Pi = S.P[(size_t)i * S.n + j], Qi = S.Q[(size_t)i * S.n + j];
            plam += Pi * lam.v[i];
            qlam += Qi * lam.v[i];
            acc[i] += Pi / ux + Qi / xl;
            gg[i] = Pi / u2 - Qi / l2;
        }
        const double xa = x - S.alfa[j], bx = S.beta[j] - x;
        const double delx = plam / u2 - qlam / l2 - epsi / xa + epsi / bx;
        const double diagx = 2.0 * (plam / (u2 * ux) + qlam / (l2 * xl)) + xsi / xa + eta / bx;
        int t = 2 * m;
        for (int i = 0; i < m; ++i) {
            acc[m + i] += gg[i] * (delx / diagx);
            for (int k = i; k < m; ++k) acc[t++] += gg[i] * gg[k] / diagx;
        }
    }
    block_reduce(acc, nv, false, 0, partials);
}

// directions (mma.py:175-182) and the step-to-boundary maxima (mma.py:184-188)
__global__ void k_mma_dir(MmaPtr S, SmallVec lam, SmallVec dlam, double epsi, double* partials) {
    double mx[4] = {-1e300, -1e300, -1e300, -1e300};
    const int m = S.m;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < S.n;
         j += (long long)gridDim.x * blockDim.x) {
        const double x = S.xs[j], xsi = S.xsi[j], eta = S.eta[j];
        const double ux = S.upp[j] - x, xl = x - S.low[j];
        const double u2 = ux * ux, l2 = xl * xl;
        double plam = S.p0[j], qlam = S.q0[j], ggd = 0.0;
        for (int i = 0; i < m; ++i) {
            const double Pi = S.P[(size_t)i * S.n + j], Qi = S.Q[(size_t)i * S.n + j];
            plam += Pi * lam.v[i];
            qlam += Qi * lam.v[i];
            ggd += (Pi / u2 - Qi / l2) * dlam.v[i];
        }
        const double xa = x - S.alfa[j], bx = S.beta[j] - x;
        const double delx = plam / u2 - qlam / l2 - epsi / xa + epsi / bx;
        const double diagx = 2.0 * (plam / (u2 * ux) + qlam / (l2 * xl)) + xsi / xa + eta / bx;
        const double dx = -delx / diagx - ggd / diagx;
        const double dxsi = -xsi + epsi / xa - xsi * dx / xa;
        const double deta = -eta + epsi / bx + eta * dx / bx;
        S.dx[j] = dx;
        S.dxsi[j] = dxsi;
        S.deta[j] = deta;
        S.bx[j] = x;
        S.bxsi[j] = xsi;
        S.beta_[j] = eta;
        mx[0] = fmax(mx[0], -1.01 * dx / xa);
        mx[1] = fmax(mx[1], 1.01 * dx / bx);
        mx[2] = fmax(mx[2], -1.01 * dxsi / xsi);
        mx[3] = fmax(mx[3], -1.01 * deta / eta);
    }
    block_reduce(mx, 4, true, 0, partials);
}

// x_new = clip(finite ? x_sub : x, alfa, beta); xold2 <- xold1 <- x
__global__ void k_mma_finish(MmaPtr S, double* __restrict__ xnew, double* __restrict__ xold1,
                             double* __restrict__ xold2, int have_old1) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < S.n;
         j += (long long)gridDim.x * blockDim.x) {
        const double x = S.x[j];
        double v = S.xs[j];
        if (!isfinite(v)) v = x;
        xnew[j] = fmin(fmax(v, S.alfa[j]), S.beta[j]);
        xold2[j] = have_old1 ? xold1[j] : x;
        xold1[j] = x;
    }
}

__global__ void k_mma_tighten(const double* __restrict__ xold1, double* low, double* upp, long long n, double f) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (long long)gridDim.x * blockDim.x) {
        low[j] = xold1[j] - f * (xold1[j] - low[j]);
        upp[j] = xold1[j] + f * (upp[j] - xold1[j]);
    }
}

// m x m (plus one) dense solve with partial pivoting (the reduced Newton system)
bool solve_small(std::vector<double>& A, std::vector<double>& b, int n) {
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
        if (A[p * n + c] == 0.0) return false;
        if (p != c) {
            for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[p * n + k]);
            std::swap(b[c], b[p]);
        }
        for (int r = c + 1; r < n; ++r) {
            const double f = A[r * n + c] / A[c * n + c];
            for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        double s = b[c];
        for (int k = c + 1; k < n; ++k) s -= A[c * n + k] * b[k];
        b[c] = s / A[c * n + c];
    }
    return true;
}

}  // namespace

}  // namespace lp

using namespace lp;

struct lp_mma {
    long long n = 0;
    int m = 0;
    double move = 0.2, c = 1000.0, asyinit = 0.5, asyincr = 1.2, asydecr = 0.7;
    int iter = 0;
    bool have_old1 = false, have_asym = false;
    hipStream_t stream = nullptr;
    DBuf x, low, upp, xmin, xmax, alfa, beta, p0, q0, P, Q, xs, xsi, eta, bx, bxsi, beta_, dx, dxsi, deta;
    DBuf xold1, xold2, stage_in, stage_df0, stage_df, stage_out, partials;
    std::vector<double> lam;
    ~lp_mma() {
        if (stream) hipStreamDestroy(stream);
    }
    MmaPtr ptrs() {
        MmaPtr S;
        S.x = x.as<double>(); S.low = low.as<double>(); S.upp = upp.as<double>();
        S.xmin = xmin.as<double>(); S.xmax = xmax.as<double>();
        S.alfa = alfa.as<double>(); S.beta = beta.as<double>();
        S.p0 = p0.as<double>(); S.q0 = q0.as<double>(); S.P = P.as<double>(); S.Q = Q.as<double>();
        S.xs = xs.as<double>(); S.xsi = xsi.as<double>(); S.eta = eta.as<double>();
        S.bx = bx.as<double>(); S.bxsi = bxsi.as<double>(); S.beta_ = beta_.as<double>();
        S.dx = dx.as<double>(); S.dxsi = dxsi.as<double>(); S.deta = deta.as<double>();
        S.n = n;
        S.m = m;
        return S;
    }
    unsigned blocks() const { return (unsigned)std::min<long long>(kMmaBlocks, (n + kBlock - 1) / kBlock); }
    // sum (k < nsum) or max (k >= nsum) of nv per-block partial rows
    std::vector<double> fetch(int nv, int nsum) {
        const unsigned nb = blocks();
        std::vector<double> h((size_t)nv * nb);
        LP_HIP(hipMemcpyAsync(h.data(), partials.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        LP_HIP(hipStreamSynchronize(stream));
        std::vector<double> out(nv);
        for (int k = 0; k < nv; ++k) {
            double v = k < nsum ? 0.0 : -1e300;
            for (unsigned b = 0; b < nb; ++b) v = k < nsum ? v + h[(size_t)k * nb + b] : std::max(v, h[(size_t)k * nb + b]);
            out[k] = v;
        }
        return out;
    }
};

namespace {

struct Sub {
    // dual / slack variables of the subproblem (host, m-sized)
    std::vector<double> y, lam, mu, s;
    double z = 1.0, zet = 1.0;
};

// full residual norm and max at the current device state and dual vars:
// device n part + host m part (mma.py:128-143)
void residual(lp_mma& M, const Sub& D, double step, double epsi, const std::vector<double>& b, double& norm,
              double& rmax) {
    const int m = M.m;
    SmallVec lv{};
    for (int i = 0; i < m; ++i) lv.v[i] = D.lam[i];
    hipLaunchKernelGGL(k_mma_resid, dim3(M.blocks()), dim3(kBlock), 0, M.stream, M.ptrs(), step, lv, epsi,
                       M.partials.as<double>());
    LP_HIP(hipGetLastError());
    std::vector<double> r = M.fetch(2 + m, 1 + m);
    double sq = r[0], mx = r[1 + m];
    const double a0 = 1.0;
    double t = a0 - D.zet;   // a = 0
    sq += t * t;
    mx = std::max(mx, std::fabs(t));
    for (int i = 0; i < m; ++i) {
        const double r2 = M.c + D.y[i] - D.mu[i] - D.lam[i];   // d = 1
        const double r4 = r[1 + i] - D.y[i] + D.s[i] - b[i];   // a = 0
        const double r7 = D.mu[i] * D.y[i] - epsi;
        const double r9 = D.lam[i] * D.s[i] - epsi;
        sq += r2 * r2 + r4 * r4 + r7 * r7 + r9 * r9;
        mx = std::max(mx, std::max(std::max(std::fabs(r2), std::fabs(r4)), std::max(std::fabs(r7), std::fabs(r9))));
    }
    const double r8 = D.zet * D.z - epsi;
    sq += r8 * r8;
    mx = std::max(mx, std::fabs(r8));
    norm = std::sqrt(sq);
    rmax = mx;
}

// primal-dual interior point on the subproblem (mma.py:115-214)
void subsolve(lp_mma& M, const std::vector<double>& b) {
    const int m = M.m;
    hipStream_t s = M.stream;
    const unsigned nb = M.blocks();
    hipLaunchKernelGGL(k_mma_start, dim3(nb), dim3(kBlock), 0, s, M.ptrs());
    LP_HIP(hipGetLastError());
    Sub D;
    D.y.assign(m, 1.0);
    D.lam.assign(m, 1.0);
    D.mu.assign(m, std::max(1.0, 0.5 * M.c));
    D.s.assign(m, 1.0);
    D.z = 1.0;
    D.zet = 1.0;
    double epsi = 1.0;
    const double a0 = 1.0, d = 1.0;
    while (epsi > 1e-7) {
        double norm, rmax;
        residual(M, D, -1.0, epsi, b, norm, rmax);
        int it = 0;
        while (rmax > 0.9 * epsi && it < 200) {
            ++it;
            SmallVec lv{};
            for (int i = 0; i < m; ++i) lv.v[i] = D.lam[i];
            hipLaunchKernelGGL(k_mma_newton, dim3(nb), dim3(kBlock), 0, s, M.ptrs(), lv, epsi,
                               M.partials.as<double>());
            LP_HIP(hipGetLastError());
            const int nv = 2 * m + m * (m + 1) / 2;
            std::vector<double> red = M.fetch(nv, nv);
            // reduced (m + 1) system [diag(s/lam + 1/diagy) + GG D^-1 GG^T, a; a^T, -zet/z]
            const int K = m + 1;
            std::vector<double> A((size_t)K * K, 0.0), rhs(K, 0.0);
            std::vector<double> diagy(m), dely(m);
            int t = 2 * m;
            for (int i = 0; i < m; ++i)
                for (int k = i; k < m; ++k) {
                    A[i * K + k] += red[t];
                    if (k != i) A[k * K + i] += red[t];
                    ++t;
                }
            for (int i = 0; i < m; ++i) {
                diagy[i] = d + D.mu[i] / D.y[i];
                dely[i] = M.c + d * D.y[i] - D.lam[i] - epsi / D.y[i];
                const double dellam = red[i] - D.y[i] - b[i] + epsi / D.lam[i];   // a = 0
                A[i * K + i] += D.s[i] / D.lam[i] + 1.0 / diagy[i];
                rhs[i] = dellam + dely[i] / diagy[i] - red[m + i];
                A[i * K + m] = 0.0;   // a = 0
                A[m * K + i] = 0.0;
            }
            const double delz = a0 - epsi / D.z;   // a = 0
            A[m * K + m] = -D.zet / D.z;
            rhs[m] = delz;
            if (!solve_small(A, rhs, K)) throw std::runtime_error("mma: singular subproblem system");
            std::vector<double> dlam(rhs.begin(), rhs.begin() + m), dy(m), dmu(m), ds(m);
            const double dz = rhs[m];
            for (int i = 0; i < m; ++i) {
                dy[i] = -dely[i] / diagy[i] + dlam[i] / diagy[i];
                dmu[i] = -D.mu[i] + epsi / D.y[i] - D.mu[i] * dy[i] / D.y[i];
                ds[i] = -D.s[i] + epsi / D.lam[i] - D.s[i] * dlam[i] / D.lam[i];
            }
            const double dzet = -D.zet + epsi / D.z - D.zet * dz / D.z;
            SmallVec dl{};
            for (int i = 0; i < m; ++i) dl.v[i] = dlam[i];
            hipLaunchKernelGGL(k_mma_dir, dim3(nb), dim3(kBlock), 0, s, M.ptrs(), lv, dl, epsi,
                               M.partials.as<double>());
            LP_HIP(hipGetLastError());
            std::vector<double> mxs = M.fetch(4, 0);
            double stm = std::max(std::max(mxs[0], mxs[1]), std::max(mxs[2], mxs[3]));
            for (int i = 0; i < m; ++i) {
                stm = std::max(stm, -1.01 * dy[i] / D.y[i]);
                stm = std::max(stm, -1.01 * dlam[i] / D.lam[i]);
                stm = std::max(stm, -1.01 * dmu[i] / D.mu[i]);
                stm = std::max(stm, -1.01 * ds[i] / D.s[i]);
            }
            stm = std::max(stm, -1.01 * dz / D.z);
            stm = std::max(stm, -1.01 * dzet / D.zet);
            double steg = 1.0 / std::max(stm, 1.0);
            const Sub B = D;
            double newnorm = norm, newmax = rmax;
            for (int k = 0; k < 50; ++k) {
                for (int i = 0; i < m; ++i) {
                    D.y[i] = B.y[i] + steg * dy[i];
                    D.lam[i] = B.lam[i] + steg * dlam[i];
                    D.mu[i] = B.mu[i] + steg * dmu[i];
                    D.s[i] = B.s[i] + steg * ds[i];
                }
                D.z = B.z + steg * dz;
                D.zet = B.zet + steg * dzet;
                residual(M, D, steg, epsi, b, newnorm, newmax);
                if (newnorm < 2.0 * norm) break;
                steg /= 2.0;
            }
            norm = newnorm;
            rmax = newmax;
        }
        epsi *= 0.1;
    }
    M.lam = D.lam;
}

}  // namespace

extern "C" {

int lp_mma_create(lp_mma** out, long long n, int m, const double* xmin, const double* xmax, double move, double c,
                  double asyinit, double asyincr, double asydecr) {
    return api_call([&]() -> int {
        if (!out || n < 1 || m < 1 || m > kMmaMaxM || !xmin || !xmax || !(move > 0)) {
            set_error("lp_mma_create: bad argument (1 <= m <= 8)");
            return LP_ERR_ARG;
        }
        auto* M = new lp_mma();
        M->n = n;
        M->m = m;
        M->move = move;
        M->c = c;
        M->asyinit = asyinit;
        M->asyincr = asyincr;
        M->asydecr = asydecr;
        LP_HIP(hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking));
        const size_t nb = (size_t)n * 8;
        for (DBuf* b : {&M->x, &M->low, &M->upp, &M->xmin, &M->xmax, &M->alfa, &M->beta, &M->p0, &M->q0, &M->xs,
                        &M->xsi, &M->eta, &M->bx, &M->bxsi, &M->beta_, &M->dx, &M->dxsi, &M->deta, &M->xold1,
                        &M->xold2})
            b->alloc(nb);
        M->P.alloc(nb * m);
        M->Q.alloc(nb * m);
        M->partials.alloc((size_t)kMmaBlocks * (2 + 3 * kMmaMaxM + kMmaMaxM * (kMmaMaxM + 1) / 2) * 8);
        LP_HIP(hipMemcpyAsync(M->xmin.p, xmin, nb, hipMemcpyHostToDevice, M->stream));
        LP_HIP(hipMemcpyAsync(M->xmax.p, xmax, nb, hipMemcpyHostToDevice, M->stream));
        LP_HIP(hipStreamSynchronize(M->stream));
        M->lam.assign(m, 0.0);
        *out = M;
        return LP_OK;
    });
}

int lp_mma_destroy(lp_mma* M) {
    delete M;
    return LP_OK;
}

int lp_mma_set_move(lp_mma* M, double move) {
    return api_call([&]() -> int {
        if (!M || !(move > 0)) {
            set_error("lp_mma_set_move: bad argument");
            return LP_ERR_ARG;
        }
        M->move = move;
        return LP_OK;
    });
}

int lp_mma_tighten(lp_mma* M, double factor) {
    return api_call([&]() -> int {
        if (!M) return LP_ERR_ARG;
        if (M->have_old1 && M->have_asym) {
            hipLaunchKernelGGL(k_mma_tighten, dim3(M->blocks()), dim3(kBlock), 0, M->stream, M->xold1.as<double>(),
                               M->low.as<double>(), M->upp.as<double>(), M->n, factor);
            LP_HIP(hipGetLastError());
            LP_HIP(hipStreamSynchronize(M->stream));
        }
        return LP_OK;
    });
}

int lp_mma_update(lp_mma* M, const double* x, double f0, const double* df0, const double* f, const double* df,
                  double* xnew, int device_ptr) {
    (void)f0;   // the subproblem uses the objective's gradient only (mma.py:92-97)
    return api_call([&]() -> int {
        if (!M || !x || !df0 || !f || !df || !xnew) return LP_ERR_ARG;
        const long long n = M->n;
        const int m = M->m;
        hipStream_t s = M->stream;
        const size_t nb = (size_t)n * 8;
        const double* dxin = x;
        const double* ddf0 = df0;
        const double* ddf = df;
        double* dout = xnew;
        if (!device_ptr) {
            M->stage_in.alloc(nb);
            M->stage_df0.alloc(nb);
            M->stage_df.alloc(nb * m);
            M->stage_out.alloc(nb);
            LP_HIP(hipMemcpyAsync(M->stage_in.p, x, nb, hipMemcpyHostToDevice, s));
            LP_HIP(hipMemcpyAsync(M->stage_df0.p, df0, nb, hipMemcpyHostToDevice, s));
            LP_HIP(hipMemcpyAsync(M->stage_df.p, df, nb * m, hipMemcpyHostToDevice, s));
            dxin = M->stage_in.as<double>();
            ddf0 = M->stage_df0.as<double>();
            ddf = M->stage_df.as<double>();
            dout = M->stage_out.as<double>();
        }
        std::vector<double> fh(m);
        for (int i = 0; i < m; ++i) {
            double v = f[i];
            if (std::isnan(v)) v = 1e6;
            if (std::isinf(v)) v = v > 0 ? 1e6 : -1e6;
            fh[i] = v;
        }
        M->iter += 1;
        const unsigned bl = M->blocks();
        const int first = (M->iter <= 2 || !M->have_asym) ? 1 : 0;
        hipLaunchKernelGGL(k_mma_asym, dim3(bl), dim3(kBlock), 0, s, dxin, M->xold1.as<double>(),
                           M->xold2.as<double>(), M->ptrs(), first, M->asyinit, M->asyincr, M->asydecr);
        LP_HIP(hipGetLastError());
        M->have_asym = true;
        hipLaunchKernelGGL(k_mma_approx, dim3(bl), dim3(kBlock), 0, s, M->ptrs(), ddf0, ddf, M->move,
                           M->partials.as<double>());
        LP_HIP(hipGetLastError());
        std::vector<double> b = M->fetch(m, m);
        for (int i = 0; i < m; ++i) b[i] -= fh[i];
        subsolve(*M, b);
        hipLaunchKernelGGL(k_mma_finish, dim3(bl), dim3(kBlock), 0, s, M->ptrs(), dout, M->xold1.as<double>(),
                           M->xold2.as<double>(), M->have_old1 ? 1 : 0);
        LP_HIP(hipGetLastError());
        M->have_old1 = true;
        if (!device_ptr) LP_HIP(hipMemcpyAsync(xnew, dout, nb, hipMemcpyDeviceToHost, s));
        LP_HIP(hipStreamSynchronize(s));
        return LP_OK;
    });
}

int lp_mma_get(lp_mma* M, int what, double* out) {
    return api_call([&]() -> int {
        if (!M || !out) return LP_ERR_ARG;
        const size_t nb = (size_t)M->n * 8;
        if (what == 0 || what == 1) {
            LP_HIP(hipMemcpyAsync(out, (what == 0 ? M->low : M->upp).p, nb, hipMemcpyDeviceToHost, M->stream));
            LP_HIP(hipStreamSynchronize(M->stream));
        } else if (what == 2) {
            std::memcpy(out, M->lam.data(), M->m * sizeof(double));
        } else if (what == 3) {
            out[0] = (double)M->iter;
        } else {
            return LP_ERR_ARG;
        }
        return LP_OK;
    });
}

}  // extern "C"
