// p-mean von Mises stress aggregate over the fine grid's cells and the parts
// of its gradient that need no solve: the explicit derivative in the cell
// densities and the adjoint right-hand side dg/du; plus the element bilinear
// form a_c^T Ke b_c the adjoint sensitivity uses.
//
// Reference: constrained.py:607-664 (the non-averaged stress branch):
//   sigma_k = C B_k u_c / pitch at the sample points (centre or 8 corners),
//   vm = sqrt(max(sigma^T V sigma, 1e-30)), s = x^q vm / lim,
//   pm = sc (sum (s / sc)^P / nc)^(1/P), sc = max(s_max, 1e-30), g = kappa pm,
//   dg/ds = kappa (s / pm)^(P-1) / nc,
//   dg/dx (explicit) = sum_k dg/ds q x^(q-1) vm / lim,
//   dg/du = sum_k dg/ds x^q / (vm lim) (V sigma)^T C B_k, zero on fixed dofs;
//   and the adjoint term lambda_c^T Ke u_c (constrained.py:659).
//
// Cells run one per thread in FP64 (the aggregate is evaluated once per
// design iteration; its cost is a few field passes). The adjoint right-hand
// side is assembled node-centrically from a per-cell buffer: each node sums
// its <= 8 cells in a fixed order (no atomics, deterministic).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "lp_internal.h"

namespace lp {

namespace {

struct PnArgs {
    const double* u;         // (nrhs, 3, Nx, Ny, Nz)
    size_t ustride;
    const double* x;         // (nx, ny, nz)
    const uint8_t* counted;  // (nx, ny, nz) bit k = sample point k counted, or null = every point
    const double* cb;        // (npts, 6, 24)
    int npts;
    double p, q, lim, kappa, nc;
};

__device__ __forceinline__ void gather_cell(const double* __restrict__ u, const Geom& g, size_t cell, double ue[24]) {
    const int Z = (int)(cell % g.nz), Y = (int)((cell / g.nz) % g.ny), X = (int)(cell / ((size_t)g.nz * g.ny));
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const size_t node = ((size_t)(X + (b & 1)) * g.Ny + (Y + ((b >> 1) & 1))) * g.Nz + (Z + ((b >> 2) & 1));
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[3 * b + c] = u[c * g.nn + node];
    }
}

// sigma = CB_k ue, sv = V sigma; returns vm
__device__ __forceinline__ double point_stress(const double* __restrict__ cbk, const double ue[24], double sv[6]) {
    double s[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double a = 0.0;
#pragma unroll
        for (int j = 0; j < 24; ++j) a += cbk[i * 24 + j] * ue[j];
        s[i] = a;
    }
    sv[0] = s[0] - 0.5 * s[1] - 0.5 * s[2];
    sv[1] = -0.5 * s[0] + s[1] - 0.5 * s[2];
    sv[2] = -0.5 * s[0] - 0.5 * s[1] + s[2];
    sv[3] = 3.0 * s[3];
    sv[4] = 3.0 * s[4];
    sv[5] = 3.0 * s[5];
    double v2 = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) v2 += sv[i] * s[i];
    return sqrt(fmax(v2, 1e-30));
}

__device__ __forceinline__ double srel_of(const PnArgs& A, double xc, double vm, bool cnt) {
    return cnt ? pow(xc, A.q) * vm / A.lim : 0.0;
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sred[kBlock / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sred[i];
    return s;
}

// MODE 0: s_max per rhs (atomic max on the bits of non-negative doubles,
// order independent); MODE 1: block partial sums of (s / sc)^P
template <int MODE>
__global__ void __launch_bounds__(kBlock) k_pnorm_scan(PnArgs A, Geom g, const double* __restrict__ sc,
                                                       double* __restrict__ out) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    double acc = 0.0;
    if (cell < g.ncell) {
        double ue[24];
        gather_cell(A.u + rhs * A.ustride, g, cell, ue);
        const double xc = A.x[cell];
        const unsigned cbits = A.counted == nullptr ? 0xffu : A.counted[cell];
        for (int k = 0; k < A.npts; ++k) {
            double sv[6];
            const double vm = point_stress(A.cb + k * 144, ue, sv);
            const double sr = srel_of(A, xc, vm, (cbits >> k) & 1u);
            if (MODE == 0)
                acc = fmax(acc, sr);
            else
                acc += pow(sr / sc[rhs], A.p);
        }
    }
    if (MODE == 0) {
        for (int o = 32; o > 0; o >>= 1) acc = fmax(acc, __shfl_xor(acc, o));
        if ((threadIdx.x & 63) == 0)
            atomicMax(reinterpret_cast<unsigned long long*>(out + rhs),
                      (unsigned long long)__double_as_longlong(acc));
    } else {
        const double s = block_sum(acc);
        if (threadIdx.x == 0) out[(size_t)rhs * gridDim.x + blockIdx.x] = s;
    }
}

// one rhs: explicit gradient gx (per cell) and the element adjoint
// right-hand sides re (ncell, 24)
__global__ void __launch_bounds__(kBlock) k_pnorm_cell(PnArgs A, Geom g, int rhs, double pm, double* __restrict__ gx,
                                                       double* __restrict__ re) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= g.ncell) return;
    double ue[24];
    gather_cell(A.u + rhs * A.ustride, g, cell, ue);
    const double xc = A.x[cell];
    const unsigned cbits = A.counted == nullptr ? 0xffu : A.counted[cell];
    double r[24];
#pragma unroll
    for (int j = 0; j < 24; ++j) r[j] = 0.0;
    double ex = 0.0;
    for (int k = 0; k < A.npts; ++k) {
        const double* cbk = A.cb + k * 144;
        double sv[6];
        const double vm = point_stress(cbk, ue, sv);
        const bool cnt = (cbits >> k) & 1u;
        const double sr = srel_of(A, xc, vm, cnt);
        const double dg = A.kappa * pow(sr / fmax(pm, 1e-30), A.p - 1.0) / A.nc;
        ex += dg * (A.q * pow(fmax(xc, 1e-12), A.q - 1.0) * vm / A.lim);
        const double coef = cnt ? dg * (pow(xc, A.q) / (vm * A.lim)) : 0.0;
#pragma unroll
        for (int j = 0; j < 24; ++j) {
            double a = 0.0;
#pragma unroll
            for (int i = 0; i < 6; ++i) a += sv[i] * cbk[i * 24 + j];
            r[j] += coef * a;
        }
    }
    if (gx) gx[(size_t)rhs * g.ncell + cell] = ex;
#pragma unroll
    for (int j = 0; j < 24; ++j) re[cell * 24 + j] = r[j];
}

// node-centric assembly of the element right-hand sides (bincount over
// edof, constrained.py:651-653), zero on fixed dofs
__global__ void __launch_bounds__(kBlock) k_pnorm_node(const double* __restrict__ re, Geom g,
                                                       const uint8_t* __restrict__ fixed, double* __restrict__ out) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nn) return;
    const int z = (int)(n % g.Nz), y = (int)((n / g.Nz) % g.Ny), x = (int)(n / ((size_t)g.Nz * g.Ny));
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < 8; ++b) {
        const int cx = x - (b & 1), cy = y - ((b >> 1) & 1), cz = z - ((b >> 2) & 1);
        if (cx < 0 || cy < 0 || cz < 0 || cx >= g.nx || cy >= g.ny || cz >= g.nz) continue;
        const double* e = re + (((size_t)cx * g.ny + cy) * g.nz + cz) * 24 + 3 * b;
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += e[c];
    }
    const int f = fixed[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) out[c * g.nn + n] = ((f >> c) & 1) ? 0.0 : s[c];
}

// a_c^T Ke b_c per cell (Ke includes the pitch)
__global__ void __launch_bounds__(kBlock) k_element_product(const double* __restrict__ a, const double* __restrict__ b,
                                                            const double* __restrict__ ke, Geom g, size_t vstride,
                                                            double* __restrict__ out) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (cell >= g.ncell) return;
    double ae[24], be[24];
    gather_cell(a + rhs * vstride, g, cell, ae);
    gather_cell(b + rhs * vstride, g, cell, be);
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 24; ++i) {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < 24; ++j) t += ke[i * 24 + j] * be[j];
        acc += ae[i] * t;
    }
    out[(size_t)rhs * g.ncell + cell] = acc;
}

inline unsigned nblk(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

const void* upload(Grid& G, DBuf& stage, const void* p, size_t bytes, bool device_ptr) {
    if (device_ptr || p == nullptr) return p;
    stage.alloc(bytes);
    LP_HIP(hipMemcpyAsync(stage.p, p, bytes, hipMemcpyHostToDevice, G.stream));
    return stage.p;
}

}  // namespace

}  // namespace lp

namespace lp {
namespace {

// verify.nodal_stress (verify.py:307-337), node-centric: one thread per node
// sums, over its <= 8 cells in a fixed order, the cell's stress at that
// corner, modulus * (CB_k u_c - thermal C [1 1 1 0 0 0]), and divides by the
// number of cells it summed (the counted ones). The corner index k of the
// node in a cell is the same for every thread of a loop step, so the 6 x 24
// CB_k rows are uniform (scalar) loads.
struct NsArgs {
    int nx, ny, nz;
    const double* u;         // (nodes, 3) node-major
    const double* mod;       // (cells) or null: mod_scalar
    double mod_scalar;
    const double* thermal;   // (cells) or null
    const uint8_t* present;  // (cells) or null: every cell
    const uint8_t* counted;  // (cells) or null: the present cells, unmasked
    const double* cb;        // (8, 6, 24) corners in fem.CORNERS order, / pitch
    double cth[6];           // C [1 1 1 0 0 0] (E = 1)
    double* out;             // (nodes, 6)
    double* vm;              // (nodes) or null
};

__global__ void __launch_bounds__(256) k_nodal_stress(NsArgs A) {
    const size_t Ny = (size_t)A.ny + 1, Nz = (size_t)A.nz + 1;
    const size_t nn = ((size_t)A.nx + 1) * Ny * Nz;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const int X = (int)(n / (Ny * Nz)), Y = (int)((n / Nz) % Ny), Z = (int)(n % Nz);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int cnt = 0, touched = 0;
#pragma unroll 1
    for (int k = 0; k < 8; ++k) {
        // the cell whose corner k is this node
        const int cx = X - (k & 1), cy = Y - ((k >> 1) & 1), cz = Z - ((k >> 2) & 1);
        if (cx < 0 || cy < 0 || cz < 0 || cx >= A.nx || cy >= A.ny || cz >= A.nz) continue;
        const size_t c = ((size_t)cx * A.ny + cy) * A.nz + cz;
        if (A.present && !A.present[c]) continue;
        ++touched;
        const bool counts = !A.counted || A.counted[c];
        if (!counts) continue;
        ++cnt;
        const double m = A.mod ? A.mod[c] : A.mod_scalar;
        double ue[24];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const size_t node = ((size_t)(cx + (b & 1)) * Ny + (cy + ((b >> 1) & 1))) * Nz + (cz + ((b >> 2) & 1));
#pragma unroll
            for (int d = 0; d < 3; ++d) ue[3 * b + d] = A.u[3 * node + d];
        }
        const double* cb = A.cb + (size_t)k * 144;
        const double th = A.thermal ? A.thermal[c] : 0.0;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double t = 0;
#pragma unroll
            for (int j = 0; j < 24; ++j) t += cb[r * 24 + j] * ue[j];
            acc[r] += (t - th * A.cth[r]) * m;
        }
    }
    // the reference divides by the cells of the stress mask (at least 1)
    // when one is given, else by every cell at the node
    const double div = A.counted ? (double)(cnt > 0 ? cnt : 1) : (double)(touched > 0 ? touched : 1);
#pragma unroll
    for (int r = 0; r < 6; ++r) A.out[6 * n + r] = acc[r] / div;
    if (A.vm) {
        const double sx = acc[0] / div, sy = acc[1] / div, sz = acc[2] / div;
        const double txy = acc[3] / div, tyz = acc[4] / div, tzx = acc[5] / div;
        A.vm[n] = sqrt(0.5 * ((sx - sy) * (sx - sy) + (sy - sz) * (sy - sz) + (sz - sx) * (sz - sx)) +
                       3.0 * (txy * txy + tyz * tyz + tzx * tzx));
    }
}

// verify._MaterialAverage (verify.py:214-305), radii up to 3.3 cells: the
// samples of each live cell (8 quarter points, or the centre past 1.5
// cells), their tensors trilinear in the cell's nodal tensors and their
// weights fill * volume / samples; then each node's weighted mean over the
// samples within the radius, one thread per node over the cells in reach
// in a fixed order (its own tensor when there are none).
__global__ void __launch_bounds__(256) k_avg_samples(int nx, int ny, int nz, int ns, const double* __restrict__ sig,
                                                     const double* __restrict__ fill, double vol, double* ssig,
                                                     double* sw) {
    const size_t nc = (size_t)nx * ny * nz;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int X = (int)(c / ((size_t)ny * nz)), Y = (int)((c / nz) % ny), Z = (int)(c % nz);
    const double f = fill ? fill[c] : 1.0;
    sw[c] = f > 0.0 ? f * vol / ns : 0.0;
    if (!(f > 0.0)) return;
    const size_t Ny = (size_t)ny + 1, Nz = (size_t)nz + 1;
    for (int smp = 0; smp < ns; ++smp) {
        double q[3];
        for (int d = 0; d < 3; ++d) q[d] = ns == 8 ? 0.25 + 0.5 * ((smp >> d) & 1) : 0.5;
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int k = 0; k < 8; ++k) {
            const int kx = k & 1, ky = (k >> 1) & 1, kz = (k >> 2) & 1;
            const double w = (kx ? q[0] : 1.0 - q[0]) * (ky ? q[1] : 1.0 - q[1]) * (kz ? q[2] : 1.0 - q[2]);
            const size_t node = ((size_t)(X + kx) * Ny + (Y + ky)) * Nz + (Z + kz);
#pragma unroll
            for (int r = 0; r < 6; ++r) acc[r] += w * sig[6 * node + r];
        }
#pragma unroll
        for (int r = 0; r < 6; ++r) ssig[(c * ns + smp) * 6 + r] = acc[r];
    }
}

__global__ void __launch_bounds__(256) k_avg_ball(int nx, int ny, int nz, int ns, double pitch, double radius,
                                                  int reach, const double* __restrict__ sig,
                                                  const double* __restrict__ ssig, const double* __restrict__ sw,
                                                  double* out) {
    const size_t Ny = (size_t)ny + 1, Nz = (size_t)nz + 1;
    const size_t nn = ((size_t)nx + 1) * Ny * Nz;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const int X = (int)(n / (Ny * Nz)), Y = (int)((n / Nz) % Ny), Z = (int)(n % Nz);
    const double r2 = radius * radius;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    double tot = 0.0;
    for (int cx = max(X - reach - 1, 0); cx <= min(X + reach, nx - 1); ++cx)
        for (int cy = max(Y - reach - 1, 0); cy <= min(Y + reach, ny - 1); ++cy)
            for (int cz = max(Z - reach - 1, 0); cz <= min(Z + reach, nz - 1); ++cz) {
                const size_t c = ((size_t)cx * ny + cy) * nz + cz;
                const double w = sw[c];
                if (!(w > 0.0)) continue;
                for (int smp = 0; smp < ns; ++smp) {
                    const double qx = ns == 8 ? 0.25 + 0.5 * (smp & 1) : 0.5;
                    const double qy = ns == 8 ? 0.25 + 0.5 * ((smp >> 1) & 1) : 0.5;
                    const double qz = ns == 8 ? 0.25 + 0.5 * ((smp >> 2) & 1) : 0.5;
                    const double dx = (cx + qx - X) * pitch, dy = (cy + qy - Y) * pitch, dz = (cz + qz - Z) * pitch;
                    if (dx * dx + dy * dy + dz * dz > r2) continue;
                    const double* sp = ssig + (c * ns + smp) * 6;
#pragma unroll
                    for (int r = 0; r < 6; ++r) acc[r] += w * sp[r];
                    tot += w;
                }
            }
#pragma unroll
    for (int r = 0; r < 6; ++r) out[6 * n + r] = tot > 0.0 ? acc[r] / tot : sig[6 * n + r];
}

__device__ __forceinline__ int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// Past 3.3 cells (centre samples): blocks of blk^3 cells counted from the
// smallest live cell index along each axis (c0), one thread per block
// merging its live cells into the summed weight, the weighted centroid (in
// cell units) and the weighted mean of the cells' centre tensors.
__global__ void __launch_bounds__(256) k_avg_blocks(int nx, int ny, int nz, int blk, int c0x, int c0y, int c0z,
                                                    int bx, int by, int bz, const double* __restrict__ sig,
                                                    const double* __restrict__ fill, double vol, double* bw,
                                                    double* bpos, double* bsig) {
    const size_t nb = (size_t)bx * by * bz;
    const size_t b = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int KX = (int)(b / ((size_t)by * bz)), KY = (int)((b / bz) % by), KZ = (int)(b % bz);
    const size_t Ny = (size_t)ny + 1, Nz = (size_t)nz + 1;
    double tw = 0.0, pos[3] = {0, 0, 0}, acc[6] = {0, 0, 0, 0, 0, 0};
    for (int X = c0x + KX * blk; X < min(c0x + (KX + 1) * blk, nx); ++X)
        for (int Y = c0y + KY * blk; Y < min(c0y + (KY + 1) * blk, ny); ++Y)
            for (int Z = c0z + KZ * blk; Z < min(c0z + (KZ + 1) * blk, nz); ++Z) {
                const size_t c = ((size_t)X * ny + Y) * nz + Z;
                const double f = fill ? fill[c] : 1.0;
                if (!(f > 0.0)) continue;
                const double w = f * vol;
                tw += w;
                pos[0] += w * (X + 0.5);
                pos[1] += w * (Y + 0.5);
                pos[2] += w * (Z + 0.5);
                for (int k = 0; k < 8; ++k) {
                    const size_t node = ((size_t)(X + (k & 1)) * Ny + (Y + ((k >> 1) & 1))) * Nz + (Z + (k >> 2));
#pragma unroll
                    for (int r = 0; r < 6; ++r) acc[r] += (0.125 * w) * sig[6 * node + r];
                }
            }
    bw[b] = tw;
    const double inv = tw > 0.0 ? 1.0 / tw : 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) bpos[3 * b + d] = pos[d] * inv;
#pragma unroll
    for (int r = 0; r < 6; ++r) bsig[6 * b + r] = acc[r] * inv;
}

// each node's weighted mean over the block samples within the radius, the
// blocks visited in a fixed order over every block whose cells reach
// within rho of the node
__global__ void __launch_bounds__(256) k_avg_block_ball(int nx, int ny, int nz, int blk, int c0x, int c0y, int c0z,
                                                        int bx, int by, int bz, double pitch, double radius,
                                                        int reach, const double* __restrict__ sig,
                                                        const double* __restrict__ bw,
                                                        const double* __restrict__ bpos,
                                                        const double* __restrict__ bsig, double* out) {
    const size_t Ny = (size_t)ny + 1, Nz = (size_t)nz + 1;
    const size_t nn = ((size_t)nx + 1) * Ny * Nz;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const int X = (int)(n / (Ny * Nz)), Y = (int)((n / Nz) % Ny), Z = (int)(n % Nz);
    const double r2 = radius * radius;
    // block K holds cells c0 + K blk .. c0 + K blk + blk - 1; a cell within
    // reach of the node lies in [X - reach - 1, X + reach]
    const int lo[3] = {max(floor_div(X - reach - 1 - c0x, blk), 0), max(floor_div(Y - reach - 1 - c0y, blk), 0),
                       max(floor_div(Z - reach - 1 - c0z, blk), 0)};
    const int hi[3] = {min(floor_div(X + reach - c0x, blk), bx - 1), min(floor_div(Y + reach - c0y, blk), by - 1),
                       min(floor_div(Z + reach - c0z, blk), bz - 1)};
    double acc[6] = {0, 0, 0, 0, 0, 0};
    double tot = 0.0;
    for (int KX = lo[0]; KX <= hi[0]; ++KX)
        for (int KY = lo[1]; KY <= hi[1]; ++KY)
            for (int KZ = lo[2]; KZ <= hi[2]; ++KZ) {
                const size_t b = ((size_t)KX * by + KY) * bz + KZ;
                const double w = bw[b];
                if (!(w > 0.0)) continue;
                const double dx = (bpos[3 * b] - X) * pitch, dy = (bpos[3 * b + 1] - Y) * pitch,
                             dz = (bpos[3 * b + 2] - Z) * pitch;
                if (dx * dx + dy * dy + dz * dz > r2) continue;
#pragma unroll
                for (int r = 0; r < 6; ++r) acc[r] += w * bsig[6 * b + r];
                tot += w;
            }
#pragma unroll
    for (int r = 0; r < 6; ++r) out[6 * n + r] = tot > 0.0 ? acc[r] / tot : sig[6 * n + r];
}

// ---- the constrained loop's averaged stress reading ----------------------
// constrained._Averaged (constrained.py:146-252) as the stress branch reads it
// (constrained.py:517-605), on a box of cells: samples at each cell's quarter
// points (rho <= 1.5 cells) or its centre, sample stress C B(s) u_e / pitch on
// the cells that moved, sample weight w0 mat; each readable node averages the
// samples within the radius. Node-centric gathers throughout: the ball sum
// of a node visits the cells in reach, the transposed sum of a sample visits
// the nodes in reach, and the adjoint right-hand side of a node visits its
// <= 8 cells; no atomics, so the sums run in a fixed order.
struct AvArgs {
    int nx, ny, nz, ns, reach;
    double pitch, radius, w0;
    const double* u;         // (nodes, 3) one load case
    const double* mat;       // (cells) material share, max(xe, 0)
    const uint8_t* readable; // (nodes)
    const double* cb;        // (ns, 6, 24) C B / pitch
    double* ss;              // (cells, ns, 6) sample stress
    uint8_t* moved;          // (cells)
    double* node;            // (nodes, 10): sig[6], rho, dinv, s_rel (von Mises), s_rel (bond)
    double* ng;              // (nodes, 9): gsig dinv [6], (gsig . sig) dinv, grho dinv, grho rho dinv
    double* hw;              // (cells, ns, 6) v * (ball^T gsig dinv)
    double* gx;              // (cells) explicit dg/dxe
    double* rhs;             // (nodes, 3)
};

struct AvLaw {
    double p, q, lim, kappa, pm, nc;
    int bond;
    double lim_b, k2, m[6], Q[36];
};

__device__ __forceinline__ double av_off(int ns, int k, int d) { return ns == 8 ? 0.25 + 0.5 * ((k >> d) & 1) : 0.5; }

__global__ void __launch_bounds__(256) k_av_samples(AvArgs A) {
    const size_t nc = (size_t)A.nx * A.ny * A.nz;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int X = (int)(c / ((size_t)A.ny * A.nz)), Y = (int)((c / A.nz) % A.ny), Z = (int)(c % A.nz);
    const size_t Ny = (size_t)A.ny + 1, Nz = (size_t)A.nz + 1;
    double ue[24];
    bool mv = false;
    for (int j = 0; j < 8; ++j) {
        const size_t node = ((size_t)(X + (j & 1)) * Ny + (Y + ((j >> 1) & 1))) * Nz + (Z + (j >> 2));
        for (int d = 0; d < 3; ++d) {
            ue[3 * j + d] = A.u[3 * node + d];
            mv = mv || ue[3 * j + d] != 0.0;
        }
    }
    A.moved[c] = mv ? 1 : 0;
    for (int k = 0; k < A.ns; ++k)
        for (int i = 0; i < 6; ++i) {
            double s = 0.0;
            if (mv)
                for (int j = 0; j < 24; ++j) s += A.cb[(k * 6 + i) * 24 + j] * ue[j];
            A.ss[(c * A.ns + k) * 6 + i] = s;
        }
}

__device__ __forceinline__ void av_vm(const double* sig, double* sv, double& f) {
    const double h = 0.5 * (sig[0] + sig[1] + sig[2]);
    sv[0] = 1.5 * sig[0] - h;
    sv[1] = 1.5 * sig[1] - h;
    sv[2] = 1.5 * sig[2] - h;
    for (int i = 3; i < 6; ++i) sv[i] = 3.0 * sig[i];
    double t = 0.0;
    for (int i = 0; i < 6; ++i) t += sv[i] * sig[i];
    f = sqrt(fmax(t, 1e-30));
}

__device__ __forceinline__ void av_bond(const AvLaw& L, const double* sig, double* sb, double& f) {
    double sn = 0.0;
    for (int i = 0; i < 6; ++i) sn += sig[i] * L.m[i];
    sn = fmax(sn, 0.0);
    for (int i = 0; i < 6; ++i) {
        double t = 0.0;
        for (int j = 0; j < 6; ++j) t += sig[j] * L.Q[j * 6 + i];
        sb[i] = sn * L.m[i] + L.k2 * t;
    }
    double t = 0.0;
    for (int i = 0; i < 6; ++i) t += sb[i] * sig[i];
    f = sqrt(fmax(t, 1e-30));
}

__global__ void __launch_bounds__(256) k_av_nodes(AvArgs A, AvLaw L) {
    const size_t Ny = (size_t)A.ny + 1, Nz = (size_t)A.nz + 1;
    const size_t nn = ((size_t)A.nx + 1) * Ny * Nz;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    double* o = A.node + 10 * n;
    if (!A.readable[n]) {
        for (int i = 0; i < 10; ++i) o[i] = 0.0;
        return;
    }
    const int X = (int)(n / (Ny * Nz)), Y = (int)((n / Nz) % Ny), Z = (int)(n % Nz);
    const double r2 = A.radius * A.radius;
    double D = 0.0, R = 0.0, S[6] = {0, 0, 0, 0, 0, 0};
    for (int cx = max(X - A.reach - 1, 0); cx <= min(X + A.reach, A.nx - 1); ++cx)
        for (int cy = max(Y - A.reach - 1, 0); cy <= min(Y + A.reach, A.ny - 1); ++cy)
            for (int cz = max(Z - A.reach - 1, 0); cz <= min(Z + A.reach, A.nz - 1); ++cz) {
                const size_t c = ((size_t)cx * A.ny + cy) * A.nz + cz;
                const double mt = A.mat[c];
                const double v = A.w0 * mt;
                if (v == 0.0) continue;
                for (int k = 0; k < A.ns; ++k) {
                    const double dx = (cx + av_off(A.ns, k, 0) - X) * A.pitch,
                                 dy = (cy + av_off(A.ns, k, 1) - Y) * A.pitch,
                                 dz = (cz + av_off(A.ns, k, 2) - Z) * A.pitch;
                    if (dx * dx + dy * dy + dz * dz > r2) continue;
                    D += v;
                    R += v * mt;
                    const double* sp = A.ss + (c * A.ns + k) * 6;
                    for (int i = 0; i < 6; ++i) S[i] += v * sp[i];
                }
            }
    const double dinv = D > 0.0 ? 1.0 / D : 0.0;
    const double rho = R * dinv;
    double sig[6], sh[6], f;
    for (int i = 0; i < 6; ++i) sig[i] = S[i] * dinv;
    for (int i = 0; i < 6; ++i) o[i] = sig[i];
    o[6] = rho;
    o[7] = dinv;
    const double rq = pow(rho, L.q);
    av_vm(sig, sh, f);
    o[8] = rq * f / L.lim;
    o[9] = 0.0;
    if (L.bond) {
        av_bond(L, sig, sh, f);
        o[9] = rq * f / L.lim_b;
    }
}

__global__ void __launch_bounds__(256) k_av_nodegrad(AvArgs A, AvLaw L) {
    const size_t nn = ((size_t)A.nx + 1) * ((size_t)A.ny + 1) * ((size_t)A.nz + 1);
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    double* g = A.ng + 9 * n;
    if (!A.readable[n]) {
        for (int i = 0; i < 9; ++i) g[i] = 0.0;
        return;
    }
    const double* o = A.node + 10 * n;
    const double rho = o[6], dinv = o[7];
    const double rq = pow(rho, L.q);
    double gsig[6] = {0, 0, 0, 0, 0, 0}, grho = 0.0, sh[6], f;
    for (int part = 0; part < (L.bond ? 2 : 1); ++part) {
        double lim;
        if (part == 0) {
            av_vm(o, sh, f);
            lim = L.lim;
        } else {
            av_bond(L, o, sh, f);
            lim = L.lim_b;
        }
        const double sr = rq * f / lim;
        const double a = L.kappa * pow(sr / fmax(L.pm, 1e-30), L.p - 1.0) / L.nc;
        const double s = a * rq / (f * lim);
        for (int i = 0; i < 6; ++i) gsig[i] += s * sh[i];
        grho += a * L.q * pow(fmax(rho, 1e-12), L.q - 1.0) * f / lim;
    }
    double gs = 0.0;
    for (int i = 0; i < 6; ++i) {
        g[i] = gsig[i] * dinv;
        gs += gsig[i] * o[i];
    }
    g[6] = gs * dinv;
    g[7] = grho * dinv;
    g[8] = grho * rho * dinv;
}

__global__ void __launch_bounds__(256) k_av_cells(AvArgs A) {
    const size_t nc = (size_t)A.nx * A.ny * A.nz;
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int X = (int)(c / ((size_t)A.ny * A.nz)), Y = (int)((c / A.nz) % A.ny), Z = (int)(c % A.nz);
    const size_t Ny = (size_t)A.ny + 1, Nz = (size_t)A.nz + 1;
    const double r2 = A.radius * A.radius;
    const double mt = A.mat[c];
    const double v = A.w0 * mt;
    const bool mv = A.moved[c] != 0;
    double t = 0.0;
    for (int k = 0; k < A.ns; ++k) {
        const double px = X + av_off(A.ns, k, 0), py = Y + av_off(A.ns, k, 1), pz = Z + av_off(A.ns, k, 2);
        double hs[6] = {0, 0, 0, 0, 0, 0}, T1 = 0.0, T2 = 0.0, T3 = 0.0;
        // nodes within rho of a point in the cell lie in [X - reach, X + 1 + reach]
        for (int nx_ = max(X - A.reach, 0); nx_ <= min(X + 1 + A.reach, A.nx); ++nx_)
            for (int ny_ = max(Y - A.reach, 0); ny_ <= min(Y + 1 + A.reach, A.ny); ++ny_)
                for (int nz_ = max(Z - A.reach, 0); nz_ <= min(Z + 1 + A.reach, A.nz); ++nz_) {
                    const double dx = (px - nx_) * A.pitch, dy = (py - ny_) * A.pitch, dz = (pz - nz_) * A.pitch;
                    if (dx * dx + dy * dy + dz * dz > r2) continue;
                    const size_t n = ((size_t)nx_ * Ny + ny_) * Nz + nz_;
                    if (!A.readable[n]) continue;
                    const double* g = A.ng + 9 * n;
                    for (int i = 0; i < 6; ++i) hs[i] += g[i];
                    T1 += g[6];
                    T2 += g[7];
                    T3 += g[8];
                }
        const double* sp = A.ss + (c * A.ns + k) * 6;
        double sdot = 0.0;
        for (int i = 0; i < 6; ++i) sdot += sp[i] * hs[i];
        t += (sdot - T1) + (2.0 * mt * T2 - T3);
        double* h = A.hw + (c * A.ns + k) * 6;
        for (int i = 0; i < 6; ++i) h[i] = mv ? v * hs[i] : 0.0;
    }
    A.gx[c] = A.w0 * t;
}

__global__ void __launch_bounds__(256) k_av_rhs(AvArgs A) {
    const size_t Ny = (size_t)A.ny + 1, Nz = (size_t)A.nz + 1;
    const size_t nn = ((size_t)A.nx + 1) * Ny * Nz;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nn) return;
    const int X = (int)(n / (Ny * Nz)), Y = (int)((n / Nz) % Ny), Z = (int)(n % Nz);
    double r[3] = {0, 0, 0};
    for (int j = 0; j < 8; ++j) {
        const int cx = X - (j & 1), cy = Y - ((j >> 1) & 1), cz = Z - (j >> 2);
        if (cx < 0 || cy < 0 || cz < 0 || cx >= A.nx || cy >= A.ny || cz >= A.nz) continue;
        const size_t c = ((size_t)cx * A.ny + cy) * A.nz + cz;
        if (!A.moved[c]) continue;
        for (int k = 0; k < A.ns; ++k) {
            const double* h = A.hw + (c * A.ns + k) * 6;
            for (int d = 0; d < 3; ++d) {
                double s = 0.0;
                for (int i = 0; i < 6; ++i) s += h[i] * A.cb[(k * 6 + i) * 24 + 3 * j + d];
                r[d] += s;
            }
        }
    }
    for (int d = 0; d < 3; ++d) A.rhs[3 * n + d] = r[d];
}

}  // namespace
}  // namespace lp

using namespace lp;

struct lp_grid : public Grid {};

extern "C" {

int lp_stress_pnorm(lp_grid* g, int nrhs, const double* u, const double* x, const uint8_t* counted,
                    const double* cb, int npts, double p, double q, double lim, double kappa, double nc, double* pm,
                    double* smax, double* rhs, double* gx, int device_ptr) {
    return api_call([&]() -> int {
        if (!g || nrhs < 1 || !u || !x || !cb || !pm || !smax || (npts != 1 && npts != 8) || !(p >= 1.0) ||
            !(q > 0.0) || !(lim > 0.0) || !(nc > 0.0)) {
            set_error("lp_stress_pnorm: bad argument");
            return LP_ERR_ARG;
        }
        Grid& G = *g;
        const Geom& g0 = G.g0;
        hipStream_t s = G.stream;
        const size_t ustride = 3 * g0.nn;
        DBuf su, sx, sc_, scb, sm, part, re, out_r, out_g;
        PnArgs A;
        A.u = static_cast<const double*>(upload(G, su, u, ustride * nrhs * 8, device_ptr));
        A.ustride = ustride;
        A.x = static_cast<const double*>(upload(G, sx, x, g0.ncell * 8, device_ptr));
        A.counted = static_cast<const uint8_t*>(upload(G, sc_, counted, g0.ncell, device_ptr));
        A.cb = static_cast<const double*>(upload(G, scb, cb, (size_t)npts * 144 * 8, false));
        A.npts = npts;
        A.p = p;
        A.q = q;
        A.lim = lim;
        A.kappa = kappa;
        A.nc = nc;
        // s_max, then sc = max(s_max, 1e-30) and the sums of (s / sc)^P
        sm.alloc(2 * nrhs * sizeof(double));
        LP_HIP(hipMemsetAsync(sm.p, 0, nrhs * sizeof(double), s));
        const dim3 grid(nblk(g0.ncell), nrhs);
        hipLaunchKernelGGL(k_pnorm_scan<0>, grid, dim3(kBlock), 0, s, A, g0, (const double*)nullptr,
                           sm.as<double>());
        LP_HIP(hipGetLastError());
        std::vector<double> h(2 * nrhs);
        LP_HIP(hipMemcpyAsync(h.data(), sm.p, nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        LP_HIP(hipStreamSynchronize(s));
        for (int r = 0; r < nrhs; ++r) {
            smax[r] = h[r];
            h[nrhs + r] = std::max(h[r], 1e-30);
        }
        LP_HIP(hipMemcpyAsync(sm.as<double>() + nrhs, h.data() + nrhs, nrhs * sizeof(double),
                              hipMemcpyHostToDevice, s));
        part.alloc((size_t)nrhs * grid.x * sizeof(double));
        hipLaunchKernelGGL(k_pnorm_scan<1>, grid, dim3(kBlock), 0, s, A, g0, (const double*)(sm.as<double>() + nrhs),
                           part.as<double>());
        LP_HIP(hipGetLastError());
        std::vector<double> hp((size_t)nrhs * grid.x);
        LP_HIP(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        LP_HIP(hipStreamSynchronize(s));
        for (int r = 0; r < nrhs; ++r) {
            double sum = 0.0;
            for (unsigned b = 0; b < grid.x; ++b) sum += hp[(size_t)r * grid.x + b];
            const double scr = h[nrhs + r];
            pm[r] = scr * std::pow(sum / nc, 1.0 / p);
        }
        if (!rhs && !gx) return LP_OK;
        // per rhs: explicit gradient and the node-assembled adjoint right-hand side
        re.alloc(g0.ncell * 24 * sizeof(double));
        double* drhs = rhs;
        double* dgx = gx;
        if (!device_ptr) {
            if (rhs) drhs = static_cast<double*>(out_r.alloc(ustride * nrhs * 8), out_r.p);
            if (gx) dgx = static_cast<double*>(out_g.alloc(g0.ncell * nrhs * 8), out_g.p);
        }
        for (int r = 0; r < nrhs; ++r) {
            hipLaunchKernelGGL(k_pnorm_cell, dim3(nblk(g0.ncell)), dim3(kBlock), 0, s, A, g0, r, pm[r], dgx,
                               re.as<double>());
            LP_HIP(hipGetLastError());
            if (drhs) {
                hipLaunchKernelGGL(k_pnorm_node, dim3(nblk(g0.nn)), dim3(kBlock), 0, s, (const double*)re.as<double>(),
                                   g0, G.fixed.as<uint8_t>(), drhs + r * ustride);
                LP_HIP(hipGetLastError());
            }
        }
        if (!device_ptr) {
            if (rhs) LP_HIP(hipMemcpyAsync(rhs, drhs, ustride * nrhs * 8, hipMemcpyDeviceToHost, s));
            if (gx) LP_HIP(hipMemcpyAsync(gx, dgx, g0.ncell * nrhs * 8, hipMemcpyDeviceToHost, s));
        }
        LP_HIP(hipStreamSynchronize(s));
        return LP_OK;
    });
}

int lp_element_product(lp_grid* g, int nrhs, const double* a, const double* b, double* out, int device_ptr) {
    return api_call([&]() -> int {
        if (!g || nrhs < 1 || !a || !b || !out) {
            set_error("lp_element_product: bad argument");
            return LP_ERR_ARG;
        }
        Grid& G = *g;
        const Geom& g0 = G.g0;
        hipStream_t s = G.stream;
        const size_t n = 3 * g0.nn;
        DBuf sa, sb, so;
        const double* da = static_cast<const double*>(upload(G, sa, a, n * nrhs * 8, device_ptr));
        const double* db = static_cast<const double*>(upload(G, sb, b, n * nrhs * 8, device_ptr));
        double* dout = device_ptr ? out : static_cast<double*>((so.alloc(g0.ncell * nrhs * 8), so.p));
        hipLaunchKernelGGL(k_element_product, dim3(nblk(g0.ncell), nrhs), dim3(kBlock), 0, s, da, db,
                           G.ke64.as<double>(), g0, n, dout);
        LP_HIP(hipGetLastError());
        if (!device_ptr) LP_HIP(hipMemcpyAsync(out, dout, g0.ncell * nrhs * 8, hipMemcpyDeviceToHost, s));
        LP_HIP(hipStreamSynchronize(s));
        return LP_OK;
    });
}

int lp_nodal_stress(int nx, int ny, int nz, double pitch, double nu, int element, const double* u,
                    const double* modulus, double modulus_scalar, const double* thermal, const uint8_t* present,
                    const uint8_t* counted, double* out, double* vm, int device_ptr) {
    return api_call([&]() -> int {
        if (nx < 1 || ny < 1 || nz < 1 || !(pitch > 0.0) || !(nu > -1.0 && nu < 0.5) || (element != 0 && element != 1) ||
            !u || !out) {
            set_error("lp_nodal_stress: bad argument");
            return LP_ERR_ARG;
        }
        const size_t nc = (size_t)nx * ny * nz;
        const size_t nn = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
        // C B at the corners (lp_hex8_cb: x outermost) into fem.CORNERS order
        // (bit0 = x), divided by the pitch
        std::vector<double> cbx(8 * 144), cb(8 * 144);
        hex8_cb_host(nu, element, 1, cbx.data());
        for (int k = 0; k < 8; ++k) {
            const int src = 4 * (k & 1) + 2 * ((k >> 1) & 1) + ((k >> 2) & 1);
            for (int i = 0; i < 144; ++i) cb[k * 144 + i] = cbx[src * 144 + i] / pitch;
        }
        const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = 1.0 / (2.0 * (1.0 + nu));
        NsArgs A{};
        A.nx = nx;
        A.ny = ny;
        A.nz = nz;
        for (int r = 0; r < 6; ++r) A.cth[r] = r < 3 ? 3.0 * lam + 2.0 * mu : 0.0;
        A.mod_scalar = modulus_scalar;
        DBuf du, dm, dt, dp, dc, dcb, dout, dvm;
        auto up = [&](DBuf& b, const void* h, size_t bytes) -> const void* {
            if (!h) return nullptr;
            if (device_ptr) return h;
            b.alloc(bytes);
            LP_HIP(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
            return b.p;
        };
        A.u = static_cast<const double*>(up(du, u, nn * 3 * 8));
        A.mod = static_cast<const double*>(up(dm, modulus, nc * 8));
        A.thermal = static_cast<const double*>(up(dt, thermal, nc * 8));
        A.present = static_cast<const uint8_t*>(up(dp, present, nc));
        A.counted = static_cast<const uint8_t*>(up(dc, counted, nc));
        dcb.alloc(cb.size() * 8);
        LP_HIP(hipMemcpy(dcb.p, cb.data(), cb.size() * 8, hipMemcpyHostToDevice));
        A.cb = dcb.as<double>();
        if (device_ptr) {
            A.out = out;
            A.vm = vm;
        } else {
            dout.alloc(nn * 6 * 8);
            A.out = dout.as<double>();
            if (vm) {
                dvm.alloc(nn * 8);
                A.vm = dvm.as<double>();
            }
        }
        hipLaunchKernelGGL(k_nodal_stress, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, nullptr, A);
        LP_HIP(hipGetLastError());
        if (!device_ptr) {
            LP_HIP(hipMemcpy(out, A.out, nn * 6 * 8, hipMemcpyDeviceToHost));
            if (vm) LP_HIP(hipMemcpy(vm, A.vm, nn * 8, hipMemcpyDeviceToHost));
        } else {
            LP_HIP(hipDeviceSynchronize());
        }
        return LP_OK;
    });
}

int lp_material_average(int nx, int ny, int nz, double pitch, double radius, const double* fill, const double* sig,
                        double* out) {
    return api_call([&]() -> int {
        if (nx < 1 || ny < 1 || nz < 1 || !(pitch > 0.0) || !(radius > 0.0) || !sig || !out) {
            set_error("lp_material_average: bad argument");
            return LP_ERR_ARG;
        }
        const double rho = radius / pitch;
        const double rblk = std::ceil(rho / 3.3);
        if (!(rho < 1e6) || rblk > 1e6) {
            set_error("lp_material_average: radius out of range");
            return LP_ERR_ARG;
        }
        const int ns = rho <= 1.5 ? 8 : 1;
        const int reach = (int)std::ceil(rho);
        const int blk = (int)rblk;
        const size_t nc = (size_t)nx * ny * nz;
        const size_t nn = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
        DBuf dsig, dfill, dss, dsw, dout;
        dsig.alloc(nn * 6 * 8);
        LP_HIP(hipMemcpy(dsig.p, sig, nn * 6 * 8, hipMemcpyHostToDevice));
        if (fill) {
            dfill.alloc(nc * 8);
            LP_HIP(hipMemcpy(dfill.p, fill, nc * 8, hipMemcpyHostToDevice));
        }
        if (blk > 1) {
            // the smallest live cell index along each axis: blocks are
            // counted from the smallest sample coordinate (verify.py:266)
            int c0[3] = {nx, ny, nz};
            for (size_t c = 0; c < nc; ++c) {
                if (fill && !(fill[c] > 0.0)) continue;
                const int idx[3] = {(int)(c / ((size_t)ny * nz)), (int)((c / nz) % ny), (int)(c % nz)};
                for (int d = 0; d < 3; ++d) c0[d] = std::min(c0[d], idx[d]);
            }
            if (c0[0] == nx) {  // no material: every node keeps its own tensor
                std::memcpy(out, sig, nn * 6 * 8);
                return LP_OK;
            }
            const int bx = (nx - c0[0] + blk - 1) / blk, by = (ny - c0[1] + blk - 1) / blk,
                      bz = (nz - c0[2] + blk - 1) / blk;
            const size_t nb = (size_t)bx * by * bz;
            DBuf dbw, dbp, dbs;
            dbw.alloc(nb * 8);
            dbp.alloc(nb * 3 * 8);
            dbs.alloc(nb * 6 * 8);
            dout.alloc(nn * 6 * 8);
            hipLaunchKernelGGL(k_avg_blocks, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, nullptr, nx, ny, nz,
                               blk, c0[0], c0[1], c0[2], bx, by, bz, dsig.as<double>(),
                               fill ? dfill.as<double>() : (const double*)nullptr, pitch * pitch * pitch,
                               dbw.as<double>(), dbp.as<double>(), dbs.as<double>());
            LP_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_avg_block_ball, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, nullptr, nx, ny,
                               nz, blk, c0[0], c0[1], c0[2], bx, by, bz, pitch, radius, reach, dsig.as<double>(),
                               dbw.as<double>(), dbp.as<double>(), dbs.as<double>(), dout.as<double>());
            LP_HIP(hipGetLastError());
            LP_HIP(hipMemcpy(out, dout.p, nn * 6 * 8, hipMemcpyDeviceToHost));
            return LP_OK;
        }
        dss.alloc(nc * ns * 6 * 8);
        dsw.alloc(nc * 8);
        dout.alloc(nn * 6 * 8);
        hipLaunchKernelGGL(k_avg_samples, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, nullptr, nx, ny, nz, ns,
                           dsig.as<double>(), fill ? dfill.as<double>() : (const double*)nullptr,
                           pitch * pitch * pitch, dss.as<double>(), dsw.as<double>());
        LP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_avg_ball, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, nullptr, nx, ny, nz, ns,
                           pitch, radius, reach, dsig.as<double>(), dss.as<double>(), dsw.as<double>(),
                           dout.as<double>());
        LP_HIP(hipGetLastError());
        LP_HIP(hipMemcpy(out, dout.p, nn * 6 * 8, hipMemcpyDeviceToHost));
        return LP_OK;
    });
}

int lp_stress_averaged(int nx, int ny, int nz, double pitch, double nu, int element, double radius, int ncase,
                       const double* u, const double* xe, const uint8_t* readable, double p, double q,
                       double limit, const double* bond, double* kappa, int it, double* pm, double* smax,
                       double* rhs, double* gx) {
    return api_call([&]() -> int {
        if (nx < 1 || ny < 1 || nz < 1 || !(pitch > 0.0) || !(radius > 0.0) || ncase < 1 || !u || !xe ||
            !readable || !(limit > 0.0) || !(p > 0.0) || !kappa || !pm || !smax || (element != 0 && element != 1) ||
            !(nu > -1.0 && nu < 0.5) || (bond && !(bond[0] > 0.0)) || radius / pitch > 1e4 || (!rhs != !gx)) {
            set_error("lp_stress_averaged: bad argument");
            return LP_ERR_ARG;
        }
        const double rho = radius / pitch;
        const int ns = rho <= 1.5 ? 8 : 1;
        const size_t nc = (size_t)nx * ny * nz;
        const size_t nn = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
        std::vector<double> cbh((size_t)ns * 6 * 24);
        hex8_cb_host(nu, element, ns == 8 ? 2 : 0, cbh.data());
        for (double& v : cbh) v /= pitch;
        AvLaw L{};
        L.p = p;
        L.q = q;
        L.lim = limit;
        L.bond = bond ? 1 : 0;
        if (bond) {
            // constrained._bond_form (constrained.py:262-273): the traction on
            // the layer plane t (3 x 6), m = n t, Q = t^T t - m m^T
            double nrm[3] = {bond[1], bond[2], bond[3]};
            const double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            if (!(len > 0.0)) {
                set_error("lp_stress_averaged: bond normal is zero");
                return LP_ERR_ARG;
            }
            for (double& v : nrm) v /= len;
            const double t[3][6] = {{nrm[0], 0, 0, nrm[1], 0, nrm[2]},
                                    {0, nrm[1], 0, nrm[0], nrm[2], 0},
                                    {0, 0, nrm[2], 0, nrm[1], nrm[0]}};
            for (int i = 0; i < 6; ++i) L.m[i] = nrm[0] * t[0][i] + nrm[1] * t[1][i] + nrm[2] * t[2][i];
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j)
                    L.Q[i * 6 + j] = t[0][i] * t[0][j] + t[1][i] * t[1][j] + t[2][i] * t[2][j] - L.m[i] * L.m[j];
            L.lim_b = limit * bond[0];
            L.k2 = bond[4] * bond[4];
        }
        size_t nread = 0;
        for (size_t i = 0; i < nn; ++i) nread += readable[i] ? 1 : 0;
        L.nc = (double)(std::max<size_t>(nread, 1) * (bond ? 2 : 1));
        std::vector<double> mat(nc);
        for (size_t i = 0; i < nc; ++i) mat[i] = std::max(xe[i], 0.0);
        DBuf du, dmat, drd, dcb, dss, dmv, dnode, dng, dhw, dgx, drhs;
        du.alloc(nn * 3 * 8);
        dmat.alloc(nc * 8);
        drd.alloc(nn);
        dcb.alloc(cbh.size() * 8);
        dss.alloc(nc * ns * 6 * 8);
        dmv.alloc(nc);
        dnode.alloc(nn * 10 * 8);
        dng.alloc(nn * 9 * 8);
        dhw.alloc(nc * ns * 6 * 8);
        dgx.alloc(nc * 8);
        drhs.alloc(nn * 3 * 8);
        LP_HIP(hipMemcpy(dmat.p, mat.data(), nc * 8, hipMemcpyHostToDevice));
        LP_HIP(hipMemcpy(drd.p, readable, nn, hipMemcpyHostToDevice));
        LP_HIP(hipMemcpy(dcb.p, cbh.data(), cbh.size() * 8, hipMemcpyHostToDevice));
        AvArgs A{};
        A.nx = nx;
        A.ny = ny;
        A.nz = nz;
        A.ns = ns;
        A.reach = (int)std::ceil(rho);
        A.pitch = pitch;
        A.radius = radius;
        A.w0 = pitch * pitch * pitch / ns;
        A.u = du.as<double>();
        A.mat = dmat.as<double>();
        A.readable = drd.as<uint8_t>();
        A.cb = dcb.as<double>();
        A.ss = dss.as<double>();
        A.moved = dmv.as<uint8_t>();
        A.node = dnode.as<double>();
        A.ng = dng.as<double>();
        A.hw = dhw.as<double>();
        A.gx = dgx.as<double>();
        A.rhs = drhs.as<double>();
        const dim3 gc((unsigned)((nc + 255) / 256)), gn((unsigned)((nn + 255) / 256)), blk(256);
        std::vector<double> node(nn * 10);
        for (int cs = 0; cs < ncase; ++cs) {
            LP_HIP(hipMemcpy(du.p, u + (size_t)cs * nn * 3, nn * 3 * 8, hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_av_samples, gc, blk, 0, nullptr, A);
            hipLaunchKernelGGL(k_av_nodes, gn, blk, 0, nullptr, A, L);
            LP_HIP(hipGetLastError());
            LP_HIP(hipMemcpy(node.data(), dnode.p, nn * 10 * 8, hipMemcpyDeviceToHost));
            // s_max and the P-mean over the readable nodes (constrained.py:562-565)
            double mx = 0.0;
            bool any = false;
            for (size_t i = 0; i < nn; ++i)
                if (readable[i]) {
                    mx = any ? std::max(mx, node[10 * i + 8]) : node[10 * i + 8];
                    any = true;
                    if (bond) mx = std::max(mx, node[10 * i + 9]);
                }
            const double sc = std::max(mx, 1e-30);
            double sum = 0.0;
            for (size_t i = 0; i < nn; ++i)
                if (readable[i]) sum += std::pow(node[10 * i + 8] / sc, p);
            if (bond)
                for (size_t i = 0; i < nn; ++i)
                    if (readable[i]) sum += std::pow(node[10 * i + 9] / sc, p);
            const double pmv = sc * std::pow(sum / L.nc, 1.0 / p);
            const double kr = mx / std::max(pmv, 1e-30);
            kappa[cs] = it > 1 ? 0.5 * (kappa[cs] + kr) : kr;
            pm[cs] = pmv;
            smax[cs] = mx;
            if (!rhs) continue;
            L.kappa = kappa[cs];
            L.pm = pmv;
            hipLaunchKernelGGL(k_av_nodegrad, gn, blk, 0, nullptr, A, L);
            hipLaunchKernelGGL(k_av_cells, gc, blk, 0, nullptr, A);
            hipLaunchKernelGGL(k_av_rhs, gn, blk, 0, nullptr, A);
            LP_HIP(hipGetLastError());
            LP_HIP(hipMemcpy(rhs + (size_t)cs * nn * 3, drhs.p, nn * 3 * 8, hipMemcpyDeviceToHost));
            LP_HIP(hipMemcpy(gx + (size_t)cs * nc, dgx.p, nc * 8, hipMemcpyDeviceToHost));
        }
        return LP_OK;
    });
}

}  // extern "C"
