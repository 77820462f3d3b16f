// Multigrid hierarchy on the device, rebuilt at every set_moduli.
//
// Reference: mg.GridElasticity.set_moduli / _coarsen / _build_coarse
// (mg.py:515-766) and _Level.diagonal (mg.py:348-371).
//
// - fine level: diagonal, alive = d > 1e-300, mask = alive & ~fixed,
//   dinv (mg.py:532-541); held = alive & fixed kept for the patch / strict.
// - first coarse level ("weights", mg.py:609-642): cell matrix
//   sum_q E_child(q) M_q, M_q = P_q^T Ke P_q, read on the fly from the fine
//   moduli; coarse cells with a child touching a held node get an explicit
//   patch matrix sum_q P_q^T (E_q Ke o dm dm^T) P_q (the fine mask applied).
// - deeper levels ("explicit", mg.py:643-666): K_c = sum_k P_k^T K_k P_k,
//   one block per coarse cell, 576 threads, FP64 arithmetic, P_k applied
//   through its 27-node sparsity (16 FMA per entry and child).
// - coarse masks: live = d > 1e-12 max d; strict: held injection and
//   kept >= 1 (mg.py:557-576).
// - coarsest: dense D K D + (I - D), symmetrised, FP64 Cholesky -> inverse
//   (mg.py:708-754; the reference factors with SuperLU on the host).
#include <algorithm>
#include <cmath>

#include "lp_internal.h"

namespace lp {

namespace {

__device__ __forceinline__ size_t cidx(int x, int y, int z, int ny, int nz) {
    return ((size_t)x * ny + y) * nz + z;
}

// weight of coarse corner r on fine corner q of child k (mg.py:57-73)
__device__ __forceinline__ double pk_w(int k, int q, int r) {
    double w = 1.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double pos = 0.5 * (((k >> d) & 1) + ((q >> d) & 1));
        w *= ((r >> d) & 1) ? pos : 1.0 - pos;
    }
    return w;
}

// ---- fine level ------------------------------------------------------------
template <typename T>
__global__ void k_fine_diag(const double* __restrict__ e, const double* __restrict__ ke, const uint8_t* __restrict__ fixed,
                            Geom g, uint8_t* mask, uint8_t* held, double* dinv64, float* dinv32) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.nn) return;
    const int z = (int)(n % g.Nz), y = (int)((n / g.Nz) % g.Ny), x = (int)(n / ((size_t)g.Nz * g.Ny));
    double d[3] = {0, 0, 0};
    for (int cx = 0; cx < 2; ++cx)
        for (int cy = 0; cy < 2; ++cy)
            for (int cz = 0; cz < 2; ++cz) {
                const int X = x - 1 + cx, Y = y - 1 + cy, Z = z - 1 + cz;
                if (X < 0 || Y < 0 || Z < 0 || X >= g.nx || Y >= g.ny || Z >= g.nz) continue;
                const int a = (1 - cx) | ((1 - cy) << 1) | ((1 - cz) << 2);
                const double ec = e[cidx(X, Y, Z, g.ny, g.nz)];
                for (int c = 0; c < 3; ++c) d[c] += ke[(3 * a + c) * 25] * ec;
            }
    const int fx = fixed[n];
    int m = 0, h = 0;
    for (int c = 0; c < 3; ++c) {
        const bool alive = d[c] > 1e-300;
        const bool fc = (fx >> c) & 1;
        if (alive && !fc) m |= 1 << c;
        if (alive && fc) h |= 1 << c;
        const double di = (alive && !fc) ? 1.0 / d[c] : 1.0;
        dinv64[c * g.nn + n] = di;
        dinv32[c * g.nn + n] = (float)di;
    }
    mask[n] = (uint8_t)m;
    held[n] = (uint8_t)h;
}

__global__ void k_count_pos(const double* __restrict__ e, size_t n, unsigned long long* cnt) {
    unsigned long long local = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        local += e[i] > 0 ? 1 : 0;
    for (int o = 32; o > 0; o >>= 1) local += __shfl_down(local, o, 64);
    if ((threadIdx.x & 63) == 0 && local) atomicAdd(cnt, local);
}

// ---- first coarse level (weights + patch) -----------------------------------
// flag coarse cells whose children (E > 0) touch a held fine node
__global__ void k_patch_flag(const double* __restrict__ fe, const uint8_t* __restrict__ held, Geom f, Geom c,
                             int* pidx, unsigned long long* cnt, unsigned long long* active_cnt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.ncell) return;
    const int Z = (int)(i % c.nz), Y = (int)((i / c.nz) % c.ny), X = (int)(i / ((size_t)c.nz * c.ny));
    bool touch = false, any = false;
    for (int k = 0; k < 8; ++k) {
        const int fx = 2 * X + (k & 1), fy = 2 * Y + ((k >> 1) & 1), fz = 2 * Z + ((k >> 2) & 1);
        if (fx >= f.nx || fy >= f.ny || fz >= f.nz) continue;
        const double ec = fe[cidx(fx, fy, fz, f.ny, f.nz)];
        if (ec != 0) any = true;
        if (!(ec > 0)) continue;
        for (int q = 0; q < 8; ++q) {
            const size_t node = cidx(fx + (q & 1), fy + ((q >> 1) & 1), fz + ((q >> 2) & 1), f.Ny, f.Nz);
            if (held[node]) touch = true;
        }
    }
    pidx[i] = touch ? 0 : -1;
    if (touch) atomicAdd(cnt, 1ull);
    if (touch || any) atomicAdd(active_cnt, 1ull);
}

__global__ void k_patch_assign(int* pidx, size_t n, int* list, unsigned int* counter) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (pidx[i] >= 0) {
        const unsigned int s = atomicAdd(counter, 1u);
        pidx[i] = (int)s;
        list[s] = (int)i;
    }
}

// one block (576 threads) per patch cell: sum_k P_k^T (E_k Ke o dm dm^T) P_k
__global__ void __launch_bounds__(576) k_patch_matrices(const int* __restrict__ list, const double* __restrict__ fe,
                                                        const uint8_t* __restrict__ fmask, const double* __restrict__ ke,
                                                        Geom f, Geom c, double* out64, float* out32) {
    __shared__ double kc[576];
    __shared__ double tt[576];
    const int cell = list[blockIdx.x];
    const int Z = cell % c.nz, Y = (cell / c.nz) % c.ny, X = cell / (c.nz * c.ny);
    const int tid = threadIdx.x;
    const int i = tid / 24, j = tid % 24;
    double acc = 0;
    for (int k = 0; k < 8; ++k) {
        const int fx = 2 * X + (k & 1), fy = 2 * Y + ((k >> 1) & 1), fz = 2 * Z + ((k >> 2) & 1);
        double v = 0;
        if (fx < f.nx && fy < f.ny && fz < f.nz) {
            const double ec = fe[cidx(fx, fy, fz, f.ny, f.nz)];
            const int qi = i / 3, ci = i % 3, qj = j / 3, cj = j % 3;
            const int mi = fmask[cidx(fx + (qi & 1), fy + ((qi >> 1) & 1), fz + ((qi >> 2) & 1), f.Ny, f.Nz)];
            const int mj = fmask[cidx(fx + (qj & 1), fy + ((qj >> 1) & 1), fz + ((qj >> 2) & 1), f.Ny, f.Nz)];
            const double di = ((mi >> ci) & 1) ? 1.0 : 0.0, dj = ((mj >> cj) & 1) ? 1.0 : 0.0;
            v = ec * ke[i * 24 + j] * di * dj;
        }
        __syncthreads();
        kc[tid] = v;
        __syncthreads();
        // tt[a][3r+cc] = sum_q K[a][3q+cc] w(k,q,r)
        {
            const int a = i, r = j / 3, cc = j % 3;
            double s = 0;
            for (int q = 0; q < 8; ++q) s += kc[a * 24 + 3 * q + cc] * pk_w(k, q, r);
            tt[tid] = s;
        }
        __syncthreads();
        {
            const int rp = i / 3, cc = i % 3;
            double s = 0;
            for (int q = 0; q < 8; ++q) s += pk_w(k, q, rp) * tt[(3 * q + cc) * 24 + j];
            acc += s;
        }
    }
    out64[(size_t)blockIdx.x * 576 + tid] = acc;
    out32[(size_t)blockIdx.x * 576 + tid] = (float)acc;
}

// weights-level diagonal
__global__ void k_weights_diag(const double* __restrict__ fe, const double* __restrict__ mq, const int* __restrict__ pidx,
                               const double* __restrict__ patch, Geom f, Geom c, double* d) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= c.nn) return;
    const int z = (int)(n % c.Nz), y = (int)((n / c.Nz) % c.Ny), x = (int)(n / ((size_t)c.Nz * c.Ny));
    double acc[3] = {0, 0, 0};
    for (int cx = 0; cx < 2; ++cx)
        for (int cy = 0; cy < 2; ++cy)
            for (int cz = 0; cz < 2; ++cz) {
                const int X = x - 1 + cx, Y = y - 1 + cy, Z = z - 1 + cz;
                if (X < 0 || Y < 0 || Z < 0 || X >= c.nx || Y >= c.ny || Z >= c.nz) continue;
                const int a = (1 - cx) | ((1 - cy) << 1) | ((1 - cz) << 2);
                const size_t ci = cidx(X, Y, Z, c.ny, c.nz);
                const int pi = pidx[ci];
                for (int cc = 0; cc < 3; ++cc) {
                    const int ii = 3 * a + cc;
                    if (pi >= 0) {
                        acc[cc] += patch[(size_t)pi * 576 + ii * 25];
                    } else {
                        double s = 0;
                        for (int q = 0; q < 8; ++q) {
                            const int fx = 2 * X + (q & 1), fy = 2 * Y + ((q >> 1) & 1), fz = 2 * Z + ((q >> 2) & 1);
                            if (fx >= f.nx || fy >= f.ny || fz >= f.nz) continue;
                            s += mq[q * 576 + ii * 25] * fe[cidx(fx, fy, fz, f.ny, f.nz)];
                        }
                        acc[cc] += s;
                    }
                }
            }
    for (int cc = 0; cc < 3; ++cc) d[cc * c.nn + n] = acc[cc];
}

// ---- explicit levels ---------------------------------------------------------
// K_c = sum_k P_k^T K_child P_k; the child's matrix from the finer level.
template <int SRC>  // 0: finer is weights, 1: finer is explicit (fp64), 2: explicit (fp32)
__global__ void __launch_bounds__(576) k_galerkin(Geom f, Geom c, const double* __restrict__ ffe /* finest E for weights */,
                                                  Geom ff, const double* __restrict__ mq, const int* __restrict__ pidx,
                                                  const double* __restrict__ patch, const double* __restrict__ k64,
                                                  const float* __restrict__ k32, const uint8_t* __restrict__ fmask,
                                                  int masked, double* out64, float* out32, unsigned long long* active_cnt) {
    __shared__ double kc[576];
    __shared__ double tt[576];
    __shared__ int anynz;
    const size_t cell = blockIdx.x;
    const int Z = (int)(cell % c.nz), Y = (int)((cell / c.nz) % c.ny), X = (int)(cell / ((size_t)c.nz * c.ny));
    const int tid = threadIdx.x;
    const int i = tid / 24, j = tid % 24;
    if (tid == 0) anynz = 0;
    double acc = 0;
    for (int k = 0; k < 8; ++k) {
        const int fx = 2 * X + (k & 1), fy = 2 * Y + ((k >> 1) & 1), fz = 2 * Z + ((k >> 2) & 1);
        double v = 0;
        if (fx < f.nx && fy < f.ny && fz < f.nz) {
            const size_t fc = cidx(fx, fy, fz, f.ny, f.nz);
            if (SRC == 0) {
                const int pi = pidx[fc];
                if (pi >= 0) {
                    v = patch[(size_t)pi * 576 + tid];
                } else {
                    for (int q = 0; q < 8; ++q) {
                        const int gx = 2 * fx + (q & 1), gy = 2 * fy + ((q >> 1) & 1), gz = 2 * fz + ((q >> 2) & 1);
                        if (gx >= ff.nx || gy >= ff.ny || gz >= ff.nz) continue;
                        v += mq[q * 576 + tid] * ffe[cidx(gx, gy, gz, ff.ny, ff.nz)];
                    }
                }
            } else if (SRC == 1) {
                v = k64[(size_t)tid * f.ncell + fc];
            } else {
                v = (double)k32[(size_t)tid * f.ncell + fc];
            }
            if (masked) {
                const int qi = i / 3, ci = i % 3, qj = j / 3, cj = j % 3;
                const int mi = fmask[cidx(fx + (qi & 1), fy + ((qi >> 1) & 1), fz + ((qi >> 2) & 1), f.Ny, f.Nz)];
                const int mj = fmask[cidx(fx + (qj & 1), fy + ((qj >> 1) & 1), fz + ((qj >> 2) & 1), f.Ny, f.Nz)];
                if (!((mi >> ci) & 1) || !((mj >> cj) & 1)) v = 0;
            }
        }
        __syncthreads();
        kc[tid] = v;
        __syncthreads();
        {
            const int a = i, r = j / 3, cc = j % 3;
            double s = 0;
            for (int q = 0; q < 8; ++q) s += kc[a * 24 + 3 * q + cc] * pk_w(k, q, r);
            tt[tid] = s;
        }
        __syncthreads();
        {
            const int rp = i / 3, cc = i % 3;
            double s = 0;
            for (int q = 0; q < 8; ++q) s += pk_w(k, q, rp) * tt[(3 * q + cc) * 24 + j];
            acc += s;
        }
    }
    if (out64) out64[(size_t)tid * c.ncell + cell] = acc;
    if (out32) out32[(size_t)tid * c.ncell + cell] = (float)acc;
    const float a32 = (float)acc;
    if ((out32 ? a32 != 0.0f : acc != 0.0)) anynz = 1;   // benign race: all writers store 1
    __syncthreads();
    if (tid == 0 && anynz) atomicAdd(active_cnt, 1ull);
}

template <typename T>
__global__ void k_explicit_diag(const T* __restrict__ k, Geom c, double* d) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= c.nn) return;
    const int z = (int)(n % c.Nz), y = (int)((n / c.Nz) % c.Ny), x = (int)(n / ((size_t)c.Nz * c.Ny));
    double acc[3] = {0, 0, 0};
    for (int cx = 0; cx < 2; ++cx)
        for (int cy = 0; cy < 2; ++cy)
            for (int cz = 0; cz < 2; ++cz) {
                const int X = x - 1 + cx, Y = y - 1 + cy, Z = z - 1 + cz;
                if (X < 0 || Y < 0 || Z < 0 || X >= c.nx || Y >= c.ny || Z >= c.nz) continue;
                const int a = (1 - cx) | ((1 - cy) << 1) | ((1 - cz) << 2);
                const size_t ci = cidx(X, Y, Z, c.ny, c.nz);
                for (int cc = 0; cc < 3; ++cc) acc[cc] += (double)k[(size_t)((3 * a + cc) * 25) * c.ncell + ci];
            }
    for (int cc = 0; cc < 3; ++cc) d[cc * c.nn + n] = acc[cc];
}

__global__ void k_max(const double* __restrict__ d, size_t n, double* out) {
    __shared__ double s[256];
    double m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = fmax(m, d[i]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // monotone max on a non-negative double via its bit pattern
        atomicMax((unsigned long long*)out, (unsigned long long)__double_as_longlong(s[0]));
    }
}

// coarse mask / dinv; strict: held injection and kept >= 1 (mg.py:557-576)
__global__ void k_coarse_mask(const double* __restrict__ d, const double* __restrict__ dmax, Geom f, Geom c,
                              const uint8_t* __restrict__ fmask, const uint8_t* __restrict__ fheld, int strict,
                              uint8_t* mask, uint8_t* held, double* dinv64, float* dinv32) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= c.nn) return;
    const int z = (int)(n % c.Nz), y = (int)((n / c.Nz) % c.Ny), x = (int)(n / ((size_t)c.Nz * c.Ny));
    const double thr = 1e-12 * fmax(*dmax, 1e-300);
    int hc = 0;
    if (strict && 2 * x < f.Nx && 2 * y < f.Ny && 2 * z < f.Nz) hc = fheld[cidx(2 * x, 2 * y, 2 * z, f.Ny, f.Nz)];
    int m = 0;
    for (int cc = 0; cc < 3; ++cc) {
        const double dv = d[cc * c.nn + n];
        bool live = dv > thr;
        if (strict) {
            double kept = 0;
            for (int a = -1; a <= 1; ++a) {
                const int fx = 2 * x + a;
                if (fx < 0 || fx >= f.Nx) continue;
                for (int b = -1; b <= 1; ++b) {
                    const int fy = 2 * y + b;
                    if (fy < 0 || fy >= f.Ny) continue;
                    for (int e = -1; e <= 1; ++e) {
                        const int fz = 2 * z + e;
                        if (fz < 0 || fz >= f.Nz) continue;
                        const double w = (a ? 0.5 : 1.0) * (b ? 0.5 : 1.0) * (e ? 0.5 : 1.0);
                        if ((fmask[cidx(fx, fy, fz, f.Ny, f.Nz)] >> cc) & 1) kept += w;
                    }
                }
            }
            live = live && !((hc >> cc) & 1) && kept >= 1.0;
        }
        if (live) m |= 1 << cc;
        const double di = live ? 1.0 / dv : 1.0;
        dinv64[cc * c.nn + n] = di;
        dinv32[cc * c.nn + n] = (float)di;
    }
    mask[n] = (uint8_t)m;
    held[n] = (uint8_t)hc;
}

// materialise the weights level's cell matrices (the FP64 and FP32
// explicit applies, the coarsest assembly); summed in FP64
template <typename T>
__global__ void k_weights_explicit(const double* __restrict__ fe, const double* __restrict__ mq,
                                   const int* __restrict__ pidx, const double* __restrict__ patch, Geom f, Geom c,
                                   T* k64) {
    const size_t cell = blockIdx.x;
    const int tid = threadIdx.x;
    const int Z = (int)(cell % c.nz), Y = (int)((cell / c.nz) % c.ny), X = (int)(cell / ((size_t)c.nz * c.ny));
    const int pi = pidx[cell];
    double v = 0;
    if (pi >= 0) {
        v = patch[(size_t)pi * 576 + tid];
    } else {
        for (int q = 0; q < 8; ++q) {
            const int fx = 2 * X + (q & 1), fy = 2 * Y + ((q >> 1) & 1), fz = 2 * Z + ((q >> 2) & 1);
            if (fx >= f.nx || fy >= f.ny || fz >= f.nz) continue;
            v += mq[q * 576 + tid] * fe[cidx(fx, fy, fz, f.ny, f.nz)];
        }
    }
    k64[(size_t)tid * c.ncell + cell] = (T)v;
}

// dense coarsest matrix (component-major dofs c*nn + node), one thread per
// (row dof, neighbour node): sum of the shared cells' blocks
template <typename T>
__global__ void k_assemble_dense(const T* __restrict__ k, Geom c, double* a) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ndof = 3 * c.nn;
    if (tid >= ndof * 27) return;
    const size_t row = tid / 27;
    const int nb = (int)(tid % 27);
    const int ci = (int)(row / c.nn);
    const size_t p = row % c.nn;
    const int pz = (int)(p % c.Nz), py = (int)((p / c.Nz) % c.Ny), px = (int)(p / ((size_t)c.Nz * c.Ny));
    const int dx = nb / 9 - 1, dy = (nb / 3) % 3 - 1, dz = nb % 3 - 1;
    const int qx = px + dx, qy = py + dy, qz = pz + dz;
    if (qx < 0 || qy < 0 || qz < 0 || qx >= c.Nx || qy >= c.Ny || qz >= c.Nz) return;
    const size_t q = cidx(qx, qy, qz, c.Ny, c.Nz);
    double v[3] = {0, 0, 0};
    for (int cx = 0; cx < 2; ++cx)
        for (int cy = 0; cy < 2; ++cy)
            for (int cz = 0; cz < 2; ++cz) {
                const int X = px - 1 + cx, Y = py - 1 + cy, Z = pz - 1 + cz;
                if (X < 0 || Y < 0 || Z < 0 || X >= c.nx || Y >= c.ny || Z >= c.nz) continue;
                // q must be a corner of this cell
                const int bx = qx - X, by = qy - Y, bz = qz - Z;
                if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
                const int a = (1 - cx) | ((1 - cy) << 1) | ((1 - cz) << 2);
                const int b = bx | (by << 1) | (bz << 2);
                const size_t cell = cidx(X, Y, Z, c.ny, c.nz);
                for (int cj = 0; cj < 3; ++cj)
                    v[cj] += (double)k[(size_t)((3 * a + ci) * 24 + 3 * b + cj) * c.ncell + cell];
            }
    for (int cj = 0; cj < 3; ++cj) a[row * ndof + cj * c.nn + q] = v[cj];
}

// a = D a D + (I - D), then 0.5 (a + a^T)
__global__ void k_mask_sym(double* a, const uint8_t* __restrict__ mask, size_t nn) {
    const size_t n = 3 * nn;
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= n * n) return;
    const size_t i = tid / n, j = tid % n;
    if (j < i) return;
    const double mi = ((mask[i % nn] >> (i / nn)) & 1) ? 1.0 : 0.0;
    const double mj = ((mask[j % nn] >> (j / nn)) & 1) ? 1.0 : 0.0;
    if (i == j) {
        a[i * n + i] = mi * a[i * n + i] * mi + (1.0 - mi);
    } else {
        const double v = 0.5 * (mi * a[i * n + j] * mj + mj * a[j * n + i] * mi);
        a[i * n + j] = v;
        a[j * n + i] = v;
    }
}

// (ncell, 24) dof indices (c * nn + node) of the level's cells: the index
// arithmetic of the operator kernels, exported for bit-exact checks
__global__ void k_cell_dofs(Geom g, int64_t* out) {
    const size_t cell = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= g.ncell) return;
    const int Z = (int)(cell % g.nz), Y = (int)((cell / g.nz) % g.ny), X = (int)(cell / ((size_t)g.nz * g.ny));
    for (int b = 0; b < 8; ++b) {
        const size_t node = cidx(X + (b & 1), Y + ((b >> 1) & 1), Z + ((b >> 2) & 1), g.Ny, g.Nz);
        for (int c = 0; c < 3; ++c) out[cell * 24 + 3 * b + c] = (int64_t)(c * g.nn + node);
    }
}

inline unsigned nblk(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

unsigned long long read_counter(DBuf& buf, hipStream_t s) {
    unsigned long long v = 0;
    LP_HIP(hipMemcpyAsync(&v, buf.p, sizeof(v), hipMemcpyDeviceToHost, s));
    LP_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace

void level_cell_dofs(const Grid& G, int level, int64_t* out_dev) {
    const Geom& g = G.lv[level].g;
    hipLaunchKernelGGL(k_cell_dofs, dim3(nblk(g.ncell)), dim3(kBlock), 0, G.stream, g, out_dev);
    LP_HIP(hipGetLastError());
}

static void alloc_vectors(Grid& G, Level& L) {
    // a 64-byte tail: the FP32 restriction reads one element past a field's
    // last row (12-byte row loads) and discards it
    const size_t n = 3 * L.g.nn * (size_t)G.max_rhs * (size_t)G.vbytes() + 64;
    L.vb.alloc(n);
    L.vx.alloc(n);
    L.vr.alloc(n);
    L.vd.alloc(n);
    L.vd2.alloc(n);
    L.vt.alloc(n);
}

int build_hierarchy(Grid& G) {
    hipStream_t s = G.stream;
    ++G.epoch;
    G.fp16_off = false;
    DBuf& counter = G.work.counter;
    DBuf& counter2 = G.work.counter2;
    DBuf& dtmp = G.work.dtmp;
    DBuf& dmaxb = G.work.dmaxb;
    counter.alloc(sizeof(unsigned long long));
    counter2.alloc(sizeof(unsigned long long));

    // ---- fine level
    Level& L0 = G.lv[0];
    L0.kind = LK_FINE;
    L0.g = G.g0;
    L0.mask.alloc(L0.g.nn);
    G.lv[0].held.alloc(L0.g.nn);
    L0.dinv64.alloc(3 * L0.g.nn * 8);
    L0.dinv32.alloc(3 * L0.g.nn * 4);
    L0.e32.alloc(L0.g.ncell * 4);
    launch_cast(L0.e64.p, 8, L0.e32.p, 4, L0.g.ncell, 0, 1, s);
    hipLaunchKernelGGL(k_fine_diag<double>, dim3(nblk(L0.g.nn)), dim3(kBlock), 0, s, L0.e64.as<double>(),
                       G.ke64.as<double>(), G.fixed.as<uint8_t>(), L0.g, L0.mask.as<uint8_t>(),
                       G.lv[0].held.as<uint8_t>(), L0.dinv64.as<double>(), L0.dinv32.as<float>());
    LP_HIP(hipGetLastError());
    LP_HIP(hipMemsetAsync(counter.p, 0, 8, s));
    hipLaunchKernelGGL(k_count_pos, dim3(std::min<size_t>(1024, nblk(L0.g.ncell))), dim3(kBlock), 0, s,
                       L0.e64.as<double>(), L0.g.ncell, counter.as<unsigned long long>());
    LP_HIP(hipGetLastError());
    long long n_active = (long long)read_counter(counter, s);
    L0.n_active = n_active;
    alloc_vectors(G, L0);
    int nl = 1;
    const bool need64 = G.opt.precision == LP_PREC_FP64;

    while (n_active > G.opt.coarse_cells && std::min({G.lv[nl - 1].g.nx, G.lv[nl - 1].g.ny, G.lv[nl - 1].g.nz}) >= 2 &&
           (G.opt.max_levels <= 0 || nl < G.opt.max_levels) && nl < kMaxLevels) {
        Level& F = G.lv[nl - 1];
        Level& C = G.lv[nl];
        int h[3] = {F.g.nx / 2, F.g.ny / 2, F.g.nz / 2};
        for (int d = 0; d < 3; ++d) h[d] = h[d] > 1 ? h[d] + h[d] % 2 : h[d];
        C.g.set(h[0], h[1], h[2]);
        C.fine = F.g;
        C.k64_apply = false;
        C.k32_apply = false;
        C.k16_apply = false;
        // FP16 cell matrices on the mixed V-cycle's Galerkin levels (param 20
        // = 1) and on the first coarse level as well (2)
        const int half = G.opt.precision == LP_PREC_MIXED ? G.tune_half_explicit : 0;
        dtmp.alloc(3 * C.g.nn * 8);
        LP_HIP(hipMemsetAsync(counter.p, 0, 8, s));
        LP_HIP(hipMemsetAsync(counter2.p, 0, 8, s));
        if (nl == 1) {
            C.kind = LK_WEIGHTS;
            C.patch_idx.alloc(C.g.ncell * sizeof(int));
            hipLaunchKernelGGL(k_patch_flag, dim3(nblk(C.g.ncell)), dim3(kBlock), 0, s, F.e64.as<double>(),
                               G.lv[0].held.as<uint8_t>(), F.g, C.g, C.patch_idx.as<int>(),
                               counter.as<unsigned long long>(), counter2.as<unsigned long long>());
            LP_HIP(hipGetLastError());
            const unsigned long long np = read_counter(counter, s);
            n_active = (long long)read_counter(counter2, s);
            C.npatch = (int)np;
            DBuf& list = C.patch_cell;
            DBuf& ctr = G.work.ctr;
            list.alloc(std::max<size_t>(1, np) * sizeof(int));
            if (G.opt.precision == LP_PREC_MIXED)
                C.patch_prod.alloc(std::max<size_t>(1, np) * 24 * sizeof(float) * (size_t)G.max_rhs);
            ctr.alloc(sizeof(unsigned int));
            LP_HIP(hipMemsetAsync(ctr.p, 0, sizeof(unsigned int), s));
            hipLaunchKernelGGL(k_patch_assign, dim3(nblk(C.g.ncell)), dim3(kBlock), 0, s, C.patch_idx.as<int>(),
                               C.g.ncell, list.as<int>(), ctr.as<unsigned int>());
            LP_HIP(hipGetLastError());
            C.patch64.alloc(std::max<size_t>(1, np) * 576 * 8);
            C.patch32.alloc(std::max<size_t>(1, np) * 576 * 4);
            if (np)
                hipLaunchKernelGGL(k_patch_matrices, dim3((unsigned)np), dim3(576), 0, s, list.as<int>(),
                                   F.e64.as<double>(), F.mask.as<uint8_t>(), G.ke64.as<double>(), F.g, C.g,
                                   C.patch64.as<double>(), C.patch32.as<float>());
            LP_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_weights_diag, dim3(nblk(C.g.nn)), dim3(kBlock), 0, s, F.e64.as<double>(),
                               G.mq64.as<double>(), C.patch_idx.as<int>(), C.patch64.as<double>(), F.g, C.g,
                               dtmp.as<double>());
            LP_HIP(hipGetLastError());
            // the FP64 path applies this level as formed cell matrices
            // sum_q w_q M_q (patch cells: the patch matrix), 1/8 of the
            // matrix-free form's flops for 4.6 KB per cell read once per apply
            C.k64_apply = need64 && G.tune_l1_explicit != 0;
            if (C.k64_apply) {
                C.k64.alloc(576 * C.g.ncell * 8);
                hipLaunchKernelGGL(k_weights_explicit<double>, dim3((unsigned)C.g.ncell), dim3(576), 0, s, F.e64.as<double>(),
                                   G.mq64.as<double>(), C.patch_idx.as<int>(), C.patch64.as<double>(), F.g, C.g,
                                   C.k64.as<double>());
                LP_HIP(hipGetLastError());
            }
            // the FP32 and mixed V-cycles likewise, in FP32 (2.3 KB per cell)
            C.k32_apply = !need64 && G.tune_l1_explicit32 != 0;
            if (C.k32_apply) {
                C.k32.alloc(576 * C.g.ncell * 4);
                hipLaunchKernelGGL(k_weights_explicit<float>, dim3((unsigned)C.g.ncell), dim3(576), 0, s,
                                   F.e64.as<double>(), G.mq64.as<double>(), C.patch_idx.as<int>(),
                                   C.patch64.as<double>(), F.g, C.g, C.k32.as<float>());
                LP_HIP(hipGetLastError());
            }
            // the mixed V-cycle: the same matrices in FP16 (1.1 KB per cell)
            C.k16_apply = half >= 2 && !C.k32_apply;
            if (C.k16_apply) {
                C.k16.alloc(576 * C.g.ncell * 2);
                C.sc16.alloc(C.g.ncell * 4);
                form_half_matrices(G, nl, nullptr, s);
            }
            LP_HIP(hipStreamSynchronize(s));   // ctr goes out of scope
        } else {
            C.kind = LK_EXPLICIT;
            if (need64) C.k64.alloc(576 * C.g.ncell * 8);
            C.k32.alloc(576 * C.g.ncell * 4);
            const bool masked = G.strict;
            if (F.kind == LK_WEIGHTS) {
                const Level& FF = G.lv[nl - 2];
                hipLaunchKernelGGL(k_galerkin<0>, dim3((unsigned)C.g.ncell), dim3(576), 0, s, F.g, C.g,
                                   FF.e64.as<double>(), FF.g, G.mq64.as<double>(), F.patch_idx.as<int>(),
                                   F.patch64.as<double>(), (const double*)nullptr, (const float*)nullptr,
                                   F.mask.as<uint8_t>(), (int)masked, need64 ? C.k64.as<double>() : nullptr,
                                   C.k32.as<float>(), counter2.as<unsigned long long>());
            } else if (need64) {
                hipLaunchKernelGGL(k_galerkin<1>, dim3((unsigned)C.g.ncell), dim3(576), 0, s, F.g, C.g,
                                   (const double*)nullptr, F.g, (const double*)nullptr, (const int*)nullptr,
                                   (const double*)nullptr, F.k64.as<double>(), (const float*)nullptr,
                                   F.mask.as<uint8_t>(), (int)masked, C.k64.as<double>(), C.k32.as<float>(),
                                   counter2.as<unsigned long long>());
            } else {
                hipLaunchKernelGGL(k_galerkin<2>, dim3((unsigned)C.g.ncell), dim3(576), 0, s, F.g, C.g,
                                   (const double*)nullptr, F.g, (const double*)nullptr, (const int*)nullptr,
                                   (const double*)nullptr, (const double*)nullptr, F.k32.as<float>(),
                                   F.mask.as<uint8_t>(), (int)masked, nullptr, C.k32.as<float>(),
                                   counter2.as<unsigned long long>());
            }
            LP_HIP(hipGetLastError());
            C.k16_apply = half >= 1;
            if (C.k16_apply) {
                C.k16.alloc(576 * C.g.ncell * 2);
                C.sc16.alloc(C.g.ncell * 4);
                form_half_matrices(G, nl, C.k32.as<float>(), s);
            }
            if (need64)
                hipLaunchKernelGGL(k_explicit_diag<double>, dim3(nblk(C.g.nn)), dim3(kBlock), 0, s, C.k64.as<double>(),
                                   C.g, dtmp.as<double>());
            else
                hipLaunchKernelGGL(k_explicit_diag<float>, dim3(nblk(C.g.nn)), dim3(kBlock), 0, s, C.k32.as<float>(),
                                   C.g, dtmp.as<double>());
            LP_HIP(hipGetLastError());
            n_active = (long long)read_counter(counter2, s);
        }
        // mask, dinv
        dmaxb.alloc(8);
        LP_HIP(hipMemsetAsync(dmaxb.p, 0, 8, s));
        hipLaunchKernelGGL(k_max, dim3(std::min<size_t>(512, nblk(3 * C.g.nn))), dim3(256), 0, s, dtmp.as<double>(),
                           3 * C.g.nn, dmaxb.as<double>());
        LP_HIP(hipGetLastError());
        C.mask.alloc(C.g.nn);
        G.lv[nl].held.alloc(C.g.nn);
        C.dinv64.alloc(3 * C.g.nn * 8);
        C.dinv32.alloc(3 * C.g.nn * 4);
        hipLaunchKernelGGL(k_coarse_mask, dim3(nblk(C.g.nn)), dim3(kBlock), 0, s, dtmp.as<double>(), dmaxb.as<double>(),
                           F.g, C.g, F.mask.as<uint8_t>(), G.lv[nl - 1].held.as<uint8_t>(), (int)G.strict,
                           C.mask.as<uint8_t>(), G.lv[nl].held.as<uint8_t>(), C.dinv64.as<double>(),
                           C.dinv32.as<float>());
        LP_HIP(hipGetLastError());
        C.n_active = n_active;
        alloc_vectors(G, C);
        ++nl;
    }
    G.nlevels = nl;

    // ---- coarsest dense inverse
    G.ncoarse = 0;
    if (nl >= 2) {
        Level& C = G.lv[nl - 1];
        const size_t ndof = 3 * C.g.nn;
        G.ncoarse = (int)ndof;
        G.cinv.alloc(ndof * ndof * 8);
        LP_HIP(hipMemsetAsync(G.cinv.p, 0, ndof * ndof * 8, s));
        if (C.kind == LK_WEIGHTS && !C.k64_apply) {
            C.k64.alloc(576 * C.g.ncell * 8);
            hipLaunchKernelGGL(k_weights_explicit<double>, dim3((unsigned)C.g.ncell), dim3(576), 0, s,
                               G.lv[nl - 2].e64.as<double>(), G.mq64.as<double>(), C.patch_idx.as<int>(),
                               C.patch64.as<double>(), C.fine, C.g, C.k64.as<double>());
            LP_HIP(hipGetLastError());
        }
        const size_t nt = ndof * 27;
        if (C.k64.p)
            hipLaunchKernelGGL(k_assemble_dense<double>, dim3(nblk(nt)), dim3(kBlock), 0, s, C.k64.as<double>(), C.g,
                               G.cinv.as<double>());
        else
            hipLaunchKernelGGL(k_assemble_dense<float>, dim3(nblk(nt)), dim3(kBlock), 0, s, C.k32.as<float>(), C.g,
                               G.cinv.as<double>());
        LP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mask_sym, dim3(nblk(ndof * ndof)), dim3(kBlock), 0, s, G.cinv.as<double>(),
                           C.mask.as<uint8_t>(), C.g.nn);
        LP_HIP(hipGetLastError());
        if (!dense_spd_inverse(G.cinv.as<double>(), (int)ndof, G.work, s)) return LP_SINGULAR;
    }
    return LP_OK;
}

}  // namespace lp
