// Level operator y = m*K(m*u) + (1-m)*u on the voxel node grid, node-centric.
//
// Reference: mg._Level.apply / apply_k (mg.py:251-346) and the two-pass
// numba kernels mgops._product_* + _scatter_into (mgops.py:66-217).
//
// MI355X design. The reference forms the (24, cells) product buffer and
// scatters it node by node. Here every thread owns one node column position
// (y, z) and marches along x: at each node it gathers the 27 neighbour nodes
// (three node planes held in registers as a sliding window, so each plane is
// loaded once per column) and sums, over the node's 8 cells, the 3 rows of
// the cell matrix that belong to the node's corner times the cell's 24
// corner values. No product buffer, no atomics, one writer per node, and the
// node's vector update (CG / smoother step) runs in the same pass as an
// epilogue. Lanes run along the flattened (y, z) plane, so every load and
// store of a wave is a contiguous run of z. Blocks are remapped so each XCD
// (8 on MI355X, round-robin block dispatch) owns a contiguous range of
// (x-chunk, yz) tiles and the halo rows stay in its own L2.
//
// Cell matrices: "fine" e_c * Ke (Ke wave-uniform, scalar loads); "weights"
// (first coarse level) sum_q w_q M_q with the 8 child moduli read from the
// finer level's E and M_q staged in LDS, or an explicit patch matrix for
// cells touching held dofs (mg.py:609-642); "explicit" K_c (576, cells) for
// deeper levels (mg.py:643-666).
//
// This header is included by one translation unit per cell-matrix kind.
#pragma once
#include "lp_internal.h"

namespace lp {

struct AGeom {
    int Nx, Ny, Nz, nx, ny, nz;
    int fnx, fny, fnz;      // finer level cells (weights kind)
    size_t nn, ncell;
    int xchunk, nyzb, nxc, nblk, per;
};

template <typename T> struct CSrc {
    const T* e;
    const T* ke;
    const T* fe;
    const T* mq;
    const int* pidx;
    const T* patch;
    const T* k;
};

AGeom make_apply_geom(const Grid& Gd, int level);

namespace apply_detail {

template <int N> struct IC { static constexpr int v = N; };

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// node epilogue: y = finished operator value at node n (all 3 components)
template <typename T>
__device__ __forceinline__ void epilogue(const EpiArgs& E, int m, size_t o, size_t nn, size_t n, const T (&yv)[3],
                                         const T (&ur)[3], double& part) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t i = o + c * nn;
        const T y_ = yv[c];
        switch (E.kind) {
            case EPI_STORE: static_cast<T*>(E.out)[i] = y_; break;
            case EPI_STORE_DOT:
                static_cast<T*>(E.out)[i] = y_;
                part += (double)ur[c] * (double)y_;
                break;
            case EPI_CHEB: {
                const T di = static_cast<const T*>(E.d)[i];
                T* X_ = static_cast<T*>(E.x);
                T* R_ = static_cast<T*>(E.r);
                X_[i] = X_[i] + di;
                const T ri = R_[i] - y_;
                R_[i] = ri;
                const T dv = static_cast<const T*>(E.dinv)[c * nn + n];
                static_cast<T*>(E.d2)[i] = (T)(E.c1 * (double)di + E.c2 * (double)dv * (double)ri);
                break;
            }
            case EPI_RESID_SCALED: {
                const T ri = static_cast<const T*>(E.b)[i] - y_;
                static_cast<T*>(E.r)[i] = ri;
                const T dv = static_cast<const T*>(E.dinv)[c * nn + n];
                static_cast<T*>(E.d2)[i] = (T)(E.c1 * (double)dv * (double)ri);
                break;
            }
            case EPI_RESID_MASK: {
                const bool mc = (m >> c) & 1;
                static_cast<T*>(E.out)[i] = mc ? (T)(static_cast<const T*>(E.b)[i] - y_) : T(0);
                break;
            }
            case EPI_JACOBI: {
                const T dv = static_cast<const T*>(E.dinv)[c * nn + n];
                const T bi = static_cast<const T*>(E.b)[i];
                static_cast<T*>(E.out)[i] = (T)((double)ur[c] + E.c1 * (double)dv * (double)(bi - y_));
                break;
            }
            case EPI_POWER: {
                const bool mc = (m >> c) & 1;
                const T dv = static_cast<const T*>(E.dinv)[c * nn + n];
                const T av = mc ? (T)(dv * y_) : T(0);
                static_cast<T*>(E.out)[i] = av;
                part += (double)ur[c] * (double)av;
                break;
            }
            case EPI_RESID: {
                const T ri = static_cast<const T*>(E.b)[i] - y_;
                static_cast<T*>(E.out)[i] = ri;
                part += (double)ri * (double)ri;
                break;
            }
            default: break;
        }
    }
}

template <typename T, int KIND>
__global__ void __launch_bounds__(256) k_apply(const T* __restrict__ u0, size_t vstride,
                                               const uint8_t* __restrict__ mask, AGeom G, CSrc<T> C,
                                               EpiArgs E) {
    __shared__ T smq[KIND == LK_WEIGHTS ? 8 * 576 : 1];
    __shared__ double sred[4];
    const int rhs = blockIdx.z;
    const int L = blockIdx.x;
    const int logical = (L & 7) * G.per + (L >> 3);
    double part = 0.0;
    const bool skip = (E.scal != nullptr && !E.scal[rhs].active) || logical >= G.nblk;
    if (KIND == LK_WEIGHTS) {
        for (int i = threadIdx.x; i < 8 * 576; i += blockDim.x) smq[i] = C.mq[i];
        __syncthreads();
    }
    const bool need_raw_all = E.kind == EPI_STORE_DOT || E.kind == EPI_JACOBI || E.kind == EPI_POWER;
    if (!skip) {
        const int xc = logical / G.nyzb;
        const int yb = logical - xc * G.nyzb;
        const int Ny = G.Ny, Nz = G.Nz, Nx = G.Nx;
        const int nyz = Ny * Nz;
        const int t = yb * blockDim.x + threadIdx.x;
        if (t < nyz) {
            const int y = t / Nz;
            const int z = t - y * Nz;
            const int x0 = xc * G.xchunk;
            const int x1 = min(x0 + G.xchunk, Nx);
            const T* __restrict__ u = u0 + (size_t)rhs * vstride;
            const size_t nn = G.nn;
            const bool vy[3] = {y > 0, true, y + 1 < Ny};
            const bool vz[3] = {z > 0, true, z + 1 < Nz};
            const bool cyv[2] = {y > 0, y < G.ny};
            const bool czv[2] = {z > 0, z < G.nz};

            T U[3][9][3];
            T Ec[3][4];

            auto load_plane = [&](auto pslot, int x) {
                constexpr int P = decltype(pslot)::v;
                const bool inx = x >= 0 && x < Nx;
#pragma unroll
                for (int dy = 0; dy < 3; ++dy)
#pragma unroll
                    for (int dz = 0; dz < 3; ++dz) {
                        const bool ok = inx && vy[dy] && vz[dz];
                        const size_t n = (size_t)x * nyz + t + (dy - 1) * Nz + (dz - 1);
                        const int m = ok ? (int)mask[n] : 0;
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            U[P][dy * 3 + dz][c] = ((m >> c) & 1) ? u[c * nn + n] : T(0);
                    }
            };
            auto load_cells = [&](auto pslot, int X) {
                constexpr int P = decltype(pslot)::v;
                if (KIND == LK_FINE) {
                    const bool inx = X >= 0 && X < G.nx;
#pragma unroll
                    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                        for (int cz = 0; cz < 2; ++cz) {
                            const bool ok = inx && cyv[cy] && czv[cz];
                            const size_t ci = ((size_t)X * G.ny + (y - 1 + cy)) * G.nz + (z - 1 + cz);
                            Ec[P][cy * 2 + cz] = ok ? C.e[ci] : T(0);
                        }
                }
            };

            auto node = [&](auto p0, auto p1, auto p2, int x) {
                constexpr int P0 = decltype(p0)::v, P1 = decltype(p1)::v, P2 = decltype(p2)::v;
                // Ke is wave-uniform and constant for the launch: read it
                // through the constant address space (scalar loads into SGPRs,
                // used directly as the FMA's scalar operand), one cell's 72
                // values at a time; the opaque pointer keeps the compiler from
                // hoisting all 576 values out of the x loop.
                unsigned long long kea = (unsigned long long)C.ke;
                T acc0 = 0, acc1 = 0, acc2 = 0;
#pragma unroll
                for (int cx = 0; cx < 2; ++cx)
#pragma unroll
                    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                        for (int cz = 0; cz < 2; ++cz) {
                            const int a = (1 - cx) | ((1 - cy) << 1) | ((1 - cz) << 2);
                            const int X = x - 1 + cx;
                            const bool cok = X >= 0 && X < G.nx && cyv[cy] && czv[cz];
                            const size_t ci = cok ? ((size_t)X * G.ny + (y - 1 + cy)) * G.nz + (z - 1 + cz) : 0;
                            T t0 = 0, t1 = 0, t2 = 0;
                            // t[i] += sum_{b,j} K(3a+i, 3b+j) u_b[j]
                            auto rows = [&](auto&& K) {
#pragma unroll
                                for (int b = 0; b < 8; ++b) {
                                    const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
                                    const int pl = cx + bx;
                                    const int w = (cy + by) * 3 + (cz + bz);
#pragma unroll
                                    for (int j = 0; j < 3; ++j) {
                                        const T uv = pl == 0 ? U[P0][w][j] : (pl == 1 ? U[P1][w][j] : U[P2][w][j]);
                                        t0 += K(3 * a + 0, 3 * b + j) * uv;
                                        t1 += K(3 * a + 1, 3 * b + j) * uv;
                                        t2 += K(3 * a + 2, 3 * b + j) * uv;
                                    }
                                }
                            };
                            if (KIND == LK_FINE) {
                                asm volatile("" : "+s"(kea));
                                const __attribute__((address_space(4))) T* kep =
                                    (const __attribute__((address_space(4))) T*)kea;
                                rows([&](int i, int j) { return kep[i * 24 + j]; });
                                const T e = cx == 0 ? Ec[P0][cy * 2 + cz] : Ec[P1][cy * 2 + cz];
                                acc0 += e * t0;
                                acc1 += e * t1;
                                acc2 += e * t2;
                            } else if (KIND == LK_EXPLICIT) {
                                if (cok) {
                                    const T* __restrict__ kc = C.k + ci;
                                    const size_t s = G.ncell;
                                    rows([&](int i, int j) { return kc[(size_t)(i * 24 + j) * s]; });
                                    acc0 += t0;
                                    acc1 += t1;
                                    acc2 += t2;
                                }
                            } else {  // weights
                                if (cok) {
                                    const int pi = C.pidx[ci];
                                    if (pi >= 0) {
                                        const T* __restrict__ kp = C.patch + (size_t)pi * 576;
                                        rows([&](int i, int j) { return kp[i * 24 + j]; });
                                        acc0 += t0;
                                        acc1 += t1;
                                        acc2 += t2;
                                    } else {
                                        const int I = X, J = y - 1 + cy, Kc = z - 1 + cz;
#pragma unroll 1
                                        for (int q = 0; q < 8; ++q) {
                                            const int fx = 2 * I + (q & 1), fy = 2 * J + ((q >> 1) & 1),
                                                      fz = 2 * Kc + ((q >> 2) & 1);
                                            const bool ok = fx < G.fnx && fy < G.fny && fz < G.fnz;
                                            const T wq = ok ? C.fe[((size_t)fx * G.fny + fy) * G.fnz + fz] : T(0);
                                            if (wq != T(0)) {
                                                t0 = 0; t1 = 0; t2 = 0;
                                                const T* mq = smq + q * 576;
                                                rows([&](int i, int j) { return mq[i * 24 + j]; });
                                                acc0 += wq * t0;
                                                acc1 += wq * t1;
                                                acc2 += wq * t2;
                                            }
                                        }
                                    }
                                }
                            }
                        }
                // finish (out = m*Ku + (1-m)*u) and epilogue at node (x, y, z)
                const size_t n = (size_t)x * nyz + t;
                const int m = mask[n];
                T yv[3] = {acc0, acc1, acc2};
                T ur[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const bool mc = (m >> c) & 1;
                    ur[c] = (!mc || need_raw_all) ? u[c * nn + n] : U[P1][4][c];
                    if (!mc) yv[c] = ur[c];
                }
                epilogue<T>(E, m, (size_t)rhs * vstride + n, nn, n, yv, ur, part);
            };

            load_plane(IC<0>{}, x0 - 1);
            load_plane(IC<1>{}, x0);
            load_cells(IC<0>{}, x0 - 1);
            int x = x0;
            while (true) {
                load_plane(IC<2>{}, x + 1);
                load_cells(IC<1>{}, x);
                node(IC<0>{}, IC<1>{}, IC<2>{}, x);
                if (++x >= x1) break;
                load_plane(IC<0>{}, x + 1);
                load_cells(IC<2>{}, x);
                node(IC<1>{}, IC<2>{}, IC<0>{}, x);
                if (++x >= x1) break;
                load_plane(IC<1>{}, x + 1);
                load_cells(IC<0>{}, x);
                node(IC<2>{}, IC<0>{}, IC<1>{}, x);
                if (++x >= x1) break;
            }
        }
    }
    if (E.partials != nullptr) {
        part = wave_sum(part);
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0) sred[wv] = part;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = 0;
            for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sred[i];
            E.partials[(size_t)rhs * gridDim.x + L] = s;
        }
    }
}

template <typename T, int KIND>
void launch_kind(const Grid& Gd, int level, const T* u, size_t vstride, int nrhs, const EpiArgs& E, hipStream_t s) {
    const Level& lv = Gd.lv[level];
    AGeom G = make_apply_geom(Gd, level);
    CSrc<T> C{};
    const bool f64 = sizeof(T) == 8;
    if (KIND == LK_FINE) {
        C.e = static_cast<const T*>(f64 ? lv.e64.p : lv.e32.p);
        C.ke = static_cast<const T*>(f64 ? Gd.ke64.p : Gd.ke32.p);
    } else if (KIND == LK_WEIGHTS) {
        const Level& fl = Gd.lv[level - 1];
        C.fe = static_cast<const T*>(f64 ? fl.e64.p : fl.e32.p);
        C.mq = static_cast<const T*>(f64 ? Gd.mq64.p : Gd.mq32.p);
        C.pidx = lv.patch_idx.as<int>();
        C.patch = static_cast<const T*>(f64 ? lv.patch64.p : lv.patch32.p);
    } else {
        C.k = static_cast<const T*>(f64 ? lv.k64.p : lv.k32.p);
    }
    dim3 grid(8 * G.per, 1, nrhs), block(kBlock);
    hipLaunchKernelGGL((k_apply<T, KIND>), grid, block, 0, s, u, vstride, lv.mask.as<uint8_t>(), G, C, E);
    LP_HIP(hipGetLastError());
}

}  // namespace apply_detail

}  // namespace lp
