// Level operator y = m*K(m*u) + (1-m)*u on the voxel node grid, node-centric.
//
// Reference: mg._Level.apply / apply_k (mg.py:251-346) and the two-pass
// numba kernels mgops._product_* + _scatter_into (mgops.py:66-217).
//
// MI355X design. The reference forms the (24, cells) product buffer and
// scatters it node by node. Here each node's value is summed where it is
// written, with no product buffer, no atomics and one writer per node, and
// the node's vector update (CG / smoother step) runs in the same pass as an
// epilogue.
//
// Fine level (cell matrix e_c * Ke): every thread owns NY nodes of one
// (y, z) column position and marches along x one cell layer at a time. A
// layer's cells (2 x (NY+1) of them) give the rows of the node below them
// (finishing it) and of the node above (starting it), so only two node
// planes are live in registers. Ke is wave-uniform and constant: it is read
// through the constant address space into SGPRs and used as the FMA's scalar
// operand; with NY = 2 each scalar value feeds two nodes. Lanes run along z,
// so every load and store of a wave is a contiguous run. Blocks are
// remapped so each XCD (8 on MI355X, round-robin dispatch) owns a contiguous
// range of (x-chunk, yz) tiles and the halo rows stay in its own L2.
//
// Coarse levels ("weights": sum_q E_child(q) M_q read on the fly, plus the
// explicit patch of cells touching held dofs, mg.py:609-642; "explicit":
// K_c stored (576, cells), mg.py:643-666): one thread per (node, cell of
// its 8) computes the node's 3 rows of that cell's matrix; the 8 lanes of a
// node reduce by shuffles. 8x the parallelism of a node per thread, which
// the small coarse grids need, and coalesced K_c reads.
//
// This header is included by one translation unit per cell-matrix kind.
#pragma once
#include <type_traits>

#include "ke_pattern.h"
#include "lp_internal.h"

namespace lp {

struct AGeom {
    int Nx, Ny, Nz, nx, ny, nz;
    int fnx, fny, fnz;      // finer level cells (weights kind)
    size_t nn, ncell;
    int xchunk, nyzb, nxc, nblk, per;
    int ny_per;             // nodes per thread along y (fine kernel)
    int ntz;                // node tiles along z (fine MFMA kernel)
    int nwt;                // wave tiles (wave-tile MFMA kernel)
    // MFMA kernels: > 0 = the right-hand sides of a tile run in consecutive
    // blocks of one XCD (their shared moduli, masks and dinv come from its
    // L2 once), the grid one-dimensional; 0 = one grid z-slice per rhs
    int rhs_inner;
};

// block -> (rhs, L) of the MFMA kernels: L is the block's index within its
// right-hand side's grid (the partials slot); with rhs_inner the 8*per
// blocks of each rhs interleave as (tile slot, rhs, XCD)
__device__ __forceinline__ void mfma_block_index(const AGeom& G, int& rhs, int& L) {
    if (G.rhs_inner > 0) {
        const int b = (int)blockIdx.x;
        const int j = b >> 3;
        rhs = j % G.rhs_inner;
        L = ((j / G.rhs_inner) << 3) | (b & 7);
    } else {
        rhs = (int)blockIdx.z;
        L = (int)blockIdx.x;
    }
}

// Ke as its KE_NV orbit values (see ke_pattern.h): kernel arguments, so
// they sit in SGPRs for the whole launch
template <typename T> struct KeVals {
    T v[KE_NV];
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
int size_xchunk(int nx, int units, int target);
constexpr int kSizeXChunkBlocks = 1024;   // blocks (FP64 / FP32 fine kernel) or waves (wave-tile kernel)

namespace apply_detail {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

__device__ __forceinline__ void block_partial(double part, EpiArgs& E, int rhs, int L, int nb = -1) {
    __shared__ double sred[8];
    part = wave_sum(part);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sred[wv] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sred[i];
        E.partials[(size_t)rhs * (nb < 0 ? (int)gridDim.x : nb) + L] = s;
    }
}

// node epilogue, in two phases so a kernel can issue the loads early:
// epi_load fetches the vectors the epilogue kind reads at node n (rhs offset
// o), epi_finish combines them with yv = the finished operator value
template <typename T> struct EpiIn {
    T v[4][3];
};

// K >= 0 fixes the epilogue kind at compile time (straight-line code, the
// prefetched values stay in registers); K < 0 reads E.kind at run time
template <typename T, int K = -1>
__device__ __forceinline__ void epi_load(const EpiArgs& E, size_t o, size_t nn, size_t n, EpiIn<T>& in) {
    const int kind = K >= 0 ? K : E.kind;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t i = o + c * nn;
        switch (kind) {
            case EPI_CHEB:
            case EPI_CHEB_LAST:
                in.v[1][c] = static_cast<const T*>(E.x)[i];
                in.v[2][c] = static_cast<const T*>(E.r)[i];
                in.v[3][c] = static_cast<const T*>(E.dinv)[c * nn + n];
                break;
            case EPI_CHEB_FIRST:
            case EPI_CHEB_ONLY:
                in.v[3][c] = static_cast<const T*>(E.dinv)[c * nn + n];
                break;
            case EPI_CHEB_LAST_R:   // r is the raw operand (in.v[2], set by the kernel)
                in.v[1][c] = static_cast<const T*>(E.x)[i];
                in.v[3][c] = static_cast<const T*>(E.dinv)[c * nn + n];
                break;
            case EPI_RESID_SCALED:
            case EPI_JACOBI:
                in.v[0][c] = static_cast<const T*>(E.b)[i];
                in.v[3][c] = static_cast<const T*>(E.dinv)[c * nn + n];
                break;
            case EPI_RESID_MASK:
            case EPI_RESID:
                in.v[0][c] = static_cast<const T*>(E.b)[i];
                break;
            case EPI_POWER:
                in.v[3][c] = static_cast<const T*>(E.dinv)[c * nn + n];
                break;
            default: break;
        }
    }
}

// AR: the type of the smoother/residual arithmetic (double for the FP32
// reference split, as the reference computes its float32 smoother with
// float64 coefficients; float inside the FP16-MFMA operator)
template <typename T, int K = -1, typename AR = double>
__device__ __forceinline__ void epi_finish(const EpiArgs& E, int m, size_t o, size_t nn, const T (&yv)[3],
                                           const T (&ur)[3], const EpiIn<T>& in, double& part) {
    const int kind = K >= 0 ? K : E.kind;
    const AR c1 = (AR)E.c1, c2 = (AR)E.c2;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t i = o + c * nn;
        const T y_ = yv[c];
        switch (kind) {
            case EPI_STORE: static_cast<T*>(E.out)[i] = y_; break;
            case EPI_STORE_DOT:
                static_cast<T*>(E.out)[i] = y_;
                part += (double)ur[c] * (double)y_;
                break;
            // the Chebyshev kinds take d = the operand itself (ur) and, for the
            // first step, b = in.v[0]
            case EPI_CHEB: {
                const T di = ur[c];
                static_cast<T*>(E.x)[i] = in.v[1][c] + di;
                const T ri = in.v[2][c] - y_;
                static_cast<T*>(E.r)[i] = ri;
                static_cast<T*>(E.d2)[i] = (T)(c1 * (AR)di + c2 * (AR)in.v[3][c] * (AR)ri);
                break;
            }
            case EPI_CHEB_FIRST: {
                const T di = ur[c];
                static_cast<T*>(E.x)[i] = di;
                const T ri = in.v[0][c] - y_;
                static_cast<T*>(E.r)[i] = ri;
                static_cast<T*>(E.d2)[i] = (T)(c1 * (AR)di + c2 * (AR)in.v[3][c] * (AR)ri);
                break;
            }
            case EPI_CHEB_ONLY: {
                const T di = ur[c];
                const T dn = (T)(c1 * (AR)di + c2 * (AR)in.v[3][c] * (AR)(in.v[0][c] - y_));
                static_cast<T*>(E.x)[i] = di + dn;
                break;
            }
            case EPI_CHEB_LAST:
            case EPI_CHEB_LAST_R: {
                const T di = ur[c];
                const T dn = (T)(c1 * (AR)di + c2 * (AR)in.v[3][c] * (AR)(in.v[2][c] - y_));
                static_cast<T*>(E.x)[i] = in.v[1][c] + di + dn;
                break;
            }
            case EPI_RESID_SCALED: {
                const T ri = in.v[0][c] - y_;
                static_cast<T*>(E.r)[i] = ri;
                static_cast<T*>(E.d2)[i] = (T)(c1 * (AR)in.v[3][c] * (AR)ri);
                break;
            }
            case EPI_RESID_MASK: {
                const bool mc = (m >> c) & 1;
                static_cast<T*>(E.out)[i] = mc ? (T)(in.v[0][c] - y_) : T(0);
                break;
            }
            case EPI_JACOBI:
                static_cast<T*>(E.out)[i] =
                    (T)((AR)ur[c] + c1 * (AR)in.v[3][c] * (AR)(in.v[0][c] - y_));
                break;
            case EPI_POWER: {
                const bool mc = (m >> c) & 1;
                const T av = mc ? (T)(in.v[3][c] * y_) : T(0);
                static_cast<T*>(E.out)[i] = av;
                part += (double)ur[c] * (double)av;
                break;
            }
            case EPI_RESID: {
                const T ri = in.v[0][c] - y_;
                static_cast<T*>(E.out)[i] = ri;
                part += (double)ri * (double)ri;
                break;
            }
            case EPI_MASKED: {
                const bool mc = (m >> c) & 1;
                static_cast<T*>(E.out)[i] = mc ? y_ : T(0);
                break;
            }
            default: break;
        }
    }
}

// the same epilogue with the kind read at run time and each kind's loads
// inline (one switch, no staging struct: nothing spills to scratch)
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
            case EPI_MASKED: {
                const bool mc = (m >> c) & 1;
                static_cast<T*>(E.out)[i] = mc ? y_ : T(0);
                break;
            }
            default: break;
        }
    }
}

template <typename T>
using cptr = const __attribute__((address_space(4))) T*;

// Raw buffer loads: 32-bit byte offsets against an SGPR base, and the
// hardware range check returns 0 for an offset past num_records, so halo
// positions outside the grid load as zero with no branch. kOOB is past any
// buffer here (all < 2 GiB per right-hand side).
constexpr unsigned kOOB = 0x80000000u;
constexpr int kRsrcWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, kRsrcWord3);
}
template <typename T> __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ int bload_u8(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return (int)__builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}

// ---------------------------------------------------------------------------
// fine level: LDS-staged cell-layer march
// ---------------------------------------------------------------------------
// A block owns 256 consecutive nodes of the flattened (y, z) plane (one per
// thread) and an x-chunk. Three x-planes of its range plus a halo of Nz+1
// nodes on either side are staged in LDS (masked u, component-major, so a
// wave's reads of one component are consecutive words: no bank conflicts);
// the plane two layers ahead is loaded into registers while the current
// layer computes, then stored, with one barrier per layer. Each thread takes
// its 4 cells of the layer one at a time (24 values live): the lo rows
// finish the node below the layer, the hi rows start the node above.
template <typename T>
struct FineSmem {
    static constexpr int kComps = 3;   // staged words per node and plane
};

//
// ZB > 0 (FP64 only): the CG's direction update fused into the operator
// (EPI_CG_APPLY): the operand is p = z + beta p_old, formed as each plane is
// staged, with z the preconditioned residual of element size ZB (E.d) and
// p_old the operand vector u; each owned node writes p (E.d2), A p (E.out)
// and its p . Ap partial, and adds alpha p_old to x (E.x) where the last
// iteration's step is pending. A right-hand side that is no longer active
// but has a pending step only updates x. The arithmetic is k_cg_update's.
// MAXE: staged entries per thread (W <= MAXE * 256; 3 for z extents up to
// 255 nodes, 4 beyond)
template <typename T, bool SYM, int ZB = 0, int MAXE = 4>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 4))) k_apply_fine(const T* __restrict__ u0, size_t vstride,
                                                    const uint8_t* __restrict__ mask, AGeom G,
                                                    const T* __restrict__ ecell, const T* __restrict__ ke,
                                                    KeVals<T> kv, EpiArgs E) {
    static_assert(ZB == 0 || sizeof(T) == 8, "the fused CG update is FP64");
    using ZT = typename std::conditional<ZB == 4, float, double>::type;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const int rhs = blockIdx.z;
    const int L = blockIdx.x;
    const int logical = (L & 7) * G.per + (L >> 3);
    double part = 0.0;
    const bool skip = (E.scal != nullptr && !E.scal[rhs].active) || logical >= G.nblk;
    if (ZB && skip && logical < G.nblk && E.scal[rhs].xpend) {
        // converged at the last step: its pending x += alpha p only
        const int xc = logical / G.nyzb;
        const int t = (logical - xc * G.nyzb) * (int)blockDim.x + (int)threadIdx.x;
        const int nyz = G.Ny * G.Nz;
        if (t < nyz) {
            const double a = E.scal[rhs].alpha;
            const double* pr = static_cast<const double*>((const void*)u0) + (size_t)rhs * vstride;
            double* xr = static_cast<double*>(E.x) + (size_t)rhs * vstride;
            for (int X = xc * G.xchunk; X < min(xc * G.xchunk + G.xchunk, G.Nx); ++X)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const size_t i = (size_t)c * G.nn + (size_t)X * nyz + t;
                    xr[i] = fma(a, pr[i], xr[i]);
                }
        }
    }
    if (!skip) {
        const int xc = logical / G.nyzb;
        const int yb = logical - xc * G.nyzb;
        const int Ny = G.Ny, Nz = G.Nz, Nx = G.Nx;
        const int nyz = Ny * Nz;
        const int tbase = yb * (int)blockDim.x;
        const int t = tbase + (int)threadIdx.x;
        const bool active = t < nyz;
        const int y = active ? t / Nz : 0;
        const int z = active ? t - y * Nz : 0;
        const int x0 = xc * G.xchunk;
        const int x1 = min(x0 + G.xchunk, Nx);
        const T* __restrict__ u = u0 + (size_t)rhs * vstride;
        const unsigned nnu = (unsigned)G.nn;
        const int halo = Nz + 1;
        const int W = (int)blockDim.x + 2 * halo;         // staged nodes per plane
        const int planeT = W * FineSmem<T>::kComps;   // component c of node e at c * W + e
        const __amdgpu_buffer_rsrc_t ru = make_rsrc(u, 3u * nnu * (unsigned)sizeof(T));
        const __amdgpu_buffer_rsrc_t rm = make_rsrc(mask, nnu);
        const __amdgpu_buffer_rsrc_t re = make_rsrc(ecell, (unsigned)G.ncell * (unsigned)sizeof(T));
        const bool cyv[2] = {active && y > 0, active && y < G.ny};
        const bool czv[2] = {active && z > 0, active && z < G.nz};
        T pre[MAXE][3];
        int pm[MAXE];
        // fused CG update: z at the staged nodes, combined at the store
        ZT pz[ZB ? MAXE : 1][3];
        const __amdgpu_buffer_rsrc_t rz =
            make_rsrc(ZB ? static_cast<const ZT*>(E.d) + (size_t)rhs * vstride : nullptr, ZB ? 3u * nnu * ZB : 0u);
        const double beta = ZB ? E.scal[rhs].beta : 0.0;

        // load plane x's staged range into registers, zero outside the grid;
        // the loads do not depend on the mask (applied at the store), so they
        // are all in flight together
        auto fetch = [&](int x) {
            const bool inx = x >= 0 && x < Nx;
#pragma unroll
            for (int k = 0; k < MAXE; ++k) {
                const int e = (int)threadIdx.x + k * (int)blockDim.x;
                const int f = tbase - halo + e;
                // outside: bit 31 set, past the buffer (an arithmetic form: a
                // select of the offset becomes branches around the loads)
                const unsigned oob = (unsigned)!(inx && e < W && f >= 0 && f < nyz) << 31;
                const unsigned n = (unsigned)(x * nyz + f);
                pm[k] = bload_u8(rm, n | oob);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    pre[k][c] = bload<T>(ru, (unsigned)((c * nnu + n) * sizeof(T)) | oob);
                    if (ZB) pz[ZB ? k : 0][c] = bload<ZT>(rz, (unsigned)((c * nnu + n) * ZB) | oob);
                }
            }
        };
        auto store = [&](int buf) {
            T* p = lds + buf * planeT;
#pragma unroll
            for (int k = 0; k < MAXE; ++k) {
                const int e = (int)threadIdx.x + k * (int)blockDim.x;
                if (e < W) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const T v = ZB ? (T)fma(beta, (double)pre[k][c], (double)pz[ZB ? k : 0][c]) : pre[k][c];
                        p[c * W + e] = ((pm[k] >> c) & 1) ? v : T(0);
                    }
                }
            }
        };

        // the 4 cells' moduli of layer X (cell (cy, cz) at 2 cy + cz), loaded
        // a layer ahead: loads complete in order, so a load issued after the
        // next plane's would wait for the whole plane
        auto ec_load = [&](int X, T (&o)[4]) {
            const bool inx = X >= 0 && X < G.nx;
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cz = 0; cz < 2; ++cz) {
                    const bool ok = inx && cyv[cy] && czv[cz];
                    const unsigned ci = (unsigned)((X * G.ny + (y - 1 + cy)) * G.nz + (z - 1 + cz));
                    o[2 * cy + cz] = bload<T>(re, ok ? ci * (unsigned)sizeof(T) : kOOB);
                }
        };
        T acc0 = 0, acc1 = 0, acc2 = 0, nx0 = 0, nx1 = 0, nx2 = 0;
        unsigned long long kea = (unsigned long long)ke;
        // the distinct Ke values in VGPRs: as kernel arguments they sit in
        // SGPRs, and with the rest of the arguments the compiler spilled them
        // to VGPR lanes and read one back (v_readlane) for nearly every FMA
        T kvv[KE_NV];
#pragma unroll
        for (int i = 0; i < KE_NV; ++i) {
            kvv[i] = kv.v[i];
            asm volatile("" : "+v"(kvv[i]));
        }
        // cell layer X between staged planes in buffers bl (X) and bh (X+1)
        auto layer = [&](int X, int bl, int bh, bool lo, const T (&ecv)[4]) {
            nx0 = 0; nx1 = 0; nx2 = 0;
            const bool inx = X >= 0 && X < G.nx;
            const T* PL = lds + bl * planeT;
            const T* PH = lds + bh * planeT;
            const int me = (int)threadIdx.x + halo;
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cz = 0; cz < 2; ++cz) {
                    // one cell's 24 values live at a time: keep the scheduler
                    // from hoisting the next cell's LDS reads above this one,
                    // or sinking this cell's arithmetic below them
                    asm volatile("" ::: "memory");
                    const T ec = ecv[2 * cy + cz];
                    if (!SYM) asm volatile("" : "+s"(kea));
                    cptr<T> Kp = (cptr<T>)kea;
                    auto K = [&](int e) -> T {
                        if (SYM) return KE_SGN[e] > 0 ? kvv[KE_IDX[e]] : (KE_SGN[e] < 0 ? -kvv[KE_IDX[e]] : T(0));
                        return Kp[e];
                    };
                    // corner by corner (3 operand values live), each row's
                    // sum still in the order j = 0..23
                    T tr[2][3] = {{0, 0, 0}, {0, 0, 0}};
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
                        const int idx = me + (cy + by - 1) * Nz + (cz + bz - 1);
                        const T* q = (bx ? PH : PL) + idx;
                        T u3[3];
#pragma unroll
                        for (int c = 0; c < 3; ++c) u3[c] = q[c * W];
#pragma unroll
                        for (int side = 0; side < 2; ++side) {
                            if (side == 0 && !lo) continue;
                            const int a = side | ((1 - cy) << 1) | ((1 - cz) << 2);
#pragma unroll
                            for (int r = 0; r < 3; ++r)
#pragma unroll
                                for (int c = 0; c < 3; ++c) tr[side][r] += K((3 * a + r) * 24 + 3 * b + c) * u3[c];
                        }
                    }
#pragma unroll
                    for (int side = 0; side < 2; ++side) {
                        if (side == 0 && !lo) continue;
                        const T t0 = tr[side][0], t1 = tr[side][1], t2 = tr[side][2];
                        if (side == 0) {
                            acc0 += ec * t0; acc1 += ec * t1; acc2 += ec * t2;
                        } else {
                            nx0 += ec * t0; nx1 += ec * t1; nx2 += ec * t2;
                        }
                    }
                    // the cell is finished here: without this tie the
                    // compiler interleaves the 4 cells of the layer and runs
                    // out of registers (one wave per SIMD)
                    asm volatile("" : "+v"(acc0), "+v"(acc1), "+v"(acc2), "+v"(nx0), "+v"(nx1), "+v"(nx2));
                }
        };

        // prologue: planes x0-1, x0, x0+1 -> buffers 0, 1, 2
        T ecc[4], ecn[4];
        ec_load(x0 - 1, ecc);
        ec_load(x0, ecn);
        fetch(x0 - 1);
        store(0);
        fetch(x0);
        store(1);
        fetch(x0 + 1);
        store(2);
        __syncthreads();
        layer(x0 - 1, 0, 1, false, ecc);
        acc0 = nx0; acc1 = nx1; acc2 = nx2;
        int bl = 1;
        for (int X = x0; X < x1; ++X) {
            const int bh = bl == 2 ? 0 : bl + 1;
            const int bn = bh == 2 ? 0 : bh + 1;          // buffer of plane X-1, refilled with X+2
            const bool more = X + 1 < x1;
            // in order of use: this node's mask, the next layer's moduli, then
            // plane X + 2 (past the chunk: a plane outside the grid, no traffic)
            const unsigned n = (unsigned)(X * nyz + t);
            const int m = bload_u8(rm, active ? n : kOOB);
#pragma unroll
            for (int k = 0; k < 4; ++k) ecc[k] = ecn[k];
            ec_load(X + 1, ecn);
            fetch(more ? X + 2 : -2);
            layer(X, bl, bh, true, ecc);
            if (ZB && active) {
                // p = z + beta p_old at the node (as staged), x += alpha p_old
                T yv[3] = {acc0, acc1, acc2};
                const CgScal& sc = E.scal[rhs];
                const bool xp = sc.xpend != 0;
                const double a = sc.alpha;
                const size_t o = (size_t)rhs * vstride + n;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double po = (double)bload<T>(ru, (unsigned)((c * nnu + n) * sizeof(T)));
                    const double zv = (double)bload<ZT>(rz, (unsigned)((c * nnu + n) * ZB));
                    const double pn = fma(beta, po, zv);
                    const size_t i = o + (size_t)c * G.nn;
                    if (!((m >> c) & 1)) yv[c] = (T)pn;
                    static_cast<double*>(E.out)[i] = yv[c];
                    static_cast<double*>(E.d2)[i] = pn;
                    part += pn * (double)yv[c];
                    if (xp) {
                        double* xr = static_cast<double*>(E.x);
                        xr[i] = fma(a, po, xr[i]);
                    }
                }
            } else if (active) {
                T yv[3] = {acc0, acc1, acc2};
                const T* q = lds + bl * planeT + ((int)threadIdx.x + halo);
                // the staged value is the operand itself where the dof is
                // free; masked-out dofs keep theirs (boundary nodes only:
                // loaded, and waited for, inside the branch)
                T ur[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) ur[c] = q[c * W];
                if (m != 7) {
                    T raw[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) raw[c] = bload<T>(ru, (unsigned)((c * nnu + n) * sizeof(T)));
                    asm volatile("" : "+v"(raw[0]), "+v"(raw[1]), "+v"(raw[2]));
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        if (!((m >> c) & 1)) ur[c] = raw[c];
                }
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    if (!((m >> c) & 1)) yv[c] = ur[c];
                epilogue<T>(E, m, (size_t)rhs * vstride + n, G.nn, n, yv, ur, part);
            }
            acc0 = nx0; acc1 = nx1; acc2 = nx2;
            if (more) {
                store(bn);
                __syncthreads();
            }
            bl = bh;
        }
    }
    if (E.partials != nullptr) block_partial(part, E, rhs, L);
}

// ---------------------------------------------------------------------------
// coarse levels: one thread per (node, cell), shuffle reduction over the 8
// ---------------------------------------------------------------------------
template <typename T, int KIND>
__global__ void __launch_bounds__(256) k_apply_cells(const T* __restrict__ u0, size_t vstride,
                                                     const uint8_t* __restrict__ mask, AGeom G, CSrc<T> C,
                                                     EpiArgs E) {
    __shared__ T smq[KIND == LK_WEIGHTS ? 8 * 576 : 1];
    const int rhs = blockIdx.z;
    const int L = blockIdx.x;
    double part = 0.0;
    const bool skip = E.scal != nullptr && !E.scal[rhs].active;
    if (KIND == LK_WEIGHTS) {
        for (int i = threadIdx.x; i < 8 * 576; i += blockDim.x) smq[i] = C.mq[i];
        __syncthreads();
    }
    const bool need_raw_all = E.kind == EPI_STORE_DOT || E.kind == EPI_JACOBI || E.kind == EPI_POWER;
    const int c8 = threadIdx.x & 7;
    const size_t nid = (size_t)L * (blockDim.x >> 3) + (threadIdx.x >> 3);
    const size_t nn = G.nn;
    const bool valid = !skip && nid < nn;
    T t0 = 0, t1 = 0, t2 = 0;
    const T* __restrict__ u = u0 + (size_t)rhs * vstride;
    int x = 0, y = 0, z = 0;
    if (valid) {
        z = (int)(nid % G.Nz);
        y = (int)((nid / G.Nz) % G.Ny);
        x = (int)(nid / ((size_t)G.Nz * G.Ny));
        const int cx = c8 & 1, cy = (c8 >> 1) & 1, cz = (c8 >> 2) & 1;
        const int X = x - 1 + cx, Y = y - 1 + cy, Z = z - 1 + cz;
        if (X >= 0 && Y >= 0 && Z >= 0 && X < G.nx && Y < G.ny && Z < G.nz) {
            const int a = (1 - cx) | ((1 - cy) << 1) | ((1 - cz) << 2);
            const size_t ci = ((size_t)X * G.ny + Y) * G.nz + Z;
            T ub[24];
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const size_t n = ((size_t)(X + (b & 1)) * G.Ny + (Y + ((b >> 1) & 1))) * G.Nz + (Z + ((b >> 2) & 1));
                const int m = mask[n];
#pragma unroll
                for (int q = 0; q < 3; ++q) ub[3 * b + q] = ((m >> q) & 1) ? u[q * nn + n] : T(0);
            }
            auto rows = [&](auto&& K, T& r0, T& r1, T& r2) {
#pragma unroll
                for (int j = 0; j < 24; ++j) {
                    r0 += K((3 * a + 0) * 24 + j) * ub[j];
                    r1 += K((3 * a + 1) * 24 + j) * ub[j];
                    r2 += K((3 * a + 2) * 24 + j) * ub[j];
                }
            };
            if (KIND == LK_EXPLICIT) {
                const T* __restrict__ kc = C.k + ci;
                const size_t s = G.ncell;
                rows([&](int e) { return kc[(size_t)e * s]; }, t0, t1, t2);
            } else {
                const int pi = C.pidx[ci];
                if (pi >= 0) {
                    const T* __restrict__ kp = C.patch + (size_t)pi * 576;
                    rows([&](int e) { return kp[e]; }, t0, t1, t2);
                } else {
#pragma unroll 1
                    for (int q = 0; q < 8; ++q) {
                        const int fx = 2 * X + (q & 1), fy = 2 * Y + ((q >> 1) & 1), fz = 2 * Z + ((q >> 2) & 1);
                        const bool ok = fx < G.fnx && fy < G.fny && fz < G.fnz;
                        const T wq = ok ? C.fe[((size_t)fx * G.fny + fy) * G.fnz + fz] : T(0);
                        if (wq != T(0)) {
                            T s0 = 0, s1 = 0, s2 = 0;
                            const T* mq = smq + q * 576;
                            rows([&](int e) { return mq[e]; }, s0, s1, s2);
                            t0 += wq * s0;
                            t1 += wq * s1;
                            t2 += wq * s2;
                        }
                    }
                }
            }
        }
    }
    // reduce the node's 8 cells (lanes c8 = 0..7 are adjacent)
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        t0 += __shfl_xor(t0, o, 64);
        t1 += __shfl_xor(t1, o, 64);
        t2 += __shfl_xor(t2, o, 64);
    }
    if (valid && c8 == 0) {
        const size_t n = nid;
        const int m = mask[n];
        T yv[3] = {t0, t1, t2};
        T ur[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const bool mc = (m >> c) & 1;
            ur[c] = (!mc || need_raw_all) ? u[c * nn + n] : T(0);
            if (!mc) yv[c] = ur[c];
        }
        epilogue<T>(E, m, (size_t)rhs * vstride + n, nn, n, yv, ur, part);
    }
    if (E.partials != nullptr) block_partial(part, E, rhs, L);
}

template <typename T, int KIND>
void launch_kind(const Grid& Gd, int level, const T* u, size_t vstride, int nrhs, const EpiArgs& E, hipStream_t s) {
    const Level& lv = Gd.lv[level];
    AGeom G = make_apply_geom(Gd, level);
    const bool f64 = sizeof(T) == 8;
    if (KIND == LK_FINE) {
        const T* e = static_cast<const T*>(f64 ? lv.e64.p : lv.e32.p);
        const T* ke = static_cast<const T*>(f64 ? Gd.ke64.p : Gd.ke32.p);
        dim3 grid(8 * G.per, 1, nrhs), block(kBlock);
        KeVals<T> kv{};
        for (int i = 0; i < KE_NV; ++i) kv.v[i] = (T)Gd.ke_orbit[i];
        const uint8_t* m = lv.mask.as<uint8_t>();
        const int W = kBlock + 2 * (G.Nz + 1);
        if (W > 4 * kBlock) {
            set_error("fine level: z extent too large for the staged plane");
            throw ArgFailure();
        }
        const size_t smem = (size_t)3 * W * FineSmem<T>::kComps * sizeof(T);
        // the kernel for (Ke symmetry, z size of the fused CG update, staged
        // entries per thread)
        auto go = [&](auto sym, auto zb, auto me) {
            hipLaunchKernelGGL((k_apply_fine<T, decltype(sym)::value, decltype(zb)::value, decltype(me)::value>), grid,
                               block, smem, s, u, vstride, m, G, e, ke, kv, E);
        };
        using I3 = std::integral_constant<int, 3>;
        using I4 = std::integral_constant<int, 4>;
        using Z0 = std::integral_constant<int, 0>;
        auto by_entries = [&](auto sym, auto zb) {
            if (W <= 3 * kBlock)
                go(sym, zb, I3{});
            else
                go(sym, zb, I4{});
        };
        auto by_sym = [&](auto zb) {
            if (Gd.ke_sym)
                by_entries(std::true_type{}, zb);
            else
                by_entries(std::false_type{}, zb);
        };
        if (E.kind == EPI_CG_APPLY) {
            if constexpr (sizeof(T) == 8) {
                if ((E.zbytes != 4 && E.zbytes != 8) || !E.scal || !E.d || !E.d2 || !E.x || !E.out) {
                    set_error("fine level: the fused CG update needs z, p, x, A p and the CG scalars");
                    throw ArgFailure();
                }
                if (E.zbytes == 4)
                    by_sym(std::integral_constant<int, 4>{});
                else
                    by_sym(std::integral_constant<int, 8>{});
            } else {
                set_error("fine level: the fused CG update runs on the FP64 operator");
                throw ArgFailure();
            }
        } else {
            by_sym(Z0{});
        }
    } else {
        CSrc<T> C{};
        if (KIND == LK_WEIGHTS) {
            const Level& fl = Gd.lv[level - 1];
            C.fe = static_cast<const T*>(f64 ? fl.e64.p : fl.e32.p);
            C.mq = static_cast<const T*>(f64 ? Gd.mq64.p : Gd.mq32.p);
            C.pidx = lv.patch_idx.as<int>();
            C.patch = static_cast<const T*>(f64 ? lv.patch64.p : lv.patch32.p);
        } else {
            C.k = static_cast<const T*>(f64 ? lv.k64.p : lv.k32.p);
        }
        dim3 grid(G.nblk, 1, nrhs), block(kBlock);
        hipLaunchKernelGGL((k_apply_cells<T, KIND>), grid, block, 0, s, u, vstride, lv.mask.as<uint8_t>(), G, C, E);
    }
    LP_HIP(hipGetLastError());
}

}  // namespace apply_detail

}  // namespace lp
