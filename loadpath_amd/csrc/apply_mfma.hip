// Fine-level operator for the mixed-precision V-cycle: the shared 24x24 Ke
// times every cell's corner vector as an FP16 MFMA contraction with FP32
// accumulation, then a node-centric sum in LDS.
//
// Reference: the same operator as mg._Level.apply on the fine level
// (mg.py:251-346; mgops._product_scaled + _scatter_into, mgops.py:80-217),
// which the reference evaluates in FP32 inside the V-cycle.
//
// Layout (per block: 256 consecutive nodes of the flattened (y, z) plane and
// an x-chunk, as the vector fine kernel):
//  - three x-planes of FP32 masked u for the block's range plus a halo of
//    Nz+1 nodes are staged in LDS (triple-buffered, next plane prefetched);
//  - per cell layer, the layer's cells covering the block's nodes (256 + Nz
//    + 1, padded to 32) are contracted 32 at a time with
//    v_mfma_f32_32x32x8_f16: A = Ke (rows padded to 32, constant in
//    registers), B = the 32 cells' corner vectors, 3 K-steps of 8. Each
//    cell's vector is taken relative to its corner-0 displacement first
//    (Ke annihilates translations, so the product is the same) so FP16
//    carries the cell's deformation, not its rigid offset;
//  - the FP32 products, scaled by each cell's modulus, go to LDS, and every
//    node sums the 3 rows of its 4 cells in the layer: the lo rows finish the
//    node below the layer, the hi rows start the node above;
//  - the epilogue (smoother / residual update) runs per node as in the
//    vector kernels.
#include "apply_impl.h"

namespace lp {

namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kCellTile = 32;

__global__ void __launch_bounds__(256) k_apply_fine_mfma(const float* __restrict__ u0, size_t vstride,
                                                         const uint8_t* __restrict__ mask, AGeom G,
                                                         const float* __restrict__ ecell,
                                                         const float* __restrict__ ke, EpiArgs E) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    using namespace apply_detail;
    const int rhs = blockIdx.z;
    const int L = blockIdx.x;
    const int logical = (L & 7) * G.per + (L >> 3);
    double part = 0.0;
    const bool skip = (E.scal != nullptr && !E.scal[rhs].active) || logical >= G.nblk;
    const bool need_raw_all = E.kind == EPI_STORE_DOT || E.kind == EPI_JACOBI || E.kind == EPI_POWER;
    if (!skip) {
        const int xc = logical / G.nyzb;
        const int yb = logical - xc * G.nyzb;
        const int Ny = G.Ny, Nz = G.Nz, Nx = G.Nx;
        const int nyz = Ny * Nz;
        const int tbase = yb * (int)blockDim.x;
        const int t = tbase + (int)threadIdx.x;
        const bool active = t < nyz;
        const int x0 = xc * G.xchunk;
        const int x1 = min(x0 + G.xchunk, Nx);
        const float* __restrict__ u = u0 + (size_t)rhs * vstride;
        const unsigned nnu = (unsigned)G.nn;
        const int halo = Nz + 1;
        const int W = (int)blockDim.x + 2 * halo;
        const int planeT = W * 3;
        const int ncell = (int)blockDim.x + halo;                     // cells touching the block's nodes
        const int ntile = (ncell + kCellTile - 1) / kCellTile;
        float* lds = reinterpret_cast<float*>(smem_raw);
        float* fl = lds + 3 * planeT;                                  // (ncell, 24) products
        const __amdgpu_buffer_rsrc_t ru = make_rsrc(u, 3u * nnu * 4u);
        const __amdgpu_buffer_rsrc_t rm = make_rsrc(mask, nnu);
        const __amdgpu_buffer_rsrc_t re = make_rsrc(ecell, (unsigned)G.ncell * 4u);
        const int lane = (int)threadIdx.x & 63;
        const int wave = (int)threadIdx.x >> 6;
        const int nwave = (int)blockDim.x >> 6;
        const int jc = lane & 31;          // cell column of the MFMA tile
        const int hh = lane >> 5;          // which half of the K slice
        // A operand: Ke row (lane & 31) (rows >= 24 zero), K columns 8s + 4hh + i
        half4 A[3];
#pragma unroll
        for (int s = 0; s < 3; ++s)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = jc, col = 8 * s + 4 * hh + i;
                A[s][i] = row < 24 ? (_Float16)ke[row * 24 + col] : (_Float16)0.0f;
            }
        constexpr int MAXE = 4;
        float pre[MAXE][3];
        auto fetch = [&](int x) {
            const bool inx = x >= 0 && x < Nx;
#pragma unroll
            for (int k = 0; k < MAXE; ++k) {
                const int e = (int)threadIdx.x + k * (int)blockDim.x;
                const int f = tbase - halo + e;
                const bool ok = inx && e < W && f >= 0 && f < nyz;
                const unsigned n = (unsigned)(x * nyz + f);
                const int m = bload_u8(rm, ok ? n : kOOB);
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    pre[k][c] = bload<float>(ru, ((m >> c) & 1) ? (c * nnu + n) * 4u : kOOB);
            }
        };
        auto store = [&](int buf) {
            float* p = lds + buf * planeT;
#pragma unroll
            for (int k = 0; k < MAXE; ++k) {
                const int e = (int)threadIdx.x + k * (int)blockDim.x;
                if (e < W) {
                    p[e * 3 + 0] = pre[k][0];
                    p[e * 3 + 1] = pre[k][1];
                    p[e * 3 + 2] = pre[k][2];
                }
            }
        };
        // products of cell layer X (planes in buffers bl = X, bh = X+1) into fl
        auto contract = [&](int X, int bl, int bh) {
            const bool inx = X >= 0 && X < G.nx;
            const float* PL = lds + bl * planeT;
            const float* PH = lds + bh * planeT;
            for (int tile = wave; tile < ntile; tile += nwave) {
                const int c = tile * kCellTile + jc;          // cell slot; its corner-0 node is staged at c
                const int p = tbase - halo + c;               // flattened (y, z) of the cell's corner 0
                const int cy = p >= 0 ? p / Nz : -1;
                const int cz = p - cy * Nz;
                const bool cok = inx && c < ncell && p >= 0 && cy < G.ny && cz < G.nz;
                const float ec = bload<float>(re, cok ? (unsigned)((X * G.ny + cy) * G.nz + cz) * 4u : kOOB);
                const int cs = c < ncell ? c : 0;             // stay inside the staged planes
                const float* q0 = PL + cs * 3;
                const float u00 = q0[0], u01 = q0[1], u02 = q0[2];
                half4 B[3];
#pragma unroll
                for (int s = 0; s < 3; ++s)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int k = 8 * s + 4 * hh + i;     // 3 * corner + component
                        const int b = k / 3, comp = k - 3 * (k / 3);
                        const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
                        const float* q = (bx ? PH : PL) + (cs + by * Nz + bz) * 3;
                        const float ref = comp == 0 ? u00 : (comp == 1 ? u01 : u02);
                        B[s][i] = (_Float16)(q[comp] - ref);
                    }
                floatx16 acc = {};
#pragma unroll
                for (int s = 0; s < 3; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x8f16(A[s], B[s], acc, 0, 0, 0);
                // lane holds rows 8*(r/4) + 4*hh + (r%4) of cell column jc
                float* dst = fl + cs * 24;
                if (c < ncell)
#pragma unroll
                for (int g = 0; g < 3; ++g) {
                    const int row = 8 * g + 4 * hh;
                    dst[row + 0] = ec * acc[4 * g + 0];
                    dst[row + 1] = ec * acc[4 * g + 1];
                    dst[row + 2] = ec * acc[4 * g + 2];
                    dst[row + 3] = ec * acc[4 * g + 3];
                }
            }
        };
        float a0 = 0, a1 = 0, a2 = 0, n0 = 0, n1 = 0, n2 = 0;
        // the node's rows from its 4 cells of the layer just contracted
        auto gather = [&](bool lo) {
            n0 = 0; n1 = 0; n2 = 0;
            const int me = (int)threadIdx.x;
#pragma unroll
            for (int cy = 0; cy < 2; ++cy)
#pragma unroll
                for (int cz = 0; cz < 2; ++cz) {
                    const float* f = fl + (me + cy * Nz + cz) * 24;
                    const int alo = ((1 - cy) << 1) | ((1 - cz) << 2);
                    if (lo) {
                        a0 += f[3 * alo + 0];
                        a1 += f[3 * alo + 1];
                        a2 += f[3 * alo + 2];
                    }
                    const int ahi = alo | 1;
                    n0 += f[3 * ahi + 0];
                    n1 += f[3 * ahi + 1];
                    n2 += f[3 * ahi + 2];
                }
        };

        fetch(x0 - 1);
        store(0);
        fetch(x0);
        store(1);
        fetch(x0 + 1);
        store(2);
        __syncthreads();
        contract(x0 - 1, 0, 1);
        __syncthreads();
        gather(false);
        a0 = n0; a1 = n1; a2 = n2;
        __syncthreads();
        int bl = 1;
        for (int X = x0; X < x1; ++X) {
            const int bh = bl == 2 ? 0 : bl + 1;
            const int bn = bh == 2 ? 0 : bh + 1;
            const bool more = X + 1 < x1;
            if (more) fetch(X + 2);
            contract(X, bl, bh);
            __syncthreads();
            gather(true);
            if (active) {
                const size_t n = (size_t)X * nyz + t;
                const int m = mask[n];
                float yv[3] = {a0, a1, a2};
                const float* q = lds + bl * planeT + ((int)threadIdx.x + halo) * 3;
                float ur[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const bool mc = (m >> c) & 1;
                    ur[c] = (!mc || need_raw_all) ? u[c * (size_t)nnu + n] : q[c];
                    if (!mc) yv[c] = ur[c];
                }
                epilogue<float>(E, m, (size_t)rhs * vstride + n, G.nn, n, yv, ur, part);
            }
            a0 = n0; a1 = n1; a2 = n2;
            if (more) store(bn);
            __syncthreads();
            bl = bh;
        }
    }
    if (E.partials != nullptr) block_partial(part, E, rhs, L);
}

}  // namespace

size_t mfma_smem_bytes(int Nz) {
    const int halo = Nz + 1;
    const int W = kBlock + 2 * halo;
    const int ncell = kBlock + halo;
    return (size_t)(3 * W * 3 + ncell * 24) * sizeof(float);
}

void launch_apply_fine_mfma(const Grid& Gd, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                            hipStream_t s) {
    const Level& lv = Gd.lv[level];
    AGeom G = make_apply_geom(Gd, level);
    const int W = kBlock + 2 * (G.Nz + 1);
    if (W > 4 * kBlock) {
        set_error("fine level: z extent too large for the staged plane");
        throw ArgFailure();
    }
    dim3 grid(8 * G.per, 1, nrhs), block(kBlock);
    hipLaunchKernelGGL(k_apply_fine_mfma, grid, block, mfma_smem_bytes(G.Nz), s, static_cast<const float*>(u), vstride,
                       lv.mask.as<uint8_t>(), G, lv.e32.as<float>(), Gd.ke32.as<float>(), E);
    LP_HIP(hipGetLastError());
}

}  // namespace lp
