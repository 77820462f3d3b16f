// Fine-level operator for the mixed-precision V-cycle: the shared 24x24 Ke
// times every cell's corner vector as an FP16 MFMA contraction with FP32
// accumulation, then a node-centric sum in LDS.
//
// Reference: the same operator as mg._Level.apply on the fine level
// (mg.py:251-346; mgops._product_scaled + _scatter_into, mgops.py:80-217),
// which the reference evaluates in FP32 inside the V-cycle.
//
// Tiling: a block owns a 15 x 15 tile of nodes in the (y, z) plane and marches
// an x-chunk of node planes. The 16 x 16 = 256 cells touching the tile are
// exactly 8 MFMA column tiles of 32 cells (two per wave); the 17 x 17 node
// patch they read is staged in LDS (three x-planes, triple-buffered, the
// next plane prefetched into registers during the contraction).
//
// Per cell layer:
//  - B (32 cells x 24) = each cell's corner vector relative to its corner 0
//    (Ke annihilates translations, so the product is unchanged) times a
//    power of two chosen per cell from its largest entry, so FP16 carries the
//    cell's deformation at full relative precision whatever its magnitude (a
//    void cell's correction can be 1e9 times a solid one's);
//  - A = Ke with its rows permuted so a lane's 12 output rows are the 4
//    corners on one side of the layer (x-low in lanes 0-31, x-high in lanes
//    32-63): 3 v_mfma_f32_32x32x8_f16 per 32 cells;
//  - the products, times E / scale (exact), go to LDS as (row, cell) with a
//    row stride of 8 mod 16 floats (conflict-free in both half-waves); every
//    node then sums its 4 cells' x-low rows (finishing the node plane) and
//    x-high rows (starting the next plane), and runs the epilogue.
#include "apply_impl.h"

namespace lp {

namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int TN = 15;              // tile nodes per side (y, z)
constexpr int TC = TN + 1;          // cells per side
constexpr int TS = TN + 2;          // staged nodes per side
constexpr int NSTAGE = TS * TS;     // 289 staged nodes per plane
constexpr int PSTR = 292;           // component stride of a staged plane (floats)
constexpr int NCELL = TC * TC;      // 256 cells = 8 tiles of 32
constexpr int FSTR = 264;           // row stride of the product buffer (= 8 mod 16)
constexpr int kThreads = 256;
static_assert(NCELL == 8 * 32, "cells per layer must fill the MFMA tiles");
static_assert(2 * kThreads >= NSTAGE, "two staging entries per thread");

// Ke row held by output row R (R < 24) of the MFMA: rows with (R >> 2) & 1 == 0
// are the x-low corners (0, 2, 4, 6), the others the x-high corners (1, 3, 5, 7);
// slot q = (R >> 3) * 4 + (R & 3) runs over corner (by + 2 bz) = q / 3, comp q % 3
__host__ __device__ constexpr int ke_row_of(int R) {
    return 3 * (2 * (((R >> 3) * 4 + (R & 3)) / 3) + ((R >> 2) & 1)) + ((R >> 3) * 4 + (R & 3)) % 3;
}

template <int EK>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) k_apply_fine_mfma(const float* __restrict__ u0, size_t vstride,
                                                              const uint8_t* __restrict__ mask, AGeom G,
                                                              const float* __restrict__ ecell,
                                                              const float* __restrict__ ke, EpiArgs E) {
    __shared__ float planes[3 * 3 * PSTR];
    __shared__ float fl[24 * FSTR];
    __shared__ float fsc[NCELL];        // per cell: E / (its FP16 operand scale)
    using namespace apply_detail;
    const int rhs = blockIdx.z;
    const int L = blockIdx.x;
    const int logical = (L & 7) * G.per + (L >> 3);
    double part = 0.0;
    const bool skip = (E.scal != nullptr && !E.scal[rhs].active) || logical >= G.nblk;
    constexpr bool need_raw_all = EK == EPI_STORE_DOT || EK == EPI_JACOBI || EK == EPI_POWER;
    if (!skip) {
        const int ntz = G.ntz;
        const int xc = logical / G.nyzb;
        const int rem = logical - xc * G.nyzb;
        const int ty = rem / ntz;
        const int y0 = ty * TN, z0 = (rem - ty * ntz) * TN;
        const int Ny = G.Ny, Nz = G.Nz, Nx = G.Nx;
        const int nyz = Ny * Nz;
        const int x0 = xc * G.xchunk;
        const int x1 = min(x0 + G.xchunk, Nx);
        const float* __restrict__ u = u0 + (size_t)rhs * vstride;
        const unsigned nnu = (unsigned)G.nn;
        const __amdgpu_buffer_rsrc_t ru = make_rsrc(u, 3u * nnu * 4u);
        const __amdgpu_buffer_rsrc_t rm = make_rsrc(mask, nnu);
        const __amdgpu_buffer_rsrc_t re = make_rsrc(ecell, (unsigned)G.ncell * 4u);
        const int tid = (int)threadIdx.x;
        const int lane = tid & 63;
        const int wave = tid >> 6;
        const int jc = lane & 31;          // cell column of the MFMA tile / output row of A
        const int hh = lane >> 5;          // K half (A, B) and output row half (D)
        half4 A[3];
#pragma unroll
        for (int s = 0; s < 3; ++s)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                A[s][i] = jc < 24 ? (_Float16)ke[ke_row_of(jc) * 24 + 8 * s + 4 * hh + i] : (_Float16)0.0f;

        // staging: entries tid and tid + 256 of the 17 x 17 patch
        float pre[2][3];
        auto fetch = [&](int x) {
            const bool inx = x >= 0 && x < Nx;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int e = tid + k * kThreads;
                const int sj = e / TS, sk = e - (e / TS) * TS;
                const int y = y0 - 1 + sj, z = z0 - 1 + sk;
                const bool ok = inx && e < NSTAGE && y >= 0 && y < Ny && z >= 0 && z < Nz;
                const unsigned n = (unsigned)(x * nyz + y * Nz + z);
                const int m = bload_u8(rm, ok ? n : kOOB);
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    pre[k][c] = bload<float>(ru, ((m >> c) & 1) ? (c * nnu + n) * 4u : kOOB);
            }
        };
        auto store = [&](int buf) {
            float* p = planes + buf * 3 * PSTR;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int e = tid + k * kThreads;
                if (e < NSTAGE) {
                    p[e] = pre[k][0];
                    p[PSTR + e] = pre[k][1];
                    p[2 * PSTR + e] = pre[k][2];
                }
            }
        };
        // products of cell layer X (node planes in buffers bl = X, bh = X + 1)
        // moduli of this lane's two cells (slots wave*32 + jc and that + 128)
        // of cell layer X; 0 outside the grid
        auto ecell_load = [&](int X, int tt) -> float {
            const int c = (wave + 4 * tt) * 32 + jc;
            const int cy = y0 - 1 + (c >> 4), cz = z0 - 1 + (c & 15);
            const bool cok = X >= 0 && X < G.nx && cy >= 0 && cy < G.ny && cz >= 0 && cz < G.nz;
            return bload<float>(re, cok ? (unsigned)((X * G.ny + cy) * G.nz + cz) * 4u : kOOB);
        };
        float ec0 = 0.0f, ec1 = 0.0f;
        auto contract = [&](int X, int bl, int bh) {
            const bool inx = X >= 0 && X < G.nx;
            const float* PL = planes + bl * 3 * PSTR;
            const float* PH = planes + bh * 3 * PSTR;
#pragma unroll 1
            for (int tt = 0; tt < 2; ++tt) {
                const int c = (wave + 4 * tt) * 32 + jc;      // cell slot
                const int cj = c >> 4, ck = c & 15;
                // a cell outside the grid has E = 0 (its staged corners are
                // finite), so its products are exact zeros
                const float ec = inx ? ec0 : 0.0f;
                ec0 = ec1;                                    // the next tile's cell
                const int s0 = cj * TS + ck;                  // staged index of corner 0
                float d[12];
                float dmax = 0.0f;
#pragma unroll
                for (int s = 0; s < 3; ++s)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int k = 8 * s + 4 * hh + i;     // 3 * corner + component
                        const int b = k / 3, comp = k - 3 * (k / 3);
                        const int bx = b & 1, by = (b >> 1) & 1, bz = (b >> 2) & 1;
                        const float* P = (bx ? PH : PL) + comp * PSTR;
                        d[4 * s + i] = P[s0 + by * TS + bz] - PL[comp * PSTR + s0];
                        dmax = fmaxf(dmax, fabsf(d[4 * s + i]));
                    }
                dmax = fmaxf(dmax, __shfl_xor(dmax, 32));
                int ex;
                (void)frexpf(dmax, &ex);                      // dmax = f 2^ex, f in [0.5, 1)
                const float sc = ldexpf(1.0f, 15 - ex);       // |d| sc < 2^15 < FP16 max
                if (hh == 0) fsc[c] = ec * ldexpf(1.0f, ex - 15);
                half4 B[3];
#pragma unroll
                for (int s = 0; s < 3; ++s)
#pragma unroll
                    for (int i = 0; i < 4; ++i) B[s][i] = (_Float16)(d[4 * s + i] * sc);
                floatx16 acc = {};
#pragma unroll
                for (int s = 0; s < 3; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x8f16(A[s], B[s], acc, 0, 0, 0);
                // lane holds output rows 8 g + 4 hh + i (g < 3): product slots
                // q = 4 g + i of its side of the layer (unscaled; the gather
                // applies the cell's fsc)
                float* dst = fl + hh * 12 * FSTR + c;
#pragma unroll
                for (int q = 0; q < 12; ++q) dst[q * FSTR] = acc[q];
            }
        };
        const bool active = tid < TN * TN;
        const int jn = tid / TN, kn = tid - (tid / TN) * TN;
        const int ny_ = y0 + jn, nz_ = z0 + kn;
        const bool own = active && ny_ < Ny && nz_ < Nz;
        float a0 = 0, a1 = 0, a2 = 0, n0 = 0, n1 = 0, n2 = 0;
        auto gather = [&](bool lo) {
            n0 = 0; n1 = 0; n2 = 0;
            if (!active) return;
#pragma unroll
            for (int by = 0; by < 2; ++by)
#pragma unroll
                for (int bz = 0; bz < 2; ++bz) {
                    const int cs = (jn + 1 - by) * TC + (kn + 1 - bz);
                    const float w = fsc[cs];
                    const float* f = fl + 3 * (by + 2 * bz) * FSTR + cs;
                    if (lo) {
                        a0 = fmaf(w, f[0], a0);
                        a1 = fmaf(w, f[FSTR], a1);
                        a2 = fmaf(w, f[2 * FSTR], a2);
                    }
                    const float* h = f + 12 * FSTR;
                    n0 = fmaf(w, h[0], n0);
                    n1 = fmaf(w, h[FSTR], n1);
                    n2 = fmaf(w, h[2 * FSTR], n2);
                }
        };

        // the epilogue's inputs for node plane X, loaded one layer ahead so
        // their latency hides behind the next contraction
        int em = 0;
        float eraw[3] = {0.0f, 0.0f, 0.0f};
        EpiIn<float> ein;
        auto epi_fetch = [&](int X) {
            if (!own || X >= x1) return;
            const size_t n = (size_t)X * nyz + (size_t)ny_ * Nz + nz_;
            em = bload_u8(rm, (unsigned)n);
#pragma unroll
            for (int c = 0; c < 3; ++c) eraw[c] = bload<float>(ru, (c * nnu + (unsigned)n) * 4u);
            epi_load<float, EK>(E, (size_t)rhs * vstride + n, G.nn, n, ein);
        };

        fetch(x0 - 1);
        store(0);
        fetch(x0);
        store(1);
        fetch(x0 + 1);
        store(2);
        epi_fetch(x0);
        ec0 = ecell_load(x0 - 1, 0);
        ec1 = ecell_load(x0 - 1, 1);
        float en0 = ecell_load(x0, 0), en1 = ecell_load(x0, 1);
        __syncthreads();
        contract(x0 - 1, 0, 1);
        ec0 = en0;
        ec1 = en1;
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
            en0 = ecell_load(X + 1, 0);
            en1 = ecell_load(X + 1, 1);
            contract(X, bl, bh);
            ec0 = en0;
            ec1 = en1;
            __syncthreads();
            gather(true);
            if (own) {
                const size_t n = (size_t)X * nyz + (size_t)ny_ * Nz + nz_;
                float yv[3] = {a0, a1, a2};
                const float* q = planes + bl * 3 * PSTR + (jn + 1) * TS + (kn + 1);
                float ur[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const bool mc = (em >> c) & 1;
                    ur[c] = (!mc || need_raw_all) ? eraw[c] : q[c * PSTR];
                    if (!mc) yv[c] = ur[c];
                }
                epi_finish<float, EK, float>(E, em, (size_t)rhs * vstride + n, G.nn, yv, ur, ein, part);
            }
            epi_fetch(X + 1);
            a0 = n0; a1 = n1; a2 = n2;
            if (more) store(bn);
            __syncthreads();
            bl = bh;
        }
    }
    if (E.partials != nullptr) block_partial(part, E, rhs, L);
}

AGeom mfma_geom(const Grid& Gd, int level) {
    AGeom G = make_apply_geom(Gd, level);
    const int nty = (G.Ny + TN - 1) / TN, ntz = (G.Nz + TN - 1) / TN;
    G.ntz = ntz;
    G.nyzb = nty * ntz;
    G.nxc = (G.Nx + G.xchunk - 1) / G.xchunk;
    G.nblk = G.nyzb * G.nxc;
    G.per = (G.nblk + 7) / 8;
    return G;
}

}  // namespace

int mfma_blocks(const Grid& Gd, int level) { return 8 * mfma_geom(Gd, level).per; }

void launch_apply_fine_mfma(const Grid& Gd, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                            hipStream_t s) {
    const Level& lv = Gd.lv[level];
    const AGeom G = mfma_geom(Gd, level);
    if ((size_t)G.Nx * G.Ny * G.Nz * 3 >= (1ull << 30)) {
        set_error("fine level: too many nodes for 32-bit buffer offsets");
        throw ArgFailure();
    }
    dim3 grid(8 * G.per, 1, nrhs), block(kThreads);
    const float* uf = static_cast<const float*>(u);
    const uint8_t* m = lv.mask.as<uint8_t>();
    const float* e = lv.e32.as<float>();
    const float* ke = Gd.ke32.as<float>();
#define LP_MFMA_CASE(K) \
    case K: hipLaunchKernelGGL(k_apply_fine_mfma<K>, grid, block, 0, s, uf, vstride, m, G, e, ke, E); break;
    switch (E.kind) {
        LP_MFMA_CASE(EPI_STORE)
        LP_MFMA_CASE(EPI_STORE_DOT)
        LP_MFMA_CASE(EPI_CHEB)
        LP_MFMA_CASE(EPI_RESID_SCALED)
        LP_MFMA_CASE(EPI_RESID_MASK)
        LP_MFMA_CASE(EPI_JACOBI)
        LP_MFMA_CASE(EPI_POWER)
        LP_MFMA_CASE(EPI_RESID)
        LP_MFMA_CASE(EPI_MASKED)
        default:
            set_error("fine level: unknown epilogue");
            throw ArgFailure();
    }
#undef LP_MFMA_CASE
    LP_HIP(hipGetLastError());
}

}  // namespace lp
