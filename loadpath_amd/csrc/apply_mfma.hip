// Level operators of the mixed-precision V-cycle as FP16 MFMA contractions
// with FP32 accumulation, then a node-centric sum in LDS:
//  - the fine level: every cell's corner vector times the shared 24x24 Ke,
//    scaled by the cell's modulus;
//  - the first coarse ("weights") level: every coarse cell's corner vector
//    times sum_q w_q M_q, M_q = P_q^T Ke P_q the child maps and w_q the
//    moduli of its 8 fine children; cells touching held dofs use their
//    explicit patch matrices (products from k_patch_prod, run first).
//
// Reference: the same operators as mg._Level.apply on the fine and first
// coarse level (mg.py:251-346; mgops._product_scaled + _scatter_into,
// mgops.py:80-217; the weights level mg.py:609-642), which the reference
// evaluates in FP32 inside the V-cycle.
//
// Tiling: a block owns a 15 x 15 tile of nodes in the (y, z) plane and marches
// an x-chunk of node planes. The 16 x 16 = 256 cells touching the tile are
// exactly 8 MFMA column tiles of 32 cells (two per wave); the 17 x 17 node
// patch they read is staged in LDS (three x-planes, triple-buffered, the
// next plane prefetched into registers during the contraction).
//
// Per cell layer:
//  - B (32 cells x 24) = each cell's corner vector relative to its corner 0
//    (Ke and every M_q annihilate translations, so the product is unchanged)
//    times a power of two chosen per cell from its largest entry, so FP16
//    carries the cell's deformation at full relative precision whatever its
//    magnitude (a void cell's correction can be 1e9 times a solid one's);
//  - A = Ke (or M_q) with its rows permuted so a lane's 12 output rows are
//    the 4 corners on one side of the layer (x-low in lanes 0-31, x-high in
//    lanes 32-63): 3 v_mfma_f32_32x32x8_f16 per 32 cells (per child q on the
//    weights level, whose products are summed with the weights w_q);
//  - the products go to LDS as (row, cell) with a row stride of 8 mod 16
//    floats (conflict-free in both half-waves), with one factor per cell
//    (E / scale); every node then sums its 4 cells' x-low rows (finishing
//    the node plane) and x-high rows (starting the next plane), and runs the
//    epilogue.
#include "apply_impl.h"

namespace lp {

namespace {

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef short short4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

constexpr int TN = 15;              // tile nodes per side (y, z)
constexpr int TC = TN + 1;          // cells per side
constexpr int TS = TN + 2;          // staged nodes per side
constexpr int NSTAGE = TS * TS;     // 289 staged nodes per plane
constexpr int PSTR = 292;           // component stride of a staged plane (floats)
constexpr int NCELL = TC * TC;      // 256 cells = 8 tiles of 32
constexpr int FSTR = 264;           // row stride of the product buffer (= 8 mod 16)
constexpr int kThreads = 256;
constexpr int kWeightsXChunk = 8;   // node planes per block on the weights level (default)
static_assert(NCELL == 8 * 32, "cells per layer must fill the MFMA tiles");
static_assert(2 * kThreads >= NSTAGE, "two staging entries per thread");

// Ke row held by output row R (R < 24) of the MFMA: rows with (R >> 2) & 1 == 0
// are the x-low corners (0, 2, 4, 6), the others the x-high corners (1, 3, 5, 7);
// slot q = (R >> 3) * 4 + (R & 3) runs over corner (by + 2 bz) = q / 3, comp q % 3
__host__ __device__ constexpr int ke_row_of(int R) {
    return 3 * (2 * (((R >> 3) * 4 + (R & 3)) / 3) + ((R >> 2) & 1)) + ((R >> 3) * 4 + (R & 3)) % 3;
}

// Ke row of product slot q (0..11) on side hh of the layer
__device__ __forceinline__ int row_of_slot(int hh, int q) { return ke_row_of(8 * (q >> 2) + 4 * hh + (q & 3)); }

struct MfmaSrc {
    const float* e;       // fine: cell moduli; weights: the finer level's cell moduli
    const float* a;       // fine: Ke (576); weights: M_q (8 x 576)
    const int* pidx;      // weights: patch slot per cell or -1
    const float* pprod;   // weights: (nrhs, npatch, 24) patch products of this apply
    int npatch;
};

template <int EK, bool WTS>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WTS ? 3 : 4, 4)))
k_apply_mfma(const float* __restrict__ u0, size_t vstride, const uint8_t* __restrict__ mask, AGeom G, MfmaSrc S,
             EpiArgs E) {
    __shared__ float planes[3 * 3 * PSTR];
    __shared__ float fl[24 * FSTR];
    __shared__ float fsc[NCELL];        // per cell: its products' common factor
    __shared__ __align__(8) _Float16 alds[WTS ? 8 * 3 * 64 * 4 : 4];
    using namespace apply_detail;
    int rhs, L;
    mfma_block_index(G, rhs, L);
    const int logical = (L & 7) * G.per + (L >> 3);
    double part = 0.0;
    const bool skip = (E.scal != nullptr && !E.scal[rhs].active) || logical >= G.nblk;
    constexpr bool need_raw_all = EK == EPI_STORE_DOT || EK == EPI_JACOBI || EK == EPI_POWER;
    // Chebyshev steps that stage their operand d = (dinv u) s from the vector
    // u they are given (b for the first step from zero, r for the last step
    // after a residual-only step)
    constexpr bool STAGE_B = EK == EPI_CHEB_FIRST || EK == EPI_CHEB_ONLY || EK == EPI_CHEB_LAST_R;
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
        const size_t ne = WTS ? (size_t)G.fnx * G.fny * G.fnz : G.ncell;
        const __amdgpu_buffer_rsrc_t re = make_rsrc(S.e, (unsigned)ne * 4u);
        const __amdgpu_buffer_rsrc_t rdv = make_rsrc(STAGE_B ? E.dinv : nullptr, STAGE_B ? 3u * nnu * 4u : 0u);
        const float s_b = (float)E.c3;
        const int tid = (int)threadIdx.x;
        const int lane = tid & 63;
        const int wave = tid >> 6;
        const int jc = lane & 31;          // cell column of the MFMA tile / output row of A
        const int hh = lane >> 5;          // K half (A, B) and output row half (D)
        half4 A[3];
        if (WTS) {
            // the 8 child maps as FP16 A operands, lane-ordered: (q, s, lane, i)
            for (int k = tid; k < 8 * 3 * 64 * 4; k += kThreads) {
                const int i = k & 3, l = (k >> 2) & 63, s = (k >> 8) % 3, q = k / 768;
                const int row = l & 31, col = 8 * s + 4 * (l >> 5) + i;
                alds[k] = row < 24 ? (_Float16)S.a[q * 576 + ke_row_of(row) * 24 + col] : (_Float16)0.0f;
            }
        } else {
#pragma unroll
            for (int s = 0; s < 3; ++s)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    A[s][i] = jc < 24 ? (_Float16)S.a[ke_row_of(jc) * 24 + 8 * s + 4 * hh + i] : (_Float16)0.0f;
        }

        // staging: entries tid and tid + 256 of the 17 x 17 patch
        float pre[2][3];
        // dinv of the staged-b kinds, multiplied in at the LDS store
        // (multiplying at the load would wait for each load in turn)
        float pdv[STAGE_B ? 2 : 1][3];
        // the plane's loads do not depend on its mask (applied at the store):
        // they are in flight together; outside the grid an offset with bit
        // 31 set reads zero (an arithmetic form: a select of the offset
        // becomes branches around the loads)
        int pm[2];
        auto fetch = [&](int x) {
            const bool inx = x >= 0 && x < Nx;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int e = tid + k * kThreads;
                const int sj = e / TS, sk = e - (e / TS) * TS;
                const int y = y0 - 1 + sj, z = z0 - 1 + sk;
                const unsigned oob = (unsigned)!(inx && e < NSTAGE && y >= 0 && y < Ny && z >= 0 && z < Nz) << 31;
                const unsigned n = (unsigned)(x * nyz + y * Nz + z);
                pm[k] = bload_u8(rm, n | oob);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const unsigned off = ((c * nnu + n) * 4u) | oob;
                    pre[k][c] = bload<float>(ru, off);
                    if (STAGE_B) pdv[STAGE_B ? k : 0][c] = bload<float>(rdv, off);
                }
            }
        };
        auto store = [&](int buf) {
            float* p = planes + buf * 3 * PSTR;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const int e = tid + k * kThreads;
                if (e < NSTAGE) {
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        p[c * PSTR + e] = ((pm[k] >> c) & 1) ? (STAGE_B ? (pdv[STAGE_B ? k : 0][c] * pre[k][c]) * s_b : pre[k][c]) : 0.0f;
                }
            }
        };
        // fine level: moduli of this lane's two cells (slots wave*32 + jc and
        // that + 128) of cell layer X, loaded a layer ahead; 0 outside the grid
        auto ecell_load = [&](int X, int tt) -> float {
            if (WTS) return 0.0f;
            const int c = (wave + 4 * tt) * 32 + jc;
            const int cy = y0 - 1 + (c >> 4), cz = z0 - 1 + (c & 15);
            const bool cok = X >= 0 && X < G.nx && cy >= 0 && cy < G.ny && cz >= 0 && cz < G.nz;
            return bload<float>(re, cok ? (unsigned)((X * G.ny + cy) * G.nz + cz) * 4u : kOOB);
        };
        float ec0 = 0.0f, ec1 = 0.0f;
        // products of cell layer X (node planes in buffers bl = X, bh = X + 1)
        auto contract = [&](int X, int bl, int bh) {
            const bool inx = X >= 0 && X < G.nx;
            const float* PL = planes + bl * 3 * PSTR;
            const float* PH = planes + bh * 3 * PSTR;
#pragma unroll 1
            for (int tt = 0; tt < 2; ++tt) {
                const int c = (wave + 4 * tt) * 32 + jc;      // cell slot
                const int cj = c >> 4, ck = c & 15;
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
                // clamp so the scale stays finite in FP32 (2^115 at most):
                // an infinite scale would turn the exact zero of corner 0
                // into NaN. Below the clamp a cell's largest difference
                // still scales to at least 2^-14 (FP16's smallest normal)
                // for every normal FP32 dmax (>= 2^-126)
                ex = max(ex, -100);
                const float sc = ldexpf(1.0f, 15 - ex);       // |d| sc < 2^15 < FP16 max
                half4 B[3];
#pragma unroll
                for (int s = 0; s < 3; ++s)
#pragma unroll
                    for (int i = 0; i < 4; ++i) B[s][i] = (_Float16)(d[4 * s + i] * sc);
                float* dst = fl + hh * 12 * FSTR + c;
                if (!WTS) {
                    // a cell outside the grid has E = 0 (its staged corners
                    // are finite), so its products are exact zeros
                    const float ec = inx ? ec0 : 0.0f;
                    ec0 = ec1;                                // the next tile's cell
                    floatx16 acc = {};
#pragma unroll
                    for (int s = 0; s < 3; ++s)
                        acc = __builtin_amdgcn_mfma_f32_32x32x8f16(A[s], B[s], acc, 0, 0, 0);
                    if (hh == 0) fsc[c] = ec * ldexpf(1.0f, ex - 15);
                    // lane holds output rows 8 g + 4 hh + i (g < 3): slots q = 4 g + i
#pragma unroll
                    for (int q = 0; q < 12; ++q) dst[q * FSTR] = acc[q];
                } else {
                    const int cy = y0 - 1 + cj, cz = z0 - 1 + ck;
                    const bool cok = inx && cy >= 0 && cy < G.ny && cz >= 0 && cz < G.nz;
                    const int cell = (X * G.ny + cy) * G.nz + cz;
                    const int pi = cok ? S.pidx[cell] : -1;
                    float y[12];
#pragma unroll
                    for (int q = 0; q < 12; ++q) y[q] = 0.0f;
#pragma unroll 2
                    for (int ch = 0; ch < 8; ++ch) {
                        const int fx = 2 * X + (ch & 1), fy = 2 * cy + ((ch >> 1) & 1), fz = 2 * cz + ((ch >> 2) & 1);
                        const bool wok = cok && pi < 0 && fx < G.fnx && fy < G.fny && fz < G.fnz;
                        const float w = bload<float>(re, wok ? (unsigned)((fx * G.fny + fy) * G.fnz + fz) * 4u : kOOB);
                        floatx16 acc = {};
#pragma unroll
                        for (int s = 0; s < 3; ++s) {
                            const half4 a = *reinterpret_cast<const half4*>(alds + ((ch * 3 + s) * 64 + lane) * 4);
                            acc = __builtin_amdgcn_mfma_f32_32x32x8f16(a, B[s], acc, 0, 0, 0);
                        }
#pragma unroll
                        for (int q = 0; q < 12; ++q) y[q] = fmaf(w, acc[q], y[q]);
                    }
                    if (pi >= 0) {
                        const float* pp = S.pprod + ((size_t)rhs * S.npatch + pi) * 24;
#pragma unroll
                        for (int q = 0; q < 12; ++q) y[q] = pp[row_of_slot(hh, q)];
                    }
                    if (hh == 0) fsc[c] = pi >= 0 ? 1.0f : ldexpf(1.0f, ex - 15);
#pragma unroll
                    for (int q = 0; q < 12; ++q) dst[q * FSTR] = y[q];
                }
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
            fetch(more ? X + 2 : -2);   // past the chunk: outside the grid, no traffic
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
                    if (STAGE_B) {
                        // d at this node (the staged operand, or computed for a
                        // masked-out dof, whose identity row returns it); b for
                        // the residual
                        const float b = eraw[c];
                        ur[c] = mc ? q[c * PSTR] : (ein.v[3][c] * b) * s_b;
                        ein.v[EK == EPI_CHEB_LAST_R ? 2 : 0][c] = b;
                    } else {
                        ur[c] = (!mc || need_raw_all) ? eraw[c] : q[c * PSTR];
                    }
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
    if (E.partials != nullptr) block_partial(part, E, rhs, L, 8 * G.per);
}

// patch cells of the weights level: (patch, row) per thread, the explicit
// 24x24 patch matrix times the cell's masked corner vector (FP32) for every
// right-hand side of the batch (the matrix row, masks and indices are read
// once); with dinv set, the operand is (dinv u) s, as the fused first
// Chebyshev step stages it
__global__ void __launch_bounds__(kBlock) k_patch_prod(const float* __restrict__ u0, size_t vstride, int nrhs,
                                                       const uint8_t* __restrict__ mask, Geom g,
                                                       const int* __restrict__ cells, const float* __restrict__ pm,
                                                       int npatch, const float* __restrict__ dinv, float s,
                                                       float* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npatch * 24) return;
    const int p = t / 24, row = t - 24 * (t / 24);
    const int cell = cells[p];
    const int Z = cell % g.nz, Y = (cell / g.nz) % g.ny, X = cell / (g.nz * g.ny);
    const float* m = pm + (size_t)p * 576 + row * 24;
    // per operand: its matrix entry (0 for a masked-out dof), offset and scale
    float w[24], f[24];
    size_t off[24];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const size_t n = ((size_t)(X + (b & 1)) * g.Ny + (Y + ((b >> 1) & 1))) * g.Nz + (Z + ((b >> 2) & 1));
        const int mk = mask[n];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const bool on = (mk >> c) & 1;
            w[3 * b + c] = on ? m[3 * b + c] : 0.0f;
            off[3 * b + c] = c * g.nn + n;
            f[3 * b + c] = (on && dinv) ? dinv[c * g.nn + n] : 1.0f;
        }
    }
    for (int rhs = 0; rhs < nrhs; ++rhs) {
        const float* u = u0 + (size_t)rhs * vstride;
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < 24; ++k) {
            const float v = w[k] != 0.0f ? u[off[k]] : 0.0f;
            acc = fmaf(w[k], dinv ? (f[k] * v) * s : v, acc);
        }
        out[((size_t)rhs * npatch + p) * 24 + row] = acc;
    }
}

// ---------------------------------------------------------------------------
// Wave-tile variant: the same contraction, with the node sums formed in
// registers. Each wave owns a 15 x 15 node tile in (y, z) (16 x 16 cells)
// and marches an x-chunk on its own (no block barriers after set-up). Lane
// (hh, r, c) = (lane >> 5, (lane >> 4) & 1, lane & 15); MFMA tile t (0..7)
// is cell rows Y0 - 1 + 2t + r, cell columns Z0 - 1 + c. The B half hh holds
// the corners with bx = hh (its cell's 4 nodes in plane X + hh), the A
// columns permuted to match, so after the contraction lane (hh, r, c) holds
// its cell's 12 outputs for the corners with bx = hh. A node then sums:
//  - in z, its cell (bz = 0 corners) and the left neighbour's (bz = 1):
//    one DPP row_shr:1 add;
//  - in y, the cell row below's by = 1 part: from the other 16-lane row of
//    the tile or, for row 0, the previous tile's row 1 (v_permlane16_swap);
//  - in x, the x-high part of the previous cell layer, kept in registers
//    per tile and brought across the wave halves with v_permlane32_swap, so
//    the lanes of half hh finish tile 2p + hh of each tile pair p.
// The staged planes are private to the wave (two, alternating).
constexpr int WN = 15;                  // nodes per wave tile side
constexpr int WS = WN + 2;              // staged nodes per side
constexpr int WPL = WS * WS;            // 289
constexpr int WCS = 292;                // component stride of a staged plane
constexpr int WPS = 3 * WCS;            // plane stride
constexpr int kWaveTiles = kThreads / 64;
constexpr int WSTAGE = (WPL + 63) / 64; // staging entries per lane

// two FP32 values to a packed pair of BF16 (round to nearest even), lo in
// the low half: a 2-vector conversion is one v_cvt_pk_bf16_f32 (a 4-vector
// one is lowered value by value with permutes)
typedef float float2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned pk_bf16(float lo, float hi) {
    const float2v v = {lo, hi};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
}

__device__ __forceinline__ float dpp_shr1(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, false));
}

// v_permlane16_swap: odd 16-lane rows of a <-> even rows of b
__device__ __forceinline__ void swap16(float& a, float& b) {
    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

// v_permlane32_swap: lanes 32-63 of a <-> lanes 0-31 of b
__device__ __forceinline__ void swap32(float& a, float& b) {
    auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}

// BF (fine level only): BF16 operands instead of the scaled FP16 ones. BF16
// has FP32's exponent range, so a cell needs no scale of its own: the
// operand is the corner differences as they are and the products times the
// cell's modulus are its outputs (8 mantissa bits instead of 11, and none of
// the per-cell maximum and power-of-two scale instructions).
template <int EK, bool WTS, bool BF = false>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3, 4)))
k_apply_wave(const float* __restrict__ u0, size_t vstride, const uint8_t* __restrict__ mask, AGeom G, MfmaSrc S,
             EpiArgs E) {
    static_assert(!(BF && WTS), "BF16 operands on the fine level only");
    __shared__ float stage[kWaveTiles * 2 * WPS];
    __shared__ __align__(8) _Float16 alds[WTS ? 8 * 3 * 64 * 4 : 4];
    using namespace apply_detail;
    int rhs, L;
    mfma_block_index(G, rhs, L);
    const int logical = (L & 7) * G.per + (L >> 3);
    const int tid = (int)threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wt = logical * kWaveTiles + wave;
    double part = 0.0;
    const bool bskip = (E.scal != nullptr && !E.scal[rhs].active) || logical >= G.nblk;
    constexpr bool need_raw_all = EK == EPI_STORE_DOT || EK == EPI_JACOBI || EK == EPI_POWER;
    constexpr bool STAGE_B = EK == EPI_CHEB_FIRST || EK == EPI_CHEB_ONLY || EK == EPI_CHEB_LAST_R;
    const int jc = lane & 31, hh = lane >> 5, r = (lane >> 4) & 1, c = lane & 15;
    half4 A[3];
    short4v Ab[3];
    // fine level: the contraction as one v_mfma_f32_32x32x16_f16 (gfx950's
    // K = 16 form, twice the rate of K = 8) and one 32x32x8; lane half hh
    // holds its 12 operands q = 0..11 as k = 8 hh + q (q < 8) and
    // 16 + 4 hh + q - 8, the Ke columns permuted to match
    half8 A16;
    half4 A8;
    if (WTS) {
        if (!bskip)
            for (int k = tid; k < 8 * 3 * 64 * 4; k += kThreads) {
                const int i = k & 3, l = (k >> 2) & 63, s = (k >> 8) % 3, q = k / 768;
                const int row = l & 31, col = 8 * s + 4 * (l >> 5) + i;
                alds[k] = row < 24 ? (_Float16)S.a[q * 576 + ke_row_of(row) * 24 + ke_row_of(col)] : (_Float16)0.0f;
            }
        __syncthreads();
    } else if (BF) {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            bf16x4 t;
#pragma unroll
            for (int i = 0; i < 4; ++i)
                t[i] = (__bf16)(jc < 24 ? S.a[ke_row_of(jc) * 24 + ke_row_of(8 * s + 4 * hh + i)] : 0.0f);
            Ab[s] = __builtin_bit_cast(short4v, t);
        }
    } else {
        // operand q of lane half hh sits at K index 8 (q >> 2) + 4 hh + (q & 3)
        // of the three-K=8 form, whose Ke column is ke_row_of of it
#pragma unroll
        for (int q = 0; q < 8; ++q)
            A16[q] = jc < 24 ? (_Float16)S.a[ke_row_of(jc) * 24 + ke_row_of(8 * (q >> 2) + 4 * hh + (q & 3))]
                             : (_Float16)0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            A8[i] = jc < 24 ? (_Float16)S.a[ke_row_of(jc) * 24 + ke_row_of(16 + 4 * hh + i)] : (_Float16)0.0f;
    }
    if (!bskip && wt < G.nwt) {
        const int xc = wt / G.nyzb;
        const int rem = wt - xc * G.nyzb;
        const int ty = rem / G.ntz;
        const int Y0 = ty * WN, Z0 = (rem - ty * G.ntz) * WN;
        const int Ny = G.Ny, Nz = G.Nz, Nx = G.Nx;
        const int nyz = Ny * Nz;
        const int x0 = xc * G.xchunk;
        const int x1 = min(x0 + G.xchunk, Nx);
        const float* __restrict__ u = u0 + (size_t)rhs * vstride;
        const unsigned nnu = (unsigned)G.nn;
        const __amdgpu_buffer_rsrc_t ru = make_rsrc(u, 3u * nnu * 4u);
        const __amdgpu_buffer_rsrc_t rm = make_rsrc(mask, nnu);
        const size_t ne = WTS ? (size_t)G.fnx * G.fny * G.fnz : G.ncell;
        const __amdgpu_buffer_rsrc_t re = make_rsrc(S.e, (unsigned)ne * 4u);
        const __amdgpu_buffer_rsrc_t rdv = make_rsrc(STAGE_B ? E.dinv : nullptr, STAGE_B ? 3u * nnu * 4u : 0u);
        const float s_b = (float)E.c3;
        float* SB = stage + wave * 2 * WPS;

        // staging: entries lane + 64 k of the 17 x 17 patch; goff = the
        // node's (y, z) offset in a plane, or -1 outside the grid
        int goff[WSTAGE];
#pragma unroll
        for (int k = 0; k < WSTAGE; ++k) {
            const int e = lane + 64 * k;
            const int sj = e / WS, sk = e - (e / WS) * WS;
            const int y = Y0 - 1 + sj, z = Z0 - 1 + sk;
            goff[k] = (e < WPL && y >= 0 && y < Ny && z >= 0 && z < Nz) ? y * Nz + z : -1;
        }
        float pre[WSTAGE][3];
        // dinv of the staged-b kinds, multiplied in when the plane is written
        // to LDS (multiplying at the load would wait for each load in turn)
        float pdv[STAGE_B ? WSTAGE : 1][3];
        auto fetch = [&](int x) {
            const bool inx = x >= 0 && x < Nx;
#pragma unroll
            for (int k = 0; k < WSTAGE; ++k) {
                const bool ok = inx && goff[k] >= 0;
                const unsigned n = (unsigned)(x * nyz + goff[k]);
                const int m = bload_u8(rm, ok ? n : kOOB);
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    const unsigned off = ((m >> cc) & 1) ? (cc * nnu + n) * 4u : kOOB;
                    pre[k][cc] = bload<float>(ru, off);
                    if (STAGE_B) pdv[STAGE_B ? k : 0][cc] = bload<float>(rdv, off);
                }
            }
        };
        auto store = [&](int buf) {
            __builtin_amdgcn_wave_barrier();
            float* P = SB + buf * WPS;
#pragma unroll
            for (int k = 0; k < WSTAGE; ++k) {
                const int e = lane + 64 * k;
                if (e < WPL) {
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc)
                        P[cc * WCS + e] = STAGE_B ? (pdv[STAGE_B ? k : 0][cc] * pre[k][cc]) * s_b : pre[k][cc];
                }
            }
            __builtin_amdgcn_wave_barrier();
        };

        // one tile's contraction on cell layer X (planes in buffers bl = X,
        // bh = X + 1): nv = the node sums of this layer's side hh, u0v = the
        // staged values at the lane's corner 0 (node (X, cell row, cell col))
        float ycar[3] = {0.0f, 0.0f, 0.0f};   // y carry: previous tile's by = 1 sums
        auto tile = [&](int X, int t, int bl, int bh, float ecell, float (&nv)[3], float (&u0v)[3]) {
            const float* PL = SB + bl * WPS;
            const float* PO = SB + (hh ? bh : bl) * WPS;
            const int s0 = (2 * t + r) * WS + c;
            float d[12];
            float dmax = 0.0f;
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) u0v[cc] = PL[cc * WCS + s0];
#pragma unroll
            for (int q = 0; q < 12; ++q) {
                const int b4 = q / 3, cc = q % 3;
                d[q] = PO[cc * WCS + s0 + (b4 & 1) * WS + (b4 >> 1)] - u0v[cc];
                if (!BF) dmax = fmaxf(dmax, fabsf(d[q]));
            }
            float y[12];
            if (BF) {
                floatx16 acc = {};
#pragma unroll
                for (int s = 0; s < 3; ++s) {
                    typedef unsigned uint2v __attribute__((ext_vector_type(2)));
                    uint2v bt;
                    bt[0] = pk_bf16(d[4 * s], d[4 * s + 1]);
                    bt[1] = pk_bf16(d[4 * s + 2], d[4 * s + 3]);
                    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(Ab[s], __builtin_bit_cast(short4v, bt), acc, 0, 0, 0);
                }
#pragma unroll
                for (int q = 0; q < 12; ++q) y[q] = acc[q] * ecell;
            } else {
            {
                // a cell's 24 operands sit in lanes l and l + 32: one scale
                float lo = dmax, hi = dmax;
                swap32(lo, hi);
                dmax = fmaxf(lo, hi);
            }
            int ex;
            (void)frexpf(dmax, &ex);
            ex = max(ex, -100);       // a finite scale (see k_apply_mfma)
            const float sc = ldexpf(1.0f, 15 - ex);
            half4 B[3];
            if (WTS) {
#pragma unroll
                for (int s = 0; s < 3; ++s)
#pragma unroll
                    for (int i = 0; i < 4; ++i) B[s][i] = (_Float16)(d[4 * s + i] * sc);
            }
            if (!WTS) {
                half8 B16;
                half4 B8;
#pragma unroll
                for (int q = 0; q < 8; ++q) B16[q] = (_Float16)(d[q] * sc);
#pragma unroll
                for (int i = 0; i < 4; ++i) B8[i] = (_Float16)(d[8 + i] * sc);
                floatx16 acc = {};
                acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(A16, B16, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x8f16(A8, B8, acc, 0, 0, 0);
                const float f = ecell * ldexpf(1.0f, ex - 15);
#pragma unroll
                for (int q = 0; q < 12; ++q) y[q] = acc[q] * f;
            } else {
                const int cy = Y0 - 1 + 2 * t + r, cz = Z0 - 1 + c;
                const bool cok = X >= 0 && X < G.nx && cy >= 0 && cy < G.ny && cz >= 0 && cz < G.nz;
                const int cell = (X * G.ny + cy) * G.nz + cz;
                const int pi = cok ? S.pidx[cell] : -1;
                // the children's weights ride in the FP16 operand: B_q = B
                // (w_q / w_max), so the 8 contractions accumulate in one
                // MFMA chain (a child 2^-24 below the cell's largest drops
                // out: far below the operand rounding)
                auto wload = [&](int ch) -> float {
                    const int fx = 2 * X + (ch & 1), fy = 2 * cy + ((ch >> 1) & 1), fz = 2 * cz + ((ch >> 2) & 1);
                    const bool wok = cok && pi < 0 && fx < G.fnx && fy < G.fny && fz < G.fnz;
                    return bload<float>(re, wok ? (unsigned)((fx * G.fny + fy) * G.fnz + fz) * 4u : kOOB);
                };
                float wmax = 0.0f;
#pragma unroll
                for (int ch = 0; ch < 8; ++ch) wmax = fmaxf(wmax, wload(ch));
                const float rw = wmax > 0.0f ? 1.0f / wmax : 0.0f;
                // two accumulation chains (even and odd children) so the
                // MFMA latency of one overlaps the other
                floatx16 acc0 = {}, acc1 = {};
#pragma unroll 1
                for (int ch = 0; ch < 8; ch += 2) {
                    const _Float16 r0 = (_Float16)(wload(ch) * rw), r1 = (_Float16)(wload(ch + 1) * rw);
#pragma unroll
                    for (int s = 0; s < 3; ++s) {
                        const half4 a0 = *reinterpret_cast<const half4*>(alds + ((ch * 3 + s) * 64 + lane) * 4);
                        const half4 a1 = *reinterpret_cast<const half4*>(alds + (((ch + 1) * 3 + s) * 64 + lane) * 4);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x8f16(a0, B[s] * r0, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x8f16(a1, B[s] * r1, acc1, 0, 0, 0);
                    }
                }
#pragma unroll
                for (int q = 0; q < 12; ++q) y[q] = acc0[q] + acc1[q];
                const float f = wmax * ldexpf(1.0f, ex - 15);
                if (pi >= 0) {
                    const float* pp = S.pprod + ((size_t)rhs * S.npatch + pi) * 24;
#pragma unroll
                    for (int q = 0; q < 12; ++q) y[q] = pp[row_of_slot(hh, q)];
                } else {
#pragma unroll
                    for (int q = 0; q < 12; ++q) y[q] *= f;
                }
            }
            }
            // slot q = 3 (by + 2 bz) + comp; z: own bz = 0 + left neighbour's bz = 1
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const float z0s = y[cc] + dpp_shr1(y[6 + cc]);
                float z1s = y[3 + cc] + dpp_shr1(y[9 + cc]);
                // y: row 1 takes row 0's by = 1 sums, row 0 the previous tile's row 1
                float pv = ycar[cc];
                swap16(pv, z1s);
                nv[cc] = z0s + (r ? pv : z1s);
                ycar[cc] = z1s;
            }
        };

        // the epilogue's inputs of one node, fetched before the pair's tiles
        // (the node's own operand values only where the epilogue needs them:
        // masked-out dofs, the kinds that read it, b of the first Chebyshev
        // step)
        struct Pend {
            int em;
            float raw[3];
            EpiIn<float> in;
            bool own;
            unsigned n;
        };
        auto epi_fetch = [&](int X, int p, Pend& P) {
            const int t = 2 * p + hh;
            const int yn = Y0 - 1 + 2 * t + r, zn = Z0 - 1 + c;
            P.own = X < x1 && c >= 1 && (t > 0 || r > 0) && yn < Ny && zn < Nz;
            P.n = (unsigned)(X * nyz + yn * Nz + zn);
            P.em = 0;
            if (!P.own) return;
            P.em = bload_u8(rm, P.n);
            if (need_raw_all || STAGE_B)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) P.raw[cc] = bload<float>(ru, (cc * nnu + P.n) * 4u);
            epi_load<float, EK>(E, (size_t)rhs * vstride + P.n, G.nn, P.n, P.in);
        };
        auto ecell_load = [&](int X, int t) -> float {
            if (WTS) return 0.0f;
            const int cy = Y0 - 1 + 2 * t + r, cz = Z0 - 1 + c;
            const bool cok = X >= 0 && X < G.nx && cy >= 0 && cy < G.ny && cz >= 0 && cz < G.nz;
            return bload<float>(re, cok ? (unsigned)((X * G.ny + cy) * G.nz + cz) * 4u : kOOB);
        };

        float car[8][3];   // x carry: the previous layer's node sums per tile
        float u0a[3], u0b[3];
        // planes x0 - 1 -> buffer 0, x0 -> 1; the layer before the chunk
        // gives the x carry (nothing before the grid)
        fetch(x0 - 1);
        store(0);
        fetch(x0);
        store(1);
        fetch(x0 + 1);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const float ec = ecell_load(x0 - 1, t);
            if (x0 > 0) {
                tile(x0 - 1, t, 0, 1, ec, car[t], u0a);
            } else {
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) car[t][cc] = 0.0f;
            }
        }
        store(0);
        Pend cur;
        float ea0 = ecell_load(x0, 0), ea1 = ecell_load(x0, 1);
        for (int X = x0; X < x1; ++X) {
            const int bl = (X - x0 + 1) & 1, bh = bl ^ 1;
            const bool more = X + 1 < x1;   // plane X + 2 is needed by layer X + 1
            if (more) fetch(X + 2);
#pragma unroll
            for (int p = 0; p < 4; ++p) {
                epi_fetch(X, p, cur);
                const float e0 = ea0, e1 = ea1;
                if (p < 3) {
                    ea0 = ecell_load(X, 2 * p + 2);
                    ea1 = ecell_load(X, 2 * p + 3);
                } else {
                    ea0 = ecell_load(X + 1, 0);
                    ea1 = ecell_load(X + 1, 1);
                }
                float na[3], nb[3];
                tile(X, 2 * p, bl, bh, e0, na, u0a);
                tile(X, 2 * p + 1, bl, bh, e1, nb, u0b);
                // x: lanes of half hh finish tile 2p + hh
                float yv[3], ur[3];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    float ca = car[2 * p][cc], nbv = nb[cc];
                    swap32(ca, nbv);
                    const float r1 = hh ? ca : na[cc];
                    const float r2 = hh ? car[2 * p + 1][cc] : nbv;
                    yv[cc] = r1 + r2;
                    car[2 * p][cc] = na[cc];
                    car[2 * p + 1][cc] = nbv;
                }
                if (cur.own) {
                    if (!(need_raw_all || STAGE_B) && cur.em != 7)
#pragma unroll
                        for (int cc = 0; cc < 3; ++cc) cur.raw[cc] = bload<float>(ru, (cc * nnu + cur.n) * 4u);
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) {
                        const bool mc = (cur.em >> cc) & 1;
                        const float st = hh ? u0b[cc] : u0a[cc];
                        if (STAGE_B) {
                            const float b = cur.raw[cc];
                            ur[cc] = mc ? st : (cur.in.v[3][cc] * b) * s_b;
                            cur.in.v[EK == EPI_CHEB_LAST_R ? 2 : 0][cc] = b;
                        } else {
                            ur[cc] = (!mc || need_raw_all) ? cur.raw[cc] : st;
                        }
                        if (!mc) yv[cc] = ur[cc];
                    }
                    epi_finish<float, EK, float>(E, cur.em, (size_t)rhs * vstride + cur.n, G.nn, yv, ur, cur.in,
                                                 part);
                }
            }
            // plane X's buffer takes plane X + 2 (loaded during this layer)
            if (more) store(bl);
        }
    }
    if (E.partials != nullptr) block_partial(part, E, rhs, L, 8 * G.per);
}

AGeom mfma_geom(const Grid& Gd, int level) {
    AGeom G = make_apply_geom(Gd, level);
    if (Gd.lv[level].kind == LK_WEIGHTS)
        G.xchunk = Gd.tune_weights_xchunk > 0 ? Gd.tune_weights_xchunk : kWeightsXChunk;
    const int nty = (G.Ny + TN - 1) / TN, ntz = (G.Nz + TN - 1) / TN;
    G.ntz = ntz;
    G.nyzb = nty * ntz;
    if (Gd.lv[level].kind == LK_FINE && Gd.tune_fine_xchunk <= 0)
        G.xchunk = size_xchunk(G.Nx, G.nyzb * (kThreads / 64), kSizeXChunkBlocks);
    G.nxc = (G.Nx + G.xchunk - 1) / G.xchunk;
    G.nblk = G.nyzb * G.nxc;
    G.per = (G.nblk + 7) / 8;
    return G;
}

template <bool WTS>
void launch_mfma(const Grid& Gd, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E, const MfmaSrc& S,
                 hipStream_t s) {
    const Level& lv = Gd.lv[level];
    AGeom G = mfma_geom(Gd, level);
    G.rhs_inner = Gd.tune_rhs_inner ? nrhs : 0;
    if ((size_t)G.Nx * G.Ny * G.Nz * 3 * sizeof(float) > (1ull << 31)) {
        set_error("level operator: too many nodes for 32-bit buffer offsets");
        throw ArgFailure();
    }
    const dim3 grid = G.rhs_inner ? dim3(8 * G.per * nrhs, 1, 1) : dim3(8 * G.per, 1, nrhs), block(kThreads);
    const float* uf = static_cast<const float*>(u);
    const uint8_t* m = lv.mask.as<uint8_t>();
#define LP_MFMA_CASE(K) \
    case K: hipLaunchKernelGGL((k_apply_mfma<K, WTS>), grid, block, 0, s, uf, vstride, m, G, S, E); break;
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
        LP_MFMA_CASE(EPI_CHEB_FIRST)
        LP_MFMA_CASE(EPI_CHEB_ONLY)
        LP_MFMA_CASE(EPI_CHEB_LAST)
        LP_MFMA_CASE(EPI_CHEB_LAST_R)
        default:
            set_error("level operator: unknown epilogue");
            throw ArgFailure();
    }
#undef LP_MFMA_CASE
    LP_HIP(hipGetLastError());
}

// wave-tile kernel: 15 x 15 node tiles per wave, kWaveTiles waves per block
AGeom wave_geom(const Grid& Gd, int level) {
    AGeom G = make_apply_geom(Gd, level);
    if (Gd.lv[level].kind == LK_WEIGHTS)
        G.xchunk = Gd.tune_weights_xchunk > 0 ? Gd.tune_weights_xchunk : kWeightsXChunk;
    const int nty = (G.Ny + WN - 1) / WN, ntz = (G.Nz + WN - 1) / WN;
    G.ntz = ntz;
    G.nyzb = nty * ntz;
    if (Gd.lv[level].kind == LK_FINE && Gd.tune_fine_xchunk <= 0)
        G.xchunk = size_xchunk(G.Nx, G.nyzb, kSizeXChunkBlocks);
    G.nxc = (G.Nx + G.xchunk - 1) / G.xchunk;
    G.nwt = G.nyzb * G.nxc;
    G.nblk = (G.nwt + kWaveTiles - 1) / kWaveTiles;
    G.per = (G.nblk + 7) / 8;
    return G;
}

template <bool WTS, bool BF = false>
void launch_wave(const Grid& Gd, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                 const MfmaSrc& S, hipStream_t s) {
    const Level& lv = Gd.lv[level];
    AGeom G = wave_geom(Gd, level);
    G.rhs_inner = Gd.tune_rhs_inner ? nrhs : 0;
    if ((size_t)G.Nx * G.Ny * G.Nz * 3 * sizeof(float) > (1ull << 31)) {
        set_error("level operator: too many nodes for 32-bit buffer offsets");
        throw ArgFailure();
    }
    const dim3 grid = G.rhs_inner ? dim3(8 * G.per * nrhs, 1, 1) : dim3(8 * G.per, 1, nrhs), block(kThreads);
    const float* uf = static_cast<const float*>(u);
    const uint8_t* m = lv.mask.as<uint8_t>();
#define LP_WAVE_CASE(K) \
    case K: hipLaunchKernelGGL((k_apply_wave<K, WTS, BF>), grid, block, 0, s, uf, vstride, m, G, S, E); break;
    switch (E.kind) {
        LP_WAVE_CASE(EPI_STORE)
        LP_WAVE_CASE(EPI_STORE_DOT)
        LP_WAVE_CASE(EPI_CHEB)
        LP_WAVE_CASE(EPI_RESID_SCALED)
        LP_WAVE_CASE(EPI_RESID_MASK)
        LP_WAVE_CASE(EPI_JACOBI)
        LP_WAVE_CASE(EPI_POWER)
        LP_WAVE_CASE(EPI_RESID)
        LP_WAVE_CASE(EPI_MASKED)
        LP_WAVE_CASE(EPI_CHEB_FIRST)
        LP_WAVE_CASE(EPI_CHEB_ONLY)
        LP_WAVE_CASE(EPI_CHEB_LAST)
        LP_WAVE_CASE(EPI_CHEB_LAST_R)
        default:
            set_error("level operator: unknown epilogue");
            throw ArgFailure();
    }
#undef LP_WAVE_CASE
    LP_HIP(hipGetLastError());
}

}  // namespace

// the wave-tile kernel on the fine level (tune_wave 1) or on both levels (2)
static bool use_wave(const Grid& Gd, int level) {
    return Gd.tune_wave == 2 || (Gd.tune_wave == 1 && Gd.lv[level].kind == LK_FINE);
}

int mfma_blocks(const Grid& Gd, int level) {
    const bool wave = use_wave(Gd, level) || (Gd.tune_bf16 && Gd.lv[level].kind == LK_FINE);
    return 8 * (wave ? wave_geom(Gd, level).per : mfma_geom(Gd, level).per);
}

bool mfma_level(const Grid& G, int level, int eb) {
    return G.fp16() && eb == 4 &&
           (G.lv[level].kind == LK_FINE || (G.lv[level].kind == LK_WEIGHTS && !G.lv[level].k32_apply) ||
            G.lv[level].k16_apply);
}

bool weights_mfma_path(const Grid& G, int level, int eb) {
    return G.fp16() && eb == 4 && level >= 1 && G.lv[level].kind == LK_WEIGHTS && !G.lv[level].k32_apply &&
           !G.lv[level].k16_apply;
}

void launch_apply_fine_mfma(const Grid& Gd, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                            hipStream_t s) {
    MfmaSrc S{};
    S.e = Gd.lv[level].e32.as<float>();
    S.a = Gd.ke32.as<float>();
    if (Gd.tune_bf16 && Gd.lv[level].kind == LK_FINE)
        launch_wave<false, true>(Gd, level, u, vstride, nrhs, E, S, s);
    else if (use_wave(Gd, level))
        launch_wave<false>(Gd, level, u, vstride, nrhs, E, S, s);
    else
        launch_mfma<false>(Gd, level, u, vstride, nrhs, E, S, s);
}

void launch_apply_weights_mfma(const Grid& Gd, int level, const void* u, size_t vstride, int nrhs,
                               const EpiArgs& E, hipStream_t s) {
    const Level& lv = Gd.lv[level];
    MfmaSrc S{};
    S.e = Gd.lv[level - 1].e32.as<float>();
    S.a = Gd.mq32.as<float>();
    S.pidx = lv.patch_idx.as<int>();
    S.npatch = lv.npatch;
    if (lv.npatch > 0) {
        const size_t need = (size_t)nrhs * lv.npatch * 24 * sizeof(float);
        if (lv.patch_prod.bytes < need) {
            set_error("weights level: patch product buffer smaller than the batch");
            throw ArgFailure();
        }
        S.pprod = lv.patch_prod.as<float>();
        const unsigned nb = (unsigned)((lv.npatch * 24 + kBlock - 1) / kBlock);
        const bool staged = E.kind == EPI_CHEB_FIRST || E.kind == EPI_CHEB_ONLY || E.kind == EPI_CHEB_LAST_R;
        hipLaunchKernelGGL(k_patch_prod, dim3(nb), dim3(kBlock), 0, s, static_cast<const float*>(u), vstride, nrhs,
                           lv.mask.as<uint8_t>(), lv.g, lv.patch_cell.as<int>(), lv.patch32.as<float>(), lv.npatch,
                           staged ? static_cast<const float*>(E.dinv) : nullptr, (float)E.c3,
                           lv.patch_prod.as<float>());
        LP_HIP(hipGetLastError());
    }
    if (use_wave(Gd, level))
        launch_wave<true>(Gd, level, u, vstride, nrhs, E, S, s);
    else
        launch_mfma<true>(Gd, level, u, vstride, nrhs, E, S, s);
}

}  // namespace lp
