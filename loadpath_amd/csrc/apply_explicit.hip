// Level operators of the "explicit" cell-matrix kind: every coarse cell holds
// its own 24x24 Galerkin matrix K_c (mg._coarsen for levels >= 2,
// mg.py:643-677).
//
// FP32 (the V-cycle precision of the fp32 and mixed paths): one thread per
// (node, one of its 8 cells) loads the 3 rows of K_c for the node's corner
// once and applies them to every right-hand side of the batch, so the
// matrices (the level's dominant traffic: 2.3 KB per cell) are read once per
// apply instead of once per right-hand side. FP64 keeps the per-rhs kernel
// of apply_impl.h.
#include "apply_impl.h"

namespace lp {

namespace {

// block sum of part into partials[rhs * gridDim.x + blockIdx.x]; the trailing
// barrier lets the caller loop over right-hand sides with one smem slot
__device__ __forceinline__ void block_partial_loop(double part, const EpiArgs& E, int rhs) {
    __shared__ double sred[kBlock / 64];
    part = apply_detail::wave_sum(part);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += sred[i];
        E.partials[(size_t)rhs * gridDim.x + blockIdx.x] = s;
    }
    __syncthreads();
}

__global__ void __launch_bounds__(kBlock) k_apply_explicit_batched(const float* __restrict__ u0, size_t vstride,
                                                                   const uint8_t* __restrict__ mask, AGeom G,
                                                                   const float* __restrict__ k, int nrhs, EpiArgs E) {
    using namespace apply_detail;
    const int c8 = threadIdx.x & 7;
    const size_t nid = (size_t)blockIdx.x * (blockDim.x >> 3) + (threadIdx.x >> 3);
    const size_t nn = G.nn;
    const bool valid = nid < nn;
    const bool need_raw_all = E.kind == EPI_STORE_DOT || E.kind == EPI_JACOBI || E.kind == EPI_POWER;
    int x = 0, y = 0, z = 0;
    bool has = false;
    int X = 0, Y = 0, Z = 0, a = 0;
    if (valid) {
        const unsigned id = (unsigned)nid, nyz = (unsigned)G.Ny * (unsigned)G.Nz;
        x = (int)(id / nyz);
        const unsigned rem = id - (unsigned)x * nyz;
        y = (int)(rem / (unsigned)G.Nz);
        z = (int)(rem - (unsigned)y * (unsigned)G.Nz);
        const int cx = c8 & 1, cy = (c8 >> 1) & 1, cz = (c8 >> 2) & 1;
        X = x - 1 + cx;
        Y = y - 1 + cy;
        Z = z - 1 + cz;
        has = X >= 0 && Y >= 0 && Z >= 0 && X < G.nx && Y < G.ny && Z < G.nz;
        a = (1 - cx) | ((1 - cy) << 1) | ((1 - cz) << 2);
    }
    // the node's 3 rows of its cell's matrix, (row, column) for column j of
    // the cell's corner vector
    float kr[72];
    const size_t ci = has ? ((size_t)X * G.ny + Y) * G.nz + Z : 0;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int j = 0; j < 24; ++j) kr[r * 24 + j] = has ? k[(size_t)((3 * a + r) * 24 + j) * G.ncell + ci] : 0.0f;
    for (int rhs = 0; rhs < nrhs; ++rhs) {
        double part = 0.0;
        const bool skip = E.scal != nullptr && !E.scal[rhs].active;
        float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
        const float* __restrict__ u = u0 + (size_t)rhs * vstride;
        if (!skip && has) {
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const size_t n = ((size_t)(X + (b & 1)) * G.Ny + (Y + ((b >> 1) & 1))) * G.Nz + (Z + ((b >> 2) & 1));
                const int m = mask[n];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const float v = ((m >> q) & 1) ? u[q * nn + n] : 0.0f;
                    t0 = fmaf(kr[3 * b + q], v, t0);
                    t1 = fmaf(kr[24 + 3 * b + q], v, t1);
                    t2 = fmaf(kr[48 + 3 * b + q], v, t2);
                }
            }
        }
#pragma unroll
        for (int o = 1; o < 8; o <<= 1) {
            t0 += __shfl_xor(t0, o, 64);
            t1 += __shfl_xor(t1, o, 64);
            t2 += __shfl_xor(t2, o, 64);
        }
        if (!skip && valid && c8 == 0) {
            const size_t n = nid;
            const int m = mask[n];
            float yv[3] = {t0, t1, t2};
            float ur[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool mc = (m >> c) & 1;
                ur[c] = (!mc || need_raw_all) ? u[c * nn + n] : 0.0f;
                if (!mc) yv[c] = ur[c];
            }
            epilogue<float>(E, m, (size_t)rhs * vstride + n, nn, n, yv, ur, part);
        }
        if (E.partials != nullptr) block_partial_loop(part, E, rhs);
    }
}

}  // namespace

int explicit_blocks(const Grid& G, int level) {
    return (int)((G.lv[level].g.nn + (kBlock / 8) - 1) / (kBlock / 8));
}

void launch_apply_explicit(const Grid& G, int level, int eb, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                           hipStream_t s) {
    if (eb == 8) {
        apply_detail::launch_kind<double, LK_EXPLICIT>(G, level, static_cast<const double*>(u), vstride, nrhs, E, s);
        return;
    }
    const Level& lv = G.lv[level];
    const AGeom A = make_apply_geom(G, level);
    hipLaunchKernelGGL(k_apply_explicit_batched, dim3(explicit_blocks(G, level)), dim3(kBlock), 0, s,
                       static_cast<const float*>(u), vstride, lv.mask.as<uint8_t>(), A, lv.k32.as<float>(), nrhs, E);
    LP_HIP(hipGetLastError());
}

}  // namespace lp
