// Level operators of the "explicit" cell-matrix kind: every coarse cell holds
// its own 24x24 Galerkin matrix K_c (mg._coarsen for levels >= 2,
// mg.py:643-677).
//
// One thread per (node, one of its 8 cells) loads the 3 rows of K_c for the
// node's corner once and applies them to every right-hand side of the
// batch, so the matrices (the level's dominant traffic: 2.3 KB per cell in
// FP32, 4.6 KB in FP64) are read once per apply instead of once per
// right-hand side. FP32 for the V-cycles of the fp32 and mixed paths; FP64
// for the all-FP64 path, whose first coarse level also runs here with its
// Sigma_q w_q M_q (patch cells: their patch matrix) formed once per
// hierarchy build.
#include "apply_impl.h"

namespace lp {

namespace {

// block = kXN nodes x their 8 cells: thread (c8, node) with c8 = t / kXN, so
// the kXN lanes of one cell slot read consecutive cells' entries (128-byte
// runs). The 8 cells' row sums meet in LDS and are added in the order of a
// butterfly over c8 (((0+1)+(2+3))+((4+5)+(6+7))); then thread (rhs, node)
// runs the epilogue.
constexpr int kXN = kBlock / 8;     // nodes per block
constexpr int kXR = kBlock / kXN;   // right-hand sides per pass (= 8)

template <typename T>
__device__ __forceinline__ void explicit_batched_body(const T* __restrict__ u0, size_t vstride,
                                                      const uint8_t* __restrict__ mask, AGeom G,
                                                      const T* __restrict__ k, int nrhs, const EpiArgs& E) {
    using namespace apply_detail;
    __shared__ T red[kXR][8][kXN][3];
    const int t = (int)threadIdx.x;
    const int c8 = t / kXN, ln = t - c8 * kXN;
    const unsigned nn = (unsigned)G.nn;
    const unsigned nid = blockIdx.x * kXN + ln;
    const bool valid = nid < nn;
    const unsigned nyz = (unsigned)G.Ny * (unsigned)G.Nz;
    const int x = valid ? (int)(nid / nyz) : 0;
    const unsigned rem = valid ? nid - (unsigned)x * nyz : 0;
    const int y = (int)(rem / (unsigned)G.Nz);
    const int z = (int)(rem - (unsigned)y * (unsigned)G.Nz);
    const int cx = c8 & 1, cy = (c8 >> 1) & 1, cz = (c8 >> 2) & 1;
    const int X = x - 1 + cx, Y = y - 1 + cy, Z = z - 1 + cz;
    const bool has = valid && X >= 0 && Y >= 0 && Z >= 0 && X < G.nx && Y < G.ny && Z < G.nz;
    const int a = (1 - cx) | ((1 - cy) << 1) | ((1 - cz) << 2);
    const size_t nc = G.ncell;
    const size_t ci = has ? ((size_t)X * G.ny + Y) * G.nz + Z : 0;
    // the node's 3 rows of its cell's matrix
    T kr[72];
#pragma unroll
    for (int e = 0; e < 72; ++e) kr[e] = has ? k[(size_t)(3 * a * 24 + e) * nc + ci] : T(0);
    unsigned off[8];
    int mk[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        off[b] = has ? ((unsigned)(X + (b & 1)) * (unsigned)G.Ny + (unsigned)(Y + ((b >> 1) & 1))) * (unsigned)G.Nz +
                           (unsigned)(Z + ((b >> 2) & 1))
                     : 0u;
        mk[b] = has ? (int)mask[off[b]] : 0;
    }
    const bool need_raw_all = E.kind == EPI_STORE_DOT || E.kind == EPI_JACOBI || E.kind == EPI_POWER;
    for (int r0 = 0; r0 < nrhs; r0 += kXR) {
        const int nr = min(kXR, nrhs - r0);
        for (int q = 0; q < nr; ++q) {
            const int rhs = r0 + q;
            const bool skip = E.scal != nullptr && !E.scal[rhs].active;
            T t0 = 0, t1 = 0, t2 = 0;
            if (!skip && has) {
                const T* __restrict__ u = u0 + (size_t)rhs * vstride;
#pragma unroll
                for (int b = 0; b < 8; ++b)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const T v = ((mk[b] >> c) & 1) ? u[(size_t)c * nn + off[b]] : T(0);
                        t0 = fma(kr[3 * b + c], v, t0);
                        t1 = fma(kr[24 + 3 * b + c], v, t1);
                        t2 = fma(kr[48 + 3 * b + c], v, t2);
                    }
            }
            red[q][c8][ln][0] = t0;
            red[q][c8][ln][1] = t1;
            red[q][c8][ln][2] = t2;
        }
        __syncthreads();
        // thread (q, node): the node's sum over its 8 cells, then the epilogue
        const int q = t / kXN;
        const int rhs = r0 + q;
        double part = 0.0;
        const bool live = q < nr && valid && !(E.scal != nullptr && !E.scal[rhs].active);
        if (live) {
            T yv[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T s01 = red[q][0][ln][c] + red[q][1][ln][c], s23 = red[q][2][ln][c] + red[q][3][ln][c];
                const T s45 = red[q][4][ln][c] + red[q][5][ln][c], s67 = red[q][6][ln][c] + red[q][7][ln][c];
                yv[c] = (s01 + s23) + (s45 + s67);
            }
            const T* __restrict__ u = u0 + (size_t)rhs * vstride;
            const int m = mask[nid];
            T ur[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool mc = (m >> c) & 1;
                ur[c] = (!mc || need_raw_all) ? u[(size_t)c * nn + nid] : T(0);
                if (!mc) yv[c] = ur[c];
            }
            epilogue<T>(E, m, (size_t)rhs * vstride + nid, nn, nid, yv, ur, part);
        }
        if (E.partials != nullptr) {
            // the kXN threads of one rhs are one half of a wave
#pragma unroll
            for (int o = kXN / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
            if (ln == 0 && q < nr) E.partials[(size_t)rhs * gridDim.x + blockIdx.x] = part;
        }
        __syncthreads();
    }
}


__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 4)))
k_apply_explicit_batched(const float* __restrict__ u0, size_t vstride, const uint8_t* __restrict__ mask, AGeom G,
                         const float* __restrict__ k, int nrhs, EpiArgs E) {
    explicit_batched_body<float>(u0, vstride, mask, G, k, nrhs, E);
}

// FP64: the node's 72 matrix entries take 144 VGPRs; two waves per SIMD
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_apply_explicit_batched64(const double* __restrict__ u0, size_t vstride, const uint8_t* __restrict__ mask, AGeom G,
                           const double* __restrict__ k, int nrhs, EpiArgs E) {
    explicit_batched_body<double>(u0, vstride, mask, G, k, nrhs, E);
}

}  // namespace

int explicit_blocks(const Grid& G, int level) {
    return (int)((G.lv[level].g.nn + kXN - 1) / kXN);
}

bool explicit_weights(const Grid& G, int level, int eb) {
    const Level& lv = G.lv[level];
    return lv.kind == LK_WEIGHTS && ((eb == 8 && lv.k64_apply) || (eb == 4 && lv.k32_apply));
}

void launch_apply_explicit(const Grid& G, int level, int eb, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                           hipStream_t s) {
    const Level& lv = G.lv[level];
    const AGeom A = make_apply_geom(G, level);
    if (eb == 8)
        hipLaunchKernelGGL(k_apply_explicit_batched64, dim3(explicit_blocks(G, level)), dim3(kBlock), 0, s,
                           static_cast<const double*>(u), vstride, lv.mask.as<uint8_t>(), A, lv.k64.as<double>(), nrhs,
                           E);
    else
        hipLaunchKernelGGL(k_apply_explicit_batched, dim3(explicit_blocks(G, level)), dim3(kBlock), 0, s,
                           static_cast<const float*>(u), vstride, lv.mask.as<uint8_t>(), A, lv.k32.as<float>(), nrhs,
                           E);
    LP_HIP(hipGetLastError());
}

}  // namespace lp
