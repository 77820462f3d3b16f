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

// ---------------------------------------------------------------------------
// FP16 cell matrices (the mixed V-cycle's coarse levels)
//
// Each cell's 576 entries are scaled by 2^k, k chosen so its largest |entry|
// lands in [2^14, 2^15) (entries down to 2^-29 of the largest stay normal
// FP16), rounded to FP16 and stored as 72 halves per corner a (rows 3a..3a+2,
// 24 columns each, row-major) in 9 groups of 8: group (a, g) of cell ci is the
// 16-byte slot (a * 9 + g) * ncell + ci, so the threads of consecutive nodes
// read consecutive 16-byte slots. sc16[ci] = 2^-k.
//
// The apply is the batched explicit kernel's (thread = node x one of its 8
// cells, the 8 cells' sums met in LDS) with the matrix entries in FP16 and
// the arithmetic in FP32, half the matrix traffic of the FP32 form. A
// rounded matrix no longer annihilates translations, so each row subtracts
// its own rounded entries' sum per component times the cell's corner-0 value:
// the product is that of the rounded matrix with the corner vector relative
// to corner 0 (as the FP16-MFMA operators form it), and a displacement's
// common offset, which can be far larger than the cell's deformation, drops
// out at FP32 rounding. Cells whose exact matrix does not annihilate
// translations (those touching held dofs: patch matrices, masked Galerkin
// levels) are flagged by a negative scale and applied without it.
constexpr int kHalfGroups = 9;      // 16-byte groups per corner (72 halves)

typedef unsigned uint4v __attribute__((ext_vector_type(4)));

// acc + h * v with h the low (HI = false) or high half of w, as FP16 read by
// the mixed-precision FMA itself: the matrix stays packed in registers (36
// instead of 72 VGPRs; a plain conversion is hoisted out of the right-hand
// side loop by the compiler, unpacking every entry)
template <bool HI>
__device__ __forceinline__ float fma_half(unsigned w, float v, float acc) {
    float r;
    if (HI)
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(w), "v"(v), "v"(acc));
    else
        asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(w), "v"(v), "v"(acc));
    return r;
}

// one 576-thread block per cell, entry tid = row * 24 + col; the weights
// level's entries sum_q w_q M_q (patch cells: the patch matrix) in FP64 as
// the FP64 and FP32 explicit forms make them, other levels from their FP32
// Galerkin matrices
template <bool WEIGHTS>
__global__ void __launch_bounds__(576) k_form_half(const double* __restrict__ fe, const double* __restrict__ mq,
                                                   const int* __restrict__ pidx, const double* __restrict__ patch,
                                                   Geom f, Geom c, const float* __restrict__ k32,
                                                   unsigned short* __restrict__ k16, float* __restrict__ sc16) {
    __shared__ double wmax[9];
    const size_t cell = blockIdx.x;
    const int tid = (int)threadIdx.x;
    double v = 0.0;
    if (WEIGHTS) {
        const int Z = (int)(cell % c.nz), Y = (int)((cell / c.nz) % c.ny), X = (int)(cell / ((size_t)c.nz * c.ny));
        const int pi = pidx[cell];
        if (pi >= 0) {
            v = patch[(size_t)pi * 576 + tid];
        } else {
            for (int q = 0; q < 8; ++q) {
                const int fx = 2 * X + (q & 1), fy = 2 * Y + ((q >> 1) & 1), fz = 2 * Z + ((q >> 2) & 1);
                if (fx >= f.nx || fy >= f.ny || fz >= f.nz) continue;
                v += mq[q * 576 + tid] * fe[((size_t)fx * f.ny + fy) * f.nz + fz];
            }
        }
    } else {
        v = (double)k32[(size_t)tid * c.ncell + cell];
    }
    // does the cell's matrix annihilate translations? Every row's entries
    // summed per component vanish (to the entries' own rounding) unless the
    // cell touches held dofs (patch matrices, strict Galerkin levels); only
    // then may the apply subtract the corner-0 translation
    __shared__ double ent[576];
    ent[tid] = v;
    __syncthreads();
    bool moves = false;
    if (tid < 72) {
        const int row = tid / 3, cc = tid - 3 * (tid / 3);
        double sum = 0.0, mag = 0.0;
        for (int b = 0; b < 8; ++b) {
            const double x = ent[row * 24 + 3 * b + cc];
            sum += x;
            mag += fabs(x);
        }
        moves = fabs(sum) > 1e-4 * mag;
    }
    const bool keep_offset = __syncthreads_or(moves);
    double a = fabs(v);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a = fmax(a, __shfl_xor(a, o, 64));
    if ((tid & 63) == 0) wmax[tid >> 6] = a;
    __syncthreads();
    double amax = 0.0;
    for (int w = 0; w < 9; ++w) amax = fmax(amax, wmax[w]);
    int ex = 0;
    if (amax > 0.0) (void)frexp(amax, &ex);            // amax = m 2^ex, m in [0.5, 1)
    const int k = min(max(15 - ex, -100), 100);
    const int row = tid / 24, col = tid - 24 * (tid / 24);
    const int ca = row / 3, e = 24 * (row - 3 * ca) + col;
    const _Float16 h = (_Float16)(float)ldexp(v, k);
    k16[((size_t)(ca * kHalfGroups + (e >> 3)) * c.ncell + cell) * 8 + (e & 7)] =
        __builtin_bit_cast(unsigned short, h);
    // the sign carries the flag: negative = apply the matrix as it is
    if (tid == 0) sc16[cell] = keep_offset ? -ldexpf(1.0f, -k) : ldexpf(1.0f, -k);
}

template <int EK>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 4)))
k_apply_half(const float* __restrict__ u0, size_t vstride, const uint8_t* __restrict__ mask, AGeom G,
             const uint4v* __restrict__ k16, const float* __restrict__ sc16, int nrhs, EpiArgs E) {
    using namespace apply_detail;
    constexpr bool need_raw_all = EK == EPI_STORE_DOT || EK == EPI_JACOBI || EK == EPI_POWER;
    // Chebyshev steps that stage their operand d = (dinv u) s from the vector
    // u they are given (b for the first step from zero, r for the last step
    // after a residual-only step)
    constexpr bool STAGE_B = EK == EPI_CHEB_FIRST || EK == EPI_CHEB_ONLY || EK == EPI_CHEB_LAST_R;
    (void)need_raw_all;
    __shared__ float red[kXR][8][kXN][3];
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
    const int a = 7 - c8;   // the node's corner in the cell
    const size_t nc = G.ncell;
    const size_t ci = has ? ((size_t)X * G.ny + Y) * G.nz + Z : 0;
    // the node's 3 rows of its cell's matrix: 72 halves in 36 words
    unsigned kw[4 * kHalfGroups];
#pragma unroll
    for (int g = 0; g < kHalfGroups; ++g) {
        const uint4v w = has ? k16[(size_t)(a * kHalfGroups + g) * nc + ci] : uint4v{0u, 0u, 0u, 0u};
#pragma unroll
        for (int j = 0; j < 4; ++j) kw[4 * g + j] = w[j];
    }
    const float sflag = has ? sc16[ci] : 0.0f;
    const float scl = fabsf(sflag);
    // cells whose matrix does not annihilate translations (negative scale):
    // no corner-0 correction
    const float rs_on = sflag < 0.0f ? 0.0f : 1.0f;
    // each row's rounded entries summed per component
    float rs[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            float sacc = 0.0f;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int e = 24 * i + 3 * b + cc;
                sacc = (e & 1) ? fma_half<true>(kw[e >> 1], 1.0f, sacc) : fma_half<false>(kw[e >> 1], 1.0f, sacc);
            }
            rs[i][cc] = sacc * rs_on;
        }
    unsigned off[8];
    int mk[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        off[b] = has ? ((unsigned)(X + (b & 1)) * (unsigned)G.Ny + (unsigned)(Y + ((b >> 1) & 1))) * (unsigned)G.Nz +
                           (unsigned)(Z + ((b >> 2) & 1))
                     : 0u;
        mk[b] = has ? (int)mask[off[b]] : 0;
    }
    const float s_b = (float)E.c3;
    float wdv[STAGE_B ? 24 : 1];
    if (STAGE_B) {
#pragma unroll
        for (int b = 0; b < 8; ++b)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc)
                wdv[STAGE_B ? 3 * b + cc : 0] =
                    ((mk[b] >> cc) & 1) ? static_cast<const float*>(E.dinv)[(size_t)cc * nn + off[b]] : 0.0f;
    }
    for (int r0 = 0; r0 < nrhs; r0 += kXR) {
        const int nr = min(kXR, nrhs - r0);
        // the 24 operand values of right-hand side r0 + q, gathered one
        // right-hand side ahead of the products (zeros for an inactive one, a
        // cell outside the grid or a masked-out dof: offsets past the buffer)
        // corners b and b + 4 (bz = 0, 1) are z-neighbours: one 8-byte load
        // each, the mask applied after it
        auto gather = [&](int q, float (&v)[24]) {
            const int rhs = min(r0 + q, nrhs - 1);
            const bool go = has && q < nr && !(E.scal != nullptr && !E.scal[rhs].active);
            const __amdgpu_buffer_rsrc_t ru = make_rsrc(u0 + (size_t)rhs * vstride, 3u * nn * 4u);
            typedef unsigned u2 __attribute__((ext_vector_type(2)));
#pragma unroll
            for (int b = 0; b < 4; ++b)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    const u2 w = __builtin_bit_cast(u2, __builtin_amdgcn_raw_buffer_load_b64(ru, go ? (cc * nn + off[b]) * 4u : kOOB, 0, 0));
                    v[3 * b + cc] = ((mk[b] >> cc) & 1) ? __uint_as_float(w.x) : 0.0f;
                    v[3 * (b + 4) + cc] = ((mk[b + 4] >> cc) & 1) ? __uint_as_float(w.y) : 0.0f;
                }
        };
        auto product = [&](int q, const float (&v)[24]) {
            if (q >= nr) return;
            float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, v0[3];
#pragma unroll
            for (int b = 0; b < 8; ++b)
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    float w = v[3 * b + cc];
                    if (STAGE_B) w = (wdv[STAGE_B ? 3 * b + cc : 0] * w) * s_b;
                    if (b == 0) v0[cc] = w;
                    const int e = 3 * b + cc;
                    if (e & 1) {
                        t0 = fma_half<true>(kw[e >> 1], w, t0);
                        t1 = fma_half<true>(kw[(24 + e) >> 1], w, t1);
                        t2 = fma_half<true>(kw[(48 + e) >> 1], w, t2);
                    } else {
                        t0 = fma_half<false>(kw[e >> 1], w, t0);
                        t1 = fma_half<false>(kw[(24 + e) >> 1], w, t1);
                        t2 = fma_half<false>(kw[(48 + e) >> 1], w, t2);
                    }
                }
            t0 -= rs[0][0] * v0[0] + rs[0][1] * v0[1] + rs[0][2] * v0[2];
            t1 -= rs[1][0] * v0[0] + rs[1][1] * v0[1] + rs[1][2] * v0[2];
            t2 -= rs[2][0] * v0[0] + rs[2][1] * v0[1] + rs[2][2] * v0[2];
            red[q][c8][ln][0] = t0 * scl;
            red[q][c8][ln][1] = t1 * scl;
            red[q][c8][ln][2] = t2 * scl;
        };
        float va[24], vb[24];
        gather(0, va);
        for (int q = 0; q < nr; q += 2) {
            gather(q + 1, vb);
            product(q, va);
            gather(q + 2, va);
            product(q + 1, vb);
        }
        __syncthreads();
        // thread (q, node): the node's sum over its 8 cells, then the epilogue
        const int q = t / kXN;
        const int rhs = r0 + q;
        double part = 0.0;
        const bool live = q < nr && valid && !(E.scal != nullptr && !E.scal[rhs].active);
        if (live) {
            float yv[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const float s01 = red[q][0][ln][cc] + red[q][1][ln][cc], s23 = red[q][2][ln][cc] + red[q][3][ln][cc];
                const float s45 = red[q][4][ln][cc] + red[q][5][ln][cc], s67 = red[q][6][ln][cc] + red[q][7][ln][cc];
                yv[cc] = (s01 + s23) + (s45 + s67);
            }
            const float* __restrict__ u = u0 + (size_t)rhs * vstride;
            const int m = mask[nid];
            const size_t o = (size_t)rhs * vstride + nid;
            EpiIn<float> in;
            epi_load<float, EK>(E, o, nn, nid, in);
            float ur[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const float raw = u[(size_t)cc * nn + nid];
                if (STAGE_B) {
                    ur[cc] = (in.v[3][cc] * raw) * s_b;
                    in.v[EK == EPI_CHEB_LAST_R ? 2 : 0][cc] = raw;
                } else {
                    ur[cc] = raw;
                }
                if (!((m >> cc) & 1)) yv[cc] = ur[cc];
            }
            epi_finish<float, EK, float>(E, m, o, nn, yv, ur, in, part);
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

// The same product with the operand staged in LDS: a block owns a brick of
// ty x tz nodes in (y, z) and marches an x-chunk of node planes. The three
// node planes a step's two cell layers touch are staged for all 8
// right-hand sides ((ty + 2) x (tz + 2) nodes each, masked; a ring of 3
// plane slots), so a thread's 24 operand values per right-hand side are LDS
// reads instead of 12 gathers from memory, and the plane the next step adds
// is loaded into registers during this one. Staging: wave w loads
// right-hand sides 2w and 2w + 1, one (rhs, component) per load with the
// staged node = lane, so each load's rhs is wave-uniform. Thread (c8, node
// slot) forms its cell's 3 rows as k_apply_half does (the same FP16 FMAs in
// the same order: the same results); thread (rhs, node slot) then sums the 8
// cells and runs the epilogue. Blocks are remapped so each XCD takes a
// contiguous run of bricks (bz fastest): neighbouring bricks share halo rows
// and the 128-byte lines of their matrix rows in that XCD's L2.
constexpr int kHBS = 64;                 // staged nodes per (rhs, component) and plane at most
struct HBrick {
    int ty, tz;       // node brick in (y, z): ty * tz <= 32, (ty + 2)(tz + 2) <= 64
    int nby, nbz, nxc;
    int xchunk;       // node planes per block
    int total, per;   // bricks, and bricks per XCD (grid = 8 * per)
};
constexpr int kHBW = kXR * 3 / (kBlock / 64);   // (rhs, component) loads per lane and plane (6)
static_assert(kXN == 32 && kXR == 8 && kBlock == 256, "brick kernel: 32 node slots x 8 cells, 8 right-hand sides");

template <int EK>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3, 4)))
k_apply_half_brick(const float* __restrict__ u0, size_t vstride, const uint8_t* __restrict__ mask, AGeom G,
                   HBrick B, const uint4v* __restrict__ k16, const float* __restrict__ sc16, int nrhs, int rbase,
                   EpiArgs E) {
    using namespace apply_detail;
    constexpr bool STAGE_B = EK == EPI_CHEB_FIRST || EK == EPI_CHEB_ONLY || EK == EPI_CHEB_LAST_R;
    __shared__ float ring[3][kXR * 3 * kHBS];   // [(rhs, component), staged node] per plane slot
    __shared__ float red[kXR][8][kXN][3];
    const int b = (int)blockIdx.x;
    const int L = (b & 7) * B.per + (b >> 3);
    const int t = (int)threadIdx.x;
    if (L >= B.total) {                         // the padding of the last XCD's run
        if (E.partials != nullptr && t % kXN == 0 && t / kXN < nrhs - rbase)
            E.partials[(size_t)(rbase + t / kXN) * gridDim.x + blockIdx.x] = 0.0;
        return;
    }
    const int c8 = t / kXN, ns = t - c8 * kXN;
    const int nbyz = B.nby * B.nbz;
    const int xc = L / nbyz, rem = L - xc * nbyz;
    const int by = rem / B.nbz, bz = rem - by * B.nbz;
    const int Y0 = by * B.ty, Z0 = bz * B.tz;
    const int Ny = G.Ny, Nz = G.Nz;
    const int x0 = xc * B.xchunk, x1 = min(x0 + B.xchunk, G.Nx);
    const int SW = B.tz + 2, NS = (B.ty + 2) * SW;
    const int nr = min(kXR, nrhs - rbase);
    const unsigned nn = (unsigned)G.nn;
    const unsigned nyz = (unsigned)Ny * (unsigned)Nz;
    const size_t nc = G.ncell;
    const float s_b = (float)E.c3;
    const __amdgpu_buffer_rsrc_t rm = make_rsrc(mask, nn);

    // staging: this lane's staged node (the same for every load)
    // the wave index as a uniform value: the staging loads' buffer
    // descriptors (one rhs per load) then stay in SGPRs
    const int lane = t & 63, wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const int sy = lane / SW, sz = lane - (lane / SW) * SW;
    const int yy = Y0 - 1 + sy, zz = Z0 - 1 + sz;
    const bool yz_ok = lane < NS && yy >= 0 && yy < Ny && zz >= 0 && zz < Nz;
    const unsigned yzo = yz_ok ? (unsigned)yy * (unsigned)Nz + (unsigned)zz : 0u;
    bool r_ok[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int q = 2 * wv + h;
        r_ok[h] = q < nr && !(E.scal != nullptr && !E.scal[rbase + q].active);
    }
    float pv[kHBW];
    int pm = 0;
    float pd[STAGE_B ? 3 : 1];
    auto fetch = [&](int xs) {
        const bool ok = yz_ok && xs >= 0 && xs < G.Nx;
        const unsigned n = (unsigned)xs * nyz + yzo;
        pm = bload_u8(rm, ok ? n : kOOB);
#pragma unroll
        for (int j = 0; j < kHBW; ++j) {
            const int h = j / 3, cc = j - 3 * h;
            const int rhs = rbase + min(2 * wv + h, nr - 1);
            const __amdgpu_buffer_rsrc_t ru = make_rsrc(u0 + (size_t)rhs * vstride, 3u * nn * 4u);
            pv[j] = bload<float>(ru, ok && r_ok[h] ? ((unsigned)cc * nn + n) * 4u : kOOB);
        }
        if (STAGE_B) {
            const __amdgpu_buffer_rsrc_t rd = make_rsrc(E.dinv, 3u * nn * 4u);
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) pd[STAGE_B ? cc : 0] = bload<float>(rd, ok ? ((unsigned)cc * nn + n) * 4u : kOOB);
        }
    };
    auto store = [&](int slot) {
        if (lane < NS) {
#pragma unroll
            for (int j = 0; j < kHBW; ++j) {
                const int h = j / 3, cc = j - 3 * h;
                float v = pv[j];
                if (STAGE_B) v = (pd[STAGE_B ? cc : 0] * v) * s_b;
                ring[slot][((2 * wv + h) * 3 + cc) * kHBS + lane] = ((pm >> cc) & 1) ? v : 0.0f;
            }
        }
    };

    // this thread's node and cell corner
    const int ly = ns / B.tz, lz = ns - (ns / B.tz) * B.tz;
    const int y = Y0 + ly, z = Z0 + lz;
    const bool node_ok = ns < B.ty * B.tz && y < Ny && z < Nz;
    const int cx = c8 & 1, cy = (c8 >> 1) & 1, cz = (c8 >> 2) & 1;
    const int a = 7 - c8;
    const int CY = y - 1 + cy, CZ = z - 1 + cz;
    const bool yz_cell = node_ok && CY >= 0 && CZ >= 0 && CY < G.ny && CZ < G.nz;
    // corner (by, bz) of the cell: staged node (ly + cy + by, lz + cz + bz)
    const int sbase = (ly + cy) * SW + (lz + cz);
    // epilogue thread (q, node slot)
    const int eq = t / kXN, es = ns;
    const int ey = Y0 + es / B.tz, ez = Z0 + es % B.tz;
    const bool e_ok = eq < nr && es < B.ty * B.tz && ey < Ny && ez < Nz &&
                      !(E.scal != nullptr && !E.scal[rbase + min(eq, nr - 1)].active);
    double part = 0.0;
    // the epilogue's inputs at node plane Xs, loaded a step ahead of their
    // use (after the step before's epilogue) so their latency hides behind a
    // whole step
    int em = 0;
    float eraw[3] = {0.0f, 0.0f, 0.0f};
    EpiIn<float> ein;
    auto epi_pre = [&](int Xs) {
        if (!e_ok || Xs >= x1) return;
        const unsigned nid = (unsigned)Xs * nyz + (unsigned)ey * (unsigned)Nz + (unsigned)ez;
        const size_t o = (size_t)(rbase + eq) * vstride + nid;
        em = mask[nid];
        epi_load<float, EK>(E, o, nn, nid, ein);
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) eraw[cc] = u0[o + (size_t)cc * nn];
    };

    // ring slots of planes X - 1, X, X + 1
    int sl0 = 0, sl1 = 1, sl2 = 2;
    fetch(x0 - 1);
    store(sl0);
    fetch(x0);
    store(sl1);
    fetch(x0 + 1);
    // the cell's 3 rows for step Xs (cell layer Xs - 1 + cx)
    auto load_k = [&](int Xs, unsigned (&kwd)[4 * kHalfGroups], float& sfl) {
        const int CX = Xs - 1 + cx;
        const bool has = yz_cell && CX >= 0 && CX < G.nx;
        const size_t ci = has ? ((size_t)CX * G.ny + CY) * G.nz + CZ : 0;
#pragma unroll
        for (int g = 0; g < kHalfGroups; ++g) {
            const uint4v w = has ? k16[(size_t)(a * kHalfGroups + g) * nc + ci] : uint4v{0u, 0u, 0u, 0u};
#pragma unroll
            for (int j = 0; j < 4; ++j) kwd[4 * g + j] = w[j];
        }
        sfl = has ? sc16[ci] : 0.0f;
    };
    unsigned kw[4 * kHalfGroups];
    float sflag;
    epi_pre(x0);
    for (int X = x0; X < x1; ++X) {
        store(sl2);
        __syncthreads();                          // planes X - 1 .. X + 1 staged; red free
        load_k(X, kw, sflag);
        if (X + 1 < x1) fetch(X + 2);             // the next step's new plane, in flight during this one
        const int CX = X - 1 + cx;
        const bool has = yz_cell && CX >= 0 && CX < G.nx;
        const float scl = fabsf(sflag);
        const float rs_on = sflag < 0.0f ? 0.0f : 1.0f;
        float rs[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                float sacc = 0.0f;
#pragma unroll
                for (int bb = 0; bb < 8; ++bb) {
                    const int e = 24 * i + 3 * bb + cc;
                    sacc = (e & 1) ? fma_half<true>(kw[e >> 1], 1.0f, sacc) : fma_half<false>(kw[e >> 1], 1.0f, sacc);
                }
                rs[i][cc] = sacc * rs_on;
            }
        // the plane slot of corners bx = 0, 1 (planes CX, CX + 1)
        const float* P0 = &ring[cx ? sl1 : sl0][sbase];
        const float* P1 = &ring[cx ? sl2 : sl1][sbase];
        for (int q = 0; q < nr; ++q) {
            float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, v0[3] = {0.0f, 0.0f, 0.0f};
            if (has && !(E.scal != nullptr && !E.scal[rbase + q].active)) {
#pragma unroll
                for (int bb = 0; bb < 8; ++bb)
#pragma unroll
                    for (int cc = 0; cc < 3; ++cc) {
                        const float* P = (bb & 1) ? P1 : P0;
                        const float w = P[(q * 3 + cc) * kHBS + ((bb >> 1) & 1) * SW + ((bb >> 2) & 1)];
                        if (bb == 0) v0[cc] = w;
                        const int e = 3 * bb + cc;
                        if (e & 1) {
                            t0 = fma_half<true>(kw[e >> 1], w, t0);
                            t1 = fma_half<true>(kw[(24 + e) >> 1], w, t1);
                            t2 = fma_half<true>(kw[(48 + e) >> 1], w, t2);
                        } else {
                            t0 = fma_half<false>(kw[e >> 1], w, t0);
                            t1 = fma_half<false>(kw[(24 + e) >> 1], w, t1);
                            t2 = fma_half<false>(kw[(48 + e) >> 1], w, t2);
                        }
                    }
                t0 -= rs[0][0] * v0[0] + rs[0][1] * v0[1] + rs[0][2] * v0[2];
                t1 -= rs[1][0] * v0[0] + rs[1][1] * v0[1] + rs[1][2] * v0[2];
                t2 -= rs[2][0] * v0[0] + rs[2][1] * v0[1] + rs[2][2] * v0[2];
            }
            red[q][c8][ns][0] = t0 * scl;
            red[q][c8][ns][1] = t1 * scl;
            red[q][c8][ns][2] = t2 * scl;
        }
        __syncthreads();                          // products in red; plane X - 1's slot is free
        if (e_ok) {
            const int q = eq, rhs = rbase + q;
            const unsigned nid = (unsigned)X * nyz + (unsigned)ey * (unsigned)Nz + (unsigned)ez;
            float yv[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const float s01 = red[q][0][es][cc] + red[q][1][es][cc], s23 = red[q][2][es][cc] + red[q][3][es][cc];
                const float s45 = red[q][4][es][cc] + red[q][5][es][cc], s67 = red[q][6][es][cc] + red[q][7][es][cc];
                yv[cc] = (s01 + s23) + (s45 + s67);
            }
            const int m = em;
            const size_t o = (size_t)rhs * vstride + nid;
            float ur[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) {
                const float raw = eraw[cc];
                if (STAGE_B) {
                    ur[cc] = (ein.v[3][cc] * raw) * s_b;
                    ein.v[EK == EPI_CHEB_LAST_R ? 2 : 0][cc] = raw;
                } else {
                    ur[cc] = raw;
                }
                if (!((m >> cc) & 1)) yv[cc] = ur[cc];
            }
            epi_finish<float, EK, float>(E, m, o, nn, yv, ur, ein, part);
        }
        epi_pre(X + 1);
        const int s = sl0;
        sl0 = sl1;
        sl1 = sl2;
        sl2 = s;
    }
    if (E.partials != nullptr) {
        // the kXN threads of one rhs are one half of a wave; the block's
        // partial covers its whole brick and x-chunk
#pragma unroll
        for (int o = kXN / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
        if (es == 0 && eq < nr) E.partials[(size_t)(rbase + eq) * gridDim.x + blockIdx.x] = part;
    }
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

bool half_explicit(const Grid& G, int level, int eb) {
    return G.fp16() && eb == 4 && level >= 1 && G.lv[level].k16_apply;
}

void form_half_matrices(const Grid& G, int level, const float* src32, hipStream_t s) {
    const Level& C = G.lv[level];
    if (C.g.ncell == 0) return;
    unsigned short* k16 = C.k16.as<unsigned short>();
    float* sc = C.sc16.as<float>();
    if (src32 == nullptr) {
        const Level& F = G.lv[level - 1];
        hipLaunchKernelGGL(k_form_half<true>, dim3((unsigned)C.g.ncell), dim3(576), 0, s, F.e64.as<double>(),
                           G.mq64.as<double>(), C.patch_idx.as<int>(), C.patch64.as<double>(), F.g, C.g,
                           (const float*)nullptr, k16, sc);
    } else {
        hipLaunchKernelGGL(k_form_half<false>, dim3((unsigned)C.g.ncell), dim3(576), 0, s, (const double*)nullptr,
                           (const double*)nullptr, (const int*)nullptr, (const double*)nullptr, C.g, C.g, src32, k16,
                           sc);
    }
    LP_HIP(hipGetLastError());
}

// the brick form's geometry when it is selected (tune_half_brick = ty * 10000
// + tz * 100 + x-chunk, 0 = off) and the level has tune_half_brick_min bricks
// or more (a brick marches its x-chunk plane by plane: on a small level too
// few blocks run for that to pay)
static bool half_brick(const Grid& G, int level, HBrick& B) {
    const int v = G.tune_half_brick;
    if (v <= 0) return false;
    const AGeom A = make_apply_geom(G, level);
    B.ty = v / 10000;
    B.tz = (v / 100) % 100;
    B.xchunk = v % 100;
    B.nby = (A.Ny + B.ty - 1) / B.ty;
    B.nbz = (A.Nz + B.tz - 1) / B.tz;
    B.nxc = (A.Nx + B.xchunk - 1) / B.xchunk;
    B.total = B.nxc * B.nby * B.nbz;
    B.per = (B.total + 7) / 8;
    return B.total >= G.tune_half_brick_min;
}

bool half_brick_shape_ok(int v) {
    if (v == 0) return true;
    const int ty = v / 10000, tz = (v / 100) % 100, xc = v % 100;
    return v > 0 && ty >= 1 && tz >= 1 && xc >= 1 && ty * tz <= kXN && (ty + 2) * (tz + 2) <= kHBS;
}

int half_blocks(const Grid& G, int level) {
    HBrick B;
    return half_brick(G, level, B) ? 8 * B.per : explicit_blocks(G, level);
}

void launch_apply_half(const Grid& G, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                       hipStream_t s) {
    const Level& lv = G.lv[level];
    const AGeom A = make_apply_geom(G, level);
    const dim3 grid(explicit_blocks(G, level)), block(kBlock);
    const float* uf = static_cast<const float*>(u);
    const uint8_t* m = lv.mask.as<uint8_t>();
    const uint4v* k = lv.k16.as<uint4v>();
    const float* sc = lv.sc16.as<float>();
    HBrick B;
    if (half_brick(G, level, B)) {
        const dim3 bgrid(8 * B.per);
        for (int r0 = 0; r0 < nrhs; r0 += kXR) {
#define LP_BRICK_CASE(K)                                                                                     \
    case K:                                                                                                  \
        hipLaunchKernelGGL((k_apply_half_brick<K>), bgrid, block, 0, s, uf, vstride, m, A, B, k, sc, nrhs, r0, \
                           E);                                                                               \
        break;
            switch (E.kind) {
                LP_BRICK_CASE(EPI_STORE)
                LP_BRICK_CASE(EPI_STORE_DOT)
                LP_BRICK_CASE(EPI_CHEB)
                LP_BRICK_CASE(EPI_RESID_SCALED)
                LP_BRICK_CASE(EPI_RESID_MASK)
                LP_BRICK_CASE(EPI_JACOBI)
                LP_BRICK_CASE(EPI_POWER)
                LP_BRICK_CASE(EPI_RESID)
                LP_BRICK_CASE(EPI_MASKED)
                LP_BRICK_CASE(EPI_CHEB_FIRST)
                LP_BRICK_CASE(EPI_CHEB_ONLY)
                LP_BRICK_CASE(EPI_CHEB_LAST)
                LP_BRICK_CASE(EPI_CHEB_LAST_R)
                default:
                    set_error("level operator: unknown epilogue");
                    throw ArgFailure();
            }
#undef LP_BRICK_CASE
            LP_HIP(hipGetLastError());
        }
        return;
    }
#define LP_HALF_CASE(K) \
    case K: hipLaunchKernelGGL((k_apply_half<K>), grid, block, 0, s, uf, vstride, m, A, k, sc, nrhs, E); break;
    switch (E.kind) {
        LP_HALF_CASE(EPI_STORE)
        LP_HALF_CASE(EPI_STORE_DOT)
        LP_HALF_CASE(EPI_CHEB)
        LP_HALF_CASE(EPI_RESID_SCALED)
        LP_HALF_CASE(EPI_RESID_MASK)
        LP_HALF_CASE(EPI_JACOBI)
        LP_HALF_CASE(EPI_POWER)
        LP_HALF_CASE(EPI_RESID)
        LP_HALF_CASE(EPI_MASKED)
        LP_HALF_CASE(EPI_CHEB_FIRST)
        LP_HALF_CASE(EPI_CHEB_ONLY)
        LP_HALF_CASE(EPI_CHEB_LAST)
        LP_HALF_CASE(EPI_CHEB_LAST_R)
        default:
            set_error("level operator: unknown epilogue");
            throw ArgFailure();
    }
#undef LP_HALF_CASE
    LP_HIP(hipGetLastError());
}

}  // namespace lp
