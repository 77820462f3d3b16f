// Grid transfers between a level and the next coarser one.
//
// Reference: mg.prolong / mg.restrict (mg.py:84-132) and the fused V-cycle
// steps mgops._residual_pad + _restrict_masked (coarse_residual) and
// mgops._prolong_add (add_correction), mgops.py:666-773.
//
// restrict: one thread per coarse node gathers its 27 fine neighbours with
// the tensor weights (1, 1/2, 1/4, 1/8); the fine residual (b - Ax)*m comes
// from the operator's epilogue, so residual, pad and restrict are one extra
// read of the fine field. prolong_add: one thread per coarse node forms its
// <= 8 fine children from their <= 8 coarse parents each, in the
// reference's summation order.
#include "lp_internal.h"

namespace lp {

namespace {

struct TG {
    int fx, fy, fz;   // fine node dims (actual field extent)
    int cx, cy, cz;   // coarse node dims
    size_t fnn, cnn;
};

// raw buffer loads: an offset past the buffer (a neighbour outside the
// grid, including a "negative" one) loads as zero with no branch
constexpr unsigned kFar = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ T bld(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ __forceinline__ float bld<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <> __device__ __forceinline__ double bld<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

template <typename T>
__global__ void k_restrict(const T* __restrict__ rf, const uint8_t* __restrict__ cmask, T* __restrict__ rc, TG g,
                           size_t fstride, size_t cstride) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= (unsigned)g.cnn) return;
    const unsigned cyz = (unsigned)g.cy * (unsigned)g.cz;
    const int X = (int)(i / cyz);
    const unsigned rem = i - (unsigned)X * cyz;
    const int Y = (int)(rem / (unsigned)g.cz);
    const int Z = (int)(rem - (unsigned)Y * (unsigned)g.cz);
    const int m = cmask ? (int)cmask[i] : 7;
    // byte offsets of the 9 fine rows' centres (z = 2Z); a row outside the
    // grid, and z = 2Z +- 1 outside it, fall past the buffer
    const bool zlo = 2 * Z - 1 >= 0, zhi = 2 * Z + 1 < g.fz;
    unsigned row[9];
#pragma unroll
    for (int a = -1; a <= 1; ++a)
#pragma unroll
        for (int b = -1; b <= 1; ++b) {
            const int x = 2 * X + a, y = 2 * Y + b;
            const bool in = x >= 0 && x < g.fx && y >= 0 && y < g.fy;
            // kFar + 16: its z neighbours stay past the buffer too
            row[(a + 1) * 3 + (b + 1)] =
                in ? (((unsigned)x * (unsigned)g.fy + (unsigned)y) * (unsigned)g.fz + (unsigned)(2 * Z)) * sizeof(T)
                   : kFar + 16u;
        }
    constexpr unsigned s1 = (unsigned)sizeof(T);
    for (int c = 0; c < 3; ++c) {
        const __amdgpu_buffer_rsrc_t r = rsrc(rf + rhs * fstride + c * g.fnn, (unsigned)(g.fnn * sizeof(T)));
        T acc = T(0);
        // the reference's order (a, b, e), the weights 1, 1/2, 1/4, 1/8; a
        // neighbour outside adds an exact zero
#pragma unroll
        for (int a = -1; a <= 1; ++a) {
            const T wa = a != 0 ? T(0.5) : T(1);
#pragma unroll
            for (int b = -1; b <= 1; ++b) {
                const T wb = wa * (b != 0 ? T(0.5) : T(1));
                const unsigned o = row[(a + 1) * 3 + (b + 1)];
                acc += bld<T>(r, zlo ? o - s1 : kFar) * (wb * T(0.5));
                acc += bld<T>(r, o) * wb;
                acc += bld<T>(r, zhi ? o + s1 : kFar) * (wb * T(0.5));
            }
        }
        rc[rhs * cstride + c * g.cnn + i] = ((m >> c) & 1) ? acc : T(0);
    }
}

// FP32: a fine row's three values (2Z - 1, 2Z, 2Z + 1) in one 12-byte load
// (a third of the loads; the same products, summed in the same order). At
// Z = 0 the load starts at 2Z; where 2Z + 1 is past the row its value is
// discarded (the V-cycle's vectors carry a tail, so the element past a
// field's last row is still inside the allocation, and the range below
// admits it).
__global__ void k_restrict_f32(const float* __restrict__ rf, const uint8_t* __restrict__ cmask,
                               float* __restrict__ rc, TG g, size_t fstride, size_t cstride) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= (unsigned)g.cnn) return;
    const unsigned cyz = (unsigned)g.cy * (unsigned)g.cz;
    const int X = (int)(i / cyz);
    const unsigned rem = i - (unsigned)X * cyz;
    const int Y = (int)(rem / (unsigned)g.cz);
    const int Z = (int)(rem - (unsigned)Y * (unsigned)g.cz);
    const int m = cmask ? (int)cmask[i] : 7;
    const bool zlo = 2 * Z - 1 >= 0, zhi = 2 * Z + 1 < g.fz;
    const unsigned sh = zlo ? 4u : 0u;
    unsigned row[9];
#pragma unroll
    for (int a = -1; a <= 1; ++a)
#pragma unroll
        for (int b = -1; b <= 1; ++b) {
            const int x = 2 * X + a, y = 2 * Y + b;
            const bool in = x >= 0 && x < g.fx && y >= 0 && y < g.fy;
            row[(a + 1) * 3 + (b + 1)] =
                in ? (((unsigned)x * (unsigned)g.fy + (unsigned)y) * (unsigned)g.fz + (unsigned)(2 * Z)) * 4u - sh
                   : kFar + 16u;
        }
    for (int c = 0; c < 3; ++c) {
        const __amdgpu_buffer_rsrc_t r = rsrc(rf + rhs * fstride + c * g.fnn, (unsigned)(g.fnn * 4 + 4));
        float acc = 0.0f;
#pragma unroll
        for (int a = -1; a <= 1; ++a) {
            const float wa = a != 0 ? 0.5f : 1.0f;
#pragma unroll
            for (int b = -1; b <= 1; ++b) {
                const float wb = wa * (b != 0 ? 0.5f : 1.0f);
                typedef unsigned u3 __attribute__((ext_vector_type(3)));
                const u3 v = __builtin_bit_cast(u3, __builtin_amdgcn_raw_buffer_load_b96(r, row[(a + 1) * 3 + (b + 1)], 0, 0));
                const float v0 = __uint_as_float(v.x), v1 = __uint_as_float(v.y), v2 = __uint_as_float(v.z);
                const float lo = zlo ? v0 : 0.0f;
                const float mid = zlo ? v1 : v0;
                const float hi = zhi ? (zlo ? v2 : v1) : 0.0f;
                acc += lo * (wb * 0.5f);
                acc += mid * wb;
                acc += hi * (wb * 0.5f);
            }
        }
        rc[rhs * cstride + c * g.cnn + i] = ((m >> c) & 1) ? acc : 0.0f;
    }
}

// one thread per coarse node: its (up to) 2 x 2 x 2 fine children
// (2X + dx, 2Y + dy, 2Z + dz), each the sum of its parents among the coarse
// node and its +x, +y, +z neighbours in the reference's summation order,
// times the weights 1 or 1/2 per odd coordinate. Consecutive threads touch
// consecutive fine pairs in z.
template <typename T, bool ADD>
__global__ void k_prolong(const T* __restrict__ xc, const uint8_t* __restrict__ cmask,
                          const uint8_t* __restrict__ fmask, T* __restrict__ xf, TG g, size_t cstride,
                          size_t fstride) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= (unsigned)g.cnn) return;
    const unsigned cyz = (unsigned)g.cy * (unsigned)g.cz;
    const unsigned X = i / cyz;
    const unsigned rem = i - X * cyz;
    const unsigned Y = rem / (unsigned)g.cz;
    const unsigned Z = rem - Y * (unsigned)g.cz;
    // the +1 neighbours (clamped at the last node, where no child uses them)
    const unsigned X1 = min(X + 1, (unsigned)g.cx - 1), Y1 = min(Y + 1, (unsigned)g.cy - 1),
                   Z1 = min(Z + 1, (unsigned)g.cz - 1);
    unsigned kc[8];
    int mc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const unsigned a = (q & 1) ? X1 : X, b = (q & 2) ? Y1 : Y, e = (q & 4) ? Z1 : Z;
        kc[q] = (a * (unsigned)g.cy + b) * (unsigned)g.cz + e;
        mc[q] = cmask ? (int)cmask[kc[q]] : 7;
    }
    // children inside the fine grid and their offsets / masks
    bool inc[8];
    unsigned kf[8];
    int mf[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        const unsigned x = 2 * X + (d & 1), y = 2 * Y + ((d >> 1) & 1), z = 2 * Z + ((d >> 2) & 1);
        inc[d] = x < (unsigned)g.fx && y < (unsigned)g.fy && z < (unsigned)g.fz;
        kf[d] = (x * (unsigned)g.fy + y) * (unsigned)g.fz + z;
        mf[d] = (ADD && fmask && inc[d]) ? (int)fmask[kf[d]] : 7;
    }
    const T* c0 = xc + rhs * cstride;
    T* f0 = xf + rhs * fstride;
    for (int c = 0; c < 3; ++c) {
        const T* cc = c0 + c * g.cnn;
        T v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = ((mc[q] >> c) & 1) ? cc[kc[q]] : T(0);
        T* fc = f0 + c * g.fnn;
#pragma unroll
        for (int d = 0; d < 8; ++d) {
            if (!inc[d]) continue;
            const int dx = d & 1, dy = (d >> 1) & 1, dz = (d >> 2) & 1;
            // parents: q with qx <= dx, qy <= dy, qz <= dz (q = qx + 2 qy + 4 qz)
            T acc = v[0];
            if (dz) acc = acc + v[4];
            if (dy) {
                acc = acc + v[2];
                if (dz) acc = acc + v[6];
            }
            if (dx) {
                T bb = v[1];
                if (dz) bb = bb + v[5];
                if (dy) {
                    bb = bb + v[3];
                    if (dz) bb = bb + v[7];
                }
                acc = acc + bb;
            }
            const T w = (dx ? T(0.5) : T(1)) * (dy ? T(0.5) : T(1)) * (dz ? T(0.5) : T(1));
            const T tmp = acc * w;
            if (ADD) {
                if ((mf[d] >> c) & 1) fc[kf[d]] = fc[kf[d]] + tmp;
            } else {
                fc[kf[d]] = tmp;
            }
        }
    }
}

// FP32 prolongation-add: the same children and sums as k_prolong, with each
// pair of z-neighbour children (dz = 0, 1: adjacent words) read and written
// as one 8-byte access (a masked-out dof is written back unchanged); where
// 2Z + 1 is past the row the dz = 0 child alone
template <bool ADD>
__global__ void k_prolong_f32(const float* __restrict__ xc, const uint8_t* __restrict__ cmask,
                              const uint8_t* __restrict__ fmask, float* __restrict__ xf, TG g, size_t cstride,
                              size_t fstride) {
    typedef float f2 __attribute__((ext_vector_type(2), aligned(4)));
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= (unsigned)g.cnn) return;
    const unsigned cyz = (unsigned)g.cy * (unsigned)g.cz;
    const unsigned X = i / cyz;
    const unsigned rem = i - X * cyz;
    const unsigned Y = rem / (unsigned)g.cz;
    const unsigned Z = rem - Y * (unsigned)g.cz;
    const unsigned X1 = min(X + 1, (unsigned)g.cx - 1), Y1 = min(Y + 1, (unsigned)g.cy - 1),
                   Z1 = min(Z + 1, (unsigned)g.cz - 1);
    unsigned kc[8];
    int mc[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const unsigned a = (q & 1) ? X1 : X, b = (q & 2) ? Y1 : Y, e = (q & 4) ? Z1 : Z;
        kc[q] = (a * (unsigned)g.cy + b) * (unsigned)g.cz + e;
        mc[q] = cmask ? (int)cmask[kc[q]] : 7;
    }
    // the 4 (dx, dy) child rows: in the grid, offset of the dz = 0 child,
    // whether the dz = 1 child exists, masks of the pair
    bool inr[4], two[4];
    unsigned kf[4];
    int mf0[4], mf1[4];
    const bool zin1 = 2 * Z + 1 < (unsigned)g.fz;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const unsigned x = 2 * X + (e & 1), y = 2 * Y + ((e >> 1) & 1);
        inr[e] = x < (unsigned)g.fx && y < (unsigned)g.fy && 2 * Z < (unsigned)g.fz;
        two[e] = inr[e] && zin1;
        kf[e] = (x * (unsigned)g.fy + y) * (unsigned)g.fz + 2 * Z;
        mf0[e] = (ADD && fmask && inr[e]) ? (int)fmask[kf[e]] : 7;
        mf1[e] = (ADD && fmask && two[e]) ? (int)fmask[kf[e] + 1] : 7;
    }
    const float* c0 = xc + rhs * cstride;
    float* f0 = xf + rhs * fstride;
    for (int c = 0; c < 3; ++c) {
        const float* cc = c0 + c * g.cnn;
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = ((mc[q] >> c) & 1) ? cc[kc[q]] : 0.0f;
        float* fc = f0 + c * g.fnn;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (!inr[e]) continue;
            const int dx = e & 1, dy = (e >> 1) & 1;
            float t[2];
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) {
                // parents: q with qx <= dx, qy <= dy, qz <= dz, the reference's order
                float acc = v[0];
                if (dz) acc = acc + v[4];
                if (dy) {
                    acc = acc + v[2];
                    if (dz) acc = acc + v[6];
                }
                if (dx) {
                    float bb = v[1];
                    if (dz) bb = bb + v[5];
                    if (dy) {
                        bb = bb + v[3];
                        if (dz) bb = bb + v[7];
                    }
                    acc = acc + bb;
                }
                const float w = (dx ? 0.5f : 1.0f) * (dy ? 0.5f : 1.0f) * (dz ? 0.5f : 1.0f);
                t[dz] = acc * w;
            }
            float* p = fc + kf[e];
            if (two[e]) {
                f2 cur = ADD ? *reinterpret_cast<const f2*>(p) : f2{0.0f, 0.0f};
                if (ADD) {
                    if ((mf0[e] >> c) & 1) cur.x = cur.x + t[0];
                    if ((mf1[e] >> c) & 1) cur.y = cur.y + t[1];
                } else {
                    cur.x = t[0];
                    cur.y = t[1];
                }
                *reinterpret_cast<f2*>(p) = cur;
            } else {
                if (ADD) {
                    if ((mf0[e] >> c) & 1) p[0] = p[0] + t[0];
                } else {
                    p[0] = t[0];
                }
            }
        }
    }
}

TG make_tg(const Geom& fine, const Geom& coarse) {
    TG g;
    g.fx = fine.Nx; g.fy = fine.Ny; g.fz = fine.Nz;
    g.cx = coarse.Nx; g.cy = coarse.Ny; g.cz = coarse.Nz;
    g.fnn = fine.nn;
    g.cnn = coarse.nn;
    return g;
}

}  // namespace

void launch_restrict_masked(const void* rf, const Geom& fine, const uint8_t* cmask, void* rc, const Geom& coarse,
                            int eb, int nrhs, size_t fstride, size_t cstride, hipStream_t s) {
    TG g = make_tg(fine, coarse);
    dim3 grid((unsigned)((g.cnn + kBlock - 1) / kBlock), nrhs);
    if (eb == 8)
        hipLaunchKernelGGL(k_restrict<double>, grid, dim3(kBlock), 0, s, (const double*)rf, cmask, (double*)rc, g,
                           fstride, cstride);
    else
        hipLaunchKernelGGL(k_restrict_f32, grid, dim3(kBlock), 0, s, (const float*)rf, cmask, (float*)rc, g, fstride,
                           cstride);
    LP_HIP(hipGetLastError());
}

void launch_prolong_add(const void* xc, const uint8_t* cmask, const Geom& coarse, const uint8_t* fmask, void* xf,
                        const Geom& fine, int eb, int nrhs, size_t cstride, size_t fstride, hipStream_t s) {
    TG g = make_tg(fine, coarse);
    dim3 grid((unsigned)((g.cnn + kBlock - 1) / kBlock), nrhs);
    if (eb == 8)
        hipLaunchKernelGGL((k_prolong<double, true>), grid, dim3(kBlock), 0, s, (const double*)xc, cmask, fmask,
                           (double*)xf, g, cstride, fstride);
    else
        hipLaunchKernelGGL((k_prolong_f32<true>), grid, dim3(kBlock), 0, s, (const float*)xc, cmask, fmask,
                           (float*)xf, g, cstride, fstride);
    LP_HIP(hipGetLastError());
}

void launch_prolong_plain(const double* c, double* f, int cx, int cy, int cz, hipStream_t s) {
    TG g;
    g.cx = cx; g.cy = cy; g.cz = cz;
    g.fx = 2 * cx - 1; g.fy = 2 * cy - 1; g.fz = 2 * cz - 1;
    g.cnn = (size_t)cx * cy * cz;
    g.fnn = (size_t)g.fx * g.fy * g.fz;
    hipLaunchKernelGGL((k_prolong<double, false>), dim3((unsigned)((g.cnn + kBlock - 1) / kBlock), 1), dim3(kBlock),
                       0, s, c, nullptr, nullptr, f, g, 0, 0);
    LP_HIP(hipGetLastError());
}

void launch_restrict_plain(const double* f, double* c, int cx, int cy, int cz, hipStream_t s) {
    TG g;
    g.cx = cx; g.cy = cy; g.cz = cz;
    g.fx = 2 * cx - 1; g.fy = 2 * cy - 1; g.fz = 2 * cz - 1;
    g.cnn = (size_t)cx * cy * cz;
    g.fnn = (size_t)g.fx * g.fy * g.fz;
    hipLaunchKernelGGL(k_restrict<double>, dim3((unsigned)((g.cnn + kBlock - 1) / kBlock), 1), dim3(kBlock), 0, s,
                       f, nullptr, c, g, 0, 0);
    LP_HIP(hipGetLastError());
}

}  // namespace lp
