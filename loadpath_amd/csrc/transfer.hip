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
        hipLaunchKernelGGL(k_restrict<float>, grid, dim3(kBlock), 0, s, (const float*)rf, cmask, (float*)rc, g,
                           fstride, cstride);
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
        hipLaunchKernelGGL((k_prolong<float, true>), grid, dim3(kBlock), 0, s, (const float*)xc, cmask, fmask,
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
