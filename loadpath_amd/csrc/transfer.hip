// Grid transfers between a level and the next coarser one.
//
// Reference: mg.prolong / mg.restrict (mg.py:84-132) and the fused V-cycle
// steps mgops._residual_pad + _restrict_masked (coarse_residual) and
// mgops._prolong_add (add_correction), mgops.py:666-773.
//
// restrict: one thread per coarse node gathers its 27 fine neighbours with
// the tensor weights (1, 1/2, 1/4, 1/8); the fine residual (b - Ax)*m comes
// from the operator's epilogue, so residual, pad and restrict are one extra
// read of the fine field. prolong_add: one thread per fine node gathers its
// <= 8 coarse parents, in the reference's summation order.
#include "lp_internal.h"

namespace lp {

namespace {

struct TG {
    int fx, fy, fz;   // fine node dims (actual field extent)
    int cx, cy, cz;   // coarse node dims
    size_t fnn, cnn;
};

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
    const T* f = rf + rhs * fstride;
    // the 27 fine neighbours' offsets and weights, shared by the components
    unsigned off[27];
    T w[27];
#pragma unroll
    for (int a = -1; a <= 1; ++a)
#pragma unroll
        for (int b = -1; b <= 1; ++b)
#pragma unroll
            for (int e = -1; e <= 1; ++e) {
                const int k = (a + 1) * 9 + (b + 1) * 3 + (e + 1);
                const int x = 2 * X + a, y = 2 * Y + b, z = 2 * Z + e;
                const bool in = x >= 0 && x < g.fx && y >= 0 && y < g.fy && z >= 0 && z < g.fz;
                off[k] = in ? ((unsigned)x * (unsigned)g.fy + (unsigned)y) * (unsigned)g.fz + (unsigned)z : 0u;
                w[k] = in ? (a != 0 ? T(0.5) : T(1)) * (b != 0 ? T(0.5) : T(1)) * (e != 0 ? T(0.5) : T(1)) : T(0);
            }
    for (int c = 0; c < 3; ++c) {
        T acc = T(0);
        const T* fc = f + c * g.fnn;
        // the same order as a loop over a, b, e skipping the outside ones
        // (adding an exact zero for those changes nothing)
#pragma unroll
        for (int k = 0; k < 27; ++k) acc += fc[off[k]] * w[k];
        rc[rhs * cstride + c * g.cnn + i] = ((m >> c) & 1) ? acc : T(0);
    }
}

template <typename T, bool ADD>
__global__ void k_prolong(const T* __restrict__ xc, const uint8_t* __restrict__ cmask,
                          const uint8_t* __restrict__ fmask, T* __restrict__ xf, TG g, size_t cstride,
                          size_t fstride) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= (unsigned)g.fnn) return;
    const unsigned fyz = (unsigned)g.fy * (unsigned)g.fz;
    const unsigned x = i / fyz;
    const unsigned rem = i - x * fyz;
    const unsigned y = rem / (unsigned)g.fz;
    const unsigned z = rem - y * (unsigned)g.fz;
    const unsigned xo = x & 1, yo = y & 1, zo = z & 1;
    const unsigned x0 = x >> 1, y0 = y >> 1, z0 = z >> 1;
    const unsigned x1 = x0 + xo, y1 = y0 + yo, z1 = z0 + zo;
    const T wx = xo ? T(0.5) : T(1), wy = yo ? T(0.5) : T(1), wz = zo ? T(0.5) : T(1);
    const int fm = fmask ? (int)fmask[i] : 7;
    // the 8 parent slots (a repeated parent for an even coordinate is
    // loaded but added as an exact zero): offsets and masks, once
    unsigned k8[8];
    int m8[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const unsigned a = (q & 1) ? x1 : x0, b = (q & 2) ? y1 : y0, e = (q & 4) ? z1 : z0;
        k8[q] = (a * (unsigned)g.cy + b) * (unsigned)g.cz + e;
        m8[q] = cmask ? (int)cmask[k8[q]] : 7;
    }
    const T* c0 = xc + rhs * cstride;
    for (int c = 0; c < 3; ++c) {
        const T* cc = c0 + c * g.cnn;
        T v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = ((m8[q] >> c) & 1) ? cc[k8[q]] : T(0);
        // the summation order of the reference: (x0: z0, z1, then y1 pair),
        // then the x1 group likewise
        T acc = v[0];
        acc = acc + (zo ? v[4] : T(0));
        acc = acc + (yo ? v[2] : T(0));
        acc = acc + ((yo & zo) ? v[6] : T(0));
        T bb = v[1];
        bb = bb + (zo ? v[5] : T(0));
        bb = bb + (yo ? v[3] : T(0));
        bb = bb + ((yo & zo) ? v[7] : T(0));
        acc = acc + (xo ? bb : T(0));
        const T tmp = acc * (wx * wy * wz);
        const size_t k = rhs * fstride + c * g.fnn + i;
        if (ADD) {
            if ((fm >> c) & 1) xf[k] = xf[k] + tmp;
        } else {
            xf[k] = tmp;
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
    dim3 grid((unsigned)((g.fnn + kBlock - 1) / kBlock), nrhs);
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
    hipLaunchKernelGGL((k_prolong<double, false>), dim3((unsigned)((g.fnn + kBlock - 1) / kBlock), 1), dim3(kBlock),
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
