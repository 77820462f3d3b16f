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
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= g.cnn) return;
    const int Z = (int)(i % g.cz);
    const int Y = (int)((i / g.cz) % g.cy);
    const int X = (int)(i / ((size_t)g.cz * g.cy));
    const int m = cmask ? (int)cmask[i] : 7;
    const T* f = rf + rhs * fstride;
    for (int c = 0; c < 3; ++c) {
        T acc = T(0);
        const T* fc = f + c * g.fnn;
#pragma unroll
        for (int a = -1; a <= 1; ++a) {
            const int x = 2 * X + a;
            if (x < 0 || x >= g.fx) continue;
            const T wa = a != 0 ? T(0.5) : T(1);
#pragma unroll
            for (int b = -1; b <= 1; ++b) {
                const int y = 2 * Y + b;
                if (y < 0 || y >= g.fy) continue;
                const T wb = wa * (b != 0 ? T(0.5) : T(1));
#pragma unroll
                for (int e = -1; e <= 1; ++e) {
                    const int z = 2 * Z + e;
                    if (z < 0 || z >= g.fz) continue;
                    acc += fc[((size_t)x * g.fy + y) * g.fz + z] * (wb * (e != 0 ? T(0.5) : T(1)));
                }
            }
        }
        rc[rhs * cstride + c * g.cnn + i] = ((m >> c) & 1) ? acc : T(0);
    }
}

template <typename T, bool ADD>
__global__ void k_prolong(const T* __restrict__ xc, const uint8_t* __restrict__ cmask,
                          const uint8_t* __restrict__ fmask, T* __restrict__ xf, TG g, size_t cstride,
                          size_t fstride) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    if (i >= g.fnn) return;
    const int z = (int)(i % g.fz);
    const int y = (int)((i / g.fz) % g.fy);
    const int x = (int)(i / ((size_t)g.fz * g.fy));
    const int x0 = x >> 1, y0 = y >> 1, z0 = z >> 1;
    const int x1 = x0 + (x & 1), y1 = y0 + (y & 1), z1 = z0 + (z & 1);
    const T wx = (x & 1) ? T(0.5) : T(1), wy = (y & 1) ? T(0.5) : T(1), wz = (z & 1) ? T(0.5) : T(1);
    const int fm = fmask ? (int)fmask[i] : 7;
    const T* c0 = xc + rhs * cstride;
    auto cidx = [&](int a, int b, int e) { return ((size_t)a * g.cy + b) * g.cz + e; };
    for (int c = 0; c < 3; ++c) {
        const T* cc = c0 + c * g.cnn;
        auto val = [&](int a, int b, int e) {
            const size_t k = cidx(a, b, e);
            const T v = cc[k];
            if (cmask) return ((cmask[k] >> c) & 1) ? v : T(0);
            return v;
        };
        T acc = val(x0, y0, z0);
        if (z & 1) acc = acc + val(x0, y0, z1);
        if (y & 1) {
            acc = acc + val(x0, y1, z0);
            if (z & 1) acc = acc + val(x0, y1, z1);
        }
        if (x & 1) {
            T bb = val(x1, y0, z0);
            if (z & 1) bb = bb + val(x1, y0, z1);
            if (y & 1) {
                bb = bb + val(x1, y1, z0);
                if (z & 1) bb = bb + val(x1, y1, z1);
            }
            acc = acc + bb;
        }
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
