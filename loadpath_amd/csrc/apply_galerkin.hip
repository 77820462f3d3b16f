// First coarse level of the mixed-precision V-cycle applied as the Galerkin
// product it is: A1 = P^T (m K0 m) P, with P the trilinear prolongation and m
// the fine mask.
//
// Reference: the weights level (mg.py:609-642) stores the same operator as
// sum_q E_child(q) M_q plus explicit patches for cells touching held dofs;
// both forms equal P^T D K0 D P (a child with E > 0 that touches no held
// node has every corner free, so D leaves it unchanged).
//
// Here: prolong the masked coarse vector onto the fine grid, apply the fine
// operator with the FP16-MFMA contraction and an output mask, then restrict
// and finish (cm * R y + (1 - cm) * u) with the level's epilogue in one
// pass. At 256x128x128 this replaces a 4.4M-thread FMA/LDS-bound kernel by
// three bandwidth-bound passes over the fine grid.
#include "apply_impl.h"

namespace lp {

namespace {

template <typename T>
__global__ void __launch_bounds__(256) k_restrict_epi(const T* __restrict__ yf, Geom f, size_t fstride,
                                                      const T* __restrict__ v0, size_t vstride,
                                                      const uint8_t* __restrict__ cmask, Geom c, EpiArgs E) {
    using namespace apply_detail;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rhs = blockIdx.y;
    double part = 0.0;
    const bool skip = E.scal != nullptr && !E.scal[rhs].active;
    if (!skip && i < c.nn) {
        const int Z = (int)(i % c.Nz);
        const int Y = (int)((i / c.Nz) % c.Ny);
        const int X = (int)(i / ((size_t)c.Nz * c.Ny));
        const T* fr = yf + (size_t)rhs * fstride;
        T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int a = -1; a <= 1; ++a) {
            const int x = 2 * X + a;
            if (x < 0 || x >= f.Nx) continue;
            const T wa = a != 0 ? T(0.5) : T(1);
#pragma unroll
            for (int b = -1; b <= 1; ++b) {
                const int y = 2 * Y + b;
                if (y < 0 || y >= f.Ny) continue;
                const T wb = wa * (b != 0 ? T(0.5) : T(1));
#pragma unroll
                for (int e = -1; e <= 1; ++e) {
                    const int z = 2 * Z + e;
                    if (z < 0 || z >= f.Nz) continue;
                    const T w = wb * (e != 0 ? T(0.5) : T(1));
                    const size_t k = ((size_t)x * f.Ny + y) * f.Nz + z;
#pragma unroll
                    for (int q = 0; q < 3; ++q) acc[q] += fr[q * f.nn + k] * w;
                }
            }
        }
        const int m = cmask[i];
        const T* v = v0 + (size_t)rhs * vstride;
        T yv[3], ur[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            ur[q] = v[q * c.nn + i];
            yv[q] = ((m >> q) & 1) ? acc[q] : ur[q];
        }
        epilogue<T>(E, m, (size_t)rhs * vstride + i, c.nn, i, yv, ur, part);
    }
    if (E.partials != nullptr) block_partial(part, E, rhs, blockIdx.x);
}

}  // namespace

bool galerkin1_path(const Grid& G, int level, int eb) {
    return G.opt.precision == LP_PREC_MIXED && eb == 4 && level == 1 && G.lv[1].kind == LK_WEIGHTS &&
           G.gal_a.p != nullptr;
}

void launch_apply_galerkin1(const Grid& G, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                            hipStream_t s) {
    const Level& C = G.lv[level];
    const Level& F = G.lv[level - 1];
    const size_t fstride = 3 * F.g.nn;
    float* fa = G.gal_a.as<float>();
    float* fb = G.gal_b.as<float>();
    launch_prolong_set(u, C.mask.as<uint8_t>(), C.g, fa, F.g, 4, nrhs, vstride, fstride, s);
    EpiArgs E2;
    E2.kind = EPI_MASKED;
    E2.out = fb;
    E2.scal = E.scal;
    launch_apply_fine_mfma(G, level - 1, fa, fstride, nrhs, E2, s);
    dim3 grid((unsigned)((C.g.nn + kBlock - 1) / kBlock), nrhs);
    hipLaunchKernelGGL(k_restrict_epi<float>, grid, dim3(kBlock), 0, s, fb, F.g, fstride, static_cast<const float*>(u),
                       vstride, C.mask.as<uint8_t>(), C.g, E);
    LP_HIP(hipGetLastError());
}

}  // namespace lp
