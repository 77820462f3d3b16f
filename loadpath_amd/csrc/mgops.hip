// The kernel-level switches of the reference's compiled CPU kernels, on the
// device: mgops.apply_masked / coarse_residual / add_correction / dot / cast
// (mgops.py:247-261, 557-585, 753-773) and femops.element_energy
// (femops.py:236-242). Each takes the reference's own arrays (a level's
// cell list, per-cell factors, float masks) and gives what the reference's
// function gives, with the same per-value operations in the same order
// (FP contraction off), so the numbers agree to the bit where the
// reference's own kernel is not compiled with fast-math (its weights
// product is). Host arrays in and out; the library stages them.
#include <algorithm>
#include <vector>

#include "lp_internal.h"

#pragma clang fp contract(off)

namespace lp {
namespace {

constexpr int kB = 256;

inline unsigned nb_of(size_t n) { return (unsigned)((n + kB - 1) / kB); }

// out at each node of the (nx+1, ny+1, nz+1) grid: the sum over the node's
// cells k = 0..7 (cell (x - k&1, y - (k>>1)&1, z - (k>>2)&1), the node its
// corner k) of the cell's rows 3k .. 3k+2 of K_c (u m)_c, the cell's product
// formed as mgops._product_* forms it; finish: out = sum m + (1 - m) u
template <typename T, int KIND>
__global__ void k_cells_apply(int nx, int ny, int nz, const long long* __restrict__ rows, long long nact,
                              const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ u,
                              const T* __restrict__ m, T* __restrict__ out, int finish) {
    const int Nx = nx + 1, Ny = ny + 1, Nz = nz + 1;
    const size_t nn = (size_t)Nx * Ny * Nz;
    const size_t node = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= nn) return;
    const int z = (int)(node % Nz), y = (int)((node / Nz) % Ny), x = (int)(node / ((size_t)Nz * Ny));
    T acc[3] = {T(0), T(0), T(0)};
    for (int k = 0; k < 8; ++k) {
        const int cx = x - (k & 1), cy = y - ((k >> 1) & 1), cz = z - ((k >> 2) & 1);
        if (cx < 0 || cx >= nx || cy < 0 || cy >= ny || cz < 0 || cz >= nz) continue;
        const long long r = rows[((size_t)cx * ny + cy) * nz + cz];
        if (r < 0) continue;
        // the cell's 24 corner values (u m), dof 3 corner + component
        T ub[24];
        for (int q = 0; q < 8; ++q) {
            const size_t nq = ((size_t)(cx + (q & 1)) * Ny + (cy + ((q >> 1) & 1))) * Nz + (cz + ((q >> 2) & 1));
            for (int d = 0; d < 3; ++d) {
                const size_t i = d * nn + nq;
                ub[3 * q + d] = m ? u[i] * m[i] : u[i];
            }
        }
        for (int d = 0; d < 3; ++d) {
            const int row = 3 * k + d;
            T f = T(0);
            if (KIND == 0) {
                // scaled: e_c Ke u_c (a = Ke, b = e)
                for (int j = 0; j < 24; ++j) f += a[row * 24 + j] * ub[j];
                f *= b[r];
            } else if (KIND == 1) {
                // weights: sum_q M_q (w_q u_c) (a = M (8, 24, 24), b = w (8, n))
                for (int q = 0; q < 8; ++q) {
                    const T wq = b[(size_t)q * nact + r];
                    for (int j = 0; j < 24; ++j) f += a[(size_t)q * 576 + row * 24 + j] * (wq * ub[j]);
                }
            } else {
                // explicit: K_c u_c (a = K (24, 24, n))
                for (int j = 0; j < 24; ++j) f += a[((size_t)row * 24 + j) * nact + r] * ub[j];
            }
            acc[d] += f;
        }
    }
    for (int d = 0; d < 3; ++d) {
        const size_t i = d * nn + node;
        T v = acc[d];
        if (finish) v = v * m[i] + (T(1) - m[i]) * u[i];
        out[i] = v;
    }
}

// c = restrict(pad((b - ax) m)) cm on the coarse grid (cx, cy, cz) nodes:
// the fine residual on b's (fx, fy, fz) nodes (in T), zero on the rest of
// the (2 cx - 1, ...) grid, gathered with the 27-point weights in the order
// of mgops._restrict_masked (mgops.py:689-716), whose sum runs in double
// (the float64 weights promote it), rounded to T before cm
template <typename T>
__global__ void k_coarse_residual(const T* __restrict__ b, const T* __restrict__ ax, const T* __restrict__ m,
                                  const T* __restrict__ cm, T* __restrict__ c, int fx, int fy, int fz, int cx, int cy,
                                  int cz) {
    const size_t cn = (size_t)cx * cy * cz;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * cn) return;
    const int d = (int)(t / cn);
    const size_t p = t % cn;
    const int Z = (int)(p % cz), Y = (int)((p / cz) % cy), X = (int)(p / ((size_t)cz * cy));
    const int px = 2 * cx - 1, py = 2 * cy - 1, pz = 2 * cz - 1;     // the padded fine grid
    const size_t fn = (size_t)fx * fy * fz;
    double acc = 0.0;
    for (int ia = -1; ia <= 1; ++ia) {
        const int x = 2 * X + ia;
        if (x < 0 || x >= px) continue;
        const double wa = ia != 0 ? 0.5 : 1.0;
        for (int ib = -1; ib <= 1; ++ib) {
            const int y = 2 * Y + ib;
            if (y < 0 || y >= py) continue;
            const double wb = wa * (ib != 0 ? 0.5 : 1.0);
            for (int ie = -1; ie <= 1; ++ie) {
                const int z = 2 * Z + ie;
                if (z < 0 || z >= pz) continue;
                T rf = T(0);
                if (x < fx && y < fy && z < fz) {
                    const size_t i = d * fn + ((size_t)x * fy + y) * fz + z;
                    rf = (b[i] - ax[i]) * m[i];
                }
                acc += (double)rf * (wb * (ie != 0 ? 0.5 : 1.0));
            }
        }
    }
    const T tmp = (T)acc;
    c[d * cn + p] = tmp * cm[d * cn + p];
}

// x += prolong(xc cm)[:, :sx, :sy, :sz] m, each fine node from its <= 8
// coarse parents in the order of mgops._prolong_add (mgops.py:722-750)
template <typename T>
__global__ void k_add_correction(const T* __restrict__ xc, const T* __restrict__ cm, const T* __restrict__ m,
                                 T* __restrict__ x, int cx, int cy, int cz, int sx, int sy, int sz) {
    const size_t fn = (size_t)sx * sy * sz, cn = (size_t)cx * cy * cz;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * fn) return;
    const int d = (int)(t / fn);
    const size_t p = t % fn;
    const int zz = (int)(p % sz), yy = (int)((p / sz) % sy), xx = (int)(p / ((size_t)sz * sy));
    const int x0 = xx >> 1, y0 = yy >> 1, z0 = zz >> 1;
    const int x1 = x0 + (xx & 1), y1 = y0 + (yy & 1), z1 = z0 + (zz & 1);
    const double wx = (xx & 1) ? 0.5 : 1.0, wy = (yy & 1) ? 0.5 : 1.0, wz = (zz & 1) ? 0.5 : 1.0;
    auto at = [&](int i, int j, int k) {
        const size_t q = d * cn + ((size_t)i * cy + j) * cz + k;
        return xc[q] * cm[q];
    };
    T a = at(x0, y0, z0);
    if (zz & 1) a = a + at(x0, y0, z1);
    if (yy & 1) {
        a = a + at(x0, y1, z0);
        if (zz & 1) a = a + at(x0, y1, z1);
    }
    if (xx & 1) {
        T bb = at(x1, y0, z0);
        if (zz & 1) bb = bb + at(x1, y0, z1);
        if (yy & 1) {
            bb = bb + at(x1, y1, z0);
            if (zz & 1) bb = bb + at(x1, y1, z1);
        }
        a = a + bb;
    }
    const T tmp = (T)((double)a * (wx * wy * wz));
    const size_t i = d * fn + p;
    x[i] += tmp * m[i];
}

// femops.element_energy: sum_j sum_k (u_j ke_jk) v_k over each element's 24
// dofs, j outer, k inner (femops.py:4-7, 236-242)
__global__ void k_element_energy(const double* __restrict__ u, const double* __restrict__ v,
                                 const long long* __restrict__ edof, long long ne, const double* __restrict__ ke,
                                 double* __restrict__ out) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const long long* dofs = edof + e * 24;
    double vv[24];
    for (int k = 0; k < 24; ++k) vv[k] = v[dofs[k]];
    double acc = 0.0;
    for (int j = 0; j < 24; ++j) {
        const double uj = u[dofs[j]];
        for (int k = 0; k < 24; ++k) acc += (uj * ke[j * 24 + k]) * vv[k];
    }
    out[e] = acc;
}

// partial sums of a . b per block (in double), a fixed grid of blocks
template <typename T>
__global__ void k_dot_partials(const T* __restrict__ a, const T* __restrict__ b, size_t n, double* partials) {
    __shared__ double sred[kB / 64];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        acc += (double)a[i] * (double)b[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kB / 64; ++w) t += sred[w];
        partials[blockIdx.x] = t;
    }
}

template <typename S, typename D>
__global__ void k_cast_plain(const S* __restrict__ a, D* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (D)a[i];
}

struct Stage {
    std::vector<DBuf*> bufs;
    ~Stage() {
        for (DBuf* b : bufs) delete b;
    }
    template <typename P> P* up(const P* host, size_t count) {
        if (!host) return nullptr;
        DBuf* b = new DBuf();
        bufs.push_back(b);
        b->alloc(std::max<size_t>(count, 1) * sizeof(P));
        if (count) LP_HIP(hipMemcpy(b->p, host, count * sizeof(P), hipMemcpyHostToDevice));
        return static_cast<P*>(b->p);
    }
    template <typename P> P* make(size_t count) {
        DBuf* b = new DBuf();
        bufs.push_back(b);
        b->alloc(std::max<size_t>(count, 1) * sizeof(P));
        return static_cast<P*>(b->p);
    }
};

template <typename T>
int cells_apply(int kind, int nx, int ny, int nz, const long long* rows, long long nact, const void* a, const void* b,
                const void* u, const void* m, void* out, int finish) {
    const size_t nc = (size_t)nx * ny * nz, nn = (size_t)(nx + 1) * (ny + 1) * (nz + 1);
    const size_t na = kind == 0 ? 576 : kind == 1 ? 8 * 576 : 576 * (size_t)nact;
    const size_t nb = kind == 0 ? (size_t)nact : kind == 1 ? 8 * (size_t)nact : 0;
    Stage S;
    const long long* dr = S.up(rows, nc);
    const T* da = S.up(static_cast<const T*>(a), na);
    const T* db = nb ? S.up(static_cast<const T*>(b), nb) : nullptr;
    const T* du = S.up(static_cast<const T*>(u), 3 * nn);
    const T* dm = S.up(static_cast<const T*>(m), m ? 3 * nn : 0);
    T* dout = S.make<T>(3 * nn);
    const unsigned g = nb_of(nn);
    if (kind == 0)
        hipLaunchKernelGGL((k_cells_apply<T, 0>), dim3(g), dim3(kB), 0, nullptr, nx, ny, nz, dr, nact, da, db, du, dm,
                           dout, finish);
    else if (kind == 1)
        hipLaunchKernelGGL((k_cells_apply<T, 1>), dim3(g), dim3(kB), 0, nullptr, nx, ny, nz, dr, nact, da, db, du, dm,
                           dout, finish);
    else
        hipLaunchKernelGGL((k_cells_apply<T, 2>), dim3(g), dim3(kB), 0, nullptr, nx, ny, nz, dr, nact, da, db, du, dm,
                           dout, finish);
    LP_HIP(hipGetLastError());
    LP_HIP(hipMemcpy(out, dout, 3 * nn * sizeof(T), hipMemcpyDeviceToHost));
    return LP_OK;
}

template <typename T>
int coarse_residual(const void* b, const void* ax, const void* m, const void* cm, void* rc, int fx, int fy, int fz,
                    int cx, int cy, int cz) {
    const size_t fn = 3 * (size_t)fx * fy * fz, cn = 3 * (size_t)cx * cy * cz;
    Stage S;
    const T* db = S.up(static_cast<const T*>(b), fn);
    const T* dax = S.up(static_cast<const T*>(ax), fn);
    const T* dm = S.up(static_cast<const T*>(m), fn);
    const T* dcm = S.up(static_cast<const T*>(cm), cn);
    T* dc = S.make<T>(cn);
    hipLaunchKernelGGL(k_coarse_residual<T>, dim3(nb_of(cn)), dim3(kB), 0, nullptr, db, dax, dm, dcm, dc, fx, fy, fz,
                       cx, cy, cz);
    LP_HIP(hipGetLastError());
    LP_HIP(hipMemcpy(rc, dc, cn * sizeof(T), hipMemcpyDeviceToHost));
    return LP_OK;
}

template <typename T>
int add_correction(const void* xc, const void* cm, const void* m, void* x, int cx, int cy, int cz, int sx, int sy,
                   int sz) {
    const size_t fn = 3 * (size_t)sx * sy * sz, cn = 3 * (size_t)cx * cy * cz;
    Stage S;
    const T* dxc = S.up(static_cast<const T*>(xc), cn);
    const T* dcm = S.up(static_cast<const T*>(cm), cn);
    const T* dm = S.up(static_cast<const T*>(m), fn);
    T* dx = S.up(static_cast<const T*>(x), fn);
    hipLaunchKernelGGL(k_add_correction<T>, dim3(nb_of(fn)), dim3(kB), 0, nullptr, dxc, dcm, dm, dx, cx, cy, cz, sx,
                       sy, sz);
    LP_HIP(hipGetLastError());
    LP_HIP(hipMemcpy(x, dx, fn * sizeof(T), hipMemcpyDeviceToHost));
    return LP_OK;
}

}  // namespace
}  // namespace lp

using namespace lp;

extern "C" {

int lp_cells_apply(int kind, int nx, int ny, int nz, const long long* rows, long long n_active, const void* a,
                   const void* b, const void* u, const void* m, void* out, int finish, int dtype_bytes) {
    return api_call([&]() -> int {
        if (kind < 0 || kind > 2 || nx < 1 || ny < 1 || nz < 1 || !rows || n_active < 0 || !a || !u || !out ||
            (kind != 2 && !b) || (finish && !m) || (dtype_bytes != 4 && dtype_bytes != 8)) {
            set_error("lp_cells_apply: bad argument");
            return LP_ERR_ARG;
        }
        return dtype_bytes == 8 ? cells_apply<double>(kind, nx, ny, nz, rows, n_active, a, b, u, m, out, finish)
                                : cells_apply<float>(kind, nx, ny, nz, rows, n_active, a, b, u, m, out, finish);
    });
}

int lp_coarse_residual(const void* b, const void* ax, const void* m, const void* cm, void* rc, int fx, int fy, int fz,
                       int cx, int cy, int cz, int dtype_bytes) {
    return api_call([&]() -> int {
        if (!b || !ax || !m || !cm || !rc || fx < 1 || fy < 1 || fz < 1 || cx < 1 || cy < 1 || cz < 1 ||
            fx > 2 * cx - 1 || fy > 2 * cy - 1 || fz > 2 * cz - 1 || (dtype_bytes != 4 && dtype_bytes != 8)) {
            set_error("lp_coarse_residual: bad argument");
            return LP_ERR_ARG;
        }
        return dtype_bytes == 8 ? coarse_residual<double>(b, ax, m, cm, rc, fx, fy, fz, cx, cy, cz)
                                : coarse_residual<float>(b, ax, m, cm, rc, fx, fy, fz, cx, cy, cz);
    });
}

int lp_add_correction(const void* xc, const void* cm, const void* m, void* x, int cx, int cy, int cz, int sx, int sy,
                      int sz, int dtype_bytes) {
    return api_call([&]() -> int {
        if (!xc || !cm || !m || !x || cx < 1 || cy < 1 || cz < 1 || sx < 1 || sy < 1 || sz < 1 ||
            sx > 2 * cx - 1 || sy > 2 * cy - 1 || sz > 2 * cz - 1 || (dtype_bytes != 4 && dtype_bytes != 8)) {
            set_error("lp_add_correction: bad argument");
            return LP_ERR_ARG;
        }
        return dtype_bytes == 8 ? add_correction<double>(xc, cm, m, x, cx, cy, cz, sx, sy, sz)
                                : add_correction<float>(xc, cm, m, x, cx, cy, cz, sx, sy, sz);
    });
}

int lp_dot(const void* a, const void* b, long long n, int dtype_bytes, double* out) {
    return api_call([&]() -> int {
        if (!a || !b || !out || n < 0 || (dtype_bytes != 4 && dtype_bytes != 8)) {
            set_error("lp_dot: bad argument");
            return LP_ERR_ARG;
        }
        const int nb = (int)std::min<long long>(1024, std::max<long long>(1, (n + 4 * kB - 1) / (4 * kB)));
        Stage S;
        double* dp = S.make<double>(nb);
        if (dtype_bytes == 8)
            hipLaunchKernelGGL(k_dot_partials<double>, dim3(nb), dim3(kB), 0, nullptr,
                               S.up(static_cast<const double*>(a), (size_t)n),
                               S.up(static_cast<const double*>(b), (size_t)n), (size_t)n, dp);
        else
            hipLaunchKernelGGL(k_dot_partials<float>, dim3(nb), dim3(kB), 0, nullptr,
                               S.up(static_cast<const float*>(a), (size_t)n),
                               S.up(static_cast<const float*>(b), (size_t)n), (size_t)n, dp);
        LP_HIP(hipGetLastError());
        std::vector<double> h(nb);
        LP_HIP(hipMemcpy(h.data(), dp, nb * 8, hipMemcpyDeviceToHost));
        double t = 0.0;
        for (double v : h) t += v;
        *out = t;
        return LP_OK;
    });
}

int lp_cast(const void* a, int a_bytes, void* out, int out_bytes, long long n) {
    return api_call([&]() -> int {
        if (!a || !out || n < 0 || (a_bytes != 4 && a_bytes != 8) || (out_bytes != 4 && out_bytes != 8)) {
            set_error("lp_cast: bad argument");
            return LP_ERR_ARG;
        }
        Stage S;
        const unsigned g = std::max(1u, std::min(2048u, nb_of((size_t)n)));
        void* dout = out_bytes == 8 ? (void*)S.make<double>((size_t)n) : (void*)S.make<float>((size_t)n);
        if (a_bytes == 8 && out_bytes == 4)
            hipLaunchKernelGGL((k_cast_plain<double, float>), dim3(g), dim3(kB), 0, nullptr,
                               S.up(static_cast<const double*>(a), (size_t)n), (float*)dout, (size_t)n);
        else if (a_bytes == 4 && out_bytes == 8)
            hipLaunchKernelGGL((k_cast_plain<float, double>), dim3(g), dim3(kB), 0, nullptr,
                               S.up(static_cast<const float*>(a), (size_t)n), (double*)dout, (size_t)n);
        else if (a_bytes == 8)
            hipLaunchKernelGGL((k_cast_plain<double, double>), dim3(g), dim3(kB), 0, nullptr,
                               S.up(static_cast<const double*>(a), (size_t)n), (double*)dout, (size_t)n);
        else
            hipLaunchKernelGGL((k_cast_plain<float, float>), dim3(g), dim3(kB), 0, nullptr,
                               S.up(static_cast<const float*>(a), (size_t)n), (float*)dout, (size_t)n);
        LP_HIP(hipGetLastError());
        LP_HIP(hipMemcpy(out, dout, (size_t)n * out_bytes, hipMemcpyDeviceToHost));
        return LP_OK;
    });
}

int lp_element_energy(const double* u, const double* v, const long long* edof, long long ne, long long ndof,
                      const double* ke, double* out) {
    return api_call([&]() -> int {
        if (!u || !edof || !ke || !out || ne < 0 || ndof < 1) {
            set_error("lp_element_energy: bad argument");
            return LP_ERR_ARG;
        }
        Stage S;
        const double* du = S.up(u, (size_t)ndof);
        const double* dv = v ? S.up(v, (size_t)ndof) : du;
        const long long* de = S.up(edof, (size_t)ne * 24);
        const double* dk = S.up(ke, 576);
        double* dout = S.make<double>((size_t)ne);
        if (ne > 0) {
            hipLaunchKernelGGL(k_element_energy, dim3(nb_of((size_t)ne)), dim3(kB), 0, nullptr, du, dv, de, ne, dk,
                               dout);
            LP_HIP(hipGetLastError());
            LP_HIP(hipMemcpy(out, dout, (size_t)ne * 8, hipMemcpyDeviceToHost));
        }
        return LP_OK;
    });
}

}  // extern "C"
