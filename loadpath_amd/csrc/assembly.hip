// The assembled route's global stiffness: fem.Stiffness (fem.py:377-526) on a
// voxel mesh, its CSR layout built once on the host and its values for
// per-cell moduli assembled on the device (femops.scatter_gather,
// femops.py:99-278).
//
// Layout (the reference's): three rows per mesh node, mesh nodes numbered
// in C order over the node grid; row 3n+i holds, for each of the node's
// neighbours (the 27 offsets in fem.OFFSETS order, kept where a cell joins
// the two), the three columns 3m+j, so the columns of a row ascend.
// Dirichlet conditions are identity rows: every slot in a fixed row or a
// fixed column is 0 and a fixed diagonal 1, so the pattern never changes.
//
// Assembly: one thread per (node, neighbour) 3x3 block, summing
// scale_c Ke[a][b] over the cells c that hold the node as corner a and the
// neighbour as corner b, corners in ascending order (the order the
// reference's scatter adds them), with the products and sums rounded
// separately (no fused multiply-add), so each value is the reference's to
// the bit. No atomics: each slot is written once.
#include <algorithm>
#include <memory>
#include <vector>

#include "lp_internal.h"

namespace lp {
namespace {

struct StiffArgs {
    int nx, ny, nz;
    long long npairs;
    const int* pair_node;         // grid node index
    const int* pair_off;          // 0..26 (fem.OFFSETS order)
    const long long* pair_base;   // data index of the block's (row 3n, first column)
    const long long* pair_row;    // row length (3 x the node's neighbours)
    const uint8_t* present;       // (cells)
    const uint8_t* fixed;         // (grid nodes, 3)
    const double* ke;             // 24 x 24, pitch included
    const double* scale;          // (cells)
    double* data;
};

__global__ void __launch_bounds__(256) k_stiff_assemble(StiffArgs A) {
    // HIP contracts a * b + c into one fused multiply-add by default; the
    // reference rounds the product and the sum separately
#pragma clang fp contract(off)
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.npairs) return;
    const size_t Ny = (size_t)A.ny + 1, Nz = (size_t)A.nz + 1;
    const size_t n = (size_t)A.pair_node[p];
    const int o = A.pair_off[p];
    const int ox = o / 9 - 1, oy = (o / 3) % 3 - 1, oz = o % 3 - 1;
    const int X = (int)(n / (Ny * Nz)), Y = (int)((n / Nz) % Ny), Z = (int)(n % Nz);
    double blk[9];
    for (int t = 0; t < 9; ++t) blk[t] = 0.0;
    for (int a = 0; a < 8; ++a) {
        const int ax = a & 1, ay = (a >> 1) & 1, az = a >> 2;
        const int cx = X - ax, cy = Y - ay, cz = Z - az;
        if (cx < 0 || cy < 0 || cz < 0 || cx >= A.nx || cy >= A.ny || cz >= A.nz) continue;
        const int bx = ax + ox, by = ay + oy, bz = az + oz;
        if (bx < 0 || by < 0 || bz < 0 || bx > 1 || by > 1 || bz > 1) continue;
        const size_t c = ((size_t)cx * A.ny + cy) * A.nz + cz;
        if (!A.present[c]) continue;
        const int b = bx | (by << 1) | (bz << 2);
        const double s = A.scale[c];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                blk[3 * i + j] += s * A.ke[(3 * a + i) * 24 + 3 * b + j];
    }
    const size_t m = n + (long long)ox * (long long)(Ny * Nz) + (long long)oy * (long long)Nz + oz;
    const long long base = A.pair_base[p], row = A.pair_row[p];
    for (int i = 0; i < 3; ++i) {
        const bool rf = A.fixed[3 * n + i] != 0;
        for (int j = 0; j < 3; ++j) {
            double v = (rf || A.fixed[3 * m + j]) ? 0.0 : blk[3 * i + j];
            if (rf && o == 13 && i == j) v = 1.0;
            A.data[base + i * row + j] = v;
        }
    }
}

}  // namespace
}  // namespace lp

using namespace lp;

struct lp_stiff {
    int nx = 0, ny = 0, nz = 0;
    long long n_dof = 0, nnz = 0, npairs = 0;
    std::vector<long long> indptr, indices;
    DBuf pair_node, pair_off, pair_base, pair_row, present, fixed, ke, scale, data;
};

extern "C" {

int lp_stiffness_create(lp_stiff** out, int nx, int ny, int nz, const uint8_t* present, const uint8_t* fixed,
                        const double* ke, long long* n_dof, long long* nnz) {
    return api_call([&]() -> int {
        if (!out || nx < 1 || ny < 1 || nz < 1 || !present || !fixed || !ke || !n_dof || !nnz ||
            (size_t)(nx + 1) * (ny + 1) * (nz + 1) >= (size_t)0x7fffffff) {
            set_error("lp_stiffness_create: bad argument");
            return LP_ERR_ARG;
        }
        auto S = std::make_unique<lp_stiff>();
        S->nx = nx;
        S->ny = ny;
        S->nz = nz;
        const size_t nc = (size_t)nx * ny * nz;
        const size_t Ny = (size_t)ny + 1, Nz = (size_t)nz + 1, nn = ((size_t)nx + 1) * Ny * Nz;
        // neighbour offsets of each node that a present cell joins (bit o of adj)
        std::vector<uint32_t> adj(nn, 0u);
        for (size_t c = 0; c < nc; ++c) {
            if (!present[c]) continue;
            const int X = (int)(c / ((size_t)ny * nz)), Y = (int)((c / nz) % ny), Z = (int)(c % nz);
            for (int a = 0; a < 8; ++a) {
                const size_t na = ((size_t)(X + (a & 1)) * Ny + (Y + ((a >> 1) & 1))) * Nz + (Z + (a >> 2));
                for (int b = 0; b < 8; ++b) {
                    const int o = ((b & 1) - (a & 1) + 1) * 9 + (((b >> 1) & 1) - ((a >> 1) & 1) + 1) * 3 +
                                  ((b >> 2) - (a >> 2) + 1);
                    adj[na] |= 1u << o;
                }
            }
        }
        // mesh node ids: the used grid nodes in C order
        std::vector<long long> id(nn, -1);
        long long used = 0;
        for (size_t g = 0; g < nn; ++g)
            if (adj[g]) id[g] = used++;
        S->n_dof = 3 * used;
        S->indptr.assign((size_t)S->n_dof + 1, 0);
        long long npairs = 0;
        for (size_t g = 0; g < nn; ++g) {
            if (!adj[g]) continue;
            const long long cnt = __builtin_popcount(adj[g]);
            npairs += cnt;
            for (int i = 0; i < 3; ++i) S->indptr[3 * id[g] + i + 1] = 3 * cnt;
        }
        for (long long r = 0; r < S->n_dof; ++r) S->indptr[r + 1] += S->indptr[r];
        S->nnz = S->indptr[S->n_dof];
        S->npairs = npairs;
        S->indices.assign((size_t)S->nnz, 0);
        std::vector<int> pnode((size_t)npairs), poff((size_t)npairs);
        std::vector<long long> pbase((size_t)npairs), prow((size_t)npairs);
        long long q = 0;
        for (size_t g = 0; g < nn; ++g) {
            if (!adj[g]) continue;
            const long long k = id[g];
            const long long row = 3LL * __builtin_popcount(adj[g]);
            const int X = (int)(g / (Ny * Nz)), Y = (int)((g / Nz) % Ny), Z = (int)(g % Nz);
            int rank = 0;
            for (int o = 0; o < 27; ++o) {
                if (!((adj[g] >> o) & 1u)) continue;
                const size_t m = ((size_t)(X + o / 9 - 1) * Ny + (Y + (o / 3) % 3 - 1)) * Nz + (Z + o % 3 - 1);
                const long long base = S->indptr[3 * k] + 3 * rank;
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) S->indices[(size_t)(base + i * row + j)] = 3 * id[m] + j;
                pnode[(size_t)q] = (int)g;
                poff[(size_t)q] = o;
                pbase[(size_t)q] = base;
                prow[(size_t)q] = row;
                ++q;
                ++rank;
            }
        }
        S->pair_node.alloc(std::max<size_t>((size_t)npairs, 1) * 4);
        S->pair_off.alloc(std::max<size_t>((size_t)npairs, 1) * 4);
        S->pair_base.alloc(std::max<size_t>((size_t)npairs, 1) * 8);
        S->pair_row.alloc(std::max<size_t>((size_t)npairs, 1) * 8);
        S->present.alloc(nc);
        S->fixed.alloc(nn * 3);
        S->ke.alloc(576 * 8);
        S->scale.alloc(nc * 8);
        S->data.alloc(std::max<size_t>((size_t)S->nnz, 1) * 8);
        if (npairs) {
            LP_HIP(hipMemcpy(S->pair_node.p, pnode.data(), (size_t)npairs * 4, hipMemcpyHostToDevice));
            LP_HIP(hipMemcpy(S->pair_off.p, poff.data(), (size_t)npairs * 4, hipMemcpyHostToDevice));
            LP_HIP(hipMemcpy(S->pair_base.p, pbase.data(), (size_t)npairs * 8, hipMemcpyHostToDevice));
            LP_HIP(hipMemcpy(S->pair_row.p, prow.data(), (size_t)npairs * 8, hipMemcpyHostToDevice));
        }
        LP_HIP(hipMemcpy(S->present.p, present, nc, hipMemcpyHostToDevice));
        LP_HIP(hipMemcpy(S->fixed.p, fixed, nn * 3, hipMemcpyHostToDevice));
        LP_HIP(hipMemcpy(S->ke.p, ke, 576 * 8, hipMemcpyHostToDevice));
        *n_dof = S->n_dof;
        *nnz = S->nnz;
        *out = S.release();
        return LP_OK;
    });
}

int lp_stiffness_pattern(lp_stiff* S, long long* indptr, long long* indices) {
    return api_call([&]() -> int {
        if (!S || !indptr || !indices) {
            set_error("lp_stiffness_pattern: bad argument");
            return LP_ERR_ARG;
        }
        std::copy(S->indptr.begin(), S->indptr.end(), indptr);
        std::copy(S->indices.begin(), S->indices.end(), indices);
        return LP_OK;
    });
}

int lp_stiffness_assemble(lp_stiff* S, const double* scale, double* data, int device_ptr) {
    return api_call([&]() -> int {
        if (!S || !scale || !data) {
            set_error("lp_stiffness_assemble: bad argument");
            return LP_ERR_ARG;
        }
        const size_t nc = (size_t)S->nx * S->ny * S->nz;
        const double* dscale = scale;
        if (!device_ptr) {
            LP_HIP(hipMemcpy(S->scale.p, scale, nc * 8, hipMemcpyHostToDevice));
            dscale = S->scale.as<double>();
        }
        double* ddata = device_ptr ? data : S->data.as<double>();
        StiffArgs A{};
        A.nx = S->nx;
        A.ny = S->ny;
        A.nz = S->nz;
        A.npairs = S->npairs;
        A.pair_node = S->pair_node.as<int>();
        A.pair_off = S->pair_off.as<int>();
        A.pair_base = S->pair_base.as<long long>();
        A.pair_row = S->pair_row.as<long long>();
        A.present = S->present.as<uint8_t>();
        A.fixed = S->fixed.as<uint8_t>();
        A.ke = S->ke.as<double>();
        A.scale = dscale;
        A.data = ddata;
        if (S->npairs) {
            hipLaunchKernelGGL(k_stiff_assemble, dim3((unsigned)((S->npairs + 255) / 256)), dim3(256), 0, nullptr,
                               A);
            LP_HIP(hipGetLastError());
        }
        if (device_ptr) {
            LP_HIP(hipDeviceSynchronize());
        } else if (S->nnz) {
            LP_HIP(hipMemcpy(data, ddata, (size_t)S->nnz * 8, hipMemcpyDeviceToHost));
        }
        return LP_OK;
    });
}

int lp_stiffness_size(lp_stiff* S, long long* n_dof, long long* nnz) {
    if (!S || !n_dof || !nnz) return LP_ERR_ARG;
    *n_dof = S->n_dof;
    *nnz = S->nnz;
    return LP_OK;
}

int lp_stiffness_destroy(lp_stiff* S) {
    delete S;
    return LP_OK;
}

}  // extern "C"
