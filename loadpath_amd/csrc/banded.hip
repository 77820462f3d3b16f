// The assembled route's direct solve on the device: what direct.DirectSolver
// (direct.py:158-246) and fem.Solver's rescue (fem.py:586-792) do on the
// host with a sparse Cholesky, for the stiffness of lp_stiffness_* (the
// reference's CSR layout).
//
// A voxel stiffness whose nodes are numbered with the longest box axis
// slowest has a narrow band: a dof couples to dofs at most w = 3 (Nm Nf + Nf
// + 1) + 2 rows away (Nm, Nf the node counts of the two faster axes). Cut the
// permuted matrix into b x b blocks, b >= w rounded up to 64; it is then block
// tridiagonal, and its Cholesky factor is block lower bidiagonal with no fill
// outside those blocks:
//
//   L_00 L_00^T = A_00
//   L_i,i-1 = A_i,i-1 L_i-1,i-1^-T,   L_ii L_ii^T = A_ii - L_i,i-1 L_i,i-1^T
//
// all dense FP64 in HBM (2 n b doubles: 27 GB for a 64 x 40 x 40 box). Each
// diagonal block is factored right-looking in 64-wide panels (the 64 x 64
// diagonal factor and its inverse in LDS, one workgroup: launch_potrf_diag),
// the triangular solves are 64-panel sweeps with those inverses, and every
// update is a plain GEMM (rocBLAS dgemm). The solve permutes the right-hand
// sides into the band order, sweeps forward and back through the blocks, and
// permutes back; iterative refinement against the CSR matrix itself (one
// SpMV and one solve per step) returns the relative residual.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "lp_internal.h"

namespace lp {
namespace {

constexpr int P = 64;   // panel width (launch_potrf_diag's block)

#define LP_RB(call)                                                                         \
    do {                                                                                    \
        const rocblas_status st_ = (call);                                                  \
        if (st_ != rocblas_status_success)                                                  \
            throw std::runtime_error(std::string("rocBLAS error ") + rocblas_status_to_string(st_)); \
    } while (0)

void gemm_rm(rocblas_handle h, bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int lda,
             const double* B, int ldb, double beta, double* C, int ldc) {
    blas_gemm_rm(h, ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

// the lower triangle of each (permuted) CSR row into the diagonal block or
// the block below the diagonal; one thread per row
__global__ void k_band_scatter(long long n, const long long* __restrict__ indptr, const long long* __restrict__ indices,
                               const double* __restrict__ data, const long long* __restrict__ inv, int b,
                               double* __restrict__ D, double* __restrict__ S) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const long long pr = inv[r];
    const long long bi = pr / b, ri = pr - bi * b;
    for (long long j = indptr[r]; j < indptr[r + 1]; ++j) {
        const long long pc = inv[indices[j]];
        if (pc > pr) continue;
        const long long bj = pc / b, cj = pc - bj * b;
        const size_t o = (size_t)ri * b + (size_t)cj;
        if (bj == bi)
            D[(size_t)bi * b * b + o] = data[j];
        else   // bj == bi - 1 (the band fits the blocks: checked on the host)
            S[(size_t)bj * b * b + o] = data[j];
    }
}

// identity rows for the padding past n
__global__ void k_band_pad(long long n, long long npad, int b, double* __restrict__ D) {
    const long long r = n + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= npad) return;
    const long long bi = r / b, ri = r - bi * b;
    D[(size_t)bi * b * b + (size_t)ri * b + ri] = 1.0;
}

// V (npad x nrhs, row-major) = F (nrhs x n) in band order; padding 0
__global__ void k_band_gather(long long n, long long npad, int nrhs, const long long* __restrict__ perm,
                              const double* __restrict__ f, double* __restrict__ v) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npad * nrhs) return;
    const long long row = t / nrhs;
    const int k = (int)(t - row * nrhs);
    v[t] = row < n ? f[(size_t)k * n + perm[row]] : 0.0;
}

// X (nrhs x n) = V back in mesh order; add: X += V
__global__ void k_band_scatter_x(long long n, int nrhs, const long long* __restrict__ perm, const double* __restrict__ v,
                                 double* __restrict__ x, int add) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nrhs) return;
    const long long row = t / nrhs;
    const int k = (int)(t - row * nrhs);
    const size_t o = (size_t)k * n + perm[row];
    x[o] = add ? x[o] + v[t] : v[t];
}

// R = F - K X (mesh order, nrhs x n) and per-block partial sums of R^2 and
// F^2 for each right-hand side (fixed order: summed on the host)
__global__ void __launch_bounds__(256) k_band_residual(long long n, int nrhs, const long long* __restrict__ indptr,
                                                        const long long* __restrict__ indices,
                                                        const double* __restrict__ data, const double* __restrict__ f,
                                                        const double* __restrict__ x, double* __restrict__ r,
                                                        double* __restrict__ parts) {
    __shared__ double sr[256], sf[256];
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    double rr = 0.0, ff = 0.0;
    if (row < n) {
        const double* xk = x + (size_t)k * n;
        double acc = 0.0;
        for (long long j = indptr[row]; j < indptr[row + 1]; ++j) acc += data[j] * xk[indices[j]];
        const double fv = f[(size_t)k * n + row];
        const double rv = fv - acc;
        r[(size_t)k * n + row] = rv;
        rr = rv * rv;
        ff = fv * fv;
    }
    sr[threadIdx.x] = rr;
    sf[threadIdx.x] = ff;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sr[threadIdx.x] += sr[threadIdx.x + o];
            sf[threadIdx.x] += sf[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        parts[((size_t)k * gridDim.x + blockIdx.x) * 2] = sr[0];
        parts[((size_t)k * gridDim.x + blockIdx.x) * 2 + 1] = sf[0];
    }
}

inline unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace
}  // namespace lp

using namespace lp;

struct lp_band {
    lp_stiff* s = nullptr;
    long long n = 0, nnz = 0, band = 0, npad = 0;
    int b = 0, nb = 0;
    size_t ncell = 0;
    DBuf perm, inv, indptr, indices, data, scale;
    DBuf D, S, Dinv, tmp, fail;
    DBuf v, f, x, r, parts;
    rocblas_handle h = nullptr;
    hipStream_t st = nullptr;
    bool factored = false;
    ~lp_band() {
        if (h) rocblas_destroy_handle(h);
        if (st) hipStreamDestroy(st);
    }
};

namespace {

// L_ii in place (lower) with its 64 x 64 diagonal inverses
void factor_block(lp_band& B, double* Di, double* Dinv_i) {
    const int b = B.b;
    for (int p = 0; p < b / P; ++p) {
        const int k = P * p;
        launch_potrf_diag(Di, b, k, P, Dinv_i + (size_t)p * P * P, B.fail.as<int>(), B.st);
        const int rest = b - k - P;
        if (rest > 0) {
            double* a21 = Di + (size_t)(k + P) * b + k;
            // L21 = A21 inv(L11)^T
            gemm_rm(B.h, false, true, rest, P, P, 1.0, a21, b, Dinv_i + (size_t)p * P * P, P, 0.0, B.tmp.as<double>(),
                    P);
            LP_HIP(hipMemcpy2DAsync(a21, (size_t)b * 8, B.tmp.p, P * 8, P * 8, rest, hipMemcpyDeviceToDevice, B.st));
            // A22 -= L21 L21^T
            gemm_rm(B.h, false, true, rest, rest, P, -1.0, a21, b, a21, b, 1.0, Di + (size_t)(k + P) * b + (k + P), b);
        }
    }
}

// S := S L^-T for the lower factor L of block i (X L^T = S by column panels)
void trsm_right_lt(lp_band& B, double* S, const double* Li, const double* Dinv_i) {
    const int b = B.b;
    for (int p = 0; p < b / P; ++p) {
        const int k = P * p;
        gemm_rm(B.h, false, true, b, P, P, 1.0, S + k, b, Dinv_i + (size_t)p * P * P, P, 0.0, B.tmp.as<double>(), P);
        LP_HIP(hipMemcpy2DAsync(S + k, (size_t)b * 8, B.tmp.p, P * 8, P * 8, b, hipMemcpyDeviceToDevice, B.st));
        const int rest = b - k - P;
        if (rest > 0)
            gemm_rm(B.h, false, true, b, rest, P, -1.0, S + k, b, Li + (size_t)(k + P) * b + k, b, 1.0, S + k + P, b);
    }
}

// V_i := L_ii^-1 V_i (forward, 64-row panels)
void lower_solve(lp_band& B, const double* Li, const double* Dinv_i, double* Vi, int nrhs) {
    const int b = B.b;
    for (int p = 0; p < b / P; ++p) {
        const int k = P * p;
        gemm_rm(B.h, false, false, P, nrhs, P, 1.0, Dinv_i + (size_t)p * P * P, P, Vi + (size_t)k * nrhs, nrhs, 0.0,
                B.tmp.as<double>(), nrhs);
        LP_HIP(hipMemcpyAsync(Vi + (size_t)k * nrhs, B.tmp.p, (size_t)P * nrhs * 8, hipMemcpyDeviceToDevice, B.st));
        const int rest = b - k - P;
        if (rest > 0)
            gemm_rm(B.h, false, false, rest, nrhs, P, -1.0, Li + (size_t)(k + P) * b + k, b, Vi + (size_t)k * nrhs,
                    nrhs, 1.0, Vi + (size_t)(k + P) * nrhs, nrhs);
    }
}

// V_i := L_ii^-T V_i (backward)
void upper_solve(lp_band& B, const double* Li, const double* Dinv_i, double* Vi, int nrhs) {
    const int b = B.b;
    for (int p = b / P - 1; p >= 0; --p) {
        const int k = P * p;
        const int rest = b - k - P;
        if (rest > 0)
            gemm_rm(B.h, true, false, P, nrhs, rest, -1.0, Li + (size_t)(k + P) * b + k, b, Vi + (size_t)(k + P) * nrhs,
                    nrhs, 1.0, Vi + (size_t)k * nrhs, nrhs);
        gemm_rm(B.h, true, false, P, nrhs, P, 1.0, Dinv_i + (size_t)p * P * P, P, Vi + (size_t)k * nrhs, nrhs, 0.0,
                B.tmp.as<double>(), nrhs);
        LP_HIP(hipMemcpyAsync(Vi + (size_t)k * nrhs, B.tmp.p, (size_t)P * nrhs * 8, hipMemcpyDeviceToDevice, B.st));
    }
}

int factor_data(lp_band& B) {
    const int b = B.b, nb = B.nb;
    const size_t bb = (size_t)b * b;
    B.factored = false;
    LP_HIP(hipMemsetAsync(B.D.p, 0, (size_t)nb * bb * 8, B.st));
    if (nb > 1) LP_HIP(hipMemsetAsync(B.S.p, 0, (size_t)(nb - 1) * bb * 8, B.st));
    LP_HIP(hipMemsetAsync(B.fail.p, 0, sizeof(int), B.st));
    hipLaunchKernelGGL(k_band_scatter, dim3(blocks_for(B.n)), dim3(256), 0, B.st, B.n, B.indptr.as<long long>(),
                       B.indices.as<long long>(), B.data.as<double>(), B.inv.as<long long>(), b, B.D.as<double>(),
                       nb > 1 ? B.S.as<double>() : nullptr);
    LP_HIP(hipGetLastError());
    if (B.npad > B.n) {
        hipLaunchKernelGGL(k_band_pad, dim3(blocks_for(B.npad - B.n)), dim3(256), 0, B.st, B.n, B.npad, b,
                           B.D.as<double>());
        LP_HIP(hipGetLastError());
    }
    double* D = B.D.as<double>();
    double* S = B.S.as<double>();
    double* Dinv = B.Dinv.as<double>();
    for (int i = 0; i < nb; ++i) {
        double* Di = D + (size_t)i * bb;
        if (i > 0) {
            const double* Sp = S + (size_t)(i - 1) * bb;
            gemm_rm(B.h, false, true, b, b, b, -1.0, Sp, b, Sp, b, 1.0, Di, b);
        }
        factor_block(B, Di, Dinv + (size_t)i * b * P);
        if (i + 1 < nb) trsm_right_lt(B, S + (size_t)i * bb, Di, Dinv + (size_t)i * b * P);
    }
    int hf = 0;
    LP_HIP(hipMemcpyAsync(&hf, B.fail.p, sizeof(int), hipMemcpyDeviceToHost, B.st));
    LP_HIP(hipStreamSynchronize(B.st));
    if (hf) {
        set_error("band Cholesky: a non-positive pivot (the structure may be a mechanism: check supports)");
        return LP_SINGULAR;
    }
    B.factored = true;
    return LP_OK;
}

// V (band order) := A^-1 V
void band_solve(lp_band& B, int nrhs) {
    const int b = B.b, nb = B.nb;
    const size_t bb = (size_t)b * b;
    const double* D = B.D.as<double>();
    const double* S = B.S.as<double>();
    const double* Dinv = B.Dinv.as<double>();
    double* V = B.v.as<double>();
    for (int i = 0; i < nb; ++i) {
        double* Vi = V + (size_t)i * b * nrhs;
        if (i > 0)
            gemm_rm(B.h, false, false, b, nrhs, b, -1.0, S + (size_t)(i - 1) * bb, b, Vi - (size_t)b * nrhs, nrhs,
                    1.0, Vi, nrhs);
        lower_solve(B, D + (size_t)i * bb, Dinv + (size_t)i * b * P, Vi, nrhs);
    }
    for (int i = nb - 1; i >= 0; --i) {
        double* Vi = V + (size_t)i * b * nrhs;
        if (i + 1 < nb)
            gemm_rm(B.h, true, false, b, nrhs, b, -1.0, S + (size_t)i * bb, b, Vi + (size_t)b * nrhs, nrhs, 1.0, Vi,
                    nrhs);
        upper_solve(B, D + (size_t)i * bb, Dinv + (size_t)i * b * P, Vi, nrhs);
    }
}

}  // namespace

extern "C" {

int lp_band_create(lp_band** out, lp_stiff* s, const long long* perm, long long* band, long long* bytes) {
    return api_call([&]() -> int {
        if (!out || !s) {
            set_error("lp_band_create: bad argument");
            return LP_ERR_ARG;
        }
        auto B = std::make_unique<lp_band>();
        B->s = s;
        if (lp_stiffness_size(s, &B->n, &B->nnz) != LP_OK || B->n < 1) {
            set_error("lp_band_create: empty stiffness");
            return LP_ERR_ARG;
        }
        const long long n = B->n;
        std::vector<long long> ip((size_t)n + 1), ix((size_t)B->nnz), pm((size_t)n), iv((size_t)n, -1);
        if (lp_stiffness_pattern(s, ip.data(), ix.data()) != LP_OK) return LP_ERR_STATE;
        for (long long i = 0; i < n; ++i) {
            const long long o = perm ? perm[i] : i;
            if (o < 0 || o >= n || iv[(size_t)o] >= 0) {
                set_error("lp_band_create: perm is not a permutation of the dofs");
                return LP_ERR_ARG;
            }
            pm[(size_t)i] = o;
            iv[(size_t)o] = i;
        }
        long long w = 0;
        for (long long r = 0; r < n; ++r)
            for (long long j = ip[(size_t)r]; j < ip[(size_t)r + 1]; ++j)
                w = std::max(w, std::abs(iv[(size_t)r] - iv[(size_t)ix[(size_t)j]]));
        B->band = w;
        const long long bsz = std::max<long long>(P, ((w + 1 + P - 1) / P) * P);
        if (bsz > 32768) {
            set_error("lp_band_create: band " + std::to_string(w) + " too wide for the dense band blocks");
            return LP_ERR_ARG;
        }
        B->b = (int)bsz;
        B->nb = (int)((n + bsz - 1) / bsz);
        B->npad = (long long)B->nb * bsz;
        const size_t bb = (size_t)bsz * bsz;
        const size_t need = ((size_t)B->nb * bb + (size_t)(B->nb - 1) * bb + (size_t)B->nb * bsz * P) * 8;
        if (bytes) *bytes = (long long)need;
        size_t fr = 0, tot = 0;
        LP_HIP(hipMemGetInfo(&fr, &tot));
        if (need + ((size_t)1 << 30) > fr) {
            set_error("lp_band_create: the band factor needs " + std::to_string(need >> 20) + " MiB, " +
                      std::to_string(fr >> 20) + " MiB free");
            return LP_ERR_ARG;
        }
        B->perm.alloc((size_t)n * 8);
        B->inv.alloc((size_t)n * 8);
        B->indptr.alloc(((size_t)n + 1) * 8);
        B->indices.alloc(std::max<size_t>((size_t)B->nnz, 1) * 8);
        B->data.alloc(std::max<size_t>((size_t)B->nnz, 1) * 8);
        LP_HIP(hipMemcpy(B->perm.p, pm.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        LP_HIP(hipMemcpy(B->inv.p, iv.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        LP_HIP(hipMemcpy(B->indptr.p, ip.data(), ((size_t)n + 1) * 8, hipMemcpyHostToDevice));
        if (B->nnz) LP_HIP(hipMemcpy(B->indices.p, ix.data(), (size_t)B->nnz * 8, hipMemcpyHostToDevice));
        B->D.alloc((size_t)B->nb * bb * 8);
        if (B->nb > 1) B->S.alloc((size_t)(B->nb - 1) * bb * 8);
        B->Dinv.alloc((size_t)B->nb * bsz * P * 8);
        B->tmp.alloc((size_t)bsz * P * 8);
        B->fail.alloc(sizeof(int));
        LP_HIP(hipStreamCreate(&B->st));
        LP_RB(rocblas_create_handle(&B->h));
        LP_RB(rocblas_set_stream(B->h, B->st));
        if (band) *band = w;
        *out = B.release();
        return LP_OK;
    });
}

int lp_band_factor(lp_band* B, const double* scale, int ncell, int device_ptr) {
    return api_call([&]() -> int {
        if (!B || !scale || ncell < 1) {
            set_error("lp_band_factor: bad argument");
            return LP_ERR_ARG;
        }
        const double* ds = scale;
        if (!device_ptr) {
            B->scale.alloc((size_t)ncell * 8);
            LP_HIP(hipMemcpy(B->scale.p, scale, (size_t)ncell * 8, hipMemcpyHostToDevice));
            ds = B->scale.as<double>();
        }
        const int rc = lp_stiffness_assemble(B->s, ds, B->data.as<double>(), 1);
        if (rc != LP_OK) return rc;
        return factor_data(*B);
    });
}

int lp_band_factor_csr(lp_band* B, const double* data, int device_ptr) {
    return api_call([&]() -> int {
        if (!B || !data) {
            set_error("lp_band_factor_csr: bad argument");
            return LP_ERR_ARG;
        }
        if (B->nnz)
            LP_HIP(hipMemcpy(B->data.p, data, (size_t)B->nnz * 8,
                             device_ptr ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
        return factor_data(*B);
    });
}

int lp_band_solve(lp_band* B, int nrhs, const double* f, double* x, int device_ptr, int refine, double* relres) {
    return api_call([&]() -> int {
        if (!B || nrhs < 1 || nrhs > 64 || !f || !x || refine < 0) {
            set_error("lp_band_solve: bad argument");
            return LP_ERR_ARG;
        }
        if (!B->factored) {
            set_error("lp_band_solve: no factor (lp_band_factor first)");
            return LP_ERR_STATE;
        }
        const long long n = B->n;
        const size_t vb = (size_t)n * nrhs * 8;
        B->v.alloc((size_t)B->npad * nrhs * 8);
        B->tmp.alloc((size_t)B->b * std::max(P, nrhs) * 8);
        const double* df = f;
        double* dx = x;
        if (!device_ptr) {
            B->f.alloc(vb);
            B->x.alloc(vb);
            LP_HIP(hipMemcpyAsync(B->f.p, f, vb, hipMemcpyHostToDevice, B->st));
            df = B->f.as<double>();
            dx = B->x.as<double>();
        }
        const long long tv = B->npad * nrhs, tx = n * nrhs;
        hipLaunchKernelGGL(k_band_gather, dim3(blocks_for(tv)), dim3(256), 0, B->st, n, B->npad, nrhs,
                           B->perm.as<long long>(), df, B->v.as<double>());
        LP_HIP(hipGetLastError());
        band_solve(*B, nrhs);
        hipLaunchKernelGGL(k_band_scatter_x, dim3(blocks_for(tx)), dim3(256), 0, B->st, n, nrhs,
                           B->perm.as<long long>(), B->v.as<double>(), dx, 0);
        LP_HIP(hipGetLastError());
        // refinement steps, then the relative residual max_k |f - K x| / |f|
        B->r.alloc(vb);
        const unsigned nbk = blocks_for(n);
        B->parts.alloc((size_t)nbk * nrhs * 2 * 8);
        std::vector<double> hp((size_t)nbk * nrhs * 2);
        double worst = 0.0;
        for (int step = 0; step <= refine; ++step) {
            hipLaunchKernelGGL(k_band_residual, dim3(nbk, nrhs), dim3(256), 0, B->st, n, nrhs,
                               B->indptr.as<long long>(), B->indices.as<long long>(), B->data.as<double>(), df, dx,
                               B->r.as<double>(), B->parts.as<double>());
            LP_HIP(hipGetLastError());
            LP_HIP(hipMemcpyAsync(hp.data(), B->parts.p, hp.size() * 8, hipMemcpyDeviceToHost, B->st));
            LP_HIP(hipStreamSynchronize(B->st));
            worst = 0.0;
            for (int k = 0; k < nrhs; ++k) {
                double rr = 0.0, ff = 0.0;
                for (unsigned q = 0; q < nbk; ++q) {
                    rr += hp[((size_t)k * nbk + q) * 2];
                    ff += hp[((size_t)k * nbk + q) * 2 + 1];
                }
                worst = std::max(worst, ff > 0.0 ? std::sqrt(rr / ff) : std::sqrt(rr));
            }
            if (step == refine) break;
            hipLaunchKernelGGL(k_band_gather, dim3(blocks_for(tv)), dim3(256), 0, B->st, n, B->npad, nrhs,
                               B->perm.as<long long>(), B->r.as<double>(), B->v.as<double>());
            LP_HIP(hipGetLastError());
            band_solve(*B, nrhs);
            hipLaunchKernelGGL(k_band_scatter_x, dim3(blocks_for(tx)), dim3(256), 0, B->st, n, nrhs,
                               B->perm.as<long long>(), B->v.as<double>(), dx, 1);
            LP_HIP(hipGetLastError());
        }
        if (!device_ptr) LP_HIP(hipMemcpyAsync(x, dx, vb, hipMemcpyDeviceToHost, B->st));
        LP_HIP(hipStreamSynchronize(B->st));
        if (relres) *relres = worst;
        return LP_OK;
    });
}

int lp_band_destroy(lp_band* B) {
    delete B;
    return LP_OK;
}

}  // extern "C"
