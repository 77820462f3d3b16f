// Coarsest-level FP64 solve: blocked Cholesky and explicit inverse.
//
// Reference: mg._build_coarse / _coarse_solve (mg.py:708-766), which factor
// the assembled coarsest operator with SuperLU on the host and solve there on
// every V-cycle (gpu.py:366-368 copies the right-hand side to the host). Here
// the factorization runs on the device each set_moduli, and the V-cycle's
// coarsest solve becomes one dense FP64 matvec with the inverse (no host
// round trip; a few thousand dofs).
//
// A = L L^T (right-looking, 64-wide panels): a one-wave kernel factors the
// diagonal block, a forward substitution per row solves the panel below it in
// place, the trailing update is a GEMM. Then every diagonal block's inverse
// at once, Linv by block rows (two GEMMs each) and A^-1 = Linv^T Linv.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <stdexcept>

#include "lp_internal.h"

namespace lp {

namespace {

constexpr int NB = 64;

// inv(L) of the lower-triangular NB x NB block in LDS (identity past its
// size), column c per lane: inv[i][c] = (delta_ic - sum_{m < i} L[i][m]
// inv[m][c]) / L[i][i], L's rows read as LDS broadcasts (four partial sums
// over m mod 4, so a row's terms are not one dependent chain)
__device__ __forceinline__ void invert_lower(const double (*L)[NB + 1], int c, double (&v)[NB]) {
#pragma unroll
    for (int row = 0; row < NB; ++row) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < row; ++m) acc[m & 3] = fma(L[row][m], v[m], acc[m & 3]);
        const double sum = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        v[row] = c <= row ? ((row == c ? 1.0 : 0.0) - sum) / L[row][row] : 0.0;
    }
}

// Factor the kb x kb diagonal block at (k, k) in place (lower); with INV,
// write inv(L_kk) to dinv (NB x NB, row-major). fail set on a non-positive
// pivot.
//
// One wave. Lane i holds row i of the block in registers (rows and columns
// past kb are the identity's, so the padded block factors to the identity).
// Column j's step (right-looking): lane j posts the pivot to LDS, every lane
// below it divides its entry by the pivot's root and posts the multiplier
// L_ij (0 above the diagonal), and each lane updates its row's entries
// m > j by L_ij L_mj with L_mj read back as an LDS broadcast (j and m
// compile-time: both loops unrolled; the entries past the diagonal are
// updated too, never read and zeroed at the end, so no lane masks). Then
// L goes through LDS to memory (and to invert_lower).
template <bool INV>
__global__ void __launch_bounds__(64) k_potrf_diag(double* a, int n, int k, int kb, double* dinv, int* fail) {
    __shared__ double L[NB][NB + 1];
    const int i = (int)threadIdx.x;
    // lane i's row: every load issued before the first is waited for
    // (indices clamped into the block, the padding selected afterwards)
    double r[NB];
    {
        const double* arow = a + (size_t)(k + min(i, kb - 1)) * n + k;
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            const double v = arow[min(m, kb - 1)];
            r[m] = i < kb && m < kb ? v : (i == m ? 1.0 : 0.0);
        }
    }
    __shared__ double col[NB + 1];
    bool bad = false;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (i == j) col[NB] = r[j];
        __syncthreads();
        const double djj = col[NB];
        const bool ok = djj > 0.0;
        bad |= !ok;
        const double piv = ok ? sqrt(djj) : 1.0;
        const double lij = i > j ? r[j] / piv : 0.0;
        col[i] = lij;
        r[j] = i > j ? lij : (i == j ? piv : r[j]);
        __syncthreads();
#pragma unroll
        for (int m = j + 1; m < NB; ++m) r[m] = fma(-lij, col[m], r[m]);
    }
    if (bad && i == 0) *fail = 1;
#pragma unroll
    for (int m = 0; m < NB; ++m) L[i][m] = m <= i ? r[m] : 0.0;
    __syncthreads();
#pragma unroll 8
    for (int e = i; e < kb * kb; e += 64) {
        const int row = e / kb, col = e - (e / kb) * kb;
        a[(size_t)(k + row) * n + k + col] = L[row][col];
    }
    if (INV) {
        double v[NB];
        invert_lower(L, i, v);
#pragma unroll
        for (int row = 0; row < NB; ++row) dinv[row * NB + i] = v[row];
    }
}

// L21 = A21 inv(L11)^T in place for the rest rows below the factored
// diagonal block at (k, k) (kb = NB): thread = one row x of the panel,
// forward substitution x_j = (a_j - sum_{m < j} x_m L_jm) / L_jj with L11 read
// as LDS broadcasts (four partial sums over m mod 4)
__global__ void __launch_bounds__(64) k_trsm_panel(double* a, int n, int k, int rest) {
    __shared__ double L[NB][NB + 1];
#pragma unroll 16
    for (int e = (int)threadIdx.x; e < NB * NB; e += 64) {
        const int row = e / NB, col = e - (e / NB) * NB;
        L[row][col] = a[(size_t)(k + row) * n + k + col];
    }
    __syncthreads();
    const int r = (int)(blockIdx.x * 64 + threadIdx.x);
    double* arow = a + (size_t)(k + NB + min(r, rest - 1)) * n + k;
    double x[NB];
#pragma unroll
    for (int m = 0; m < NB; ++m) x[m] = arow[m];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < j; ++m) acc[m & 3] = fma(x[m], L[j][m], acc[m & 3]);
        x[j] = (x[j] - ((acc[0] + acc[1]) + (acc[2] + acc[3]))) / L[j][j];
    }
    if (r < rest)
#pragma unroll
        for (int m = 0; m < NB; ++m) arow[m] = x[m];
}

// inv(L_bb) of every factored diagonal block b (grid = the blocks, one wave
// each) into dinv (NB x NB per block) and into the diagonal blocks of linv
// (n x n)
__global__ void __launch_bounds__(64) k_trtri_blocks(const double* a, int n, double* dinv, double* linv) {
    __shared__ double L[NB][NB + 1];
    const int b = (int)blockIdx.x, k = b * NB, kb = min(NB, n - k);
#pragma unroll 16
    for (int e = (int)threadIdx.x; e < NB * NB; e += 64) {
        const int row = e / NB, col = e - (e / NB) * NB;
        L[row][col] = row < kb && col < kb ? (col <= row ? a[(size_t)(k + row) * n + k + col] : 0.0)
                                           : (row == col ? 1.0 : 0.0);
    }
    __syncthreads();
    const int c = (int)threadIdx.x;
    double v[NB];
    invert_lower(L, c, v);
    double* Db = dinv + (size_t)b * NB * NB;
#pragma unroll
    for (int row = 0; row < NB; ++row) {
        Db[row * NB + c] = v[row];
        if (row < kb && c < kb) linv[(size_t)(k + row) * n + k + c] = v[row];
    }
}

__global__ void k_zero_upper(double* a, int n) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    if (j > i) a[e] = 0.0;
}

void gemm(void* h, int M, int N, int K, double alpha, const double* A, int lda, int ta, const double* B, int ldb,
          int tb, double beta, double* C, int ldc) {
    blas_gemm_rm(h, ta != 0, tb != 0, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
}

// y (n, nrhs stride) = Ainv b, FP64 accumulate; one wave per row
template <typename T>
__global__ void k_matvec(const double* __restrict__ ainv, int n, const T* __restrict__ b, T* __restrict__ x,
                         size_t stride) {
    const int row = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const int lane = threadIdx.x & 63;
    const int rhs = blockIdx.y;
    if (row >= n) return;
    const double* ar = ainv + (size_t)row * n;
    const T* bb = b + rhs * stride;
    double acc = 0;
    for (int j = lane; j < n; j += 64) acc += ar[j] * (double)bb[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (lane == 0) x[rhs * stride + row] = (T)acc;
}

}  // namespace

Scratch::~Scratch() { blas_release(blas); }

void blas_gemm_rm(void* h, bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int lda,
                  const double* B, int ldb, double beta, double* C, int ldc) {
    if (M <= 0 || N <= 0) return;
    // row-major C = op(A) op(B) is the column-major C^T = op(B)^T op(A)^T
    const rocblas_status st = rocblas_dgemm(static_cast<rocblas_handle>(h),
                                            tb ? rocblas_operation_transpose : rocblas_operation_none,
                                            ta ? rocblas_operation_transpose : rocblas_operation_none, N, M, K, &alpha,
                                            B, ldb, A, lda, &beta, C, ldc);
    if (st != rocblas_status_success)
        throw std::runtime_error(std::string("rocBLAS dgemm: ") + rocblas_status_to_string(st));
}

void* blas_handle(void*& slot, hipStream_t s) {
    if (!slot) {
        rocblas_handle h = nullptr;
        if (rocblas_create_handle(&h) != rocblas_status_success) throw std::runtime_error("rocblas_create_handle failed");
        slot = h;
    }
    if (rocblas_set_stream(static_cast<rocblas_handle>(slot), s) != rocblas_status_success)
        throw std::runtime_error("rocblas_set_stream failed");
    return slot;
}

void blas_release(void* h) {
    if (h) rocblas_destroy_handle(static_cast<rocblas_handle>(h));
}

bool dense_spd_inverse(double* a, int n, Scratch& w, hipStream_t s) {
    void* bh = blas_handle(w.blas, s);
    const int nblk = (n + NB - 1) / NB;
    DBuf& dinv = w.dn_dinv;
    DBuf& fail = w.dn_fail;
    DBuf& linv = w.dn_linv;
    DBuf& t = w.dn_t;
    dinv.alloc((size_t)nblk * NB * NB * 8);
    fail.alloc(sizeof(int));
    linv.alloc((size_t)n * n * 8);
    t.alloc((size_t)n * n * 8);
    LP_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), s));
    LP_HIP(hipMemsetAsync(linv.p, 0, (size_t)n * n * 8, s));
    double* D = dinv.as<double>();
    for (int b = 0; b < nblk; ++b) {
        const int k = b * NB, kb = std::min(NB, n - k);
        hipLaunchKernelGGL(k_potrf_diag<false>, dim3(1), dim3(64), 0, s, a, n, k, kb, (double*)nullptr,
                           fail.as<int>());
        LP_HIP(hipGetLastError());
        const int rest = n - k - kb;
        if (rest > 0) {
            // panel L21 = A21 inv(L11)^T in place (rest x NB: a block with
            // rows below it is a full one)
            if (kb != NB) throw std::logic_error("dense factor: a partial diagonal block with rows below it");
            hipLaunchKernelGGL(k_trsm_panel, dim3((unsigned)((rest + 63) / 64)), dim3(64), 0, s, a, n, k, rest);
            LP_HIP(hipGetLastError());
            // trailing A22 -= L21 L21^T
            const double* l21 = a + (size_t)(k + kb) * n + k;
            gemm(bh, rest, rest, kb, -1.0, l21, n, 0, l21, n, 1, 1.0, a + (size_t)(k + kb) * n + (k + kb), n);
        }
    }
    hipLaunchKernelGGL(k_zero_upper, dim3((unsigned)(((size_t)n * n + 255) / 256)), dim3(256), 0, s, a, n);
    LP_HIP(hipGetLastError());
    // inv(L_ii) of every diagonal block at once (into D and Linv's diagonal),
    // then Linv by block rows: Linv[i, :i] = -inv(L_ii) (L[i, :i] Linv[:i, :i])
    double* Li = linv.as<double>();
    hipLaunchKernelGGL(k_trtri_blocks, dim3((unsigned)nblk), dim3(64), 0, s, a, n, D, Li);
    LP_HIP(hipGetLastError());
    for (int b = 1; b < nblk; ++b) {
        const int k = b * NB, kb = std::min(NB, n - k);
        double* Db = D + (size_t)b * NB * NB;
        // T (kb x k) = L[k:k+kb, :k] Linv[:k, :k]
        gemm(bh, kb, k, k, 1.0, a + (size_t)k * n, n, 0, Li, n, 0, 0.0, t.as<double>(), n);
        // Linv[k:k+kb, :k] = -inv(L_kk) T
        gemm(bh, kb, k, kb, -1.0, Db, NB, 0, t.as<double>(), n, 0, 0.0, Li + (size_t)k * n, n);
    }
    // A^-1 = Linv^T Linv
    gemm(bh, n, n, n, 1.0, Li, n, 1, Li, n, 0, 0.0, a, n);
    int hf = 0;
    LP_HIP(hipMemcpyAsync(&hf, fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
    LP_HIP(hipStreamSynchronize(s));
    return hf == 0;
}

void launch_potrf_diag(double* a, int lda, int k, int kb, double* dinv, int* fail, hipStream_t s) {
    hipLaunchKernelGGL(k_potrf_diag<true>, dim3(1), dim3(64), 0, s, a, lda, k, kb, dinv, fail);
    LP_HIP(hipGetLastError());
}

void launch_dense_matvec(const double* ainv, int n, const void* b, void* x, int eb, int nrhs, size_t stride,
                         hipStream_t s) {
    dim3 grid((n + 3) / 4, nrhs);
    if (eb == 8)
        hipLaunchKernelGGL(k_matvec<double>, grid, dim3(256), 0, s, ainv, n, (const double*)b, (double*)x, stride);
    else
        hipLaunchKernelGGL(k_matvec<float>, grid, dim3(256), 0, s, ainv, n, (const float*)b, (float*)x, stride);
    LP_HIP(hipGetLastError());
}

}  // namespace lp
