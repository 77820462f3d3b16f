// Coarsest-level FP64 solve: blocked Cholesky and explicit inverse.
//
// Reference: mg._build_coarse / _coarse_solve (mg.py:708-766), which factor
// the assembled coarsest operator with SuperLU on the host and solve there on
// every V-cycle (gpu.py:366-368 copies the right-hand side to the host). Here
// the factorization runs on the device each set_moduli, and the V-cycle's
// coarsest solve becomes one dense FP64 matvec with the inverse (no host
// round trip; a few thousand dofs).
//
// A = L L^T (right-looking, 64-wide panels): a one-workgroup kernel factors
// the diagonal block in LDS and inverts it; the panel is a GEMM with that
// inverse; the trailing update a GEMM. Then Linv by block rows and
// A^-1 = Linv^T Linv.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <stdexcept>

#include "lp_internal.h"

namespace lp {

namespace {

constexpr int NB = 64;

// Factor the kb x kb diagonal block at (k, k) in place (lower), write
// inv(L_kk) to dinv (NB x NB, row-major). fail set on a non-positive pivot.
//
// One wave. Lane i holds row i of the block in registers (rows and columns
// past kb are the identity's, so the padded block factors to the identity).
// Column j's step (right-looking): lane j posts the pivot to LDS, every lane
// below it divides its entry by the pivot's root and posts the multiplier
// L_ij (0 above the diagonal), and each lane updates its row's entries
// m > j by L_ij L_mj with L_mj read back as an LDS broadcast (j and m
// compile-time: both loops unrolled; the entries past the diagonal are
// updated too, never read and zeroed at the end, so no lane masks). L then goes
// to LDS and lane c forms column c of inv(L) row by row, reading L's rows as
// broadcasts: inv[i][c] = (delta_ic - sum_{m < i} L[i][m] inv[m][c]) / L[i][i].
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
    // (four partial sums over m mod 4, so the row's terms are not one
    // dependent chain)
    const int c = i;
    double v[NB];
#pragma unroll
    for (int row = 0; row < NB; ++row) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int m = 0; m < row; ++m) acc[m & 3] = fma(L[row][m], v[m], acc[m & 3]);
        const double sum = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        v[row] = c <= row ? ((row == c ? 1.0 : 0.0) - sum) / L[row][row] : 0.0;
    }
#pragma unroll
    for (int row = 0; row < NB; ++row) dinv[row * NB + c] = v[row];
}

__global__ void k_copy2d(const double* src, int lds, double* dst, int ldd, int rows, int cols) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const int i = e / cols, j = e % cols;
    dst[(size_t)i * ldd + j] = src[(size_t)i * lds + j];
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
    DBuf& panel = w.dn_panel;
    DBuf& linv = w.dn_linv;
    DBuf& t = w.dn_t;
    dinv.alloc((size_t)nblk * NB * NB * 8);
    fail.alloc(sizeof(int));
    panel.alloc((size_t)n * NB * 8);
    linv.alloc((size_t)n * n * 8);
    t.alloc((size_t)n * n * 8);
    LP_HIP(hipMemsetAsync(fail.p, 0, sizeof(int), s));
    LP_HIP(hipMemsetAsync(linv.p, 0, (size_t)n * n * 8, s));
    double* D = dinv.as<double>();
    for (int b = 0; b < nblk; ++b) {
        const int k = b * NB, kb = std::min(NB, n - k);
        double* Db = D + (size_t)b * NB * NB;
        hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(64), 0, s, a, n, k, kb, Db, fail.as<int>());
        LP_HIP(hipGetLastError());
        const int rest = n - k - kb;
        if (rest > 0) {
            // panel L21 = A21 inv(L11)^T  (rest x kb)
            gemm(bh, rest, kb, kb, 1.0, a + (size_t)(k + kb) * n + k, n, 0, Db, NB, 1, 0.0, panel.as<double>(), NB);
            const int cnt = rest * kb;
            hipLaunchKernelGGL(k_copy2d, dim3((cnt + 255) / 256), dim3(256), 0, s, panel.as<double>(), NB,
                               a + (size_t)(k + kb) * n + k, n, rest, kb);
            LP_HIP(hipGetLastError());
            // trailing A22 -= L21 L21^T
            gemm(bh, rest, rest, kb, -1.0, panel.as<double>(), NB, 0, panel.as<double>(), NB, 1, 1.0,
                 a + (size_t)(k + kb) * n + (k + kb), n);
        }
    }
    hipLaunchKernelGGL(k_zero_upper, dim3((unsigned)(((size_t)n * n + 255) / 256)), dim3(256), 0, s, a, n);
    LP_HIP(hipGetLastError());
    // Linv by block rows: Linv[i, :i] = -inv(L_ii) (L[i, :i] Linv[:i, :i]); Linv[i,i] = inv(L_ii)
    double* Li = linv.as<double>();
    for (int b = 0; b < nblk; ++b) {
        const int k = b * NB, kb = std::min(NB, n - k);
        double* Db = D + (size_t)b * NB * NB;
        hipLaunchKernelGGL(k_copy2d, dim3((kb * kb + 255) / 256), dim3(256), 0, s, Db, NB, Li + (size_t)k * n + k, n,
                           kb, kb);
        LP_HIP(hipGetLastError());
        if (k > 0) {
            // T (kb x k) = L[k:k+kb, :k] Linv[:k, :k]
            gemm(bh, kb, k, k, 1.0, a + (size_t)k * n, n, 0, Li, n, 0, 0.0, t.as<double>(), n);
            // Linv[k:k+kb, :k] = -inv(L_kk) T
            gemm(bh, kb, k, kb, -1.0, Db, NB, 0, t.as<double>(), n, 0, 0.0, Li + (size_t)k * n, n);
        }
    }
    // A^-1 = Linv^T Linv
    gemm(bh, n, n, n, 1.0, Li, n, 1, Li, n, 0, 0.0, a, n);
    int hf = 0;
    LP_HIP(hipMemcpyAsync(&hf, fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
    LP_HIP(hipStreamSynchronize(s));
    return hf == 0;
}

void launch_potrf_diag(double* a, int lda, int k, int kb, double* dinv, int* fail, hipStream_t s) {
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(64), 0, s, a, lda, k, kb, dinv, fail);
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
