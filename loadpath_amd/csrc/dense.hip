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
#include <algorithm>

#include "lp_internal.h"

namespace lp {

namespace {

constexpr int NB = 64;

// Factor the kb x kb diagonal block at (k, k) in place (lower), write
// inv(L_kk) to dinv (NB x NB, row-major). fail set on a non-positive pivot.
__global__ void __launch_bounds__(256) k_potrf_diag(double* a, int n, int k, int kb, double* dinv, int* fail) {
    __shared__ double s[NB][NB + 1];
    __shared__ double inv[NB][NB + 1];
    const int tid = threadIdx.x;
    for (int e = tid; e < kb * kb; e += blockDim.x) {
        const int i = e / kb, j = e % kb;
        s[i][j] = a[(size_t)(k + i) * n + k + j];
    }
    __syncthreads();
    for (int j = 0; j < kb; ++j) {
        if (tid == 0) {
            const double d = s[j][j];
            if (!(d > 0.0)) {
                *fail = 1;
                s[j][j] = 1.0;
            } else {
                s[j][j] = sqrt(d);
            }
        }
        __syncthreads();
        const double piv = s[j][j];
        for (int i = j + 1 + tid; i < kb; i += blockDim.x) s[i][j] /= piv;
        __syncthreads();
        for (int e = tid; e < (kb - j - 1) * (kb - j - 1); e += blockDim.x) {
            const int i = j + 1 + e / (kb - j - 1), c = j + 1 + e % (kb - j - 1);
            if (c <= i) s[i][c] -= s[i][j] * s[c][j];
        }
        __syncthreads();
    }
    // inv(L): column c solved by thread c
    if (tid < kb) {
        const int c = tid;
        for (int i = 0; i < kb; ++i) {
            double v = (i == c) ? 1.0 : 0.0;
            for (int m = c; m < i; ++m) v -= s[i][m] * inv[m][c];
            inv[i][c] = (i < c) ? 0.0 : v / s[i][i];
        }
    }
    __syncthreads();
    for (int e = tid; e < kb * kb; e += blockDim.x) {
        const int i = e / kb, j = e % kb;
        a[(size_t)(k + i) * n + k + j] = j <= i ? s[i][j] : 0.0;
        dinv[i * NB + j] = inv[i][j];
    }
}

// C (M x N, ldc) = beta C + alpha * op(A) op(B); A: M x K (or K x M if ta),
// B: K x N (or N x K if tb). Row-major. 64x64 tile per block, 256 threads,
// 4x4 outputs per thread, K chunks of 16 staged in LDS.
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
                                              int ta, const double* __restrict__ B, int ldb, int tb, double beta,
                                              double* C, int ldc) {
    __shared__ double As[16][64 + 1];
    __shared__ double Bs[16][64 + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            const int kk = e / 64, mm = e % 64;
            const int gm = m0 + mm, gk = k0 + kk;
            double va = 0;
            if (gm < M && gk < K) va = ta ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            As[kk][mm] = va;
            const int gn = n0 + mm;
            double vb = 0;
            if (gn < N && gk < K) vb = tb ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
            Bs[kk][mm] = vb;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 16; ++kk) {
            double av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] += av[i] * bv[j];
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gm = m0 + ty + 16 * i, gn = n0 + tx + 16 * j;
            if (gm < M && gn < N) {
                double* c = C + (size_t)gm * ldc + gn;
                *c = (beta == 0.0 ? 0.0 : beta * *c) + alpha * acc[i][j];
            }
        }
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

void gemm(hipStream_t s, int M, int N, int K, double alpha, const double* A, int lda, int ta, const double* B, int ldb,
          int tb, double beta, double* C, int ldc) {
    if (M <= 0 || N <= 0) return;
    dim3 grid((N + 63) / 64, (M + 63) / 64);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, M, N, K, alpha, A, lda, ta, B, ldb, tb, beta, C, ldc);
    LP_HIP(hipGetLastError());
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

bool dense_spd_inverse(double* a, int n, Scratch& w, hipStream_t s) {
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
        hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(256), 0, s, a, n, k, kb, Db, fail.as<int>());
        LP_HIP(hipGetLastError());
        const int rest = n - k - kb;
        if (rest > 0) {
            // panel L21 = A21 inv(L11)^T  (rest x kb)
            gemm(s, rest, kb, kb, 1.0, a + (size_t)(k + kb) * n + k, n, 0, Db, NB, 1, 0.0, panel.as<double>(), NB);
            const int cnt = rest * kb;
            hipLaunchKernelGGL(k_copy2d, dim3((cnt + 255) / 256), dim3(256), 0, s, panel.as<double>(), NB,
                               a + (size_t)(k + kb) * n + k, n, rest, kb);
            LP_HIP(hipGetLastError());
            // trailing A22 -= L21 L21^T
            gemm(s, rest, rest, kb, -1.0, panel.as<double>(), NB, 0, panel.as<double>(), NB, 1, 1.0,
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
            gemm(s, kb, k, k, 1.0, a + (size_t)k * n, n, 0, Li, n, 0, 0.0, t.as<double>(), n);
            // Linv[k:k+kb, :k] = -inv(L_kk) T
            gemm(s, kb, k, kb, -1.0, Db, NB, 0, t.as<double>(), n, 0, 0.0, Li + (size_t)k * n, n);
        }
    }
    // A^-1 = Linv^T Linv
    gemm(s, n, n, n, 1.0, Li, n, 1, Li, n, 0, 0.0, a, n);
    int hf = 0;
    LP_HIP(hipMemcpyAsync(&hf, fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
    LP_HIP(hipStreamSynchronize(s));
    return hf == 0;
}

void launch_potrf_diag(double* a, int lda, int k, int kb, double* dinv, int* fail, hipStream_t s) {
    hipLaunchKernelGGL(k_potrf_diag, dim3(1), dim3(256), 0, s, a, lda, k, kb, dinv, fail);
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
