// Internal declarations shared by the liblpmi.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/loadpath_mi.h"

namespace lp {

void set_error(const std::string& msg);

#define LP_HIP(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            ::lp::set_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " \
                            + __FILE__ + ":" + std::to_string(__LINE__) + ": " #call); \
            throw ::lp::HipFailure();                                                \
        }                                                                            \
    } while (0)

struct HipFailure {};
struct ArgFailure {};

// C-ABI entry guard: HIP failures, argument failures and other exceptions
// become status codes (the message is already in the error slot for the
// first two)
template <typename F> int api_call(F&& f) {
    try {
        return f();
    } catch (const HipFailure&) {
        return -1;
    } catch (const ArgFailure&) {
        return -2;
    } catch (const std::exception& e) {
        set_error(e.what());
        return -3;
    }
}

// Owning device allocation.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        if (n <= bytes && p) return;
        release();
        if (n == 0) return;
        LP_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

enum LevelKind { LK_FINE = 0, LK_WEIGHTS = 1, LK_EXPLICIT = 2 };

constexpr int kMaxLevels = 16;
constexpr int kBlock = 256;

// Geometry of one level: cells (nx, ny, nz), nodes (nx+1, ny+1, nz+1).
struct Geom {
    int nx = 0, ny = 0, nz = 0;
    int Nx = 0, Ny = 0, Nz = 0;
    size_t ncell = 0, nn = 0;
    void set(int x, int y, int z) {
        nx = x; ny = y; nz = z;
        Nx = x + 1; Ny = y + 1; Nz = z + 1;
        ncell = (size_t)x * y * z;
        nn = (size_t)Nx * Ny * Nz;
    }
};

struct Level {
    int kind = LK_FINE;
    Geom g;
    Geom fine;                  // finer level's geometry (weights kind reads its E)
    DBuf mask;                  // uint8 per node, bit c = component c free and live
    DBuf dinv64, dinv32;        // 3*nn
    DBuf e64, e32;              // fine: cell moduli
    DBuf patch_idx;             // weights: int per cell, -1 or slot
    DBuf patch64, patch32;      // weights: (npatch, 576)
    DBuf patch_cell;            // weights: int per patch, its cell
    DBuf patch_prod;            // weights, mixed path: FP32 (max_rhs, npatch, 24) patch products
    int npatch = 0;
    DBuf k64, k32;              // explicit: (576, ncell)
    bool k64_apply = false;     // weights: k64 holds sum_q w_q M_q and the FP64 operator applies it
    bool k32_apply = false;     // weights: k32 holds them in FP32 for the FP32 / mixed V-cycle
    // mixed V-cycle, coarse levels: the cell matrices in FP16 with a power-of-two
    // scale per cell (k16: (corner, 16-byte group, cell) of 8 halves, the
    // corner's 3 rows x 24 columns row-major in 9 groups; sc16: per cell)
    DBuf k16, sc16;
    bool k16_apply = false;
    double lmax = 0.0;
    long long n_active = 0;
    // V-cycle work vectors (elem size 8 or 4) for max_rhs right-hand sides
    DBuf vb, vx, vr, vd, vd2, vt;
    DBuf lmax_vec;              // power-iteration vector (warm start)
    DBuf held;                  // uint8 per node, bit c = component c held and alive (strict / patch)
    bool lmax_warm = false;
};

// Scratch that the per-iteration paths (hierarchy rebuild, lambda_max,
// coarsest factorization) reuse: it only grows, so a SIMP iteration makes
// no device allocation (hipFree synchronizes the whole device, which would
// serialize grids running concurrently on their own streams).
struct Scratch {
    DBuf counter, counter2, dtmp, dmaxb, ctr;
    DBuf lmax_sc, lmax_av, lmax_tmp;
    DBuf dn_dinv, dn_fail, dn_panel, dn_linv, dn_t;
    void* blas = nullptr;       // rocBLAS handle for the dense coarsest factor (its grid's stream)
    Scratch() = default;
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    ~Scratch();
};

struct Grid {
    lp_options opt{};
    Geom g0;
    double ke[576];
    DBuf ke64, ke32;            // device copies of Ke
    DBuf mq64, mq32;            // 8 x 576 first-coarse-level matrices P_k^T Ke P_k
    DBuf pk;                    // 8 x 576 child prolongations (double)
    DBuf fixed;                 // uint8 per node, bit c = component c held
    int nlevels = 0;
    Level lv[kMaxLevels];
    bool strict = false;
    bool built = false;
    hipStream_t stream = nullptr;
    // coarsest dense inverse
    int ncoarse = 0;            // dofs of the coarsest level (3 * nodes)
    DBuf cinv;                  // ncoarse^2 doubles
    DBuf cwork;                 // factor workspace
    // CG vectors (FP64), max_rhs each
    DBuf x, r, p, ap, b;        // FP64 CG state (z = M r is never stored in FP64)
    DBuf p2;                    // the next direction, written by the fused CG operator (swapped with p)
    DBuf x0;                    // the warm start kept for a repeat in the strict hierarchy
    // mixed path, FP32 CG with reliable FP64 updates: direction, its product
    // and the solution since the last update (FP32); G.x holds the FP64 sum
    DBuf p32, ap32, xlo32;
    int tune_cg32 = 0;          // the mixed path's CG in FP32 with reliable updates (1) or in FP64 (0)
    double cg32_delta = 0.1;    // a reliable update when |r| falls below delta x its largest since the last
    bool cg32_flex_replaced = false;   // the flexible beta also at a replaced step
    long long cg32_updates = 0; // reliable updates so far
    // the STALL fallback of GridElasticity._solve inside lp_solve and the
    // device loops (1); 0 when the caller runs it around the solve, as the
    // reference's GridElasticity does around a backend's solve; 2 caps the
    // first attempt at one iteration (exercises the strict repeat)
    int stall = 1;
    int strict_switches = 0;    // solves that moved to the strict hierarchy
    // reductions
    DBuf partials;              // doubles
    size_t partials_cap = 0;
    DBuf scal;                  // per-rhs scalar block (CgScal)
    DBuf status_h;              // unused placeholder
    void* host_pinned = nullptr;
    size_t host_pinned_bytes = 0;
    // staging for host pointers
    DBuf stage_a, stage_b;
    Scratch work;
    int max_rhs = 1;
    bool ke_sym = false;        // Ke has the cube's symmetry (ke_pattern.h): orbit values used
    double ke_orbit[32] = {};
    int tune_fine_ny = 1;       // fine-level nodes per thread along y (1 or 2)
    int tune_fine_xchunk = 0;   // fine-level nodes per thread along x (0: from the grid's size)
    int tune_weights_xchunk = 8; // weights-level node planes per block (FP16-MFMA kernels)
    int tune_fused_cheb = 1;    // MFMA levels: Chebyshev vector passes fused into the operator
    int tune_graph = 1;         // PCG replays the V-cycle from a captured hipGraph
    int tune_wave = 1;          // FP16-MFMA levels: wave-tile kernel on the fine level (1), on both (2), on none (0)
    int tune_rhs_inner = 1;     // FP16-MFMA kernels: a tile's right-hand sides in consecutive blocks of one XCD
    int tune_bf16 = 0;          // mixed path, fine level: BF16 operands (1) or scaled FP16 ones (0)
    int tune_l1_explicit = 1;   // FP64 path: first coarse level as formed cell matrices (1) or matrix-free weights (0)
    int tune_l1_explicit32 = 0; // FP32 / mixed V-cycles: the same in FP32 (1) or matrix-free weights (0)
    int tune_half_explicit = 2; // mixed V-cycle: every coarse level (2) / the Galerkin levels (1) as FP16 cell matrices, or none (0)
    int tune_cg32_restart = 0;  // FP32 CG: restart at each reliable update (iterative refinement)
    int tune_cg_fused = 1;      // FP64 CG: direction and x updates fused into the next operator launch
    int tune_half_brick = 40804; // FP16 cell-matrix levels: LDS-staged bricks (ty * 10000 + tz * 100 + x-chunk), 0 = off
    int tune_half_brick_min = 512; // ... on levels of at least this many bricks
    int tune_cheb_r = 1;        // fused degree-2 smoothers from nonzero x: residual step stores r only
    int tune_flex = -1;         // flexible (Polak-Ribiere) beta: 1 on, 0 off, -1 on for the mixed path only
    bool flex() const { return tune_flex < 0 ? opt.precision == LP_PREC_MIXED : tune_flex != 0; }
    // mixed path: a solve the FP16 levels fail is repeated with FP32
    // operators on every level (1), not repeated (0), or forced after one
    // iteration (2, exercises the repeat); fp16_off holds for the rest of
    // the current hierarchy, fallbacks counts the repeats
    int tune_fallback = 1;
    bool fp16_off = false;
    int fallbacks = 0;
    bool fp16() const { return opt.precision == LP_PREC_MIXED && !fp16_off; }
    // the captured V-cycle: valid for one batch size and one epoch (every
    // hierarchy build, lambda_max estimate or tuning change starts a new one,
    // since the graph holds the level pointers and the Chebyshev coefficients)
    hipGraphExec_t vgraph = nullptr;
    int vgraph_nrhs = 0;
    long long vgraph_epoch = -1;
    long long epoch = 0;
    int vbytes() const { return opt.precision == LP_PREC_FP64 ? 8 : 4; }
    // launch profiling of one level operator (HIP events on the grid's stream)
    struct Prof {
        bool on = false;
        int level = 0, eb = 0;
        std::vector<hipEvent_t> ev;      // pairs
        std::vector<double> bytes;       // algorithmic bytes per recorded launch
        size_t used = 0;
        long long dropped = 0;
    };
    mutable Prof prof;
    ~Grid() {
        for (hipEvent_t e : prof.ev) hipEventDestroy(e);
        if (vgraph) hipGraphExecDestroy(vgraph);
    }
};

// Per right-hand-side CG scalars, device resident.
struct CgScal {
    double rz, pap, alpha, beta, rr, bnorm, tol, tmp;
    int active, converged, failed, iters;
    int xpend;   // alpha of this iteration not yet added to x (x += alpha p)
};

// ---- element.cpp -------------------------------------------------------
void hex8_ke_host(double nu, int element, double* ke);
void hex8_cb_host(double nu, int element, int points, double* out /* npts*6*24 */);
void child_prolongations_host(double* pk /* 8*576 */);

// ---- apply.hip ---------------------------------------------------------
// Epilogue selectors for the level operator y = A u at each node.
enum EpiKind {
    EPI_STORE = 0,        // out = y
    EPI_STORE_DOT = 1,    // out = y; partial += u . y          (p . Ap)
    EPI_CHEB = 2,         // x += d; r -= y; d2 = c1 d + c2 dinv r
    EPI_RESID_SCALED = 3, // r = b - y; d = s dinv r
    EPI_RESID_MASK = 4,   // out = (b - y) * m
    EPI_JACOBI = 5,       // out = u + w dinv (b - y)
    EPI_POWER = 6,        // out = dinv * y * m; partial += u . out
    EPI_RESID = 7,        // out = b - y; partial += out . out (double)
    EPI_MASKED = 8,       // out = m ? y : 0 (the masked operator m K m u, no identity rows)
    // fused Chebyshev steps (FP16-MFMA level operators only). The first step
    // from x = 0 stages its operand d = (dinv b) s from b (u = b); the last
    // step also adds the next direction to x, so no separate update pass:
    EPI_CHEB_FIRST = 9,   // d = (dinv b) s: x = d; r = b - y; d2 = c1 d + c2 dinv r
    EPI_CHEB_ONLY = 10,   // d = (dinv b) s: x = d + c1 d + c2 dinv (b - y)
    EPI_CHEB_LAST = 11,   // x += d + c1 d + c2 dinv (r - y)
    // the fine FP64 operator only: u = p_old; the operand is p = z + beta p_old
    // (z = d, of element size zbytes); d2 = p; out = A p; partial += p . Ap;
    // x += alpha p_old where the CG step is pending (alpha, beta from scal)
    EPI_CG_APPLY = 12,
    // the last Chebyshev step of a smoother whose residual step stored r
    // only: the operand is r, staged as d = (dinv r) s (s = c3), and the
    // operand itself is the epilogue's r (FP16-MFMA level operators only)
    EPI_CHEB_LAST_R = 13, // d = (dinv r) s: x += d + c1 d + c2 dinv (r - y)
};

struct EpiArgs {
    int kind = EPI_STORE;
    void* out = nullptr;
    const void* b = nullptr;
    void* x = nullptr;
    void* r = nullptr;
    const void* d = nullptr;
    void* d2 = nullptr;
    const void* dinv = nullptr;
    double c1 = 0, c2 = 0;
    double c3 = 0;                // operand scale s of EPI_CHEB_FIRST / EPI_CHEB_ONLY / EPI_CHEB_LAST_R
    double* partials = nullptr;   // per (rhs, block)
    const CgScal* scal = nullptr; // skip inactive rhs when set
    const double* scale_dev = nullptr;  // optional per-rhs scale read on device
    int zbytes = 0;               // EPI_CG_APPLY: element size of z (4 or 8)
};

// Launch y = A_level u with epilogue, for nrhs vectors of stride vstride.
// Returns the number of blocks per rhs (for partial sums).
int launch_apply(const Grid& G, int level, int elem_bytes, const void* u, size_t vstride, int nrhs,
                 const EpiArgs& epi, hipStream_t s);
// blocks (= reduction partials per rhs) of one operator launch at element size eb
int apply_blocks(const Grid& G, int level, int eb);
int mfma_blocks(const Grid& G, int level);
// blocks (= partials per rhs) of the vector fine-level kernel (launch_apply_fine) at level 0
int fine_vector_blocks(const Grid& G);
// blocks (= partials per rhs) of an FP32 explicit-level launch
int explicit_blocks(const Grid& G, int level);
// the FP64 operator of a weights level applies its formed cell matrices (k64)
bool explicit_weights(const Grid& G, int level, int eb);
// the mixed V-cycle's coarse level applied from its FP16 cell matrices
bool half_explicit(const Grid& G, int level, int eb);
void launch_apply_half(const Grid& G, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                       hipStream_t s);
// blocks (= partials per rhs) of an FP16 cell-matrix launch (the brick form
// when tune_half_brick selects it)
int half_blocks(const Grid& G, int level);
// tune_half_brick's value is a brick the kernel takes
bool half_brick_shape_ok(int v);
// FP16 cell matrices with a per-cell scale from the FP64 weights-level sums
// (src32 == nullptr) or from an FP32 (576, cells) array
void form_half_matrices(const Grid& G, int level, const float* src32, hipStream_t s);
void launch_apply_fine(const Grid& G, int level, int eb, const void* u, size_t vstride, int nrhs,
                       const EpiArgs& E, hipStream_t s);
void launch_apply_weights(const Grid& G, int level, int eb, const void* u, size_t vstride, int nrhs,
                          const EpiArgs& E, hipStream_t s);
void launch_apply_explicit(const Grid& G, int level, int eb, const void* u, size_t vstride, int nrhs,
                           const EpiArgs& E, hipStream_t s);
// fine level of the mixed-precision V-cycle: FP16 MFMA Ke contraction, FP32 vectors
void launch_apply_fine_mfma(const Grid& G, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                            hipStream_t s);

// ---- vec.hip -------------------------------------------------------------
void ensure_partials(Grid& G, size_t n);
// sum partials[rhs*nb + i] -> dst[rhs] (double)
void launch_reduce(const double* partials, int nb, int nrhs, double* dst, size_t dst_stride_bytes,
                   hipStream_t s);
int dot_blocks(size_t n);
void launch_dot(const void* a, const void* b, int elem_bytes, size_t n, size_t stride, int nrhs,
                double* partials, const CgScal* scal, hipStream_t s);
void launch_cast(const void* src, int src_bytes, void* dst, int dst_bytes, size_t n, size_t stride,
                 int nrhs, hipStream_t s);
void launch_fill(void* dst, int elem_bytes, size_t n, double v, hipStream_t s);
// scale[rhs] = max |v_rhs| (1 if zero)
// z = (double) v and the partial sums of r . z (skipping inactive rhs)
void launch_cast_dot(const void* v, int vb, const double* r, double* z, size_t n, size_t stride, int nrhs,
                     double* partials, const CgScal* scal, hipStream_t s, const double* ap = nullptr);
// flexible beta from the two sums of launch_cast_dot(..., ap)
void launch_cg_beta_flex(Grid& G, int nrhs, int nb, hipStream_t s);
// v /= sqrt(sum of dotp[0, nbd)), that sum into *sv, and (prev given) the
// sum of prev[0, nbp) into *sp, in one launch (one right-hand side)
void launch_norm_scale(void* v, int elem_bytes, size_t n, const double* dotp, int nbd, double* sv,
                       const double* prev, int nbp, double* sp, hipStream_t s);
void launch_scale_by_dev(void* v, int elem_bytes, size_t n, size_t stride, int nrhs,
                         const double* num, size_t num_stride, int mode, hipStream_t s);
void launch_cheb_start(const void* b, const void* dinv, double s_, void* x, void* r, void* d,
                       int elem_bytes, size_t n, size_t stride, int nrhs, hipStream_t st);
void launch_axpy_inplace(void* x, const void* d, int elem_bytes, size_t n, size_t stride, int nrhs,
                         hipStream_t st);
void launch_jacobi_start(const void* b, const void* dinv, double w, void* x, int elem_bytes,
                         size_t n, size_t stride, int nrhs, hipStream_t st);
// x += a p; r -= a ap; rv = r in the V-cycle precision (vb bytes); partial rr
// r -= alpha Ap with |r|^2 partials and rv = r in the V-cycle precision
void launch_cg_step(Grid& G, int nrhs, void* rv, int vb, hipStream_t s);
// x += alpha p (pending rhs), then p = z + beta p (active rhs) unless xonly
void launch_cg_update(Grid& G, int nrhs, const void* z, int zb, bool xonly, hipStream_t s);
// finalize kernels on CgScal
void launch_cg_finalize(Grid& G, int nrhs, int stage, int nb, hipStream_t s);
// warm start: which 0 = partials of |b - s Ax|^2 (Ax in G.ap), 1 = apply the decision
void launch_warm(Grid& G, int nrhs, int which, hipStream_t s);
// FP32 CG with reliable FP64 updates: xlo += alpha p, r -= alpha ap (|r|^2 partials)
void launch_cg32_step(Grid& G, int nrhs, float* r, hipStream_t s);
// partials of r . z and (ap != null) z . ap, in the flexible beta's layout
void launch_dot2_f32(Grid& G, int nrhs, const float* r, const float* z, const float* ap, hipStream_t s);
// p = z + beta p (first: p = z)
void launch_cg32_update(Grid& G, int nrhs, const float* z, bool first, hipStream_t s);
// x += xlo, xlo = 0
void launch_cg32_fold(Grid& G, int nrhs, hipStream_t s);
// rv = (float) r (active rhs)
void launch_cast_active(Grid& G, int nrhs, const double* r, float* rv, hipStream_t s);

// ---- transfer.hip --------------------------------------------------------
void launch_restrict_masked(const void* rf, const Geom& fine, const uint8_t* cmask, void* rc,
                            const Geom& coarse, int elem_bytes, int nrhs, size_t fstride,
                            size_t cstride, hipStream_t s);
void launch_prolong_add(const void* xc, const uint8_t* cmask, const Geom& coarse, const uint8_t* fmask,
                        void* xf, const Geom& fine, int elem_bytes, int nrhs, size_t cstride,
                        size_t fstride, hipStream_t s);
void launch_prolong_plain(const double* c, double* f, int cx, int cy, int cz, hipStream_t s);
// mixed path, weights level: the FP16-MFMA operator (patch products first)
void launch_apply_weights_mfma(const Grid& G, int level, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                               hipStream_t s);
bool weights_mfma_path(const Grid& G, int level, int eb);
// the level's operator at element size eb is an FP16-MFMA kernel (fused Chebyshev kinds allowed)
bool mfma_level(const Grid& G, int level, int eb);
void launch_restrict_plain(const double* f, double* c, int cx, int cy, int cz, hipStream_t s);

// ---- hierarchy.hip -------------------------------------------------------
int build_hierarchy(Grid& G);             // returns LP_OK or LP_SINGULAR
void level_cell_dofs(const Grid& G, int level, int64_t* out_dev);

// ---- dense.hip -----------------------------------------------------------
// Factor the SPD n x n matrix a (lower Cholesky, in place) and overwrite it
// with its inverse. Returns false on a non-positive pivot.
bool dense_spd_inverse(double* a, int n, Scratch& w, hipStream_t s);
// Row-major C (M x N, ldc) = alpha op(A) op(B) + beta C by rocBLAS dgemm on
// handle h (the library's plain GEMM); blas_handle creates the handle in
// slot on first use and binds it to stream s.
void blas_gemm_rm(void* h, bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int lda,
                  const double* B, int ldb, double beta, double* C, int ldc);
void* blas_handle(void*& slot, hipStream_t s);
void blas_release(void* h);
// Factor the kb x kb diagonal block at (k, k) of the row-major matrix a
// (leading dimension lda; kb <= 64) in place, lower; inv(L_kk) to dinv
// (64 x 64 row-major); *fail set on a non-positive pivot.
void launch_potrf_diag(double* a, int lda, int k, int kb, double* dinv, int* fail, hipStream_t s);
void launch_dense_matvec(const double* ainv, int n, const void* b, void* x, int elem_bytes, int nrhs,
                         size_t stride, hipStream_t s);

// ---- label.hip ------------------------------------------------------------
// 26-connected components of a device cell mask: labels = the root (smallest
// cell index) of each cell's component, -1 off the mask
void label26_roots(int nx, int ny, int nz, const uint8_t* mask, int* labels, int* flag, hipStream_t s);

// ---- solver.hip -----------------------------------------------------------
void vcycle(Grid& G, int level, int nrhs);           // on lv[level].vb -> lv[level].vx
void pcg(Grid& G, int nrhs, bool warm, int* iters, int* converged);
// CG iterations the plain hierarchy gets before the strict one takes over (mg.py:928-930)
constexpr int kStall = 100;
// pcg with the STALL -> strict fallback (mg.py:950-965); LP_OK or a rebuild's status
int solve_stall(Grid& G, int nrhs, bool warm, int* iters, int* converged);
// set_moduli's rebuild from the fine moduli in lv[0].e64, lambda_max included
int rebuild_hierarchy(Grid& G);

}  // namespace lp
