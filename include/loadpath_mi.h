/*
 * loadpath_mi.h — C-ABI of the MI355X (gfx950) voxel-SIMP library liblpmi.so.
 *
 * This is the drop-in boundary for loadpath's voxel SIMP hot path. Every
 * entry point replaces one reference interface, cited as file:line into the
 * loadpath repository. The Python host package (loadpath_amd) binds these
 * with ctypes; a loadpath-side binding is a thin ctypes wrapper over the same
 * symbols.
 *
 * Conventions
 *  - Plain pointers and sizes only. Node-grid fields are component-major
 *    (3, nx+1, ny+1, nz+1), C order, exactly mg.GridElasticity.to_grid's
 *    layout (mg.py:495-499); several right-hand sides are stacked in front:
 *    (nrhs, 3, nx+1, ny+1, nz+1). Cell fields are (nx, ny, nz), C order
 *    (mg.py:508-511). nx, ny, nz are the even-padded cell counts
 *    (mg.py:438-441).
 *  - `device_ptr` != 0: pointers are device pointers (e.g. a torch tensor's
 *    data_ptr()) used in place; 0: host pointers, the library copies.
 *  - Every call returns an int status: LP_OK, LP_NOT_CONVERGED,
 *    LP_SINGULAR, or a negative error; lp_last_error() gives the message
 *    (thread-local). A handle is not re-entrant: one host thread per handle,
 *    as mgops.py:29-30 requires of a GridElasticity.
 */
#ifndef LOADPATH_MI_H
#define LOADPATH_MI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LP_OK 0
#define LP_NOT_CONVERGED 1   /* mg.py:944-946 RuntimeError("linear solve did not converge") */
#define LP_SINGULAR 2        /* mg.py:580-586 singular coarsest factor -> strict hierarchy */
#define LP_ERR_HIP (-1)
#define LP_ERR_ARG (-2)
#define LP_ERR_STATE (-3)

/* precision of the V-cycle (the outer CG is always FP64) */
#define LP_PREC_FP64 0       /* FP64 everywhere ("plain FP64") */
#define LP_PREC_FP32 1       /* FP64 CG + FP32 V-cycle: the reference split (mg.py:28-32, gpu.py:483-521) */
#define LP_PREC_MIXED 2      /* FP64 CG + FP32 V-cycle whose two finest operators are FP16-MFMA contractions */

#define LP_SMOOTH_CHEBYSHEV 0  /* mg.py:796-835 */
#define LP_SMOOTH_JACOBI 1     /* damped Jacobi (configs' V(2,2), omega 0.6) */

typedef struct lp_options {
    double rtol;        /* ||r|| <= rtol ||b||, GridElasticity(rtol=) mg.py:428 */
    int maxiter;        /* GridElasticity(maxiter=) */
    int coarse_cells;   /* GridElasticity(coarse_cells=) mg.py:556 */
    int degree;         /* Chebyshev degree / Jacobi sweeps, GridElasticity(degree=) */
    int smoother;       /* LP_SMOOTH_* */
    double omega;       /* damped-Jacobi weight */
    int precision;      /* LP_PREC_* */
    int max_rhs;        /* right-hand sides solved together (batched PCG) */
    int lmax_iters;     /* power iterations for lambda_max (mg.py:771, 12) */
    int max_levels;     /* 0 = no cap; else stop coarsening at this many levels */
} lp_options;

typedef struct lp_grid lp_grid;

/* ---- element ----------------------------------------------------------- */
/* fem.hex8_ke(nu, element) (fem.py:220-224): unit-cube 24x24 at E = 1.
 * element: 0 = "im" (Wilson bubbles condensed), 1 = "plain" trilinear. */
int lp_hex8_ke(double nu, int element, double* ke_out /* 576 */);

/* C B at the stress sample points, E = 1, unit pitch (constrained.py:383-387):
 * points 0 = the centre, fem.hex8_b_center (fem.py:307-308), 1 point;
 * points 1 = the 8 corners, x outermost, fem.hex8_b with the condensed
 * bubbles (fem.py:227-237); points 2 = the 8 quarter points (natural
 * +-0.5, fem.CORNERS order: bit 0 = x) of constrained._Averaged
 * (constrained.py:176-183), with the bubbles. out (npts, 6, 24); divide
 * by the pitch. */
int lp_hex8_cb(double nu, int element, int points, double* out);

/* ---- grid handle: mg.GridElasticity (mg.py:411-511) --------------------- */
/* fixed_grid: (3, nx+1, ny+1, nz+1) uint8 (mg.py:447-451); ke: 576 doubles
 * already scaled by the pitch (mg.py:446). */
int lp_grid_create(lp_grid** out, int nx, int ny, int nz, const uint8_t* fixed_grid,
                   const double* ke, const lp_options* opt);
void lp_grid_destroy(lp_grid* g);

/* GridElasticity.set_moduli(e) (mg.py:515-591): per-cell moduli on the
 * padded cell grid; rebuilds every level (Galerkin coarsening, masks,
 * lambda_max, coarsest Cholesky) on the device. Returns LP_SINGULAR if the
 * coarsest factor fails even in the strict hierarchy.
 * flags: bit 0 = e is a device pointer; bit 1 = defer lambda_max (the
 * caller then calls lp_estimate_lmax per level). */
int lp_set_moduli(lp_grid* g, const double* e_cells, int flags);

/* GridElasticity.rtol / .maxiter, which the reference's callers set before
 * a solve (simp.py:734-746; the delivery solves at 1e-6, simp.py:2481):
 * the tolerance and cap of the next lp_solve calls. */
int lp_set_solve_tol(lp_grid* g, double rtol, int maxiter);

/* GridElasticity._estimate_lmax (mg.py:771-794): 12 power iterations of
 * D^-1 A at `level`, times 1.3. start: host (3, nodes) start vector, or NULL
 * for the level's last vector (warm start) or a fixed hash start. */
int lp_estimate_lmax(lp_grid* g, int level, const double* start);

/* Hierarchy description: levels, their cell dims (3 ints each) and flags. */
int lp_grid_info(lp_grid* g, int* nlevels, int* dims, int* strict, double* lmax);

/* _Level.apply (mg.py:317-346) / mgops.apply_masked (mgops.py:247-261):
 * out = m*K(m*u) + (1-m)*u at `level`. dtype_bytes 8: the FP64 operator
 * (level 0 only), 4: the FP32 V-cycle copy. */
int lp_apply(lp_grid* g, int level, int dtype_bytes, const void* u, void* out, int nrhs, int device_ptr);

/* GridElasticity.precondition (mg.py:862-864): one V-cycle, FP64 in/out. */
int lp_precondition(lp_grid* g, const double* r, double* z, int nrhs, int device_ptr);

/* GridElasticity._pcg / solve_many on node grids (mg.py:872-965): b, x
 * (nrhs, 3, nodes) FP64, already masked; warm != 0 starts from x (the
 * energy-optimal warm start, mg.py:881-896). iters/converged per rhs. */
int lp_solve(lp_grid* g, int nrhs, const double* b, double* x, int warm, int device_ptr,
             int* iters, int* converged);

/* GridElasticity.strain_energy (mg.py:975-988) / femops.element_energy
 * (femops.py:236-242): se (nrhs, nx, ny, nz) = u_c^T Ke u_c. */
int lp_strain_energy(lp_grid* g, int nrhs, const double* u, double* se, int device_ptr);

/* b = f masked to the free live dofs of the current hierarchy (mg.py:938-
 * 939: b = to_grid(f) * levels[0].mask), nrhs node-grid fields: the
 * right-hand side lp_solve takes, formed on the device. */
int lp_mask_rhs(lp_grid* g, int nrhs, const double* f, double* b, int device_ptr);

/* The bilinear form a_c^T Ke b_c per cell (the adjoint term of
 * constrained.py:659, lambda_e^T Ke u_e): out (nrhs, nx, ny, nz). */
int lp_element_product(lp_grid* g, int nrhs, const double* a, const double* b, double* out, int device_ptr);

/* p-mean von Mises stress aggregate, constrained.py:607-664 (the
 * non-averaged branch), one load case per rhs on the fine grid:
 *   sigma_k = cb_k u_c (cb from lp_hex8_cb / pitch, npts 1 or 8),
 *   vm = sqrt(max(sigma^T V sigma, 1e-30)), s = x_c^q vm / lim (0 at a
 *   point not counted), pm = sc (sum (s / sc)^p / nc)^(1/p), sc = max(s_max,
 *   1e-30). u (nrhs, 3, nodes), x (nx, ny, nz) the material share, counted
 * (nx, ny, nz) uint8 with bit k = sample point k counted (0xff: the whole
 * cell; per point for the contact band, constrained.py:349-358), or NULL
 * (every point), nc the number of counted samples. Outputs pm
 * and s_max (host, nrhs); optional rhs (nrhs, 3, nodes) = d(kappa pm)/du,
 * zero on fixed dofs (the adjoint right-hand side), and gx (nrhs, nx, ny,
 * nz) = the explicit d(kappa pm)/dx at fixed u. The full gradient is
 * gx - dE/dx * lp_element_product(lambda, u) with K lambda = rhs. */
int lp_stress_pnorm(lp_grid* g, int nrhs, const double* u, const double* x, const unsigned char* counted,
                    const double* cb, int npts, double p, double q, double lim, double kappa, double nc,
                    double* pm, double* smax, double* rhs, double* gx, int device_ptr);

/* Level data for parity checks: what = 0 mask bits (uint8, per node),
 * 1 dinv (double, 3 x nodes), 2 cell dof indices (int64, (ncells, 24), the
 * index arithmetic every apply kernel uses; fem.HexMesh.edof on the grid),
 * 3 held bits (uint8 per node: alive and fixed; alive = mask | held). */
int lp_level_data(lp_grid* g, int level, int what, void* out);

/* mg.prolong / mg.restrict (mg.py:84-132) on device fields of 3 components:
 * coarse (3, cx, cy, cz) <-> fine (3, 2cx-1, 2cy-1, 2cz-1), FP64. */
int lp_prolong(const double* coarse, double* fine, int cx, int cy, int cz, int device_ptr);
int lp_restrict(const double* fine, double* coarse, int cx, int cy, int cz, int device_ptr);

/* ---- SIMP loop on the device (benchmarks/reference.py:214-257 and
 *      simp.py:1143-1189, 1897-1904, 2247-2381) ---------------------------- */
typedef struct lp_simp lp_simp;

#define LP_FILTER_TOP3D 0     /* reach ceil(r)-1, row sums (reference.py:188-211) */
#define LP_FILTER_PADDED 1    /* reach floor(r), full kernel sum (simp.py:1143-1189) */

typedef struct lp_simp_stats {
    double compliance;    /* sum_i sum_c E_c ce_c of the iteration's design */
    double change;        /* max |xnew - x| */
    double volume;        /* mean xphys after the update (production loop:
                             the blueprint's mean over design cells before it) */
    int cg_iters;         /* max PCG iterations over the load cases */
    int cg_iters_total;   /* PCG iterations summed over the load cases */
    int oc_evals;         /* bisection evaluations (production loop: volume
                             evaluations of the multiplier root) */
    double seconds;
    /* production loop only (lp_simp_set_loop), else 0 */
    double beta;          /* projection sharpness of this iteration */
    double move;          /* the step limit it used */
    double step_fit;      /* where along the last step the measured minimum
                             sat, min(., 99); -1 when not measured */
    double lambda;        /* the volume multiplier */
    int vol_held;         /* the root was bracketed */
    int converged;        /* the loop's stop test passed after this iteration */
} lp_simp_stats;

/* loads: (nrhs, 3, nodes) FP64 node-grid load vectors (masked inside).
 * active: (nx, ny, nz) uint8 design cells or NULL (passive cells stay 0). */
int lp_simp_create(lp_simp** out, lp_grid* g, int nrhs, const double* loads, const uint8_t* active,
                   double volfrac, double rmin, int filter_kind, double penal, double emin, double move,
                   int warm_start);
void lp_simp_destroy(lp_simp* s);
/* n OC iterations; stats_out holds n records. */
int lp_simp_iterate(lp_simp* s, int n, lp_simp_stats* stats_out);
/* what: 0 x, 1 xphys, 2 dc (filtered), 3 ce (summed over cases): (nx,ny,nz) doubles */
int lp_simp_get(lp_simp* s, int what, double* out);
int lp_simp_set_x(lp_simp* s, const double* x);
/* The design update of lp_simp_iterate: 0 optimality criteria with the
 * volume bisection (reference.py:240-251, the default); 1 MMA (mma.py) on
 * 10 c / c_first with the volume constraint sum xphys / (vf n) - 1 <= 0, both
 * gradients through the density filter (the normalization of loadpath's MMA
 * loops on compliance-type objectives, thermal.py:1035-1043). Boxes without
 * passive cells. */
int lp_simp_set_optimizer(lp_simp* s, int optimizer);

/* loadpath's production loop, simp._optimize_run (simp.py:1640-2449) for a
 * box with the plain compliance objective: blueprint xp = P(H x / hs, beta,
 * eta), analysed (eroded) design xe = P(., beta, eta + erode), preserved
 * cells 1, E = emin + xe^penal (1 - emin); dc = H (dc_xp dP_e / hs),
 * dv = H (w dP / hs) (simp.py:1994-2204); the measured step
 * (simp.py:2228-2245); the OC step with the volume multiplier bracketed in
 * log lambda from the last one and landed by Illinois false position to
 * 1e-13 of the target (simp.py:2303-2373); beta doubling on the schedule
 * and the stop tests (simp.py:2413-2449). Fields as SimpOptions
 * (simp.py:899-1000); schedule: iterations per doubling by level
 * (beta_schedule) or NULL. */
typedef struct lp_loop_opts {
    double beta_start;    /* 1 */
    double beta_max;      /* 32 */
    int beta_every;       /* 10 */
    double eta;           /* 0.5 */
    double erode;         /* 0.15 */
    double move;          /* 0.2 */
    int adaptive_move;    /* 1: the measured step; 0: move, move / 2 past beta 4 */
    double tol;           /* 0.01 */
    double stall_rtol;    /* 2e-3 */
    int stall_window;     /* 6 */
    double emin;          /* 1e-6 (simp.py E_MIN) */
    double penal;         /* 3 */
    const int* schedule;  /* NULL or schedule_len entries */
    int schedule_len;
} lp_loop_opts;
/* Switch s to the production loop and restart it: x = volfrac, preserved
 * cells (preserve[c] != 0, NULL: none) 1, passive cells 0. case_weights:
 * nrhs weights of the load cases' compliances, NULL: all 1. Each
 * lp_simp_iterate call then runs loop iterations; stats.converged marks
 * the iteration after which the loop stops. */
int lp_simp_set_loop(lp_simp* s, const lp_loop_opts* o, const uint8_t* preserve, const double* case_weights);

/* SIMP moduli and their derivative (simp.py:1861, 1897-1904): e = emin +
 * x^penal (1 - emin), de = penal x^(penal-1) (1 - emin) (de may be NULL). */
int lp_simp_moduli(long long n, const double* x, double penal, double emin, double* e, double* de, int device_ptr);

/* The production loop's design ties (simp.py:1650, 2223-2226, 2288-2289,
 * 2296-2304): mirror planes through the box centre across the given axes
 * (0 x, 1 y, 2 z; domain.mirror, applied in this order as
 * simp._symmetrize does, simp.py:1476-1482) and extrusion columns along
 * extrusion_axis (-1: none; manufacture.tie_groups(domain, extrusion),
 * manufacture.py:413-427, and simp._extrude, simp.py:1453-1463: each
 * column's design cells take their mean). The sensitivities dc and dv are
 * tied after the filter, every trial design of the volume root before its
 * volume is read (void cells emptied again after the columns). */
int lp_simp_set_ties(lp_simp* s, const int* mirror_axes, int n_mirror, int extrusion_axis);

/* The multigrid path's live-cell reduction of simp._Analysis._solve_mg
 * (simp.py:795-826): alive = xe > ALIVE (1e-3) or an anchor cell; while
 * fewer than REDUCE_BELOW (0.8) of the cells are alive, the cells not alive
 * or not 26-connected to a live cell that touches a support (_connected,
 * simp.py:264-282) get modulus 0; a load on a dof no kept cell holds, or a
 * reduced solve that fails, sends the iteration to the whole grid (the
 * latter from a zero start). support_touch, anchor: (nx, ny, nz) uint8 per
 * cell (the cell touches a supported node; the cell touches a supported or
 * loaded node or is preserved, simp.py:154-155), host arrays; NULL for
 * either switches the reduction off. */
int lp_simp_set_reduction(lp_simp* s, const uint8_t* support_touch, const uint8_t* anchor);
/* Counters of the reduction since lp_simp_set_reduction: iterations solved
 * on the reduced system, whole-grid repeats after a failed reduced solve or
 * build, and iterations sent to the whole grid by a load on a dead dof. */
int lp_simp_loop_info(lp_simp* s, int* reduced_iters, int* full_retries, int* full_for_loads);

/* Heaviside projection simp._project / _dproject (simp.py:1443-1450):
 * out = (tanh(b e) + tanh(b (x - e))) / (tanh(b e) + tanh(b (1 - e))), dout
 * its derivative in x; either output may be NULL. */
int lp_project(long long n, const double* xt, double beta, double eta, double* out, double* dout, int device_ptr);
/* displacements of the last solve, (nrhs, 3, nodes) doubles */
int lp_simp_displacements(lp_simp* s, double* out);

/* density filter alone: out = H v (unnormalised cone stencil on the cell
 * grid, zero outside), simp.GridFilter.__matmul__ (simp.py:1164-1172). */
int lp_filter_apply(int nx, int ny, int nz, double radius, int reach, const double* v, double* out,
                    int device_ptr);
/* simp.GridFilter(domain, radius) @ v (simp.py:1148-1172) on the domain's
 * box of nx x ny x nz cells of side pitch: reach floor(radius / pitch),
 * weights radius - |d| pitch, summed in (dx, dy, dz) order, zero outside
 * the box; cells outside the mesh carry v = 0 (the caller scatters). */
int lp_grid_filter(int nx, int ny, int nz, double pitch, double radius, const double* v, double* out,
                   int device_ptr);

/* verify.nodal_stress / nodal_von_mises (verify.py:197-206, 307-337) on a
 * voxel grid of nx x ny x nz cells: each present cell's stress at its 8
 * corners, modulus * C (B u_e / pitch, thermal taken off the normal
 * strains), averaged at each node over its cells (over its counted cells,
 * at least 1, when counted is given: verify's stress_mask). element 0 "im",
 * 1 "plain". u: (nodes, 3) node-major, nodes (nx+1)(ny+1)(nz+1) in C order
 * (the dof vector of a full-box HexMesh); modulus (cells) or NULL for
 * modulus_scalar; thermal (cells) or NULL; present / counted (cells) uint8
 * or NULL (every cell). out (nodes, 6) = (sx, sy, sz, txy, tyz, tzx);
 * vm (nodes) or NULL = fem.von_mises of it (fem.py:311-314). Nodes no
 * present cell touches read 0. */
int lp_nodal_stress(int nx, int ny, int nz, double pitch, double nu, int element, const double* u,
                    const double* modulus, double modulus_scalar, const double* thermal, const uint8_t* present,
                    const uint8_t* counted, double* out, double* vm, int device_ptr);

/* verify._MaterialAverage (verify.py:214-305) applied to nodal tensors: sig
 * (nodes, 6) node-major on the (nx+1)(ny+1)(nz+1) grid (lp_nodal_stress's
 * output), fill (cells) the filled share of each cell (0 = no material;
 * NULL: every cell full). Each live cell's samples (its 8 quarter points
 * while radius <= 1.5 cells, else its centre) carry the tensor trilinear in
 * its corners and the weight fill * pitch^3 / samples; past 3.3 cells the
 * samples of each block of ceil(radius / 3.3 cells)^3 cells (counted from
 * the smallest live cell along each axis) merge into one at their weighted
 * centroid. out (nodes, 6) = each node's weighted mean over the samples
 * within radius, or its own tensor when none is. Host arrays. */
int lp_material_average(int nx, int ny, int nz, double pitch, double radius, const double* fill, const double* sig,
                        double* out);

/* The constrained loop's material-averaged stress reading and its partials
 * (constrained._Averaged, constrained.py:146-252, as the stress branch reads
 * it at 517-605; no thin walls), on a box of nx x ny x nz cells of side
 * pitch. u: ncase displacement fields (nodes, 3) node-major on the
 * (nx+1)(ny+1)(nz+1) grid; xe (cells) the eroded densities (material
 * max(xe, 0); 0 off the part); readable (nodes) the nodes read. Samples at
 * each cell's quarter points while radius <= 1.5 cells, else its centre,
 * with C B (element 0 "im", 1 "plain") on the cells that moved; bond NULL
 * or {ratio, normal x, y, z, k} (Limits.bond). kappa: in the previous
 * normalization per case, out this one (constrained._normalize at
 * iteration it). Out per case: pm, smax; rhs (ncase, nodes, 3) the adjoint
 * right-hand side (fixed dofs not zeroed) and gx (ncase, cells) the
 * explicit dg/dxe, both NULL to skip. Host arrays. */
int lp_stress_averaged(int nx, int ny, int nz, double pitch, double nu, int element, double radius, int ncase,
                       const double* u, const double* xe, const uint8_t* readable, double p, double q,
                       double limit, const double* bond, double* kappa, int it, double* pm, double* smax,
                       double* rhs, double* gx);

/* ---- kernel-level switches: mgops / femops on the device --------------- */
/* mgops.apply_masked(kind, u, a, b, cells, m, out, finish) (mgops.py:247-
 * 261): out = K (u m) over a level's cells, and with finish out = out m +
 * (1 - m) u. kind 0 "scaled" (a = Ke (24, 24), b = e (n_active)), 1
 * "weights" (a = M_q (8, 24, 24), b = w (8, n_active)), 2 "explicit" (a =
 * K (24, 24, n_active), b unused); rows (nx, ny, nz) int64 the Cells.rows
 * map (a cell's column, or -1); u, m, out (3, nx+1, ny+1, nz+1); m NULL:
 * no mask (finish needs one). dtype_bytes 4 or 8 for every field. */
int lp_cells_apply(int kind, int nx, int ny, int nz, const long long* rows, long long n_active, const void* a,
                   const void* b, const void* u, const void* m, void* out, int finish, int dtype_bytes);
/* mgops.coarse_residual(b, ax, m, cm, coarse_shape) (mgops.py:753-763):
 * rc (3, cx, cy, cz) = restrict(((b - ax) m) padded to (3, 2cx-1, 2cy-1,
 * 2cz-1)) cm; b, ax, m (3, fx, fy, fz). */
int lp_coarse_residual(const void* b, const void* ax, const void* m, const void* cm, void* rc, int fx, int fy, int fz,
                       int cx, int cy, int cz, int dtype_bytes);
/* mgops.add_correction(xc, cm, m, x) (mgops.py:766-773): x (3, sx, sy, sz)
 * += prolong(xc cm)[:, :sx, :sy, :sz] m; xc, cm (3, cx, cy, cz). */
int lp_add_correction(const void* xc, const void* cm, const void* m, void* x, int cx, int cy, int cz, int sx, int sy,
                      int sz, int dtype_bytes);
/* mgops.dot(a, b) (mgops.py:578-585): sum a b over n values, in double. */
int lp_dot(const void* a, const void* b, long long n, int dtype_bytes, double* out);
/* mgops.cast(a, dtype) (mgops.py:564-575): out = a converted, n values. */
int lp_cast(const void* a, int a_bytes, void* out, int out_bytes, long long n);
/* femops.element_energy(u, v, edof, ke) (femops.py:236-242): out (ne) =
 * sum_j sum_k (u_j ke_jk) v_k over each element's 24 dofs edof (ne, 24)
 * int64 (j outer, k inner, the reference's order); v NULL: u. */
int lp_element_energy(const double* u, const double* v, const long long* edof, long long ne, long long ndof,
                      const double* ke, double* out);

/* ---- the assembled route: fem.Stiffness (fem.py:377-526) ---------------- */
/* The CSR stiffness of the voxel mesh of the present cells of an nx x ny x
 * nz box, its layout the reference's (rows 3n+i over the mesh nodes in C
 * order, for each neighbour in fem.OFFSETS order the columns 3m+j; Dirichlet
 * rows and columns zeroed, fixed diagonal 1), built once on the host.
 * present (cells) u8; fixed ((nx+1)(ny+1)(nz+1), 3) u8 node-major; ke the
 * 24x24 element matrix at E = 1, pitch included. Out: n_dof, nnz. */
typedef struct lp_stiff lp_stiff;
int lp_stiffness_create(lp_stiff** out, int nx, int ny, int nz, const uint8_t* present, const uint8_t* fixed,
                        const double* ke, long long* n_dof, long long* nnz);
/* indptr (n_dof + 1) and indices (nnz), int64, host arrays. */
int lp_stiffness_pattern(lp_stiff* s, long long* indptr, long long* indices);
/* femops.scatter_gather (femops.py:99-278) and Stiffness.assemble: data (nnz)
 * for per-cell moduli scale (box cells, C order), each value the sum of
 * scale_c Ke blocks in the reference's order with separate roundings (equal
 * to the reference's to the bit). device_ptr = 1: scale and data are device
 * pointers. */
int lp_stiffness_assemble(lp_stiff* s, const double* scale, double* data, int device_ptr);
int lp_stiffness_destroy(lp_stiff* s);
/* n_dof and nnz of a stiffness. */
int lp_stiffness_size(lp_stiff* s, long long* n_dof, long long* nnz);

/* The assembled route's direct solve on the device: the factorization that
 * direct.DirectSolver (direct.py:158-246) and fem.Solver's rescue
 * (fem.py:586-792) run on the host, as a block-tridiagonal Cholesky of the
 * band of the stiffness in the node order perm (n_dof int64, new position ->
 * mesh dof; NULL: mesh order). The band (max |i - j| over the pattern in
 * that order) sets the block size b (>= band + 1, a multiple of 64); the
 * factor takes (2 n_dof b) doubles of device memory (bytes out). Fails with
 * LP_ERR_ARG when that does not fit. */
typedef struct lp_band lp_band;
int lp_band_create(lp_band** out, lp_stiff* s, const long long* perm, long long* band, long long* bytes);
/* Assemble the stiffness for per-cell moduli scale (ncell = nx ny nz box
 * cells, as lp_stiffness_assemble) on the device and factor it; LP_SINGULAR
 * on a non-positive pivot (a mechanism). */
int lp_band_factor(lp_band* b, const double* scale, int ncell, int device_ptr);
/* Factor the CSR values data (nnz, the stiffness's pattern) instead. */
int lp_band_factor_csr(lp_band* b, const double* data, int device_ptr);
/* x (nrhs x n_dof, right-hand side major) = K^-1 f, then refine steps of
 * iterative refinement against the CSR matrix (x += K^-1 (f - K x));
 * relres = max over the right-hand sides of |f - K x| / |f| after them. */
int lp_band_solve(lp_band* b, int nrhs, const double* f, double* x, int device_ptr, int refine, double* relres);
int lp_band_destroy(lp_band* b);

/* 26-connected components of a cell mask, scipy.ndimage.label(mask,
 * structure=ones((3,3,3))) as simp._Analysis._connected and _prune_dead use
 * it (simp.py:264-282, 1276-1282): labels (nx, ny, nz) int32, 0 off the
 * mask, components 1..count numbered by their first cell in C order.
 * Host arrays. */
int lp_label26(int nx, int ny, int nz, const uint8_t* mask, int* labels, int* count);

/* lp_label26 without a device allocation. No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). mask
 * (nx, ny, nz) uint8 and labels (nx, ny, nz) int32 are DEVICE pointers
 * (labels also holds the parent links while it runs); count is a host
 * pointer. The round flag and the numbering's per-block root counts live
 * in a static device array of 1 MiB per GPU, held until the process ends,
 * so nx * ny * nz is limited to LP_LABEL26_NOALLOC_MAX_CELLS (larger grids
 * are rejected with LP_ERR_ARG). Runs on the null stream of the current
 * device and returns when done; every call on one GPU shares that array,
 * so calls that overlap on one GPU (several host threads) race. Same
 * labels and count as lp_label26. */
#define LP_LABEL26_NOALLOC_MAX_CELLS (1 << 26)
int lp_label26_noalloc(int nx, int ny, int nz, const uint8_t* mask, int* labels, int* count);

/* ---- diagnostics -------------------------------------------------------- */
const char* lp_last_error(void);
int lp_version(void);
/* average device time (ms) of the most recent timed kernel calls: see
 * lp_time_apply; used by bench.py for the roofline figures. */
int lp_time_apply(lp_grid* g, int level, int dtype_bytes, int nrhs, int reps, double* ms_per_apply);

/* Launch profiling (no reference counterpart: measurement for the
 * roofline). While enabled, every launch of the operator at (level,
 * dtype_bytes) is bracketed by HIP events on the grid's stream (up to 8192
 * launches); lp_profile_stats returns their count, summed device time and
 * summed algorithmic bytes (node vectors read and written by the launch,
 * cell moduli, mask at one bit per dof). Returns LP_NOT_CONVERGED if
 * launches were dropped past the event pool. */
int lp_profile(lp_grid* g, int level, int dtype_bytes, int enable);
int lp_profile_stats(lp_grid* g, long long* launches, double* total_ms, double* total_bytes);

/* The method of moving asymptotes, mma.MMA (mma.py:31-214): n design
 * variables, 1 <= m <= 8 constraints; xmin, xmax host arrays (n). update
 * takes the design x, the objective's value and gradient (f0, df0 (n)), the
 * constraint values f (m, host) and gradients df (m, n), and writes the new
 * design; x, df0, df and xnew are device pointers when device_ptr is set.
 * tighten pulls the asymptotes toward the last design (mma.py:52-58).
 * get: what 0 the lower asymptotes (n), 1 the upper (n), 2 the subproblem's
 * multipliers lambda (m), 3 the iteration count (1 value). */
typedef struct lp_mma lp_mma;
int lp_mma_create(lp_mma** out, long long n, int m, const double* xmin, const double* xmax, double move, double c,
                  double asyinit, double asyincr, double asydecr);
int lp_mma_destroy(lp_mma* mma);
int lp_mma_tighten(lp_mma* mma, double factor);
/* MMA.move = move (the move limit of the next update; the constrained loop
 * halves and grows it between iterations, constrained.py:783-795). */
int lp_mma_set_move(lp_mma* mma, double move);
int lp_mma_update(lp_mma* mma, const double* x, double f0, const double* df0, const double* f, const double* df,
                  double* xnew, int device_ptr);
int lp_mma_get(lp_mma* mma, int what, double* out);

/* The autotuner (no reference counterpart): times the fine-level operator
 * of the V-cycle (FP16 MFMA on the mixed path, the vector kernel otherwise)
 * on nrhs right-hand sides for each x-chunk in {4, 8, 12, 16, 24, 32}, keeps the
 * fastest and returns it with its time per apply; on the mixed path it then
 * times the weights level's operator for 2, 4, 8 and 16 node planes per
 * block and keeps the fastest (lp_tune param 8). */
int lp_autotune(lp_grid* g, int nrhs, int reps, int* xchunk, double* ms_best);

/* Launch-shape tuning (no reference counterpart: the autotuner's knob).
 * param 0: fine-level nodes per thread along y (1 or 2); 1: fine-level
 * nodes per thread along x (x-chunk, 1..64; until set, 16, 8 or 4 from the
 * grid's size, so that a launch has about 1,024 blocks or waves; returns
 * the previous setting, 0 for that default); 2: on the FP16-MFMA levels of
 * the mixed path, fuse the Chebyshev vector passes into the operator
 * launches (1, the default) or run them as separate passes (0); 3: replay
 * the PCG's V-cycle from a captured hipGraph (1, the default with the
 * Chebyshev smoother) or launch it kernel by kernel (0); 4: the CG's beta,
 * flexible Polak-Ribiere (1), the reference's Fletcher-Reeves (0) or
 * flexible on the mixed path only (-1, the default): the FP16-rounded
 * V-cycle is not exactly symmetric; 5: on the mixed path, repeat a solve
 * that breaks down or reaches the cap with FP32 operators on every level
 * until the next hierarchy build (1, the default), do not (0), or force
 * the repeat after one iteration (2, for tests). Returns the previous
 * value. Param 6 returns the number of such repeats so far (value
 * ignored). Param 7: which FP16-MFMA level operators run as the wave-tile
 * kernel (node sums formed in registers across lanes) rather than the
 * block-tile kernel (node sums through LDS): the fine level (1, the
 * default; the faster choice on each level), both levels (2) or neither
 * (0). Param 8: node planes per block of the weights level's FP16-MFMA
 * kernel (1-64; 8 by default); value 0 returns the current one unchanged.
 * Param 9: the STALL fallback of GridElasticity._solve (mg.py:950-965) in
 * lp_solve and the device SIMP loops: the plain hierarchy gets 100 CG
 * iterations, then the strict one is built and every right-hand side solved
 * again from its start (1, the default), or the solve runs to maxiter on
 * the hierarchy as built (0: a caller that runs the fallback itself, as the
 * reference's GridElasticity does around a backend's solve); 2 caps the
 * plain hierarchy's attempt at one iteration (for tests). Param 10: the
 * strict flag, GridElasticity.strict (mg.py:561-586); the next
 * lp_set_moduli builds the hierarchy it names. Param 11 returns the number
 * of solves that moved to the strict hierarchy (value ignored). Param 12:
 * the all-FP64 path's first coarse level is applied as its cell matrices
 * sum_q w_q M_q, formed at each hierarchy build (1, the default), or
 * matrix-free as sum_q w_q (M_q u) (0); from the next lp_set_moduli.
 * Param 13 (mixed-precision grids): the outer CG in FP32 with reliable FP64
 * updates (1) or in FP64 (0, the default). Param 14 returns the number of
 * reliable updates so far (value ignored). Param 15: the reliable-update
 * threshold in thousandths (1-999; negative: the flexible beta at a
 * replaced step). Param 16: the fine level's MFMA operands in BF16 without
 * per-cell scales (1) or scaled FP16 (0, the default). Param 17: the MFMA
 * kernels run a tile's right-hand sides in consecutive blocks of one XCD
 * (1, the default) or the right-hand side as the slow block index (0).
 * Param 18: the FP32 and mixed V-cycles' first coarse level applied as its
 * formed FP32 cell matrices by the batched explicit kernel (1) or
 * matrix-free (0, the default); from the next lp_set_moduli. Param 19
 * (with param 13): restart the FP32 PCG from each recomputed FP64 residual
 * (1: iterative refinement as gpu.py:429-481's solve_fp32 does it, with
 * the inner tolerance param 15's threshold, 3 for its 3e-3) or carry the
 * direction on (0, the default). Param 20: the mixed V-cycle's coarse
 * levels applied as FP16 cell matrices with a power-of-two scale per cell,
 * formed at each hierarchy build: every coarse level, the first one's
 * sum_q w_q M_q included (2, the default), the Galerkin levels below it
 * only, the first coarse level through the FP16-MFMA weights kernel (1),
 * or FP32 cell matrices on the Galerkin levels (0); from the next
 * lp_set_moduli. Param 21: the FP64 CG's
 * direction and solution updates (p = z + beta p, x += alpha p) done by the
 * next iteration's operator launch as it stages p (1, the default; the same
 * arithmetic, bitwise the same iterates) or by a separate pass (0).
 * Param 22: the FP16 cell-matrix levels' operand staged in LDS per brick of
 * ty x tz nodes marching x-chunks of node planes (value ty * 10000 + tz *
 * 100 + x-chunk, with ty * tz <= 32 and (ty + 2)(tz + 2) <= 64; 40804,
 * 4 x 8 nodes and 4 planes, by default; the same FMAs in the same order:
 * bitwise the same products) or gathered per node (0). Param 23: the brick form only on levels of at least this many
 * bricks (512 by default; smaller levels keep the gathers). Param 24: the
 * mixed V-cycle's fused Chebyshev smoothers of degree 2 from a nonzero x:
 * the residual launch stores r only and the last step stages its operand
 * d = (dinv r) / theta from r (1, the default: one vector write and one
 * read fewer per right-hand side and level), or the residual launch stores
 * r and d (0). */
int lp_tune(lp_grid* g, int param, int value);

#ifdef __cplusplus
}
#endif
#endif /* LOADPATH_MI_H */
