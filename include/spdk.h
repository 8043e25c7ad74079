/*
 * spdk.h — C ABI of the MI355X spectral-difference kernel library (libspdk.so).
 *
 * This is the drop-in boundary for the per-time-step spectral-difference
 * update of spd_K. Each stage entry point replaces one free function of the
 * reference kernel layer declared in src/prototypes.hpp:13-105 (cited per
 * entry), with the same argument meaning; the fused entry points replace a
 * whole integrator step of src/hydro_ader.hpp:291-325.
 *
 * Conventions
 *  - All array arguments are device pointers to FP64 data in the reference's
 *    C-order layouts: SD arrays (t, var, k, j, i, kk, jj, ii) and FV arrays
 *    (var, k, j, i) (src/define.hpp:124-125). Shapes travel as plain ints.
 *  - Every call returns 0 on success or a negative SPDK_E_* code; calls are
 *    asynchronous on the context's stream unless stated otherwise.
 *  - One context per host thread; a context is not reentrant.
 */
#ifndef SPDK_H_
#define SPDK_H_
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define SPDK_OK 0
#define SPDK_E_SHAPE (-1)   /* array shapes inconsistent with the call       */
#define SPDK_E_HIP (-2)     /* a HIP runtime call or launch failed           */
#define SPDK_E_ARG (-3)     /* bad argument (null pointer, unknown enum ...) */
#define SPDK_E_NOGPU (-4)   /* no usable gfx950 device                       */
#define SPDK_E_CONFIG (-5)  /* configuration rejected (maps to exit(1))      */

/* boundary types (src/define.hpp:44) */
#define SPDK_BC_PERIODIC 0
#define SPDK_BC_GRADFREE 1
#define SPDK_BC_REFLECTIVE 2

/* operator matrices uploaded at context creation (src/hydro_ader.hpp:122-141) */
#define SPDK_M_SP_TO_FP 0  /* (p+2) x (p+1) */
#define SPDK_M_FP_TO_SP 1  /* (p+1) x (p+2) */
#define SPDK_M_DFP_TO_SP 2 /* (p+1) x (p+2) */
#define SPDK_M_SP_TO_CV 3  /* (p+1) x (p+1) */
#define SPDK_M_CV_TO_SP 4  /* (p+1) x (p+1) */
#define SPDK_M_FP_TO_CV 5  /* (p+1) x (p+2) */
#define SPDK_M_INVADER 6   /*  T x T        */
#define SPDK_M_COUNT 7

/* Shape of an SD array: (nt, nv, Nz, Ny, Nx, nz, ny, nx); N* count elements
 * including ghosts, n* points per element (structs.hpp:104-129). */
typedef struct spdk_sd_shape {
    int nt, nv, Nz, Ny, Nx, nz, ny, nx;
} spdk_sd_shape;

/* Shape of an FV array: (nv, Nz, Ny, Nx) (fv_structs.hpp:68-93). */
typedef struct spdk_fv_shape {
    int nv, Nz, Ny, Nx;
} spdk_fv_shape;

/* Run parameters the kernels need (src/global.hpp:41-58). */
typedef struct spdk_params {
    int p;              /* polynomial degree                                   */
    int active[3];      /* axis x,y,z active                                   */
    int bc[3];          /* SPDK_BC_* per axis                                  */
    int n_ader;         /* time slices T: p+1 (ADER) or 1 (SSP-RK)             */
    int rk_order;       /* SSP-RK stages (1..3) when n_ader == 1               */
    double gamma, cfl;
    double g[3];        /* constant gravity (source term)                      */
    double nu;          /* kinematic viscosity (viscous terms when > 0)        */
    double h[3];        /* element widths dx, dy, dz                           */
    double nad_tol;     /* NAD band width                                      */
    int nad_delta;      /* band scaled by local range                          */
    int nad_moore;      /* Moore (box) neighbourhood, else face neighbours     */
    int sed;            /* smooth-extrema relaxation (applied when p>1)        */
    int blending;       /* fractional theta blending                           */
    double vel[3];      /* induction advection velocity                        */
    double beta;        /* bulk viscosity coefficient (hydro/beta, -2/3 nu)    */
} spdk_params;

/* Host-side operator data copied to the device at creation. Matrices are
 * row-major, sizes as listed by SPDK_M_*; wt has n_ader entries; fv_faces[d]
 * has fv_nfaces[d] entries and fv_centers[d] fv_ncells[d] entries (the FV
 * sub-grid of axis d, structs.hpp:40-75). */
typedef struct spdk_operators {
    const double* mat[SPDK_M_COUNT];
    const double* wt;
    const double* fv_faces[3];
    const double* fv_centers[3];
    int fv_nfaces[3];
    int fv_ncells[3];
} spdk_operators;

typedef struct spdk_ctx spdk_ctx;
typedef struct spdk_mlp spdk_mlp; /* per-element MLP of the learned Picard guess */

/* ---- context ------------------------------------------------------------ */
/* Creates a context on the current HIP device. `stream` is a hipStream_t (NULL
 * = a stream created and owned by the context). Replaces the operator set-up
 * of src/hydro_ader.hpp:76-186 (matrix upload). */
int spdk_ctx_create(const spdk_params* prm, const spdk_operators* ops, void* stream, spdk_ctx** out);
int spdk_ctx_destroy(spdk_ctx* ctx);
int spdk_ctx_sync(spdk_ctx* ctx);          /* blocks until the stream drains   */
void* spdk_ctx_stream(spdk_ctx* ctx);      /* the hipStream_t kernels run on   */
const char* spdk_last_error(void);         /* message of the last failure      */
int spdk_device_info(char* name, size_t len, int* cus);

/* HIP-event timing of the dominant kernel (the Picard sweep of the fused path,
 * the flux-point interpolation of the stage path): enable, reset, read the
 * accumulated device milliseconds and launch count. */
int spdk_timing_enable(spdk_ctx* ctx, int on);
int spdk_timing_reset(spdk_ctx* ctx);
int spdk_timing_read(spdk_ctx* ctx, double* ms_total, int* launches, char* name, size_t len);

/* ---- device memory helpers (plain bytes) --------------------------------- */
int spdk_malloc(void** ptr, size_t bytes);
int spdk_free(void* ptr);
int spdk_memset0(spdk_ctx* ctx, void* ptr, size_t bytes);
int spdk_memcpy_h2d(spdk_ctx* ctx, void* dst, const void* src, size_t bytes);  /* blocking */
int spdk_memcpy_d2h(spdk_ctx* ctx, void* dst, const void* src, size_t bytes);  /* blocking */
int spdk_memcpy_d2d(spdk_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- SD stage kernels (hydro) -------------------------------------------- */
/* transform_a_to_b_1d / _slice (src/transforms.cpp:120-177): one-axis
 * contraction with matrix `mat` (SPDK_M_*) along `dim`; t_src < 0 maps every
 * slice, else slice t_src of A into slice 0 of B. */
int spdk_transform_1d(spdk_ctx* ctx, const double* A, spdk_sd_shape sa, double* B, spdk_sd_shape sb, int mat, int dim,
                      int t_src);
/* transform_a_to_b (src/transforms.cpp:99-117): x, y, z sweeps over the active
 * axes through scratch T (shape of B). When the active axes are x, x-y or
 * x-y-z with the matrix's point count along each (up to 6), all sweeps of an
 * element run in one pass and T is not touched; the result is the same bits. */
int spdk_transform(spdk_ctx* ctx, const double* A, double* B, double* T, spdk_sd_shape s, int mat);
/* compute_conservatives / compute_primitives, SD (src/hydro.cpp:47-72, :108-132) */
int spdk_conservatives(spdk_ctx* ctx, const double* W, double* U, spdk_sd_shape s);
int spdk_primitives(spdk_ctx* ctx, const double* U, double* W, spdk_sd_shape s);
/* compute_fluxes (src/hydro.cpp:167-206): Euler flux normal to `dim` */
int spdk_hydro_flux(spdk_ctx* ctx, const double* U, double* F, spdk_sd_shape s, int dim);
/* boundaries, SD traces (src/boundary.cpp:48-113): ghost interface points of
 * an array staggered along `dim` */
int spdk_fill_sd_traces(spdk_ctx* ctx, double* U, spdk_sd_shape s, int dim);
/* sd_riemann_solver (src/hydro.cpp:368-418): LLF at element faces along `dim`,
 * written in place over F on both sides */
int spdk_riemann_llf(spdk_ctx* ctx, const double* U, double* F, spdk_sd_shape s, int dim);
/* sd_riemann_solver with its viscous flag (src/hydro.cpp:368-418): the LLF
 * face fluxes as spdk_riemann_llf; with viscous != 0 both copies of each
 * face state in U are then replaced by riemann_wind(u_L, u_R, 0.5)
 * (src/hydro.cpp:265-268, :402-409), the interface state the viscous
 * gradients are taken from. */
int spdk_sd_riemann(spdk_ctx* ctx, double* U, double* F, spdk_sd_shape s, int dim, int viscous);
/* compute_gradient (src/hydro.cpp:452-482): dU (solution points, n_ader
 * slices, all variables) of the active elements from the flux-point array
 * U_fp along dim, du = sum_ll U(ll) dfp_to_sp(id, ll) / h_dim. Ghost elements
 * of dU are not written. */
int spdk_sd_gradient(spdk_ctx* ctx, const double* U_fp, spdk_sd_shape s_fp, double* dU, spdk_sd_shape s_sp, int dim);
/* compute_viscous_flux + viscous_fluxes (src/hydro.cpp:484-555): at every
 * flux point along dim (ghost elements included) the gradients dU1 (along
 * dim), dU2 (along dim+1), dU3 (along dim+2, cyclic) interpolated from the
 * solution points along dim, the primitives of U_fp there, and the viscous
 * flux (prm.nu, prm.beta) written over U_fp. */
int spdk_viscous_flux(spdk_ctx* ctx, double* U_fp, spdk_sd_shape s_fp, const double* dU1, const double* dU2,
                      const double* dU3, spdk_sd_shape s_sp, int dim);
/* sd_rusanov_solver (src/hydro.cpp:420-450): at every element face along
 * dim, riemann_wind(-0.5) of (the right element's face value, the left
 * element's) added to F on both sides. */
int spdk_sd_rusanov(spdk_ctx* ctx, const double* U_fp, double* F, spdk_sd_shape s, int dim);
/* copy_ader (src/hydro_ader.hpp:428-444): zero-order Picard guess */
int spdk_copy_ader(spdk_ctx* ctx, const double* U, double* U_ader, spdk_sd_shape s_ader);
/* update_prediction (src/transforms.cpp:263-325); Fy/Fz may be NULL on
 * inactive axes. s = shape of U (slice 0); sf[d] = shapes of the F arrays. */
int spdk_ader_predict(spdk_ctx* ctx, const double* U, double* U_ader, const double* Fx, const double* Fy,
                      const double* Fz, spdk_sd_shape s, int q, double dt);
/* update_solution (src/transforms.cpp:327-374) */
int spdk_ader_correct(spdk_ctx* ctx, double* U, const double* U_ader, const double* Fx, const double* Fy,
                      const double* Fz, spdk_sd_shape s, int q, double dt);
/* combine_solution (src/transforms.cpp:252-261): U <- a*U0 + (1-a)*U */
int spdk_rk_combine(spdk_ctx* ctx, double* U, const double* U0, spdk_sd_shape s, double a);
/* compute_dt (src/hydro.cpp:208-262): min over active elements; writes the
 * result to *dt_host (blocking). */
int spdk_min_dt(spdk_ctx* ctx, const double* W_cv, spdk_sd_shape s, double* dt_host);

/* ---- FV fallback stage kernels -------------------------------------------- */
/* face_integral (src/transforms.cpp:435-481); T is an SD scratch of the F_fp
 * shape with nt = 1 */
int spdk_fv_face_integral(spdk_ctx* ctx, const double* F_fp, spdk_sd_shape s_fp, double* F, spdk_fv_shape sf,
                          double* T, int t_id, int dim);
/* fv_update_solution (src/transforms.cpp:483-543) */
int spdk_fv_update(spdk_ctx* ctx, double* U_new, double* U_old, double* U_cv, spdk_sd_shape s_cv, const double* Fx,
                   const double* Fy, const double* Fz, spdk_fv_shape sc, int t_id, double dt, int update);
/* fv_update_solution's update = 0 call as FV_Update_solution makes it
 * (src/hydro_ader.hpp:500-586), followed by compute_primitives of U_new and
 * U_old: U_new is written, and on the active cells W_new = prim(U_new) and
 * W_old = prim(U_cv); U_old is not written. Fill the ghosts of W_new and
 * W_old (spdk_fill_fv_ghosts) where the reference fills those of U_new and
 * U_old before converting: the result is the same bits (a ghost is a copy of
 * an active cell, the reflective wall negating the normal momentum, which the
 * conversion maps to the exactly negated velocity). Six variables. */
int spdk_fv_update_prims(spdk_ctx* ctx, double* U_new, double* W_new, double* W_old, const double* U_cv,
                         spdk_sd_shape s_cv, const double* Fx, const double* Fy, const double* Fz, spdk_fv_shape sc,
                         int t_id, double dt);
/* fv_update_solution's update = 1 call as FV_Update_solution makes it
 * (src/hydro_ader.hpp:500-586): after the update = 0 call of the same slice
 * and the blend of the fallback fluxes with weights theta (FV shape, one
 * variable), with U_cv unchanged since. Cells with a nonzero theta in
 * themselves or a face neighbour are recomputed; for the others U_new from
 * the update = 0 call is the result (their face fluxes are unchanged) and
 * only U_cv = U_new is written (with gravity every cell is recomputed: the
 * update = 1 form takes the source from the density it has just written).
 * U_old, which already holds U_cv, is not written. Same bits as the
 * update = 1 call. theta has the FV cell shape sc (its first variable is
 * read). The three FV updates check that s_cv and sc agree: six variables,
 * and per active axis sc's cells = active elements x points + 2 FV ghosts
 * (SPDK_E_SHAPE otherwise). */
int spdk_fv_update_blended(spdk_ctx* ctx, double* U_new, double* U_cv, spdk_sd_shape s_cv, const double* Fx,
                           const double* Fy, const double* Fz, const double* theta, spdk_fv_shape sc, int t_id,
                           double dt);
/* FV_Update_solution (src/hydro_ader.hpp:500-575) with the cell values kept
 * in the FV cell layout (shape sc = the context's FV cells, 6 variables;
 * variable 5 is never touched) across the ADER nodes, in two buffers that
 * take turns. Same bits as the U_cv sequence above.
 * from_sd: U_fv = transform_a_to_b(U_sp, sp -> cv) of the active elements
 *   (src/hydro_ader.hpp:501), written at their FV cell positions.
 * update_prims: the update = 0 call reading U_fv: U_next, W_new =
 *   prim(U_next), W_old = prim(U_fv) (spdk_fv_update_prims); of W_new only
 *   the variables in the bit mask w_new_vars are written (bit v: variable v;
 *   trouble detection and PAD read density and pressure, 0x11).
 * update_blended: the update = 1 call: rewrites U_next only in cells whose
 *   theta or a face neighbour's theta is nonzero (every cell with gravity);
 *   the caller then swaps U_fv and U_next (U_cv = U_new without a copy).
 * to_sd: U_sp = transform_a_to_b(cv -> sp) (src/hydro_ader.hpp:570) — ghost
 *   elements by the sp -> cv -> sp round trip — and, when W_cv / dt_next are
 *   given, W_cv = cv(prim(U_sp)) and the next dt (compute_primitives,
 *   transform_a_to_b, compute_dt: src/hydro_ader.hpp:235-243).
 * The tile transforms need leading active axes (x, x-y, x-y-z), one slice and
 * at most 6 points per axis (SPDK_E_SHAPE otherwise). */
int spdk_fv_state_from_sd(spdk_ctx* ctx, const double* U_sp, spdk_sd_shape s, double* U_fv, spdk_fv_shape sc);
int spdk_fv_state_update_prims(spdk_ctx* ctx, double* U_next, double* W_new, double* W_old, const double* U_fv,
                               const double* Fx, const double* Fy, const double* Fz, spdk_fv_shape sc, int t_id,
                               double dt, int w_new_vars);
int spdk_fv_state_update_blended(spdk_ctx* ctx, double* U_next, const double* U_fv, const double* Fx, const double* Fy,
                                 const double* Fz, const double* theta, spdk_fv_shape sc, int t_id, double dt);
int spdk_fv_state_to_sd(spdk_ctx* ctx, const double* U_fv, spdk_fv_shape sc, double* U_sp, spdk_sd_shape s, double* W_cv,
                        double* dt_next);
/* boundaries, FV cell-centred fields (src/boundary.cpp:123-190), all active axes x->y->z */
int spdk_fill_fv_ghosts(spdk_ctx* ctx, double* U, spdk_fv_shape s);
/* compute_primitives, FV (src/hydro.cpp:134-153) */
int spdk_fv_primitives(spdk_ctx* ctx, const double* U, double* W, spdk_fv_shape s);
/* detect_troubles (src/trouble_detection.cpp:261-290): NAD, SED per axis,
 * relax, PAD. alpha_* are persistent per-axis scratch (zero-initialised). */
int spdk_detect_troubles(spdk_ctx* ctx, const double* W_new, const double* W_old, double* troubles, double* alpha_x,
                         double* alpha_y, double* alpha_z, spdk_fv_shape s, int pad, int limit_mask);
/* detect_troubles with the SED coefficients kept as one 16-bit word per FV
 * cell (bit v*3+d: alpha_d of variable v >= 1) instead of three FP64 arrays:
 * relax_NAD only tests whether the smallest alpha around a cell is below 1,
 * so the troubles are the same as spdk_detect_troubles' for finite states.
 * abits (FV cell count words) must be zero before the first call; mask covers
 * variables 0..4. */
int spdk_detect_troubles_bits(spdk_ctx* ctx, const double* W_new, const double* W_old, double* troubles,
                              uint16_t* abits, spdk_fv_shape s, int pad, int mask);
/* apply_blending + blending_ring, or theta_from_troubles when blending is off
 * (src/trouble_detection.cpp:191-242). theta_tmp is scratch: with all three
 * axes active both passes run in one tiled kernel that does not write it. */
int spdk_blend_theta(spdk_ctx* ctx, const double* troubles, double* theta_tmp, double* theta, spdk_fv_shape s_tr);
/* fallback_fluxes (src/hydro.cpp:753-788): MUSCL-Hancock + HLLC blended into
 * the FV face fluxes; dt_sub = w_t * dt */
int spdk_muscl_hllc_blend(spdk_ctx* ctx, const double* W, const double* theta, double* Fx, double* Fy, double* Fz,
                          spdk_fv_shape s, double dt_sub);

/* ---- induction stage kernels (staggered face B, edge E) -------------------- */
/* The reference's induction solver is 3-D only (src/induction_ader.hpp:106-110).
 * These entries also run with z inactive (one point per element along z, no
 * ghosts): a derivative along an inactive axis is zero and an interpolation
 * along it takes the value as is; with all axes active the arithmetic is the
 * reference's. */
/* rotational_a_to_b (src/induction.cpp:94-137): B_dim = d1 A2 / d1 - d2 A1 / d2 */
int spdk_B_curl(spdk_ctx* ctx, const double* A1, spdk_sd_shape s1, const double* A2, spdk_sd_shape s2, double* B,
                spdk_sd_shape sb, double d1, double d2, int dim);
/* compute_E (src/induction.cpp:139-208): edge E normal to `dim` from face B
 * interpolated sp->fp; E has 8 variables [E, b1, b2, v1, v2, 0, b1, b2] */
int spdk_induction_E(spdk_ctx* ctx, double* E, spdk_sd_shape se, const double* B1, spdk_sd_shape s1, const double* B2,
                     spdk_sd_shape s2, int dim);
/* E_riemann_solver (src/induction.cpp:509-560): upwind by max(v_L, v_R) of
 * variable v1, in place on both sides of faces along `dim` */
int spdk_E_riemann(spdk_ctx* ctx, double* E, spdk_sd_shape s, int dim, int v1);
/* update_B_prediction (src/induction.cpp:303-365), active elements */
int spdk_B_predict(spdk_ctx* ctx, const double* B, double* B_ader, spdk_sd_shape s_ader, const double* E1,
                   spdk_sd_shape s1, const double* E2, spdk_sd_shape s2, double d1, double d2, double dt, int dim);
/* update_B_solution (src/induction.cpp:367-420), active elements */
int spdk_B_correct(spdk_ctx* ctx, double* B, spdk_sd_shape sb, const double* E1, spdk_sd_shape s1, const double* E2,
                   spdk_sd_shape s2, double d1, double d2, double dt, int dim);
/* compute_B2_cv (src/induction.cpp:422-465): cv averages of B and |B|^2 */
int spdk_B2_cv(spdk_ctx* ctx, double* B2, spdk_sd_shape s, const double* Bx, spdk_sd_shape sx, const double* By,
               spdk_sd_shape sy, const double* Bz, spdk_sd_shape sz);
/* ---- fused induction step ---------------------------------------------------
 * One ADER step of the induction solver without the FV update (the stage
 * sequence copy_ader, p+1 x Solve_E with p x update_B_prediction, then
 * update_B_solution: src/induction_ader.hpp:343-359) in p+2 launches, bit-exact
 * with the stage entries: x-y (z inactive) or x-y-z active, p = 1..4 (1..3 in
 * 3-D), periodic or gradfree boundaries. Face B in the stage layout, updated
 * in place over the active elements. The workspace (size from
 * spdk_induction_fused_workspace_bytes, ghost-inclusive element counts; 0 when
 * the configuration is not covered) holds two copies of the raw edge field of
 * every slice. */
size_t spdk_induction_fused_workspace_bytes(spdk_ctx* ctx, const int N[3]);
int spdk_induction_step_fused(spdk_ctx* ctx, double* Bx, spdk_sd_shape sx, double* By, spdk_sd_shape sy, double* Bz,
                              spdk_sd_shape sz, void* workspace, double dt);
/* The same step in mixed precision (no reference counterpart): raw edge E,
 * face solves, interpolations and predictor slices in FP32; dB/dt of the
 * corrector in FP64 from the FP32 edge values and B accumulated in FP64, so
 * div B stays at FP64 round-off. Workspace: half the FP64 one. */
size_t spdk_induction_mixed_workspace_bytes(spdk_ctx* ctx, const int N[3]);
int spdk_induction_step_mixed(spdk_ctx* ctx, double* Bx, spdk_sd_shape sx, double* By, spdk_sd_shape sy, double* Bz,
                              spdk_sd_shape sz, void* workspace, double dt);
/* No reference counterpart (the reference writes B2_cv only): max |div B| at
 * the solution points of the active elements, each face component
 * differentiated fp->sp along its own axis, and the magnetic energy, the sum
 * of |B|^2 (B2 variable 3, from spdk_B2_cv) x cv volume over active elements.
 * Blocks until out[0] = max |div B| and out[1] = energy are on the host. */
int spdk_induction_diagnostics(spdk_ctx* ctx, const double* Bx, spdk_sd_shape sx, const double* By, spdk_sd_shape sy,
                               const double* Bz, spdk_sd_shape sz, const double* B2, spdk_sd_shape s2, double* out);
/* compute_B_cv_from_cf (src/induction.cpp:467-500): FV cell B from face averages */
int spdk_B_cv_from_cf(spdk_ctx* ctx, double* B, spdk_fv_shape sf, const double* Bx, spdk_sd_shape sx, const double* By,
                      spdk_sd_shape sy, const double* Bz, spdk_sd_shape sz);
/* edge_integral (src/induction.cpp:602-654): edge-averaged E of slice t_id */
int spdk_edge_integral(spdk_ctx* ctx, const double* E_ep, spdk_sd_shape s, double* E, spdk_fv_shape sf, int t_id,
                       int dim);
/* fv_update_B_solution (src/induction.cpp:656-720) */
int spdk_fv_update_B(spdk_ctx* ctx, double* B_new, double* B_old, spdk_fv_shape sn, double* B, spdk_sd_shape sb,
                     const double* E1, spdk_fv_shape s1, const double* E2, spdk_fv_shape s2, int t_id, double dt, int dim,
                     int update);
/* boundaries, FV with alignment (src/boundary.cpp:123-190): alignment 1 and
 * a_dim mark a field staggered along a_dim; all active axes x->y->z */
int spdk_fill_fv_ghosts_aligned(spdk_ctx* ctx, double* U, spdk_fv_shape s, int alignment, int a_dim);

/* ---- fused element-tile path (hydro) -------------------------------------- */
/* Device workspace for the fused path of a mesh with ghost-inclusive element
 * counts N[3]; all scratch lives in caller memory. The fused path covers the
 * inviscid update, gravity included (0 and an error for nu > 0). */
size_t spdk_fused_workspace_bytes(spdk_ctx* ctx, const int N[3]);
/* One full SD step on U_sp (1,6,Nz,Ny,Nx,n,n,n) with the plain FP64 arithmetic
 * of ADER_step/RK_step + update_solution (src/hydro_ader.hpp:291-325),
 * followed by compute_primitives, sp->cv and compute_dt. Writes the next dt
 * to *dt_next (blocking). When W_cv / F_x are non-NULL the step also writes
 * the cv primitives and the x flux-point fluxes of its last sweep (outputs).
 * Bit-identical to the stage sequence. */
int spdk_hydro_step_fused(spdk_ctx* ctx, double* U_sp, spdk_sd_shape s, void* workspace, double dt, double* W_cv,
                          double* F_x, double* dt_next);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): spdk_hydro_step_fused with its workspace taken from a
 * static per-GPU array of spdk_fused_noalloc_max_bytes() bytes (64 MiB; each
 * GPU the library runs on holds it until the process ends). The call
 * allocates nothing; meshes whose workspace (spdk_fused_workspace_bytes)
 * exceeds the array are rejected with SPDK_E_SHAPE. Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (other threads,
 * other streams) race. Same kernels and results as spdk_hydro_step_fused. */
size_t spdk_fused_noalloc_max_bytes(void);
int spdk_hydro_step_fused_noalloc(spdk_ctx* ctx, double* U_sp, spdk_sd_shape s, double dt, double* W_cv, double* F_x,
                                  double* dt_next);
/* The predictor part of ADER_step only (src/hydro_ader.hpp:291-303: copy_ader
 * and the p Solve_fluxes + update_prediction pairs) with the fused tile
 * kernel, bit-identical to the stage sequence, its T slices written to
 * U_ader in the stage layout (T, >= 5 variables, elements, points). The
 * caller then runs the last flux evaluation and the FV update with the
 * fallback (src/hydro_ader.hpp:500-586) through the stage entries. */
int spdk_fused_predict(spdk_ctx* ctx, double* U_sp, spdk_sd_shape s, void* workspace, double dt, double* U_ader,
                       spdk_sd_shape s_ader);
/* The predictor part of ADER_step and the last Solve_fluxes
 * (src/hydro_ader.hpp:278-303) with the fused tile kernel: the final sweep
 * writes, instead of the corrector, the flux-point fluxes of every slice and
 * active axis (Riemann fluxes at element faces included) to F_x / F_y / F_z in
 * the stage flux-point layout of s_ader with p + 2 points along the axis
 * (slots past the physical variables zero). Bit-identical to the stage
 * sequence; the caller then runs the FV update with the fallback
 * (src/hydro_ader.hpp:500-586). Uses the fused workspace. */
int spdk_fused_predict_fluxes(spdk_ctx* ctx, double* U_sp, spdk_sd_shape s, void* workspace, double dt, double* F_x,
                              double* F_y, double* F_z, spdk_sd_shape s_ader);
/* As spdk_fused_predict_fluxes, with the face_integral of every slice and
 * active axis (src/transforms.cpp:378-481, called per ADER node by
 * src/hydro_ader.hpp:500-510) formed in the same sweep: F_fv[t * 3 + d]
 * receives the sub-cell face fluxes of slice t along axis d in the FV
 * layout of the cell shape sc (6 variables, ghost layers included) with one
 * more face along d; only the faces face_integral writes are written
 * (variable 5 is left as it is). F_x (optional, may be NULL) receives the x
 * flux-point fluxes in the stage layout of s_ader (the F_ader_fp_x output).
 * Bit-identical to the stage flux evaluation followed by
 * spdk_fv_face_integral. */
int spdk_fused_predict_fv_fluxes(spdk_ctx* ctx, double* U_sp, spdk_sd_shape s, void* workspace, double dt,
                                 double* const* F_fv, spdk_fv_shape sc, double* F_x, spdk_sd_shape s_ader);
/* Solve_fluxes of given ADER slices (src/hydro_ader.hpp:278-289: sp->fp,
 * fluxes, trace fill, LLF) with the fused tile kernel: U_ader in the stage
 * layout (T, >= 5 variables, elements, points), fluxes written as by
 * spdk_fused_predict_fluxes. The workspace holds the slices' face traces:
 * spdk_fused_fluxes_workspace_bytes(N) bytes (0: no tile variant). */
size_t spdk_fused_fluxes_workspace_bytes(spdk_ctx* ctx, const int N[3]);
int spdk_fused_fluxes(spdk_ctx* ctx, const double* U_ader, spdk_sd_shape s_ader, void* workspace, double* F_x,
                      double* F_y, double* F_z);

/* ---- mixed-precision ADER path (hydro) ------------------------------------ */
/* The ADER step of src/hydro_ader.hpp:291-305 with the space-time state, its
 * face traces and the transform / divergence / predictor contractions in
 * FP32 (as deviations from each element's FP64 mean), and fluxes, Riemann
 * solves, the corrector accumulation, cv transform and dt in FP64. Constant
 * gravity is a source in both: the predictor's in the FP32 deviations (the
 * mean's part rounded once from FP64), the corrector's summed in FP64 from
 * the final slices; nu > 0 is refused. Not bit-identical to the reference;
 * its error against the FP64 path is measured per configuration. No
 * reference counterpart beyond the step it replaces. */
#define SPDK_GUESS_ZERO 0     /* U_ader = U at every node (copy_ader)             */
#define SPDK_GUESS_EXTERNAL 1 /* caller-filled predictor buffer                   */
#define SPDK_GUESS_MLP 2      /* the network attached with spdk_mixed_attach_mlp  */
typedef struct spdk_picard {
    int residual;   /* 0: p predictor sweeps then the corrector (the reference's
                       count); 1: sweep the elements that have not passed the
                       residual test (and the elements reading their traces)
                       until all pass or reach the cap                       */
    int guess;      /* SPDK_GUESS_*                                            */
    double tol;     /* residual test, per interior element:
                       max|d_out - d_in| / max_v|mean_v| < tol                 */
    int max_sweeps; /* cap of predictor sweeps per element; <= 0: the fixed
                       count p. An element reaching the cap keeps its state
                       (from the zero-order guess and the cap p, that is the
                       fixed-count predictor); from an external or learned
                       guess it restarts from the zero-order guess with the
                       fixed count p, unless guess_keep                        */
    int guess_keep; /* external or learned guess: 1 = an element reaching the
                       cap keeps its state, as from the zero-order guess (no
                       restart). Loses accuracy where the guess is poor: the
                       KH-trained network on the 3-D sine wave gives 130x the
                       FP64 error with it, 1.001x with the restart            */
} spdk_picard;
typedef struct spdk_picard_stats {
    long long steps;     /* accumulated by spdk_hydro_step_mixed                 */
    long long fallbacks; /* steps in which some element reached the cap
                            unconverged (restarted, from a guess)              */
    double sweeps;       /* sweeps, each weighted by the fraction of elements
                            it updated (the corrector counts 1)               */
    double last_residual;
} spdk_picard_stats;
size_t spdk_mixed_workspace_bytes(spdk_ctx* ctx, const int N[3]);
/* Element means, FP32 deviations and their traces for U_sp; must precede a
 * step with SPDK_GUESS_EXTERNAL (the guess is a deviation from these means). */
int spdk_mixed_prepare(spdk_ctx* ctx, const double* U_sp, spdk_sd_shape s, void* workspace);
/* Pointers into the workspace: the (T,5,E,TPE) FP32 predictor buffer a guess
 * is written to, the (5,E) FP64 element means and the (5,E,TPE) FP32
 * deviations of U_sp. Any may be NULL. */
int spdk_mixed_predictor_buffer(spdk_ctx* ctx, void* workspace, const int N[3], float** D, double** mu, float** d_old);
/* One step; pc NULL means the fixed count from the zero-order guess. Writes
 * the next dt to *dt_next (blocking; the residual form also synchronises
 * once per sweep to read the residual). */
int spdk_hydro_step_mixed(spdk_ctx* ctx, double* U_sp, spdk_sd_shape s, void* workspace, double dt, double* W_cv,
                          double* F_x, const spdk_picard* pc, spdk_picard_stats* stats, double* dt_next);
/* The predictor part of that step only (the Picard iteration of
 * src/hydro_ader.hpp:291-305 up to, not including, the last flux evaluation
 * and update_solution): the guess (zero-order or learned) and the fixed
 * count or the residual test as pc says; the predictor buffer then holds the
 * space-time predictor for spdk_mixed_export_predictor, ahead of the FP64
 * flux evaluation and the FV update with trouble detection and the MUSCL
 * fallback (FV_Update_solution, src/hydro_ader.hpp:500-586). U_sp is read,
 * not written. Stats count the sweeps plus 1 for that flux evaluation. */
int spdk_mixed_predict(spdk_ctx* ctx, double* U_sp, spdk_sd_shape s, void* workspace, double dt, const spdk_picard* pc,
                       spdk_picard_stats* stats);
/* spdk_mixed_predict followed, in the same call, by the last flux evaluation
 * (src/hydro_ader.hpp:278-289: face Riemann fluxes in FP64, interior
 * flux-point fluxes as FP32 deviations from the element mean's flux) and the
 * face_integral of every slice and active axis (src/transforms.cpp:378-481;
 * transverse contraction in FP32, on MFMA at 3-D p=3 unless valu != 0, the
 * mean flux added in FP64): F_fv[t * 3 + d] receives the sub-cell face fluxes
 * of slice t along axis d in the FV layout of the cell shape sc (6
 * variables, ghost layers included, one more face along d); each face is
 * written once, so the FV update on them conserves exactly. F_x (optional)
 * receives the x flux-point fluxes (F_ader_fp_x output). The FV update with
 * trouble detection and the fallback (src/hydro_ader.hpp:510-586) follows
 * through the stage entries. No counterpart for the mixed arithmetic. */
int spdk_mixed_predict_fv_fluxes(spdk_ctx* ctx, double* U_sp, spdk_sd_shape s, void* workspace, double dt,
                                 const spdk_picard* pc, spdk_picard_stats* stats, double* const* F_fv,
                                 spdk_fv_shape sc, double* F_x, int valu);
/* Histogram of the predictor sweeps each element needed to pass the residual
 * test, accumulated over residual-form steps (bin k = k sweeps counted from
 * the step's first sweep, a restart included; bin 31 = never). Copies nbins (<= 32) counts; reset clears. */
int spdk_mixed_histogram(spdk_ctx* ctx, void* workspace, const int N[3], unsigned long long* hist, int nbins,
                         int reset);
/* Learned guess. Feature and target sizes of this degree and dimension:
 * din = (d+2) n^d + 6 and dout = T (d+2) n^d (the density, the active
 * momenta and the energy of every point; targets for every time node). */
int spdk_mixed_guess_dims(spdk_ctx* ctx, int* din, int* dout);
/* Attach (NULL: detach) the network used by SPDK_GUESS_MLP; it is moved to the
 * context's stream. Sizes must match spdk_mixed_guess_dims. */
int spdk_mixed_attach_mlp(spdk_ctx* ctx, spdk_mlp* net);
/* Training data: nsweeps Picard sweeps from the zero-order guess without the
 * corrector (the predictor buffer then holds the nsweeps-th iterate), and the
 * rows [e0, e0 + nrows) of features (X, nrows x din) and of targets from the
 * current predictor buffer (Y, nrows x dout); X or Y may be NULL. */
int spdk_mixed_iterate(spdk_ctx* ctx, double* U_sp, spdk_sd_shape s, void* workspace, double dt, int nsweeps);
/* The predictor buffer (after spdk_mixed_iterate or spdk_mixed_predict) as FP64 space-time slices in
 * the stage layout of U_ader (T, >= 5 variables, elements, points; the
 * reference's U_ader_sp, src/hydro_ader.hpp:291-305): U = mean + deviation,
 * further variable slots 0. With it the FV update with trouble detection and
 * the MUSCL fallback (spdk_fv_* entries) runs on a mixed-precision predictor. */
int spdk_mixed_export_predictor(spdk_ctx* ctx, void* workspace, double* U_ader, spdk_sd_shape s_ader);
int spdk_mixed_guess_data(spdk_ctx* ctx, void* workspace, const int N[3], double dt, int e0, int nrows, float* X,
                          long ldx, float* Y, long ldy);

/* ---- per-element MLP for the Picard initial guess -------------------------- */
/* No reference counterpart: the learned guess that replaces copy_ader
 * (src/hydro_ader.hpp:428-444) where the caller asks for it. An MLP
 * in -> h1 -> h2 -> out with GELU on the hidden layers, BF16 MFMA products and
 * FP32 accumulation, FP32 master weights and Adam. Dimensions are padded to
 * multiples of 64 internally; all row arrays below are device FP32 with the
 * given leading dimension (elements). Runs on `stream` (hipStream_t, NULL =
 * the null stream). */
int spdk_mlp_create(const int dims[4], int batch_cap, void* stream, spdk_mlp** out);
int spdk_mlp_destroy(spdk_mlp* net);
int spdk_mlp_dims(spdk_mlp* net, int padded[4]);
int spdk_mlp_set_stream(spdk_mlp* net, void* stream);
/* Parameters as one host FP32 array: W1 (h1 x in, row-major), b1, W2, b2, W3, b3. */
size_t spdk_mlp_param_count(spdk_mlp* net);
int spdk_mlp_set_params(spdk_mlp* net, const float* host); /* also resets Adam */
int spdk_mlp_get_params(spdk_mlp* net, float* host);
int spdk_mlp_forward(spdk_mlp* net, const float* X, long ldx, int B, float* Y, long ldy);
/* Gradients of mean((f(X) - T)^2) (mean over B x out) in the parameter layout. */
int spdk_mlp_gradients(spdk_mlp* net, const float* X, long ldx, const float* T, long ldt, int B, float* host_grad,
                       float* loss);
int spdk_mlp_train_step(spdk_mlp* net, const float* X, long ldx, const float* T, long ldt, int B, float lr,
                        float* loss);
/* dst[r][:] = src[idx[r]][:] (device rows of `cols` floats; idx device ints). */
int spdk_gather_rows(const float* src, long ld, const int* idx, int n, int cols, float* dst, void* stream);
/* flags[r] (device bytes) = 1 if every value of row r is finite and below maxabs in magnitude. */
int spdk_rows_finite(const float* Y, long ld, int rows, int cols, float maxabs, unsigned char* flags, void* stream);

/* ---- whole-run driver ------------------------------------------------------ */
typedef struct spdk_run_result {
    int steps;                 /* completed time steps                           */
    long long dofs;            /* active solution points                         */
    double evolution_seconds;  /* evolution loop, IO excluded                    */
    double dof_updates_per_s;  /* dofs * steps / evolution_seconds               */
    double t_final;            /* simulated time reached                         */
    double kernel_ms_total;    /* device time of the dominant kernel (HIP events)*/
    int kernel_launches;       /* launches of that kernel                        */
    char kernel_name[64];
    long long picard_steps;    /* mixed path: steps, element-weighted sweeps,    */
    double picard_sweeps;      /*   zero-guess fallbacks and the histogram of    */
    long long picard_fallbacks;/*   spdk_mixed_histogram (residual form)         */
    long long picard_hist[32];
} spdk_run_result;

/* Runs the host driver (src/main.cpp:86-257) on an argv-style command line:
 * "-i file.athinput" plus "block/name=value" overrides; outputs go to outdir
 * (SPD_OUTPUT_DIR semantics). Extra keys: job/path = fused|stages,
 * job/precision = fp64|mixed, job/picard = fixed|residual, job/picard_tol,
 * job/picard_max, job/guess = zero|mlp, job/mlp_weights, time/nlim, job/warmup,
 * job/timing. */
int spdk_run_argv(int argc, const char* const* argv, const char* outdir, spdk_run_result* res);

#ifdef __cplusplus
}
#endif
#endif
