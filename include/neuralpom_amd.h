/*
 * neuralpom_amd -- C ABI of the MI355X explicit POM step (forward + adjoint).
 *
 * The boundary the reference's Python core would bind to (through ctypes) in
 * place of its eager PyTorch operators.  Plain pointers, sizes and a HIP
 * stream; no torch types.  All fields use the reference's layout:
 *   3-D fields (im, jm, 16) row-major, k fastest  (src/neuralpom/cores/explicit/base.py:162-174)
 *   2-D fields (im, jm) row-major                  (base.py:120-160)
 * An ensemble packs members back to back: member m's 3-D field starts at
 * element off3[m], its 2-D field at off2[m] (one allocation per field).
 *
 * Every entry point returns 0 on success or a positive error code:
 *   POM_E_ARG (invalid argument), POM_E_HIP (HIP runtime error),
 *   POM_E_UNSUPPORTED (configuration not supported by this build).
 */
#ifndef NEURALPOM_AMD_H
#define NEURALPOM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POM_OK 0
#define POM_E_ARG 1
#define POM_E_HIP 2
#define POM_E_UNSUPPORTED 3
#define POM_E_VMAX 4
#define POM_E_SCRATCH 5   /* input larger than an entry point's static scratch */

#define POM_F32 0
#define POM_F64 1

/* Dynamic state fields of one internal step.  Names are the reference's
 * POM2K_Core attribute names (explicit/base.py:120-189); the step reads the
 * *_in table and writes the *_out table (src/neuralpom/cores/explicit/
 * time_loop.py:20-709 updates them in place or by rebinding). */
typedef enum {
  /* 3-D */
  POM_U = 0, POM_UB, POM_V, POM_VB, POM_T, POM_TB, POM_S, POM_SB,
  POM_Q2, POM_Q2B, POM_Q2L, POM_Q2LB, POM_KM, POM_KH, POM_KQ, POM_AAM, POM_RHO, POM_W, POM_L,
  POM_ADVX, POM_ADVY, POM_DRHOX, POM_DRHOY,
  POM_N3D,
  /* 2-D */
  POM_EL = POM_N3D, POM_ELB, POM_ELF, POM_UA, POM_UAB, POM_UAF, POM_VA, POM_VAB, POM_VAF,
  POM_D, POM_DT, POM_ET, POM_ETB, POM_ETF, POM_EGF, POM_EGB, POM_UTF, POM_UTB, POM_VTF, POM_VTB,
  POM_WUBOT, POM_WVBOT, POM_ADVUA, POM_ADVVA, POM_ADX2D, POM_ADY2D, POM_DRX2D, POM_DRY2D, POM_AAM2D,
  POM_FLUXUA, POM_FLUXVA, POM_TPS,
  POM_E_ATMOS, POM_VFLUXF, POM_VFLUXB, POM_SWRAD, POM_WTSURF, POM_WSSURF,
  POM_NSTATE
} pom_field;

/* Static geometry and forcing (not differentiated).  Entries from
 * POM_NGEOM_BASE on belong to the open-boundary tide policy
 * (src/neuralpom/cores/case_adapters/coastal_boundary.py:9-67, implicit family)
 * and are read only when prm->boundary == POM_BOUNDARY_TIDE (may be null
 * otherwise):
 *   POM_G_TIDE_NODE  1 at tide nodes, 0 elsewhere
 *   POM_G_TIDE_AMP   tide amplitude at the nodes (m)
 *   POM_G_TIDE_PHASE tide phase at the nodes, in radians (phase_deg * pi / 180)
 *   POM_G_ETA_FIXED  1 at interior points the eta solve holds fixed (tide
 *                    nodes, dry cells); the edges are always held
 *   POM_G_TBND, POM_G_SBND  exterior T / S of open edges, (im, jm, 16) with
 *                    each edge cell holding its tbe/tbw/tbn/tbs value
 *                    (explicit/base.py:181-189; j edges take precedence at corners) */
typedef enum {
  POM_G_H = 0, POM_G_DX, POM_G_DY, POM_G_ART, POM_G_ARU, POM_G_ARV, POM_G_FSM, POM_G_DUM, POM_G_DVM,
  POM_G_COR, POM_G_CBC, POM_G_WUSURF, POM_G_WVSURF, POM_G_TSURF, POM_G_SSURF, /* 2-D */
  POM_G_RMEAN, POM_G_TCLIM, POM_G_SCLIM,                                      /* 3-D */
  POM_NGEOM_BASE,
  POM_G_TIDE_NODE = POM_NGEOM_BASE, POM_G_TIDE_AMP, POM_G_TIDE_PHASE, POM_G_ETA_FIXED, /* 2-D */
  POM_G_TBND, POM_G_SBND,                                                              /* 3-D */
  POM_NGEOM
} pom_geom;

typedef struct {
  int32_t im, jm;
  int64_t off2, off3;
} pom_member_t;

#define POM_CORE_EXPLICIT 0
#define POM_CORE_IMPLICIT 1
#define POM_PRECOND_JACOBI 0
#define POM_PRECOND_SPECTRAL 1
#define POM_BOUNDARY_CLOSED 0   /* closed basin (implicit_p0_fixed/boundary/closed_basin.py) */
#define POM_BOUNDARY_TIDE 1     /* open edges with tide nodes (case_adapters/coastal_boundary.py:9-67) */

/* Scalars of the step (explicit/base.py:46-118, v1/core_spec.py:8-47).
 * The semi-implicit core (core = POM_CORE_IMPLICIT) also reads the eta-solver
 * settings (implicit_p0_fixed/core.py:62-119, numerics/double_gyre.py:186-191):
 * eta_max_iter, eta_precond, eta_tol (absolute), eta_rel_tol, and the
 * post-solve Shapiro weight gamma_baro (0 in production, core_spec.py:24-29),
 * and the lateral boundary policy: boundary = POM_BOUNDARY_CLOSED, or
 * POM_BOUNDARY_TIDE with tide_wt for the open-boundary tide policy of the case
 * adapter (coastal_boundary.py:9-67): tide elevation at the nodes, radiation
 * conditions on the depth-mean transports with rfe rfw rfn rfs, the eta solve
 * holding the edges, dry cells and nodes fixed, and the open forms of bcond
 * 3-6 (kernels/coastal_explicit_baroclinic.py:496-928, exterior T / S from
 * POM_G_TBND / POM_G_SBND).  The explicit core takes POM_BOUNDARY_TIDE as the
 * adapter's explicit family (coastal_boundary.py:11-30): on every substep the
 * nodes take the tide at the substep's time (tide_coef, tide_sec,
 * tide_substep0 below) and the other edge points keep their previous
 * elevation, the radiation conditions act on uaf / vaf, and bcond 3-6 take the
 * open forms; that family is forward-only (the adjoint entry points return
 * POM_E_UNSUPPORTED, as the reference validates no adjoint for it,
 * case_adapters/registry.py:39-40). */
typedef struct {
  double dti, dte, dti2, dte2, ispi, isp2i;
  double alpha, smoth, horcon, grav, rhoref, kappa, umol, tprni, small;
  double tbias, sbias, hmax, rfe, rfw, rfn, rfs;
  int32_t isplit, ispadv, ntp, nbct, nbcs, iext;
  double vmaxl;
  int32_t core, eta_max_iter, eta_precond, boundary;
  double eta_tol, eta_rel_tol, gamma_baro;
  double tide_wt;   /* tide policy: (speed_deg_per_h * pi / 180) * t_s / 3600 of the step */
  /* eta_capture = 1: the eta solves launch exactly eta_max_iter iterations and
   * never synchronise with the host (the reference's capture-safe solver,
   * implicit_p0_fixed/core.py:504-555, runtime_cuda_graph_fast_path); a member
   * that has converged keeps its iterate (its kernels are no-ops), so results
   * equal the default mode's.  0: the host reads the device's count of
   * unconverged members every few iterations and stops when it is zero. */
  int32_t eta_capture, reserved0;
  /* spatial Shapiro filters of the semi-implicit core's closed-basin numerics
   * (implicit_p0_fixed/numerics/double_gyre.py:78-184): weights of the new
   * T, S and of the filtered internal velocity (0 in production,
   * v1/core_spec.py:24-29); read only with core = POM_CORE_IMPLICIT and the
   * closed basin. */
  double tracer_shapiro, velocity_shapiro;
  /* lower bound of the semi-implicit core's Smagorinsky viscosity
   * (implicit_p0_fixed/time_loop.py:186-188; 0 = none, the production value) */
  double aam_min;
  /* tide policy, explicit core: substep n of the step (1..isplit) sets the
   * nodes to amp cos(tide_coef * sec_n / 3600 - phase) with
   *   sec_n = tide_sec + dte * (tide_substep0 + n - 1),
   * tide_coef = speed_deg_per_h * pi / 180, tide_sec = time * 86400 - dti of
   * the step and tide_substep0 the adapter's substep counter before the step
   * (coastal_boundary.py:9-14; the counter is 0 after install and after
   * every reset of the case registry, registry.py:57-58) */
  double tide_coef, tide_sec;
  int32_t tide_substep0, reserved1;
} pom_params_t;

typedef struct pom_batch pom_batch;

/* Member table of an ensemble.  Copies the table to the device (one small
 * allocation owned by the handle).  dtype: POM_F32 or POM_F64.  Each member
 * is at least 4 x 4 columns, one 3-D field of a member is below 2 GiB
 * (im * jm * 16 * sizeof(element) < 2^31) and one 3-D field of the whole
 * batch below 4 GiB (the kernels address a batch field with 32-bit byte
 * offsets from its base pointer); otherwise POM_E_ARG. */
int pom_batch_create(const pom_member_t* members, int32_t n_members, int32_t dtype, pom_batch** out);
int pom_batch_destroy(pom_batch* b);

/* Launch shapes of a batch's point kernels.  A launch over n blocks of 256
 * points takes n workgroups, or, once n exceeds *_min_groups, ceil(n / bpg)
 * of them (at least *_min_groups), each striding over the blocks:
 *   map2_*:  the fused barotropic substep (default 4 blocks, 768 minimum);
 *   flat2_*: 2-D maps and adjoint gathers (default 1, 768);
 *   flat3_*: 3-D maps and adjoint gathers (default 1, 768).
 * Results do not depend on them.  Every value is in [1, 2^20].  The tuner
 * (neuralpom_amd/autotune.py) stores the best shapes per workload in
 * neuralpom_amd/tuning.json, which the Python ensembles apply.  No
 * counterpart in the reference (its kernels are torch's). */
typedef struct {
  int32_t map2_blocks_per_group, map2_min_groups;
  int32_t flat2_blocks_per_group, flat2_min_groups;
  int32_t flat3_blocks_per_group, flat3_min_groups;
} pom_tuning_t;
int pom_batch_tune(pom_batch* b, const pom_tuning_t* t);
int pom_batch_get_tuning(const pom_batch* b, pom_tuning_t* t);

/* Workspace bytes needed by pom_explicit_step (keep=0) and by
 * pom_explicit_step_adj.  With the tide policy (prm->boundary =
 * POM_BOUNDARY_TIDE) the adjoint workspace, pom_explicit_step_adj,
 * pom_explicit_step_taped and pom_explicit_step_reverse return
 * POM_E_UNSUPPORTED (a forward-only family). */
int pom_explicit_workspace(const pom_batch* b, const pom_params_t* prm, int32_t adjoint, size_t* bytes);

/* One internal step (time_loop.py:75-696) for every member.
 *   ramp:    per-member ramp factor (device, n_members)  (time_loop.py:79-88)
 *   vgrid:   per-member (z, zz, dz, dzz), device [n_members][4][16]
 *   internal: 1 to run the 3-D internal mode, 0 for the first step from
 *            time0 = 0 (time_loop.py:438)
 *   vamax_out: device scalar receiving max |vaf| over all substeps
 * Replaces POM2K_Core.run_time_loop's loop body; entries of state_out that the
 * step leaves unchanged are copied from state_in.  The state_out arrays are
 * distinct from each other and from every state_in and geom array (the step
 * reads its inputs after writing outputs). */
int pom_explicit_step(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                      const void* const* state_in, void* const* state_out, const void* const* geom,
                      int32_t internal, void* vamax_out, void* workspace, size_t ws_bytes, void* stream);

/* Vector-Jacobian product of one step: recomputes the forward from state_in
 * into the workspace, then accumulates
 *   grad_in[f] += (d state_out / d state_in[f])^T cot_out
 * for every field f with a non-null grad_in[f] (zero-initialised by the caller
 * or holding a running sum).  Null cot_out[f] entries are zero cotangents.
 * The cot_out arrays are read only and are read until the call's last
 * kernel: they must not overlap any grad_in array.
 * Differentiable inputs: u ub v vb t tb s sb q2 q2b q2l q2lb km kh kq aam rho
 * el elb ua uab va vab d dt et etb egb utf utb vtf vtb wubot wvbot. */
int pom_explicit_step_adj(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                          const void* const* state_in, const void* const* geom, int32_t internal,
                          const void* const* cot_out, void* const* grad_in, void* workspace, size_t ws_bytes,
                          void* stream);

/* The same vector-Jacobian product in two calls, as autograd's forward and
 * backward split it (the reference differentiates its step with torch
 * autograd, which keeps the forward's intermediates for the backward,
 * src/neuralpom/cores/explicit/time_loop.py:20-709 under torch.enable_grad):
 *   pom_explicit_step_taped: the step of pom_explicit_step (same outputs and
 *     vamax_out) that also leaves its intermediates in the workspace
 *     (pom_explicit_workspace with adjoint = 1);
 *   pom_explicit_step_reverse: the products of pom_explicit_step_adj for the
 *     last taped step in that workspace, without recomputing the forward;
 *     state_out is the taped call's (the reverse reads, never writes, it).
 * Between the two calls the workspace, state_in, state_out and geom must be
 * left as they were and prm, ramp, vgrid and internal must be the same.
 * Results equal pom_explicit_step and pom_explicit_step_adj bitwise. */
int pom_explicit_step_taped(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                            const void* const* state_in, void* const* state_out, const void* const* geom,
                            int32_t internal, void* vamax_out, void* workspace, size_t ws_bytes, void* stream);
int pom_explicit_step_reverse(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                              const void* const* state_in, const void* const* state_out, const void* const* geom,
                              int32_t internal, const void* const* cot_out, void* const* grad_in, void* workspace,
                              size_t ws_bytes, void* stream);

/* ---- Semi-implicit (P0-fixed) core: src/neuralpom/cores/implicit_p0_fixed/
 * time_loop.py:38-688 with the closed-basin double-gyre policy
 * (boundary/closed_basin.py, forcing/surface.py, numerics/double_gyre.py).
 * prm->core must be POM_CORE_IMPLICIT.  The barotropic mode is predictor ->
 * eta solve (batched PCG, Jacobi or spectral DST-I preconditioner, core.py:
 * 460-773) -> corrector; every member stops at its own convergence.
 *
 * pom_implicit_prepare: builds the spectral-preconditioner cache (sine
 *   matrices and inverse eigenvalues at the interior medians of dx, dy, h,
 *   core.py:730-760) from the member geometry; synchronises the stream.
 *   Replaces POM2K_Core._get_spectral_eta_preconditioner_cache and the
 *   acceleration cache of core.py:121-199 (run once after initialisation).
 * pom_implicit_step: one internal step (replaces the loop body of
 *   implicit_p0_fixed/time_loop.py:95-688).  eta_iters_out (device, may be
 *   null) receives each member's PCG iteration count.  The vamax check is a
 *   device scalar as in pom_explicit_step.
 * pom_implicit_step_adj: vector-Jacobian product of one step with the
 *   reference's eta-solve adjoint (second PCG on the area-weighted transpose
 *   plus the depth term, core.py:24-59, 231-252, 645-675).  Same conventions
 *   as pom_explicit_step_adj. */
int pom_implicit_cache_bytes(const pom_batch* b, size_t* bytes);
int pom_implicit_prepare(const pom_batch* b, const pom_params_t* prm, const void* const* geom, void* cache,
                         size_t cache_bytes, void* stream);
int pom_implicit_workspace(const pom_batch* b, const pom_params_t* prm, int32_t adjoint, size_t* bytes);
int pom_implicit_step(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* ramp,
                      const void* vgrid, const void* const* state_in, void* const* state_out, const void* const* geom,
                      int32_t internal, void* vamax_out, int32_t* eta_iters_out, void* workspace, size_t ws_bytes,
                      void* stream);
int pom_implicit_step_adj(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* ramp,
                          const void* vgrid, const void* const* state_in, const void* const* geom, int32_t internal,
                          const void* const* cot_out, void* const* grad_in, void* workspace, size_t ws_bytes,
                          void* stream);
/* pom_implicit_step_taped / pom_implicit_step_reverse: the split of
 * pom_implicit_step_adj, with the conventions of pom_explicit_step_taped /
 * pom_explicit_step_reverse (cache unchanged between the calls as well). */
int pom_implicit_step_taped(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* ramp,
                            const void* vgrid, const void* const* state_in, void* const* state_out,
                            const void* const* geom, int32_t internal, void* vamax_out, int32_t* eta_iters_out,
                            void* workspace, size_t ws_bytes, void* stream);
int pom_implicit_step_reverse(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* ramp,
                              const void* vgrid, const void* const* state_in, const void* const* state_out,
                              const void* const* geom, int32_t internal, const void* const* cot_out,
                              void* const* grad_in, void* workspace, size_t ws_bytes, void* stream);

/* The semi-implicit core's elliptic operator and eta solve on their own, for
 * callers that invoke or override them outside the step (case adapters,
 * src/neuralpom/cores/case_adapters/coastal_boundary.py:62-64; tests).
 * geom is the POM_NGEOM table (dx, dy, art, dum, dvm are read); the operator
 * coefficient is 4 dti^2 grav alpha from prm.
 *
 * pom_elliptic_apply: out = A(d) eta = eta - coeff div(d grad eta) / art on
 *   interior cells, out = eta on the edge ring.  Replaces
 *   POM2K_Core.apply_elliptic_operator (implicit_p0_fixed/core.py:591-643).
 * pom_elliptic_apply_adj: grad_eta += A(d)^T cot_out and
 *   grad_d += (d(A(d) eta)/dd)^T cot_out (either may be null).
 * pom_eta_workspace: workspace bytes of pom_eta_solve / pom_eta_solve_adj.
 * pom_eta_solve: x = A(d)^-1 rhs by the step's batched PCG (core.py:460-589,
 *   solve_implicit_eta_cg :775-793): start from the interior of x0 (null:
 *   zero), rhs edge values ignored, x zero on the edge ring; absolute
 *   tolerance prm->eta_tol, relative prm->eta_rel_tol, at most
 *   prm->eta_max_iter iterations, preconditioner prm->eta_precond (the
 *   spectral one needs the pom_implicit_prepare cache).  iters_out (device
 *   int32 [n_members], may be null) receives each member's iteration count.
 * pom_eta_solve_adj: backward of the solve (_ImplicitEtaSolveFunction,
 *   core.py:24-59, 231-252, 645-675): writes grad_rhs = lambda =
 *   area A^-1 (cot_x / area) (max-normalised right-hand side, absolute
 *   tolerance scaled alike) and accumulates grad_d += -(d(A x)/dd)^T lambda
 *   (grad_d may be null); x is the forward solution. */
int pom_elliptic_apply(const pom_batch* b, const pom_params_t* prm, const void* const* geom, const void* eta,
                       const void* d, void* out, void* stream);
int pom_elliptic_apply_adj(const pom_batch* b, const pom_params_t* prm, const void* const* geom, const void* eta,
                           const void* d, const void* cot_out, void* grad_eta, void* grad_d, void* stream);
int pom_eta_workspace(const pom_batch* b, size_t* bytes);
int pom_eta_solve(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* const* geom,
                  const void* rhs, const void* d, const void* x0, void* x, int32_t* iters_out, void* workspace,
                  size_t ws_bytes, void* stream);
int pom_eta_solve_adj(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* const* geom,
                      const void* x, const void* d, const void* cot_x, void* grad_rhs, void* grad_d,
                      int32_t* iters_out, void* workspace, size_t ws_bytes, void* stream);

/* pom_eta_solve_noalloc: pom_eta_solve without a workspace argument and
 *   without any allocation (no hipMalloc / hipFree / pinned host memory, no
 *   host synchronisation), for callers that cannot allocate: code under HIP
 *   graph capture, tight serving loops.  No counterpart in the reference.
 *   The work arrays live in a static device array of POM_NOALLOC_SCRATCH_BYTES
 *   per GPU, held until the process ends; a batch whose eta workspace
 *   (pom_eta_workspace) exceeds it returns POM_E_SCRATCH.  The solve runs in
 *   the capture-safe mode (exactly prm->eta_max_iter iterations launched;
 *   converged members keep their iterate), so x equals pom_eta_solve's.
 *   Every call on one GPU shares that GPU's array: calls that overlap in time
 *   on one GPU (other streams, other host threads) race and must be
 *   serialised by the caller. */
#define POM_NOALLOC_SCRATCH_BYTES (64u << 20)
int pom_eta_solve_noalloc(const pom_batch* b, const pom_params_t* prm, const void* cache, const void* const* geom,
                          const void* rhs, const void* d, const void* x0, void* x, int32_t* iters_out, void* stream);

/* Operator-level entry points: the reference's operator methods that callers
 * invoke outside the time loop (src/neuralpom/hybrid/sparse_coupling.py:10-15
 * calls dens through its eos(), :69-82 calls baropg).  Same kernels as the step.
 *
 * pom_dens: equation of state (src/neuralpom/cores/explicit/momentum.py:467-541)
 *   rho_out(k) = EOS(t, s, depth) * fsm for k < 15; level 15 copies rho_in.
 * pom_dens_adj: b_t += dEOS/dt^T b_rho_out, b_s += ..., b_rho_in += level-15
 *   part (null outputs are skipped). */
int pom_dens(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* t, const void* s,
             const void* rho_in, const void* h, const void* fsm, void* rho_out, void* stream);
int pom_dens_adj(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* t, const void* s,
                 const void* rho_in, const void* h, const void* fsm, const void* b_rho_out, void* b_t, void* b_s,
                 void* b_rho_in, void* stream);

/* pom_baropg: sigma-coordinate baroclinic pressure gradient
 * (src/neuralpom/cores/explicit/momentum.py:543-700).
 *   in  = {rho, rmean, dt, dx, dy, dum, dvm}  (rmean 3-D, the rest of dt..dvm 2-D)
 *   out = {rho_out, drhox, drhoy}
 * pom_baropg_adj: cot_out = {b_rho_out, b_drhox, b_drhoy} (entries may be
 *   null), grad_in = {b_rho, b_dt} accumulated (entries may be null). */
int pom_baropg(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
               const void* const* in, void* const* out, void* stream);
int pom_baropg_adj(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid,
                   const void* const* in, const void* const* cot_out, void* const* grad_in, void* stream);

/* pom_profq: the Mellor-Yamada 2.5 closure on its own, the reference's
 *   operator method profq (src/neuralpom/cores/explicit/tracers.py:158-431,
 *   positional signature :158-163, returned tuple :431, consumed at
 *   time_loop.py:509-518), on the step's kernels (k_profq, then the km / kh
 *   edge copy).
 *   in  = {uf (q2 forward value), vf (q2l forward value), q2, q2b, q2lb, km, kh,
 *          kq, u, v, t, s, rho (3-D), wusurf, wvsurf, wubot, wvbot, etf, dt, h,
 *          fsm (2-D)}                                               (21 entries)
 *   out = {q2 new (uf), q2l new (vf), l, kq, km, kh (3-D, required),
 *          sm, sh, cc, a, c (3-D), dh (2-D) (the work arrays the reference
 *          returns; each may be null)}                              (12 entries)
 *   Scalars from prm: dti2, umol, grav, kappa, small, tbias, sbias, rhoref,
 *   core (selects the wall-term guard, 1e-20 explicit / 1e-8 implicit).
 * pom_profq_adj: grad_in[k] += (d out[0..5] / d in[k])^T cot_out[0..5] for the
 *   differentiable inputs k = uf .. rho, wubot, wvbot, etf, dt (entries of
 *   grad_in and cot_out may be null; wusurf, wvsurf, h, fsm are passive).
 *   Recomputes the forward.
 * pom_profq_workspace: workspace bytes of both. */
int pom_profq_workspace(const pom_batch* b, size_t* bytes);
int pom_profq(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* in,
              void* const* out, void* workspace, size_t ws_bytes, void* stream);
int pom_profq_adj(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* in,
                  const void* const* cot_out, void* const* grad_in, void* workspace, size_t ws_bytes, void* stream);

/* pom_proft: vertical diffusion of one tracer, the reference's operator method
 *   proft (tracers.py:890-1010, signature :890-892, returned tuple :1010,
 *   consumed at time_loop.py:573-581).
 *   in  = {f, kh (3-D), wfsurf, fsurf, etf, h (2-D)}, nbc = 1 (flux) or 3
 *         (prescribed surface value); 2 and 4 (shortwave absorption) return
 *         POM_E_UNSUPPORTED.
 *   out = {f new (required), a, c, ee, gg (3-D), dh (2-D) (work arrays; each
 *         may be null)}.
 * pom_proft_adj: grad_in = {f, kh, etf} += the transposed sweep of cot_f (each
 *   may be null; wfsurf and fsurf are passive). */
int pom_proft(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* in, int32_t nbc,
              void* const* out, void* stream);
int pom_proft_adj(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* in,
                  int32_t nbc, const void* cot_f, void* const* grad_in, void* stream);

/* pom_bcond: one lateral boundary condition of the explicit core alone,
 *   POM2K_Core.bcond(idx, elf, uaf, vaf, uf, vf, w) (cores/explicit/
 *   baroclinic.py:496-928), as callers outside the time loop apply it (the
 *   restart path: cores/implicit_p0_fixed/checkpointing.py:55-77).  fields / out per idx:
 *     1: {elf}                          -> {elf}   (edge copy, * fsm)
 *     2: {uaf, vaf, el}                 -> {uaf, vaf}   (radiation forms with rfe..rfs and ramp, * dum / dvm)
 *     3: {uf, vf, u, v}                 -> {uf, vf}   (edge velocities, * dum / dvm)
 *     4: {uf, vf, t, s, u, v, w, dt}    -> {uf, vf} = new t, s   (upstream edge tracers, * fsm)
 *     5: {w}                            -> {w}   (levels 0-14 * fsm)
 *     6: {uf, vf, q2, q2l, u, v}        -> {uf, vf} = new q2, q2l   (upstream edge turbulence, + 1e-10)
 *   (dt is 2-D, the others 3-D except elf, uaf, vaf, el); geom is the
 *   pom_geom table.  With prm->core = POM_CORE_IMPLICIT (closed basin) the
 *   semi-implicit core's policy applies instead, reading the first entries of
 *   the same lists (implicit_p0_fixed/baroclinic.py:505-510 ->
 *   boundary/closed_basin.py:27-87): 1 elf * fsm, 2 uaf * dum, vaf * dvm
 *   (edge ring zero), 3 uf * dum, vf * dvm (edge ring zero), 4 edge cells of
 *   uf, vf copied from the nearest interior cell, 5 w * fsm, 6 uf, vf * fsm.
 *   The tide policy's bcond (either core) returns POM_E_UNSUPPORTED: its
 *   bcond 1 advances the adapter's substep counter, so it runs inside the
 *   step only.
 * pom_bcond_adj: grad_fields[q] += d(out)/d(fields[q])^T cot_out (entries
 *   may be null). */
int pom_bcond(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid, int32_t idx,
              const void* const* fields, const void* const* geom, void* const* out, void* stream);
int pom_bcond_adj(const pom_batch* b, const pom_params_t* prm, const void* ramp, const void* vgrid, int32_t idx,
                  const void* const* fields, const void* const* geom, const void* const* cot_out,
                  void* const* grad_fields, void* stream);

/* ---- State movement around the step (no physics).
 *
 * pom_pack_input: the corrector network's input, 101 channels per member in
 *   NCHW (src/neuralpom/hybrid/state.py:181-226, pack order constants.py:2-7):
 *   fields = {u, ub, v, vb, t, s (3-D, all 16 levels), el, ua, uab, va, vab};
 *   member m's planes (101 x im x jm) start at element 101 * off2[m] of out.
 *   mean / scale (each may be null as a whole; entries null together) are
 *   per-cell statistics of one member's shape: channel = (x - mean) / scale,
 *   as StateNormalizer.normalize (state.py:156-179); with statistics the
 *   members must share (im, jm).
 * pom_pack_input_adj: grad_fields[f] += d(pack)/d(field f)^T cot_out
 *   (null entries skipped).
 * pom_restrict4: volume-weighted conservative 4 x 4 restriction of a fine
 *   batch onto a coarse one (member m to member m, coarse im = (im - 2)/4 + 2,
 *   likewise jm): coarse interior cell (I, J) averages fine interior cells
 *   4(I-1)+1 .. 4I with weights art*mask (volume_weighted = 0) or
 *   art*mask*(h + el)[*dz(k) for 3-D fields] (volume_weighted = 1); the coarse
 *   edge ring is 0.  The fine-to-coarse projection of the reference's training
 *   data (paper/paper_zh.md:625), whose code is not in the reference.
 * pom_restrict4_adj: grad_field += the restriction's transpose applied to
 *   cot_out (a fine interior cell gets w / W of its coarse cell's cotangent);
 *   with volume weights and el given, grad_el (may be null) += the
 *   cotangent through the weights.  Arguments as pom_restrict4. 
 * pom_copy_arrays: n device-to-device copies (sizes in bytes, multiples of 4)
 *   in one launch: device-resident snapshots of a state for restart and the
 *   per-step recompute, replacing the host round trip of capture / restore
 *   (src/neuralpom/v1/checkpoint.py:51-95). */
int pom_pack_input(const pom_batch* b, const void* const* fields, const void* const* mean,
                   const void* const* scale, void* out, void* stream);
int pom_pack_input_adj(const pom_batch* b, const void* const* scale, const void* cot_out, void* const* grad_fields,
                       void* stream);
int pom_restrict4(const pom_batch* fine, const pom_batch* coarse, int32_t three_d, int32_t volume_weighted,
                  const void* field, const void* art, const void* mask, const void* h, const void* el,
                  const void* fine_vgrid, void* out, void* stream);
int pom_restrict4_adj(const pom_batch* fine, const pom_batch* coarse, int32_t three_d, int32_t volume_weighted,
                      const void* field, const void* art, const void* mask, const void* h, const void* el,
                      const void* fine_vgrid, const void* cot_out, void* grad_field, void* grad_el, void* stream);
int pom_copy_arrays(int32_t n, void* const* dst, const void* const* src, const int64_t* bytes, void* stream);

/* ---- The corrector hook between physics intervals, fused
 * (src/neuralpom/hybrid/sparse_coupling.py:18-67: volume_weights,
 * budget_project, paired_limit, apply_joint_correction).
 *
 * pom_correction_workspace: device scratch bytes of the two calls below.
 * pom_joint_correction: for every member, raw = the network's 60 output planes
 *   (u, v, t, s on levels 0-14) of im x jm starting at element 60 * off2[m];
 *   scales = {u, v, t, s} (nonzero), cap > 0; geom = {dum, dvm, fsm, art, h, et}
 *   (2-D); vgrid as for the step (dz is its row 2).  Velocity increments lose
 *   their dz-weighted column mean and are masked; tracer increments lose their
 *   layer mean weighted by w = art fsm (h + et) dz (wet cells only); each pair
 *   is scaled by one factor min(cap / max(peak, cap), 1), peak = max |x / scale|
 *   over the pair.  state_in / state_out = {u, ub, v, vb, t, tb, s, sb}: out =
 *   in + increment on levels 0-14, level 15 copied.  diag (device [M][3], may
 *   be null) = (tracer factor, t budget, s budget) as the reference returns
 *   them.  The density refresh that follows is pom_dens.
 * pom_joint_correction_adj: grad_raw (written, the raw layout) and grad_et
 *   (accumulated, may be null) for cotangents cot_out of the 8 outputs
 *   (entries may be null) and cot_diag [M][3] (may be null); the state inputs'
 *   cotangents are cot_out themselves.  At a peak attained by several elements
 *   the factor's cotangent is shared evenly by all of them. */
int pom_correction_workspace(const pom_batch* b, size_t* bytes);
int pom_joint_correction(const pom_batch* b, const void* raw, const void* const* geom, const void* vgrid,
                         const double* scales, double cap, void* work, const void* const* state_in,
                         void* const* state_out, void* diag, void* stream);
int pom_joint_correction_adj(const pom_batch* b, const void* raw, const void* const* geom, const void* vgrid,
                             const double* scales, double cap, void* work, const void* const* cot_out,
                             const void* cot_diag, void* grad_raw, void* grad_et, void* stream);

/* ---- The density-75 projection of the DND policy, fused
 * (src/neuralpom/evaluation/density.py:14-64: eos_gradient,
 * constrained_amplitude, apply_density75), run after pom_joint_correction.
 *
 * pom_density75_workspace: device scratch bytes of pom_density75.
 * pom_density75: for every member, x = (after - before) / scale for T and S
 *   on levels 0-14 (scales = {t, s}, nonzero); d = g / max(|g|, 1e-30) with
 *   g = (drho/dT scale_t, drho/dS scale_s), the partials of the model EOS
 *   (prm: tbias, sbias, grav, rhoref) at `before` in float64; w = art fsm
 *   (h + et) dz; a = x . d.  Per level the Gram matrix and moments of d
 *   (weights normalised by the level sum) give c through the 2 x 2
 *   pseudo-inverse that drops eigenvalues at or below 1e-10 of the largest;
 *   adjusted = a - d . c; proposed = (x - 0.25 adjusted d) scale; one factor
 *   min(cap / max(peak, cap), 1), peak = max |proposed / scale| over T and S;
 *   out = before + factor proposed on levels 0-14 (level 15 from before) for
 *   T, its history level, S and its history level.  A member whose x is all
 *   zero gets out = after (the reference returns early).  geom = {fsm, art,
 *   h, et} (2-D); before / after / state_out = {t, tb, s, sb}; vgrid as for
 *   the step (zz row 1, dz row 2).  diag (device double [M][3], may be null)
 *   = (factor, max |constraint residual|, 1 if x was nonzero else 0; factor 1
 *   and residual 0 then).  The density refresh that follows is pom_dens. */
int pom_density75_workspace(const pom_batch* b, size_t* bytes);
int pom_density75(const pom_batch* b, const pom_params_t* prm, const void* vgrid, const void* const* geom,
                  const double* scales, double cap, void* work, const void* const* before,
                  const void* const* after, void* const* state_out, void* diag, void* stream);

/* Tracing: record HIP events around every launch of one stage on its stream
 * (a stage >= 0 starts a new record; -1 stops recording and keeps what was
 * recorded).  pom_profile_read waits for the recorded events and returns the
 * summed kernel time and the number of launches since the last read.  No counterpart in the reference (it times host wall clock only,
 * src/neuralpom/v1/engine.py:341-362). */
typedef enum {
  POM_STAGE_PROFQ = 0,      /* k_profq: Mellor-Yamada closure, forward */
  POM_STAGE_PROFQ_ADJ = 1,  /* k_profq_adj: its reverse sweep */
  POM_STAGE_EXTUV = 2,      /* depth-mean momentum substep (one launch per substep) */
  POM_STAGE_ADVCT = 3,      /* 3-D momentum advection */
  POM_NSTAGE
} pom_stage;
int pom_profile_stage(int32_t stage);
int pom_profile_read(double* total_ms, int32_t* count);

#ifdef __cplusplus
}
#endif
#endif
