/*
 * bfe.h — C ABI of the MI355X (gfx950) explicit time-march of BristolFE-v2.
 *
 * The library replaces the time-marching loop of
 *   code/internal/fn_explicit_dynamic_solver_v6.m:1-221
 * which code/fn_BristolFE_v2.m:120-122 calls with the assembled K, C, M.
 * A MEX shim (or any FFI) passes exactly those inputs: compressed sparse rows
 * of K and C (both symmetric, so MATLAB's CSC arrays Jc/Ir/Pr can be passed
 * unchanged as row pointers/columns/values), the lumped mass m = sum(M)
 * (v6:122-123), and dt = time(2) - time(1) (v6:93).
 *
 * Conventions
 *   - Every index crossing this ABI is 0-based (row pointers int64, column
 *     and DOF indices int32), except in the host mesh-preprocessing helpers,
 *     which take element tables as the model holds them (1-based int64).
 *   - Dense time data are column-major, as MATLAB stores them:
 *       frcs  : frc_rows x n_t,  element (r, t) at frcs[r + frc_rows * t]
 *       hist  : q x n_t
 *       field : n x n_frames
 *   - The caller allocates every output buffer; the handle owns device
 *     memory and one HIP stream. Inputs are borrowed for the call only.
 *   - Every function returns a bfe_status; nothing throws across the ABI.
 *     bfe_last_error() returns the message of the calling thread's last
 *     failure.
 *   - Handles are independent: calls on different handles may run from
 *     different threads; calls on one handle must not overlap.
 */
#ifndef BFE_H
#define BFE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bfe_status {
  BFE_OK = 0,
  BFE_ERR_ARG = 1,     /* malformed input (shape, index range, non-square) */
  BFE_ERR_HIP = 2,     /* HIP runtime failure (no device, launch error)    */
  BFE_ERR_NOMEM = 3,   /* device or host allocation failed                 */
  BFE_ERR_NUMERIC = 4, /* zero lumped mass, singular interface block,      */
                       /* non-finite values in the march                   */
  BFE_ERR_INTERNAL = 5
} bfe_status;

/* solver_mode strings of code/fn_BristolFE_v2.m:28 and
 * fn_explicit_dynamic_solver_v6.m:130-139 */
#define BFE_MODE_VEL_AT_CURRENT 0   /* 'vel at curent time step' (default) */
#define BFE_MODE_VEL_AT_LAST_HALF 1 /* 'vel at last half time step'        */

/* solver_precision of fn_FE_entry_point.m:31 and v6:142-158 */
#define BFE_PREC_DOUBLE 0
#define BFE_PREC_SINGLE 1

typedef struct bfe_plan_s* bfe_handle;

/* Static description of a plan, for logging and roofline accounting. */
typedef struct bfe_info {
  int64_t n;             /* DOF                                          */
  int64_t nnz_k;         /* stored entries of K (after dropping zeros)   */
  int64_t nnz_sell;      /* entries of the SELL-64 operator incl. padding */
  int64_t n_interface;   /* rows coupled through off-diagonal C          */
  int64_t nnz_interface; /* entries of the interface coupling C_off      */
  int64_t nnz_xinv;      /* entries of the truncated X_II^-1             */
  int32_t damped;        /* 1 if any row carries a damping term          */
  int32_t mode;
  int32_t precision;
  int32_t device;
  int64_t device_bytes;  /* operator + coefficient bytes on the device   */
  int32_t operator_kind; /* BFE_OPERATOR_CSR or BFE_OPERATOR_EBE (march) */
  int32_t ebe_slots;     /* (template, vertex) slots, summed over groups  */
  int32_t ebe_templates; /* element templates, summed over groups         */
  int32_t ebe_pattern;   /* mesh pattern of the first group, 0: none      */
  int64_t ebe_elements;  /* elements of the element operator             */
  double ebe_check;      /* set-up check: relative step difference to CSR */
  int32_t ebe_classes;   /* neighbour classes of the mesh pattern         */
  int32_t ebe_groups;    /* element groups of the element operator       */
  int32_t multi_group;   /* rows per gather group of the last batched run */
  int32_t ebe_sparse;    /* element groups on a structurally sparse product
                            (zero template entries skipped; equal to rounding) */
  int32_t multi_operator; /* operator of the last batched run: BFE_OPERATOR_* */
  int32_t grad_pairs;    /* DOFs per node of the node-pair form of the gradient
                            elements (1 or 2; BFE_TUNE_ADJOINT_PAIRS), 0: none */
} bfe_info;

/* Timing of the last bfe_run / bfe_run_batched, measured with HIP events
 * on the handle's stream around the time loop only (the interval the
 * reference brackets at v6:162-220). */
typedef struct bfe_run_stats {
  double loop_ms;        /* HIP-event time of the whole time loop        */
  int64_t steps;         /* time steps executed (n_t - ti_start + 1)     */
  int64_t launches;      /* kernel launches inside the loop              */
  int64_t columns;       /* right-hand sides marched together            */
  int64_t nonfinite;     /* non-finite values found in the history output */
                         /* (an unstable dt diverges silently in v6)      */
  double step_kernel_ms; /* summed HIP-event time of the sampled step-    */
                         /* kernel launches (see bfe_set_kernel_timing)   */
  int64_t step_kernel_samples;
} bfe_run_stats;

/*
 * bfe_create — set up a march.
 * Replaces v6:122-159 (lumped mass, A/B/D formation, device upload).
 *   n          number of global DOF (size of K, C, M)
 *   k_*        CSR of K (n x n); duplicate (row, col) entries are summed
 *   m          lumped mass per DOF, sum(M) (v6:122); every entry non-zero
 *   c_*        CSR of C (n x n); may be NULL/empty (C = 0). Diagonal entries
 *              are absorbing-layer damping; off-diagonal entries are the
 *              fluid-solid coupling of ASI2D2 elements
 *              (code/internal/elements/fn_el_ASI2D2.m:60-76)
 *   dt         time step
 *   mode       BFE_MODE_*
 *   precision  BFE_PREC_*; single keeps FP64 setup and marches in FP32
 *   device     HIP device ordinal
 * For mode 'vel at curent time step' the implicit part
 * X = I + dt/2 M^-1 C (v6:136-138) is applied exactly: diagonal rows are
 * folded into per-row coefficients, the interface block X_II is inverted on
 * the host in FP64 and applied per step (entries below 1e-18 of a column's
 * largest are dropped).
 */
bfe_status bfe_create(int64_t n, const int64_t* k_rowptr, const int32_t* k_col,
                      const double* k_val, const double* m,
                      const int64_t* c_rowptr, const int32_t* c_col,
                      const double* c_val, double dt, int32_t mode,
                      int32_t precision, int32_t device, bfe_handle* out);

/*
 * bfe_run — march n_t time samples from rest.
 * Replaces the loop v6:162-218 and its outputs:
 *   frc_idx[p], frcs (frc_rows x n_t)  forcing DOFs and signals; frc_rows is
 *              1 (one signal broadcast to every DOF, v6:182 with a scalar
 *              right-hand side) or p. Duplicate DOFs: the last wins, as in
 *              MATLAB's f(idx) = v.
 *   hist_idx[q], hist_out (q x n_t)    monitored DOFs, u after each update
 *              (v6:195-197); columns before ti_start stay 0.
 *   field_every  frames at ti = 1, 1+N, ... (v6:112-114); <= 0 means none.
 *   field_is_vel 1: frame = (u - u_prev)/dt, 0: frame = u (v6:64-69, 200-208)
 *   field_out (n x n_frames) may be NULL when field_every <= 0.
 *   ti_start_out  1-based first step with non-zero forcing (v6:163-166);
 *              n_t + 1 when the forcing is all zero (nothing is marched).
 */
bfe_status bfe_run(bfe_handle h, int32_t n_t, const int32_t* frc_idx, int32_t p,
                   const double* frcs, int32_t frc_rows, const int32_t* hist_idx,
                   int32_t q, double* hist_out, int32_t field_every,
                   int32_t field_is_vel, double* field_out,
                   int32_t* ti_start_out);

/*
 * Per-element field output on the device: the reduction of
 * code/internal/fn_get_field_output.m:64-79, called at
 * code/fn_BristolFE_v2.m:145, moved next to the march.
 *
 * bfe_set_field_elements — the mesh tables of the reduction:
 *   node_dof (n_nodes x max_df)  gl_lookup - 1: global DOF of (node, dof), -1 if absent
 *   els (n_el x npe)             element nodes, 0-based, -1 where a row is shorter
 * The lumped mass is the one given to bfe_create.
 *
 * bfe_run_field_ke — bfe_run with velocity frames every field_every steps
 * reduced on the device to the 'KE' field: per element and frame
 * sqrt(mean over the element's nodes of sum over the node's DOFs of
 * |v^2 m| / 2). el_out is n_el x n_frames (column-major); the n x n_frames
 * velocity frames stay on the device.
 */
bfe_status bfe_set_field_elements(bfe_handle h, int64_t n_nodes, int32_t max_df,
                                  const int32_t* node_dof, int64_t n_el, int32_t npe,
                                  const int32_t* els);
bfe_status bfe_run_field_ke(bfe_handle h, int32_t n_t, const int32_t* frc_idx, int32_t p,
                            const double* frcs, int32_t frc_rows, const int32_t* hist_idx,
                            int32_t q, double* hist_out, int32_t field_every, double* el_out,
                            int32_t* ti_start_out);

/*
 * The other per-element fields of code/internal/fn_get_field_output.m:28-58
 * on the device (the field_output_type strings of v6:64-69 and
 * fn_BristolFE_v2.m:145):
 *   BFE_FIELD_KE    'KE' (as bfe_run_field_ke; velocity frames)
 *   BFE_FIELD_MEAN  'mean(ui)' / 'mean(vi)': element mean of DOF column
 *                   comp (0-based column of gl_lookup; an absent DOF is 0)
 *   BFE_FIELD_DIV   'div(u)' / 'div(v)': sum over the element's nodes of
 *                   components 1..dim dotted with the unit vector from the
 *                   element centre to the node; 0 on an element with an
 *                   absent node
 *   BFE_FIELD_CURL  'curl(u)' / 'curl(v)': as DIV with the unit vectors
 *                   turned to (y, -x); 2D only (the reference errors on 3D)
 * field_is_vel picks velocity (1) or displacement (0) frames, as v6 does
 * from the type string. DIV and CURL need bfe_set_field_geometry: the node
 * coordinates, n_nodes x dim row-major (dim <= the DOFs per node of
 * bfe_set_field_elements). el_out is n_el x n_frames (column-major).
 */
#define BFE_FIELD_KE 0
#define BFE_FIELD_MEAN 1
#define BFE_FIELD_DIV 2
#define BFE_FIELD_CURL 3
bfe_status bfe_set_field_geometry(bfe_handle h, int32_t dim, const double* nds);
bfe_status bfe_run_field(bfe_handle h, int32_t n_t, const int32_t* frc_idx, int32_t p,
                         const double* frcs, int32_t frc_rows, const int32_t* hist_idx, int32_t q,
                         double* hist_out, int32_t field_every, int32_t field_type,
                         int32_t field_is_vel, int32_t comp, double* el_out,
                         int32_t* ti_start_out);

/*
 * bfe_run_batched — march n_src independent load cases on the same operator
 * in one launch sequence (sparse x multivector). Replaces the sequential
 * per-step loop of code/fn_BristolFE_v2.m:72-123 as driven by
 * subdoms/fn_run_main_model.m:101-115 (one step per transmitter).
 *   frc_idx[p], frc_src[p]  forcing DOF and the source column it belongs to
 *   frcs (frc_rows x n_t)   signals; frc_rows = 1 broadcasts one signal to
 *                           every forcing entry, else frc_rows = p
 *   hist_idx[q]             monitored DOFs, recorded for every source
 *   hist_out (q x n_src x n_t)  element (k, s, t) at k + q*(s + n_src*t)
 * All sources start at the same ti_start (the first step at which any
 * forcing is non-zero). No field output.
 */
bfe_status bfe_run_batched(bfe_handle h, int32_t n_t, int32_t n_src,
                           const int32_t* frc_idx, const int32_t* frc_src,
                           int32_t p, const double* frcs, int32_t frc_rows,
                           const int32_t* hist_idx, int32_t q, double* hist_out,
                           int32_t* ti_start_out);

/*
 * Differentiation (no counterpart in the reference, which has no adjoint).
 * Parameters: a stiffness scale s_e and a density rho_e per element, with
 *   K = sum_e s_e K_e,  M = sum_e rho_e Mbar_e,  C_diag = sum_e mu_e rho_e Mbar_e
 * (K_e including the absorbing-layer reduction of
 * fn_build_global_matrices_v5.m:116, C from :115), evaluated at s_e = 1 and
 * the model's densities. The adjoint is the same recurrence marched backward;
 * gradients are accumulated in FP64.
 *
 * bfe_set_gradient_elements — element data of the parameters:
 *   el_dofs (n_el x kmax, row-major) 0-based global DOFs, -1 for unused slots
 *   el_K    (n_el x kmax x kmax)     K_e at s_e = 1
 *   el_m    (n_el x kmax)            diagonal of dM/drho_e (lumped mass per unit density)
 *   el_mu   (n_el)                   damping per unit mass, C_e = mu_e M_e
 */
bfe_status bfe_set_gradient_elements(bfe_handle h, int32_t n_el, int32_t kmax,
                                     const int32_t* el_dofs, const double* el_K,
                                     const double* el_m, const double* el_mu);

/*
 * bfe_run_checkpointed — bfe_run (no field output) that also keeps, on the
 * device and under record `slot`, the two states at the start of every
 * ckpt_every-step segment (counted back from the last step) and all states of
 * the last segment, for one following bfe_adjoint on the same slot. Several
 * slots (e.g. one per transmitter) may be held at once.
 */
bfe_status bfe_run_checkpointed(bfe_handle h, int32_t slot, int32_t n_t, const int32_t* frc_idx,
                                int32_t p, const double* frcs, int32_t frc_rows,
                                const int32_t* hist_idx, int32_t q, double* hist_out,
                                int32_t ckpt_every, int32_t* ti_start_out);

/*
 * bfe_run_checkpointed_batched — bfe_run_batched (arguments as there, at
 * most 32 sources) that also keeps, under record `slot`, the two states
 * (n x n_src each) at the start of every ckpt_every-step segment and all
 * states of the last segment, for one following bfe_adjoint, which then
 * returns the gradient of J summed over the sources. It marches the CSR
 * operator (the one the adjoint's recomputation uses).
 */
bfe_status bfe_run_checkpointed_batched(bfe_handle h, int32_t slot, int32_t n_t, int32_t n_src,
                                        const int32_t* frc_idx, const int32_t* frc_src, int32_t p,
                                        const double* frcs, int32_t frc_rows,
                                        const int32_t* hist_idx, int32_t q, double* hist_out,
                                        int32_t ckpt_every, int32_t* ti_start_out);

/*
 * bfe_adjoint — gradients of a scalar J(hist) of the checkpointed run held
 * under `slot` (which it consumes).
 *   hist_grad  dJ/dhist, the layout of that run's hist_out: q x n_t
 *              (bfe_run_checkpointed), q x n_src x n_t
 *              (bfe_run_checkpointed_batched)
 *   g_stiffness, g_density (n_el)  dJ/ds_e and dJ/drho_e (summed over the
 *              sources of a batched run)
 * The reverse sweep recomputes each segment from its checkpoint in the same
 * launches that march the adjoint over the following segment (for a batched
 * run: one multi-source march of 2 n_src columns per step).
 */
bfe_status bfe_adjoint(bfe_handle h, int32_t slot, const double* hist_grad,
                       double* g_stiffness, double* g_density);

/*
 * bfe_set_tuning — launch shapes and operator choice (no counterpart in the
 * reference): wavefronts per workgroup of the multi-source kernel (1, 2, 4
 * or 8; default 2), the XCD-aware workgroup order (default on), and the keys
 * below. Results do not depend on any of
 * them beyond rounding; the autotuner picks the multi-source shape by timing.
 */
#define BFE_TUNE_MULTI_ROWS_PER_BLOCK 1
#define BFE_TUNE_XCD_REMAP 2
/* rows each wavefront of the multi-source march handles together (1, 2, 4,
 * 8 or 16; default 4): consecutive rows of one 64-row slice, interleaved */
#define BFE_TUNE_MULTI_ROWS_PER_WAVE 6
/* rows per gather group of the multi-source march (1, 2, 4 or 8, at most
 * the rows per wavefront; 0, the default: automatic). With G > 1 each
 * aligned group of G rows is stored over the union of its rows' columns
 * (built on the device at the first batched run that uses it), so one
 * gather of a neighbour serves G rows: G = 2 for the two rows of a 2D solid
 * node, G = 4 for two neighbouring nodes. Automatic picks 2 when that saves
 * at least a third of the gathers, 4 when that saves a further fifth of
 * those, else 1 (never more than the rows per wavefront). bfe_info
 * multi_group reports the group of the last batched run. Results agree to
 * rounding. */
#define BFE_TUNE_MULTI_GROUP 8
/* 1: bfe_run captures its whole time loop (step kernels, interface solve,
 * last snapshot) into one hipGraph and launches that, instead of launching
 * step by step; for small meshes whose steps are launch-latency bound.
 * Plain marches only (not bfe_run_checkpointed); no kernel sampling.
 * Default 0. */
#define BFE_TUNE_GRAPH 7
/* 1: bfe_run marches the CSR operator with one cooperative launch of a
 * persistent kernel that runs every time step (grid barrier between steps,
 * and after the interface solve). Results are identical to the step-by-step
 * march. On MI355X a grid barrier (about 16 us) costs more than a kernel
 * boundary, so BFE_TUNE_GRAPH is the faster choice for small meshes. Plain
 * marches on the CSR operator only (not with an element operator in use, not
 * bfe_run_checkpointed); no kernel sampling. Default 0. */
#define BFE_TUNE_PERSISTENT 9
/* at most this many workgroups in the persistent march's launch (0, the
 * default: as many as the device holds at once, at most one per row block).
 * With fewer workgroups than row blocks each workgroup sweeps blocks
 * lb, lb + grid, ... between the grid barriers; results do not change. */
#define BFE_TUNE_PERSISTENT_GRID 10
/* operator of the batched march (bfe_run_batched) once an element operator
 * is attached: BFE_OPERATOR_EBE marches the multi-source element product
 * (k_step_ebe_multi: one wavefront per node, lane = source, each wavefront
 * marching a run of consecutive nodes) when every element group has a
 * mesh-pattern kernel, BFE_OPERATOR_NODE the same kernel on the node-block
 * form (below) when every group is a triangle pattern, else the row-grouped
 * CSR operator is used; BFE_OPERATOR_CSR (the default) always uses the CSR
 * operator. bfe_info multi_operator reports the operator of the last
 * batched run. Results agree to rounding. */
#define BFE_TUNE_MULTI_OPERATOR 11
/* consecutive nodes each wavefront of the multi-source element product
 * marches (1 to 1024; default 4) */
#define BFE_TUNE_MULTI_NODES_PER_WAVE 12
/* 1 (default): along a run of regular nodes the multi-source element
 * product keeps the neighbour rows that the next node shares in registers
 * and gathers only the others; 0: every node gathers all its neighbour rows.
 * Results are identical. */
#define BFE_TUNE_MULTI_SLIDE 13
/* 1 (default): the batched adjoint (bfe_run_checkpointed_batched +
 * bfe_adjoint) accumulates node-pair correlations of the adjoint and forward
 * states over the sweep and forms the element gradients from them once, when
 * the gradient elements' vertices are nodes of 1 or 2 consecutive DOFs with
 * few neighbours (2D triangles); 0: per-element sums every segment
 * (k_grad_multi). Results agree to rounding. */
#define BFE_TUNE_ADJOINT_PAIRS 14
/* operator of the single-source march once an element operator is attached
 * (default: the element operator) */
#define BFE_TUNE_OPERATOR 3
#define BFE_OPERATOR_CSR 0
#define BFE_OPERATOR_EBE 1
/* batched march only: the node-block form of the element operator, per node
 * the DOF blocks of K for each neighbour class (summed once from the element
 * templates on the device, at the first batched run that asks for it); K u1
 * is then sum_j B_j(v) u1(neighbour j), 28 products per 2D solid node against
 * the slot product's 84. Triangle patterns only; other groups fall back to
 * the CSR operator. */
#define BFE_OPERATOR_NODE 2
/* slots of the element operator whose loads are issued together in the
 * generic gather kernel (2, 4 or 8; default 8), and whether a mesh-pattern
 * kernel is used when the mesh matches one (0 or 1; default 1) */
#define BFE_TUNE_EBE_GROUP 4
#define BFE_TUNE_EBE_KERNEL 5
bfe_status bfe_set_tuning(bfe_handle h, int32_t key, int32_t value);

/*
 * bfe_set_element_operator — attach an element-by-element form of K to the
 * handle for the single-source march (bfe_run). Replaces the product K u of
 * v6:185 (B = X^-1 (2I - dt^2 M^-1 K), v6:137) by a gather over the
 * elements around each node, with the element matrices of
 * fn_build_global_matrices_v5.m:105-128 stored as a few templates:
 *   K = sum_e el_scale[e] * P_e' tmpl[el_tmpl[e]] P_e
 * el_scale carries the absorbing-layer stiffness reduction (v5:116) and any
 * per-element factor; P_e maps local DOF (vertex l, component c), index
 * l*dpn + c, to global DOF el_nodes[e*npe + l]*dpn + c. The handle's DOFs
 * must be numbered node-major with dpn DOFs on every node (n = n_nodes*dpn).
 *   el_nodes (n_el x npe, 0-based), el_tmpl (n_el), el_scale (n_el)
 *   tmpl (n_tmpl x (npe*dpn)^2, row-major)
 * Shapes: npe = 3 with dpn 1, 2 or 3 (triangles), npe = 4 with dpn = 3 (tets).
 * At most one element of each template may have a given node as its local
 * vertex l (true of structured meshes), and at most 64 (template, vertex)
 * pairs may occur. Before accepting it, the library marches one step of both
 * operators from the same random state and rejects the element operator
 * (BFE_ERR_ARG) if the results differ by more than check_tol relative
 * (<= 0: 1e-9 in double, 1e-3 in single precision). The CSR operator stays
 * in use for the checkpointed/adjoint runs (bfe_run_batched: see
 * BFE_TUNE_MULTI_OPERATOR).
 * No counterpart in the reference (which always forms sparse B).
 */
bfe_status bfe_set_element_operator(bfe_handle h, int64_t n_nodes, int32_t dpn, int64_t n_el,
                                    int32_t npe, const int32_t* el_nodes, const int32_t* el_tmpl,
                                    const double* el_scale, int32_t n_tmpl, const double* tmpl,
                                    double check_tol);

/*
 * bfe_set_element_groups — the element operator for meshes whose nodes carry
 * different DOF sets (e.g. CPE3 solid nodes with u1, u2 and AC2D3 fluid
 * nodes with p, fn_build_global_matrices_v5.m:83-128): one group per element
 * type, each with its own node numbering 0..n_nodes-1 and
 *   node_dof (n_nodes)  global DOF of component 0 of each node; the node's
 *                       dpn components are node_dof[j] + c
 * (NULL: node j's DOFs are j*dpn + c, a single group only). Every global DOF
 * must belong to exactly one group (K has no entry between groups; the
 * ASI2D2 coupling is in C). The other fields are those of
 * bfe_set_element_operator; the same set-up check applies. At most 8 groups.
 */
typedef struct bfe_element_group {
  int64_t n_nodes;
  int32_t dpn;
  int32_t npe;
  int64_t n_el;
  const int32_t* el_nodes;  /* n_el x npe, group-local node indices */
  const int32_t* el_tmpl;   /* n_el                                 */
  const double* el_scale;   /* n_el                                 */
  int32_t n_tmpl;
  int32_t reserved;
  const double* tmpl;       /* n_tmpl x (npe*dpn)^2                 */
  const int32_t* node_dof;  /* n_nodes, or NULL                     */
} bfe_element_group;

bfe_status bfe_set_element_groups(bfe_handle h, int32_t n_groups, const bfe_element_group* groups,
                                  double check_tol);

/*
 * bfe_set_kernel_timing — sample the duration of the main step kernel.
 * No counterpart in the reference (its only timing is the wall clock of
 * v6:220). every > 0 brackets every every-th step-kernel launch with HIP
 * events on the handle's stream; the next runs report the sum and count in
 * bfe_run_stats. every <= 0 switches sampling off (the default).
 */
bfe_status bfe_set_kernel_timing(bfe_handle h, int32_t every);

/*
 * Native global assembly for large meshes (host, FP64, multi-threaded).
 * Replaces code/internal/fn_build_global_matrices_v5.m:50-159 for meshes
 * of CPE3 (type 0), AC2D3 (1) and C3D4 (2) elements without interface
 * elements, with the same DOF numbering, absorbing modifications and column
 * compaction; K is produced row by row, so memory stays O(nnz).
 *   nds (n_nodes x dim, row-major), els (n_el x nper, 0-based, -1 unused),
 *   el_type, el_mat (0-based), el_abs (absorbing index or NULL),
 *   mat_D (n_mat x 36, 6x6 row-major; AC2D3 reads D[0] = bulk modulus),
 *   mat_rho (n_mat), dof_mask (bit d-1 = DOF d in use; 0 = all).
 * bfe_assemble_pattern returns the sizes and a work object that
 * bfe_assemble_fill consumes: CSR of K (rowptr n_dof+1, col/kval nnz),
 * lumped mass m and damping diagonal c (n_dof), gl_lookup (n_nodes x max_df,
 * row-major, 1-based, 0 = absent).
 */
bfe_status bfe_assemble_pattern(int64_t n_nodes, int32_t dim, const double* nds, int64_t n_el,
                                int32_t nper, const int32_t* els, const int32_t* el_type,
                                const int32_t* el_mat, const double* el_abs, int32_t n_mat,
                                const double* mat_D, const double* mat_rho, int32_t dof_mask,
                                double damping_power_law, double max_damping,
                                double max_stiffness_reduction, int64_t* n_dof, int64_t* nnz,
                                int32_t* max_df, void** work);
bfe_status bfe_assemble_fill(void* work, int64_t* rowptr, int32_t* col, double* kval, double* m,
                             double* c, int64_t* gl_lookup);
const char* bfe_assemble_last_error(void);

/*
 * Device-side global assembly (GPU `device`, FP64). Replaces
 * code/internal/fn_build_global_matrices_v5.m:50-159 with
 * fn_reduce_global_matrices.m:7-16 and fn_create_fast_lookup.m:4-8 for
 * meshes of CPE3 (type 0), AC2D3 (1), C3D4 (2) and ASI2D2 fluid-solid
 * interface elements (3). Inputs as bfe_assemble_pattern; el_mat of an
 * ASI2D2 element is ignored (the reference gives interface elements no
 * material, v5:92-95). Node -> element incidence, element matrices and the
 * row-by-row gather of K and C (each entry summed in ascending element order)
 * all run on the GPU. bfe_assemble_device returns the sizes and a work object
 * that holds the result in device memory; bfe_assemble_device_fill copies it
 * out and releases it: CSR of K (k_rowptr n_dof+1, k_col/k_val nnz_k), CSR of
 * C (absorbing diagonal and ASI2D2 couplings; c_rowptr, c_col/c_val nnz_c),
 * the lumped mass m (n_dof) and gl_lookup (n_nodes x max_df, row-major,
 * 1-based, 0 = absent). bfe_assemble_device_free releases a work object that
 * is not filled. Errors: bfe_assemble_last_error.
 */
bfe_status bfe_assemble_device(int32_t device, int64_t n_nodes, int32_t dim, const double* nds,
                               int64_t n_el, int32_t nper, const int32_t* els,
                               const int32_t* el_type, const int32_t* el_mat, const double* el_abs,
                               int32_t n_mat, const double* mat_D, const double* mat_rho,
                               int32_t dof_mask, double damping_power_law, double max_damping,
                               double max_stiffness_reduction, int64_t* n_dof, int64_t* nnz_k,
                               int64_t* nnz_c, int32_t* max_df, void** work);
bfe_status bfe_assemble_device_fill(void* work, int64_t* k_rowptr, int32_t* k_col, double* k_val,
                                    int64_t* c_rowptr, int32_t* c_col, double* c_val, double* m,
                                    int64_t* gl_lookup);
bfe_status bfe_assemble_device_free(void* work);

/*
 * Element templates of the element-by-element operator, found on the GPU
 * (device). The reference assembles K from every element's own matrix
 * (code/internal/fn_build_global_matrices_v5.m:105-128); on a structured mesh
 * most elements are translates of a few shapes, and this finds them. Elements
 * (els: n_el x npe, 0-based; el_mat: material per element) share a template
 * when their node offsets from vertex 1, rounded to 2^-24 of the largest
 * offset, and their materials agree: key = sum of the rounded offsets and the
 * material times mult ((npe-1)*dim + 1 odd multipliers), wrapping 64-bit.
 * Templates are numbered in ascending order of the key (as a signed integer);
 * first[t] is the smallest element index of template t, el_tmpl the template
 * of every element, and jfactor (2D triangles only, else NULL) J_first / J_e
 * with J from absolute coordinates (fn_el_CPE3.m:56). More than max_templates
 * shapes: BFE_OK with *n_tmpl = 0. Two shapes with one key: BFE_ERR_ARG.
 */
bfe_status bfe_element_templates(int32_t device, int64_t n_nodes, int32_t dim, const double* nds,
                                 int64_t n_el, int32_t npe, const int32_t* els,
                                 const int32_t* el_mat, const int64_t* mult, int32_t max_templates,
                                 int32_t* el_tmpl, int64_t* first, int32_t* n_tmpl,
                                 double* jfactor);

/*
 * Subdomain coupling operations (GPU `device`, FP64).
 *
 * bfe_convolve replaces code/internal/fn_convolve.m:22-34: every signal of f
 * (n_sig x n_t, row-major, time contiguous) is convolved with g through FFTs
 * of length 2 n_t (g zero-padded or truncated to that length) and the first
 * n_t samples are kept in h (n_sig x n_t). g is one signal of n_g samples
 * (g_per_signal = 0) or one per signal (n_sig x n_g, g_per_signal = 1). The
 * reference's ifft keeps a rounding-level imaginary part; h is the real part.
 *
 * bfe_disps_to_forces replaces code/fn_convert_disps_to_forces_v2.m:1-29:
 * the forces M a + C v + K d of a displacement history disp (n x n_t,
 * column-major, as MATLAB holds it) on the DOFs of layers 2 and 3 (lyrs, n
 * values), with the displacements of layers 3, 4 (in_or_out = 0, 'in') or
 * 1, 2 (1, 'out') set to zero first; m is the lumped mass (sum(M)) and
 * solver_mode a BFE_MODE_* value choosing the velocity difference. n_force
 * must equal the number of layer-2 and layer-3 DOFs; force_idx receives them
 * (0-based, ascending) and force their forces (n_force x n_t, column-major).
 */
bfe_status bfe_convolve(int32_t device, int64_t n_sig, int32_t n_t, const double* f, int32_t n_g,
                        const double* g, int32_t g_per_signal, double* h);
bfe_status bfe_disps_to_forces(int32_t device, int64_t n, const int64_t* k_rowptr,
                               const int32_t* k_col, const double* k_val,
                               const int64_t* c_rowptr, const int32_t* c_col,
                               const double* c_val, const double* m, double dt, int32_t n_t,
                               const double* disp, const int32_t* lyrs, int32_t in_or_out,
                               int32_t solver_mode, int64_t n_force, int32_t* force_idx,
                               double* force);

/*
 * bfe_disps_to_forces_noalloc — bfe_disps_to_forces that allocates nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Arguments, results and errors are those of
 * bfe_disps_to_forces, computed by the same kernel. Every device array of the
 * call (the staged K, C, m, layers, force rows and displacements, and the
 * forces) lives in one static 64 MiB device array per GPU, held from the
 * library's load until the process ends; a call whose arrays need more
 * (8 (4 n + 2 (nnz_K + nnz_C) + n n_t + n_force n_t) bytes, each sub-array
 * rounded up to 256 bytes) returns BFE_ERR_ARG. Calls on one GPU share that
 * array, so calls that overlap on one GPU (from several threads) race; calls
 * on different GPUs do not.
 */
bfe_status bfe_disps_to_forces_noalloc(int32_t device, int64_t n, const int64_t* k_rowptr,
                                       const int32_t* k_col, const double* k_val,
                                       const int64_t* c_rowptr, const int32_t* c_col,
                                       const double* c_val, const double* m, double dt,
                                       int32_t n_t, const double* disp, const int32_t* lyrs,
                                       int32_t in_or_out, int32_t solver_mode, int64_t n_force,
                                       int32_t* force_idx, double* force);

/*
 * Host mesh preprocessing (C++, multi-threaded), for models of millions of
 * elements. Element tables are taken as the model holds them: n_el x nper,
 * row-major, 1-based node numbers, 0 = absent. Points are n_pts x 2,
 * row-major.
 *   bfe_element_centres        code/fn_calc_element_centres.m (out n_el x dim)
 *   bfe_points_in_polygon      inpolygon as code/fn_elements_in_region.m:6
 *                              uses it: 1 inside or on the polygon (vx, vy:
 *                              n_v vertices), else 0
 *   bfe_dist_points_to_polygon code/fn_dist_point_to_bdry_2D.m: distance to
 *                              the nearest vertex and to the two edges there
 */
bfe_status bfe_element_centres(int64_t n_nodes, int32_t dim, const double* nds, int64_t n_el,
                               int32_t nper, const int64_t* els, double* out);
bfe_status bfe_points_in_polygon(int64_t n_pts, const double* pts, int32_t n_v,
                                 const double* vx, const double* vy, uint8_t* out);
bfe_status bfe_dist_points_to_polygon(int64_t n_pts, const double* pts, int32_t n_b,
                                      const double* bdry, double* out);
/*
 *   bfe_tri_grid               the node grid and triangles of
 *                              code/fn_rectangular_structured_mesh.m:22-43
 *                              (iso = 0) and code/fn_isometric_structured_mesh.m:22-75
 *                              (iso = 1) before their region clip: x (nx),
 *                              y (ny) the linspace axes, shift = el_size / 4
 *                              (isometric row offset); node k = iy + ny ix;
 *                              nds (nx ny x 2), els (n_el x 3, 1-based) in the
 *                              reference's order; with nds = els = NULL only
 *                              *n_el is set
 *   bfe_remove_unused_nodes    code/internal/fn_remove_unused_nodes.m: renumbers
 *                              els in place (used nodes in increasing order),
 *                              old_nds (n_nodes) receives each new node's old
 *                              number, *n_used their count
 *   bfe_get_edges              code/fn_get_edges.m:1-35: the unique edges
 *                              (lower, higher node; ed, *n_ed x 2) in
 *                              increasing order and the elements on edge i,
 *                              largest first, at el[ptr[i] .. ptr[i+1]);
 *                              ed, ptr and el hold n_el x nper entries at most
 */
bfe_status bfe_tri_grid(int32_t iso, int64_t nx, int64_t ny, const double* x, const double* y,
                        double shift, double* nds, int64_t* els, int64_t* n_el);
bfe_status bfe_remove_unused_nodes(int64_t n_nodes, int64_t n_el, int32_t nper, int64_t* els,
                                   int64_t* old_nds, int64_t* n_used);
bfe_status bfe_get_edges(int64_t n_el, int32_t nper, const int64_t* els, int64_t* n_ed,
                         int64_t* ed, int64_t* ptr, int64_t* el);

bfe_status bfe_info_get(bfe_handle h, bfe_info* info);
bfe_status bfe_last_run_stats(bfe_handle h, bfe_run_stats* stats);
bfe_status bfe_destroy(bfe_handle h);
const char* bfe_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BFE_H */
