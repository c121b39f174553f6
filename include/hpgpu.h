/*
 * hpgpu.h -- C-ABI of the MI355X residual/relaxation library for tri_hp physics.
 *
 * One context per solution block.  It mirrors the reference's per-block virtual
 * interface (tri_hp/tri_hp.h:207-252, reached from block::cycle through
 * multigrid_interface, tri_mesh/block.h:233-247).  All arrays passed in or out use
 * the reference layouts (tri_hp/allocate.cpp:105-107): vertex v(npnt,NV), side
 * s(nseg,sm0,NV), interior i(ntri,im0,NV), row-major, variable index fastest.
 *
 * Return codes: 0 ok, >0 numerical error count (the reference's int error of
 * setup_preconditioner), <0 runtime/HIP error.  Nothing throws across the ABI.
 * A context is used by one host thread at a time (one block per thread, as in
 * tri_mesh/blocks.cpp:220-240); distinct contexts are independent.
 */
#ifndef HPGPU_H
#define HPGPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hpg_ctx hpg_ctx;

/* Mesh arrays of tri_mesh (tri_mesh/tri_mesh.h:34-98): side j of a triangle runs
 * from vertex (j+1)%3 to (j+2)%3 and tri_sgn is +1 when the segment is stored in
 * that direction (tri_mesh/input.cpp:363-381).  Edge boundaries list their
 * segments in order (edge_bdry::seg); ebdry_essential[b] is a bit mask of the
 * variables that boundary holds fixed (hp_edge_bdry::essential_indices,
 * tri_hp/hp_boundary.h:36-38). */
typedef struct {
    int npnt, nseg, ntri;
    const double *pnts;      /* npnt x 2 */
    const int *seg_pnt;      /* nseg x 2 */
    const int *tri_pnt;      /* ntri x 3 */
    const int *tri_seg;      /* ntri x 3 */
    const int *tri_sgn;      /* ntri x 3 */
    int nebd;
    const int *ebdry_start;  /* nebd + 1 offsets into ebdry_seg */
    const int *ebdry_seg;
    const int *ebdry_essential;
} hpg_tri_mesh;

/* HPG_PHYSICS_INS: tri_hp_ins, unknowns (u, v, p), NV = 3.
 * HPG_PHYSICS_CNS: tri_hp_cns, primitive unknowns (p, u, v, RT), NV = 4
 * (tri_hp/cns/tri_hp_cns.h:19-72); every array below then carries 4 values per
 * vertex, side mode and interior mode.  Boundaries whose essential mask is 0b1110
 * (u, v, RT fixed) are the reference's cns "inflow" boundaries
 * (tri_hp/cns/bdry_cns.h:56-66) with the constant state bc[]. */
enum { HPG_PHYSICS_INS = 0, HPG_PHYSICS_CNS = 1 };
enum { HPG_FP64 = 0, HPG_FP32_DEFECT = 1 };
enum { HPG_PRCN_ANALYTIC = 0, HPG_PRCN_LEARNED = 1 };

/* Input-map values the hot path reads (tri_hp/ins/allocate.cpp:13-44,
 * tri_mesh/blocks.cpp:1126-1141, tri_hp/allocate.cpp:275-276, :198). */
typedef struct {
    int physics;             /* HPG_PHYSICS_INS or HPG_PHYSICS_CNS */
    int log2p;               /* 0,1,2 -> p = 1,2,4 */
    double rho, mu, adis;    /* b0_rho (ins), b0_mu, b0_dissipation */
    double bc_pressure;      /* ins: pressure of the freestream ibc on edge fluxes */
    double fadd;             /* hp_fadd */
    int nstage;              /* <= 7 */
    double alpha[8], beta[8];/* stage coefficients; entry nstage is set to 1 */
    double cfl[4];           /* per log2p */
    /* cns only (tri_hp/cns/allocate.cpp:32-45, tet_hp/cns/allocate.cpp:29-50): gamma,
     * prandtl, Rgas, and the inflow / boundary-flux state of the ibc: tri (p, u, v, RT)
     * in bc[0..3]; tet (unused, u, v, w, RT) in bc[0..4] */
    double gamma, prandtl, Rgas;
    double bc[5];
} hpg_params;

/* replaces tri_hp::init / tri_hp_ins::init (tri_hp/allocate.cpp, tri_hp/ins/allocate.cpp:13-44) */
int hpg_create(const hpg_tri_mesh *mesh, const hpg_params *prm, hpg_ctx **out);
int hpg_destroy(hpg_ctx *ctx);

/* state exchange with the host-owned ug arrays (tri_hp::ug, tri_hp/tri_hp.h:71-75) */
int hpg_set_state(hpg_ctx *ctx, const double *ugv, const double *ugs, const double *ugi);
int hpg_get_state(hpg_ctx *ctx, double *ugv, double *ugs, double *ugi);
/* residual exchange (tri_hp::hp_global::res, tri_hp/tri_hp.h:118-121) */
int hpg_get_residual(hpg_ctx *ctx, double *resv, double *ress, double *resi);
int hpg_set_residual(hpg_ctx *ctx, const double *resv, const double *ress, const double *resi);
/* diagonal preconditioner (vprcn, sprcn, tprcn of tri_hp::hp_global, tri_hp.h:138-139) */
int hpg_get_preconditioner(hpg_ctx *ctx, double *vprcn, double *sprcn, double *tprcn);
int hpg_set_tprcn(hpg_ctx *ctx, const double *tprcn);
/* cns block preconditioners, 16 doubles (row-major 4x4) per entity: the vertex and
 * side sums and the per-element tpreconditioner (tri_hp_cns::hp_cns_global::
 * vpreconditioner/spreconditioner/tpreconditioner, tri_hp/cns/tri_hp_cns.h:33) */
int hpg_get_block_preconditioner(hpg_ctx *ctx, double *vpc, double *spc, double *tpc);

/* replaces tri_hp::rsdl(int stage) (tri_hp/nstage.cpp:12-110) with
 * tri_hp_ins::element_rsdl (tri_hp/ins/rsdl.cpp:17-436) or tri_hp_cns::element_rsdl
 * (tri_hp/cns/rsdl.cpp:16-535) and the edge-boundary fluxes
 * (tri_hp/hp_boundary.cpp:973-1049); stage == nstage is the full residual */
int hpg_rsdl(hpg_ctx *ctx, int stage);
/* replaces tri_hp_ins::setup_preconditioner (tri_hp/ins/setup_preconditioner.cpp:9-310)
 * or tri_hp_cns::setup_preconditioner (tri_hp/cns/setup_preconditioner.cpp:7-239)
 * + tri_hp::setup_preconditioner (tri_hp/setup_preconditioner.cpp:9-96); returns the
 * error count (non-positive element size or NaN in the solution) */
int hpg_setup_preconditioner(hpg_ctx *ctx);
/* replaces tri_hp::minvrt (tri_hp/setup_preconditioner.cpp:102-373) */
int hpg_minvrt(hpg_ctx *ctx);
/* replaces tri_hp::update (tri_hp/nstage.cpp:256-427) or, for cns, tri_hp_cns::update
 * (tri_hp/cns/update.cpp:276-357: minvrt, inflow modify_boundary_residual,
 * project_new_variables, stage update) */
int hpg_update(hpg_ctx *ctx);
/* h-multigrid: make `coarse` (a context created on the coarse mesh with log2p = 0)
 * the coarse neighbour of `fine` (replaces tri_mesh::connect and the ccnnct/fcnnct
 * tables, tri_mesh/tri_mesh.h:197-201, tri_mesh/blocks.cpp:1245-1260).  cc_tri/cc_wt:
 * for every fine vertex the coarse triangle holding it and its 3 barycentric weights;
 * fc_tri/fc_wt: for every coarse vertex the fine triangle and weights.  Afterwards
 * hpg_mg_restrict(coarse) restricts the fine residual and state onto the coarse mesh
 * (tri_hp/movco.cpp:41-102) and hpg_mg_prolongate(coarse) adds the coarse correction
 * to the fine state (tri_hp/movfin.cpp:16-82).  The fine context must outlive it. */
int hpg_link_fine(hpg_ctx *coarse, hpg_ctx *fine, const int *cc_tri, const double *cc_wt, const int *fc_tri,
                  const double *fc_wt);
/* replaces block::cycle (tri_mesh/blocks.cpp:1267-1348) over a stack of grids: grids[0]
 * the fine context at its finest p-level, grids[k] (k >= 1) the coarse contexts linked
 * with hpg_link_fine(grids[k], grids[k-1], ...); multigrid level lvl runs on grid
 * clamp(lvl - extra_finest_levels, 0, ngrid - 1) (:1274-1275), the first
 * extra_finest_levels levels lowering p on the fine mesh.  vw cycles per level (2: W),
 * itercrsn updates before and iterrfne after each coarse correction, each after a
 * setup_preconditioner.  The whole cycle is enqueued without host round trips: the
 * setup_preconditioner error counts accumulate on the device and the return value is
 * the number of grids that reported an error during the cycle (0: none; the reference
 * stops at the first, this form finishes the cycle); maxres (may be NULL) receives the
 * entry level's maxres after its first smoothing, as block::cycle reads it.  At most 16
 * grids.  The second call with unchanged settings (the stack, the cycle shape, and per
 * grid the precision, preconditioner mode and MLP, the CFL list and the p-level)
 * captures the cycle into a HIP graph on grids[0]'s stream; later such calls replay it,
 * with the same kernels and arguments as the enqueued form. */
int hpg_mg_cycle(hpg_ctx *const *grids, int ngrid, int extra_finest_levels, int vw, int itercrsn, int iterrfne,
                 double *maxres);
/* No reference counterpart: enable (default) or disable the graph capture of
 * hpg_mg_cycle on this fine context (disabled: every cycle is enqueued call by call). */
int hpg_set_cycle_graph(hpg_ctx *ctx, int enable);
/* replaces tri_hp::maxres (tri_hp/allocate.cpp:496-525); per_var may be NULL */
int hpg_maxres(hpg_ctx *ctx, double *per_var, double *out);
/* raw-residual judge: max |res| over every vertex, side and interior coefficient
 * that is not held by an essential boundary condition (no reference counterpart;
 * the convergence measure the benchmarks use) */
int hpg_resnorm(hpg_ctx *ctx, double *out);
/* replace tri_hp::mg_restrict / mg_prolongate for p-levels (tri_hp/movco.cpp:4-40,
 * tri_hp/movfin.cpp:7-14) */
int hpg_mg_restrict(hpg_ctx *ctx);
int hpg_mg_prolongate(hpg_ctx *ctx);
int hpg_get_log2p(hpg_ctx *ctx, int *log2p);

/* precision and preconditioner selection (no reference counterpart: additions
 * behind the same update() entry point).  In HPG_FP32_DEFECT mode hpg_update on
 * the finest level is one defect-correction iteration: the FP64 full residual
 * R64(u0) (left in the residual arrays), then `ninner` multistage iterations of the
 * reference's scheme in FP32 (FP32 MFMA element residual, FP32 minvrt) on
 * F(v) = R32(v) + (1/c - 1) R64(u0) from v = fl32(u0), then u0 += c (v - fl32(u0))
 * in FP64.  F(fl32(u0) + d) ~ R64(u0)/c + J(u0) d, so the inner solve approximates
 * the defect equation for the scaled correction d = delta / c; c keeps |d| near
 * 1e-3 |u0|, where FP32 residual differences resolve J(u0) d.  Coarse p-levels stay
 * FP64.  The FP64 residual stays the judge, so the converged solution is FP64's. */
int hpg_set_precision(hpg_ctx *ctx, int mode);
int hpg_set_preconditioner(hpg_ctx *ctx, int mode);
int hpg_set_defect_inner(hpg_ctx *ctx, int ninner);
/* Replaces the pseudo-time CFL numbers of the first n p-levels (cfl[log2p], n <= 4):
 * the per-level "cfl" list of the reference's input map (tri_hp/allocate.cpp:275-276),
 * changed between cycles instead of at construction.  Returns -2 for n outside 1..4 or
 * a CFL that is negative or not finite. */
int hpg_set_cfl(hpg_ctx *ctx, const double *cfl, int n);
/* Curved boundary sides (ins): the side-mode coefficients of the coordinates of each
 * listed boundary side, crds (nside, sm0, 2) in the side's direction, the array
 * hp_edge_bdry::crds that the reference's curv_init fills (tri_hp/hp_boundary.cpp:
 * 621-675) and crdtocht / crdtocht1d read (tri_hp/gtol.cpp:204-320).  The elements that
 * own one then take the curved branch of element_rsdl (tri_hp/ins/rsdl.cpp:83-257) and
 * of setup_preconditioner (tri_hp/ins/setup_preconditioner.cpp:44-90), and their
 * boundary edges the curved normals (tri_hp/hp_boundary.cpp:973-1015), on every level
 * with p >= 2 (coarser p-levels use the leading modes; at p = 1 every element is
 * straight, as tri_hp::metric::setinfo, tri_hp/metric.cpp:45-70).  nside = 0 makes every
 * side straight again.  Returns -2 for a side that is not a boundary side, a side listed
 * twice or non-finite coefficients; the element-Jacobian and linearised entry points
 * return -2 while curved sides are set at p >= 2. */
int hpg_set_curved(hpg_ctx *ctx, int nside, const int *sides, const double *crds);
/* Boundary flux of each edge boundary (ins), the ebdry objects that
 * tri_hp_ins::getnewedgeobject creates (tri_hp/ins/getnewbdry.cpp:70-110):
 *   HPG_EDGE_GENERIC        inflow, outflow, symmetry (bdry_ins::generic::flux,
 *                           tri_hp/ins/bdry_ins.h:34-49; the three differ only in their
 *                           essential masks: u and v, none, the direction `dir`);
 *   HPG_EDGE_EULER          bdry_ins::euler (bdry_ins.h:96-113);
 *   HPG_EDGE_CHARACTERISTIC bdry_ins::characteristic (tri_hp/ins/bdry.cpp:245-320).
 * types[b] for edge boundary b (mesh order); ibc = the freestream state (u, v, p) of
 * the initial/boundary-condition function (ibc_ins::freestream, tri_hp/ins/getnewibc.cpp:
 * 16-51: (flowspeed cos a, flowspeed sin a, 0)); its pressure replaces
 * hpg_params.bc_pressure.  Default: every boundary generic, velocity (0, 0).  Returns
 * -2 for an unknown type, fewer types than edge boundaries or a non-finite state. */
enum { HPG_EDGE_GENERIC = 0, HPG_EDGE_EULER = 1, HPG_EDGE_CHARACTERISTIC = 2 };
/* Periodic boundaries (ins), the reference's prdc edge boundaries (tri_mesh/tri_boundary.h:
 * 113-148, getnewbdry.cpp:144-146): pairs[2k], pairs[2k+1] are two edge boundaries, side j
 * of the first matched with side n-1-j of the second as the comm boundaries pair them
 * (tri_mesh/tri_boundary.cpp:344-470); the second must be the first moved by one
 * translation.  The reference keeps a partial residual on each copy and averages the
 * preconditioner sums and the solved vertex and side values across the pair
 * (tri_hp/setup_preconditioner.cpp:24-36, 171-176, 255-258, 331-334); here every copy
 * holds the assembled residual, which gives the same update.  Residuals read back hold
 * the assembled value on every copy (the reference's are the partial sums of each side of
 * the pair).  The state's copies are made equal (from the first boundary's sides) here
 * and by every hpg_set_state.  The periodic boundaries carry no flux and no essential
 * mask.  Call before hpg_set_edge_flux and hpg_set_curved, once; not on h-multigrid
 * levels.  The linearised and Jacobian-probe entry points return -2 on periodic
 * contexts.  Returns -2 for boundaries that do not pair (side counts, a translation). */
int hpg_set_periodic(hpg_ctx *ctx, int npair, const int *pairs);
int hpg_set_edge_flux(hpg_ctx *ctx, int nebd, const int *types, const double *ibc);
/* J(u) delta at the current state through the FP32 linearised kernels (finest
 * level; overwrites the residual arrays with R(u)).  No reference counterpart;
 * the directional derivative of tri_hp::rsdl(nstage). */
int hpg_apply_jacobian(hpg_ctx *ctx, const double *dv, const double *ds, const double *di,
                       double *outv, double *outs, double *outi);
/* Test support, no reference counterpart: the full residual of tri_hp::rsdl(nstage)
 * (tri_hp/nstage.cpp:12-110) evaluated by the kernels of the defect correction's inner
 * solve (finest level): ins in FP32 at fl32(u) (element and edge kernels); cns with the
 * element residual in FP32 arithmetic and the boundary fluxes in FP64 (the residual
 * arrays are overwritten). */
int hpg_residual_fp32(hpg_ctx *ctx, double *resv, double *ress, double *resi);
/* Replaces tri_hp::element_jacobian (tri_hp/nstage.cpp:131-165, fixed mesh) for a batch
 * of `nel` distinct elements `tind` on the current level, ins or cns: forward
 * differences of element_rsdl(tind, 0) with dw(var) = eps_r * sum of |uht(var)| over
 * the vertex modes + eps_a (the reference's eps_r = eps_a = 1e-8,
 * tri_hp/hp_boundary.h:19-23).  K (host, nel x N x N, N = tm * NV, row-major) holds
 * K(i*NV + n, mode*NV + var) = d(lf_re + lf_im)(n)(i) / d uht(var)(mode).  The state
 * is left bit-identical; the residual arrays are untouched.  Returns -1 with an
 * error for an index out of range or listed twice.  After any other negative return
 * (a HIP failure in the middle of the column loop) the state may still hold one
 * perturbed coefficient per batch element: set it again before further use. */
int hpg_element_jacobian(hpg_ctx *ctx, const int *tind, int nel, double eps_r, double eps_a, double *K);

/* Learned preconditioner (no reference counterpart).  An MLP 11 -> 64 -> 64 -> 64 -> 9
 * (GELU, bfloat16 weights, FP32 accumulation) maps element features (vertex offsets,
 * mean u, v, p, log viscous scale, log analytic pseudo-time step) to log scalings
 * y[type][var] of the analytic element scaling tprcn of
 * tri_hp_ins::setup_preconditioner (tri_hp/ins/setup_preconditioner.cpp:202-219) for
 * the vertex (type 0), side (1) and interior (2) rows of the diagonal preconditioner;
 * with HPG_PRCN_LEARNED and an MLP attached, hpg_setup_preconditioner on every p-level
 * of the context assembles vprcn from tprcn * exp(clamp(y[0], -2, 2)), sprcn and the
 * side couplings of minvrt from the type-1 scaling and the interior solve from the
 * type-2 scaling (the features carry the level through the analytic pseudo-time step;
 * each coarse-mesh context of a multigrid stack takes the MLP and the mode on its own).
 * Training data come from the method itself:
 * the features and the measured element-Jacobian diagonal per mode type (sum over the
 * type's modes of |J_e(mn, mn)| per variable n).  The MLP must not be trained while a
 * context that uses it runs. */
typedef struct hpg_mlp hpg_mlp;
int hpg_mlp_create(unsigned seed, hpg_mlp **out);
int hpg_mlp_destroy(hpg_mlp *m);
int hpg_mlp_num_params(void);
int hpg_mlp_get_weights(hpg_mlp *m, float *w);
int hpg_mlp_set_weights(hpg_mlp *m, const float *w);
int hpg_mlp_forward(hpg_mlp *m, const float *X, int n, float *Y);
int hpg_mlp_loss_grad(hpg_mlp *m, const float *X, const float *T, int n, float *loss, float *grad);
int hpg_mlp_train(hpg_mlp *m, const float *X, const float *T, int n, int epochs, int batch, float lr, float *last_loss);
/* No counterpart in the reference: an allocation-free form of hpg_mlp_loss_grad for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * training loops, harnesses that forbid hipMalloc).  X (n x 11), T (n x 9), loss (1)
 * and grad (hpg_mlp_num_params()) are DEVICE pointers of the caller; the per-block
 * partial gradients live in a 16 MiB static device array that the library holds on
 * each GPU until the process ends.  Returns -2 with nothing enqueued when
 * n > HPG_MLP_NOALLOC_MAX_SAMPLES.  Loss and gradient are bitwise those of
 * hpg_mlp_loss_grad; the call returns after its work is done.  Calls that overlap on
 * one GPU (from several threads or MLPs) share the array and race. */
#define HPG_MLP_NOALLOC_MAX_SAMPLES 27712
int hpg_mlp_loss_grad_noalloc(hpg_mlp *m, const float *X, const float *T, int n, float *loss, float *grad);
int hpg_set_mlp(hpg_ctx *ctx, hpg_mlp *m);
/* study support: log scalings y (ntri x 9, [type][var]) used in place of the MLP's output
 * while HPG_PRCN_LEARNED is selected (y = NULL returns to the MLP) */
int hpg_set_prcn_log_scalings(hpg_ctx *ctx, const float *y);
int hpg_mlp_features(hpg_ctx *ctx, float *out);             /* ntri x 11 */
int hpg_measure_jacobian_diag(hpg_ctx *ctx, double *out);   /* ntri x 3 x 3 ([type][var]), finest level */

/* blocks until the context's stream is idle */
int hpg_synchronize(hpg_ctx *ctx);
/* Measurement support, no reference counterpart: bracket every element-residual
 * kernel launch with HIP events on the context's stream (enable != 0; clears the
 * record), and read back the summed duration, launch count and elements processed
 * since enabling of one kind of launch: `full` = 1 the FP64 full (beta > 0) residual,
 * 0 the FP64 inviscid-only one, 3 / 2 the same for the FP32 kernels of the
 * defect-correction inner solve. */
int hpg_kernel_timing(hpg_ctx *ctx, int enable);
int hpg_kernel_timing_read(hpg_ctx *ctx, int full, double *total_ms, long long *nlaunch, long long *nelem);
/* text of the last error on this context (static storage, never NULL) */
const char *hpg_last_error(hpg_ctx *ctx);

/* Test support, no reference counterpart.
 * hpg_basis_operator: host-built operator of the p basis by name ("P", "Pr", "Ps",
 *   "I", "Ir", "Is", "Dr", "Ds", "xg", "wx", "eg", "we", "g1", "dg1", "vdiag", "sfmv",
 *   "ifmb", "vfms", "sdiag", "sfms", "bfmi", "idiag", "mass"); returns the length,
 *   copies into out when non-NULL; no device work (tri_basis tables,
 *   tri_basis/initialize.cpp:161-958).
 * hpg_selftest_mfma: one v_mfma_f64_16x16x4f64 on A (16x4) and B (4x16), row-major;
 *   out[lane*4 + i] is accumulator register i of lane `lane`. */
int hpg_basis_operator(int p, const char *name, double *out);
int hpg_selftest_mfma(const double *A, const double *B, double *out);
/* the same for v_mfma_f32_16x16x4f32 */
int hpg_selftest_mfma_f32(const float *A, const float *B, float *out);
/* test support, no reference counterpart: one v_mfma_f32_16x16x16_bf16 on A and B
 * (16x16 row-major, rounded to bfloat16 on the device); out[lane*4 + i] is
 * accumulator register i of lane `lane` (the layout the MLP inference kernel uses). */
int hpg_selftest_mfma_bf16(const float *A, const float *B, float *out);
/* test support, no reference counterpart: `reps` plain device copies of nbytes
 * (multiple of 16) between two fresh buffers; the known-size kernel that HBM
 * traffic counters of the other kernels are scaled against. */
int hpg_selftest_copy(long long nbytes, int reps, double *ms_per_copy);

/* ---- tet_hp_ins (straight tetrahedra, p = 1, 2, NV = 4: u, v, w, p) and tet_hp_cns
 * (NV = 5: p, u, v, w, RT; tet_hp/cns/tet_hp_cns.h:19-66) ----------------------------
 * One context per tetrahedral block.  Mesh arrays of tet_mesh (tet_mesh/tet_mesh.h:
 * 73-137): edge i of a tet joins its local vertices (2,3), (3,1), (2,1), (1,0), (2,0),
 * (3,0) with sign +1 when the edge's first vertex is the first of that pair
 * (tet_mesh/setup.cpp:230-275); tets positively oriented; boundary faces ordered so
 * that (p2 - p1) x (p0 - p1) points out, edge j of a face opposite its vertex j.
 * Every boundary face is an inflow face: ins (tet_hp/ins/bdry_ins.h:118-183) u, v, w
 * essential on its vertices and edges, continuity flux rho u.n through it; cns
 * (tet_hp/cns/bdry_cns.h:70-98, bdry.cpp:129-150) u, v, w, RT essential, the five
 * fluxes with the solution's pressure and the boundary state bc[1..4].  The cns path
 * uses the Weiss-Smith preconditioner (preconditioner = 1, the reference's default). */
typedef struct hpg_tet_ctx hpg_tet_ctx;
typedef struct {
    int npnt, nseg, ntet, nbf;
    const double *pnts;      /* npnt x 3 */
    const int *seg_pnt;      /* nseg x 2 */
    const int *tet_pnt;      /* ntet x 4 */
    const int *tet_seg;      /* ntet x 6 */
    const int *tet_sgn;      /* ntet x 6 */
    const int *bf_pnt;       /* nbf x 3 */
    const int *bf_seg;       /* nbf x 3 */
    const int *bf_sgn;       /* nbf x 3 */
} hpg_tet_mesh;
/* replaces tet_hp::init / tet_hp_ins::init / tet_hp_cns::init (tet_hp/allocate.cpp,
 * tet_hp/ins/allocate.cpp, tet_hp/cns/allocate.cpp); prm: physics HPG_PHYSICS_INS or
 * HPG_PHYSICS_CNS, log2p 0 or 1, rho (ins), mu, adis, fadd, nstage, alpha, beta, cfl;
 * cns also gamma, prandtl, Rgas, bc[1..4] */
int hpg_tet_create(const hpg_tet_mesh *mesh, const hpg_params *prm, hpg_tet_ctx **out);
int hpg_tet_destroy(hpg_tet_ctx *ctx);
/* tet_hp::ug (tet_hp/tet_hp.h:89-94): v(npnt,NV), e(nseg,em0,NV), em0 = p - 1 at log2p */
int hpg_tet_set_state(hpg_tet_ctx *ctx, const double *ugv, const double *uge);
int hpg_tet_get_state(hpg_tet_ctx *ctx, double *ugv, double *uge);
/* tet_hp::hp_global::res (tet_hp/tet_hp.h:120-130) */
int hpg_tet_get_residual(hpg_tet_ctx *ctx, double *resv, double *rese);
int hpg_tet_set_residual(hpg_tet_ctx *ctx, const double *resv, const double *rese);
/* vprcn, eprcn (inverted), iprcn (per element), NV values each, and tau: ins
 * tet_hp_ins::tau (ntet x 4), cns tet_hp_cns::tau (ntet x 5 x 5, adis Tau / jcb) */
int hpg_tet_get_preconditioner(hpg_tet_ctx *ctx, double *vprcn, double *eprcn, double *iprcn, double *tau);
/* cns: hp_cns_gbl->vpreconditioner (npnt x 5 x 5) and epreconditioner (nseg x 5 x 5)
 * (tet_hp/cns/setup_preconditioner.cpp:152-183) */
int hpg_tet_get_cns_blocks(hpg_tet_ctx *ctx, double *vpc, double *epc);
/* replaces tet_hp::rsdl(stage) (tet_hp/update.cpp:9-131, tet_hp.h:229) with
 * tet_hp_ins::element_rsdl (tet_hp/ins/rsdl.cpp:15-785) or tet_hp_cns::element_rsdl
 * (tet_hp/cns/rsdl.cpp:16-843), straight, steady */
int hpg_tet_rsdl(hpg_tet_ctx *ctx, int stage);
/* replaces tet_hp_ins::setup_preconditioner (tet_hp/ins/setup_preconditioner.cpp:7-225)
 * or tet_hp_cns::setup_preconditioner (tet_hp/cns/setup_preconditioner.cpp:9-489), then
 * tet_hp::setup_preconditioner (tet_hp/setup_preconditioner.cpp:428-525); returns the
 * error count (cns: the reference aborts on a NaN or a non-positive element size) */
int hpg_tet_setup_preconditioner(hpg_tet_ctx *ctx);
/* replaces tet_hp::minvrt (tet_hp/setup_preconditioner.cpp:11-351) or
 * tet_hp_cns::minvrt (tet_hp/cns/update.cpp:146-347), p <= 2 */
int hpg_tet_minvrt(hpg_tet_ctx *ctx);
/* replaces tet_hp::update (tet_hp/update.cpp:181-417) or tet_hp_cns::update
 * (tet_hp/cns/update.cpp:18-143) */
int hpg_tet_update(hpg_tet_ctx *ctx);
/* Replaces tet_hp::element_jacobian (tet_hp/update.cpp:134-178) for a batch of `nel`
 * distinct tetrahedra on the current level (ins or cns, p <= 2): forward differences
 * of element_rsdl(tind, 0) with one dw for all variables, eps_r * (sum of |uht(n)(i)|
 * over the four vertices and all variables) + eps_a (the reference's eps_r = 1e-6,
 * eps_a = 1e-10).  K (host, nel x N x N, N = tm * NV) holds
 * K(i*NV + n, mode*NV + var).  The state is left bit-identical; the residual arrays
 * are untouched (the element records are overwritten).  The batch is split into
 * vertex-disjoint sets by colouring; each column evaluates only the set's elements,
 * so every batch element is evaluated tm * NV + 1 times, whatever the mesh size.  After a
 * negative return other than a rejected batch the state may hold perturbed
 * coefficients: set it again before further use. */
int hpg_tet_element_jacobian(hpg_tet_ctx *ctx, const int *tind, int nel, double eps_r, double eps_a, double *K);
/* replaces tet_hp::maxres (tet_hp/allocate.cpp:530-559) */
int hpg_tet_maxres(hpg_tet_ctx *ctx, double *per_var, double *out);
/* max |res| over vertex and edge modes, essential components excluded */
int hpg_tet_resnorm(hpg_tet_ctx *ctx, double *out);
/* FP32 defect correction of tet_hp::update (no reference counterpart; the form of
 * hpg_set_precision for tri contexts): with HPG_FP32_DEFECT each hpg_tet_update on the
 * finest level evaluates the FP64 residual R64(u0), runs `ninner` multistage iterations
 * whose element residuals (stages after the first) are evaluated in FP32 arithmetic on
 * F(v) = R(v) + (1/c - 1) R64(u0) from v = fl32(u0), with the FP64 preconditioner and
 * stage update, then sets u = u0 + c (v - fl32(u0)); c keeps |v - fl32(u0)| near
 * 1e-3 |u0|.  The FP64 residual stays the judge of convergence. */
int hpg_tet_set_precision(hpg_tet_ctx *ctx, int mode, int ninner);
/* test support: the full residual with the element residual in FP32 arithmetic (the
 * inner-solve kernels); overwrites the residual arrays */
int hpg_tet_residual_fp32(hpg_tet_ctx *ctx, double *resv, double *rese);
/* replace tet_hp::mg_restrict / mg_prolongate, p-levels (tet_hp/mg_restrict.cpp:4-25,
 * tet_hp/mg_prolongate.cpp:7-14) */
int hpg_tet_mg_restrict(hpg_tet_ctx *ctx);
int hpg_tet_mg_prolongate(hpg_tet_ctx *ctx);
int hpg_tet_get_log2p(hpg_tet_ctx *ctx, int *log2p);
int hpg_tet_synchronize(hpg_tet_ctx *ctx);
const char *hpg_tet_last_error(hpg_tet_ctx *ctx);
/* per-launch HIP-event timing of the tet element kernel (full = beta > 0 stages) */
int hpg_tet_kernel_timing(hpg_tet_ctx *ctx, int on);
int hpg_tet_kernel_timing_read(hpg_tet_ctx *ctx, int full, double *total_ms, long long *launches,
                               long long *elements);

#ifdef __cplusplus
}
#endif
#endif
