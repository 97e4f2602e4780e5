/*
 * gospl_hip.h — C ABI of the MI355X (gfx950) implementation of goSPL's
 * per-timestep surface-process hot path.
 *
 * Plain pointers and sizes only.  Unless stated otherwise every array pointer
 * is a DEVICE pointer owned by the caller, and every per-vertex array is in
 * the mesh's device (space-filling-curve) order.  Receiver tables are
 * column-major (nrcv, N): entry k of vertex i is at [k * N + i].
 * All functions return 0 on success and nonzero on error; the message is
 * available from gospl_last_error().  Kernels are enqueued on `stream`;
 * functions that must return a data-dependent count to the host
 * (gospl_pit_info, gospl_flowacc's level count) synchronise that stream.
 *
 * Each declaration names the goSPL interface it replaces.  goSPL reaches
 * these through the f2py module `gospl._fortran` (fortran/functions.pyf) and
 * PETSc KSP solves inside its Python mixins.
 */
#ifndef GOSPL_HIP_H
#define GOSPL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gospl_mesh gospl_mesh_t;
typedef struct ihipStream_t* gospl_stream_t; /* a hipStream_t (NULL = default stream) */

/* Message of the last failing call on this thread. */
const char* gospl_last_error(void);

/*
 * Host function.  FV geometry of a triangulation: replaces `definetin`
 * (fortran/functions.pyf:297-308, functions.F90:2715-2940) as called from
 * UnstMesh._meshStructure (gospl/mesher/unstructuredmesh.py:150-160).
 * HOST arrays: coords (nb,3), cells_nodes (n,3), cells_edges (n,3),
 * edges_nodes (m,2), circumcenter (n,3), all row-major; outputs ngbID,
 * eLgt, vDist (nb,12) row-major with -1 / 0 padding, nnb (nb), area (nb).
 * curved = 1 for a sphere (setcurvedmesh, functions.pyf:44-46).
 */
int gospl_fv_geometry(const double* coords, int64_t nb, const int32_t* cells_nodes,
                      const int32_t* cells_edges, int64_t n, const int32_t* edges_nodes,
                      int64_t m, const double* circumcenter, int curved, int32_t* ngbID,
                      int32_t* nnb, double* eLgt, double* vDist, double* area);

/*
 * Mesh handle: replaces the Fortran module state `meshparams`
 * (functions.F90:10-36) that `definetin` fills.  The geometry arrays stay
 * owned by the caller and must outlive the handle; the handle owns O(N)
 * device workspace.  nbr/elgt/vdist are (W, N) column-major ELL tables.
 * Neighbour relations must be mutual (they are for a triangulation whose
 * vertices have at most 12 neighbours); the call fails otherwise.
 */
int gospl_mesh_create(int64_t N, int32_t W, int32_t flat, const int32_t* nbr,
                      const int32_t* deg, const double* elgt, const double* vdist,
                      const double* area, const int32_t* gid, const int8_t* outlet,
                      gospl_mesh_t** out);
int gospl_mesh_destroy(gospl_mesh_t* mesh);

/*
 * The caller changed, in place, mesh arrays it lent to the handle (e.g. the
 * areas, as goSPL's `updatearea`, functions.F90:2699-2713): drop what the
 * handle derived from them and keeps between calls (the hillslope rows).
 * No counterpart in goSPL (its operator is rebuilt from its own arrays).
 */
int gospl_mesh_touch(gospl_mesh_t* mesh);

/*
 * Multiple-flow-direction receivers: replaces `mfdreceivers`
 * (functions.pyf:146-156, functions.F90:1544-1671).  elev (N) -> rcv, dist,
 * wgt (nrcv, N).  nrcv <= 12.
 */
int gospl_mfdreceivers(gospl_mesh_t* mesh, int32_t nrcv, double flowexp, const double* elev,
                       double sealevel, int32_t* rcv, double* dist, double* wgt,
                       gospl_stream_t stream);

/*
 * FAMesh._buildFlowDirection (gospl/flow/flowplex.py:480-544): receivers on
 * h, then (down != 0) flat-depression routing along flatdirs where
 * h == lfill and the cell has no downslope neighbour, then flat-mesh outlets
 * as self-receivers, then lsink = no receiver weight & not outlet & not sea
 * (sea = lfill <= sealevel).  pitids/flatdirs are read only when down != 0.
 */
int gospl_flow_direction(gospl_mesh_t* mesh, int32_t nrcv, double flowexp, const double* h,
                         double sealevel, const double* lfill, int32_t down,
                         const int32_t* pitids, const int32_t* flatdirs, int32_t* rcv,
                         double* dist, double* wgt, int8_t* lsink, gospl_stream_t stream);

/*
 * Depression filling: replaces PITFill._performFilling and the level shift
 * of fillElevation (gospl/flow/pitfilling.py:557-702, 864-870), i.e. the
 * Fortran `edge_tile`, `fill_tile`, `graph_nodes`, `combine_edges`,
 * `fill_edges`, `fill_depressions` (functions.pyf:179-241).
 * h (N) -> lfill (N) = ((filled h - level) + level) with cells whose fill is
 * below sealevel reset to (h - level) + level.
 */
int gospl_fill_depressions(gospl_mesh_t* mesh, const double* h, double level,
                           double sealevel, double* lfill, gospl_stream_t stream);

/*
 * fillElevation (gospl/flow/pitfilling.py:849-870) with the level chosen on
 * the device, so nothing is read back: level = max(min(h) (+ 1e-3 when
 * add_eps, goSPL's sphere case), floor_level) with Python's max, floor_level
 * = sealevel + oFill as goSPL computes it; then as gospl_fill_depressions.
 */
int gospl_fill_elevation(gospl_mesh_t* mesh, const double* h, int32_t add_eps, double floor_level,
                         double sealevel, double* lfill, gospl_stream_t stream);

/*
 * Depression information: replaces PITFill._pitInformation and the
 * depth-volume curves of fillElevation (gospl/flow/pitfilling.py:704-847,
 * 872-891), i.e. the Fortran `label_pits`, `pits_cons`, `fill_dir`,
 * `nghb_dir`, `getpitvol` (functions.pyf:172-177, 243-277) plus the
 * numpy spill-point, flat-receiver and pit-parameter code.
 * In: hl (N) unfilled elevation, lfill (N) from gospl_fill_depressions.
 * Out: lfill (marine cells reset to hl), pitids (N, 1..npits or -1),
 * flatdirs (N), ptdir (N), lspill (N int8).  Per-pit arrays have capacity
 * `cap` rows (index 0 unused, pits 1..npits): spill (cap), pitparams
 * (3, cap) = volume, max fill, max depth; filled_lvl (5, cap) and
 * filled_vol (5, cap); rows past npits are not written.  *npits (HOST)
 * receives the pit count; the call fails if npits + 1 > cap.  Per-pit sums are reduced in a fixed order, so
 * results are reproducible run to run (their last bits differ from numpy's
 * group-by order).
 */
int gospl_pit_info(gospl_mesh_t* mesh, const double* hl, double* lfill, double sealevel,
                   int32_t* pitids, int32_t* flatdirs, int32_t* ptdir, int8_t* lspill,
                   int64_t cap, int32_t* spill, double* pitparams, double* filled_lvl,
                   double* filled_vol, int64_t* npits, gospl_stream_t stream);

/*
 * Rows 0..npits of gospl_pit_info's per-pit tables (same buffers and `cap`)
 * read back to HOST memory `out` in one transfer: 14 * (npits + 1) doubles,
 * one block of npits + 1 after the other = spill (as f64, exact), pitparams
 * rows 0-2, filled_lvl rows 0-4, filled_vol rows 0-4 (goSPL's pitParams,
 * filled_lvl and filled_vol with pitInfo's spill ids, pitfilling.py:521-555,
 * 872-891).  Synchronises `stream`; pinned `out` makes the copy direct.
 * No counterpart in goSPL (its tables are host arrays already).
 */
int gospl_pit_tables(gospl_mesh_t* mesh, int64_t npits, int64_t cap, const int32_t* spill,
                     const double* pitparams, const double* filled_lvl,
                     const double* filled_vol, double* out, gospl_stream_t stream);

/*
 * Flow accumulation solve: replaces FAMesh._solve_KSP on the flow matrix
 * (I - W)^T built by matrixFlow (gospl/flow/flowplex.py:323-478, 843) with an
 * exact sweep over the receiver DAG.  rcv/wgt (nrcv, N), b (N) -> q (N).
 * *nlevels (HOST, may be NULL) receives the DAG depth when the sweep runs
 * level by level, and 0 for the dataflow sweep (the default for up to three
 * receivers; GOSPL_DAG_CHAIN=0 at mesh creation selects the level sweep for
 * this call and gospl_spl).  Both give the same bits.  Receivers built with
 * flat routing on a partly filled surface can hold a cycle; the vertices the
 * sweep cannot reach (the cycle and everything below it) are then solved by
 * the fixed-point iteration q <- b + W^T q restricted to them, to a relative
 * change of 1e-14, as goSPL's iterative solve converges on a draining cycle.
 * *unsolved (HOST, may be NULL) receives 0, or the number of those vertices
 * when the iteration does not converge (an undrained cycle, on which goSPL's
 * solvers give up, flowplex.py:178-322; q is then not meaningful there).
 */
int gospl_flowacc(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcv, const double* wgt,
                  const double* b, double* q, int32_t* nlevels, int64_t* unsolved,
                  gospl_stream_t stream);

/*
 * Stream-power erosion: replaces SPL._eroMats + SPL._getEroDepRate and the
 * state update of SPL.erodepSPL (gospl/eroder/SPL.py:45-114, 260-366,
 * 402-438) for n = 1 and no deposition (G = 0).
 * kfac (N, may be NULL) = per-vertex K multiplier (sedfactor * rain^d);
 * sea = lfill <= sealevel.  Updates h, cumed; writes eb (N) =
 * erosion/deposition rate with outlets (flat meshes) and lsink cells zeroed,
 * and eblocal = (eb*dt)/dt.  *nlevels (HOST, may be NULL) as for
 * gospl_flowacc.
 */
int gospl_spl(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcvi, const double* wgti,
              const double* disti, const double* fal, const double* kfac, double K, double m,
              double dt, double sealevel, const double* lfill, const int8_t* lsink, double* h,
              double* cumed, double* eb, double* eblocal, int32_t* nlevels,
              gospl_stream_t stream);

/*
 * Non-linear stream power, detachment-limited (n != 1, G = 0): replaces
 * nlSPL._solveNL and the update of erodepSPLnl (gospl/eroder/nlSPL.py:321-396,
 * 398-456) with local_spl / local_spl_coeff (functions.F90:1469-1542).
 * F(x)_i = x_i - h_i + Kbr_i S_i(x)^n = 0 with S_i the weighted positive slope to
 * the receivers (w > 0) and Kbr = K kfac dt FAL^m elimiter (0 where lfill <=
 * sealevel; elimiter = dh/(dh + 0.01) of the largest drop dh to a receiver).
 * The system is lower-triangular in the receiver order, so each Newton step
 * is one exact sweep; Newton runs until max|step| <= tol max(max|x|, 1)
 * (goSPL's SNES stops at rtol/atol 1e-6), at most maxit steps, failing
 * otherwise.  Then the state update of gospl_spl.  *iters / *step_rel (HOST,
 * may be NULL) receive the Newton steps and the last relative step.
 * Synchronises `stream` once per step.
 */
int gospl_spl_nl(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                 const double* disti, const double* fal, const double* kfac, double K, double m, double n,
                 double dt, double sealevel, const double* lfill, const int8_t* lsink, double* h,
                 double* cumed, double* eb, double* eblocal, double tol, int32_t maxit, int32_t* iters,
                 double* step_rel, gospl_stream_t stream);

/*
 * Transport-limited stream power (G > 0, n = 1): replaces
 * SPL._coupledEDSystem and the G > 0 branch of SPL._getEroDepRate
 * (gospl/eroder/SPL.py:116-258, 306-333) and the update of erodepSPL.
 * goSPL's nested system in elevation h and sediment flux q,
 *   [eMat - diag(fDep)  -diag(fDep dt/area)] [h]   [(1 - fDep) hOld]
 *   [diag(area/dt)       fMati             ] [q] = [hOld area / dt ]
 * with eMat the rows of gospl_spl, fMati = (I - W)^T of the same receivers
 * (rcvi, wgti) and fDep = G area / max(FAL, G area / 0.99) (0 where FAL = 0,
 * where lfill <= sealevel and at flat-mesh outlets, at most 0.99), is reduced
 * to its elevation block and solved by restarted GMRES right-preconditioned
 * by the stream-power sweep (each iteration one exact sweep in each
 * direction of the receiver DAG), from hOld, until the relative residual of
 * the elevation block is <= rtol or maxit iterations (goSPL: TFQMR with a
 * Schur fieldsplit to rtol 1e-8; on non-convergence goSPL warns and keeps
 * the iterate, as this entry point does: *relres > rtol then tells).
 * *iters / *relres (HOST, may be NULL).  qs (N, may be NULL) receives the
 * sediment flux q of the solution.  Then the state update of gospl_spl.
 * Synchronises `stream` several times per iteration.
 */
int gospl_spl_dep(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                  const double* disti, const double* fal, const double* kfac, double K, double m, double G,
                  double dt, double sealevel, const double* lfill, const int8_t* lsink, double* h,
                  double* cumed, double* eb, double* eblocal, double rtol, int32_t maxit, int32_t* iters,
                  double* relres, double* qs, gospl_stream_t stream);

/*
 * Non-linear transport-limited stream power (G > 0, n != 1): replaces
 * nlSPL._solveNL_ed / _form_residual_ed and the update of erodepSPLnl
 * (gospl/eroder/nlSPL.py:82-125, 255-320, 398-456).  The root of
 *   F(h) = (h - hOld)(1 - fDep) + Kbr S(h)^n - fDep dt max(Qt(h), 0) / area,
 *   Qt(h) = fMati^-1 ((hOld - h) area / dt),
 * (Kbr, S as gospl_spl_nl; fDep as gospl_spl_dep) by Newton steps, each
 * linear system J d = -F (the triangular slope part plus the flux coupling
 * where Qt > 0) solved by the GMRES of gospl_spl_dep to a relative residual
 * of 1e-13, until max|d| <= tol max(max|h|, 1) (goSPL's SNES: L-BFGS to
 * rtol/atol 1e-6), at most maxit steps, failing otherwise.  *iters / *inner /
 * *step_rel (HOST, may be NULL): Newton steps, GMRES iterations in all and
 * the last relative step.  Then the state update of gospl_spl.
 */
int gospl_spl_nl_dep(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                     const double* disti, const double* fal, const double* kfac, double K, double m, double n,
                     double G, double dt, double sealevel, const double* lfill, const int8_t* lsink, double* h,
                     double* cumed, double* eb, double* eblocal, double tol, int32_t maxit, int32_t* iters,
                     int32_t* inner, double* step_rel, gospl_stream_t stream);

/*
 * Refined solves (the reduced-cost mode; no counterpart in goSPL, which
 * solves these systems with FGMRES + ILU(0) to rtol 1e-8, flowplex.py:341-408).
 * Iterative refinement: each pass takes the FP64 residual r = rhs - A x of the
 * FP64 iterate x and solves the correction A d = r by one dataflow sweep that
 * accumulates in 64-bit fixed point (one integer atomic per dependency, order
 * independent, so runs are bit-reproducible), x += d; passes stop once
 * ||r||_2 <= rtol ||rhs||_2 or after maxpass corrections.  *passes (HOST, may
 * be NULL) receives the corrections run, or -1 when the solve fell back to
 * the exact sweep of gospl_flowacc / gospl_spl (more than three receivers, a
 * receiver cycle or a fixed-point overflow); *relres (HOST) the final
 * relative residual.  Synchronises `stream`.
 *  - gospl_flowacc_refine: as gospl_flowacc; q is IN/OUT: the initial guess
 *    (goSPL warm-starts from the previous solution, flowplex.py:395-396).
 *  - gospl_spl_refine: as gospl_spl, the iterate starting at the old h.
 */
int gospl_flowacc_refine(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcv, const double* wgt,
                         const double* b, double* q, double rtol, int32_t maxpass,
                         int32_t* passes, double* relres, int64_t* unsolved,
                         gospl_stream_t stream);
int gospl_spl_refine(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                     const double* disti, const double* fal, const double* kfac, double K,
                     double m, double dt, double sealevel, const double* lfill,
                     const int8_t* lsink, double* h, double* cumed, double* eb, double* eblocal,
                     double rtol, int32_t maxpass, int32_t* passes, double* relres,
                     gospl_stream_t stream);

/*
 * Linear hillslope diffusion: replaces hillSLP.getHillslope/_hillSlope(0)
 * (gospl/sed/hillslope.py:171-233, 614-649) with `sethillslopecoeff`
 * (functions.pyf:10-14, functions.F90:901-938) and its Richardson/ILU(0)
 * solve.  Diffusivity cda on land, cdm at sea (h <= sealevel), times dt.
 * Updates h, cumed and eblocal (+= dh/dt).  sweeps > 0: that many Jacobi
 * sweeps (the Python host derives the count from the operator's contraction
 * bound rho = max s/(1+s)); sweeps <= 0: sweeps until the error bound
 * rho/(1-rho) * max|change| is below 2^-52 * max|h| (checked on the device
 * every 32 sweeps, one read-back each), failing past 10^6 sweeps, where goSPL
 * stops its Richardson/ILU(0) solve at rtol 1e-8 (hillslope.py:347-371).
 * The operator rows are kept in the handle between calls: a call rebuilds
 * the rows where the sea flag of the vertex or of a neighbour changed, all
 * of them when cda, cdm or dt change (goSPL keeps its operator until the
 * coastline moves, hillslope.py:184-200); results equal a full rebuild's.
 */
int gospl_hillslope(gospl_mesh_t* mesh, double cda, double cdm, double dt, double sealevel,
                    int32_t sweeps, double* h, double* cumed, double* eblocal,
                    gospl_stream_t stream);

/*
 * Refined hillslope solve (the reduced-cost mode of gospl_hillslope; no
 * counterpart in goSPL): passes of r = h - M x in FP64 (x starts at h),
 * `sweeps32` FP32 Jacobi sweeps of M d = r on an FP32 copy of the kept
 * operator, x += d, until ||r||_2 <= rtol ||h||_2 or maxpass corrections.
 * sweeps32 <= 0 runs the converged FP64 solve of gospl_hillslope instead
 * (*passes = -1).  *passes / *relres (HOST, may both be NULL: then nothing is
 * read back and the stream is not synchronised) receive the corrections run
 * and the last relative residual.
 */
int gospl_hillslope_refine(gospl_mesh_t* mesh, double cda, double cdm, double dt, double sealevel,
                           int32_t sweeps32, double rtol, int32_t maxpass, double* h, double* cumed,
                           double* eblocal, int32_t* passes, double* relres, gospl_stream_t stream);

/*
 * Adjoints (reverse-mode derivatives) of one step's solves with the step's
 * receivers, weights, distances, discharge, fill and sea/sink flags held at
 * their forward values.  No counterpart in goSPL, which has no derivatives;
 * they are the transposes of the systems goSPL assembles in
 * FAMesh.matrixFlow (flowplex.py:429-478), SPL._eroMats (SPL.py:45-114) and
 * hillSLP._hillSlope (hillslope.py:235-293).  All arrays are on the device.
 *  - gospl_flowacc_adjoint: y solves (I - W) y = lam, the transpose of the
 *    flow system (I - W^T) q = b of gospl_flowacc (self receivers dropped as
 *    there): one sweep from the outlets up.  Errors on a receiver cycle.
 *  - gospl_spl_adjoint: reverse of gospl_spl.  With A(h) x = h the forward
 *    system (h the pre-erosion elevation) and h' = x except where the forward
 *    kept h (flat-mesh outlets, lsink cells), given lam = dL/dh' it writes
 *    grad_h = dL/dh (through x and through A's dependence on h_i - h_r),
 *    grad_logk (may be NULL) = dL/dlog10(kfac) per vertex and grad_fal (may
 *    be NULL) = dL/dfal.  x (may be NULL: then solved again) is the forward
 *    solution.  One sweep from the sources down for A^T mu = lam (plus one
 *    the other way when x is NULL).
 *  - gospl_hillslope_frozen: h_out = the `sweeps` Jacobi sweeps of
 *    gospl_hillslope on h_in with the sea flags given per vertex (sea[i] != 0:
 *    diffusivity cdm) instead of taken from h_in; the handle's kept operator
 *    is not touched.
 *  - gospl_hillslope_adjoint: out = M^-T lam by `sweeps` Jacobi sweeps of the
 *    transposed operator (same spectrum, so the forward's count applies),
 *    with M the operator of gospl_hillslope_frozen for the same sea flags.
 * sweeps < 1 is an error for both hillslope entries.
 */
int gospl_flowacc_adjoint(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcv, const double* wgt,
                          const double* lam, double* y, gospl_stream_t stream);
int gospl_spl_adjoint(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                      const double* disti, const double* fal, const double* kfac, double K, double m,
                      double dt, double sealevel, const double* lfill, const int8_t* lsink,
                      const double* h, const double* x, const double* lam, double* grad_h,
                      double* grad_logk, double* grad_fal, gospl_stream_t stream);
int gospl_hillslope_frozen(gospl_mesh_t* mesh, double cda, double cdm, double dt, const int8_t* sea,
                           int32_t sweeps, const double* h_in, double* h_out, gospl_stream_t stream);
int gospl_hillslope_adjoint(gospl_mesh_t* mesh, double cda, double cdm, double dt, const int8_t* sea,
                            int32_t sweeps, const double* lam, double* out, gospl_stream_t stream);

/*
 * Hillslope operator coefficients: replaces `sethillslopecoeff`
 * (functions.pyf:10-14, functions.F90:901-938).  kd (N) = diffusivity * dt
 * per vertex; coeff (13, N): row 0 the diagonal 1 + sum v, row 1 + p the
 * coefficient -v of neighbour p (0 past the vertex's degree), with
 * v = (0.5 * (kd_i + kd_n) / area_i) * vDist / eLgt.  Rows of zero-area
 * vertices are all zero.  Flat-mesh outlets are not special-cased here (that
 * is the operator assembly's job, hillslope.py:253-255).
 */
int gospl_hillslope_coeff(gospl_mesh_t* mesh, const double* kd, double* coeff,
                          gospl_stream_t stream);

/*
 * Runoff source of flowAccumulation (gospl/flow/flowplex.py:794-795):
 * b = rain * area with sea (lfill <= sealevel) zeroed.
 */
int gospl_rain_source(gospl_mesh_t* mesh, const double* rain, const double* lfill,
                      double sealevel, double* b, gospl_stream_t stream);

/*
 * Pit-overflow cascade pieces (FAMesh._distributeDownstream,
 * gospl/flow/flowplex.py:576-686):
 *  - gospl_pit_inflow: inv[p] = sum of fa over lsink cells of pit p (npit rows),
 *    in a fixed order.  It reuses the grouping of cells by pit that the last
 *    gospl_pit_info built for the same `pitids` buffer and pit count (so
 *    `pitids` must not be changed in between), else it builds one, which
 *    synchronises `stream`.
 *  - gospl_pit_water: per-pit water levels; full[p] != 0 sets water = lfill,
 *    else if level[p] is not NaN and water <= level[p], water = level[p].
 *  - gospl_cascade_seed: one pass's device update from the host decisions
 *    (flowplex.py:602-666) in ONE device buffer `stage` of 2*npit + 2*nspill
 *    doubles: full[p] = stage[p] != 0 and level[p] = stage[npit + p] applied
 *    as gospl_pit_water (pits p >= npit untouched), then rhs = 0 except
 *    rhs[stage[2 npit + j]] = stage[2 npit + nspill + j] / dt for j < nspill
 *    (distinct spill vertex ids, exact in f64; ids outside [0, N) are
 *    skipped).  Replaces gospl_pit_water plus the host-side scatter of the
 *    spilled volumes into the next solve's right-hand side.
 *  - gospl_cascade_add: nfa = q*dt; fal += (nfa masked where hl < water)/dt.
 *  - gospl_flow_finalize (flowplex.py:906-933): zero fal at filled pit
 *    cells, fillfal = fal with 0.1*max(fal) there (when has_pits), clamp >= 0.
 */
int gospl_pit_inflow(gospl_mesh_t* mesh, const double* fa, const int8_t* lsink,
                     const int32_t* pitids, int64_t npit, double* inv, gospl_stream_t stream);
int gospl_pit_water(gospl_mesh_t* mesh, const int32_t* pitids, const int8_t* full,
                    const double* level, const double* lfill, double* water,
                    gospl_stream_t stream);
int gospl_cascade_seed(gospl_mesh_t* mesh, const double* stage, int64_t npit, int64_t nspill,
                       const int32_t* pitids, const double* lfill, double* water, double dt,
                       double* rhs, gospl_stream_t stream);
int gospl_cascade_add(gospl_mesh_t* mesh, const double* q, const double* hl,
                      const double* water, double dt, double* nfa, double* fal,
                      gospl_stream_t stream);
int gospl_flow_finalize(gospl_mesh_t* mesh, const double* hl, const double* water,
                        const int32_t* pitids, int32_t has_pits, double* fal,
                        double* fillfal, gospl_stream_t stream);

/*
 * FAMesh.flowAccumulation (gospl/flow/flowplex.py:752-945) in one call: the
 * fill and pit information of fillElevation (pitfilling.py:849-903), the
 * step's receivers (the main set, flowplex.py:786-791), the runoff source, the
 * flow solve, the pit-overflow cascade with goSPL's decisions taken on the
 * host in native code (flowplex.py:576-686, 849-904: the same decisions as a
 * Python driver over the entry points above, bit for bit, including numpy's
 * pairwise summation of the spilled volumes), and the end of the call
 * (flowplex.py:906-933).  No counterpart in goSPL, whose driver is Python.
 * Buffers: every pointer of gospl_flow_buffers is a DEVICE array of N
 * (per-vertex) or `cap` (per-pit) rows; the handle keeps the host tables and
 * staging it needs (grown on demand), and gospl_flow_tables returns the
 * per-pit tables of the last call (HOST, gospl_pit_tables' layout, n1 =
 * npits + 1 rows; valid until the next call).  Results: water = waterFilled - h (goSPL's waterFilled after the
 * call), the main set in rcv0/dist0/wgt0/lsink0, the cascade's last rebuild
 * in set 1; res->last_set says which set goSPL's live arrays (rcvID, lsink)
 * name.  Returns 2 when the main flow solve meets an undrained receiver cycle
 * (goSPL raises there); a cascade solve that does not converge is zeroed, as
 * goSPL does.  Synchronises `stream` (the pit count, the tables, each pass's
 * inflows).
 */
typedef struct {
  int32_t nrcv;
  double flowexp;          /* 1.1 (inputparser.py:194) */
  double sealevel;
  double floor_level;      /* sealevel + oFill */
  int32_t add_eps;         /* 1 on a sphere (fillElevation's +1e-3) */
  double dt;
  double maxarea;          /* the largest cell area */
  int32_t max_passes;      /* 100 */
  int32_t patience;        /* 3 */
  double rel_improve;      /* 1e-3 */
} gospl_flow_params;
typedef struct {
  const double* h;
  const double* rain;
  double* lfill;
  int32_t* pitids;
  int32_t* flatdirs;
  int32_t* ptdir;
  int8_t* lspill;
  int32_t* rcv0;
  double* dist0;
  double* wgt0;
  int8_t* lsink0;
  int32_t* rcv1;
  double* dist1;
  double* wgt1;
  int8_t* lsink1;
  double* b;
  double* fal;
  double* fillfal;
  double* water;
  double* hl;
  double* q;
  double* nfa;
  double* rhs;
  double* fa;
  int64_t cap;
  int32_t* spill;
  double* pitparams;
  double* flvl;
  double* fvol;
  double* inv;
} gospl_flow_buffers;
typedef struct {
  int64_t npits;
  int32_t cascade_passes;
  int32_t flow_solves;
  int64_t undrained;
  int32_t last_set;
} gospl_flow_result;
int gospl_flow_accumulation(gospl_mesh_t* mesh, const gospl_flow_params* params,
                            const gospl_flow_buffers* buffers, gospl_flow_result* result,
                            gospl_stream_t stream);
int gospl_flow_tables(gospl_mesh_t* mesh, const double** tables, int64_t* n1);
/* numpy's pairwise summation of n doubles (HOST), the order np.sum uses. */
double gospl_np_sum(const double* a, int64_t n);

/*
 * gospl_flow_finalize for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops): no mesh handle, and no hipMalloc, hipFree or
 * allocating library call.  No counterpart in goSPL.  Same inputs, outputs and
 * results (bit for bit) as gospl_flow_finalize on an N-vertex mesh; N must be
 * in 1..2^31-1 (the call fails otherwise).  Its scratch is a fixed 8 KiB
 * device array (1024 partial maxima) that each GPU holds until the process
 * ends and that every call on that GPU shares: calls that overlap on one GPU
 * (different streams or threads) race.
 */
int gospl_flow_finalize_noalloc(int64_t N, const double* hl, const double* water,
                                const int32_t* pitids, int32_t has_pits, double* fal,
                                double* fillfal, gospl_stream_t stream);

/*
 * Ensembles in one mesh (no counterpart in goSPL, which runs one model per
 * process): a union mesh of independent members, `seg` (device, N int32)
 * giving each vertex's member 0..nseg-1, or a negative value for padding
 * vertices.  The union's gid must stay a permutation of 0..N-1 (each member's
 * ids plus its offset), which orders pits and spill candidates as each
 * member's own ids do.  Each entry reproduces, per member, what its
 * single-mesh counterpart computes on that member alone:
 *  - gospl_mesh_set_tiebreak: the ids mfdreceivers' slope perturbation
 *    val*(1 - 1e-15 gid) uses (functions.F90:1618-1629), device N int32 owned
 *    by the caller (the members' own ids); NULL restores gid.
 *  - gospl_fill_elevation_seg: gospl_fill_elevation with each member's own
 *    level max(min h (+1e-3), floor_level) (pitfilling.py:849-870); padding
 *    vertices get floor_level.
 *  - gospl_flow_finalize_seg: gospl_flow_finalize with each member's own pit
 *    flag (has_pits, device int8[nseg]) and discharge maximum.
 *  - gospl_spl_batch: gospl_spl with the stream-power exponent m per vertex
 *    (mvec, device N); per-member K goes through kfac.
 *  - gospl_hillslope_refine_seg: gospl_hillslope_refine with per-member FP32
 *    sweep counts (sweeps32, HOST int32[nseg], each >= 1), residual norms and
 *    stop tests.  offsets (HOST int64[nseg+1]) are the members' vertex ranges,
 *    each a run of whole 256-vertex workgroups (padding at its end), from 0 to
 *    N.  passes (HOST int32[nseg], may be NULL: then nothing is read back)
 *    receives each member's correction count.
 */
int gospl_mesh_set_tiebreak(gospl_mesh_t* mesh, const int32_t* tgid);

/*
 * Launch shape of the receiver-DAG dataflow sweeps (no counterpart in
 * goSPL): chain_blocks workgroups (<= the number that fit on the device at
 * once; <= 0 keeps), chain_share = pool size from which a wavefront hands
 * batches to the global queue (< 0 keeps), chain_pollers = idle wavefronts
 * that stay to poll it (< 0 keeps).  Results do not depend on the shape.
 * gospl_dag_get_shape: out (HOST int32[5]) = chain_blocks, chain_share,
 * chain_pollers, the resident limit, the compute-unit count.
 */
int gospl_dag_set_shape(gospl_mesh_t* mesh, int32_t chain_blocks, int32_t chain_share,
                        int32_t chain_pollers);
int gospl_dag_get_shape(gospl_mesh_t* mesh, int32_t* out);
int gospl_fill_elevation_seg(gospl_mesh_t* mesh, const double* h, int32_t add_eps, double floor_level,
                             double sealevel, const int32_t* seg, int32_t nseg, double* lfill,
                             gospl_stream_t stream);
int gospl_flow_finalize_seg(gospl_mesh_t* mesh, const double* hl, const double* water,
                            const int32_t* pitids, const int32_t* seg, int32_t nseg,
                            const int8_t* has_pits, double* fal, double* fillfal, gospl_stream_t stream);
int gospl_spl_batch(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcvi, const double* wgti,
                    const double* disti, const double* fal, const double* kfac, double K,
                    const double* mvec, double dt, double sealevel, const double* lfill,
                    const int8_t* lsink, double* h, double* cumed, double* eb, double* eblocal,
                    int32_t* nlevels, gospl_stream_t stream);
int gospl_hillslope_refine_seg(gospl_mesh_t* mesh, double cda, double cdm, double dt, double sealevel,
                               const int32_t* seg, int32_t nseg, const int64_t* offsets,
                               const int32_t* sweeps32, double rtol, int32_t maxpass, double* h,
                               double* cumed, double* eblocal, int32_t* passes, gospl_stream_t stream);

/*
 * Sediment-mode depression information: replaces PITFill._pitInformation with
 * sed=True (gospl/flow/pitfilling.py:704-847, 885-887 as fillElevation(sed=
 * True) calls it from SEDMesh.sedChange, gospl/sed/sedplex.py:1075).  As
 * gospl_pit_info, except that pits are labelled above the fill `level`
 * (label_pits(level, lFill), :718-719) instead of sea level and the per-pit
 * sums do not clip elevations at sea level (:828-832, 885-887).  `lfill`
 * comes from gospl_fill_depressions(h, level, sealevel = -INFINITY), which
 * keeps the filled values of submerged cells (:690-693).
 */
int gospl_pit_info_sed(gospl_mesh_t* mesh, const double* hl, double* lfill, double level, double sealevel,
                       int32_t* pitids, int32_t* flatdirs, int32_t* ptdir, int8_t* lspill, int64_t cap,
                       int32_t* spill, double* pitparams, double* filled_lvl, double* filled_vol,
                       int64_t* npits, gospl_stream_t stream);

/*
 * Continental sediment routing (SEDMesh, gospl/sed/sedplex.py).  Per-pit
 * arrays are indexed by pit id (rows 0..npit-1, row 0 unused); per-pit sums
 * use gospl_pit_inflow's fixed-order grouping of the last gospl_pit_info(_sed)
 * call for the same `pitids`.
 *  - gospl_sed_source: b = (-eb) * area, the eroded volume rate that
 *    _getSedFlux routes (:109-131).
 *  - gospl_sed_stray: flow-terminal cells outside every pit (lsink &
 *    pitID < 0) drop their volume, kept in closed_depo when `closed` (no
 *    draining border, :236-251).
 *  - gospl_sed_add: one routed pass of _distributeSediment (:424-431): vnext
 *    = q (0 at draining borders), vsedl += (hl < sedfilled ? 0 : q) / dt.
 *  - gospl_sed_closed: h and cumed += closed_depo / area where positive
 *    (:491-500).
 *  - gospl_sed_bottomup: _bottomUpDelta (:525-560): delta = max(0,
 *    min(target[pit], lfill) - hl) on cells of pits with a finite target,
 *    else 0.
 *  - gospl_sed_tilt: _addPitMicroTilt (:584-655): on cells of selected pits
 *    (psel[pit]) with delta > 0, delta = max(0, delta + 1e-6 (d - mean_pit d))
 *    with d the distance to the pit's spill point (spill_xyz, 3 per pit) and
 *    the mean over those cells (also returned in `mean`, npit rows).
 *  - gospl_sed_inlets: _identifyPitInlets (:657-697) on the selected pits:
 *    inlet[c] = 1 where c receives (rcv, nrcv x N) from a donor outside its
 *    pit; count[p] = inlets of pit p.
 *  - gospl_sed_lake_depth: sel = cells of selected pits, wd = max(0,
 *    target[pit] - hl) there, denom[p] = sum of wd * area (:833-871).
 *  - gospl_sed_lake_init: the initial deposit of _diffuseLargePit (:873-898):
 *    ca[pit] wd / max(denom, 1e-30) where denom > 0, plus cb[pit] /
 *    (count[pit] area) at inlets with count > 0.
 *  - gospl_sed_pit_volume: vol[p] = sum of delta * area over sel cells of p.
 *  - gospl_sed_scale: delta *= scale[pit] on sel cells, 0 elsewhere.
 *  - gospl_sed_apply: cumed += delta; h += delta (:1041-1046).
 */
int gospl_sed_source(gospl_mesh_t* mesh, const double* eb, double* b, gospl_stream_t stream);
int gospl_sed_stray(gospl_mesh_t* mesh, const int8_t* lsink, const int32_t* pitids, int32_t closed,
                    double* vsed, double* closed_depo, gospl_stream_t stream);
int gospl_sed_add(gospl_mesh_t* mesh, const double* q, const double* hl, const double* sedfilled, double dt,
                  double* vnext, double* vsedl, gospl_stream_t stream);
int gospl_sed_closed(gospl_mesh_t* mesh, const double* closed_depo, double* h, double* cumed,
                     gospl_stream_t stream);
int gospl_sed_bottomup(gospl_mesh_t* mesh, const int32_t* pitids, const double* target, const double* lfill,
                       const double* hl, double* delta, gospl_stream_t stream);
int gospl_sed_tilt(gospl_mesh_t* mesh, const double* xyz, const int32_t* pitids, int64_t npit,
                   const int8_t* psel, const double* spill_xyz, double* delta, double* mean,
                   gospl_stream_t stream);
int gospl_sed_inlets(gospl_mesh_t* mesh, int32_t nrcv, const int32_t* rcv, const int32_t* pitids, int64_t npit,
                     const int8_t* psel, int8_t* inlet, double* count, gospl_stream_t stream);
int gospl_sed_lake_depth(gospl_mesh_t* mesh, const int32_t* pitids, int64_t npit, const int8_t* psel,
                         const double* target, const double* hl, int8_t* sel, double* wd, double* denom,
                         gospl_stream_t stream);
int gospl_sed_lake_init(gospl_mesh_t* mesh, const int32_t* pitids, const int8_t* sel, const int8_t* inlet,
                        const double* ca, const double* cb, const double* denom, const double* count,
                        const double* wd, double* delta, gospl_stream_t stream);
int gospl_sed_pit_volume(gospl_mesh_t* mesh, const double* delta, const int8_t* sel, const int32_t* pitids,
                         int64_t npit, double* vol, gospl_stream_t stream);
int gospl_sed_scale(gospl_mesh_t* mesh, const int32_t* pitids, const int8_t* sel, const double* scale,
                    double* delta, gospl_stream_t stream);
int gospl_sed_apply(gospl_mesh_t* mesh, const double* delta, double* h, double* cumed, gospl_stream_t stream);

/*
 * Marine deposition of river sediment: SEAMesh.seaChange
 * (gospl/sed/seaplex.py:738-861) and its helpers.  All arrays on the device;
 * xyz is (N, 3) row-major; rcv / wgt are (12, N).
 *  - gospl_coast_distance: _distanceCoasts / _globalCoastsTree (:51-135).
 *    The coastline points are VTK's contour of z at `level` on the triangle
 *    mesh (one point per edge with one end below the level and the other at
 *    or above it, x_lo + t (x_hi - x_lo), t = (level - z_lo) / (z_hi - z_lo));
 *    dist = Euclidean distance of each vertex with mask != 0 (mask may be
 *    NULL: all) to the nearest point (goSPL: a k-d tree query), 0 for the
 *    others and everywhere when there is no point.  *npoints (HOST, may be
 *    NULL): the point count.  Synchronises `stream`.
 *  - gospl_eps_fill: epsfill (functions.F90:2633-2697): cells below `level`
 *    keep h; a cell at or above it reached from a seed (a cell below the
 *    level next to one at or above it) becomes max(h, the next double above
 *    the least fill of its neighbours), as goSPL's priority flood leaves it;
 *    unreached cells keep h.  *passes (HOST, may be NULL).  Synchronises.
 *  - gospl_diffuse_fixed: the implicit linear diffusion (I + L(kd)) x = rhs of
 *    hillSLP._buildDiffMat / _solveSmooth (hillslope.py:131-150, 235-257,
 *    373-402) with per-vertex kd (sethillslopecoeff, identity rows at
 *    flat-mesh outlets), by the Chebyshev semi-iteration over the Jacobi
 *    sweep from x0 = rhs: x1 = sweep(x0), x_{k+1} = w_k (sweep(x_k) -
 *    x_{k-1}) + x_{k-1}, w_1 = 1 / (1 - rho^2 / 2), w_k = 1 / (1 - rho^2
 *    w_{k-1} / 4), rho = max over rows of (diag - 1) / diag; the count is
 *    the least k with 2 sigma^k <= 2^-49 (1 - rho), sigma = rho / (1 +
 *    sqrt(1 - rho^2)), at most maxit (goSPL: Richardson to rtol 1e-8).  x
 *    must not alias rhs or kd.  *sweeps (HOST, may be NULL).  Synchronises.
 *  - gospl_marine_receivers: mfdrcvrs (functions.F90:1673-1785) as _matOcean
 *    calls it (:198-219): rcv = the strictly lower neighbours on fillz in
 *    neighbour order then the vertex itself, wgt = 1 / their count in single
 *    precision as the Fortran's `1. / kk` (0 on the padding), sink = 1 where
 *    there is none; flat-mesh outlets: no receivers, not sinks.
 *  - gospl_marine_prep: shelf capacity cap = (clinoH - hl) area with
 *    clinoH = sealevel - 1e-3 - cdist clinslp, hl where hl >= sealevel or
 *    above clinoH, and the incoming volume vin = vsed dt on cells with
 *    lfill <= sealevel, 0 elsewhere or where negative (:761-776).
 *  - gospl_marine_route: _distOcean (:404-619): vdep = the deposited volume
 *    of the vin routed through (rcv, wgt) against the capacities cap; the
 *    passes, at most 5000, stop once the volume still moving is at most
 *    max(1, 1e-6 of its initial total).  *passes (HOST, may be NULL).
 *    Synchronises `stream` once per 16 passes.
 *  - gospl_marine_dep: _depMarineSystem (:240-402), the coupled solve of the
 *    routed volumes vol with Gmar > 0: goSPL's nested system with A00 =
 *    I - diag(fDepm), A01 = -diag(fDepm dt / area), A10 = diag(area / dt),
 *    A11 = (I - W)^T of (rcv, wgt) (fDepm = min(Gmar, 0.99), 0 where h >
 *    sealevel; right-hand sides (1 - fDepm) h and vol / dt + h area / dt),
 *    reduced to its elevation block and solved by the GMRES of gospl_spl_dep
 *    (identity preconditioner); voldep = max((hnew - h) area, 0).
 *  - gospl_marine_thickness: dh = vdep / area (0 where area is 0), values
 *    below floor_m set to 0 (seaplex.py:814-818).
 *  - gospl_marine_apply: cumed += t; h += t; eblocal += t / dt (:836-847).
 */
int gospl_coast_distance(gospl_mesh_t* mesh, const double* xyz, const double* z, double level,
                         const int8_t* mask, double* dist, int64_t* npoints, gospl_stream_t stream);
int gospl_eps_fill(gospl_mesh_t* mesh, const double* h, double level, double* fillz, int32_t* passes,
                   gospl_stream_t stream);
int gospl_diffuse_fixed(gospl_mesh_t* mesh, const double* kd, const double* rhs, double* x, int32_t maxit,
                        int32_t* sweeps, gospl_stream_t stream);
int gospl_marine_receivers(gospl_mesh_t* mesh, const double* fillz, int32_t* rcv, double* wgt, int8_t* sink,
                           gospl_stream_t stream);
int gospl_marine_prep(gospl_mesh_t* mesh, const double* hl, const double* cdist, const double* lfill,
                      const double* vsed, double sealevel, double clinslp, double dt, double* cap, double* vin,
                      gospl_stream_t stream);
int gospl_marine_route(gospl_mesh_t* mesh, const int32_t* rcv, const double* wgt, const int8_t* sink,
                       const double* cap, const double* vin, double* vdep, int32_t* passes, gospl_stream_t stream);
int gospl_marine_dep(gospl_mesh_t* mesh, const int32_t* rcv, const double* wgt, const double* h, const double* vol,
                     double Gmar, double sealevel, double dt, double* voldep, double rtol, int32_t maxit,
                     int32_t* iters, double* relres, gospl_stream_t stream);
int gospl_marine_thickness(gospl_mesh_t* mesh, const double* vdep, double floor_m, double* dh,
                           gospl_stream_t stream);
int gospl_marine_apply(gospl_mesh_t* mesh, const double* t, double dt, double* h, double* cumed, double* eblocal,
                       gospl_stream_t stream);

/*
 * Non-linear thickness diffusion of a deposit over one step: replaces
 * hillSLP._diffuseImplicit in its lagged-diffusivity form
 * (_diffuseImplicitPicard, gospl/sed/hillslope.py:925-1076; PETSc KSP
 * solves of the 13-column jacobiancoeff operator, functions.F90:1421-1467).
 * h = bed + dh; nsub backward-Euler sub-steps of dt / nsub, each with npic
 * Picard updates solving (I + dt_sub L(Cd(h))) h' = h_start with Cd =
 * min_diff + cd0 (1 - exp(-dexp max(h - bed, 0 below 0.1))) on mask cells,
 * 0 elsewhere; ndepo = max(h - bed, 0).  Each linear solve is the
 * Chebyshev semi-iteration over the Jacobi sweep, for the count at which
 * its error bound (from the first sweep's change and the contraction bound
 * rho = max (diag - 1) / diag) is at most 2^-52 of the solution's magnitude
 * (goSPL's Richardson stops at rtol 1e-8); *sweeps (HOST, may be NULL)
 * receives the sweeps used.  Synchronises `stream` once per linear solve.
 */
int gospl_nl_diffusion(gospl_mesh_t* mesh, const double* dh, const int8_t* mask, double cd0, double min_diff,
                       double dexp, double dt, int32_t nsub, int32_t npic, const double* bed, double* ndepo,
                       int64_t* sweeps, gospl_stream_t stream);

/*
 * Instrumentation (no counterpart in goSPL): when enabled, HIP events on the
 * launch stream bracket every DAG sweep kernel (flow accumulation and SPL).
 * Returns the accumulated milliseconds and call counts since the previous
 * call (HOST outputs, may be NULL), resets them, and sets the enable flag.
 */
int gospl_sweep_timing(gospl_mesh_t* mesh, int32_t enable, double* fa_ms, int64_t* fa_calls,
                       double* spl_ms, int64_t* spl_calls);

#ifdef __cplusplus
}
#endif

#endif /* GOSPL_HIP_H */
