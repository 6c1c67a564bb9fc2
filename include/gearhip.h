/*
 * gearhip -- C ABI of the MI355X (gfx950) batched CROWN bound / branch /
 * prune path.
 *
 * Conventions (all entry points):
 *   - every pointer argument is a device pointer unless stated otherwise;
 *     buffers are caller-owned (the library allocates device memory only
 *     for program plans, in gh_dag_plan_create, and frees it in
 *     gh_dag_plan_destroy);
 *   - arrays are dense row-major, FP64 unless the name says f32;
 *   - work is enqueued on `stream` (0 = the null stream) and is
 *     asynchronous; the return value reports argument and launch errors
 *     only;
 *   - return codes: GH_OK, GH_EINVAL (bad size or null pointer),
 *     GH_EUNSUPPORTED (a size this build has no kernel for), GH_ELAUNCH.
 *
 * Reference citations are paths relative to the reference's solver/ tree.
 */
#ifndef GEARHIP_H
#define GEARHIP_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#include "gearhip_program.h"

#ifdef __cplusplus
extern "C" {
#endif

#define GH_OK 0
#define GH_EINVAL 1
#define GH_EUNSUPPORTED 2
#define GH_ELAUNCH 3

/* Library version, major * 10000 + minor * 100 + patch (host only). */
int gh_version(void);

/*
 * Problem setup for the QCQP bound: packs Q (S, n, n) once into the MFMA
 * operand order the bound kernels stream (the immutable per-program plan
 * of this ABI).  `packed` holds gh_qcqp_pack_bytes(n, S) bytes (0 when the
 * size is unsupported).  Supported: 1 <= n <= 64.
 */
size_t gh_qcqp_pack_bytes(int n, int S);
int gh_qcqp_pack_f64(int n, int S, const double *Q, void *packed,
                     hipStream_t stream);

/*
 * CROWN lower bounds of a QCQP over B boxes, FP64.
 *
 * Replaces, for programs of the form f_s(x) = x^T Q_s x + c_s^T x + d_s
 * (s = 0 objective, s >= 1 constraints), the call
 *   BoundedModule.compute_bounds(x=(BoundedTensor(x, PerturbationLpNorm(
 *       x_L=xL, x_U=xU)),), C=eye(S), method='CROWN', bound_upper=False,
 *       return_A=True, needed_A_dict={out: {in}})
 * (auto_LiRPA/bound_general.py:1275-1593, backward_bound.py:880-1206), as
 * bound by input_split/bounding.py:316-326.
 *
 *   packed: gh_qcqp_pack_f64 of Q (S, n, n); Cm (S, n), d (S);
 *   xL, xU (B, n): boxes;
 *   lb (B, S): lower bounds (out);
 *   lA (B, R, n), lbias (B, R) with R = plane_rows: the linear lower planes
 *   lA x + lbias before concretization of the leading R rows (out, either
 *   may be NULL; 1 <= R <= S when one is given).  R = S is the reference's
 *   lA; R = 1 (the objective row) is what gh_child_keys_f64 reads.
 * Supported: 1 <= n <= 64.  Results are bit-identical to the FP64 oracle.
 */
int gh_qcqp_crown_f64(int B, int n, int S, const void *packed,
                      const double *Cm, const double *d, const double *xL,
                      const double *xU, double *lb, double *lA, double *lbias,
                      int plane_rows, hipStream_t stream);
/* gh_qcqp_crown_f64 on the boxes rows[b] (b < B, int64) of xL, xU: a
 * caller's picked domains bound where its storage holds them, without
 * gathering them first (pick_out_batch, branching_domains.py:831-962,
 * then the bound).  lb, lA, lbias are indexed by b as above. */
int gh_qcqp_crown_rows_f64(int B, int n, int S, const void *packed, const double *Cm,
                           const double *d, const double *xL, const double *xU,
                           const int64_t *rows, double *lb, double *lA, double *lbias,
                           int plane_rows, hipStream_t stream);

/*
 * FP32-sound form of gh_qcqp_crown_f64: the same relaxation evaluated in
 * FP32 on MFMA with a rigorous outward error term, so each bound is a valid
 * lower bound and not above the FP64 bound of gh_qcqp_crown_f64.  McCormick
 * choices within the FP32/FP64 difference of a tie are settled with the
 * FP64 path's exact arithmetic for that element; flag[b] = 1 records that
 * this happened for box b (diagnostic only).  This form is new surface (the
 * reference has FP64 only, CV/gear/config.py:219-221).
 *
 *   packed: gh_qcqp_pack_f32 of Q (gh_qcqp_pack_f32_bytes(n, S) bytes);
 *   Q (S, n, n) FP64 (for the tie cases), Cm, d, xL, xU as
 *   gh_qcqp_crown_f64 (FP64; boxes are rounded outward to FP32);
 *   lb (B, S) FP64 (FP32 values); flag (B) int32;
 *   lA (B, R, n), lbias (B, R) of the leading R = plane_rows rows (either
 *   may be NULL): the lower planes, lA x + lbias <= f_s(x) on the box with
 *   lbias lowered by the planes' rigorous error (for gh_lp_dual_bound_f64
 *   with R = S, for gh_child_keys_f64 with any R).
 * Supported: 1 <= n <= 64.
 */
size_t gh_qcqp_pack_f32_bytes(int n, int S);
int gh_qcqp_pack_f32(int n, int S, const double *Q, void *packed,
                     hipStream_t stream);
int gh_qcqp_crown_f32s(int B, int n, int S, const void *packed,
                       const double *Q, const double *Cm, const double *d,
                       const double *xL, const double *xU, double *lb,
                       int32_t *flag, double *lA, double *lbias, int plane_rows,
                       hipStream_t stream);
/* gh_qcqp_crown_f32s on the boxes rows[b] of xL, xU (as
 * gh_qcqp_crown_rows_f64). */
int gh_qcqp_crown_rows_f32s(int B, int n, int S, const void *packed, const double *Q,
                            const double *Cm, const double *d, const double *xL,
                            const double *xU, const int64_t *rows, double *lb, int32_t *flag,
                            double *lA, double *lbias, int plane_rows, hipStream_t stream);

/*
 * FP32-sound CROWN lower bound of a QCQP of any size.  Same contract and
 * the same guarantee as gh_qcqp_crown_f32s (a valid lower bound, not above
 * the FP64 bound; near-tie McCormick choices settled with the FP64 path's
 * arithmetic, flag[b] = 1 when that happened), arranged for large n: per
 * row and tile of boxes, one fused kernel forms the products Q_s x0,
 * |Q_s| ep and |Q_s| |x0| on FP32 MFMA and relaxes each element in FP64 in
 * its epilogue, a second forms Q_s^T beta + Ainit and the concretization
 * sums, with error terms that hold for any summation order.
 *   plan: gh_qcqp_large_plan_bytes(n, S) bytes, filled once by
 *   gh_qcqp_large_prepare from Q (S, n, n) FP64;
 *   Q64: the same Q (FP64, for the near-tie cases); Cm, d, xL, xU, lb, flag
 *   as gh_qcqp_crown_f32s (lb holds FP64 values);
 *   ws: workspace; gh_qcqp_large_workspace_bytes(n, S, k) bytes hold k
 *   boxes, and B boxes are bounded in chunks of as many as ws holds.
 * New surface (the reference has FP64 only, CV/gear/config.py:219-221).
 */
size_t gh_qcqp_large_plan_bytes(int n, int S);
int gh_qcqp_large_prepare(int n, int S, const double *Q, void *plan, hipStream_t stream);
size_t gh_qcqp_large_workspace_bytes(int n, int S, int B);
int gh_qcqp_crown_f32s_large(int B, int n, int S, const void *plan, const double *Q64,
                             const double *Cm, const double *d, const double *xL,
                             const double *xU, double *lb, int32_t *flag, void *ws,
                             size_t ws_bytes, hipStream_t stream);

/*
 * ALM-PGD primal search, fused: N restarts of the reference's
 * ALMPGDSolver.solve_with_restarts (complete_verifier/nlp_solver/
 * alm.py:1268-2065; inner Adam _solve_subproblem_adam :551-788,
 * AdamClipping_NLP.step :134-160, ALMConfig :163-268) on a QCQP, every
 * outer and inner iteration in one launch.  Each restart runs as the
 * reference runs one restart of one domain (batch 1, 1 restart): its inner
 * loop ends at its own stationarity (gradient norm <= eta_k, :764-776) or
 * L plateau (:678-699); its outer loop at its own objective patience or
 * first-feasibility patience (:1800-1849); the multipliers, rho and the
 * tolerances eps_k / eta_k follow :1851-1914 (update_multipliers_always
 * selects the root-mode rule, else the gated BaB-mode rule).  The reference
 * couples restarts only through those early exits, whose .all() over the
 * batch is not reproduced.
 *   plan: gh_alm_plan_bytes(n, S) bytes, filled once by gh_alm_prepare from
 *   Q (S, n, n) (the symmetric parts (Q + Q^T) / 2);
 *   Cm (S, n), d (S), rhs (S): objective / constraint data (rhs[0] unused);
 *   lo, hi (N, n): each restart's box; x (N, n): starting points in, last
 *   iterates out; best_x (N, n), best_obj (N): each restart's best feasible
 *   point (violation <= feas_tol) and objective (+inf if none);
 *   cfg: HOST pointer;
 *   trace (may be NULL): N x outer_max_iter records of
 *   gh_alm_trace_width(S) doubles: rho, eps_k, eta_k after the update,
 *   the violation, inner steps taken, a flag (0 updated, 1 stopped at this
 *   iteration before the update, 2 not run), the last inner gradient norm,
 *   then mu (S - 1) after the update and g = f - rhs (S - 1) before it.
 * Supported: n <= 1024, S <= 1024.
 */
typedef struct {
    double rho0, kappa, max_rho, inner_lr, feas_tol, beta1, beta2, adam_eps;
    double alpha, beta, eps_final, eta_final;
    double inner_lr_outer_decay, inner_lr_min_factor, obj_improvement_tol, inner_plateau_rtol;
    int32_t outer_max_iter, inner_steps, scale_eps_with_bias, update_multipliers_always;
    int32_t obj_patience, first_feas_patience, inner_plateau_check, inner_plateau_patience;
} gh_alm_config;
size_t gh_alm_plan_bytes(int n, int S);
int gh_alm_trace_width(int S);
int gh_alm_prepare(int n, int S, const double *Q, void *plan, hipStream_t stream);
int gh_alm_pgd_f64(int N, int n, int S, const void *plan, const double *Cm, const double *d,
                   const double *rhs, const double *lo, const double *hi, double *x,
                   double *best_x, double *best_obj, const gh_alm_config *cfg, double *trace,
                   hipStream_t stream);

/*
 * FP64 LP dual certificate per box: a rigorous lower bound of
 *   min c^T x + c0  s.t.  G x <= h,  xL <= x <= xU
 * with c, c0 the objective row's CROWN plane (lA[b, 0, :], lbias[b, 0]) and
 * G, h the constraint rows' planes against their thresholds
 * (G = lA[b, 1:, :], h = thr[1:] - lbias[b, 1:]), as bounding.py forms it
 * (complete_verifier/input_split/bounding.py:489-520) for lp_bound
 * (complete_verifier/nlp_solver/dual_ascent.py:459-561).  Multipliers
 * y >= 0 come from `iters` iterations of diagonally preconditioned PDHG
 * (dual_ascent.py:312-432); the bound is the weak-duality value at the best
 * checkpoint, evaluated with the error model of rigorous_dual.py:1-54 from
 * FP64 operations only, so it is a certificate whatever the iteration
 * count.  Rows with a non-finite threshold or plane are inactive (y = 0).
 * With a finite objective threshold thr[0] (the incumbent), a box stops
 * iterating at the first checkpoint whose certified value reaches thr[0]:
 * the prune drops such a box whatever its exact bound; every other box runs
 * all `iters` iterations.
 *   lA (B, S, n), lbias (B, S) from gh_qcqp_crown_f64; thr (S); xL, xU (B, n);
 *   lb (B): the certified bound (-inf if none); y (B, S - 1): its
 *   multipliers (may be NULL).
 * Supported: n <= 512, S - 1 <= 512.
 */
int gh_lp_dual_bound_f64(int B, int n, int S, const double *lA, const double *lbias,
                         const double *thr, const double *xL, const double *xU, int iters,
                         double *lb, double *y, hipStream_t stream);

/*
 * Widest-variable branching: idx[b] = argmax_i (xU[b,i] - xL[b,i]), the
 * lowest index on ties.  Replaces input_split_branching(...,
 * branching_method='naive') (complete_verifier/input_split/
 * branching_heuristics.py:59-86).  idx is int32 (B).
 */
int gh_branch_widest_f64(int B, int n, const double *xL, const double *xU,
                         int32_t *idx, hipStream_t stream);

/*
 * Bisection of each box on dimension idx[b] at (xL + xU) / 2: child 0 is
 * the upper half, child 1 the lower half; children are written
 * [all child 0; all child 1] into oL, oU (2B, n).  Replaces
 * input_split_parallel(..., split_depth=1, split_partitions=2)
 * (complete_verifier/input_split/split.py:26-122).
 */
int gh_split_f64(int B, int n, const double *xL, const double *xU,
                 const int32_t *idx, double *oL, double *oU,
                 hipStream_t stream);

/*
 * Constrained concretization: root[b][h] = a lower bound of
 * min obj[b][h] . x over the box [xL[b], xU[b]] intersected with the cuts
 * A[b][k] . x + bc[b][k] <= 0, by the sequential one-row dual sweep.
 * Replaces constraints_solving with a batch-specific objective, after
 * sort_out_constr_batches(no_return_inf=True) and
 * drop_redundant_constraints(max_rows) (auto_LiRPA/concretize_func.py:
 * 269-850; its Triton sweep concretize_triton.py:159-217), as
 * PerturbationLpNorm.concretize_matrix calls it (auto_LiRPA/
 * perturbations.py:241-263) for clip_type 'complete'.
 *   obj (B, H, X); A (B, K, X), bc (B, K) (may be NULL when K = 0);
 *   xL, xU (B, X); max_rows: keep at most this many binding cuts per box,
 *   the largest worst-case values (<= 0: no cap);
 *   root (B, H) out: never below the box bound of obj over [xL, xU] (the
 *   larger of the two is returned, as the reference does);
 *   flags (B) out: bit 0 = some cut misses the box (the domain is
 *   infeasible: its bounds may be set to +inf), bit 1 = every cut holds on
 *   the whole box (root is the box bound); nkept (B) out (may be NULL):
 *   cuts the sweep visited.
 * Supported: 1 <= X <= 64, 0 <= K <= 256, min(K, max_rows) <= 64.
 * Bit-identical to the oracle's restatement (its sums run in a stated
 * order); the error term counts the box's own kept cuts.
 */
int gh_constrained_concretize_f64(int B, int H, int K, int X, const double *obj,
                                  const double *A, const double *bc, const double *xL,
                                  const double *xU, int max_rows, double *root,
                                  int32_t *flags, int32_t *nkept, hipStream_t stream);

/*
 * The bound form of gh_constrained_concretize_f64, for the planes of a CROWN
 * pass: lb[b][h] = max(lb[b][h], lbias[b][h] widened and added to the
 * constrained root of lA[b][h] as the box concretization adds it
 * (auto_LiRPA/concretize_bounds.py:280-304)), and +inf on every row of a
 * box that a cut misses (input_split/bounding.py:345-350).  lb (B, H) holds
 * the box bounds on entry (e.g. gh_qcqp_crown_f64's lb) and the result on
 * return; lA (B, H, X), lbias (B, H); box b's cuts are A + b lda
 * ((K, X) row-major) and bc + b ldb (K) (lda >= K X, ldb >= K: the cut rows
 * may sit inside a wider per-domain record); the rest as above.
 */
int gh_constrained_bound_f64(int B, int H, int K, int X, const double *lA,
                             const double *lbias, const double *A, int64_t lda,
                             const double *bc, int64_t ldb, const double *xL,
                             const double *xU, int max_rows, double *lb, int32_t *flags,
                             int32_t *nkept, hipStream_t stream);

/*
 * Domain clipping: shrink each box [xL[b], xU[b]] along every coordinate to
 * the part where some point may still be unverified by the planes
 * lA[p][s] x + lbias[p][s] against thr[s] (p = b % P: the 2K children of a
 * split share their parent's planes), rounding every candidate bound
 * outward so no exactly feasible point is removed; a box that one plane
 * already certifies is emptied (oL[b][0] = nextafter(xU[b][0], +inf)).
 * Replaces clip_domains(x_L, x_U, thresholds, lA, lbias, num_iters=1)
 * (complete_verifier/input_split/clip.py:147-503) of clip_type 'complete'.
 *   lA (P, S, n), lbias (P, S); thr (S) (+inf: an inactive row);
 *   xL, xU (B, n); oL, oU (B, n) out (may alias xL, xU);
 *   state (B) out: 0 left as is (non-finite plane data), 1 clipped,
 *   2 certified and emptied, 3 clipped to empty.
 * Supported: 1 <= n <= 64.  Bit-identical to the oracle's restatement.
 */
int gh_clip_domains_f64(int B, int P, int n, int S, const double *lA, const double *lbias,
                        const double *thr, const double *xL, const double *xU, double *oL,
                        double *oU, int32_t *state, hipStream_t stream);

/*
 * Child keys of a split (solving mode): for the 2K children of K parents
 * (layout of gh_split_f64: child k belongs to parent k mod K),
 * keys[k][0] = max(plb[p][0], the minimum of the parent's objective plane
 * lA[p][0] x + lbias[p][0] over the child box, lowered by
 * (n + 4) u (sum |a| (|c| + r) + |b|) and one step toward -inf), and
 * keys[k][s] = plb[p][s] for s >= 1.  Replaces _raise_child_keys /
 * _plane_minimum_down (complete_verifier/input_split/
 * batch_branch_and_bound.py:189-225) of the reordered loop.
 *   plb (K, S): the parents' bounds; lA (K, R, n), lbias (K, R): their
 *   planes of the leading R >= 1 rows (row 0 is read); cL, cU (2K, n):
 *   children; keys (2K, S) out.  Both sums run in the order the oracle's
 *   plane_minimum_down states (64 lane-strided partials, then an xor tree).
 */
int gh_child_keys_f64(int K, int n, int S, int R, const double *plb, const double *lA,
                      const double *lbias, const double *cL, const double *cU, double *keys,
                      hipStream_t stream);

/*
 * gh_branch_widest_f64, gh_split_f64 and gh_child_keys_f64 in one pass over
 * K parents (the reordered loop's decision, split and child-key steps,
 * batch_branch_and_bound.py:470-534): cL, cU (2K, n) and keys (2K, S) out,
 * equal to the three calls in sequence.  lA / lbias may be NULL (no
 * planes: keys are the parents' bounds).
 */
int gh_branch_split_keys_f64(int K, int n, int S, int R, const double *pL, const double *pU,
                             const double *plb, const double *lA, const double *lbias,
                             double *cL, double *cU, double *keys, hipStream_t stream);

/*
 * Smart branching of B boxes (input_split_branching with method 'sb',
 * complete_verifier/input_split/branching_heuristics.py:25-125, 176-200,
 * with the GEAR presets' options sb_sum, touch_zero_score 0): idx[b] =
 * argmax_i sum_s max(|lA[b,s,i]|, coeff_thresh) (xU - xL)[i] / 2 (rows
 * summed in order, lowest index on ties); an objective threshold of +inf
 * counts as the box's own bound, and any non-finite plane coefficient,
 * bound or threshold in the batch sends every box to the widest input;
 * with root widths root_w (n, or NULL) and min_width_ratio > 0 a choice
 * narrower, relative to the root, than min_width_ratio times the widest
 * relative width becomes that widest input (the exhaustive guard).
 *   xL, xU (B, n), or rows (B) int64 into them (may be NULL); lb (B, S);
 *   thr (S); lA (B, S, n); idx (B) int32 out; flag: one device int32 of
 *   scratch.
 */
int gh_branch_sb_f64(int B, int n, int S, const double *xL, const double *xU,
                     const int64_t *rows, const double *lb, const double *thr, const double *lA,
                     double coeff_thresh, const double *root_w, double min_width_ratio,
                     int32_t *idx, int32_t *flag, hipStream_t stream);

/* Bytes of device workspace gh_filter_append_f64 needs for B boxes. */
size_t gh_filter_workspace_bytes(int B);

/*
 * Prune verified boxes and append the survivors to a frontier.  A box is
 * dropped when some row has finite lb and threshold with lb > thr, when
 * the objective row has lb >= thr, when the objective lb is +inf with a
 * finite threshold, or when its box is invalid (xL > xU, all finite).
 * Replaces UnsortedInputDomainList.filter_verified_domains /
 * get_remaining_index (complete_verifier/input_split/
 * branching_domains.py:238-558) followed by add (:560-830).
 *
 *   lb (B, S), thr (S): thresholds shared by all boxes (row 0 the
 *   incumbent, +inf if none; rows >= 1 rhs + constr_viol_tol);
 *   cL, cU (B, n): boxes;
 *   FL, FU (cap, n), Flb (cap, S): frontier; survivors are written in
 *   ascending box order starting at row `base` (host value);
 *   total (device int64): number of survivors (out);
 *   disc_min (device double, in/out): lowered to the smallest finite
 *   objective threshold of a dropped box
 *   (branching_domains.py:489-496);
 *   ws: gh_filter_workspace_bytes(B) bytes.
 * The caller guarantees base + B <= cap.
 */
int gh_filter_append_f64(int B, int S, int n, const double *lb,
                         const double *thr, const double *cL,
                         const double *cU, double *FL, double *FU,
                         double *Flb, int64_t base, int64_t cap,
                         int64_t *total, double *disc_min, void *ws,
                         hipStream_t stream);

/* Bytes of device workspace gh_split_append_f64 needs for K parents. */
size_t gh_split_append_workspace_bytes(int K);

/*
 * The reordered step's decision, split, child keys, prune and append in
 * one call (batch_branch_and_bound.py:470-534 followed by
 * UnsortedInputDomainList.add, branching_domains.py:238-830): equal to
 * gh_branch_split_keys_f64 on K parents, then gh_filter_append_f64 on the
 * 2K children (base, cap, total, disc_min as there), then
 * gh_fill_order_i64(2K, total, base, dead, order_out), without writing
 * the children anywhere but the frontier: each survivor's box is rebuilt
 * from its parent.  Parent p's box is row prows[p] of pL, pU (row p when
 * prows is NULL), so the parents can be read in place in the frontier as
 * long as the append does not overwrite them; plb, lA, lbias are indexed
 * by p.  ws: gh_split_append_workspace_bytes(K) bytes.  The caller
 * guarantees base + 2K <= cap.  Frontiers whose domains carry a
 * payload use the separate calls.
 */
int gh_split_append_f64(int K, int n, int S, int R, const double *pL, const double *pU,
                        const int64_t *prows, const double *plb, const double *lA,
                        const double *lbias,
                        const double *thr, double *FL, double *FU, double *Flb, int64_t base,
                        int64_t cap, int64_t *order_out, int64_t dead, int64_t *total,
                        double *disc_min, void *ws, hipStream_t stream);

/*
 * Companion of gh_filter_append_f64 for domains that carry a payload (the
 * cut rows of clip_type 'complete', stored with each domain by
 * UnsortedInputDomainList.add, complete_verifier/input_split/
 * branching_domains.py:560-830): called right after gh_filter_append_f64
 * with the same B, ws and base, it copies the payload of each survivor k,
 * row k % P of pay (P, W) (the children of a split share their parent's),
 * to row base + rank of Fpay (cap, W).
 */
int gh_filter_payload_f64(int B, int W, int P, const double *pay, const void *ws,
                          double *Fpay, int64_t base, int64_t cap, hipStream_t stream);

/*
 * out (device double, in/out) = min(out, min_b x[b * stride]) over
 * b < N.  Used for the global lower bound over the frontier's objective
 * row (batch_branch_and_bound.py:620-634).
 */
int gh_min_strided_f64(int64_t N, int64_t stride, const double *x,
                       double *out, hipStream_t stream);

/*
 * Frontier with a logical order over stable physical rows (rows never move
 * when the frontier is re-sorted; only their int64 slot ids do):
 * gh_frontier_keys_f64: keys[i] = max_s (lb[order[i]][s] - thr[s]) (NaN if
 * a term is NaN), the sort key of UnsortedInputDomainList.sort with
 * sort_heuristic 'lb_rhs' (branching_domains.py:1171-1241);
 * gh_min_rows_f64: *out = min(*out, min_i x[idx[i] * stride]), the global
 * lower bound over the logical rows (batch_branch_and_bound.py:620-634).
 */
int gh_frontier_keys_f64(int64_t N, int S, const int64_t *order, const double *lb,
                         const double *thr, double *keys, hipStream_t stream);
int gh_min_rows_f64(int64_t N, int64_t stride, const int64_t *idx, const double *x, double *out,
                    hipStream_t stream);
/* Stable merge of two runs sorted descending by key (every NaN first, as
 * torch's descending sort; -0 equals +0), A before B on equal keys: the
 * frontier's re-sort when only the appended children are unsorted, equal
 * to the stable sort of the whole (UnsortedInputDomainList.sort,
 * branching_domains.py:1171-1241).  kO, vO (nA + nB) out. */
int gh_merge_desc_f64(int64_t nA, const double *kA, const int64_t *vA, int64_t nB,
                      const double *kB, const int64_t *vB, double *kO, int64_t *vO,
                      hipStream_t stream);
/* Stable sort of the N rows named by order[0..N) by their sort key
 * max_s (lb[row][s] - thr[s]) descending (every NaN first, as torch's
 * descending sort; -0 equals +0; equal keys keep their input order): the
 * keys and slot ids in sorted order into kOut, vOut (N each); kTmp, vTmp
 * (N each) are scratch.  Equal to gh_frontier_keys_f64 followed by torch's
 * stable descending argsort and the gathers of keys and slot ids
 * (UnsortedInputDomainList.sort, branching_domains.py:1171-1241); a key
 * written out is the row's key up to the NaN payload and the sign of zero. */
int gh_frontier_sort_f64(int64_t N, int S, const int64_t *order, const double *lb,
                         const double *thr, double *kOut, int64_t *vOut, double *kTmp,
                         int64_t *vTmp, hipStream_t stream);
/* Slot ids of B children just appended by gh_filter_append_f64 at
 * physical row hw: out[t] = hw + t for the *total survivors, `dead` (the
 * frontier's dead record) for the rest, so the host need not read the
 * survivor count before the next step. */
int gh_fill_order_i64(int B, const int64_t *total, int64_t hw, int64_t dead, int64_t *out,
                      hipStream_t stream);
/* Rows idx[i] (i < N) of up to three row-major arrays s_k of widths w_k
 * (a width 0 skips the array) into contiguous d_k (N, w_k): the picked
 * records (x_L, x_U, lb) in one launch (pick_out_batch,
 * branching_domains.py:831-962). */
int gh_gather_rows_f64(int64_t N, const int64_t *idx, int w0, const double *s0, double *d0,
                       int w1, const double *s1, double *d1, int w2, const double *s2,
                       double *d2, hipStream_t stream);

/*
 * gh_split_append_f64 with the split input of each parent given (idx (K)
 * int32, e.g. from gh_branch_sb_f64) instead of its widest input.
 */
int gh_split_append_idx_f64(int K, int n, int S, int R, const double *pL, const double *pU,
                            const int64_t *prows, const double *plb, const double *lA,
                            const double *lbias, const double *thr, const int32_t *idx,
                            double *FL, double *FU, double *Flb, int64_t base, int64_t cap,
                            int64_t *order_out, int64_t dead, int64_t *total, double *disc_min,
                            void *ws, hipStream_t stream);

/*
 * gh_split_append_f64 without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * scratch lives in a fixed device array of the library (3.7 MB per GPU,
 * held until the process ends), so K may be at most
 * gh_split_append_noalloc_max_parents() (131,072 parents, 262,144
 * children); larger K returns GH_EUNSUPPORTED and launches nothing.  The
 * call allocates nothing.  Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other streams or threads) race and must be
 * serialized by the caller.  Results equal gh_split_append_f64's bit for
 * bit.
 */
int gh_split_append_f64_noalloc(int K, int n, int S, int R, const double *pL, const double *pU,
                                const int64_t *prows, const double *plb, const double *lA,
                                const double *lbias, const double *thr, double *FL, double *FU,
                                double *Flb, int64_t base, int64_t cap, int64_t *order_out,
                                int64_t dead, int64_t *total, double *disc_min,
                                hipStream_t stream);
int gh_split_append_noalloc_max_parents(void);

/* ------------------------------------------------ expression programs */

/*
 * Program plan: the immutable, device-resident form of one or more
 * expression programs (gearhip_program.h), built once per program set.
 * Replaces the per-model graph that BoundedModule.__init__ builds
 * (auto_LiRPA/bound_general.py:50-160) for programs in the node vocabulary
 * of gearhip_program.h.  `progs` and the arrays they point to are HOST
 * memory; the plan uploads what the kernels read to the current device.
 * Returns GH_EUNSUPPORTED for programs past the plan limits (a row with
 * more than 8192 nodes).
 */
typedef struct gh_dag_plan gh_dag_plan;
int gh_dag_plan_create(int nprog, const gh_prog *progs, gh_dag_plan **plan);
int gh_dag_plan_destroy(gh_dag_plan *plan);
/* Rows (jobs) in the plan and per-lane state slots of the largest row. */
int gh_dag_plan_info(const gh_dag_plan *plan, int64_t *jobs, int64_t *max_slots,
                     int64_t *tape_bytes);
/* Per program (host arrays of nprog, each may be NULL), per box bound:
 * bytes of per-box state the bound reads and writes (A propagation and
 * relaxation data, as the reference's backward pass materializes them),
 * bytes of box and bound I/O, and backward-pass node visits. */
int gh_dag_plan_stats(const gh_dag_plan *plan, int64_t *state_bytes, int64_t *io_bytes,
                      int64_t *steps);
/* Per program (host arrays of nprog, each may be NULL), per box bound: the
 * algorithmic FP64 operations of every row's bound (add, sub, mul 1, fma 2;
 * a relaxation ~10 per node, the A scaling 4 and the rounding enclosure 2
 * per pushed A entry and side, concretization 4 per input and side,
 * interval propagation 2 per node; intermediate-bound passes included) and
 * the transcendental evaluations, counted from the tape. */
int gh_dag_plan_work(const gh_dag_plan *plan, int64_t *flops, int64_t *transcendentals);

/* One program's batch of boxes (device pointers).  lA, lbias and undef may
 * be NULL: each row's lower plane lA x + lbias over the program's inputs
 * (before concretization, as BoundedModule.compute_bounds returns it in
 * A_dict with return_A, auto_LiRPA/backward_bound.py:1093-1111) and the
 * rows whose relaxation was undefined (their plane is A = 0, bias -inf,
 * :445-463; their lb is the interval fallback). */
typedef struct {
    const double *xL, *xU;   /* (cap, n) */
    double *lb;              /* (cap, R) out */
    const int64_t *count;    /* device scalar: boxes to bound, <= cap */
    double *lA;              /* (cap, R, n) out, or NULL */
    double *lbias;           /* (cap, R) out, or NULL */
    uint8_t *undef;          /* (cap, R) out, or NULL */
} gh_dag_batch;

/* Workspace bytes for gh_dag_crown_f64 with `waves` resident wavefronts. */
size_t gh_dag_workspace_bytes(const gh_dag_plan *plan, int waves);

/*
 * CROWN lower bounds of every row of every program over its batch, FP64,
 * in one persistent launch: `waves` wavefronts (64 boxes each) pull
 * (row, 64-box chunk) items from a device queue, largest rows first, until
 * the queue drains.  Same contract as gh_qcqp_crown_f64:
 *   compute_bounds(..., method='CROWN', bound_upper=False, return_A)
 * (auto_LiRPA/bound_general.py:1275-1593): intermediate bounds of every
 * input of a relaxed node by CROWN intersected with intervals
 * (bound_general.py:1032-1222), the operators' relaxations
 * (operators/bivariate.py, activations.py, s_shaped.py, convex_concave.py,
 * trigonometric.py), A67 rounding enclosure (backward_bound.py:550-823),
 * box concretization (concretize_bounds.py:267-321) and the interval
 * fallback of rows with no relaxation (backward_bound.py:826-877).
 *   batches: device array of nprog gh_dag_batch, in plan order;
 *   ws: gh_dag_workspace_bytes(plan, waves) bytes;
 *   lds_slots (0 .. 64): per-lane slots of each row's A area (the passes'
 *   A values, the most accessed state) kept in LDS, lds_slots x 512 bytes
 *   per wavefront; the rest of the state stays in ws.
 * Counts are read on the device, so a captured launch follows the batch
 * sizes at replay.  Calls sharing `ws` must be ordered on one stream.
 */
int gh_dag_crown_f64(const gh_dag_plan *plan, const gh_dag_batch *batches, void *ws,
                     size_t ws_bytes, int waves, int lds_slots, hipStream_t stream);
/* The largest A area of a row in the plan (slots): lds_slots past it
 * change nothing. */
int gh_dag_plan_aslots(const gh_dag_plan *plan, int64_t *max_aslots);

#ifdef __cplusplus
}
#endif

#endif /* GEARHIP_H */
