/* qphix_mi355x.h -- entry points of libqphixmilc_hip.so that have no
 * counterpart in the reference's MILC interface.
 *
 * They are additions for callers that want the MI355X-specific solvers
 * (mixed precision with reliable updates, batched right-hand sides), the
 * fermion force, or control over synchronisation and diagnostics.  Plain C,
 * pointers and sizes only.
 */
#ifndef _QPHIX_MI355X_H
#define _QPHIX_MI355X_H

#include <qphix_int.h>
#include <qphix_f3.h>
#include <qphix_d3.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- run-time control and diagnostics ------------------------------------ */

/* on != 0 (default): a contract violation or HIP failure prints and exits,
 * as the reference's MYASSERT does.  on == 0: the failure is recorded and the
 * call returns (QPHIX_FAIL where the signature allows). */
void QPHIX_MI355X_set_abort_on_error(int on);
/* Number of failures recorded since the last clear; copies the last message. */
int QPHIX_MI355X_last_error(char *buf, int len);
void QPHIX_MI355X_clear_error(void);

/* on != 0 (default): every call has finished on the GPU when it returns.
 * on == 0: Dslash calls return once queued on the library's stream; solvers
 * and extract_* still synchronise before returning host data. */
void QPHIX_MI355X_set_sync(int on);
void QPHIX_MI355X_sync(void);
/* The hipStream_t every kernel of the library runs on. */
void *QPHIX_MI355X_stream(void);
int QPHIX_MI355X_device_info(int *n_cu, double *hbm_gib);

/* Long-link storage: 0 = automatic (12 reals when every long link equals
 * +-conj(row0 x row1) in its third row, else 18), 18 = always 18.
 * Applies to links created afterwards. */
void QPHIX_MI355X_set_long_recon(int mode);
/* Gauge-field copy read by the gauge force: 0 = automatic (rows 0-1 only,
 * row 2 rebuilt as conj(row0 x row1), when every link is SU(3) to rounding;
 * else all 3 rows), 18 = always 3 rows.  Applies to fields created afterwards;
 * extract_G always returns the links as given. */
void QPHIX_MI355X_set_gauge_recon(int mode);
/* Gauge force kernel: 1 (default) = built from stored plaquette staples
 * (57 SU(3) products per link; used when G keeps its two-row copy, i.e. its
 * links are SU(3) to rounding, and the staple buffer fits in device memory),
 * 0 = every loop walked link by link (140 products per link; any field). */
void QPHIX_MI355X_set_gauge_force_staples(int on);
/* Rows the gauge force reads per link of G: 12 or 18 reals. */
int QPHIX_D3_gauge_links_recon(QPHIX_D3_GaugeField *G);
int QPHIX_F3_gauge_links_recon(QPHIX_F3_GaugeField *G);
/* on != 0 (default): the solvers (CG, multi-shift, mixed and batched) capture
 * their batch of iterations once per solve as a hipGraph and replay it;
 * on == 0: plain launches (the same kernels in the same order, hence
 * identical results). */
void QPHIX_MI355X_set_graphs(int on);
/* Test hook for the out-of-memory paths: after `after` more successful device
 * allocations, every allocation by the library fails as if device memory
 * were exhausted (the error is recorded as for a real failure); a negative
 * value turns this off (the default). */
void QPHIX_MI355X_debug_fail_alloc(int after);
/* Single-reduction CG steps, the FEWSUMS variant of the reference's CG
 * (ks_d_congrad_fn.cpp:450-541, a compile-time option there): the p.Ap
 * reduction also sums Re<Ap|r>, |Ap|^2 and |r|^2, the new |r|^2 follows as
 * |r|^2 + 2a Re<Ap|r> + a^2 |Ap|^2, and x, r and p are updated in one pass.
 * mode bit 0: QPHIX_*3_asqtad_invert (when relresid is off); bit 1: the sloppy
 * steps of QPHIX_D3_asqtad_invert_mixed.  Default 2 (QPHIX_FEWSUMS in the
 * environment sets the initial value). */
void QPHIX_MI355X_set_fewsums(int mode);
/* on != 0 (default; QPHIX_MRHS_LDS=0 in the environment sets the initial
 * value to 0): multi-RHS Dslash launches with one, two or four sources per
 * work-item (the autotuner's choice) stage each wavefront's links through LDS
 * (loaded once per site instead of once per work-item); on == 0: every
 * work-item loads its own links.  Both give identical results. */
void QPHIX_MI355X_set_mrhs_lds(int on);
/* Both multi-shift solvers (QPHIX_*3_asqtad_invert_multi and the batched
 * QPHIX_D3_asqtad_invert_multi_rhs) update the shifted systems' x_j and p_j
 * `steps` steps at a time (each read and written once per group of steps; the
 * base system every step): 8 (the default and the largest) down to 1 (every
 * step; 0 and negative values mean 1).  All give identical results. */
void QPHIX_MI355X_set_msx_fuse(int steps);
/* Links of the mixed solvers' sloppy operator (QPHIX_D3_asqtad_invert_mixed):
 * 32 (default) = FP32, 16 = 16-bit fixed point (per link a float scale, the
 * largest |element|, and int16 elements; ~1.5e-5 relative per element), FP32
 * arithmetic either way.  The FP64 reliable updates keep the result at the
 * requested true residual. */
void QPHIX_MI355X_set_sloppy_link_bits(int bits);
/* Long links of the FP32 sloppy operator (link bits 32): 8 (default) = each
 * long link as 8 reals (U01, U02, U10 and the half-angle tangents of the
 * phases of U00 and U20; the rest follows from unitarity and det = +-1), used
 * when the FP64 links store 12-real long links and every long link has rows 0
 * and 1 orthonormal to 1e-6 and |U01|^2 + |U02|^2 >= 1e-10, else the 12-real
 * form; 12 = always the FP64 copy's
 * form.  ~1.5e-7 relative error per element typical, up to ~5e-5 at rare
 * links where |U00| or |U20| is tiny; the FP64 reliable updates keep the
 * result at the requested true residual.  Read at each solve. */
void QPHIX_MI355X_set_sloppy_long_recon(int recon);
/* One application of the sloppy Dslash (the D3 links' FP32 or 16-bit copy,
 * link_bits 32 or 16) on FP32 vectors: out = D in on `parity`, repeated reps
 * times; returns milliseconds per application (-1 on a contract violation). */
double QPHIX_D3_asqtad_dslash_sloppy(QPHIX_D3_FermionLinksAsqtad *L, QPHIX_F3_ColorVector *out,
                                     QPHIX_F3_ColorVector *in, QPHIX_evenodd_t parity, int link_bits,
                                     int reps);
/* Storage chosen for L: 12 or 18. */
int QPHIX_D3_asqtad_links_recon(QPHIX_D3_FermionLinksAsqtad *L);
int QPHIX_F3_asqtad_links_recon(QPHIX_F3_FermionLinksAsqtad *L);

/* Back-link convention probe (needs both parities).  Returns the largest
 * |back_mu(x) - s * U_mu(x - hop)^dagger| over all back links, for
 * s = -1 in out[0] and s = +1 in out[1], fat links in out[0..1] and long
 * links in out[2..3]. */
void QPHIX_D3_asqtad_links_backcheck(QPHIX_D3_FermionLinksAsqtad *L, double out[4]);

/* ---- solvers ----------------------------------------------------------------
 * Mixed-precision CG with reliable updates for the system of
 * QPHIX_D3_asqtad_invert (ks_d_congrad_fn.cpp:221-619 solves it in one
 * precision): every CG step runs with FP32 links and FP32 (sloppy_bits = 32)
 * or FP16-stored, FP32-computed (sloppy_bits = 16) vectors; the solution, the
 * true residual b - Mx (recomputed with the FP64 Dslash whenever the sloppy
 * residual has dropped by 10x, or on apparent convergence) and all reductions
 * are FP64.  Stops, as the reference's CG does (ks_d_congrad_fn.cpp:424-427),
 * when the true squared relative residual is below res_arg->resid and the
 * relative residue (ks_blas_utils_c.cpp:110-149) is below res_arg->relresid,
 * a target <= 0 counting as met; both are evaluated on the FP64 true residual
 * at each reliable update (with only relresid set, updates happen every 10x
 * drop of the sloppy residual).  After inv_arg->max * inv_arg->nrestart
 * iterations it stops with info->status = QPHIX_FAIL.  out is the initial
 * guess.  res_arg->final_rsq/size_r and final_rel/size_relr return the last
 * true values; final_restart the number of reliable updates; final_iter
 * counts sloppy steps plus true-residual evaluations.  Returns the
 * iterations of the last parity solved. */
int QPHIX_D3_asqtad_invert_mixed(QPHIX_info_t *info, QPHIX_D3_FermionLinksAsqtad *asqtad,
                                 QPHIX_invert_arg_t *inv_arg, QPHIX_resid_arg_t *res_arg,
                                 double mass, QPHIX_D3_ColorVector *out, QPHIX_D3_ColorVector *in,
                                 int sloppy_bits);

/* ---- domain decomposition on one GPU ------------------------------------------
 * asqtad_dslash of a lattice cut into domains along the directions of
 * cut_mask (bit mu, extent >= 8 along each), with every domain exchanging its
 * 3-deep faces with itself (self-loopback, the reference's ASSUME_MULTINODE
 * build, ks_boundary.cpp:199,250-256): the faces of `in` are packed
 * (ks_boundary.cpp:440-556), the interior sites computed from local
 * neighbours, and the boundary sites from local neighbours plus the received
 * faces (:558-690).  The result is bit-identical to asqtad_dslash.  reps > 0
 * returns the average milliseconds of reps calls; reps = 0 makes one call
 * and returns 0. */
double QPHIX_D3_asqtad_dslash_faces(QPHIX_D3_FermionLinksAsqtad *asqtad, QPHIX_D3_ColorVector *out,
                                    QPHIX_D3_ColorVector *in, QPHIX_evenodd_t parity, int cut_mask, int reps);
double QPHIX_F3_asqtad_dslash_faces(QPHIX_F3_FermionLinksAsqtad *asqtad, QPHIX_F3_ColorVector *out,
                                    QPHIX_F3_ColorVector *in, QPHIX_evenodd_t parity, int cut_mask, int reps);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  asqtad_dslash_faces (reps = 0) with its face buffer in
 * a fixed 64 MiB device array per GPU, taken at QPHIX_init and held until
 * the process ends: no hipMalloc, hipFree or allocating library call.
 * Faces needing more than 64 MiB (sizeof(Real) * 36 * sum over cut mu of
 * the sites of one parity in a slice orthogonal to mu) are rejected with
 * QPHIX_FAIL and a recorded error before anything is launched.  Returns
 * QPHIX_SUCCESS or QPHIX_FAIL.  All calls on one GPU share the array: calls
 * that overlap in time (other streams, concurrently replayed graphs) race.
 * With QPHIX_MI355X_set_sync(0) the call returns once queued on the
 * library's stream. */
int QPHIX_D3_asqtad_dslash_faces_noalloc(QPHIX_D3_FermionLinksAsqtad *asqtad, QPHIX_D3_ColorVector *out,
                                         QPHIX_D3_ColorVector *in, QPHIX_evenodd_t parity, int cut_mask);
int QPHIX_F3_asqtad_dslash_faces_noalloc(QPHIX_F3_FermionLinksAsqtad *asqtad, QPHIX_F3_ColorVector *out,
                                         QPHIX_F3_ColorVector *in, QPHIX_evenodd_t parity, int cut_mask);

/* ---- fermion links from the gauge field ---------------------------------------
 * Asqtad coefficients per path, in the physical sign convention (links
 * without staggered phases): fat link = one_link U + three_staple (each of 6
 * staples) + five_staple (each of 24 ordered five-link staples) + seven_staple
 * (each of 48) + lepage (each of 6 two-step staples); naik multiplies
 * U U U.  Asqtad: {5/8, -1/24, 1/16, 1/64, 1/384, -1/16}; fat7 alone:
 * {1/8, 0, 1/16, 1/64, 1/384, 0}.  Tadpole factors go into the coefficients. */
typedef struct {
  double one_link, naik, three_staple, five_staple, seven_staple, lepage;
} QPHIX_MI355X_asqtad_coeffs_t;
/* Fat and Naik links of the thin SU(3) field G (no phases), built on the GPU
 * with MILC's staggered phases and (flags bit 0 clear) the antiperiodic time
 * boundary folded in, and MILC's double-store back links: the handle equals
 * asqtad_create_L_from_raw of those links (MILC forms them on the host and
 * the reference's own rephase/long-link kernels are disabled,
 * ks_dslash_complete_specialization.cpp:238-275).  Arithmetic in FP64. */
QPHIX_D3_FermionLinksAsqtad *QPHIX_D3_asqtad_create_L_from_G(QPHIX_info_t *info,
                                                             const QPHIX_MI355X_asqtad_coeffs_t *coeffs,
                                                             QPHIX_D3_GaugeField *G, int flags);
QPHIX_F3_FermionLinksAsqtad *QPHIX_F3_asqtad_create_L_from_G(QPHIX_info_t *info,
                                                             const QPHIX_MI355X_asqtad_coeffs_t *coeffs,
                                                             QPHIX_F3_GaugeField *G, int flags);
/* HISQ links (Follana et al., Phys. Rev. D 75 (2007) 054502): fat7, projection
 * to U(3), then the second-level smearing {1, -(1 + naik_eps)/24, 1/16, 1/64,
 * 1/384, -1/8} of the projected field; phases and boundary as above. */
QPHIX_D3_FermionLinksAsqtad *QPHIX_D3_hisq_create_L_from_G(QPHIX_info_t *info, QPHIX_D3_GaugeField *G,
                                                           double naik_eps, int flags);
QPHIX_F3_FermionLinksAsqtad *QPHIX_F3_hisq_create_L_from_G(QPHIX_info_t *info, QPHIX_F3_GaugeField *G,
                                                           double naik_eps, int flags);

/* ---- gauge-field update and molecular dynamics --------------------------------
 * U_mu(x) <- exp(eps A_mu(x)) U_mu(x) for every link of G (both parities), A
 * the traceless anti-Hermitian momentum of F (H = -iA, MILC's 8-real layout).
 * The reference updates the gauge field inside update_h_u_gforce_imp
 * (gauge_force_imp.cpp:221-369, update_mg_body with epsilonU = eps, body
 * absent); here it is a call of its own, MILC's update_u.  The exponential is
 * exact to rounding (Cayley-Hamilton form, scaling and squaring), so U stays
 * SU(3); a two-row copy kept for the force kernel is updated with it. */
void QPHIX_D3_update_gauge(QPHIX_info_t *info, QPHIX_D3_GaugeField *G, QPHIX_D3_Force *F, double eps);
void QPHIX_F3_update_gauge(QPHIX_info_t *info, QPHIX_F3_GaugeField *G, QPHIX_F3_Force *F, float eps);
/* nsteps molecular-dynamics steps of size eps for the pure-gauge Symanzik
 * action of coeffs (the force of QPHIX_*3_symanzik_1loop_gauge_force, the
 * update of QPHIX_*3_update_gauge), conserving
 * H = -1/2 sum Tr A^2 - sum_loops c Re Tr U_loop to O(eps^2).
 * scheme 0: leapfrog; 1: Omelyan's 2nd-order minimum-norm integrator
 * (lambda = 0.1931833275...).  G and F are updated in place. */
void QPHIX_D3_gauge_md(QPHIX_info_t *info, QPHIX_D3_GaugeField *G, QPHIX_D3_Force *F,
                       QPHIX_gauge_coeffs_t *coeffs, double eps, int nsteps, int scheme);
void QPHIX_F3_gauge_md(QPHIX_info_t *info, QPHIX_F3_GaugeField *G, QPHIX_F3_Force *F,
                       QPHIX_gauge_coeffs_t *coeffs, float eps, int nsteps, int scheme);

/* Fermion-force outer products (no reference counterpart: milc-qphix has no
 * fermion force).  For the pseudofermion action
 *     S = sum_k coef[k] Re phi_k^dag X_k,  X_k = (4 m_k^2 - D_eo D_oe)^{-1} phi_k,
 * given the solutions X[k] (their even halves; e.g. the outputs of
 * asqtad_invert_multi, several sources being more terms), writes for every
 * site y (both parities) and direction mu, in raw MILC layout (72 reals per
 * site, even sites first, [mu][row][col][re,im]):
 *     oprod_fat [y][mu] = eps(y) sum_k coef[k] psi_k(y+mu)  psi_k(y)^dag
 *     oprod_long[y][mu] = eps(y) sum_k coef[k] psi_k(y+3mu) psi_k(y)^dag
 * where psi_k = X_k on even sites and D_oe X_k on odd sites (D from asqtad)
 * and eps = +1 on even, -1 on odd sites.  Then, for the forward fat and long
 * links U as passed to asqtad_create_L_from_raw (phases folded in; back links
 * -U^dag(y-mu) as MILC double-stores them),
 *     dS = 2 Re sum_{y,mu} Tr[dU_fat oprod_fat] + 2 Re sum Tr[dU_long oprod_long].
 * Either output may be NULL.  Accumulation is in FP64 in both precisions.
 * info: final_sec, final_flop, status. */
void QPHIX_D3_asqtad_fermion_force_oprod(QPHIX_info_t *info, QPHIX_D3_FermionLinksAsqtad *asqtad,
                                         int nterms, const double *coef, QPHIX_D3_ColorVector *X[],
                                         double *oprod_fat, double *oprod_long);
void QPHIX_F3_asqtad_fermion_force_oprod(QPHIX_info_t *info, QPHIX_F3_FermionLinksAsqtad *asqtad,
                                         int nterms, const double *coef, QPHIX_F3_ColorVector *X[],
                                         float *oprod_fat, float *oprod_long);
/* Average ms of the device part of fermion_force_oprod (the D_oe X_k and the
 * outer-product kernel; outputs kept in HBM, no export); oprod_ms (optional)
 * receives the outer-product kernel's own average ms. */
double QPHIX_D3_asqtad_fermion_force_time(QPHIX_D3_FermionLinksAsqtad *asqtad, int nterms,
                                          const double *coef, QPHIX_D3_ColorVector *X[], int reps,
                                          double *oprod_ms);
double QPHIX_F3_asqtad_fermion_force_time(QPHIX_F3_FermionLinksAsqtad *asqtad, int nterms,
                                          const double *coef, QPHIX_F3_ColorVector *X[], int reps,
                                          double *oprod_ms);

/* Multi-shift CG on nrhs sources at once (ks_multicg_offset.cpp:205-689
 * solves one source in one precision).  Solves (4 mass[j]^2 - D D) out[r*nshift+j]
 * = in[r] for every source r and shift j, each to the squared relative true
 * residual res_arg[r*nshift+j]->resid (res_arg[0]->resid where that is 0;
 * a resulting target <= 0 is rejected).  relresid is not used, as in the
 * reference's multi-shift (ks_multicg_offset.cpp:282 reads it, no test uses it).
 * Sources advance in blocks of rhs_block (1, 2, 4 or 8; 0 = automatic) through
 * one multi-right-hand-side FP32 Dslash per half-step, with FP64 solutions and
 * FP64 reliable updates of each source's base system; afterwards every shifted
 * solution whose own true residual misses the target is finished by mixed-
 * precision CG.  out is overwritten.  final_iter = multi-shift steps of that
 * source + refinement iterations; final_restart = reliable updates.
 * Returns the largest multi-shift step count. */
int QPHIX_D3_asqtad_invert_multi_rhs(QPHIX_info_t *info, QPHIX_D3_FermionLinksAsqtad *asqtad,
                                     QPHIX_invert_arg_t *inv_arg, QPHIX_resid_arg_t *res_arg[],
                                     double *mass, int nshift, int nrhs,
                                     QPHIX_D3_ColorVector *out[], QPHIX_D3_ColorVector *in[],
                                     int rhs_block);

/* Dslash on nvec (1..8) FP32 vectors in one launch: each link is read once
 * and applied to all of them (ks_long_dslash.cpp:564-668 handles one vector).
 * out[k] = D in[k] on the sites of `parity`, as QPHIX_F3_asqtad_dslash. */
void QPHIX_F3_asqtad_dslash_multi(QPHIX_F3_FermionLinksAsqtad *asqtad, QPHIX_F3_ColorVector *out[],
                                  QPHIX_F3_ColorVector *in[], int nvec, QPHIX_evenodd_t parity);

/* ---- timing ----------------------------------------------------------------
 * Runs `reps` single-parity Dslash applications back to back on the library's
 * stream and returns the average kernel time in milliseconds measured with
 * HIP events on that stream. */
double QPHIX_D3_asqtad_dslash_time(QPHIX_D3_FermionLinksAsqtad *L, QPHIX_D3_ColorVector *out,
                                   QPHIX_D3_ColorVector *in, QPHIX_evenodd_t parity, int reps);
double QPHIX_F3_asqtad_dslash_time(QPHIX_F3_FermionLinksAsqtad *L, QPHIX_F3_ColorVector *out,
                                   QPHIX_F3_ColorVector *in, QPHIX_evenodd_t parity, int reps);
/* Average ms of the multi-vector Dslash kernel alone (vectors packed once). */
double QPHIX_F3_asqtad_dslash_multi_time(QPHIX_F3_FermionLinksAsqtad *asqtad,
                                         QPHIX_F3_ColorVector *out[], QPHIX_F3_ColorVector *in[],
                                         int nvec, QPHIX_evenodd_t parity, int reps);
/* Runs (once per lattice and Naik storage, then cached) the autotuner of the
 * multi-RHS Dslash that the multi-vector Dslash and the multi-RHS solver use.
 * us (optional, 12 doubles) receives microseconds per source for block sizes
 * NR = 1, 2, 4, 8 x sources per work-item K = 1, 2, 4 (0 where K > NR);
 * best_k (optional) the best K of the best block size.  Returns the best
 * block size, or 0 on failure.  No reference counterpart (MI355X extension). */
int QPHIX_F3_asqtad_mrhs_tune(QPHIX_F3_FermionLinksAsqtad *asqtad, double *us, int *best_k);
/* Plain device-to-device copy of `bytes` bytes, `reps` times; average ms.
 * Used to calibrate achieved HBM bandwidth and profiler byte counters. */
double QPHIX_MI355X_copy_time(size_t bytes, int reps);

#ifdef __cplusplus
}
#endif

#endif /* _QPHIX_MI355X_H */
