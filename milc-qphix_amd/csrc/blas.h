// blas.h -- solver vector kernels (blas.hip) and their launchers.
#pragma once
#include "reduce.h"

namespace qphix {

constexpr int BLAS_BLOCK = 256;
// Blocks of the vector kernels that sum the previous kernel's partials
// themselves (consumer-side reductions).  Every block reads all of them (at
// 32^4, 2048 Dslash partials x 4 sums: 64 KB per block), so larger blocks read
// fewer bytes but leave fewer wavefronts resident.  cons_block() threads per
// block: 512 by default (QPHIX_CONS_BLOCK = 256 or 1024 in the environment),
// compiled for up to CONS_BLOCK_MAX.  Measured at 32^4 (mixed CG, FP32 /
// FP16-stored vectors): 256 -> 0.190 / 0.180 s, 512 -> 0.187 / 0.179 s; 1024
// made the single-reduction update 36.5 us against 21.5 us (half the
// wavefronts resident).
constexpr int CONS_BLOCK_MAX = 1024;
int cons_block();
inline int cons_nb(int vh) { return (vh + cons_block() - 1) / cons_block(); }

ReduceCtl reduce_ctl(SolverState *st, int op);
double *reduce_result_ptr();  // where reductions without a state leave their sums

template <typename T> void blas_norm2(const T *v, SolverState *st, int op);
template <typename T>
void blas_cg_update(T *x, const T *p, T *r, const T *ttt, SolverState *st, const int *gate);
template <typename T> void blas_cg_axpy(const T *r, T *p, const SolverState *st, const int *gate);
// One CG step after a p.Ap Dslash launched with consumer_finalize (nb_pkp
// blocks of partials): update x, r and |r|^2, then p, each kernel summing the
// previous one's partials itself (no finalize launches).
template <typename T>
void blas_cg_step_c(T *x, T *p, T *r, const T *ttt, SolverState *st, int *gate, int nb_pkp);
double *partials_b();
// One single-reduction (FEWSUMS) CG step after an EPI_PKP_FS Dslash with
// nb_pkp blocks of partials: x, r and p in one pass (k_cg_step_fs).
template <typename T>
void blas_cg_step_fs(T *x, T *p, T *r, const T *ttt, SolverState *st, int *gate, int nb_pkp);
// Which solvers take single-reduction steps (fewsums_mode(), common.h): bit 0
// the one-precision CG (QPHIX_*3_asqtad_invert, when relresid is off), bit 1
// the sloppy steps of the mixed-precision CG.
template <typename T> void blas_ms_update_r(T *r, const T *ttt, SolverState *st, const int *gate);
template <typename T>
void blas_ms_update_pm(T *const *x, T *const *p, const T *r, int nshift, const SolverState *st,
                       const int *gate);
template <typename T>
void blas_ms_update_dest(T *const *x, T *const *p, int nshift, const SolverState *st);
// update_pm with the shifted systems up to st->fuse steps at a time (rprev:
// fuse - 1 saved residuals); flush = true applies what is still pending.
template <typename T>
void blas_ms_update_pm_def(T *const *x, T *const *p, const T *r, T *rprev, int nshift, SolverState *st,
                           const int *gate, bool flush);
void set_int(int *p, int v);

} // namespace qphix
