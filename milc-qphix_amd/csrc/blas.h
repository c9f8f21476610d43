// blas.h -- solver vector kernels (blas.hip) and their launchers.
#pragma once
#include "reduce.h"

namespace qphix {

constexpr int BLAS_BLOCK = 256;

ReduceCtl reduce_ctl(SolverState *st, int op);
double *reduce_result_ptr();  // where reductions without a state leave their sums

template <typename T> void blas_norm2(const T *v, SolverState *st, int op);
template <typename T>
void blas_cg_update(T *x, const T *p, T *r, const T *ttt, SolverState *st, const int *gate);
template <typename T> void blas_cg_axpy(const T *r, T *p, const SolverState *st, const int *gate);
template <typename T> void blas_ms_update_r(T *r, const T *ttt, SolverState *st, const int *gate);
template <typename T>
void blas_ms_update_pm(T *const *x, T *const *p, const T *r, int nshift, const SolverState *st,
                       const int *gate);
template <typename T>
void blas_ms_update_dest(T *const *x, T *const *p, int nshift, const SolverState *st);
void set_int(int *p, int v);

} // namespace qphix
