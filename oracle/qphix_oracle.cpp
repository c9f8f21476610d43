/* qphix_oracle.cpp -- CPU oracle for the staggered hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; tests/, the smoke check and the CPU-baseline leg of the
 * benchmark use it as the checker.  It restates, in plain scalar C++ with
 * OpenMP site loops, what the reference computes for:
 *
 *   geometry + neighbour tables    qphix_misc.cpp:46-230, 494-509
 *   raw <-> per-site link slots    qphix_d3.cpp:970-1085 (C4: slot 2mu back, 2mu+1 forward)
 *   Dslash site loop               ks_long_dslash.cpp:564-668 (16 neighbour terms, summed)
 *   CG                             ks_d_congrad_fn.cpp:62-188 (BLAS), 221-619 (driver)
 *   multi-shift CG                 ks_multicg_offset.cpp:69-184 (BLAS), 205-689 (driver)
 *   relative residue               ks_blas_utils_c.cpp:110-149
 *   fermion-force outer products   no reference counterpart (derived from the action)
 *
 * Layout is the reference's VECLEN=1 layout without the XY_PAD site (padding
 * never changes a result): site index ind = xh + Nxh*(y + Ny*(z + Nz*t)) within
 * a parity, i.e. the raw MILC index of that parity.
 *
 * The reference's kernel body (milc-qphix-codegen, absent) would fix the
 * rounding order of the 16-term sum; the oracle sums in the order
 * mu = X,Y,Z,T and, within mu, fat fwd, fat back, long fwd, long back.
 * Parity pins for every function here are the closed forms and invariants in
 * tests/test_oracle.py (free-field spectrum, anti-Hermiticity, gauge
 * covariance, plane-wave CG/multi-shift solutions, dense solves at 4^4).
 *
 * Deliberate differences from the reference, shared with the product:
 *  - EVENODD solves the even half with the even operator, then the odd half
 *    with the odd one.  The reference's drivers take parity = 3 & 1 = 1 for the
 *    even half (ks_d_congrad_fn.cpp:285,240-249), i.e. the odd-checkerboard
 *    operator on even-site data; that is not reproduced.
 *  - multi-shift `finished[]` starts at 0 (the reference mallocs it without
 *    initialising, ks_multicg_offset.cpp:332).
 */
#include <cmath>
#include <complex>
#include <initializer_list>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { P_EVEN = 2, P_ODD = 1, P_EVENODD = 3 };

struct Geom {
  int nx = 0, ny = 0, nz = 0, nt = 0, nxh = 0;
  long vh = 0;                    /* sites per parity */
  std::vector<unsigned> neigh[2]; /* [cb][16*ind + k], qphix_misc.cpp:88-221 */
};
Geom G;

/* qphix_misc.cpp:88-221 restated with periodic (and double) wrap.  cb is the
 * parity of the site the table belongs to (0 even, 1 odd). */
void build_neighbours() {
  const int Vxh = G.nxh, Vy = G.ny, Vz = G.nz, Vt = G.nt;
  const long Pxy = (long)Vxh * Vy, Pxyz = Pxy * Vz;
  auto wrap = [](long v, long n) { v %= n; return v < 0 ? v + n : v; };
  for (int cb = 0; cb < 2; ++cb) {
    G.neigh[cb].assign(16 * G.vh, 0);
    for (int t = 0; t < Vt; ++t)
      for (int z = 0; z < Vz; ++z)
        for (int y = 0; y < Vy; ++y)
          for (int x = 0; x < Vxh; ++x) {
            const int xodd = (y + z + t + cb) & 1;
            const long ind = t * Pxyz + z * Pxy + (long)y * Vxh + x;
            const long row = ind - x;
            unsigned *n = &G.neigh[cb][16 * ind];
            const int dx1[2] = {xodd - 1, xodd};
            const int dx3[2] = {xodd - 2, xodd + 1};
            for (int fb = 0; fb < 2; ++fb) {
              n[0 + fb] = (unsigned)(row + wrap(x + dx1[fb], Vxh));
              n[8 + fb] = (unsigned)(row + wrap(x + dx3[fb], Vxh));
            }
            const int dir_len[3] = {Vy, Vz, Vt};
            const int coord[3] = {y, z, t};
            const long stride[3] = {Vxh, Pxy, Pxyz};
            for (int d = 0; d < 3; ++d) {
              const long base = ind - coord[d] * stride[d];
              for (int fb = 0; fb < 2; ++fb) {
                const int s = fb ? 1 : -1;
                n[2 + 2 * d + fb] = (unsigned)(base + wrap(coord[d] + s, dir_len[d]) * stride[d]);
                n[10 + 2 * d + fb] = (unsigned)(base + wrap(coord[d] + 3 * s, dir_len[d]) * stride[d]);
              }
            }
          }
  }
}

/* Links of one output parity in the reference's per-site slot order:
 * [ind][slot 0..7][row][col][re,im] for fat and long (qcd_data_types.h:36-44). */
template <typename T> struct Links {
  std::vector<T> fat[2], lng[2];
};

template <typename T>
Links<T> *links_create(const T *fat, const T *lng, const T *fatback, const T *lngback) {
  auto *L = new Links<T>;
  for (int cb = 0; cb < 2; ++cb) {
    L->fat[cb].assign(144 * G.vh, 0);
    L->lng[cb].assign(144 * G.vh, 0);
    const long off = cb ? 72 * G.vh : 0; /* raw odd block at 72*even_sites */
#pragma omp parallel for
    for (long i = 0; i < G.vh; ++i)
      for (int mu = 0; mu < 4; ++mu)
        for (int e = 0; e < 18; ++e) {
          const long src = off + 72 * i + 18 * mu + e;
          L->fat[cb][144 * i + 18 * (2 * mu + 1) + e] = fat[src];
          L->fat[cb][144 * i + 18 * (2 * mu) + e] = fatback[src];
          L->lng[cb][144 * i + 18 * (2 * mu + 1) + e] = lng[src];
          L->lng[cb][144 * i + 18 * (2 * mu) + e] = lngback[src];
        }
  }
  return L;
}

/* acc += U * v, U 3x3 complex row-major [r][c][re,im], v [c][re,im] */
template <typename T> inline void mat_vec_acc(const T *U, const T *v, T *acc) {
  for (int r = 0; r < 3; ++r) {
    T re = 0, im = 0;
    for (int c = 0; c < 3; ++c) {
      const T ur = U[6 * r + 2 * c], ui = U[6 * r + 2 * c + 1];
      re += ur * v[2 * c] - ui * v[2 * c + 1];
      im += ur * v[2 * c + 1] + ui * v[2 * c];
    }
    acc[2 * r] += re;
    acc[2 * r + 1] += im;
  }
}

/* ks_long_dslash.cpp:564-668: for every site of output parity cb,
 * out = sum_k slot_k * in[neigh_k].  `in` and `out` are half-lattice arrays
 * of 6 reals per site. */
template <typename T> void dslash(const Links<T> *L, T *out, const T *in, int cb) {
  const unsigned *nt = G.neigh[cb].data();
  const T *gf = L->fat[cb].data();
  const T *gl = L->lng[cb].data();
#pragma omp parallel for schedule(static)
  for (long i = 0; i < G.vh; ++i) {
    T acc[6] = {0, 0, 0, 0, 0, 0};
    for (int mu = 0; mu < 4; ++mu) {
      mat_vec_acc(gf + 144 * i + 18 * (2 * mu + 1), in + 6 * (long)nt[16 * i + 2 * mu + 1], acc);
      mat_vec_acc(gf + 144 * i + 18 * (2 * mu), in + 6 * (long)nt[16 * i + 2 * mu], acc);
      mat_vec_acc(gl + 144 * i + 18 * (2 * mu + 1), in + 6 * (long)nt[16 * i + 8 + 2 * mu + 1], acc);
      mat_vec_acc(gl + 144 * i + 18 * (2 * mu), in + 6 * (long)nt[16 * i + 8 + 2 * mu], acc);
    }
    for (int e = 0; e < 6; ++e) out[6 * i + e] = acc[e];
  }
}

/* Reductions accumulate in the vector type T, as the reference's OpenMP
 * reductions do (`fptype rsq = 0.0`, ks_d_congrad_fn.cpp:65,90,139,163), in a
 * fixed blocked order so results do not depend on the thread count. */
template <typename T, typename F> double blocked_sum(long n, F term) {
  const long B = 4096;
  const long nb = (n + B - 1) / B;
  std::vector<T> part(nb);
#pragma omp parallel for schedule(static)
  for (long b = 0; b < nb; ++b) {
    T s = 0;
    const long e = (b + 1) * B < n ? (b + 1) * B : n;
    for (long i = b * B; i < e; ++i) s += term(i);
    part[b] = s;
  }
  T tot = 0;
  for (long b = 0; b < nb; ++b) tot += part[b];
  return (double)tot;
}

/* ks_blas_utils_c.cpp:110-149.  Per site and colour the reference squares and
 * adds the real and imaginary parts in the vector type (fptype) and widens the
 * result to double (calc_soalen_rel_residue, :118-121); sites with r = 0 are
 * skipped, sites with x = 0 count 1. */
template <typename T> double relative_residue(const T *r, const T *x) {
  double res = 0;
  for (long i = 0; i < G.vh; ++i) {
    double num = 0, den = 0;
    for (int c = 0; c < 3; ++c) {
      const T *rc = r + 6 * i + 2 * c, *xc = x + 6 * i + 2 * c;
      num += (double)(T)(rc[0] * rc[0] + rc[1] * rc[1]);
      den += (double)(T)(xc[0] * xc[0] + xc[1] * xc[1]);
    }
    if (num == 0) continue;
    res += den == 0 ? 1.0 : num / den;
  }
  return std::sqrt(2 * res / ((double)G.nx * G.ny * G.nz * G.nt));
}

struct InvertOut {
  double final_rsq, final_rel, size_r, size_relr, final_flop;
  int final_iter, final_restart, ret;
};

/* ks_d_congrad_fn.cpp:221-619 restated.  `x` and `b` are full raw arrays
 * (even block then odd block); parity is the API parity.
 *
 * fewsums = 1 restates the FEWSUMS variant (ks_d_congrad_fn.cpp:260-263, 402,
 * 450-541) with the sums the C++ port leaves out taken from MILC's code kept
 * in its comments (:119-128, :150-158): the p.Ap pass also sums
 * c_tr = Re<ttt|r> and c_tt = <ttt|ttt>; the x/r pass sums the actual |r|^2,
 * used as oldrsq one step later; rsq = oldrsq + 2 a c_tr + a^2 c_tt replaces
 * the summed |r|^2 for the stopping test and b. */
template <typename T>
void invert(const Links<T> *L, int parity_api, int niter, int max_restarts, double resid,
            double relresid, T mass, T *x_full, const T *b_full, InvertOut *o, int fewsums = 0) {
  const long n = 6 * G.vh;
  std::vector<T> ttt(n), tmp(n), p(n), r(n);
  const T rsqmin = (T)resid, relrsqmin = (T)relresid;
  const int max_cg = max_restarts * niter;
  const T msq_x4 = (T)(4.0 * mass * mass);
  int total_iters = 0, iteration = 0, nrestart = 0;
  double rsq = 0, relrsq = 0, oldrsq, pkp, actual_rsq = 0, c_tr = 0, c_tt = 0;
  double size_r = 0, size_relr = 0, final_rsq = 0, final_rel = 0;
  int todo[2], ntodo = 0;
  if (parity_api == P_EVENODD) { todo[0] = 0; todo[1] = 1; ntodo = 2; }
  else { todo[0] = parity_api == P_EVEN ? 0 : 1; ntodo = 1; }
  for (int pass = 0; pass < ntodo; ++pass) {
    const int parity = todo[pass], other = 1 - parity;
    T *xd = x_full + (parity ? n : 0);
    const T *bs = b_full + (parity ? n : 0);
    const double source_norm = blocked_sum<T>(n, [&](long i) { return bs[i] * bs[i]; });
    nrestart = 0;
    total_iters += iteration;
    iteration = 0;
    size_r = 0;
    size_relr = 0;
    if (source_norm == 0.0) {
      for (long i = 0; i < n; ++i) xd[i] = 0;
      final_rel = 0;
      final_rsq = 0;
      continue;
    }
    while (true) {
      if ((iteration % niter == 0) ||
          ((rsqmin <= 0 || rsqmin > size_r) && (relrsqmin <= 0 || relrsqmin > size_relr))) {
        dslash(L, tmp.data(), xd, other);
        dslash(L, ttt.data(), tmp.data(), parity);
        /* calc_rsq, ks_d_congrad_fn.cpp:82-99 */
        rsq = blocked_sum<T>(n, [&](long i) {
          ttt[i] = ttt[i] + (-msq_x4) * xd[i];
          r[i] = bs[i] + ttt[i];
          p[i] = r[i];
          return r[i] * r[i];
        });
        if (relrsqmin > 0) relrsq = relative_residue(r.data(), xd);
        actual_rsq = rsq;
        final_rsq = (T)rsq / source_norm;
        final_rel = (T)relrsq;
        iteration++;
        if (iteration >= max_cg || nrestart >= max_restarts ||
            ((rsqmin <= 0 || rsqmin > final_rsq) && (relrsqmin <= 0 || relrsqmin > final_rel)))
          break;
        if (iteration != 0) nrestart++;
      }
      oldrsq = fewsums ? actual_rsq : rsq;
      dslash(L, tmp.data(), p.data(), other);
      dslash(L, ttt.data(), tmp.data(), parity);
      /* calc_pkp, ks_d_congrad_fn.cpp:159-171 */
      pkp = blocked_sum<T>(n, [&](long i) {
        ttt[i] = ttt[i] + (-msq_x4) * p[i];
        return p[i] * ttt[i];
      });
      if (fewsums) {
        c_tr = blocked_sum<T>(n, [&](long i) { return ttt[i] * r[i]; });
        c_tt = blocked_sum<T>(n, [&](long i) { return ttt[i] * ttt[i]; });
      }
      iteration++;
      const T a = (T)(-rsq / pkp);
      /* calc_rsq2, ks_d_congrad_fn.cpp:131-147 */
      const double summed = blocked_sum<T>(n, [&](long i) {
        xd[i] = xd[i] + a * p[i];
        r[i] = r[i] + a * ttt[i];
        return r[i] * r[i];
      });
      if (fewsums) {
        actual_rsq = summed;
        rsq = oldrsq + 2.0 * a * c_tr + (double)a * a * c_tt;  // :531
      } else {
        rsq = summed;
      }
      if (relrsqmin > 0) relrsq = relative_residue(r.data(), xd);
      size_r = (T)rsq / source_norm;
      size_relr = (T)relrsq;
      final_rsq = size_r;
      const T b = (T)rsq / oldrsq;
      /* axpy, ks_d_congrad_fn.cpp:178-188 */
#pragma omp parallel for
      for (long i = 0; i < n; ++i) p[i] = r[i] + b * p[i];
    }
  }
  total_iters += iteration;
  o->final_rsq = final_rsq;
  o->final_rel = final_rel;
  o->size_r = size_r;
  o->size_relr = size_relr;
  o->final_iter = total_iters;
  o->final_restart = nrestart;
  o->final_flop = 1187.0 * total_iters * 2.0 * G.vh;
  o->ret = iteration;
}

/* ks_multicg_offset.cpp:205-689 restated.  x[j] full raw arrays. */
template <typename T>
void invert_multi(const Links<T> *L, int parity_api, int niter, int max_restarts, double resid0,
                  const T *mass, int num_offsets, T **x_full, const T *b_full, InvertOut *o) {
  o->ret = 0;
  if (num_offsets == 0) return;
  const long n = 6 * G.vh;
  std::vector<T> ttt(n), tmp(n), r(n);
  std::vector<std::vector<T>> pm(num_offsets, std::vector<T>(n));
  std::vector<T> shifts(num_offsets), zeta_i(num_offsets), zeta_im1(num_offsets),
      zeta_ip1(num_offsets), beta_i(num_offsets), beta_im1(num_offsets), alpha(num_offsets);
  std::vector<int> finished(num_offsets, 0);
  const T rsqmin = (T)resid0;
  const int max_cg = max_restarts * niter;
  double offset_low = 1.0e+20;
  int j_low = -1;
  for (int j = 0; j < num_offsets; ++j) {
    shifts[j] = (T)(mass[j] * mass[j] * 4.0);
    if (shifts[j] < offset_low) { offset_low = shifts[j]; j_low = j; }
  }
  for (int j = 0; j < num_offsets; ++j)
    if (j != j_low) shifts[j] -= shifts[j_low];
  int iteration = 0, total_iters = 0;
  double rsq = 0, oldrsq, pkp, source_norm = 0;
  int todo[2], ntodo;
  if (parity_api == P_EVENODD) { todo[0] = 0; todo[1] = 1; ntodo = 2; }
  else { todo[0] = parity_api == P_EVEN ? 0 : 1; ntodo = 1; }
  for (int pass = 0; pass < ntodo; ++pass) {
    const int parity = todo[pass], other = 1 - parity;
    const long voff = parity ? n : 0;
    const T *bs = b_full + voff;
    iteration = 0;
    int num_now = num_offsets;
    source_norm = blocked_sum<T>(n, [&](long i) { return bs[i] * bs[i]; });
    for (int j = 0; j < num_offsets; ++j) {
      std::memset(x_full[j] + voff, 0, sizeof(T) * n);
      std::memcpy(pm[j].data(), bs, sizeof(T) * n);
      finished[j] = 0;
    }
    std::memcpy(r.data(), bs, sizeof(T) * n);
    rsq = source_norm;
    if (source_norm == 0.0) continue;
    const double rsqstop = rsqmin * source_norm;
    for (int j = 0; j < num_offsets; ++j) {
      zeta_im1[j] = zeta_i[j] = 1.0;
      beta_im1[j] = -1.0;
      alpha[j] = 0.0;
    }
    bool converged = false;
    do {
      oldrsq = rsq;
      dslash(L, tmp.data(), pm[j_low].data(), other);
      dslash(L, ttt.data(), tmp.data(), parity);
      const T sh = (T)(-1.0 * offset_low);
      T *pl = pm[j_low].data();
      pkp = blocked_sum<T>(n, [&](long i) {
        ttt[i] = ttt[i] + sh * pl[i];
        return pl[i] * ttt[i];
      });
      iteration++;
      total_iters++;
      beta_i[j_low] = (T)(-rsq / pkp);
      zeta_ip1[j_low] = 1.0;
      for (int j = 0; j < num_now; ++j) {
        if (j == j_low) continue;
        zeta_ip1[j] = zeta_i[j] * zeta_im1[j] * beta_im1[j_low];
        const T c1 = beta_i[j_low] * alpha[j_low] * (zeta_im1[j] - zeta_i[j]);
        const T c2 = zeta_im1[j] * beta_im1[j_low] * (1.0 + shifts[j] * beta_i[j_low]);
        if (c1 + c2 != (T)0.0) zeta_ip1[j] /= c1 + c2;
        else { zeta_ip1[j] = 0.0; finished[j] = 1; }
        if (zeta_i[j] != (T)0.0) beta_i[j] = beta_i[j_low] * zeta_ip1[j] / zeta_i[j];
        else { zeta_ip1[j] = 0.0; beta_i[j] = 0.0; finished[j] = 1; }
        if (finished[j] == 1 && j == num_now - 1 && j > j_low) num_now--;
      }
      const T bl = beta_i[j_low];
      rsq = blocked_sum<T>(n, [&](long i) {
        r[i] = r[i] + bl * ttt[i];
        return r[i] * r[i];
      });
      if (rsq <= rsqstop) {
        /* update_dest, ks_multicg_offset.cpp:142-159 */
        for (int j = 0; j < num_now; ++j) {
          T *xd = x_full[j] + voff;
          const T *pj = pm[j].data();
          const T bj = beta_i[j];
#pragma omp parallel for
          for (long i = 0; i < n; ++i) xd[i] = xd[i] + bj * pj[i];
        }
        converged = true;
        break;
      }
      alpha[j_low] = (T)(rsq / oldrsq);
      for (int j = 0; j < num_now; ++j) {
        if (j == j_low) continue;
        if (zeta_i[j] * beta_i[j_low] != (T)0.0)
          alpha[j] = alpha[j_low] * zeta_ip1[j] * beta_i[j] / (zeta_i[j] * beta_i[j_low]);
        else { alpha[j] = 0.0; finished[j] = 1; }
      }
      /* update_pm, ks_multicg_offset.cpp:161-184 */
      for (int j = 0; j < num_now; ++j) {
        T *xd = x_full[j] + voff;
        T *pj = pm[j].data();
        const T bj = beta_i[j], zj = zeta_ip1[j], aj = alpha[j];
#pragma omp parallel for
        for (long i = 0; i < n; ++i) {
          xd[i] = xd[i] + bj * pj[i];
          const T tt = r[i] * zj;
          pj[i] = tt + aj * pj[i];
        }
      }
      for (int j = 0; j < num_now; ++j) {
        beta_im1[j] = beta_i[j];
        zeta_im1[j] = zeta_i[j];
        zeta_i[j] = zeta_ip1[j];
      }
    } while (iteration < max_cg);
    (void)converged;
  }
  o->final_rsq = source_norm != 0 ? (T)rsq / source_norm : 0.0;
  o->size_r = o->final_rsq;
  o->final_iter = iteration;
  o->final_restart = 0;
  o->final_flop = (1205.0 + 15.0 * num_offsets) * total_iters * 2.0 * G.vh;
  o->ret = iteration;
}

// ---------------------------------------------------------------------------
// Symanzik one-loop gauge action and force, loop-centric.
//
// The reference's kernel bodies for this (gforce_imp_body / update_mg_body,
// gauge_force_imp_complete_specialization.cpp:24-87) are in the absent codegen
// repository; its orchestration (gauge_force_imp.cpp:136-369) fixes only the
// arguments: coefficients for the plaquette, 1x2 rectangle and parallelogram
// loops and a step eps applied to the momentum.  The oracle defines
//   S(U) = - sum over geometric loops P of c_P Re Tr U_P     (each loop once)
//   mom_mu(x) <- mom_mu(x) - eps * TA( W_mu(x) ),
//   W_mu(x) = sum over loops P through link (x,mu) of c_P * (P traced from x,
//             first step +mu),  TA(M) = (M - M^dag)/2 - Tr(M - M^dag)/6,
// so that dS/dt under U -> exp(tX) U equals -Re Tr(X W) (pinned by finite
// differences in tests).  Loops are enumerated geometrically per site:
// plaquettes (mu<nu), rectangles 2x1 (long side mu, mu != nu) and the 16
// three-dimensional six-link loops (four skew hexagons per elementary cube),
// and each loop adds its contribution to every link it contains.
// ---------------------------------------------------------------------------
typedef std::complex<double> cd;
struct M3 {
  cd a[9];
};
inline M3 mmul(const M3 &x, const M3 &y) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cd s = 0;
      for (int k = 0; k < 3; ++k) s += x.a[3 * i + k] * y.a[3 * k + j];
      r.a[3 * i + j] = s;
    }
  return r;
}
inline M3 madj(const M3 &x) {
  M3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.a[3 * i + j] = std::conj(x.a[3 * j + i]);
  return r;
}
inline M3 meye() {
  M3 r;
  for (int i = 0; i < 9; ++i) r.a[i] = (i % 4 == 0) ? 1.0 : 0.0;
  return r;
}

struct GLat {
  int L[4];
  long vh;
  long raw(const int *x) const {
    int c[4];
    for (int d = 0; d < 4; ++d) c[d] = ((x[d] % L[d]) + L[d]) % L[d];
    const int p = (c[0] + c[1] + c[2] + c[3]) & 1;
    return p * vh + c[0] / 2 + (long)(L[0] / 2) * (c[1] + (long)L[1] * (c[2] + (long)L[2] * c[3]));
  }
};

inline M3 get_link(const double *U, const GLat &g, const int *x, int mu) {
  const double *p = U + 72 * g.raw(x) + 18 * mu;
  M3 r;
  for (int e = 0; e < 9; ++e) r.a[e] = cd(p[2 * e], p[2 * e + 1]);
  return r;
}

// A loop: start vertex and signed steps (+d+1 forward, -(d+1) backward).
struct Loop {
  int x[4];
  int n;
  int step[6];
  double c;
};

template <typename F> void for_each_loop(const GLat &g, const double *coef, F f) {
  int x[4];
  for (x[3] = 0; x[3] < g.L[3]; ++x[3])
    for (x[2] = 0; x[2] < g.L[2]; ++x[2])
      for (x[1] = 0; x[1] < g.L[1]; ++x[1])
        for (x[0] = 0; x[0] < g.L[0]; ++x[0]) {
          Loop l;
          for (int d = 0; d < 4; ++d) l.x[d] = x[d];
          for (int mu = 0; mu < 4; ++mu)
            for (int nu = 0; nu < 4; ++nu) {
              if (nu == mu) continue;
              const int m = mu + 1, n = nu + 1;
              if (mu < nu && coef[0] != 0) {
                l.n = 4; l.c = coef[0];
                int s[4] = {m, n, -m, -n};
                for (int k = 0; k < 4; ++k) l.step[k] = s[k];
                f(l);
              }
              if (coef[1] != 0) {
                l.n = 6; l.c = coef[1];
                int s[6] = {m, m, n, -m, -m, -n};
                for (int k = 0; k < 6; ++k) l.step[k] = s[k];
                f(l);
              }
            }
          if (coef[2] == 0) continue;
          for (int mu = 0; mu < 4; ++mu)
            for (int nu = mu + 1; nu < 4; ++nu)
              for (int rho = nu + 1; rho < 4; ++rho) {
                const int m = mu + 1, n = nu + 1, r = rho + 1;
                const int hex[4][6] = {{m, n, r, -m, -n, -r},
                                       {m, r, n, -m, -r, -n},
                                       {n, m, r, -n, -m, -r},
                                       {n, -m, r, -n, m, -r}};
                for (int h = 0; h < 4; ++h) {
                  Loop q;
                  for (int d = 0; d < 4; ++d) q.x[d] = x[d];
                  if (h == 3) q.x[mu] += 1;  // the fourth hexagon starts at x + mu
                  q.n = 6; q.c = coef[2];
                  for (int k = 0; k < 6; ++k) q.step[k] = hex[h][k];
                  f(q);
                }
              }
        }
}

// Link matrices along a loop: M_k for step k, and the vertices v_k.
inline void loop_links(const double *U, const GLat &g, const Loop &l, M3 *Mk, int (*v)[4]) {
  int y[4] = {l.x[0], l.x[1], l.x[2], l.x[3]};
  for (int k = 0; k < l.n; ++k) {
    for (int d = 0; d < 4; ++d) v[k][d] = y[d];
    const int s = l.step[k], d = (s > 0 ? s : -s) - 1;
    if (s > 0) {
      Mk[k] = get_link(U, g, y, d);
      y[d] += 1;
    } else {
      y[d] -= 1;
      Mk[k] = madj(get_link(U, g, y, d));
    }
  }
}

double gauge_action(const double *U, const GLat &g, const double *coef, const int *box_x,
                    int box_r, int link_mu) {
  double S = 0;
  for_each_loop(g, coef, [&](const Loop &l) {
    if (box_x) {
      for (int d = 0; d < 4; ++d) {
        int dd = ((l.x[d] - box_x[d]) % g.L[d] + g.L[d]) % g.L[d];
        dd = dd > g.L[d] / 2 ? g.L[d] - dd : dd;
        if (dd > box_r) return;
      }
    }
    M3 Mk[6];
    int v[6][4];
    loop_links(U, g, l, Mk, v);
    if (box_x && link_mu >= 0) {  // only loops that contain link (box_x, link_mu)
      const long target = g.raw(box_x);
      bool hit = false;
      for (int k = 0; k < l.n && !hit; ++k) {
        const int s = l.step[k], d = (s > 0 ? s : -s) - 1;
        if (d != link_mu) continue;
        hit = g.raw(s > 0 ? v[k] : v[(k + 1) % l.n]) == target;
      }
      if (!hit) return;
    }
    M3 P = Mk[0];
    for (int k = 1; k < l.n; ++k) P = mmul(P, Mk[k]);
    S -= l.c * (P.a[0].real() + P.a[4].real() + P.a[8].real());
  });
  return S;
}

void gauge_force_W(const double *U, const GLat &g, const double *coef, std::vector<M3> &W) {
  W.assign(8 * g.vh * 4, M3{});
  for_each_loop(g, coef, [&](const Loop &l) {
    M3 Mk[6];
    int v[6][4];
    loop_links(U, g, l, Mk, v);
    for (int k = 0; k < l.n; ++k) {
      const int s = l.step[k], d = (s > 0 ? s : -s) - 1;
      // cyclic product starting at k (forward link) or at k+1 (backward link)
      const int k0 = s > 0 ? k : (k + 1) % l.n;
      M3 P = Mk[k0];
      for (int j = 1; j < l.n; ++j) P = mmul(P, Mk[(k0 + j) % l.n]);
      long site;
      if (s > 0) site = g.raw(v[k]);
      else site = g.raw(v[(k + 1) % l.n]);
      if (s < 0) P = madj(P);
      M3 &w = W[4 * site + d];
      for (int e = 0; e < 9; ++e) w.a[e] += l.c * P.a[e];
    }
  });
}

/* Fermion-force outer products.  No reference counterpart (milc-qphix has no
 * fermion force); the quantity is fixed by the action
 *   S = sum_k c_k Re phi^dag X_k,  X_k = (sigma_k - D_eo D_oe)^{-1} phi,
 * whose derivative with respect to a forward link U(y,mu) (back link
 * -U(y-mu)^dag, as in `dslash` with MILC's double-store) is
 *   dS = 2 Re Tr[dU(y,mu) O(y,mu)],  O(y,mu) = eps(y) sum_k c_k psi_k(y+mu) psi_k(y)^dag,
 * psi_k = X_k on even sites and D_oe X_k on odd ones, eps = +1 even, -1 odd;
 * Naik links likewise with y+3mu.  The pin is the exact-difference FD of S in
 * tests/test_oracle_force.py.  X[k] are full raw vectors (even half used);
 * o1, o3 raw 72-real/site arrays ([mu][row][col][re,im], even sites first). */
void fermion_oprod(const Links<double> *L, int nt, const double *c, const double *const *X,
                   double *o1, double *o3) {
  const long vh = G.vh;
  std::vector<std::vector<double>> Y(nt);
  for (int k = 0; k < nt; ++k) {
    Y[k].assign(6 * vh, 0.0);
    dslash(L, Y[k].data(), X[k], 1);
  }
  for (int cb = 0; cb < 2; ++cb) {
    const unsigned *ng = G.neigh[cb].data();
    const double eps = cb ? -1.0 : 1.0;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < vh; ++i)
      for (int mu = 0; mu < 4; ++mu) {
        double a1[18] = {0}, a3[18] = {0};
        for (int k = 0; k < nt; ++k) {
          const double *here = cb ? Y[k].data() + 6 * i : X[k] + 6 * i;
          const double *other = cb ? X[k] : Y[k].data();
          const double *f1 = other + 6 * (long)ng[16 * i + 2 * mu + 1];
          const double *f3 = other + 6 * (long)ng[16 * i + 8 + 2 * mu + 1];
          for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) {
              const double br = here[2 * q], bi = -here[2 * q + 1]; /* conj(psi(y)_q) */
              a1[6 * r + 2 * q] += c[k] * (f1[2 * r] * br - f1[2 * r + 1] * bi);
              a1[6 * r + 2 * q + 1] += c[k] * (f1[2 * r] * bi + f1[2 * r + 1] * br);
              a3[6 * r + 2 * q] += c[k] * (f3[2 * r] * br - f3[2 * r + 1] * bi);
              a3[6 * r + 2 * q + 1] += c[k] * (f3[2 * r] * bi + f3[2 * r + 1] * br);
            }
        }
        const long o = 72 * (cb * vh + i) + 18 * mu;
        for (int e = 0; e < 18; ++e) {
          o1[o + e] = eps * a1[e];
          o3[o + e] = eps * a3[e];
        }
      }
  }
}

} // namespace

extern "C" {

/* S(U) for a full raw gauge field (72 reals/site, even sites first).
 * coef = {plaquette, rectangle, parallelogram}.  box != NULL restricts the
 * sum to loops starting within Chebyshev distance box_r of box (periodic);
 * with link_mu >= 0 as well, only loops containing link (box, link_mu). */
double oracle_gauge_action(const int *latsize, const double *U, const double *coef,
                           const int *box, int box_r, int link_mu) {
  GLat g;
  for (int d = 0; d < 4; ++d) g.L[d] = latsize[d];
  g.vh = (long)latsize[0] * latsize[1] * latsize[2] * latsize[3] / 2;
  return gauge_action(U, g, coef, box, box_r, link_mu);
}

/* mom <- mom - eps * TA(W) for every link; mom is a raw anti-Hermitian array
 * (72 reals/site).  Also returns W (72 reals/site) when W_out != NULL. */
void oracle_gauge_force(const int *latsize, const double *U, const double *coef, double eps,
                        double *mom, double *W_out) {
  GLat g;
  for (int d = 0; d < 4; ++d) g.L[d] = latsize[d];
  g.vh = (long)latsize[0] * latsize[1] * latsize[2] * latsize[3] / 2;
  std::vector<M3> W;
  gauge_force_W(U, g, coef, W);
  for (long l = 0; l < 8 * g.vh; ++l) {
    const M3 &w = W[l];
    M3 ta;
    cd tr = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) ta.a[3 * i + j] = 0.5 * (w.a[3 * i + j] - std::conj(w.a[3 * j + i]));
    tr = (ta.a[0] + ta.a[4] + ta.a[8]) / 3.0;
    for (int i = 0; i < 3; ++i) ta.a[4 * i] -= tr;
    for (int e = 0; e < 9; ++e) {
      mom[18 * l + 2 * e] -= eps * ta.a[e].real();
      mom[18 * l + 2 * e + 1] -= eps * ta.a[e].imag();
      if (W_out) {
        W_out[18 * l + 2 * e] = w.a[e].real();
        W_out[18 * l + 2 * e + 1] = w.a[e].imag();
      }
    }
  }
}


/* U <- exp(eps A) U for every link (update_mg_body's gauge update with
 * epsilonU = eps, gauge_force_imp.cpp:185,334; MILC's update_u).  U and mom are
 * raw 72-real/site arrays (mom anti-Hermitian).  The exponential is the plain
 * matrix Taylor series with scaling and squaring (Moler & Van Loan 2003): the
 * argument is halved until its max-row-sum norm is below 1/2, 30 terms are
 * summed and the result squared back. */
static M3 mexp(M3 X) {
  double nrm = 0;
  for (int i = 0; i < 3; ++i) {
    double r = 0;
    for (int j = 0; j < 3; ++j) r += std::abs(X.a[3 * i + j]);
    nrm = r > nrm ? r : nrm;
  }
  int s = 0;
  while (nrm > 0.5) {
    nrm *= 0.5;
    ++s;
  }
  const double sc = std::ldexp(1.0, -s);
  for (int e = 0; e < 9; ++e) X.a[e] *= sc;
  M3 E = meye(), term = meye();
  for (int n = 1; n <= 30; ++n) {
    term = mmul(term, X);
    for (int e = 0; e < 9; ++e) term.a[e] /= (double)n;
    for (int e = 0; e < 9; ++e) E.a[e] += term.a[e];
  }
  for (int k = 0; k < s; ++k) E = mmul(E, E);
  return E;
}

static void gauge_update(const GLat &g, double *U, const double *mom, double eps) {
#pragma omp parallel for schedule(static)
  for (long l = 0; l < 8 * g.vh; ++l) {
    M3 A, u;
    for (int e = 0; e < 9; ++e) {
      A.a[e] = eps * cd(mom[18 * l + 2 * e], mom[18 * l + 2 * e + 1]);
      u.a[e] = cd(U[18 * l + 2 * e], U[18 * l + 2 * e + 1]);
    }
    const M3 r = mmul(mexp(A), u);
    for (int e = 0; e < 9; ++e) {
      U[18 * l + 2 * e] = r.a[e].real();
      U[18 * l + 2 * e + 1] = r.a[e].imag();
    }
  }
}

static GLat glat(const int *latsize) {
  GLat g;
  for (int d = 0; d < 4; ++d) g.L[d] = latsize[d];
  g.vh = (long)latsize[0] * latsize[1] * latsize[2] * latsize[3] / 2;
  return g;
}

void oracle_gauge_update(const int *latsize, double *U, const double *mom, double eps) {
  gauge_update(glat(latsize), U, mom, eps);
}

void oracle_matrix_exp(const double *X, double *out) {
  M3 m;
  for (int e = 0; e < 9; ++e) m.a[e] = cd(X[2 * e], X[2 * e + 1]);
  const M3 r = mexp(m);
  for (int e = 0; e < 9; ++e) {
    out[2 * e] = r.a[e].real();
    out[2 * e + 1] = r.a[e].imag();
  }
}

/* Pure-gauge molecular dynamics, nsteps of size eps: scheme 0 leapfrog (force
 * half step, then gauge and force steps, force half step at the end), scheme 1
 * Omelyan's 2MN (lambda = 0.1931833275037836; Omelyan, Mryglod & Folk 2003),
 * adjacent force steps merged.  Force steps mom <- mom - e TA(W) as
 * oracle_gauge_force, gauge steps U <- exp(e A) U. */
void oracle_gauge_md(const int *latsize, double *U, double *mom, const double *coef, double eps, int nsteps,
                     int scheme) {
  const GLat g = glat(latsize);
  auto force = [&](double e) { oracle_gauge_force(latsize, U, coef, e, mom, nullptr); };
  if (scheme == 0) {
    for (int k = 0; k < nsteps; ++k) {
      force(k == 0 ? 0.5 * eps : eps);
      gauge_update(g, U, mom, eps);
    }
    if (nsteps > 0) force(0.5 * eps);
  } else {
    const double lam = 0.1931833275037836;
    for (int k = 0; k < nsteps; ++k) {
      force(k == 0 ? lam * eps : 2 * lam * eps);
      gauge_update(g, U, mom, 0.5 * eps);
      force((1 - 2 * lam) * eps);
      gauge_update(g, U, mom, 0.5 * eps);
    }
    if (nsteps > 0) force(lam * eps);
  }
}

/* Fermion links from the thin gauge field, path by path (the product is
 * multiplied link by link along each path, no nesting).  Paths for link
 * (x, mu), with signed directions d != +-mu: the link itself (c1); staples
 * [d, mu, -d] (c3); ordered [d, e, mu, -e, -d], e != +-d (c5); ordered
 * [d, e, f, mu, -f, -e, -d], all distinct directions (c7); [d, d, mu, -d, -d]
 * (cL); Naik [mu, mu, mu] (cN).  Coefficients in the physical sign convention
 * of Orginos-Toussaint-Sugar / Lepage (fat7: 1/8, 1/16, 1/64, 1/384).  HISQ
 * (Follana et al. 2007): fat7, U(3) projection W = V (V^dag V)^(-1/2) (here
 * from a Jacobi eigen-decomposition of V^dag V), then {1, 1/16, 1/64, 1/384,
 * cL -1/8} and Naik -(1 + eps)/24 of W.  Phases: MILC's eta (C11) and -1 on
 * t-links whose path crosses the last time slice when apbc. */
static M3 path_product(const double *U, const GLat &g, const int *x, const int *steps, int n) {
  int y[4] = {x[0], x[1], x[2], x[3]};
  M3 P = meye();
  for (int k = 0; k < n; ++k) {
    const int s = steps[k], d = (s > 0 ? s : -s) - 1;
    if (s > 0) {
      P = mmul(P, get_link(U, g, y, d));
      y[d] += 1;
    } else {
      y[d] -= 1;
      P = mmul(P, madj(get_link(U, g, y, d)));
    }
  }
  return P;
}

static void smear_paths(const double *U, const GLat &g, const double *c, double *fat, double *lng) {
  const double c1 = c[0], cn = c[1], c3 = c[2], c5 = c[3], c7 = c[4], cl = c[5];
  const long V = 2 * g.vh;
#pragma omp parallel for schedule(dynamic, 16)
  for (long s = 0; s < V; ++s) {
    // coordinates of raw site s
    const long p = s / g.vh, i = s % g.vh;
    int x[4];
    long r = i / (g.L[0] / 2);
    const int xh = (int)(i % (g.L[0] / 2));
    x[1] = (int)(r % g.L[1]);
    r /= g.L[1];
    x[2] = (int)(r % g.L[2]);
    x[3] = (int)(r / g.L[2]);
    x[0] = 2 * xh + (int)((x[1] + x[2] + x[3] + p) & 1);
    for (int mu = 0; mu < 4; ++mu) {
      const int m = mu + 1;
      M3 acc{};
      auto add = [&](double w, std::initializer_list<int> st) {
        if (w == 0) return;
        int steps[8], n = 0;
        for (int v : st) steps[n++] = v;
        const M3 P = path_product(U, g, x, steps, n);
        for (int e = 0; e < 9; ++e) acc.a[e] += w * P.a[e];
      };
      add(c1, {m});
      for (int d = -4; d <= 4; ++d) {
        if (d == 0 || d == m || d == -m) continue;
        add(c3, {d, m, -d});
        add(cl, {d, d, m, -d, -d});
        for (int e = -4; e <= 4; ++e) {
          if (e == 0 || e == m || e == -m || e == d || e == -d) continue;
          add(c5, {d, e, m, -e, -d});
          for (int f = -4; f <= 4; ++f) {
            if (f == 0 || f == m || f == -m || f == d || f == -d || f == e || f == -e) continue;
            add(c7, {d, e, f, m, -f, -e, -d});
          }
        }
      }
      for (int e = 0; e < 9; ++e) {
        fat[72 * s + 18 * mu + 2 * e] = acc.a[e].real();
        fat[72 * s + 18 * mu + 2 * e + 1] = acc.a[e].imag();
      }
      if (lng) {
        const int st[3] = {m, m, m};
        const M3 N = path_product(U, g, x, st, 3);
        for (int e = 0; e < 9; ++e) {
          lng[72 * s + 18 * mu + 2 * e] = cn * N.a[e].real();
          lng[72 * s + 18 * mu + 2 * e + 1] = cn * N.a[e].imag();
        }
      }
    }
  }
}

/* Eigen-decomposition of a Hermitian 3x3 matrix by cyclic complex Jacobi
 * rotations: H = Q diag(w) Q^dag. */
static void herm_eig(M3 H, double w[3], M3 &Q) {
  Q = meye();
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = i + 1; j < 3; ++j) off += std::norm(H.a[3 * i + j]);
    if (off < 1e-34) break;
    for (int p = 0; p < 3; ++p)
      for (int q = p + 1; q < 3; ++q) {
        const cd hpq = H.a[3 * p + q];
        const double a = std::abs(hpq);
        if (a < 1e-300) continue;
        const cd ph = hpq / a;  // H_pq = a e^{i phi}
        const double hpp = H.a[4 * p].real(), hqq = H.a[4 * q].real();
        const double theta = 0.5 * std::atan2(2 * a, hpp - hqq);
        const double cs = std::cos(theta), sn = std::sin(theta);
        // rotation R with columns p, q: R_pp = c, R_qp = s conj(ph), R_pq = -s ph, R_qq = c
        M3 R = meye();
        R.a[4 * p] = cs;
        R.a[4 * q] = cs;
        R.a[3 * q + p] = sn * std::conj(ph);
        R.a[3 * p + q] = -sn * ph;
        H = mmul(madj(R), mmul(H, R));
        Q = mmul(Q, R);
      }
  }
  for (int k = 0; k < 3; ++k) w[k] = H.a[4 * k].real();
}

static void u3_project(const GLat &g, double *Vf) {
#pragma omp parallel for schedule(static)
  for (long l = 0; l < 8 * g.vh; ++l) {
    M3 V;
    for (int e = 0; e < 9; ++e) V.a[e] = cd(Vf[18 * l + 2 * e], Vf[18 * l + 2 * e + 1]);
    double w[3];
    M3 Q;
    herm_eig(mmul(madj(V), V), w, Q);
    M3 D{};
    for (int k = 0; k < 3; ++k) D.a[4 * k] = 1.0 / std::sqrt(w[k]);
    const M3 Wm = mmul(V, mmul(Q, mmul(D, madj(Q))));
    for (int e = 0; e < 9; ++e) {
      Vf[18 * l + 2 * e] = Wm.a[e].real();
      Vf[18 * l + 2 * e + 1] = Wm.a[e].imag();
    }
  }
}

static void fold_phases(const GLat &g, double *f, int hop, int apbc) {
  for (long s = 0; s < 2 * g.vh; ++s) {
    const long p = s / g.vh, i = s % g.vh;
    int x[4];
    long r = i / (g.L[0] / 2);
    const int xh = (int)(i % (g.L[0] / 2));
    x[1] = (int)(r % g.L[1]);
    r /= g.L[1];
    x[2] = (int)(r % g.L[2]);
    x[3] = (int)(r / g.L[2]);
    x[0] = 2 * xh + (int)((x[1] + x[2] + x[3] + p) & 1);
    const int eta[4] = {(x[3] & 1) ? -1 : 1, ((x[3] + x[0]) & 1) ? -1 : 1, ((x[3] + x[0] + x[1]) & 1) ? -1 : 1, 1};
    for (int mu = 0; mu < 4; ++mu) {
      double sg = eta[mu];
      if (mu == 3 && apbc && x[3] + hop >= g.L[3]) sg = -sg;
      for (int e = 0; e < 18; ++e) f[72 * s + 18 * mu + e] *= sg;
    }
  }
}

/* coef = {one_link, naik, three_staple, five_staple, seven_staple, lepage};
 * hisq != 0 ignores coef and builds HISQ links with Naik factor 1 + naik_eps.
 * fat, lng: raw forward links (72 reals/site, even sites first), phased. */
void oracle_links_from_gauge(const int *latsize, const double *U, const double *coef, int hisq, double naik_eps,
                             int apbc, double *fat, double *lng) {
  const GLat g = glat(latsize);
  if (!hisq) {
    smear_paths(U, g, coef, fat, lng);
  } else {
    std::vector<double> W(72 * 2 * g.vh);
    const double fat7[6] = {1.0 / 8, 0, 1.0 / 16, 1.0 / 64, 1.0 / 384, 0};
    smear_paths(U, g, fat7, W.data(), nullptr);
    u3_project(g, W.data());
    const double lvl2[6] = {1.0, -(1.0 + naik_eps) / 24.0, 1.0 / 16, 1.0 / 64, 1.0 / 384, -1.0 / 8};
    smear_paths(W.data(), g, lvl2, fat, lng);
  }
  fold_phases(g, fat, 1, apbc);
  fold_phases(g, lng, 3, apbc);
}

/* Lattice set-up; the reference's minimum-size checks qphix_misc.cpp:494-497. */
int oracle_setup(const int *latsize) {
  if (latsize[0] % 4 || latsize[1] % 4 || latsize[2] % 4 || latsize[3] % 4) return 1;
  G.nx = latsize[0]; G.ny = latsize[1]; G.nz = latsize[2]; G.nt = latsize[3];
  G.nxh = G.nx / 2;
  G.vh = (long)G.nxh * G.ny * G.nz * G.nt;
  build_neighbours();
  return 0;
}

/* Neighbour table of one parity, 16 entries per site (test access). */
void oracle_neighbours(int cb, unsigned *out) {
  std::memcpy(out, G.neigh[cb].data(), sizeof(unsigned) * 16 * G.vh);
}

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void *oracle_links_create_d(const double *f, const double *l, const double *fb, const double *lb) {
  return links_create<double>(f, l, fb, lb);
}
void *oracle_links_create_f(const float *f, const float *l, const float *fb, const float *lb) {
  return links_create<float>(f, l, fb, lb);
}
void oracle_links_destroy_d(void *L) { delete (Links<double> *)L; }
void oracle_links_destroy_f(void *L) { delete (Links<float> *)L; }

/* cb = output parity (0 even, 1 odd); half-lattice vectors */
void oracle_dslash_d(void *L, double *out, const double *in, int cb) {
  dslash((Links<double> *)L, out, in, cb);
}
void oracle_dslash_f(void *L, float *out, const float *in, int cb) {
  dslash((Links<float> *)L, out, in, cb);
}

double oracle_relative_residue_d(const double *r, const double *x) { return relative_residue(r, x); }

int oracle_invert_fewsums_d(void *L, int parity, int niter, int nrestart, double resid, double mass, double *x,
                            const double *b, double *out_d, int *out_i) {
  InvertOut o{};
  invert((Links<double> *)L, parity, niter, nrestart, resid, 0.0, mass, x, b, &o, 1);
  out_d[0] = o.final_rsq; out_d[1] = o.final_rel; out_d[2] = o.size_r; out_d[3] = o.size_relr;
  out_d[4] = o.final_flop;
  out_i[0] = o.final_iter; out_i[1] = o.final_restart;
  return o.ret;
}
int oracle_invert_fewsums_f(void *L, int parity, int niter, int nrestart, double resid, float mass, float *x,
                            const float *b, double *out_d, int *out_i) {
  InvertOut o{};
  invert((Links<float> *)L, parity, niter, nrestart, resid, 0.0, mass, x, b, &o, 1);
  out_d[0] = o.final_rsq; out_d[1] = o.final_rel; out_d[2] = o.size_r; out_d[3] = o.size_relr;
  out_d[4] = o.final_flop;
  out_i[0] = o.final_iter; out_i[1] = o.final_restart;
  return o.ret;
}
double oracle_relative_residue_f(const float *r, const float *x) { return relative_residue(r, x); }

int oracle_invert_d(void *L, int parity, int niter, int nrestart, double resid, double relresid,
                    double mass, double *x, const double *b, double *out_d, int *out_i) {
  InvertOut o{};
  invert((Links<double> *)L, parity, niter, nrestart, resid, relresid, mass, x, b, &o);
  out_d[0] = o.final_rsq; out_d[1] = o.final_rel; out_d[2] = o.size_r; out_d[3] = o.size_relr;
  out_d[4] = o.final_flop;
  out_i[0] = o.final_iter; out_i[1] = o.final_restart;
  return o.ret;
}
int oracle_invert_f(void *L, int parity, int niter, int nrestart, double resid, double relresid,
                    float mass, float *x, const float *b, double *out_d, int *out_i) {
  InvertOut o{};
  invert((Links<float> *)L, parity, niter, nrestart, resid, relresid, mass, x, b, &o);
  out_d[0] = o.final_rsq; out_d[1] = o.final_rel; out_d[2] = o.size_r; out_d[3] = o.size_relr;
  out_d[4] = o.final_flop;
  out_i[0] = o.final_iter; out_i[1] = o.final_restart;
  return o.ret;
}

int oracle_invert_multi_d(void *L, int parity, int niter, int nrestart, double resid0,
                          const double *mass, int nmass, double **x, const double *b,
                          double *out_d, int *out_i) {
  InvertOut o{};
  invert_multi((Links<double> *)L, parity, niter, nrestart, resid0, mass, nmass, x, b, &o);
  out_d[0] = o.final_rsq; out_d[1] = o.final_flop;
  out_i[0] = o.final_iter;
  return o.ret;
}
int oracle_invert_multi_f(void *L, int parity, int niter, int nrestart, double resid0,
                          const float *mass, int nmass, float **x, const float *b,
                          double *out_d, int *out_i) {
  InvertOut o{};
  invert_multi((Links<float> *)L, parity, niter, nrestart, resid0, mass, nmass, x, b, &o);
  out_d[0] = o.final_rsq; out_d[1] = o.final_flop;
  out_i[0] = o.final_iter;
  return o.ret;
}

void oracle_fermion_oprod_d(void *L, int nt, const double *c, const double *const *X, double *o1,
                            double *o3) {
  fermion_oprod((const Links<double> *)L, nt, c, X, o1, o3);
}

} /* extern "C" */
