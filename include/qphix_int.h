/* qphix_int.h -- precision-independent types and init/finalize.
 *
 * Replaces the reference's qphix_int.h (qphix_int.h:13-96).  The structs below
 * must stay layout-identical to the reference's so that MILC objects compiled
 * against either header interoperate.
 */
#ifndef _QPHIX_INT_H
#define _QPHIX_INT_H

#include <sys/types.h>

#ifdef __cplusplus
extern "C" {
#endif

/* qphix_int.h:13-17 */
typedef enum {
  QPHIX_SUCCESS = 0,
  QPHIX_FAIL = 1,
  QPHIX_MEM_ERROR
} QPHIX_status_t;

/* qphix_int.h:19-23.  Raw arrays keep all even sites first, then all odd. */
typedef enum {
  QPHIX_EVEN = 0x02,
  QPHIX_ODD = 0x01,
  QPHIX_EVENODD = 0x03
} QPHIX_evenodd_t;

typedef float QPHIX_F_Real;
typedef double QPHIX_D_Real;

#define QPHIX_EXTENDED_LAYOUT

/* qphix_int.h:30-42.  Only a single-node layout (machsize all 1) is accepted. */
typedef struct {
  int (*node_number)(const int coords[]);
  size_t (*node_index)(const int coords[]);
  int latdim;
  int *latsize;
  int machdim;
  int *machsize;
  int this_node;
  size_t sites_on_node;
  size_t even_sites_on_node;
  void *mpi_comm;
} QPHIX_layout_t;
#define QPHIX_LAYOUT_ZERO ((QPHIX_layout_t){NULL,NULL,0,NULL,0,NULL,0,0,0,NULL})

/* qphix_int.h:44-50 */
typedef struct {
  double final_sec;       /* (out) wall time of the call in seconds */
  double final_flop;      /* (out) flop count by the reference's model */
  QPHIX_status_t status;  /* (out) */
  int count1, count2;     /* (out) generic counters */
} QPHIX_info_t;
#define QPHIX_INFO_ZERO ((QPHIX_info_t){0,0,QPHIX_SUCCESS,0,0})

/* qphix_int.h:52-58 */
typedef struct {
  double plaquette;
  double rectangle;
  double parallelogram;
  double adjoint_plaquette;   /* accepted and ignored, as in the reference */
} QPHIX_gauge_coeffs_t;
#define QPHIX_GAUGE_COEFFS_ZERO ((QPHIX_gauge_coeffs_t){0.,0.,0.,0.})

/* qphix_int.h:60-69.  The default initialiser keeps the reference's order
 * (it does not match the member order; preserved on purpose). */
typedef struct {
  QPHIX_evenodd_t parity;
  int max;                 /* (in) iterations between true-residual restarts */
  int restart;             /* (in) unused by the solvers, as in the reference */
  int nrestart;            /* (in) restarts allowed; max*nrestart caps the total */
  QPHIX_evenodd_t evenodd;
} QPHIX_invert_arg_t;
#define QPHIX_INVERT_ARG_DEFAULT ((QPHIX_invert_arg_t){2000,1000,5,QPHIX_EVENODD})

/* qphix_int.h:71-85.  resid is a squared relative residual |r|^2/|b|^2. */
typedef struct {
  double resid;
  double final_rsq;
  double relresid;
  double final_rel;
  double size_r;
  double size_relr;
  int final_iter;
  int final_restart;
} QPHIX_resid_arg_t;
#define QPHIX_RESID_ARG_DEFAULT ((QPHIX_resid_arg_t){1e-6,0,0,0,0,0})

/* qphix_int.h:91-96 (definitions: qphix_misc.cpp:427-594) */
QPHIX_status_t QPHIX_init_F(QPHIX_layout_t *layout);
QPHIX_status_t QPHIX_init_D(QPHIX_layout_t *layout);
QPHIX_status_t QPHIX_init(QPHIX_layout_t *layout);
QPHIX_status_t QPHIX_finalize();
QPHIX_status_t QPHIX_finalize_F();
QPHIX_status_t QPHIX_finalize_D();

#ifdef __cplusplus
}
#endif

#define QPHIX_OPP_PAR(p) ((p)==QPHIX_EVENODD ? (p) : (p)^11)
#define QPHIX_OPP_DIR(d) ((d)&1 ? (d)-1 : (d)+1)

#endif /* _QPHIX_INT_H */
