/* qphix_f3.h -- Single-precision declarations of the MILC-facing interface.
 *
 * Replaces the reference's qphix_f3.h (qphix_f3.h:12-100).  Every entry
 * point below is implemented by libqphixmilc_hip.so on one MI355X (gfx950):
 * handles own device memory, raw arrays stay caller-owned host memory in
 * MILC's even-sites-first order, and every call has finished on the GPU when
 * it returns.
 */
#ifndef _QPHIX_F3
#define _QPHIX_F3

#ifdef __cplusplus
extern "C" {
#endif

typedef struct QPHIX_F3_ColorVector_struct QPHIX_F3_ColorVector;
typedef struct QPHIX_F3_FermionLinksAsqtad_struct QPHIX_F3_FermionLinksAsqtad;
typedef struct QPHIX_F3_GaugeField_struct QPHIX_F3_GaugeField;
typedef struct QPHIX_F3_Force_struct QPHIX_F3_Force;

/* Colour vectors: 6 reals per site, [colour][re,im]; odd block at 6*even_sites.
 * qphix_f3.h:18,21 (reference definitions qphix_f3.cpp:79-174, 490-559) */
QPHIX_F3_ColorVector *QPHIX_F3_create_V_from_raw(QPHIX_F_Real *src, QPHIX_evenodd_t evenodd);
void QPHIX_F3_extract_V_to_raw(QPHIX_F_Real *dest, QPHIX_F3_ColorVector *src, QPHIX_evenodd_t evenodd);
void QPHIX_F3_destroy_V(QPHIX_F3_ColorVector *V);

/* Asqtad/HISQ links: 72 reals per site, [dir X,Y,Z,T][row][col][re,im].
 * fat/lng are the forward links U_mu(x); fatback/lngback are the links used
 * for the -mu / -3mu hops at the same site, applied as given (MILC's
 * double-store convention passes -U_mu(x-mu)^dagger).
 * qphix_f3.h (create/extract L); reference qphix_f3.cpp:970-1206 */
QPHIX_F3_FermionLinksAsqtad *QPHIX_F3_asqtad_create_L_from_raw(QPHIX_F_Real *fat, QPHIX_F_Real *lng,
    QPHIX_F_Real *fatback, QPHIX_F_Real *lngback, QPHIX_evenodd_t evenodd);
void QPHIX_F3_asqtad_extract_L_to_raw(QPHIX_F_Real *fat, QPHIX_F_Real *lng,
    QPHIX_F3_FermionLinksAsqtad *src, QPHIX_evenodd_t evenodd);
void QPHIX_F3_asqtad_destroy_L(QPHIX_F3_FermionLinksAsqtad *L);

/* Momentum (anti-Hermitian, 72 reals per site) and gauge field (72 reals per
 * site).  Reference qphix_f3.cpp:177-488, 562-949, 1208-1245 */
QPHIX_F3_Force *QPHIX_F3_create_F_from_raw(QPHIX_F_Real *fwdrawmom, QPHIX_evenodd_t evenodd);
void QPHIX_F3_extract_F_to_raw(QPHIX_F_Real *rawdest, QPHIX_F3_Force *src, QPHIX_evenodd_t evenodd);
void QPHIX_F3_destroy_F(QPHIX_F3_Force *field);
QPHIX_F3_GaugeField *QPHIX_F3_create_G_from_raw(QPHIX_F_Real *fwdrawgauge, QPHIX_evenodd_t evenodd);
void QPHIX_F3_extract_G_to_raw(QPHIX_F_Real *rawdest, QPHIX_F3_GaugeField *src, QPHIX_evenodd_t evenodd);
void QPHIX_F3_destroy_G(QPHIX_F3_GaugeField *field);

/* Symanzik one-loop gauge force: mom <- mom + eps * F(U), loops weighted by
 * coeffs->plaquette/rectangle/parallelogram.  Reference gauge_force_imp.cpp:136-199 */
void QPHIX_F3_symanzik_1loop_gauge_force(QPHIX_info_t *info, QPHIX_F3_GaugeField *gauge,
    QPHIX_F3_Force *force, QPHIX_gauge_coeffs_t *coeffs, QPHIX_F_Real eps);

/* out(x) = sum of the 16 link*neighbour terms for every site x of `parity`
 * (not EVENODD); out must have the opposite parity to in.
 * Reference qphix_f3.cpp:1247-1259 -> ks_long_dslash.cpp:564-668 */
void QPHIX_F3_asqtad_dslash(QPHIX_F3_FermionLinksAsqtad *asqtad, QPHIX_F3_ColorVector *out,
    QPHIX_F3_ColorVector *in, QPHIX_evenodd_t parity);

/* CG on (4 m^2 - D_pq D_qp) out = in.  Returns the iterations of the last
 * parity solved; res_arg->final_iter holds the total.
 * Reference ks_d_congrad_fn.cpp:221-619 */
int QPHIX_F3_asqtad_invert(QPHIX_info_t *info, QPHIX_F3_FermionLinksAsqtad *asqtad,
    QPHIX_invert_arg_t *inv_arg, QPHIX_resid_arg_t *res_arg, QPHIX_F_Real mass,
    QPHIX_F3_ColorVector *out, QPHIX_F3_ColorVector *in);

/* Multi-shift CG: (4 mass[j]^2 - D D) out[j] = in for every j; stops on
 * res_arg[0]->resid.  Reference ks_multicg_offset.cpp:205-689 */
int QPHIX_F3_asqtad_invert_multi(QPHIX_info_t *info, QPHIX_F3_FermionLinksAsqtad *asqtad,
    QPHIX_invert_arg_t *inv_arg, QPHIX_resid_arg_t *res_arg[], QPHIX_F_Real *mass,
    int num_offsets, QPHIX_F3_ColorVector *out[], QPHIX_F3_ColorVector *in);

#if QPHIX_Precision == 'F'
#include <qphix_f3_generic.h>
#endif

#ifdef __cplusplus
}
#endif

#endif /* _QPHIX_F3 */
