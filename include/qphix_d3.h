/* qphix_d3.h -- Double-precision declarations of the MILC-facing interface.
 *
 * Replaces the reference's qphix_d3.h (qphix_d3.h:12-100).  Every entry
 * point below is implemented by libqphixmilc_hip.so on one MI355X (gfx950):
 * handles own device memory, raw arrays stay caller-owned host memory in
 * MILC's even-sites-first order, and every call has finished on the GPU when
 * it returns.
 */
#ifndef _QPHIX_D3
#define _QPHIX_D3

#ifdef __cplusplus
extern "C" {
#endif

typedef struct QPHIX_D3_ColorVector_struct QPHIX_D3_ColorVector;
typedef struct QPHIX_D3_FermionLinksAsqtad_struct QPHIX_D3_FermionLinksAsqtad;
typedef struct QPHIX_D3_GaugeField_struct QPHIX_D3_GaugeField;
typedef struct QPHIX_D3_Force_struct QPHIX_D3_Force;

/* Colour vectors: 6 reals per site, [colour][re,im]; odd block at 6*even_sites.
 * qphix_d3.h:18,21 (reference definitions qphix_d3.cpp:79-174, 490-559) */
QPHIX_D3_ColorVector *QPHIX_D3_create_V_from_raw(QPHIX_D_Real *src, QPHIX_evenodd_t evenodd);
void QPHIX_D3_extract_V_to_raw(QPHIX_D_Real *dest, QPHIX_D3_ColorVector *src, QPHIX_evenodd_t evenodd);
void QPHIX_D3_destroy_V(QPHIX_D3_ColorVector *V);

/* Asqtad/HISQ links: 72 reals per site, [dir X,Y,Z,T][row][col][re,im].
 * fat/lng are the forward links U_mu(x); fatback/lngback are the links used
 * for the -mu / -3mu hops at the same site, applied as given (MILC's
 * double-store convention passes -U_mu(x-mu)^dagger).
 * qphix_d3.h (create/extract L); reference qphix_d3.cpp:970-1206 */
QPHIX_D3_FermionLinksAsqtad *QPHIX_D3_asqtad_create_L_from_raw(QPHIX_D_Real *fat, QPHIX_D_Real *lng,
    QPHIX_D_Real *fatback, QPHIX_D_Real *lngback, QPHIX_evenodd_t evenodd);
void QPHIX_D3_asqtad_extract_L_to_raw(QPHIX_D_Real *fat, QPHIX_D_Real *lng,
    QPHIX_D3_FermionLinksAsqtad *src, QPHIX_evenodd_t evenodd);
void QPHIX_D3_asqtad_destroy_L(QPHIX_D3_FermionLinksAsqtad *L);

/* Momentum (anti-Hermitian, 72 reals per site) and gauge field (72 reals per
 * site).  Reference qphix_d3.cpp:177-488, 562-949, 1208-1245 */
QPHIX_D3_Force *QPHIX_D3_create_F_from_raw(QPHIX_D_Real *fwdrawmom, QPHIX_evenodd_t evenodd);
void QPHIX_D3_extract_F_to_raw(QPHIX_D_Real *rawdest, QPHIX_D3_Force *src, QPHIX_evenodd_t evenodd);
void QPHIX_D3_destroy_F(QPHIX_D3_Force *field);
QPHIX_D3_GaugeField *QPHIX_D3_create_G_from_raw(QPHIX_D_Real *fwdrawgauge, QPHIX_evenodd_t evenodd);
void QPHIX_D3_extract_G_to_raw(QPHIX_D_Real *rawdest, QPHIX_D3_GaugeField *src, QPHIX_evenodd_t evenodd);
void QPHIX_D3_destroy_G(QPHIX_D3_GaugeField *field);

/* Symanzik one-loop gauge force: mom <- mom + eps * F(U), loops weighted by
 * coeffs->plaquette/rectangle/parallelogram.  Reference gauge_force_imp.cpp:136-199 */
void QPHIX_D3_symanzik_1loop_gauge_force(QPHIX_info_t *info, QPHIX_D3_GaugeField *gauge,
    QPHIX_D3_Force *force, QPHIX_gauge_coeffs_t *coeffs, QPHIX_D_Real eps);

/* out(x) = sum of the 16 link*neighbour terms for every site x of `parity`
 * (not EVENODD); out must have the opposite parity to in.
 * Reference qphix_d3.cpp:1247-1259 -> ks_long_dslash.cpp:564-668 */
void QPHIX_D3_asqtad_dslash(QPHIX_D3_FermionLinksAsqtad *asqtad, QPHIX_D3_ColorVector *out,
    QPHIX_D3_ColorVector *in, QPHIX_evenodd_t parity);

/* CG on (4 m^2 - D_pq D_qp) out = in.  Returns the iterations of the last
 * parity solved; res_arg->final_iter holds the total.
 * Reference ks_d_congrad_fn.cpp:221-619 */
int QPHIX_D3_asqtad_invert(QPHIX_info_t *info, QPHIX_D3_FermionLinksAsqtad *asqtad,
    QPHIX_invert_arg_t *inv_arg, QPHIX_resid_arg_t *res_arg, QPHIX_D_Real mass,
    QPHIX_D3_ColorVector *out, QPHIX_D3_ColorVector *in);

/* Multi-shift CG: (4 mass[j]^2 - D D) out[j] = in for every j; stops on
 * res_arg[0]->resid.  Reference ks_multicg_offset.cpp:205-689 */
int QPHIX_D3_asqtad_invert_multi(QPHIX_info_t *info, QPHIX_D3_FermionLinksAsqtad *asqtad,
    QPHIX_invert_arg_t *inv_arg, QPHIX_resid_arg_t *res_arg[], QPHIX_D_Real *mass,
    int num_offsets, QPHIX_D3_ColorVector *out[], QPHIX_D3_ColorVector *in);

#if QPHIX_Precision == 'D'
#include <qphix_d3_generic.h>
#endif

#ifdef __cplusplus
}
#endif

#endif /* _QPHIX_D3 */
