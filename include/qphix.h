/* qphix.h -- umbrella header of the MILC-facing C interface.
 *
 * Replaces the reference's qphix.h (qphix.h:1-64).  A translation unit picks
 * the precision of the generic names (QPHIX_asqtad_invert, ...) by defining
 * QPHIX_Precision ('F' or 'D') or QPHIX_PrecisionInt (1 or 2) before the
 * include; single precision is the default.  Both precision-specific name
 * sets (QPHIX_F3_*, QPHIX_D3_*) are always declared.
 */
#ifndef _QPHIX_H
#define _QPHIX_H

#if !defined(QPHIX_Precision) && !defined(QPHIX_PrecisionInt)
#define QPHIX_PrecisionInt 1
#endif

#if defined(QPHIX_PrecisionInt) && !defined(QPHIX_Precision)
#if QPHIX_PrecisionInt == 1
#define QPHIX_Precision 'F'
#elif QPHIX_PrecisionInt == 2
#define QPHIX_Precision 'D'
#else
#error "QPHIX_PrecisionInt must be 1 (single) or 2 (double)"
#endif
#elif defined(QPHIX_Precision) && !defined(QPHIX_PrecisionInt)
#if QPHIX_Precision == 'F'
#define QPHIX_PrecisionInt 1
#elif QPHIX_Precision == 'D'
#define QPHIX_PrecisionInt 2
#else
#error "QPHIX_Precision must be 'F' or 'D'"
#endif
#else
#if (QPHIX_Precision == 'F' && QPHIX_PrecisionInt != 1) || \
    (QPHIX_Precision == 'D' && QPHIX_PrecisionInt != 2)
#error "QPHIX_Precision and QPHIX_PrecisionInt disagree"
#endif
#endif

#if QPHIX_PrecisionInt == 1
#define QPHIX_PrecisionLetter F
#else
#define QPHIX_PrecisionLetter D
#endif

#include <qphix_int.h>
#include <qphix_f3.h>
#include <qphix_d3.h>

#endif /* _QPHIX_H */
