// MFMA complex contraction (BF16 / FP16 split operands, FP32 accumulate).
// Placeholder dispatch: precisions without an MFMA kernel fall back to the
// FP32 VALU contraction.

#include "compas/device/common.h"
#include "compas/internal.h"

namespace compas {

void complex_contraction_mfma(
    const CompasContext& ctx,
    const cfloat* A,
    const cfloat* B,
    cfloat* out,
    ContractionShape s,
    ContractionPrecision precision,
    bool* handled) {
    (void)ctx;
    (void)A;
    (void)B;
    (void)out;
    (void)s;
    (void)precision;
    *handled = false;
}

}  // namespace compas
