// FP32 instantiation of the advection passes.  Built with -ffp-contract=fast
// (multiply-adds fused), since this form is not matched bit for bit.
#include "advection_impl.h"

int adv_dispatch_f32(const Geo &g, const adv::Prof &pr, const adv::Fields &fl, int nsc, bool mom, int so, int mo,
                     hipStream_t st) {
    return adv::dispatch<adv::PolicyF32>(g, pr, fl, nsc, mom, so, mo, st);
}
