// FP64 instantiation of the advection passes.  Built with -ffp-contract=off:
// no multiply-add is fused, so the arithmetic matches the reference's
// (-std=c99 builds do not contract) to the last bit.
#include "advection_impl.h"

int adv_dispatch_f64(const Geo &g, const adv::Prof &pr, const adv::Fields &fl, int nsc, bool mom, int so, int mo,
                     hipStream_t st) {
    return adv::dispatch<adv::PolicyF64>(g, pr, fl, nsc, mom, so, mo, st);
}
