// Launchers of the Legendre stage kernel (legendre.hpp).
//
// launch_stage picks the precision (FP64 or k digit slices) and the members
// per workgroup at run time.  The kernel instances themselves live in
// stage_fp64.hip and stage_split{4,5,6,7}.hip (explicit instantiations of
// launch_stage_ns, one translation unit per precision, compiled in
// parallel); callers see only the declaration below.
#pragma once

#include "legendre.hpp"

namespace aeo {

inline ColK col_consts(const aeo_plan_s* p, double R, const double* lap, double f10) {
    return ColK{p->L, R, p->d_CpT, p->d_CmT, lap, f10};
}

// One launch of leg_stage_kernel<KS, Epi, mb, ANA, SYN, NS> over every m and
// member group (mb in {1, 2, 4, 8}).
template <class KS, class Epi, bool ANA, bool SYN, int NS>
int launch_stage_ns(const aeo_plan_s* p, int mb, const StageArgs& a, const Epi& ep,
                    hipStream_t s);

// ns = 0: FP64 contractions; ns in {4, 5, 6, 7}: split digits (plan tables must match).
template <class KS, class Epi, bool ANA, bool SYN>
inline int launch_stage(const aeo_plan_s* p, int mb, const StageArgs& a, const Epi& ep,
                        hipStream_t s, int ns = 0) {
    if (ns != 0) {
        AEO_REQUIRE(p->nslices == ns, AEO_ERR_STATE,
                    "split-precision stage: the plan's digit tables have a different slice count");
    }
    switch (ns) {
        case 0: return launch_stage_ns<KS, Epi, ANA, SYN, 0>(p, mb, a, ep, s);
        case 4: return launch_stage_ns<KS, Epi, ANA, SYN, 4>(p, mb, a, ep, s);
        case 5: return launch_stage_ns<KS, Epi, ANA, SYN, 5>(p, mb, a, ep, s);
        case 6: return launch_stage_ns<KS, Epi, ANA, SYN, 6>(p, mb, a, ep, s);
        case 7: return launch_stage_ns<KS, Epi, ANA, SYN, 7>(p, mb, a, ep, s);
        default:
            set_error("split-precision stage: nslices must be 4, 5, 6 or 7");
            return AEO_ERR_UNSUPPORTED;
    }
}

}  // namespace aeo
