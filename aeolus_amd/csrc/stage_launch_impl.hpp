// Definition of launch_stage_ns and the list of stage-kernel combinations the
// library uses; included only by the per-precision instance files.
#pragma once

#include "stage_launch.hpp"

namespace aeo {

template <class KS, class Epi, int MB, bool ANA, bool SYN, int NS>
static int launch_stage_mb(const aeo_plan_s* p, const StageArgs& a, const Epi& ep,
                           hipStream_t s) {
    const size_t lds = StageLds<KS, Epi, MB, NS>::bytes(p->L, p->nh, ANA, SYN);
    if (lds > 163840) armed_events() = ArmedEvents{};
    AEO_REQUIRE(lds <= 163840, AEO_ERR_UNSUPPORTED,
                "Legendre stage kernel: LDS demand exceeds 160 KiB (reduce members per block)");
    auto kern = leg_stage_kernel<KS, Epi, MB, ANA, SYN, NS>;
    if (lds > 65536) {
        AEO_CHECK_HIP(hipFuncSetAttribute((const void*)kern,
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    dim3 grid(a.ys_mmajor ? p->L + 1 : stage_grid_x(p->L + 1), 1, (a.nmem + MB - 1) / MB);
    launch_k(kern, grid, dim3(kStageThreads), lds, s, p->gdev(), a, ep);
    AEO_LAUNCH_CHECK();
    return AEO_OK;
}

template <class KS, class Epi, bool ANA, bool SYN, int NS>
int launch_stage_ns(const aeo_plan_s* p, int mb, const StageArgs& a, const Epi& ep,
                    hipStream_t s) {
    switch (mb) {
        case 1: return launch_stage_mb<KS, Epi, 1, ANA, SYN, NS>(p, a, ep, s);
        case 2: return launch_stage_mb<KS, Epi, 2, ANA, SYN, NS>(p, a, ep, s);
        case 4: return launch_stage_mb<KS, Epi, 4, ANA, SYN, NS>(p, a, ep, s);
        default: return launch_stage_mb<KS, Epi, 8, ANA, SYN, NS>(p, a, ep, s);
    }
}

#define AEO_STAGE_INSTANCE(KS, Epi, ANA, SYN, NS)                                             \
    template int launch_stage_ns<KS, Epi, ANA, SYN, NS>(const aeo_plan_s*, int,               \
                                                        const StageArgs&, const Epi&, hipStream_t);

// Every (synthesis kind, epilogue, analysis, synthesis) combination the
// launchers in gauss.hip use, at precision NS.
#define AEO_STAGE_INSTANCES(NS)                                         \
    AEO_STAGE_INSTANCE(KindRaw, EpiNoneS, false, true, NS)             \
    AEO_STAGE_INSTANCE(KindNone, EpiStore1S, true, false, NS)          \
    AEO_STAGE_INSTANCE(KindJac, EpiNoneS, false, true, NS)             \
    AEO_STAGE_INSTANCE(KindBve, EpiNoneS, false, true, NS)             \
    AEO_STAGE_INSTANCE(KindSwe, EpiNoneS, false, true, NS)             \
    AEO_STAGE_INSTANCE(KindBve, EpiBveS, true, true, NS)               \
    AEO_STAGE_INSTANCE(KindBve, EpiBveS, true, false, NS)              \
    AEO_STAGE_INSTANCE(KindSwe, EpiSweS, true, true, NS)               \
    AEO_STAGE_INSTANCE(KindSwe, EpiSweS, true, false, NS)              \
    AEO_STAGE_INSTANCE(KindBveAdj, EpiNoneS, false, true, NS)          \
    AEO_STAGE_INSTANCE(KindBveAdj, EpiBveAdjS, true, true, NS)         \
    AEO_STAGE_INSTANCE(KindBveAdj, EpiBveAdjS, true, false, NS)

}  // namespace aeo
