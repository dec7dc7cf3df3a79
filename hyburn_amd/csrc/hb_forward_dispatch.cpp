// hb_forward_dispatch.cpp -- routes a forward launch to the translation unit
// holding its kernels (hb_forward.hip is compiled once per unit and mode).
#include "hb_internal.h"

namespace hb {

#define HB_DECL_UNIT(n)                                                                          \
    hipError_t fwd_unit_fast_##n(int op, const LaunchDesc* d, hipStream_t s, int arg, int* out); \
    hipError_t fwd_unit_exact_##n(int op, const LaunchDesc* d, hipStream_t s, int arg, int* out);
HB_DECL_UNIT(0) HB_DECL_UNIT(1) HB_DECL_UNIT(2) HB_DECL_UNIT(3) HB_DECL_UNIT(4)
HB_DECL_UNIT(5) HB_DECL_UNIT(6)
#undef HB_DECL_UNIT

namespace {

using UnitFn = hipError_t (*)(int, const LaunchDesc*, hipStream_t, int, int*);
constexpr int kUnits = 7;
const UnitFn kFast[kUnits] = {fwd_unit_fast_0, fwd_unit_fast_1, fwd_unit_fast_2, fwd_unit_fast_3, fwd_unit_fast_4,
                              fwd_unit_fast_5, fwd_unit_fast_6};
const UnitFn kExact[kUnits] = {fwd_unit_exact_0, fwd_unit_exact_1, fwd_unit_exact_2, fwd_unit_exact_3,
                               fwd_unit_exact_4, fwd_unit_exact_5, fwd_unit_exact_6};

// Unit of a launch: one per instantiated (model, pre channels) pair. -1: no
// kernel for this combination (every launch group holds a single model).
int unit_of(const LaunchDesc& d) {
    if (d.model < 0) return -1;
    switch (d.model * 4 + d.np) {
    case HB_G2DO * 4 + 1: return 0;
    case HB_MPR * 4 + 2: return 1;
    case HB_MPR * 4 + 1: return 1;
    case HB_JANSEN_RIT * 4 + 1: return 2;
    case HB_WILSON_COWAN * 4 + 1: return 3;
    case HB_KURAMOTO * 4 + 1: return 4;
    case HB_KURAMOTO * 4 + 2: return 5;
    case HB_RWW * 4 + 1: return 6;
    default: return -1;
    }
}

hipError_t run(const UnitFn* t, int op, const LaunchDesc& d, hipStream_t s, int* out) {
    const int u = unit_of(d);
    if (u < 0) return hipErrorInvalidValue;
    return t[u](op, &d, s, 0, out);
}

hipError_t attrs(const UnitFn* t, int bytes) {
    for (int u = 0; u < kUnits; ++u) {
        hipError_t e = t[u](2, nullptr, nullptr, bytes, nullptr);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

hipError_t launch_forward_fast(const LaunchDesc& d, hipStream_t s) { return run(kFast, 0, d, s, nullptr); }
hipError_t launch_forward_exact(const LaunchDesc& d, hipStream_t s) { return run(kExact, 0, d, s, nullptr); }
hipError_t forward_occupancy_fast(const LaunchDesc& d, int* blocks) { return run(kFast, 1, d, nullptr, blocks); }
hipError_t forward_set_attributes_fast(int max_lds) { return attrs(kFast, max_lds); }
hipError_t forward_set_attributes_exact(int max_lds) { return attrs(kExact, max_lds); }

}  // namespace hb
