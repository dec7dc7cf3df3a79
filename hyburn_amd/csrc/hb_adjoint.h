// hb_adjoint.h -- descriptors shared by the adjoint planner and its kernels.
#pragma once

#include "hb_internal.h"

namespace hb {

// One instance of an adjoint plan. Device pointers into the plan's arena.
struct alignas(16) AdjInst {
    int32_t model, integrator, n_nodes, n_steps;
    int32_t np, cfn, engine, has_coupling;
    int32_t seg_len, n_seg, ring_depth, aring_depth;   // S, ceil(T/S), forward and adjoint ring depths
    int32_t ring_lds, aring_lds, bold_period, n_win;
    int32_t loss_kind, bold_fused, rc_fused, pad2;   // bold_fused: BOLD forward in helper waves of the blocked
                                                     // forward; rc_fused: segment recompute in helper waves of
                                                     // the blocked reverse (double tape)
    float dt, dt_half, dt_sixth, pad3;
    float mc[kMC];
    float cc[kCC];
    float nscale[8];
    double dt_ms;
    const float* x0;            // [N][NV]
    const float* node_mc;       // [N][kMC] or null (uniform mc)
    const float* node_jac;      // [N][kMC] d mc / d p[grad_param], or null
    const int32_t* indptr;      // CSR by target
    const uint2* edges;         // .x = src | delay << 16, .y = weight bits
    const int32_t* indptr_t;    // CSC by source
    const uint2* edges_t;       // .x = tgt | delay << 16, .y = weight bits
    const float* noise;         // raw N(0,1) [T][N][NV] or null
    float* ring;                // HBM ring (forward, then adjoint) when not in LDS
    float* glog;                // [T][N][np] gathered coupling sums
    float* ckpt;                // [n_seg][N][NV]
    float* tape;                // [S + 1][N][NV], two of them with rc_fused (segments by parity)
    float* zlog;                // [n_win][N] BOLD window inputs
    double* bstate;             // [n_win][N][4] Balloon-Windkessel states (FP64)
    double* bold;               // [n_win][N] native BOLD samples (FP64)
    double* dsig;               // [n_win][N] dL/dBOLD
    double* bscan;              // [16][N][20] chunk maps of the BOLD adjoint
    float* seedw;               // [n_win][N] dL/d(window input) / count
    const float* target;        // FC [N][N], BOLD [n_win][N] or order parameter [T]
    double* fc_work;            // [3N + 2N^2]
    float* ord_work;            // [S + 1][4]
    float* fin;                 // [N][NV] final state
    float* out_loss;            // [1]
    float* out_gcc;             // [1] dL / d cc[0]
    float* out_gp;              // [N] or null
    float* out_gx0;             // [N][NV] or null
};

// One launch group: instances of one model / coupling kind / nodes-per-thread.
struct AdjLaunch {
    const AdjInst* insts;       // device array (contiguous for the group)
    int32_t n;
    int32_t model;
    int32_t np;                 // pre channels stored per node
    int32_t npt;                // nodes per thread (1 or 2)
    bool blocked;               // delay-blocked kernels (min delay >= 8, npt 1)
    bool lring;                 // blocked kernels: rings in LDS
    bool bold_fused;            // blocked forward carries the BOLD monitor in helper waves (launched 2 x block)
    bool rc_fused;              // blocked reverse recomputes segments in helper waves (launched 2 x block)
    int32_t ic;                 // integrator class of every instance (0 Euler, 1 Heun, 2 RK4)
    int32_t cfk;                // HB_CF_LINEAR when every instance couples linearly (blocked kernels), else -1
    int32_t block;
    int32_t lds_fwd, lds_rev;   // dynamic LDS bytes
};

hipError_t adjoint_set_attributes(int max_lds);
// per-model kernel units (hb_adjoint.hip built with HB_ADJ_UNIT = model)
hipError_t adj_unit_0(int op, const AdjLaunch* L, hipStream_t s, int arg);
hipError_t adj_unit_1(int op, const AdjLaunch* L, hipStream_t s, int arg);
hipError_t adj_unit_2(int op, const AdjLaunch* L, hipStream_t s, int arg);
hipError_t adj_unit_3(int op, const AdjLaunch* L, hipStream_t s, int arg);
hipError_t adj_unit_4(int op, const AdjLaunch* L, hipStream_t s, int arg);
hipError_t adj_unit_5(int op, const AdjLaunch* L, hipStream_t s, int arg);
// stage 0: forward save; 1: BOLD forward; 2: loss; 3: BOLD reverse; 4: reverse
hipError_t launch_adjoint_stage(int stage, const AdjLaunch& L, hipStream_t s);

}  // namespace hb
