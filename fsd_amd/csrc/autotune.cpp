// Autotuner of launch shapes: times, on the context's own positions and lists,
// every variant of the kernels that compute the same operator and keeps the
// fastest (fsd_tuning).  Coordinate search over the wave-space knobs (spreading
// kernel, contraction kernel, fused or 3-D x transforms), then an exhaustive
// search over lanes per particle x workgroup size for the real-space sum and
// for the R_FU product.  Each variant is timed with HIP events on the
// context's stream: two untimed applies, then the mean of five.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>

#include "fsd_internal.h"

#define RC(call)                \
    do {                        \
        const int r_ = (call);  \
        if (r_ < 0) return r_;  \
    } while (0)

namespace {

// device scratch and the two events, released on every return
struct Scratch {
    float *gF = nullptr, *gU = nullptr, *U = nullptr, *F = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    ~Scratch()
    {
        for (float *p : {gF, gU, U, F})
            if (p) hipFree(p);
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
    }
};

template <class Fn>
int time_us(fsd_ctx *ctx, const Scratch &s, Fn &&apply, float *us)
{
    for (int w = 0; w < 2; ++w) RC(apply());
    constexpr int REP = 5;
    FSD_HIP(hipEventRecord(s.e0, ctx->stream));
    for (int r = 0; r < REP; ++r) RC(apply());
    FSD_HIP(hipEventRecord(s.e1, ctx->stream));
    FSD_HIP(hipEventSynchronize(s.e1));
    float ms = 0.f;
    FSD_HIP(hipEventElapsedTime(&ms, s.e0, s.e1));
    *us = 1000.0f * ms / REP;
    return 0;
}

}  // namespace

// The neighbour list, the lubrication pairs and the spreading set-up must be current
// for pos.  Leaves ctx->tune at the fastest variants (tuned = 1).
int autotune_run(fsd_ctx *ctx, const float4 *pos)
{
    const int64_t N = ctx->N;
    Scratch s;
    FSD_HIP(hipMalloc(&s.gF, sizeof(float) * 11 * N));
    FSD_HIP(hipMalloc(&s.gU, sizeof(float) * 11 * N));
    FSD_HIP(hipMalloc(&s.U, sizeof(float) * 6 * N));
    FSD_HIP(hipMalloc(&s.F, sizeof(float) * 6 * N));
    FSD_HIP(hipEventCreate(&s.e0));
    FSD_HIP(hipEventCreate(&s.e1));
    RC(k_fill(ctx, 11 * N, 1.0f, s.gF));
    RC(k_fill(ctx, 6 * N, 1.0f, s.U));
    fsd_tuning &t = ctx->tune;
    // wave space: one knob at a time from the defaults
    auto wave = [&]() { return k_mobility(ctx, pos, s.gF, s.gU, 2, nullptr); };
    float best = FLT_MAX, us = 0.f;
    RC(time_us(ctx, s, wave, &best));
    const bool can_spread2 = ctx->tile == 8 && ctx->pse.P == 6;
    const bool can_contract_tiled = ctx->contract_lds > 0;
    const bool can_fx = ctx->fx_B > 0 && ctx->plan_fwd2 && ctx->plan_inv2;
    int *knobs[3] = {&t.spread, &t.contract, &t.xpass};
    const bool alt[3] = {can_spread2, can_contract_tiled, can_fx};   // the default (0) is a real choice
    for (int k = 0; k < 3; ++k) {
        if (!alt[k]) { *knobs[k] = 1; continue; }   // only the general kernel applies
        *knobs[k] = 1;
        RC(time_us(ctx, s, wave, &us));
        if (us < best) best = us;
        else *knobs[k] = 0;
    }
    t.us[0] = best;
    // real space
    best = FLT_MAX;
    int bg = t.real_g, bb = t.real_blk;
    for (int g : {4, 8, 16})
        for (int b : {128, 256}) {
            t.real_g = g; t.real_blk = b;
            RC(time_us(ctx, s, [&]() { return k_mobility(ctx, pos, s.gF, s.gU, 1, nullptr); }, &us));
            if (us < best) { best = us; bg = g; bb = b; }
        }
    t.real_g = bg; t.real_blk = bb; t.us[1] = best;
    // R_FU
    best = FLT_MAX;
    bg = t.rfu_g; bb = t.rfu_blk;
    for (int g : {8, 16})
        for (int b : {128, 256}) {
            t.rfu_g = g; t.rfu_blk = b;
            RC(time_us(ctx, s, [&]() { return k_rfu_apply(ctx, s.U, s.F, nullptr, 1.0f); }, &us));
            if (us < best) { best = us; bg = g; bb = b; }
        }
    t.rfu_g = bg; t.rfu_blk = bb; t.us[2] = best;
    t.tuned = 1;
    FSD_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}
