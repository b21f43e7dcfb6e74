// ddyn_ctx.h -- the context behind the opaque ddyn_ctx of include/ddyn.h,
// shared by the library's HIP sources (ddyn_hip.hip: state, steps, adjoint;
// ddyn_image.hip: images and SEDs).  Not part of the C ABI.
#ifndef DDYN_CTX_H
#define DDYN_CTX_H
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ddyn.h"

struct ImageData;   // ddyn_image.hip

struct ddyn_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    ddyn_params prm{};
    bool have_params = false;
    int64_t n_grav = 0, n_dust = 0;
    // full state, n_all entries
    double *x = nullptr, *y = nullptr, *z = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *mu = nullptr, *N = nullptr;
    int32_t* id = nullptr;
    // massive stage buffers, n_grav entries
    double4* p64[4] = {nullptr, nullptr, nullptr, nullptr};
    float4* p32[4] = {nullptr, nullptr, nullptr, nullptr};
    // second set for the overlapped schedule (grains present): swapped with p64/p32 every step
    double4* q64[4] = {nullptr, nullptr, nullptr, nullptr};
    float4* q32[4] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t gstream = nullptr;   // grain kernel stream of the overlapped schedule
    hipStream_t gstream2 = nullptr;  // second grain stream: the upper half of the grain blocks
    hipEvent_t ev_stage[2] = {nullptr, nullptr}, ev_dust[2] = {nullptr, nullptr};
    hipEvent_t ev_dust2[2] = {nullptr, nullptr};
    double* stage = nullptr;     // 9 * n_grav: vr, Sr, Sa
    double4* small_stg = nullptr;  // star + planets' stage positions of up to 1,024 steps (split small-system form)
    double* part = nullptr;      // nslice * 3 * n_grav
    int nslice = 0, slice_len = 0, tpt = 1;
    double* acc = nullptr;       // 3 * n_all accel scratch
    // adjoint tape (ddyn_adjoint_record): a checkpoint at the start of every
    // segment (7 * n_all each), the last segment's per-step states (7 * n_all)
    // and stage-1..3 massive positions (3 * n_grav), and the final state
    double *tape_ck = nullptr, *tape_seg = nullptr, *tape_fin = nullptr;
    double4* tape_stg = nullptr;
    double tape_dt = 0.0;
    int64_t tape_n = 0, tape_ck_every = 0;
    bool have_tape = false;
    bool tape_last = false;      // tape_seg / tape_stg still hold the last segment
    int32_t* rem = nullptr;      // n_dust accel scratch
    // grain coefficients
    double *beta_p = nullptr, *betasw_p = nullptr;
    bool use_beta_p = false, have_betasw_p = false;
    float *beta_t = nullptr, *betasw_t = nullptr, *q_t = nullptr;
    int32_t n_bins = 0;
    // timing
    bool timing = false;
    bool fuse_small = true;      // k_small_steps for star + planets only systems
    struct Ev { int cls; hipEvent_t a, b; };
    std::vector<Ev> events;      // recorded, not yet read
    std::vector<Ev> event_pool;  // read, reused by later launches
    int64_t t_launch[DDYN_KCLASS_COUNT] = {0, 0, 0};
    double t_ms[DDYN_KCLASS_COUNT] = {0, 0, 0};
    // images and SEDs (ddyn_image.hip): the optical tables and the scratch of
    // the per-pixel sums; freed by ddyn_image_release
    ImageData* img = nullptr;
};

// Frees the context's imaging tables and scratch (ddyn_image.hip); called by ddyn_destroy.
void ddyn_image_release(ddyn_ctx* c);

#endif
