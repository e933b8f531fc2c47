// ssmx_common.h - host error reporting and device helpers shared by the
// propagator (ssmx_propagate.hip) and adjoint (ssmx_adjoint.hip) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/ssmx.h"

namespace ssmx {

constexpr int kWave = 64;

// One message buffer per host thread, shared by every translation unit.
inline char *error_buffer()
{
    static thread_local char buf[512] = "";
    return buf;
}

inline int set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(error_buffer(), 512, fmt, ap);
    va_end(ap);
    return code;
}

// Value of `v` held by lane ^ 1 (DPP quad_perm [1, 0, 3, 2]).
__device__ __forceinline__ double swap_pair(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xB1, 0xF,
                                            0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xB1, 0xF,
                                            0xF, false);
    return __hiloint2double(hi, lo);
}

// Broadcast lane `src` (wave-uniform) of `v` to the whole wavefront.
__device__ __forceinline__ double read_lane(double v, int src)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, kWave);
    return v;
}

// Largest v over the wavefront, returned wave-uniform (scalar register).
__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return __builtin_amdgcn_readfirstlane(v);
}

// Exchange barrier of one wavefront: LDS writes before it are visible to
// LDS reads after it (LDS instructions of a wavefront execute in order).
__device__ __forceinline__ void wave_exchange()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Node time of global output interval i, fine division r, node c, in the
// reference's order without contraction:
//   ((t0 + (r dt) / F) + dt i) + (dt c) / F
__device__ inline double node_time(double t0, double dt, int fine, int r,
                                   long i, double c)
{
#pragma clang fp contract(off)
    const double offset = t0 + ((double)r * dt) / (double)fine;
    const double start = offset + dt * (double)i;
    return start + (dt * c) / (double)fine;
}

// Waveform coefficient g at time t; `cursor` only moves forward, so each
// lane's node times must increase between calls.
__device__ inline double wave_coef(const ssmx_wave *wv,
                                   const ssmx_segment *segs, int g, double t,
                                   int &cursor)
{
#pragma clang fp contract(off)
    double c = wv->base[g];
    const int begin = wv->seg_begin[g];
    const int count = wv->seg_count[g];
    while (cursor < count && !(t < segs[begin + cursor].t_end))
        ++cursor;
    if (cursor < count) {
        const ssmx_segment sg = segs[begin + cursor];
        if (!(t < sg.t_start))
            c += sg.amp * cos(sg.omega * t + sg.phase);
    }
    return c;
}

// Same value as wave_coef for any order of times (the cursor moves both
// ways).  Also returns the segment record holding t (or -1) and the
// derivative of c_g(t) with respect to that segment's amplitude.
__device__ inline double wave_coef_seek(const ssmx_wave *wv,
                                        const ssmx_segment *segs, int g,
                                        double t, int &cursor, int &record,
                                        double &damp)
{
#pragma clang fp contract(off)
    double c = wv->base[g];
    const int begin = wv->seg_begin[g];
    const int count = wv->seg_count[g];
    record = -1;
    damp = 0.0;
    if (count == 0)
        return c;
    if (cursor >= count)
        cursor = count - 1;
    if (cursor < 0)
        cursor = 0;
    while (cursor > 0 && t < segs[begin + cursor].t_start)
        --cursor;
    while (cursor < count - 1 && !(t < segs[begin + cursor].t_end))
        ++cursor;
    const ssmx_segment sg = segs[begin + cursor];
    if (!(t < sg.t_start) && t < sg.t_end) {
        damp = cos(sg.omega * t + sg.phase);
        c += sg.amp * damp;
        record = begin + cursor;
    }
    return c;
}

}  // namespace ssmx
