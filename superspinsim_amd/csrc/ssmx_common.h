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

// Value of `v` from another lane of the same quad by a DPP quad
// permutation: CTRL packs the source lane of quad positions 0..3 in bit
// pairs (quad_perm [s0, s1, s2, s3] = s0 | s1 << 2 | s2 << 4 | s3 << 6).
template <int CTRL>
__device__ __forceinline__ double quad_perm(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF,
                                            0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF,
                                            0xF, false);
    return __hiloint2double(hi, lo);
}

// Value of `v` held by lane ^ 1.
__device__ __forceinline__ double swap_pair(double v)
{
    return quad_perm<0xB1>(v);  // [1, 0, 3, 2]
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

// The quad layout: slot 1 of row r is column r ^ 1, slots 2 and 3 are the
// even and the odd row of the other pair of r's quad, (r & ~1) ^ 2 and
// (r & ~1) ^ 3.  True (wave-uniform) when every lane's columns follow it;
// slot 1 is then read by DPP and slots 2, 3 by one 16-byte LDS read.
template <int SLOTS>
__device__ __forceinline__ bool quad_layout(const int (&col)[SLOTS],
                                            int lane, bool partner)
{
    if constexpr (SLOTS < 4) {
        return false;
    } else {
        const int other = (lane & ~1) ^ 2;
        const bool mine = col[1] == (lane ^ 1) && col[2] == other &&
                          col[3] == (other | 1);
        return partner &&
               __builtin_amdgcn_readfirstlane(__all(mine ? 1 : 0)) != 0;
    }
}

// y = A v with lane r holding ELL row r of A (values a[], columns col[]) and
// component r of v: one Taylor product.  Slot 0 is the lane's own
// component.  With QUAD (the quad layout), slot 1 comes from lane r ^ 1 by
// DPP and slots 2, 3 (the other pair of the quad) by one 16-byte LDS read;
// otherwise slot 1 is lane ^ 1 by DPP when `partner` is set.  The other
// slots are gathered from `xbuf` (one double per lane of the wavefront,
// 16-byte aligned) after every lane has written its component;
// LDS instructions of a wavefront execute in order, so a wavefront-scope
// fence is the only barrier.  Even and odd slots are summed in two chains in
// slot order, so the rounding depends on the row's values and columns only,
// not on how the components were fetched.
template <int SLOTS, bool QUAD>
__device__ __forceinline__ double ell_apply(const double (&a)[SLOTS],
                                            const int (&col)[SLOTS],
                                            double v, double *xbuf, int lane,
                                            bool partner)
{
    double g[SLOTS];
    g[0] = v;
    if constexpr (QUAD) {
        // One LDS instruction for the two slots instead of two gathers.
        // Reading them by DPP (two moves per double) instead was measured
        // no faster than the gathers: the moves cost as much issue as the
        // LDS instructions they remove.
        static_assert(SLOTS >= 4, "a quad row has at least 4 slots");
        g[1] = swap_pair(v);
        xbuf[lane] = v;
        wave_exchange();
        const double2 o =
            reinterpret_cast<const double2 *>(xbuf)[(lane >> 1) ^ 1];
        g[2] = o.x;
        g[3] = o.y;
#pragma unroll
        for (int sl = 4; sl < SLOTS; ++sl)
            g[sl] = xbuf[col[sl]];
        wave_exchange();
    } else {
        xbuf[lane] = v;
        wave_exchange();
        if (SLOTS > 1)
            g[1] = partner ? swap_pair(v) : xbuf[col[1]];
#pragma unroll
        for (int sl = 2; sl < SLOTS; ++sl)
            g[sl] = xbuf[col[sl]];
        wave_exchange();
    }
    double y0 = a[0] * g[0];
    double y1 = 0.0;
#pragma unroll
    for (int sl = 1; sl < SLOTS; ++sl) {
        if (sl & 1)
            y1 = fma(a[sl], g[sl], y1);
        else
            y0 = fma(a[sl], g[sl], y0);
    }
    return y0 + y1;
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
