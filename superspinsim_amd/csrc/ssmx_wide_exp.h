// ssmx_wide_exp.h - the exponential x <- exp(A) x of one row per thread,
// shared by the wide propagator (ssmx_wide.hip) and the wide adjoint's
// recomputation (ssmx_adjoint_wide.hip), so that both give the same states
// bit for bit.
//
// Several wavefronts hold one instance, thread r owning row r.  Each Taylor
// term is written to an LDS exchange buffer of the instance and combined
// over the row's ELL slots after one workgroup barrier: slot 0 is the
// diagonal, slot 1 the partner r ^ 1 (DPP), other slots are gathered.  Two
// buffers alternate, kBuf doubles apart, so the gather addresses are fixed
// and the buffer is an immediate offset.  Generator values are staged in
// LDS as [m][slot / 2][row][slot % 2].
#pragma once

#include "ssmx_common.h"

namespace {

using ssmx::kWave;
using ssmx::read_lane;
using ssmx::swap_pair;

// largest workgroup: 16 wavefronts (1024 threads)
constexpr int kMaxBlockWaves = 16;
// stride between an instance's two exchange buffers (doubles): one buffer
// holds up to 512 rows; 4 KiB fits a ds_read immediate offset
constexpr int kBuf = SSMX_MAX_N_WIDE;

// 1 / q! for the Taylor loop, q = 0..40 (lane q reads entry q).
__constant__ ssmx::InvFactorials kInvW = ssmx::make_inv_factorials();

// Wave-uniform inputs of one exponential.
struct ExpArgs {
    const double *lv;       // staged values [m][slot / 2][row][slot % 2]
    int NP, M, base;        // rows, matrices, first factor lane
    int substeps, taylor;
    double inv_sub;
    bool partner;
};

// y = A v for this thread's row, reading buffer P (0 or 1): v is written to
// the row's own entry, then WS slots are combined (slot 0 diagonal, slot 1
// partner, others gathered).  Two accumulation chains (even and odd
// slots).  The store, the barrier and the gathers come before any VALU work
// of the product (the partner's DPP moves, the caller's accumulation of the
// previous term).
template <int SLOTS, int WS, int P>
__device__ __forceinline__ double wide_product(
    const double (&a)[WS], double v, double *own,
    const double *const (&gp)[SLOTS], bool partner)
{
    own[P * kBuf] = v;
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
    double g[WS];
    g[0] = v;
    if (!partner)
        g[1] = gp[1][P * kBuf];
#pragma unroll
    for (int sl = 2; sl < WS; ++sl)
        g[sl] = gp[sl][P * kBuf];
    __builtin_amdgcn_sched_barrier(0);
    if (partner)
        g[1] = swap_pair(v);
    double y0 = a[0] * g[0];
    double y1 = a[1] * g[1];
#pragma unroll
    for (int sl = 2; sl < WS; ++sl) {
        if (sl & 1)
            y1 = fma(a[sl], g[sl], y1);
        else
            y0 = fma(a[sl], g[sl], y0);
    }
    return y0 + y1;
}

// Taylor terms t = T..M of one exponential, unrolled: y_t = A y_(t-1)
// (unscaled powers) on buffer (P0 + t - 1) & 1 and acc <- fma(y_t, 1 / t!,
// acc), added after the next product has stored y_t (the same operations in
// the same order as the loop in wide_exp, so the same bits).  On entry
// `term` is y_(T-1), not yet added when T > 1.
template <int SLOTS, int WS, int M, int P0, int T = 1>
__device__ __forceinline__ void wide_chain(const double (&a)[WS],
                                           double &term, double &acc,
                                           double *own,
                                           const double *const (&gp)[SLOTS],
                                           bool partner)
{
    if constexpr (T <= M) {
        const double y = wide_product<SLOTS, WS, (P0 + T - 1) & 1>(
            a, term, own, gp, partner);
        if constexpr (T > 1) {
            acc = fma(term, ssmx::inv_factorial(T - 1), acc);
            asm volatile("" : "+v"(acc));  // not sunk to the chain's end
        }
        term = y;
        wide_chain<SLOTS, WS, M, P0, T + 1>(a, term, acc, own, gp, partner);
    } else {
        acc = fma(term, ssmx::inv_factorial(M), acc);
    }
}

// x <- exp(A_j) x for this thread's row: assemble the row of A_j from the
// staged values (slots < WS), then s substeps of the degree-m Taylor
// action.  `parity` is the buffer the next product writes.  With UNROLL,
// degree kUnrolledTaylor runs unrolled (wide_chain): a v_readlane of 1 / t!
// would stall every product.  Both forms give the same bits.
template <int SLOTS, int WS, bool UNROLL>
__device__ __forceinline__ double wide_exp(double x, const ExpArgs &e,
                                           int row, double factor,
                                           double inv_lane, double *own,
                                           const double *const (&gp)[SLOTS],
                                           int &parity)
{
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    double a[WS];
#pragma unroll
    for (int sl = 0; sl < WS; ++sl)
        a[sl] = 0.0;
    for (int m = 0; m < e.M; ++m) {
        const double f = read_lane(factor, e.base + m);
        const double2 *v =
            reinterpret_cast<const double2 *>(e.lv +
                                              (size_t)m * SLOTP * e.NP) +
            row;
        double2 vv[WS / 2];
#pragma unroll
        for (int pr = 0; pr < WS / 2; ++pr)
            vv[pr] = v[pr * e.NP];
#pragma unroll
        for (int pr = 0; pr < WS / 2; ++pr) {
            a[2 * pr] = fma(f, vv[pr].x, a[2 * pr]);
            a[2 * pr + 1] = fma(f, vv[pr].y, a[2 * pr + 1]);
        }
    }
    if (e.substeps > 1) {
#pragma unroll
        for (int sl = 0; sl < WS; ++sl)
            a[sl] *= e.inv_sub;
    }
    // y_t = A y_(t-1) unscaled; acc <- fma(y_t, 1 / t!, acc)
    auto taylor_term = [&](double y, int t, double &term, double &acc) {
        term = y;
        acc = fma(y, read_lane(inv_lane, t), acc);
    };
    for (int sub = 0; sub < e.substeps; ++sub) {
        double term = x;
        double acc = x;
        if (UNROLL && e.taylor == ssmx::kUnrolledTaylor) {
            constexpr int kM = ssmx::kUnrolledTaylor;
            if (parity)
                wide_chain<SLOTS, WS, kM, 1>(a, term, acc, own, gp,
                                             e.partner);
            else
                wide_chain<SLOTS, WS, kM, 0>(a, term, acc, own, gp,
                                             e.partner);
            parity ^= kM & 1;
            x = acc;
            continue;
        }
        // products alternate buffers; the loop body is one pair (buffer 0,
        // then 1), so both buffer offsets are immediates
        int t = 1;
        if (parity && t <= e.taylor) {
            taylor_term(wide_product<SLOTS, WS, 1>(a, term, own, gp,
                                                   e.partner),
                        t, term, acc);
            ++t;
            parity = 0;
        }
        for (; t + 1 <= e.taylor; t += 2) {
            taylor_term(wide_product<SLOTS, WS, 0>(a, term, own, gp,
                                                   e.partner),
                        t, term, acc);
            taylor_term(wide_product<SLOTS, WS, 1>(a, term, own, gp,
                                                   e.partner),
                        t + 1, term, acc);
        }
        if (t <= e.taylor) {
            taylor_term(wide_product<SLOTS, WS, 0>(a, term, own, gp,
                                                   e.partner),
                        t, term, acc);
            parity = 1;
        }
        x = acc;
    }
    return x;
}

// wide_exp for the smallest even slot class WS >= ws (wave-uniform).
template <int SLOTS, bool UNROLL, int WS = 2>
__device__ __forceinline__ double wide_exp_dispatch(
    int ws, double x, const ExpArgs &e, int row, double factor,
    double inv_lane, double *own, const double *const (&gp)[SLOTS],
    int &parity)
{
    if constexpr (WS >= SLOTS) {
        return wide_exp<SLOTS, SLOTS, UNROLL>(x, e, row, factor, inv_lane,
                                              own, gp, parity);
    } else {
        if (ws <= WS)
            return wide_exp<SLOTS, WS, UNROLL>(x, e, row, factor, inv_lane,
                                               own, gp, parity);
        return wide_exp_dispatch<SLOTS, UNROLL, WS + 2>(
            ws, x, e, row, factor, inv_lane, own, gp, parity);
    }
}

}  // namespace
