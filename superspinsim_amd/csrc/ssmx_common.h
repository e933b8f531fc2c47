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

// Padding column of an unused slot s of lane r: r itself for even s, its
// partner for odd s (the row's own aligned pair); the slot's value is 0.
__device__ __forceinline__ int pad_column(int s, int lane)
{
    return (s & 1) ? (lane | 1) : (lane & ~1);
}

// Number of leading slot pairs (2 + 2k, 3 + 2k) that hold the two columns
// 2p, 2p + 1 of an aligned pair in every lane's row, when slot 1 is the
// partner r ^ 1 (else 0); wave-uniform.  Such slot pairs are read whole:
// by DPP when only the first holds one (the other pair of the row's quad),
// else by one 16-byte LDS read each.
template <int SLOTS>
__device__ __forceinline__ int fetched_pairs(const int (&col)[SLOTS],
                                             int lane, bool partner)
{
    int mine = 0;
    if (partner && SLOTS > 1 && col[1] == (lane ^ 1)) {
#pragma unroll
        for (int s = 2; s + 1 < SLOTS; s += 2) {
            if (mine == (s - 2) / 2 && (col[s] & 1) == 0 &&
                col[s + 1] == col[s] + 1)
                ++mine;
        }
    } else {
        return 0;
    }
    // the wavefront's minimum
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(mine, off, kWave);
        mine = o < mine ? o : mine;
    }
    return __builtin_amdgcn_readfirstlane(mine);
}

// y = A v with lane r holding ELL row r of A (values a[], columns col[]) and
// component r of v: one Taylor product.  Slot 0 is the lane's own
// component.  Slot 1 is lane r ^ 1 by DPP when `partner` is set (always
// when NPAIR > 0).  The first NPAIR slot pairs (2 + 2k, 3 + 2k) hold aligned
// column pairs (fetched_pairs): with NPAIR = 1 the other pair of the quad,
// read by DPP; otherwise each from one 16-byte LDS read.  The other slots
// are gathered one by one.  The LDS reads see `xbuf` (one
// double per lane of the wavefront, 16-byte aligned) after every lane has
// written its component; LDS instructions of a wavefront execute in order,
// so a wavefront-scope fence is the only barrier.  The slots are summed in
// one fma chain in slot order (the register-held slots first, the LDS
// gathers last, so one fma follows the last gather), and the rounding
// depends on the row's values and columns only, not on how the components
// were fetched.  `pre` runs after the gathers are issued and before the
// arithmetic: a caller may fill a[] there.
// No work between a product's gathers and its arithmetic.
struct NoPrologue {
    __device__ void operator()() const {}
};

template <int SLOTS, int NPAIR, typename Pre = NoPrologue>
__device__ __forceinline__ double ell_apply(const double (&a)[SLOTS],
                                            const int (&col)[SLOTS],
                                            double v, double *xbuf, int lane,
                                            bool partner, Pre &&pre = Pre{})
{
    static_assert(NPAIR >= 0 && 2 + 2 * NPAIR <= (SLOTS > 2 ? SLOTS : 2),
                  "fetched pairs exceed the slots");
    double g[SLOTS];
    g[0] = v;
    xbuf[lane] = v;
    // The store, then every LDS read, before any VALU work of the product
    // (its DPP moves and multiplies, the caller's accumulation of the
    // previous term): the round trip starts as soon as v is ready and the
    // reads queue right behind the store.
    __builtin_amdgcn_sched_barrier(0);
    wave_exchange();
    const double2 *x2 = reinterpret_cast<const double2 *>(xbuf);
    if constexpr (NPAIR != 1) {
#pragma unroll
        for (int k = 0; k < NPAIR; ++k) {
            const double2 o = x2[col[2 + 2 * k] >> 1];
            g[2 + 2 * k] = o.x;
            g[3 + 2 * k] = o.y;
        }
    }
    if (SLOTS > 1 && !(NPAIR > 0 || partner))
        g[1] = xbuf[col[1]];
#pragma unroll
    for (int sl = 2 + 2 * NPAIR; sl < SLOTS; ++sl)
        g[sl] = xbuf[col[sl]];
    __builtin_amdgcn_sched_barrier(0);
    // work the caller puts behind the round trip (the first product of an
    // exponential assembles the exponent's row here: a[] is read below)
    pre();
    if (SLOTS > 1 && (NPAIR > 0 || partner))
        g[1] = swap_pair(v);
    if constexpr (NPAIR == 1) {
        // the quad layout: the other pair of the quad is lanes
        // (r & ~1) ^ 2 and its odd neighbour, in reach of DPP.  With the
        // unrolled Taylor chain the four moves ran 1 % faster than the
        // 16-byte read (with the Taylor loop, 6 % slower).
        g[2] = quad_perm<0x0A>(v);  // [2, 2, 0, 0]
        g[3] = quad_perm<0x5F>(v);  // [3, 3, 1, 1]
    }
    wave_exchange();
    double y = a[0] * g[0];
#pragma unroll
    for (int sl = 1; sl < SLOTS; ++sl)
        y = fma(a[sl], g[sl], y);
    return y;
}

// In a wave-uniform branch: keeps it a branch.  Without it the compiler
// may compute the guarded work on every pass and select the result (a
// v_cndmask per register), issue the headline kernels cannot spare.
__device__ __forceinline__ void keep_branch()
{
    asm volatile("");
}

// The lane's row of A = sum_m f_m G_m over SLOTS slots from the staged
// values `vals` ([m][slot / 2][lane][slot % 2], 64 lanes), f_m = lane
// base + m of `factor`.  Each slot accumulates fma(f_m, v, .) from m = 0,
// the first from 0.0 (the same bits as zeroing the row first, without the
// moves).
template <int SLOTS>
__device__ __forceinline__ void assemble_row(double (&a)[SLOTS],
                                             const double *vals,
                                             double factor, int base,
                                             int M, int lane)
{
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    {
        const double f = read_lane(factor, base);
        const double2 *v = reinterpret_cast<const double2 *>(vals) + lane;
#pragma unroll
        for (int pr = 0; pr < SLOTP / 2; ++pr) {
            const double2 vv = v[pr * kWave];
            a[2 * pr] = fma(f, vv.x, 0.0);
            if (2 * pr + 1 < SLOTS)
                a[2 * pr + 1] = fma(f, vv.y, 0.0);
        }
    }
    for (int m = 1; m < M; ++m) {
        const double f = read_lane(factor, base + m);
        const double2 *v =
            reinterpret_cast<const double2 *>(vals + m * SLOTP * kWave) +
            lane;
#pragma unroll
        for (int pr = 0; pr < SLOTP / 2; ++pr) {
            const double2 vv = v[pr * kWave];
            a[2 * pr] = fma(f, vv.x, a[2 * pr]);
            if (2 * pr + 1 < SLOTS)
                a[2 * pr + 1] = fma(f, vv.y, a[2 * pr + 1]);
        }
    }
}

// assemble_row with the MG generator rows held in registers (g[m][slot]):
// the same operations in the same order, without the LDS reads.
template <int SLOTS, int MG>
__device__ __forceinline__ void assemble_row_reg(double (&a)[SLOTS],
                                                 const double (&g)[MG][SLOTS],
                                                 double factor, int base)
{
    {
        const double f = read_lane(factor, base);
#pragma unroll
        for (int sl = 0; sl < SLOTS; ++sl)
            a[sl] = fma(f, g[0][sl], 0.0);
    }
#pragma unroll
    for (int m = 1; m < MG; ++m) {
        const double f = read_lane(factor, base + m);
#pragma unroll
        for (int sl = 0; sl < SLOTS; ++sl)
            a[sl] = fma(f, g[m][sl], a[sl]);
    }
}

// assemble_row with the first MG generator rows in registers (g[m][slot])
// and rows MG..M-1 read from the staged values: the same operations in the
// same order.
template <int SLOTS, int MG>
__device__ __forceinline__ void assemble_row_mix(double (&a)[SLOTS],
                                                 const double (&g)[MG][SLOTS],
                                                 const double *vals,
                                                 double factor, int base,
                                                 int M, int lane)
{
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    assemble_row_reg<SLOTS, MG>(a, g, factor, base);
    for (int m = MG; m < M; ++m) {
        const double f = read_lane(factor, base + m);
        const double2 *v =
            reinterpret_cast<const double2 *>(vals + m * SLOTP * kWave) +
            lane;
#pragma unroll
        for (int pr = 0; pr < SLOTP / 2; ++pr) {
            const double2 vv = v[pr * kWave];
            a[2 * pr] = fma(f, vv.x, a[2 * pr]);
            if (2 * pr + 1 < SLOTS)
                a[2 * pr + 1] = fma(f, vv.y, a[2 * pr + 1]);
        }
    }
}

// 1 / t! in double, the same value wherever it is formed: the product
// 2 * 3 * ... * t rounded in that order (exact up to t = 18), then 1 / it.
__host__ __device__ constexpr double inv_factorial(int t)
{
    double f = 1.0;
    for (int k = 2; k <= t; ++k)
        f *= (double)k;
    return 1.0 / f;
}

// inv_factorial(0 .. SSMX_MAX_TAYLOR), for per-lane tables of the Taylor
// loop (lane t holds 1 / t!).
struct InvFactorials {
    double v[SSMX_MAX_TAYLOR + 1];
};

__host__ __device__ constexpr InvFactorials make_inv_factorials()
{
    InvFactorials r{};
    for (int t = 0; t <= SSMX_MAX_TAYLOR; ++t)
        r.v[t] = inv_factorial(t);
    return r;
}

// inv_factorial(T) built by two scalar moves at the point of use: as plain
// constants the compiler hoists all of them out of the step loop, where
// they spill and come back through v_readlane, whose SGPR result stalls the
// next multiply.
template <int T>
__device__ __forceinline__ double scalar_inv_factorial()
{
    constexpr uint64_t bits = __builtin_bit_cast(uint64_t, inv_factorial(T));
    uint32_t lo, hi;
    asm volatile("s_mov_b32 %0, %1" : "=s"(lo) : "i"((uint32_t)bits));
    asm volatile("s_mov_b32 %0, %1" : "=s"(hi) : "i"((uint32_t)(bits >> 32)));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// 1 / T! as an operand: 1 / 1! and 1 / 2! stay plain constants (inline
// operands), the others are built where they are used.
template <int T>
__device__ __forceinline__ double inv_factorial_operand()
{
    if constexpr (T <= 2)
        return inv_factorial(T);
    else
        return scalar_inv_factorial<T>();
}

// Terms T..M of taylor_action's unrolled chain, stopping after term m
// (template recursion, so every 1 / T! is a compile-time constant).  On
// entry `term` is y_(T-1), not yet added to `acc` when T > 1: it is added
// after the product that consumes it has issued its LDS store.
template <int SLOTS, int NPAIR, int M, int T = 1>
__device__ __forceinline__ void taylor_chain(const double (&a)[SLOTS],
                                             const int (&col)[SLOTS],
                                             double &term, double &acc,
                                             double *xbuf, int lane,
                                             bool partner, int m)
{
    if constexpr (T <= M) {
        if (T > m) {
            if constexpr (T > 1)
                acc = fma(term, inv_factorial_operand<T - 1>(), acc);
            return;
        }
        const double y =
            ell_apply<SLOTS, NPAIR>(a, col, term, xbuf, lane, partner);
        if constexpr (T > 1) {
            acc = fma(term, inv_factorial_operand<T - 1>(), acc);
            asm volatile("" : "+v"(acc));  // not sunk to the chain's end
        }
        term = y;
        taylor_chain<SLOTS, NPAIR, M, T + 1>(a, col, term, acc, xbuf, lane,
                                             partner, m);
    } else {
        acc = fma(term, inv_factorial_operand<M>(), acc);
    }
}

// exp(A) x by its Taylor series of degree m: y_t = A y_(t-1) from y_0 = x
// (the unscaled powers, each product rounded as ell_apply rounds it) and
// acc <- fma(y_t, 1 / t!, acc) for t = 1..m, one rounding per term, so
// every kernel that repeats the chain gets the same bits.  Carrying the
// powers unscaled saves the multiply of each term by 1 / t on the VALU and
// on the chain's critical path.  M > 0: the degree is M, the chain is
// unrolled and the 1 / t! are plain constants.  M < 0: the chain is unrolled
// up to -M terms and stops after the runtime degree m <= -M (one scalar
// branch per term), each 1 / t! built at its use (scalar_inv_factorial).
// M = 0: a loop over the runtime degree m, 1 / t! broadcast from lane t of
// inv_lane (which holds 1 / lane!).  All three give the same doubles in
// the same order; the unrolled forms avoid the loop's v_readlane, whose
// SGPR result stalls the dependent chain.  Each term is added to the sum
// after the next product has stored it to LDS (same operations, same
// order), so the accumulation is off the chain's critical path.
template <int SLOTS, int NPAIR, int M, typename Pre = NoPrologue>
__device__ __forceinline__ double taylor_action(const double (&a)[SLOTS],
                                                const int (&col)[SLOTS],
                                                double x, double *xbuf,
                                                int lane, bool partner, int m,
                                                double inv_lane,
                                                Pre &&pre = Pre{})
{
    double term = x;
    double acc = x;
    if constexpr (M > 0) {
        // `pre` (which may fill a[]) runs inside the first product, behind
        // the LDS store of x and its gathers
        const double y1 = ell_apply<SLOTS, NPAIR>(a, col, term, xbuf, lane,
                                                  partner, pre);
        term = y1;
#pragma unroll
        for (int t = 2; t <= M; ++t) {
            const double y = ell_apply<SLOTS, NPAIR>(a, col, term, xbuf,
                                                     lane, partner);
            acc = fma(term, inv_factorial(t - 1), acc);
            // each term is added inside the product after it; without the
            // empty asm the compiler sinks all M additions to the end of
            // the chain, holding every term in registers until then
            asm volatile("" : "+v"(acc));
            term = y;
        }
        acc = fma(term, inv_factorial(M), acc);
    } else if constexpr (M < 0) {
        pre();
        // The degree through an opaque scalar move: the per-term exit tests
        // then stay scalar compares inside the chain, where the compiler
        // would otherwise hoist all -M of them out of the step loop as lane
        // masks and spill them (v_readlane reloads in the chain).
        int mm;
        asm volatile("s_mov_b32 %0, %1" : "=s"(mm) : "s"(m));
        taylor_chain<SLOTS, NPAIR, -M>(a, col, term, acc, xbuf, lane,
                                       partner, mm);
    } else {
        pre();
        for (int t = 1; t <= m; ++t) {
            const double y = ell_apply<SLOTS, NPAIR>(a, col, term, xbuf,
                                                     lane, partner);
            if (t > 1)
                acc = fma(term, read_lane(inv_lane, t - 1), acc);
            term = y;
        }
        acc = fma(term, read_lane(inv_lane, m), acc);
    }
    return acc;
}

// The Taylor degree with a chain unrolled for exactly that degree (the
// degree the 1e-16 tolerance gives at the 20 ps steps of the NV
// workloads): taylor_action<.., kUnrolledTaylor>, the adjoint's tail and
// the wide kernels.
constexpr int kUnrolledTaylor = 11;

// Other degrees 1..kUnrolledMax run the chain with per-term exits
// (taylor_action<.., -kUnrolledMax>) in the 64-row forward and queue
// kernels.
constexpr int kUnrolledMax = 16;

// sum_s w[j,s] c(t_s) for a coefficient c = 1 at every node, in the order
// and roundings of the kernels' step factors (fma(w, c, cw) over s).
__device__ __forceinline__ double unit_weight_sum(const double *w, int j,
                                                  int n_node)
{
    double cw = 0.0;
    for (int s = 0; s < n_node; ++s)
        cw = fma(w[j * n_node + s], 1.0, cw);
    return cw;
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
