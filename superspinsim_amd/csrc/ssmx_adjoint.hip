// ssmx_adjoint.hip - reverse-mode gradient of the integrated readout of the
// batched commutator-free Magnus propagator (gfx950).
//
// Forward (ssmx_propagate, non-rotating frame, substeps = 1): per step k,
//   x_{k+1} = E_{k,0} E_{k,1} ... E_{k,J-1} x_k,  E_{kj} = T_m(A_kj),
// T_m the degree-m Taylor polynomial, A_kj = sum_M f_{kj,M} G_M, and the
// readout F = sum_k w_k obs . x_k (trapezoid weights).  The adjoint runs the
// steps backwards:
//   lambda_T = (h/2) s,   lambda_k = Phi_k^T lambda_{k+1} + w_k s,
// with s = sum_o seed_o obs_o, and for each exponential (input y, output
// cotangent mu) accumulates the exponent gradient
//   Abar = sum_{i + t < m_a} c_it u_i w_t^T,  u_i = (A^T)^i mu / i!,
//   w_t = A^t y / t!,  c_it = i! t! / (i + t + 1)!,
// (m_a = SSMX_ADJ_MAX_TAYLOR, a compile-time degree: the loops over the
// adjoint and forward terms are fully unrolled and branch-free), which is
// the adjoint action of the block-triangular augmented exponential
// exp([[A, E], [0, A]]) whose upper block is the Frechet derivative of exp
// at A in direction E: the powers of A (forward terms w_t) and of A^T
// (adjoint terms u_i) are shared between the exponential and its
// derivative.  mu is then mapped through T_{m_a}(A)^T = sum_i u_i.
// Both chains carry unscaled powers, v_i = (A^T)^i mu and z_t = A^t y, so no
// multiply by 1 / i sits on either chain: Abar = sum_i v_i q_i^T with
// q_i = sum_t z_t / (i + t + 1)! (c_it u_i w_t^T = v_i z_t^T / (i + t + 1)!)
// and mu <- sum_i v_i / i!.
//
// Work split: one workgroup of two wavefronts per instance.  Wavefront 0
// (the recompute wavefront) re-runs segment p - 1 of ckpt_every steps from
// its checkpoint and stores every exponential's input y into the
// workspace, while wavefront 1 (the adjoint wavefront) runs the adjoint of
// segment p, reading those inputs in reverse; a workgroup barrier ends each
// phase.  Inside the adjoint wavefront the forward terms w_t of the next
// exponential are computed in the same loop as the adjoint terms u_i of the
// current one: two independent dependency chains per lane.
//
// Abar is accumulated on a transposed ELL pattern (lane c, slot -> row r):
// entry Abar[r][c] = sum_i v_i[r] q_i[c], so the gathers of v_i that the
// product A^T v_i needs also feed the accumulation.

#include "ssmx_common.h"

#include <cstdlib>

namespace {

using ssmx::kUnrolledTaylor;
using ssmx::kWave;
using ssmx::node_time;
using ssmx::read_lane;
using ssmx::set_error;
using ssmx::swap_pair;
using ssmx::wave_coef;
using ssmx::wave_coef_seek;
using ssmx::wave_exchange;
using ssmx::wave_sum;

constexpr int kMaxAdj = SSMX_ADJ_MAX_TAYLOR;
constexpr int kMaxSlots = 32;
// Forward-term handoff ring between the two wavefronts of an instance.
constexpr int kRing = 3;
// Bound on the polls of one handoff wait; on overrun the kernel finishes
// and reports NaN gradients instead of waiting on.  The environment
// variable SSMX_ADJOINT_MAX_POLLS lowers it (tests force the overrun path
// from the host so that every kernel still exits normally).
constexpr long kMaxPolls = 10000000;

// 1 / q!, q = 0..40; the recompute uses the same table and recursion as
// the propagator so that it reproduces the forward states.
__constant__ ssmx::InvFactorials kInvA =
    ssmx::make_inv_factorials();

struct AdjParams {
    ssmx_adjoint_desc d;
    const int32_t *col_f;
    const double *val_f;
    const int32_t *col_t;
    const double *val_t;
    const double *dval_t;
    const double *w;
    const double *nodes;
    const double *coef;
    const ssmx_wave *waves;
    const ssmx_segment *segs;
    const double *ckpt;
    const double *obs;
    const double *seed;
    double *work;
    double *grad_theta;
    double *grad_seg;
    double *lambda_out;
    long max_polls;  // bound on one handoff wait
};

// Scalar factor of matrix m in exponential j of global step `step`
// (non-rotating): h * sum_s w[j,s] c_m(t_s), in the propagator's order.
// `seek` selects the two-way segment cursor.
template <bool SEEK>
__device__ __forceinline__ double step_factor(const AdjParams &p, const double *coef_i,
                              const ssmx_wave *wv, long step, int j, int m,
                              int &cursor)
{
    const ssmx_adjoint_desc &d = p.d;
    const int S = d.n_node, K = d.n_gen;
    const long i_loc = step / d.fine;
    const int r = (int)(step - i_loc * d.fine);
    const long ig = (long)d.first_out + i_loc;
    double cw = 0.0;
    if (wv) {
        for (int s = 0; s < S; ++s) {
            const double t = node_time(d.t0, d.dt, d.fine, r, ig, p.nodes[s]);
            double c;
            if (SEEK) {
                int rec;
                double damp;
                c = wave_coef_seek(wv, p.segs, m, t, cursor, rec, damp);
            } else {
                c = wave_coef(wv, p.segs, m, t, cursor);
            }
            cw = fma(p.w[j * S + s], c, cw);
        }
    } else {
        const double *row = coef_i + ((size_t)(i_loc * d.fine + r) * S) * K;
        for (int s = 0; s < S; ++s)
            cw = fma(p.w[j * S + s], row[s * K + m], cw);
    }
    return d.h * cw;
}

// Segment record of generator g holding node s of global step `st` (or -1)
// and d c_g / d amp there.
__device__ __forceinline__ double amp_seek(const AdjParams &p,
                                        const ssmx_wave *wv, long st, int s,
                                        int g, int &cursor, int &rec)
{
    const ssmx_adjoint_desc &d = p.d;
    const long i_loc = st / d.fine;
    const int r = (int)(st - i_loc * d.fine);
    const double t = node_time(d.t0, d.dt, d.fine, r,
                               (long)d.first_out + i_loc, p.nodes[s]);
    double damp;
    wave_coef_seek(wv, p.segs, g, t, cursor, rec, damp);
    return damp;
}

// q_i = sum_{t < MA - i} z_t / (i + t + 1)! from the unscaled forward
// terms z_t = A^t y (every weight a compile-time constant); the stored-input
// and the recomputing wavefronts both form it here, in this order.
template <int MA>
__device__ __forceinline__ void forward_weights(const double (&z)[MA],
                                                double (&q)[MA])
{
#pragma unroll
    for (int i = 0; i < MA; ++i) {
        double qv = ssmx::inv_factorial(i + 1) * z[0];
#pragma unroll
        for (int t = 1; t + i < MA; ++t)
            qv = fma(ssmx::inv_factorial(i + t + 1), z[t], qv);
        q[i] = qv;
    }
}

// y = A v on an ELL pattern (slot 0 = own lane, slot 1 = partner lane
// r ^ 1 by DPP or a gathered column, slots 2.. gathered through `xbuf`).
// Same order of operations as the propagator's product.
template <int SL>
__device__ __forceinline__ double ell_product(const double (&a)[SL],
                                              const int (&col)[SL], double v,
                                              double *xbuf, int lane,
                                              bool use_partner)
{
    double g[SL];
    g[0] = v;
    xbuf[lane] = v;
    // the store and the gathers before any VALU work of the product
    __builtin_amdgcn_sched_barrier(0);
    wave_exchange();
    if (SL > 1 && !use_partner)
        g[1] = xbuf[col[1]];
#pragma unroll
    for (int sl = 2; sl < SL; ++sl)
        g[sl] = xbuf[col[sl]];
    __builtin_amdgcn_sched_barrier(0);
    if (SL > 1 && use_partner)
        g[1] = swap_pair(v);
    wave_exchange();
    double y = a[0] * g[0];
#pragma unroll
    for (int sl = 1; sl < SL; ++sl)
        y = fma(a[sl], g[sl], y);
    return y;
}

// a[sl] = sum_m f_m V_m[sl] with f_m broadcast from lane base + m of
// `factor` and V staged in LDS as [m][slot / 2][lane][slot % 2].
template <int SL>
__device__ __forceinline__ void assemble(double (&a)[SL], const double *lv,
                                         double factor, int base, int M,
                                         int lane)
{
    constexpr int SLP = SL + (SL & 1);
#pragma unroll
    for (int sl = 0; sl < SL; ++sl)
        a[sl] = 0.0;
    for (int m = 0; m < M; ++m) {
        const double f = read_lane(factor, base + m);
        const double2 *v =
            reinterpret_cast<const double2 *>(lv + m * SLP * kWave) + lane;
#pragma unroll
        for (int pr = 0; pr < SLP / 2; ++pr) {
            const double2 vv = v[pr * kWave];
            a[2 * pr] = fma(f, vv.x, a[2 * pr]);
            if (2 * pr + 1 < SL)
                a[2 * pr + 1] = fma(f, vv.y, a[2 * pr + 1]);
        }
    }
}

// Two wavefronts per instance (one workgroup of 128 threads):
//  wavefront 0 re-runs segment p - 1 from its checkpoint, storing every
//    exponential's input in the workspace, and in the same loop evaluates
//    the forward terms z_t = A^t y of segment p's exponentials in reverse
//    order (both are chains of products on the forward pattern, so one
//    16-byte exchange per term serves both); it reduces them to
//    q_i = sum_t z_t / (i + t + 1)! (forward_weights) and hands q over
//    through a ring of LDS slots;
//  wavefront 1 runs the adjoint terms v_i of segment p in reverse order and
//    accumulates the exponent gradient on the transposed pattern.
// q travels through a ring of kRing LDS slots with two counters (produced,
// consumed): wavefront 0 may run up to kRing exponentials ahead of
// wavefront 1.  A workgroup barrier per segment ends a phase, after which
// the recomputed inputs are read back; both wavefronts pass one barrier per
// phase.
// a[sl] = sum_m f_m g[m][sl] with the generator values g in registers; the
// same order of operations as assemble().
template <int SL, int MG>
__device__ __forceinline__ void assemble_reg(double (&a)[SL],
                                             const double (&g)[MG][SL],
                                             double factor, int base)
{
#pragma unroll
    for (int sl = 0; sl < SL; ++sl)
        a[sl] = 0.0;
#pragma unroll
    for (int m = 0; m < MG; ++m) {
        const double f = read_lane(factor, base + m);
#pragma unroll
        for (int sl = 0; sl < SL; ++sl)
            a[sl] = fma(f, g[m][sl], a[sl]);
    }
}

// MG = 0: generator values staged in LDS (any n_gen); MG = n_gen (1..3):
// each wavefront keeps its generator values in registers, which removes
// the LDS reads of every assembly.
// MA: adjoint degree (<= kMaxAdj).  STORED: wavefront 0 reads every
// exponential's input from `ckpt` (a forward run with SSMX_STORE_INPUTS)
// and runs only the forward terms, no recomputation.
template <int SF, int ST, int MG, int MA = kMaxAdj, bool STORED = false>
__global__ __attribute__((amdgpu_flat_work_group_size(2 * kWave, 2 * kWave),
                          amdgpu_waves_per_eu(ST <= 16 ? 2 : 1))) void
adjoint_kernel(AdjParams p)
{
    static_assert(MA >= 1 && MA <= kMaxAdj, "adjoint degree");
    constexpr int SFP = SF + (SF & 1);
    constexpr int STP = ST + (ST & 1);
    extern __shared__ double lds[];
    const ssmx_adjoint_desc &d = p.d;
    const int M = d.n_gen;
    // LDS: exchange of wavefront 0 (64 x 16 B), exchange of wavefront 1
    // (64 x 8 B), handoff [2][MA][64], staged forward / transposed values
    double *x0 = lds;
    double *x1 = lds + 2 * kWave;
    double *hand = lds + 3 * kWave;
    double *lf = hand + kRing * MA * kWave;
    double *lt = lf + (MG == 0 ? M * SFP * kWave : 0);
    // handoff counters after the staged values
    int *produced =
        reinterpret_cast<int *>(lt + (MG == 0 ? M * STP * kWave : 0));
    int *consumed = produced + 1;
    // set by either wavefront when one of its handoff waits overran: the
    // ring may then hold values the other side never saw or overwrote
    int *overrun = produced + 2;
    if (threadIdx.x == 0) {
        *produced = 0;
        *consumed = 0;
        *overrun = 0;
    }

    if constexpr (MG == 0) {
        for (int e = threadIdx.x; e < M * SFP * kWave; e += blockDim.x) {
            const int half = e & 1;
            const int row = (e >> 1) % kWave;
            const int pair = (e / (2 * kWave)) % (SFP / 2);
            const int m = e / (kWave * SFP);
            const int slot = 2 * pair + half;
            lf[e] = slot < d.slots
                        ? p.val_f[((size_t)m * d.slots + slot) * kWave + row]
                        : 0.0;
        }
        for (int e = threadIdx.x; e < M * STP * kWave; e += blockDim.x) {
            const int half = e & 1;
            const int row = (e >> 1) % kWave;
            const int pair = (e / (2 * kWave)) % (STP / 2);
            const int m = e / (kWave * STP);
            const int slot = 2 * pair + half;
            lt[e] = slot < d.slots_t
                        ? p.val_t[((size_t)m * d.slots_t + slot) * kWave +
                                  row]
                        : 0.0;
        }
    }
    __syncthreads();

    const int inst = blockIdx.x;
    const int role = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
    const int lane = threadIdx.x & (kWave - 1);
    const int n = d.n;
    const int J = d.n_exp;
    const int S = d.n_node;
    const bool use_partner = d.partner != 0;
    const long total = (long)d.n_out * d.fine;
    const int K = d.ckpt_every;
    const int P = (int)((total + K - 1) / K);
    const bool wave_mode = (d.flags & SSMX_COEF_WAVE) != 0;
    const double *coef_i =
        wave_mode ? nullptr : p.coef + (size_t)inst * d.coef_inst_stride;
    const ssmx_wave *wv = wave_mode ? p.waves + inst : nullptr;
    const size_t seg_stride = (size_t)K * J * kWave;

    if (role == 0) {
        // ------------- recomputation + forward terms -------------
        int cf[SF];
        cf[0] = lane;
#pragma unroll
        for (int sl = 1; sl < SF; ++sl)
            cf[sl] = sl < d.slots ? p.col_f[sl * kWave + lane] : lane;
        const double inv_lane = kInvA.v[lane <= SSMX_MAX_TAYLOR ? lane : 0];
        const int JM = J * M;
        const int B = kWave / JM;          // steps per coefficient batch
        const int my_b = lane / JM;
        const int my_j = (lane % JM) / M;
        const int my_m = lane % M;
        const int m_fwd = d.taylor;
        int cur_r = 0, cur_v = 0;
        double2 *xp = reinterpret_cast<double2 *>(x0);
        double gf[MG > 0 ? MG : 1][SF];
        if constexpr (MG > 0) {
#pragma unroll
            for (int m = 0; m < MG; ++m)
#pragma unroll
                for (int sl = 0; sl < SF; ++sl)
                    gf[m][sl] = sl < d.slots
                                    ? p.val_f[((size_t)m * d.slots + sl) *
                                                  kWave +
                                              lane]
                                    : 0.0;
        }
        long handed = 0;                   // exponentials handed over
        int slot = 0;                      // ring slot of the next one
        bool broken = false;

        if constexpr (STORED) {
            // Forward terms only: the input of every exponential comes
            // from the forward run (ckpt = [inst][step][j][64]).  Same
            // phases as the adjoint wavefront (one barrier each).
            (void)inv_lane;
            (void)m_fwd;
            (void)cur_r;
            const double *yin = p.ckpt + (size_t)inst * total * J * kWave;
            for (int q = 0; q <= P; ++q) {
                if (q >= 1) {
                    const int vseg = P - q;
                    const long vk0 = (long)vseg * K;
                    const long vk1 = vk0 + K < total ? vk0 + K : total;
                    const long EV = (vk1 - vk0) * J;
                    long vk = vk1 - 1;     // reverse order: (vk, 0), (vk, 1)
                    int vj = 0;
                    long vkb = vk + B;     // batch top (forces a load)
                    double vfac = 0.0;
                    // inputs prefetched two exponentials ahead (the loads
                    // stream from HBM; positions before the segment's first
                    // step are clamped to it, their values unused)
                    auto next = [&](long &k, int &j) {
                        if (++j == J) {
                            j = 0;
                            --k;
                        }
                        if (k < vk0) {
                            k = vk0;
                            j = J - 1;
                        }
                    };
                    long pk = vk;
                    int pj = vj;
                    double ypre = yin[((size_t)pk * J + pj) * kWave + lane];
                    next(pk, pj);
                    double ypre2 = yin[((size_t)pk * J + pj) * kWave + lane];
                    for (long e = 0; e < EV; ++e) {
                        double aV[SF];
                        if (vk <= vkb - B) {
                            vkb = vk;
                            const long st = vk - my_b;
                            vfac = (my_b < B && st >= vk0)
                                       ? step_factor<true>(p, coef_i, wv, st,
                                                           my_j, my_m, cur_v)
                                       : 0.0;
                        }
                        const int vbase = ((int)(vkb - vk) * J + vj) * M;
                        if constexpr (MG > 0)
                            assemble_reg<SF, MG>(aV, gf, vfac, vbase);
                        else
                            assemble<SF>(aV, lf, vfac, vbase, M, lane);
                        double W[MA];
                        W[0] = ypre;
                        ypre = ypre2;
                        next(pk, pj);
                        ypre2 = yin[((size_t)pk * J + pj) * kWave + lane];
                        // z_t = A^t y (unscaled), one 8-byte exchange
                        // per term
#pragma unroll
                        for (int t = 1; t < MA; ++t)
                            W[t] = ell_product<SF>(aV, cf, W[t - 1], x0,
                                                   lane, use_partner);
                        double Q[MA];
                        forward_weights<MA>(W, Q);
                        // wait for a free slot, fill it, publish it
                        for (long poll = 0;
                             !broken &&
                             handed - __hip_atomic_load(
                                          consumed, __ATOMIC_ACQUIRE,
                                          __HIP_MEMORY_SCOPE_WORKGROUP) >=
                                 kRing;
                             ++poll) {
                            if (poll >= p.max_polls) {
                                broken = true;
                                if (lane == 0)
                                    __hip_atomic_store(
                                        overrun, 1, __ATOMIC_RELEASE,
                                        __HIP_MEMORY_SCOPE_WORKGROUP);
                            } else {
                                __builtin_amdgcn_s_sleep(1);
                            }
                        }
                        double *hb = hand + slot * MA * kWave;
#pragma unroll
                        for (int i = 0; i < MA; ++i)
                            hb[i * kWave + lane] = Q[i];
                        if (++slot == kRing)
                            slot = 0;
                        ++handed;
                        if (lane == 0)
                            __hip_atomic_store(produced, (int)handed,
                                               __ATOMIC_RELEASE,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                        if (++vj == J) {
                            vj = 0;
                            --vk;
                        }
                    }
                }
                __syncthreads();           // end of phase
            }
            return;
        }

        for (int q = 0; q <= P; ++q) {
            const bool r_seg = q < P;      // recompute segment P - 1 - q
            const bool v_seg = q >= 1;     // forward terms of segment P - q
            const int rseg = r_seg ? P - 1 - q : 0;
            const int vseg = v_seg ? P - q : 0;
            const long rk0 = (long)rseg * K;
            const long rk1 = r_seg ? (rk0 + K < total ? rk0 + K : total) : rk0;
            const long vk0 = (long)vseg * K;
            const long vk1 = v_seg ? (vk0 + K < total ? vk0 + K : total) : vk0;
            const long ER = (rk1 - rk0) * J;
            const long EV = (vk1 - vk0) * J;
            double *wr = p.work + ((size_t)inst * 2 + (rseg & 1)) * seg_stride;
            const double *wvin =
                p.work + ((size_t)inst * 2 + (vseg & 1)) * seg_stride;

            double x = (r_seg && lane < n)
                           ? p.ckpt[((size_t)inst * P + rseg) * n + lane]
                           : 0.0;
            long rk = rk0;                 // recompute: step, exponential,
            int rj = J - 1;                // position in the batch
            int rb = 0;
            cur_r = 0;                     // forward-only cursor: restart
            double rfac = 0.0;
            long vk = vk1 - 1;             // forward terms, reverse order
            int vj = 0;
            long vkb = vk + B;             // batch top (forces a load)
            double vfac = 0.0;
            long pk = vk;                  // prefetched input (clamped)
            int pj = vj;
            double ypre = EV > 0 ? wvin[((size_t)(pk - vk0) * J + pj) *
                                            kWave + lane]
                                 : 0.0;
            const long EMAX = ER > EV ? ER : EV;
            for (long e = 0; e < EMAX; ++e) {
                const bool r_on = e < ER;
                const bool v_on = e < EV;
                double aR[SF], aV[SF];
#pragma unroll
                for (int sl = 0; sl < SF; ++sl) {
                    aR[sl] = 0.0;
                    aV[sl] = 0.0;
                }
                if (r_on) {
                    if (rj == J - 1) {
                        if (rb == B)
                            rb = 0;
                        if (rb == 0) {
                            const long mine = rk + my_b;
                            rfac = (my_b < B && mine < rk1)
                                       ? step_factor<false>(p, coef_i, wv,
                                                            mine, my_j, my_m,
                                                            cur_r)
                                       : 0.0;
                        }
                    }
                    wr[((size_t)(rk - rk0) * J + rj) * kWave + lane] = x;
                    if constexpr (MG > 0)
                        assemble_reg<SF, MG>(aR, gf, rfac, (rb * J + rj) * M);
                    else
                        assemble<SF>(aR, lf, rfac, (rb * J + rj) * M, M, lane);
                }
                double W[MA];
                W[0] = ypre;
                if (v_on) {
                    if (vk <= vkb - B) {
                        vkb = vk;
                        const long st = vk - my_b;
                        vfac = (my_b < B && st >= vk0)
                                   ? step_factor<true>(p, coef_i, wv, st,
                                                       my_j, my_m, cur_v)
                                   : 0.0;
                    }
                    const int vbase = ((int)(vkb - vk) * J + vj) * M;
                    if constexpr (MG > 0)
                        assemble_reg<SF, MG>(aV, gf, vfac, vbase);
                    else
                        assemble<SF>(aV, lf, vfac, vbase, M, lane);
                    // prefetch the next exponential's input (clamped to
                    // the segment's last one)
                    if (++pj == J) {
                        pj = 0;
                        --pk;
                    }
                    if (pk < vk0) {
                        pk = vk0;
                        pj = J - 1;
                    }
                    ypre = wvin[((size_t)(pk - vk0) * J + pj) * kWave + lane];
                }
                // fused chains: recompute term t and forward term t
                double term = x;
                double acc = x;
#pragma unroll
                for (int t = 1; t < MA; ++t) {
                    xp[lane] = make_double2(term, W[t - 1]);
                    ssmx::wave_exchange();
                    double gR[SF], gV[SF];
                    gR[0] = term;
                    gV[0] = W[t - 1];
                    if (use_partner) {
                        gR[1] = swap_pair(term);
                        gV[1] = swap_pair(W[t - 1]);
                    } else {
                        const double2 g = xp[cf[1]];
                        gR[1] = g.x;
                        gV[1] = g.y;
                    }
#pragma unroll
                    for (int sl = 2; sl < SF; ++sl) {
                        const double2 g = xp[cf[sl]];
                        gR[sl] = g.x;
                        gV[sl] = g.y;
                    }
                    ssmx::wave_exchange();
                    // the propagator's product and roundings: same terms,
                    // same order; y_t = A y_(t-1) unscaled, acc += y_t / t!
                    // in one fma
                    double y = aR[0] * gR[0];
                    double v = aV[0] * gV[0];
#pragma unroll
                    for (int sl = 1; sl < SF; ++sl) {
                        y = fma(aR[sl], gR[sl], y);
                        v = fma(aV[sl], gV[sl], v);
                    }
                    term = y;
                    if (t <= m_fwd)
                        acc = fma(term, ssmx::inv_factorial(t), acc);
                    W[t] = v;
                }
                // the rest of the recomputation's chain: for the common
                // degree with constant 1 / t! (a v_readlane result would
                // stall each product), else the loop
                if (m_fwd == kUnrolledTaylor) {
#pragma unroll
                    for (int t = MA; t <= kUnrolledTaylor; ++t) {
                        term = ell_product<SF>(aR, cf, term, x0, lane,
                                               use_partner);
                        acc = fma(term, ssmx::inv_factorial(t), acc);
                    }
                } else {
                    for (int t = MA; t <= m_fwd; ++t) {
                        term = ell_product<SF>(aR, cf, term, x0, lane,
                                               use_partner);
                        acc = fma(term, read_lane(inv_lane, t), acc);
                    }
                }
                if (r_on) {
                    x = acc;
                    if (rj == 0) {
                        rj = J - 1;
                        ++rk;
                        ++rb;
                    } else {
                        --rj;
                    }
                }
                if (v_on) {
                    double Q[MA];
                    forward_weights<MA>(W, Q);
                    // wait for a free slot, fill it, publish it
                    for (long poll = 0;
                         !broken &&
                         handed - __hip_atomic_load(
                                      consumed, __ATOMIC_ACQUIRE,
                                      __HIP_MEMORY_SCOPE_WORKGROUP) >=
                             kRing;
                         ++poll) {
                        if (poll >= p.max_polls) {
                            broken = true;
                            if (lane == 0)
                                __hip_atomic_store(
                                    overrun, 1, __ATOMIC_RELEASE,
                                    __HIP_MEMORY_SCOPE_WORKGROUP);
                        } else {
                            __builtin_amdgcn_s_sleep(1);
                        }
                    }
                    double *hb = hand + slot * MA * kWave;
#pragma unroll
                    for (int i = 0; i < MA; ++i)
                        hb[i * kWave + lane] = Q[i];
                    if (++slot == kRing)
                        slot = 0;
                    ++handed;
                    if (lane == 0)
                        __hip_atomic_store(produced, (int)handed,
                                           __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
                    if (++vj == J) {
                        vj = 0;
                        --vk;
                    }
                }
            }
            __syncthreads();               // end of phase
        }
        return;
    }

    // ------------- adjoint terms -------------
    // The adjoint chain bounds the kernel (the forward-term wavefront has
    // slack): when both are ready to issue, this one goes first.
    __builtin_amdgcn_s_setprio(2);
    int ct[ST];
    ct[0] = lane;
#pragma unroll
    for (int sl = 1; sl < ST; ++sl)
        ct[sl] = sl < d.slots_t ? p.col_t[sl * kWave + lane] : lane;
    const int grad_gen = p.grad_seg ? d.grad_gen : -1;
    const int theta_gen = d.theta_gen;
    double seedrow = 0.0;
    for (int o = 0; o < d.n_obs; ++o)
        seedrow = fma(p.seed[(size_t)inst * d.n_obs + o],
                      p.obs[o * kWave + lane], seedrow);
    double lam = (0.5 * d.h) * seedrow;
    double Sd[ST];
#pragma unroll
    for (int sl = 0; sl < ST; ++sl)
        Sd[sl] = 0.0;
    int cur_rec = -1;
    int cursor_a = 0;
    const double *lt_grad = lt + (grad_gen >= 0 ? grad_gen : 0) * STP * kWave;
    double gt[MG > 0 ? MG : 1][ST];
    if constexpr (MG > 0) {
#pragma unroll
        for (int m = 0; m < MG; ++m)
#pragma unroll
            for (int sl = 0; sl < ST; ++sl)
                gt[m][sl] =
                    sl < d.slots_t
                        ? p.val_t[((size_t)m * d.slots_t + sl) * kWave + lane]
                        : 0.0;
    }
    // Amplitude gradient of the current segment record: per-lane partial
    // sums of <Abar, G_grad_gen>, reduced over the wavefront only when the
    // record changes.
    double amp_acc = 0.0;
    long taken = 0;                     // exponentials taken over
    int slot = 0;                       // ring slot of the next one
    bool broken = false;
    // after an overrun every record this instance adds to is poisoned
    auto flush = [&]() __attribute__((always_inline)) {
        if (cur_rec >= 0) {
            const double v =
                broken ? __builtin_nan("") : wave_sum(amp_acc);
            if (lane == 0)
                p.grad_seg[cur_rec] += v;
            amp_acc = 0.0;
        }
    };
    const int JM = J * M;
    const int WB = JM + S;              // batch lanes per step
    const int BA = kWave / WB;          // steps per batch
    const int my_b = lane / WB;
    const int my_r = lane % WB;

    for (int q = 0; q <= P; ++q) {
        if (q >= 1) {
            const int seg = P - q;
            const long k0 = (long)seg * K;
            const long k1 = k0 + K < total ? k0 + K : total;
            const long E = (k1 - k0) * J;
            long kb = 0;                // batch holds kb, kb - 1, ...
            double fac = 0.0, damp_l = 0.0;
            int rec_l = -1;
            auto load_batch = [&](long top) __attribute__((always_inline)) {
                kb = top;
                const long st = top - my_b;
                fac = 0.0;
                damp_l = 0.0;
                rec_l = -1;
                if (my_b < BA && st >= k0) {
                    if (my_r < JM) {
                        fac = step_factor<true>(p, coef_i, wv, st, my_r / M,
                                                my_r % M, cursor_a);
                    } else if (grad_gen >= 0) {
                        damp_l = amp_seek(p, wv, st, my_r - JM, grad_gen,
                                          cursor_a, rec_l);
                    }
                }
            };
            long ck = k1 - 1;
            int cj = 0;
            load_batch(ck);
            for (long e = 0; e < E; ++e) {
                if (ck <= kb - BA)
                    load_batch(ck);
                const int base = (int)(kb - ck) * WB;
                double a_t[ST];
                if constexpr (MG > 0)
                    assemble_reg<ST, MG>(a_t, gt, fac, base + cj * M);
                else
                    assemble<ST>(a_t, lt, fac, base + cj * M, M, lane);
                const double f_dep =
                    theta_gen >= 0 ? read_lane(fac, base + cj * M + theta_gen)
                                   : 0.0;
                // wait until q of this exponential is published
                for (long poll = 0;
                     !broken &&
                     __hip_atomic_load(produced, __ATOMIC_ACQUIRE,
                                       __HIP_MEMORY_SCOPE_WORKGROUP) <=
                         taken;
                     ++poll) {
                    if (poll >= p.max_polls) {
                        broken = true;
                        if (lane == 0)
                            __hip_atomic_store(overrun, 1, __ATOMIC_RELEASE,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
                    } else {
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                const double *hb = hand + slot * MA * kWave;
                if (++slot == kRing)
                    slot = 0;
                double Q[MA];
#pragma unroll
                for (int i = 0; i < MA; ++i)
                    Q[i] = hb[i * kWave + lane];
                // the slot is free once its values are in registers
                ssmx::wave_exchange();
                __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
                // a producer that overran marks it before overwriting a
                // slot, so values read after that carry the mark
                if (__hip_atomic_load(overrun, __ATOMIC_ACQUIRE,
                                      __HIP_MEMORY_SCOPE_WORKGROUP))
                    broken = true;
                ++taken;
                if (lane == 0)
                    __hip_atomic_store(consumed, (int)taken, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_WORKGROUP);

                double u = lam;           // v_0 = mu
                double lam_new = lam;
                double T[ST];
#pragma unroll
                for (int sl = 0; sl < ST; ++sl)
                    T[sl] = 0.0;
                // the stored-input variant always has the partner slot
                const bool partner_slot = STORED || use_partner;
                double gp[ST];          // the previous term's gathers
#pragma unroll
                for (int i = 0; i < MA; ++i) {
                    x1[lane] = u;
                    // the store and the gathers before any VALU work
                    __builtin_amdgcn_sched_barrier(0);
                    ssmx::wave_exchange();
                    double gu[ST];
                    gu[0] = u;
                    if (!partner_slot)
                        gu[1] = x1[ct[1]];
#pragma unroll
                    for (int sl = 2; sl < ST; ++sl)
                        gu[sl] = x1[ct[sl]];
                    __builtin_amdgcn_sched_barrier(0);
                    if (partner_slot)
                        gu[1] = swap_pair(u);
                    // The previous term's accumulation, issued while this
                    // term's gathers are in flight (it needs none of them).
                    // The empty asm keeps each accumulation here: T is only
                    // read in conditional blocks below, and sinking the
                    // chain there would keep every term's gathers alive.
                    if (i > 0) {
#pragma unroll
                        for (int sl = 0; sl < ST; ++sl) {
                            T[sl] = fma(gp[sl], Q[i - 1], T[sl]);
                            asm volatile("" : "+v"(T[sl]));
                        }
                    }
                    ssmx::wave_exchange();
                    double u0 = a_t[0] * gu[0];
                    double u1 = 0.0;
#pragma unroll
                    for (int sl = 1; sl < ST; ++sl) {
                        if (sl & 1)
                            u1 = fma(a_t[sl], gu[sl], u1);
                        else
                            u0 = fma(a_t[sl], gu[sl], u0);
                    }
                    // v_(i+1) = A^T v_i, unscaled; mu_new += v / (i+1)!
                    u = u0 + u1;
                    lam_new = fma(u, ssmx::inv_factorial(i + 1), lam_new);
#pragma unroll
                    for (int sl = 0; sl < ST; ++sl)
                        gp[sl] = gu[sl];
                }
#pragma unroll
                for (int sl = 0; sl < ST; ++sl) {
                    T[sl] = fma(gp[sl], Q[MA - 1], T[sl]);
                    asm volatile("" : "+v"(T[sl]));
                }
                lam = lam_new;

                // exponent gradient -> accumulators
                if (theta_gen >= 0) {
#pragma unroll
                    for (int sl = 0; sl < ST; ++sl)
                        Sd[sl] = fma(f_dep, T[sl], Sd[sl]);
                }
                if (grad_gen >= 0) {
                    // <Abar, G_grad_gen> restricted to this lane's column
                    double tg = 0.0;
                    if constexpr (MG > 0) {
#pragma unroll
                        for (int m = 0; m < MG; ++m)
                            if (m == grad_gen) {
#pragma unroll
                                for (int sl = 0; sl < ST; ++sl)
                                    tg = fma(T[sl], gt[m][sl], tg);
                            }
                    } else {
                        const double2 *g2 =
                            reinterpret_cast<const double2 *>(lt_grad) + lane;
#pragma unroll
                        for (int pr = 0; pr < STP / 2; ++pr) {
                            const double2 gg = g2[pr * kWave];
                            tg = fma(T[2 * pr], gg.x, tg);
                            if (2 * pr + 1 < ST)
                                tg = fma(T[2 * pr + 1], gg.y, tg);
                        }
                    }
#pragma unroll
                    for (int s = 0; s < SSMX_MAX_NODE; ++s) {
                        if (s < S) {
                            const int rec =
                                __builtin_amdgcn_readlane(rec_l, base + JM + s);
                            if (rec >= 0) {
                                const double dw =
                                    (d.h * p.w[cj * S + s]) *
                                    read_lane(damp_l, base + JM + s);
                                if (rec != cur_rec) {
                                    flush();
                                    cur_rec = rec;
                                }
                                amp_acc = fma(dw, tg, amp_acc);
                            }
                        }
                    }
                }
                if (cj == J - 1)
                    lam = fma(ck == 0 ? 0.5 * d.h : d.h, seedrow, lam);
                if (++cj == J) {
                    cj = 0;
                    --ck;
                }
            }
        }
        __syncthreads();                   // end of phase
    }

    flush();
    if (broken) {
        // a handoff wait overran its bound: report NaN instead of values
        const double nan = __builtin_nan("");
#pragma unroll
        for (int sl = 0; sl < ST; ++sl)
            Sd[sl] = nan;
        lam = nan;
    }
    if (p.grad_theta) {
        for (int t = 0; t < d.n_theta; ++t) {
            double v = 0.0;
#pragma unroll
            for (int sl = 0; sl < ST; ++sl)
                if (sl < d.slots_t)
                    v = fma(Sd[sl],
                            p.dval_t[((size_t)t * d.slots_t + sl) * kWave +
                                     lane],
                            v);
            v = wave_sum(v);
            if (lane == 0)
                p.grad_theta[(size_t)inst * d.n_theta + t] = v;
        }
    }
    if (p.lambda_out && lane < n)
        p.lambda_out[(size_t)inst * n + lane] = lam;
}

template <int SF, int ST, int MG = 0, int MA = kMaxAdj, bool STORED = false>
int launch(const AdjParams &p, hipStream_t stream)
{
    constexpr int SFP = SF + (SF & 1);
    constexpr int STP = ST + (ST & 1);
    const int M = p.d.n_gen;
    // exchange buffers, handoff ring, staged values, three int counters
    const size_t staged = MG == 0 ? (size_t)M * (SFP + STP) * kWave : 0;
    const size_t lds = (staged + 3 * kWave + (size_t)kRing * kMaxAdj * kWave +
                        2) *
                       sizeof(double);
    if (lds > 160 * 1024)
        return set_error(SSMX_EINVAL,
                         "adjoint tables need %zu bytes of LDS (> 160 KiB)",
                         lds);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void *>(
                &adjoint_kernel<SF, ST, MG, MA, STORED>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return set_error(SSMX_ELAUNCH, "hipFuncSetAttribute: %s",
                             hipGetErrorString(e));
    }
    hipLaunchKernelGGL((adjoint_kernel<SF, ST, MG, MA, STORED>),
                       dim3(p.d.n_inst),
                       dim3(2 * kWave), lds, stream, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(SSMX_ELAUNCH, "adjoint launch: %s",
                         hipGetErrorString(e));
    return SSMX_OK;
}

// (forward, transposed) slot counts with a compiled kernel, cheapest first;
// a problem runs on the first pair that holds both patterns.
constexpr int kPairs[][2] = {{4, 4},   {4, 7},   {7, 7},   {4, 10},
                             {7, 10},  {10, 10}, {7, 16},  {10, 16},
                             {16, 16}, {16, 24}, {24, 24}, {32, 32}};

// Stored-input variants: the planned adjoint degree rounded up to 8, 9 or
// 10.
template <int SF, int ST, int MG>
int launch_stored(const AdjParams &p, hipStream_t s)
{
    if (p.d.taylor_adj <= 8)
        return launch<SF, ST, MG, 8, true>(p, s);
    if (p.d.taylor_adj == 9)
        return launch<SF, ST, MG, 9, true>(p, s);
    return launch<SF, ST, MG, 10, true>(p, s);
}

int dispatch(const AdjParams &p, hipStream_t s)
{
    int pick = -1;
    for (int i = 0; i < (int)(sizeof(kPairs) / sizeof(kPairs[0])); ++i)
        if (p.d.slots <= kPairs[i][0] && p.d.slots_t <= kPairs[i][1]) {
            pick = i;
            break;
        }
    if (p.d.flags & SSMX_STORE_INPUTS) {
        // compiled for two or three register-held matrices, at most 7
        // forward and 10 adjoint slots
        const int M = p.d.n_gen;
        if ((M == 2 || M == 3) && pick >= 0 && pick <= 4) {
            const bool ten = p.d.slots_t > 7;
            if (M == 2)
                return ten ? launch_stored<7, 10, 2>(p, s)
                           : launch_stored<7, 7, 2>(p, s);
            return ten ? launch_stored<7, 10, 3>(p, s)
                       : launch_stored<7, 7, 3>(p, s);
        }
        return set_error(SSMX_EINVAL,
                         "stored-input adjoint needs 2 or 3 matrices and at "
                         "most 7 / 10 slots (have %d, %d / %d)", M,
                         p.d.slots, p.d.slots_t);
    }
    // two or three matrices: generator values in registers
    const int M = p.d.n_gen;
    if (M == 2 || M == 3) {
        const bool two = M == 2;
        switch (pick) {
        case 0: return two ? launch<4, 4, 2>(p, s) : launch<4, 4, 3>(p, s);
        case 1: return two ? launch<4, 7, 2>(p, s) : launch<4, 7, 3>(p, s);
        case 2: return two ? launch<7, 7, 2>(p, s) : launch<7, 7, 3>(p, s);
        case 3: return two ? launch<4, 10, 2>(p, s) : launch<4, 10, 3>(p, s);
        case 4: return two ? launch<7, 10, 2>(p, s) : launch<7, 10, 3>(p, s);
        default: break;
        }
    }
    switch (pick) {
    case 0: return launch<4, 4>(p, s);
    case 1: return launch<4, 7>(p, s);
    case 2: return launch<7, 7>(p, s);
    case 3: return launch<4, 10>(p, s);
    case 4: return launch<7, 10>(p, s);
    case 5: return launch<10, 10>(p, s);
    case 6: return launch<7, 16>(p, s);
    case 7: return launch<10, 16>(p, s);
    case 8: return launch<16, 16>(p, s);
    case 9: return launch<16, 24>(p, s);
    case 10: return launch<24, 24>(p, s);
    case 11: return launch<32, 32>(p, s);
    default: break;
    }
    return set_error(SSMX_EINVAL, "no adjoint kernel for slots %d / %d",
                     p.d.slots, p.d.slots_t);
}

int validate(const ssmx_adjoint_desc *d)
{
    if (!d)
        return set_error(SSMX_EINVAL, "null descriptor");
    if (d->n < 1 || d->n > SSMX_MAX_N_WIDE)
        return set_error(SSMX_EINVAL, "n=%d outside 1..%d", d->n,
                         SSMX_MAX_N_WIDE);
    if (d->n_inst < 1)
        return set_error(SSMX_EINVAL, "n_inst=%d < 1", d->n_inst);
    if (d->n_gen < 1 || d->n_gen > SSMX_MAX_GEN)
        return set_error(SSMX_EINVAL, "n_gen=%d outside 1..%d", d->n_gen,
                         SSMX_MAX_GEN);
    if (d->slots < 1 || d->slots > kMaxSlots)
        return set_error(SSMX_EINVAL, "slots=%d outside 1..%d", d->slots,
                         kMaxSlots);
    if (d->slots_t < 1 || d->slots_t > kMaxSlots)
        return set_error(SSMX_EINVAL, "slots_t=%d outside 1..%d", d->slots_t,
                         kMaxSlots);
    if (d->n_exp < 1 || d->n_exp > SSMX_MAX_EXP)
        return set_error(SSMX_EINVAL, "n_exp=%d outside 1..%d", d->n_exp,
                         SSMX_MAX_EXP);
    if (d->n_node < 1 || d->n_node > SSMX_MAX_NODE)
        return set_error(SSMX_EINVAL, "n_node=%d outside 1..%d", d->n_node,
                         SSMX_MAX_NODE);
    if (d->n_exp * d->n_gen > kWave ||
        d->n_exp * d->n_gen + d->n_node > kWave)
        return set_error(SSMX_EINVAL,
                         "n_exp * n_gen + n_node = %d exceeds one batch",
                         d->n_exp * d->n_gen + d->n_node);
    if (d->fine < 1 || d->n_out < 1)
        return set_error(SSMX_EINVAL, "fine=%d, n_out=%d", d->fine, d->n_out);
    if (d->taylor < 1 || d->taylor > SSMX_MAX_TAYLOR)
        return set_error(SSMX_EINVAL, "taylor=%d outside 1..%d", d->taylor,
                         SSMX_MAX_TAYLOR);
    if (d->taylor_adj < 1 || d->taylor_adj > SSMX_ADJ_MAX_TAYLOR)
        return set_error(SSMX_EINVAL, "taylor_adj=%d outside 1..%d",
                         d->taylor_adj, SSMX_ADJ_MAX_TAYLOR);
    if (d->n_obs < 1 || d->n_obs > SSMX_MAX_OBS)
        return set_error(SSMX_EINVAL, "n_obs=%d outside 1..%d", d->n_obs,
                         SSMX_MAX_OBS);
    if (d->flags & ~(SSMX_COEF_WAVE | SSMX_STORE_INPUTS))
        return set_error(SSMX_EINVAL,
                         "flags=%d: the adjoint supports SSMX_COEF_WAVE and "
                         "SSMX_STORE_INPUTS only (non-rotating frame, "
                         "integral readout)",
                         d->flags);
    if ((d->flags & SSMX_STORE_INPUTS) && d->n > SSMX_MAX_N)
        return set_error(SSMX_EINVAL,
                         "stored inputs are supported for n <= %d",
                         SSMX_MAX_N);
    if (d->first_out < 0)
        return set_error(SSMX_EINVAL, "first_out=%d < 0", d->first_out);
    if (d->partner != 0 && d->partner != 1)
        return set_error(SSMX_EINVAL, "partner flag must be 0 or 1");
    if (d->ckpt_every < 1)
        return set_error(SSMX_EINVAL, "ckpt_every=%d < 1", d->ckpt_every);
    if (d->n_theta < 0 || d->n_theta > SSMX_ADJ_MAX_THETA)
        return set_error(SSMX_EINVAL, "n_theta=%d outside 0..%d", d->n_theta,
                         SSMX_ADJ_MAX_THETA);
    if (d->theta_gen < -1 || d->theta_gen >= d->n_gen)
        return set_error(SSMX_EINVAL, "theta_gen=%d outside -1..%d",
                         d->theta_gen, d->n_gen - 1);
    if (d->n_theta > 0 && d->theta_gen < 0)
        return set_error(SSMX_EINVAL, "derivative tables need theta_gen");
    if (d->grad_gen < -1 || d->grad_gen >= d->n_gen)
        return set_error(SSMX_EINVAL, "grad_gen=%d outside -1..%d",
                         d->grad_gen, d->n_gen - 1);
    if (d->grad_gen >= 0 && !(d->flags & SSMX_COEF_WAVE))
        return set_error(SSMX_EINVAL,
                         "segment gradients need waveform coefficients");
    if (!(d->h > 0.0) || !(d->dt > 0.0))
        return set_error(SSMX_EINVAL, "h and dt must be positive");
    if (d->n > SSMX_MAX_N) {
        // several wavefronts per instance (ssmx_adjoint_wide.hip)
        if (d->partner != 1)
            return set_error(SSMX_EINVAL, "n > 64 needs partner slots");
        if (d->grad_gen >= 0)
            return set_error(SSMX_EINVAL,
                             "n > 64: segment amplitude gradients are not "
                             "supported (grad_gen must be -1)");
        if (d->reserved < 0 || d->reserved > d->slots)
            return set_error(SSMX_EINVAL,
                             "reserved (matrix slots) = %d outside 0..%d",
                             d->reserved, d->slots);
    } else if (d->reserved != 0) {
        return set_error(SSMX_EINVAL, "reserved must be 0 for n <= 64");
    }
    return SSMX_OK;
}

}  // namespace

int ssmx_adjoint_wide(const ssmx_adjoint_desc *desc, const int32_t *ell_col,
                      const double *ell_val, const int32_t *ell_col_t,
                      const double *dval_t, const double *cf_weights,
                      const double *cf_nodes, const void *coef,
                      const ssmx_segment *segments, const double *ckpt,
                      const double *obs, const double *seed,
                      double *workspace, double *grad_theta,
                      double *lambda_out, hipStream_t stream);

extern "C" {

int ssmx_adjoint_workspace_bytes(const ssmx_adjoint_desc *desc,
                                 uint64_t *bytes)
{
    ssmx::error_buffer()[0] = '\0';
    int rc = validate(desc);
    if (rc)
        return rc;
    if (!bytes)
        return set_error(SSMX_EINVAL, "null output pointer");
    if (desc->n > SSMX_MAX_N) {
        // inputs of every exponential of one checkpoint segment
        const uint64_t np = (uint64_t)(desc->n + kWave - 1) / kWave * kWave;
        *bytes = (uint64_t)desc->n_inst * (uint64_t)desc->ckpt_every *
                 (uint64_t)desc->n_exp * np * sizeof(double);
        return SSMX_OK;
    }
    if (desc->flags & SSMX_STORE_INPUTS) {
        *bytes = 0;                     // the inputs come from the forward
        return SSMX_OK;
    }
    *bytes = (uint64_t)desc->n_inst * 2u * (uint64_t)desc->ckpt_every *
             (uint64_t)desc->n_exp * kWave * sizeof(double);
    return SSMX_OK;
}

int ssmx_adjoint(const ssmx_adjoint_desc *desc, const int32_t *ell_col,
                 const double *ell_val, const int32_t *ell_col_t,
                 const double *ell_val_t, const double *dval_t,
                 const double *cf_weights, const double *cf_nodes,
                 const void *coef, const ssmx_segment *segments,
                 const double *ckpt, const double *obs, const double *seed,
                 double *workspace, double *grad_theta, double *grad_seg,
                 double *lambda_out, void *stream)
{
    ssmx::error_buffer()[0] = '\0';
    int rc = validate(desc);
    if (rc)
        return rc;
    const bool wave_mode = (desc->flags & SSMX_COEF_WAVE) != 0;
    const bool wide = desc->n > SSMX_MAX_N;
    const bool stored = (desc->flags & SSMX_STORE_INPUTS) != 0;
    if (!ell_col || !ell_val || !ell_col_t || (!ell_val_t && !wide) ||
        !cf_weights || !cf_nodes || !coef || !ckpt || !obs || !seed ||
        (!workspace && !stored))
        return set_error(SSMX_EINVAL, "null required pointer");
    if (desc->n_theta > 0 && grad_theta && !dval_t)
        return set_error(SSMX_EINVAL, "n_theta > 0 needs dval_t");
    if (wave_mode && !segments)
        return set_error(SSMX_EINVAL, "waveform mode needs segments");
    if (wide)
        return ssmx_adjoint_wide(desc, ell_col, ell_val, ell_col_t, dval_t,
                                 cf_weights, cf_nodes, coef, segments, ckpt,
                                 obs, seed, workspace, grad_theta,
                                 lambda_out,
                                 static_cast<hipStream_t>(stream));

    AdjParams p;
    p.d = *desc;
    p.col_f = ell_col;
    p.val_f = ell_val;
    p.col_t = ell_col_t;
    p.val_t = ell_val_t;
    p.dval_t = dval_t;
    p.w = cf_weights;
    p.nodes = cf_nodes;
    p.coef = wave_mode ? nullptr : static_cast<const double *>(coef);
    p.waves = wave_mode ? static_cast<const ssmx_wave *>(coef) : nullptr;
    p.segs = segments;
    p.ckpt = ckpt;
    p.obs = obs;
    p.seed = seed;
    p.work = workspace;
    p.grad_theta = desc->n_theta > 0 ? grad_theta : nullptr;
    p.grad_seg = desc->grad_gen >= 0 ? grad_seg : nullptr;
    p.lambda_out = lambda_out;
    p.max_polls = kMaxPolls;
    if (const char *env = std::getenv("SSMX_ADJOINT_MAX_POLLS")) {
        char *end = nullptr;
        const long v = std::strtol(env, &end, 10);
        if (end != env && v >= 0 && v < kMaxPolls)
            p.max_polls = v;
    }
    hipStream_t s = static_cast<hipStream_t>(stream);

    return dispatch(p, s);
}

}  // extern "C"
