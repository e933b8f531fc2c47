// ssmx_adjoint_wide.hip - adjoint of the trapezoid-integrated readout for
// superoperators wider than one wavefront (65 <= n <= 512; gfx950), e.g.
// the 21-level NV + 14N model (n = 441 or 171).
//
// Same quantities as ssmx_adjoint.hip (include/ssmx.h): per instance the
// adjoint state lambda_0 = dLoss/d rho0 and
//   grad_theta[t] = sum_{k,j} f_{kj,theta_gen} <Abar_kj, D_t>,
//   Abar = sum_{i + t < MA} u_i w_t^T / (i + t + 1)!,
//   u_i = (A^T)^i mu,  w_t = A^t y  (unscaled powers),
// for each exponential exp(A) with input y and output cotangent mu.
//
// One workgroup of W = ceil(n / 64) wavefronts per instance, thread r
// owning row r (as the wide propagator).  Per checkpoint segment, last
// first:
//  1. recomputation from the checkpoint with the propagator's own
//     exponential (ssmx_wide_exp.h, bit-identical states), storing the
//     input of every exponential in the workspace;
//  2. the reverse sweep: per exponential (last applied first) the chain
//     u_i on the transposed pattern (lambda <- exp(A)^T lambda =
//     sum u_i / i!, and v_t = sum_i u_i / (i + t + 1)!, row-local), then
//     the chain w_t on the
//     forward pattern extended by the derivative tables' non-zeros, where
//     each exchange of w_t also adds f_theta v_t[r] w_t[c] to the row's
//     accumulator B[r][c]; finally lambda += w_k s (readout weight).
// At the end grad_theta[t] = sum_{r,c} B[r][c] D_t[r][c] (one reduction per
// parameter).  Only the matrices' own slots are staged in LDS; the values
// of A^T are read from them through per-thread source offsets, found once
// at launch by scanning row c of the forward pattern for column r.
//
// No counterpart in the reference (it has no gradients); it differentiates
// the path ssmx_propagate replaces (generate_simulator.py:1945-2116).

#include "ssmx_wide_exp.h"

namespace {

using ssmx::inv_factorial;
using ssmx::node_time;
using ssmx::set_error;
using ssmx::wave_coef_seek;
using ssmx::wave_max;
using ssmx::wave_sum;

// one instance per workgroup: at most 512 rows = 8 wavefronts
constexpr int kMaxWaves = SSMX_MAX_N_WIDE / kWave;

struct AWParams {
    ssmx_adjoint_desc d;
    int slots_f;            // staged (matrix) slots
    int zero_entry;         // offset of a staged value that is 0 in every
                            // plane
    const int32_t *col;     // [slots][NP] extended forward pattern
    const double *val;      // [M][slots][NP]
    const int32_t *col_t;   // [slots_t][NP] transposed pattern
    const double *dval;     // [n_theta][slots][NP]
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
    double *lambda_out;
};

// h sum_s w[j,s] c_m(t_s) of propagator step `step` in the propagator's
// rounding order (non-rotating frame), for steps visited in any order.
__device__ double adj_factor(const AWParams &p, const double *coef_i,
                             const ssmx_wave *wv, long step, int j, int m,
                             int &cursor)
{
    const ssmx_adjoint_desc &d = p.d;
    const int S = d.n_node, K = d.n_gen;
    const long i_loc = step / d.fine;
    const int r = (int)(step - i_loc * d.fine);
    const long ig = (long)d.first_out + i_loc;
    double cw = 0.0;
    for (int s = 0; s < S; ++s) {
        double c;
        if (wv) {
            int rec;
            double damp;
            c = wave_coef_seek(
                wv, p.segs, m,
                node_time(d.t0, d.dt, d.fine, r, ig, p.nodes[s]), cursor,
                rec, damp);
        } else {
            c = coef_i[((size_t)(i_loc * d.fine + r) * S + s) * K + m];
        }
        cw = fma(p.w[j * S + s], c, cw);
    }
    return d.h * cw;
}

// Workspace access bypasses the vector L1 (agent-scope relaxed atomics are
// plain loads and stores with coherence bits), so a thread always reads
// back its own latest store of a slot that earlier segments also used.
__device__ __forceinline__ void ws_store(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double ws_load(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave-uniform inputs of one reverse exponential.
struct RevArgs {
    const double *lv;   // staged matrix values [m][slot / 2][row][slot % 2]
    int NP, M, base;    // rows, matrices, first factor lane
    int row;
    double factor;      // this lane's factor (lane j M + m)
    double fth;         // factor of the matrix the derivative tables follow
    double *own;        // the row's entry in exchange buffer 0
};

// Step I (1..MA) of the transposed chain on buffer (I - 1) & 1: the
// unscaled power u <- A^T u, lambda += u / I!, v_t += u / (I + t + 1)! for
// t < MA - I.  The factorials are compile-time constants (no v_readlane,
// whose SGPR result stalls the next VALU instruction), and no multiply by
// 1 / I sits in the chain.
template <int SF, int WT, int MA, int I>
__device__ __forceinline__ void u_chain(double &u, double &acc,
                                        double (&v)[MA],
                                        const double (&at)[WT],
                                        const RevArgs &r,
                                        const double *const (&gpt)[SF])
{
    u = wide_product<SF, WT, (I - 1) & 1>(at, u, r.own, gpt, true);
    acc = fma(u, inv_factorial(I), acc);
#pragma unroll
    for (int t = 0; t < MA - I; ++t)
        v[t] = fma(inv_factorial(I + t + 1), u, v[t]);
    if constexpr (I < MA)
        u_chain<SF, WT, MA, I + 1>(u, acc, v, at, r, gpt);
}

// Exchange T (0..MA-1) of the forward-term chain on buffer (MA + T) & 1:
// B += f_theta v_T w[c] over the WE extended slots, then
// w <- A w (unscaled) over the WA matrix slots (not after the last).
template <int SE, int WE, int WA, int MA, int T>
__device__ __forceinline__ void w_chain(double &w, const double (&v)[MA],
                                        double (&B)[SE],
                                        const double (&a)[WA],
                                        const RevArgs &r,
                                        const double *const (&gpe)[SE])
{
    constexpr int P = (MA + T) & 1;
    r.own[P * kBuf] = w;
    __syncthreads();
    double g[WE];
    g[0] = w;
    g[1] = swap_pair(w);
#pragma unroll
    for (int sl = 2; sl < WE; ++sl)
        g[sl] = gpe[sl][P * kBuf];
    const double cv = r.fth * v[T];
#pragma unroll
    for (int sl = 0; sl < WE; ++sl)
        B[sl] = fma(cv, g[sl], B[sl]);
    if constexpr (T + 1 < MA) {
        double y0 = a[0] * g[0];
        double y1 = a[1] * g[1];
#pragma unroll
        for (int sl = 2; sl < WA; ++sl) {
            if (sl & 1)
                y1 = fma(a[sl], g[sl], y1);
            else
                y0 = fma(a[sl], g[sl], y0);
        }
        w = y0 + y1;
        w_chain<SE, WE, WA, MA, T + 1>(w, v, B, a, r, gpe);
    }
}

// Transposed half of one reverse exponential for a wavefront whose rows
// use at most WT transposed slots: lambda <- exp(A)^T lambda and the
// row-local v_t.  Starts on buffer 0, ends before buffer MA & 1.
template <int SF, int WT, int MA>
__device__ __forceinline__ void reverse_u(double &lam, double (&v)[MA],
                                          const RevArgs &r,
                                          const double *const (&gpt)[SF],
                                          const uint32_t (&src)[SF / 2])
{
    double at[WT];
#pragma unroll
    for (int sl = 0; sl < WT; ++sl)
        at[sl] = 0.0;
    for (int m = 0; m < r.M; ++m) {
        const double f = read_lane(r.factor, r.base + m);
        const double *plane = r.lv + (size_t)m * SF * r.NP;
        double vv[WT];
#pragma unroll
        for (int sl = 0; sl < WT; ++sl)
            vv[sl] = plane[(src[sl >> 1] >> (16 * (sl & 1))) & 0xFFFF];
#pragma unroll
        for (int sl = 0; sl < WT; ++sl)
            at[sl] = fma(f, vv[sl], at[sl]);
    }
    double u = lam;
    double acc = lam;
#pragma unroll
    for (int t = 0; t < MA; ++t)
        v[t] = inv_factorial(t + 1) * u;
    u_chain<SF, WT, MA, 1>(u, acc, v, at, r, gpt);
    lam = acc;
}

// Forward-term half for a wavefront whose rows use at most WE extended
// slots: B += f_theta sum_t v_t w_t^T on the extended pattern.
template <int SF, int SE, int WE, int MA>
__device__ __forceinline__ void reverse_w(double y, const double (&v)[MA],
                                          double (&B)[SE],
                                          const RevArgs &r,
                                          const double *const (&gpe)[SE])
{
    constexpr int WA = WE < SF ? WE : SF;
    double a[WA];
#pragma unroll
    for (int sl = 0; sl < WA; ++sl)
        a[sl] = 0.0;
    for (int m = 0; m < r.M; ++m) {
        const double f = read_lane(r.factor, r.base + m);
        const double2 *pv =
            reinterpret_cast<const double2 *>(r.lv +
                                              (size_t)m * SF * r.NP) +
            r.row;
        double2 vv[WA / 2];
#pragma unroll
        for (int pr = 0; pr < WA / 2; ++pr)
            vv[pr] = pv[pr * r.NP];
#pragma unroll
        for (int pr = 0; pr < WA / 2; ++pr) {
            a[2 * pr] = fma(f, vv[pr].x, a[2 * pr]);
            a[2 * pr + 1] = fma(f, vv[pr].y, a[2 * pr + 1]);
        }
    }
    double w = y;
    w_chain<SE, WE, WA, MA, 0>(w, v, B, a, r, gpe);
}

// reverse_u for the smallest even class WT >= wt (wave-uniform)
template <int SF, int MA, int WT = 2>
__device__ __forceinline__ void reverse_u_dispatch(
    int wt, double &lam, double (&v)[MA], const RevArgs &r,
    const double *const (&gpt)[SF], const uint32_t (&src)[SF / 2])
{
    if constexpr (WT >= SF) {
        reverse_u<SF, SF, MA>(lam, v, r, gpt, src);
    } else {
        if (wt <= WT)
            reverse_u<SF, WT, MA>(lam, v, r, gpt, src);
        else
            reverse_u_dispatch<SF, MA, WT + 2>(wt, lam, v, r, gpt, src);
    }
}

// reverse_w for the smallest even class WE >= we (wave-uniform)
template <int SF, int SE, int MA, int WE = 2>
__device__ __forceinline__ void reverse_w_dispatch(
    int we, double y, const double (&v)[MA], double (&B)[SE],
    const RevArgs &r, const double *const (&gpe)[SE])
{
    if constexpr (WE >= SE) {
        reverse_w<SF, SE, SE, MA>(y, v, B, r, gpe);
    } else {
        if (we <= WE)
            reverse_w<SF, SE, WE, MA>(y, v, B, r, gpe);
        else
            reverse_w_dispatch<SF, SE, MA, WE + 2>(we, y, v, B, r, gpe);
    }
}

template <int SF, int SE, int MA>
__global__ __launch_bounds__(kWave *kMaxWaves) void adjoint_wide_kernel(
    AWParams p, int W)
{
    extern __shared__ double lds[];
    const ssmx_adjoint_desc &d = p.d;
    const int NP = W * kWave;
    const int row = threadIdx.x;
    const int lane = row & (kWave - 1);
    const int wave = row / kWave;
    const int inst = blockIdx.x;
    const int n = d.n;
    const int M = d.n_gen;
    const int J = d.n_exp;
    const int slots = d.slots;
    const bool wave_mode = (d.flags & SSMX_COEF_WAVE) != 0;

    // LDS: staged matrix values [m][slot / 2][row][slot % 2] (SF slots),
    // two exchange buffers kBuf apart, one partial sum per wavefront
    double *lv = lds;
    double *xb0 = lv + (size_t)M * SF * NP;
    double *red = xb0 + 2 * kBuf;
    const int staged = p.slots_f < SF ? p.slots_f : SF;
    for (int e = row; e < M * SF * NP; e += blockDim.x) {
        const int half = e & 1;
        const int r = (e >> 1) % NP;
        const int pair = (e / (2 * NP)) % (SF / 2);
        const int m = e / (NP * SF);
        const int slot = 2 * pair + half;
        lv[e] = slot < staged ? p.val[((size_t)m * slots + slot) * NP + r]
                              : 0.0;
    }
    __syncthreads();

    // extended forward pattern: gathers of the reverse sweep; its first SF
    // slots also serve the recomputation.  Slot 0's entry holds the row's
    // used slots (bits 0-7) and, if set, its matrix slots (bits 8-15).
    const int head = p.col[row];
    const int cnt_e = head & 0xFF;
    const int cnt_f = (head >> 8) & 0xFF;
    const double *gpe[SE];
    const double *gpf[SF];
#pragma unroll
    for (int sl = 0; sl < SE; ++sl) {
        const int c = sl == 0 ? row
                              : (sl < slots ? p.col[sl * NP + row] : row);
        gpe[sl] = xb0 + (c & (kBuf - 1));
    }
#pragma unroll
    for (int sl = 0; sl < SF; ++sl)
        gpf[sl] = gpe[sl];
    // slot classes of this wavefront: the matrices' widest row (capped at
    // the staged slots; columns beyond carry derivative entries only) for
    // the recomputation, the extended pattern for the forward-term chain
    int ws_f = wave_max(cnt_f > 0 ? cnt_f : cnt_e);
    ws_f = ws_f < staged ? ws_f : staged;
    const int ws_e = wave_max(cnt_e < slots ? cnt_e : slots);

    // transposed pattern: gathers, and where each A^T value lives among the
    // staged values (row c, slot s' of A with column r)
    const int cnt_t = p.col_t[row];
    const int ws_t = wave_max(cnt_t < d.slots_t ? cnt_t : d.slots_t);
    // packed offsets (two per register) of the A^T values in an m-plane of
    // the staged values; unused and missing entries name p.zero_entry
    const double *gpt[SF];
    uint32_t src[SF / 2];
#pragma unroll
    for (int sl = 0; sl < SF / 2; ++sl)
        src[sl] = 0;
#pragma unroll
    for (int sl = 0; sl < SF; ++sl) {
        int c = row;
        int s = -1;
        if (sl == 0) {
            s = 0;
        } else if (sl < d.slots_t && sl < cnt_t) {
            c = p.col_t[sl * NP + row];
            if (c < 0 || c >= NP)
                c = row;
            else if (c == (row ^ 1))
                s = 1;
            else
                for (int k = 2; k < staged; ++k)
                    if (p.col[k * NP + c] == row) {
                        s = k;
                        break;
                    }
        }
        gpt[sl] = xb0 + (c & (kBuf - 1));
        const uint32_t off = s < 0 ? (uint32_t)p.zero_entry
                                   : (uint32_t)((s >> 1) * 2 * NP + c * 2 +
                                                (s & 1));
        src[sl >> 1] |= off << (16 * (sl & 1));
    }

    double srow = 0.0;
    for (int o = 0; o < d.n_obs; ++o)
        srow = fma(p.seed[(size_t)inst * d.n_obs + o], p.obs[o * NP + row],
                   srow);
    const double inv_lane = kInvW.v[lane <= SSMX_MAX_TAYLOR ? lane : 0];

    double B[SE];
#pragma unroll
    for (int sl = 0; sl < SE; ++sl)
        B[sl] = 0.0;
    double lam = 0.5 * d.h * srow;

    const long T = (long)d.n_out * d.fine;
    const int K = d.ckpt_every;
    const long P = (T + K - 1) / K;
    const int JM = J * M;
    const int my_j = lane / M;
    const int my_m = lane % M;
    int cursor = 0;
    const double *coef_i =
        wave_mode ? nullptr : p.coef + (size_t)inst * d.coef_inst_stride;
    const ssmx_wave *wv = wave_mode ? p.waves + inst : nullptr;
    const int tg = d.theta_gen;
    ExpArgs e;
    e.lv = lv;
    e.NP = NP;
    e.M = M;
    e.base = 0;
    e.substeps = 1;
    e.taylor = d.taylor;
    e.inv_sub = 1.0;
    e.partner = true;
    int parity = 0;
    double *own = xb0 + row;
    double *ws = p.work + (size_t)inst * K * J * NP;

    for (long seg = P - 1; seg >= 0; --seg) {
        const long k0 = seg * K;
        const long k1 = T < k0 + K ? T : k0 + K;
        // 1. recomputation from the checkpoint
        double x = row < n ? p.ckpt[((size_t)inst * P + seg) * n + row]
                           : 0.0;
        for (long k = k0; k < k1; ++k) {
            const double factor =
                lane < JM ? adj_factor(p, coef_i, wv, k, my_j, my_m, cursor)
                          : 0.0;
            for (int j = J - 1; j >= 0; --j) {
                ws_store(ws + ((size_t)(k - k0) * J + j) * NP + row, x);
                e.base = j * M;
                // the loop form (same bits as the propagator's unrolled
                // chain): this kernel is at its register limit
                x = wide_exp_dispatch<SF, false>(ws_f, x, e, row, factor,
                                                 inv_lane, own, gpf, parity);
            }
        }
        if (parity) {
            // each reverse exponential starts on buffer 0 (and, with 2 MA
            // exchanges, ends ready for buffer 0): finish every read of it
            __syncthreads();
            parity = 0;
        }
        // 2. reverse sweep
        for (long k = k1 - 1; k >= k0; --k) {
            const double factor =
                lane < JM ? adj_factor(p, coef_i, wv, k, my_j, my_m, cursor)
                          : 0.0;
            for (int j = 0; j < J; ++j) {
                const double y =
                    ws_load(ws + ((size_t)(k - k0) * J + j) * NP + row);
                RevArgs r;
                r.lv = lv;
                r.NP = NP;
                r.M = M;
                r.base = j * M;
                r.row = row;
                r.factor = factor;
                r.fth = tg >= 0 ? read_lane(factor, j * M + tg) : 0.0;
                r.own = own;
                double v[MA];
                reverse_u_dispatch<SF, MA>(ws_t, lam, v, r, gpt, src);
                reverse_w_dispatch<SF, SE, MA>(ws_e, y, v, B, r, gpe);
            }
            lam = fma(k == 0 ? 0.5 * d.h : d.h, srow, lam);
        }
    }

    if (p.lambda_out && row < n)
        p.lambda_out[(size_t)inst * n + row] = lam;
    if (p.grad_theta) {
        for (int t = 0; t < d.n_theta; ++t) {
            double part = 0.0;
#pragma unroll
            for (int sl = 0; sl < SE; ++sl)
                if (sl < slots)
                    part = fma(B[sl],
                               p.dval[((size_t)t * slots + sl) * NP + row],
                               part);
            part = wave_sum(part);
            if (lane == 0)
                red[wave] = part;
            __syncthreads();
            if (row == 0) {
                double s = 0.0;
                for (int k = 0; k < W; ++k)
                    s += red[k];
                p.grad_theta[(size_t)inst * d.n_theta + t] = s;
            }
            __syncthreads();
        }
    }
}

// Offset (within one staged plane) of a value that is 0 for every matrix:
// slot SF - 1 when the matrices use fewer slots, else slot 0 of a padding
// row; -1 if there is none.
int zero_entry(int SF, int slots_f, int n, int NP)
{
    if (slots_f < SF)
        return ((SF - 1) >> 1) * 2 * NP + ((SF - 1) & 1);
    if (NP > n)
        return (NP - 1) * 2;
    return -1;
}

template <int SF, int SE, int MA>
int launch(AWParams p, hipStream_t stream)
{
    const int W = (p.d.n + kWave - 1) / kWave;
    const int NP = W * kWave;
    p.zero_entry = zero_entry(SF, p.slots_f, p.d.n, NP);
    if (p.zero_entry < 0 || p.zero_entry >= 65536)
        return set_error(SSMX_EINVAL,
                         "n=%d: no zero staged entry for %d matrix slots",
                         p.d.n, p.slots_f);
    const size_t lds =
        ((size_t)p.d.n_gen * SF * NP + 2 * kBuf + W) * sizeof(double);
    if (lds > 160 * 1024)
        return set_error(SSMX_EINVAL,
                         "n=%d: %d matrices x %d staged slots need %zu "
                         "bytes of LDS (> 160 KiB)",
                         p.d.n, p.d.n_gen, SF, lds);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void *>(&adjoint_wide_kernel<SF, SE, MA>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return set_error(SSMX_ELAUNCH, "hipFuncSetAttribute: %s",
                             hipGetErrorString(e));
    }
    hipLaunchKernelGGL((adjoint_wide_kernel<SF, SE, MA>), dim3(p.d.n_inst),
                       dim3(NP), lds, stream, p, W);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(SSMX_ELAUNCH, "wide adjoint launch: %s",
                         hipGetErrorString(e));
    return SSMX_OK;
}

template <int SF, int SE>
int pick_degree(const AWParams &p, hipStream_t s)
{
    // the planned degree, at least 8 (a higher degree only adds accuracy)
    if (p.d.taylor_adj <= 8)
        return launch<SF, SE, 8>(p, s);
    if (p.d.taylor_adj == 9)
        return launch<SF, SE, 9>(p, s);
    return launch<SF, SE, 10>(p, s);
}

}  // namespace

// Entry from ssmx_adjoint (ssmx_adjoint.hip) for n > 64; the common
// descriptor checks and pointers were validated there.
int ssmx_adjoint_wide(const ssmx_adjoint_desc *desc, const int32_t *ell_col,
                      const double *ell_val, const int32_t *ell_col_t,
                      const double *dval_t, const double *cf_weights,
                      const double *cf_nodes, const void *coef,
                      const ssmx_segment *segments, const double *ckpt,
                      const double *obs, const double *seed,
                      double *workspace, double *grad_theta,
                      double *lambda_out, hipStream_t stream)
{
    const bool wave_mode = (desc->flags & SSMX_COEF_WAVE) != 0;
    AWParams p;
    p.d = *desc;
    p.slots_f = desc->reserved > 0 ? desc->reserved : desc->slots;
    p.col = ell_col;
    p.val = ell_val;
    p.col_t = ell_col_t;
    p.dval = dval_t;
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
    p.lambda_out = lambda_out;
    const int sf = p.slots_f > desc->slots_t ? p.slots_f : desc->slots_t;
    const int se = desc->slots;
    const int NP = (desc->n + kWave - 1) / kWave * kWave;
    // staged slots SF (matrix and transposed) and extended slots SE; a
    // variant needs a zero staged entry (zero_entry)
    auto fits = [&](int SF) {
        return sf <= SF && zero_entry(SF, p.slots_f, desc->n, NP) >= 0;
    };
    if (fits(8) && se <= 16)
        return pick_degree<8, 16>(p, stream);
    if (fits(12) && se <= 18)
        return pick_degree<12, 18>(p, stream);
    if (fits(12) && se <= 20)
        return pick_degree<12, 20>(p, stream);
    if (fits(16) && se <= 22)
        return pick_degree<16, 22>(p, stream);
    if (fits(16) && se <= 32)
        return pick_degree<16, 32>(p, stream);
    return set_error(SSMX_EINVAL,
                     "n > 64 adjoint supports at most 16 matrix / transposed "
                     "slots (15 when n is a multiple of 64) and 32 extended "
                     "slots (got %d / %d / %d)",
                     p.slots_f, desc->slots_t, desc->slots);
}
