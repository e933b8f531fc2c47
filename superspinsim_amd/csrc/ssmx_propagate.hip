// ssmx_propagate.hip - batched commutator-free Magnus propagator for gfx950.
//
// One wavefront (64 lanes) advances one instance (one sweep point, one
// protocol) through n_out * fine propagator steps.  Lane r owns row r of the
// superoperator and component r of the vectorised density matrix; n <= 64.
//
// Per step of size h (reference: generate_simulator.py:1954-2015):
//   for j = J-1 .. 0:                      (CF product order, :1259-1294)
//     A_j = h sum_g c~_jg G_g, c~_jg = sum_s w[j,s] c[s,g]     (:278-448)
//       or, rotating frame, A_j = sum_s sum_g (c[s,g] h w[j,s]) G~[s,g]
//                                                             (:450-532)
//     x <- exp(A_j) x
//   rotating frame: x <- exp(L0 h) x (2x2 / 1x1 blocks, :902-1021)
// The reference forms exp(A_j) as a matrix by 2R = 70 residual squarings of
// a first-order Taylor step and multiplies whole operators; here the action
// exp(A_j) x is evaluated directly by a degree-m Taylor series on the state
// (optionally in 2^k substeps), which is exact to the same tolerance and
// needs m sparse matrix-vector products instead of 70 dense products.
//
// Data movement per Taylor term (the hot loop):
//  - slot 0 of every ELL row is the diagonal: no exchange;
//  - slot 1 is the partner column r ^ 1 (the Re/Im pair of a coherence, or
//    a 2x2 eigen-block): exchanged with a DPP quad permutation (VALU);
//  - slots 2.. are arbitrary columns: each lane writes its component to a
//    per-wavefront LDS buffer once, then reads its columns with ds_read_b64
//    (LDS instructions of one wavefront execute in order, so no barrier is
//    needed; half as many LDS instructions as two 32-bit ds_bpermute per
//    double);
//  - in the quad layout (slots 2, 3 of every row are the other pair of its
//    quad) those two come by DPP as well; in the pair-fetch layout (slots
//    2 + 2k, 3 + 2k of every row are the two columns of an aligned pair
//    2p, 2p + 1) each slot pair comes from one 16-byte LDS read
//    (ssmx::ell_apply).  The host picks the basis order; the kernel
//    recognises the layout from the column table.
// Coefficients are evaluated lane-parallel in batches: lane e of a batch
// computes the scalar factor of (step, exponential j, matrix m); the factors
// are broadcast to the wavefront with readlane when the step runs.
// Generator values live in VGPRs when they fit, else in LDS shared by the
// workgroup's wavefronts.

#include "ssmx_common.h"

namespace {

using ssmx::assemble_row;
using ssmx::assemble_row_mix;
using ssmx::assemble_row_reg;
using ssmx::ell_apply;
using ssmx::keep_branch;
using ssmx::kUnrolledMax;
using ssmx::kUnrolledTaylor;
using ssmx::kWave;
using ssmx::node_time;
using ssmx::pad_column;
using ssmx::fetched_pairs;
using ssmx::read_lane;
using ssmx::set_error;
using ssmx::taylor_action;
using ssmx::unit_weight_sum;
using ssmx::swap_pair;
using ssmx::wave_coef;
using ssmx::wave_max;
using ssmx::wave_sum;

constexpr int kWavesPerBlock = 8;
// readout rows staged in LDS (obs readouts hold 1..2)
constexpr int kObsRows = 2;

// 1 / q! for the Taylor loop, q = 0..40 (lane q reads entry q).
__constant__ ssmx::InvFactorials kInv =
    ssmx::make_inv_factorials();

struct Params {
    ssmx_propagate_desc d;
    const int32_t *ell_col;
    const double *ell_val;
    const double *w;
    const double *nodes;
    const double *coef;
    const ssmx_wave *waves;
    const ssmx_segment *segs;
    const double *rho0;
    const double *eig;
    const double *obs;
    const double *obs_offset;
    double *out;
    double *state_out;
};

// Scalar factor of matrix m in exponential j at global step `step`:
//   non-rotating (m = g):      h * sum_s w[j,s] c_g(t_s)
//   rotating (m = s * K + g):  c_g(t_s) * (h * w[j,s])
__device__ double step_factor(const Params &p, long step, int j, int m,
                              int &cursor)
{
    const ssmx_propagate_desc &d = p.d;
    const int S = d.n_node, K = d.n_gen;
    const long i_loc = step / d.fine;
    const int r = (int)(step - i_loc * d.fine);
    const long ig = (long)d.first_out + i_loc;
    const bool wave_mode = (d.flags & SSMX_COEF_WAVE) != 0;
    const bool rotating = (d.flags & SSMX_ROTATING) != 0;
    const double *row = nullptr;
    if (!wave_mode)
        row = p.coef + (size_t)0 + ((size_t)(i_loc * d.fine + r) * S) * K;
    if (rotating) {
        const int s = m / K, g = m % K;
        const double c =
            wave_mode ? wave_coef(p.waves, p.segs, g,
                                  node_time(d.t0, d.dt, d.fine, r, ig,
                                            p.nodes[s]),
                                  cursor)
                      : row[s * K + g];
        return c * (d.h * p.w[j * S + s]);
    }
    double cw = 0.0;
    for (int s = 0; s < S; ++s) {
        const double c =
            wave_mode ? wave_coef(p.waves, p.segs, m,
                                  node_time(d.t0, d.dt, d.fine, r, ig,
                                            p.nodes[s]),
                                  cursor)
                      : row[s * K + m];
        cw = fma(p.w[j * S + s], c, cw);
    }
    return d.h * cw;
}

// The lane's ELL columns.  Slot 0 is the diagonal by definition
// (ell_col[0][row] holds the row's used-slot count); slot 1 is the partner
// r ^ 1 when `partner` is set; other slots hold column indices.  Unused
// slots point at the row itself with value 0, so every slot is read without
// predication.  The empty asm with a memory clobber keeps each call's loads
// its own: the kernel reads the columns to choose the layout variant, the
// variant again for its loops, and no copy stays live (or spilled) across
// the choice.
template <int SLOTS>
__device__ __forceinline__ void load_columns(const Params &p, int lane,
                                             int (&col)[SLOTS])
{
    asm volatile("" ::: "memory");
    col[0] = lane;
#pragma unroll
    for (int s = 1; s < SLOTS; ++s)
        col[s] = s < p.d.slots ? p.ell_col[s * kWave + lane]
                               : pad_column(s, lane);
}

// One instance on one wavefront, after the kernel has staged the generator
// values (`vals`, [m][slot / 2][lane][slot % 2]) and loaded the lane's ELL
// columns.  NPAIR: leading slot pairs fetched whole (ell_apply).  MG > 0:
// the MG (= M) generator rows are read from LDS once into registers;
// MG < 0: the first -MG rows are, the others are read at each exponent.
// UNIT0 (MG = 2 or -2, two exponentials, one substep, lab frame, waveform
// mode):
// matrix 0's coefficient is 1 at every node (the folded constant
// generators), so its share of each exponent, fma(f0_j, G0, 0), is formed
// once per launch; f0_j is computed as step_factor computes it, so the row
// is the same bits.  The
// caller checks that f0_0 == f0_1 (CF4:2's two weight rows hold the same
// two numbers, and their sums are equal), so one row serves both
// exponentials.
template <int SLOTS, int NPAIR, int TAY, int MG = 0, bool STORE = false,
          bool UNIT0 = false>
__device__ __forceinline__ void propagate_instance(const Params &p,
                                                   const double *vals,
                                                   double *xbuf, int M,
                                                   int inst)
{
    const ssmx_propagate_desc &d = p.d;
    const int lane = threadIdx.x & (kWave - 1);
    // the columns read again here, so that the kernel's copy (read to
    // choose the layout variant) is dead before the loops
    int col[SLOTS];
    load_columns<SLOTS>(p, lane, col);
    const bool rotating = !UNIT0 && (d.flags & SSMX_ROTATING) != 0;
    const int n = d.n;
    const int J = d.n_exp;
    const bool wave_mode = (d.flags & SSMX_COEF_WAVE) != 0;
    const bool read_full = (d.flags & SSMX_READ_FULL) != 0;
    const bool read_int = (d.flags & SSMX_READ_INTEGRAL) != 0;
    const bool use_partner = d.partner != 0;

    constexpr int MGR = MG < 0 ? -MG : MG;
    // UNIT0 with three or more matrices: when each row of the wavefront is
    // non-zero in at most one of matrices 1..M-1 (a packed system's copies
    // of a varying generator, one per member block), the lane's own matrix
    // 1 + beta is the one held in g[1], and the exponent row is
    // fma(f_(1+beta), G_(1+beta), cst): the other matrices would add exact
    // zeros to it.  Wave-uniform; found once from the staged values.
    int beta = 0;
    bool disjoint = false;
    if constexpr (UNIT0 && MG < 0) {
        constexpr int SLOTP = SLOTS + (SLOTS & 1);
        int planes = 0;
        for (int m = 1; m < M; ++m) {
            bool nz = false;
#pragma unroll
            for (int sl = 0; sl < SLOTS; ++sl)
                nz |= vals[((m * (SLOTP / 2) + sl / 2) * kWave + lane) * 2 +
                           (sl & 1)] != 0.0;
            if (nz) {
                beta = m - 1;
                ++planes;
            }
        }
        disjoint = wave_max(planes) <= 1;
        if (!disjoint)
            beta = 0;
    }
    double g[MGR > 0 ? MGR : 1][SLOTS];
    if constexpr (MGR > 0) {
        constexpr int SLOTP = SLOTS + (SLOTS & 1);
#pragma unroll
        for (int m = 0; m < MGR; ++m) {
            const int plane = m == 1 ? 1 + beta : m;
#pragma unroll
            for (int sl = 0; sl < SLOTS; ++sl)
                g[m][sl] =
                    vals[((plane * (SLOTP / 2) + sl / 2) * kWave + lane) * 2 +
                         (sl & 1)];
        }
    }
    static_assert(!UNIT0 || MG == 2 || MG == -2,
                  "UNIT0 needs matrices 0 and 1 in registers");
    double cst[SLOTS];
    if constexpr (UNIT0) {
        const double f0 = d.h * unit_weight_sum(p.w, 0, d.n_node);
#pragma unroll
        for (int sl = 0; sl < SLOTS; ++sl)
            cst[sl] = fma(f0, g[0][sl], 0.0);
    }

    // Per-instance views.
    Params q = p;
    if (wave_mode)
        q.waves = p.waves + inst;
    else
        q.coef = p.coef + (size_t)inst * d.coef_inst_stride;

    // Lane t holds 1/t: the Taylor recursion broadcasts it with readlane
    // (a scalar memory load or a division would sit in the dependent
    // chain of every term).
    const double inv_lane = kInv.v[lane <= SSMX_MAX_TAYLOR ? lane : 0];

    double x = lane < n ? p.rho0[(size_t)inst * n + lane] : 0.0;

    double eig_own = 0.0, eig_part = 0.0;
    if (rotating) {
        eig_own = p.eig[lane];
        eig_part = p.eig[kWave + lane];
    }
    // observable k on state v: its row is read from LDS at each readout
    // (once per output interval; held in registers across the step loop
    // the rows would spill), its wave-uniform offset once
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    const double *obs_rows = vals + M * SLOTP * kWave + kWavesPerBlock * kWave;
    double off0 = 0.0, off1 = 0.0;
    if (!read_full && p.obs_offset) {
        off0 = p.obs_offset[(size_t)inst * d.n_obs];
        if (d.n_obs > 1)
            off1 = p.obs_offset[(size_t)inst * d.n_obs + 1];
    }
    auto observe = [&](int k, double v) {
        return wave_sum(obs_rows[k * kWave + lane] * v) +
               (k == 0 ? off0 : off1);
    };

    // Lane-parallel coefficient batches: entry e = (b * J + j) * M + m.
    // UNIT0 reads the factors of matrices 1..M-1 only (matrix 0's share is
    // the constant row): entry e = (b * J + j) * (M - 1) + m - 1, more steps
    // per batch
    const int Mb = UNIT0 ? (MG == 2 ? 1 : M - 1) : M;
    const int JM = J * Mb;
    const int B = kWave / JM;  // steps per batch (validated >= 1)
    const long total = (long)d.n_out * d.fine;
    const int my_b = lane / JM;
    const int my_j = (lane % JM) / Mb;
    const int my_m = UNIT0 ? 1 + lane % Mb : lane % M;
    int cursor = 0;
    double factor = 0.0;

    // UNIT0 is dispatched with one substep only (no scaling of the row,
    // no substep loop in the step)
    const int substeps = UNIT0 ? 1 : d.substeps;
    const double inv_sub = 1.0 / (double)substeps;
    double xint = 0.0;
    const double half_h = 0.5 * d.h;

    // Checkpoints: with ckpt_every = K > 0, state_out receives the state
    // before steps 0, K, 2K, ... instead of the final state.
    const int ckpt_every = (!STORE && p.state_out) ? d.ckpt_every : 0;
    const long n_ckpt = ckpt_every > 0 ? (total + ckpt_every - 1) / ckpt_every
                                       : 0;
    int ckpt_count = 0;
    long ckpt_index = 0;

    long step = 0;
    int b = 0;
    for (int i = 0; i < d.n_out; ++i) {
        for (int r = 0; r < d.fine; ++r, ++step, ++b) {
            if (b == B)
                b = 0;
            if (ckpt_every > 0) {
                if (ckpt_count == 0 && lane < n)
                    p.state_out[((size_t)inst * n_ckpt + ckpt_index) * n +
                                lane] = x;
                if (ckpt_count == 0)
                    ++ckpt_index;
                if (++ckpt_count == ckpt_every)
                    ckpt_count = 0;
            }
            if (b == 0) {
                const long mine = step + my_b;
                factor = (my_b < B && mine < total)
                             ? step_factor(q, mine, my_j, my_m, cursor)
                             : 0.0;
            }
            const double x_prev = x;
            // exponential j with its row a assembled
            auto exponential = [&](int j, double (&a)[SLOTS]) {
                if (substeps > 1) {
                    keep_branch();
#pragma unroll
                    for (int sl = 0; sl < SLOTS; ++sl)
                        a[sl] *= inv_sub;
                }
                if constexpr (STORE) {
                    // SSMX_STORE_INPUTS: every exponential's input, for the
                    // adjoint (stride 64 per exponential)
                    p.state_out[(((size_t)inst * total + step) * J + j) *
                                    kWave +
                                lane] = x;
                }
                // the adjoint's recomputation repeats these roundings
                for (int sub = 0; sub < substeps; ++sub)
                    x = taylor_action<SLOTS, NPAIR, TAY>(
                        a, col, x, xbuf, lane, use_partner, d.taylor,
                        inv_lane);
            };
            if constexpr (UNIT0) {
                // one substep: the row of A_j is assembled inside the first
                // product, behind the LDS store of x and its gathers
#pragma unroll
                for (int jj = 1; jj >= 0; --jj) {
                    double a[SLOTS];
                    auto assemble = [&]() {
                        const int base = (b * 2 + jj) * Mb;
                        if constexpr (MG < 0) {
                            if (disjoint) {
                                // the lane's block factor, across lanes
                                const double f =
                                    __shfl(factor, base + beta, kWave);
#pragma unroll
                                for (int sl = 0; sl < SLOTS; ++sl)
                                    a[sl] = fma(f, g[1][sl], cst[sl]);
                                return;
                            }
                        }
                        const double f = read_lane(factor, base);
#pragma unroll
                        for (int sl = 0; sl < SLOTS; ++sl)
                            a[sl] = fma(f, g[1][sl], cst[sl]);
                        if constexpr (MG < 0) {
                            // matrices 2..M-1 from the staged values, in
                            // assemble_row_mix's order; a factor of 0 (the
                            // generator off in this step) would add exact
                            // zeros (the row is never -0)
                            for (int m = 2; m < M; ++m) {
                                const double fm =
                                    read_lane(factor, base + m - 1);
                                if (fm == 0.0)
                                    continue;
                                const double2 *v =
                                    reinterpret_cast<const double2 *>(
                                        vals + m * SLOTP * kWave) +
                                    lane;
#pragma unroll
                                for (int pr = 0; pr < SLOTP / 2; ++pr) {
                                    const double2 vv = v[pr * kWave];
                                    a[2 * pr] = fma(fm, vv.x, a[2 * pr]);
                                    if (2 * pr + 1 < SLOTS)
                                        a[2 * pr + 1] =
                                            fma(fm, vv.y, a[2 * pr + 1]);
                                }
                            }
                        }
                    };
                    if constexpr (STORE)
                        p.state_out[(((size_t)inst * total + step) * J +
                                     jj) *
                                        kWave +
                                    lane] = x;
                    x = taylor_action<SLOTS, NPAIR, TAY>(
                        a, col, x, xbuf, lane, use_partner, d.taylor,
                        inv_lane, assemble);
                }
            } else {
                for (int j = J - 1; j >= 0; --j) {
                    double a[SLOTS];
                    if constexpr (MG > 0)
                        assemble_row_reg<SLOTS, MG>(a, g, factor,
                                                    (b * J + j) * M);
                    else if constexpr (MG < 0)
                        assemble_row_mix<SLOTS, MGR>(a, g, vals, factor,
                                                     (b * J + j) * M, M,
                                                     lane);
                    else
                        assemble_row(a, vals, factor, (b * J + j) * M, M,
                                     lane);
                    exponential(j, a);
                }
            }
            if (rotating) {
                const double xp = swap_pair(x);
                x = x + fma(eig_own, x, eig_part * xp);
            }
            if (read_int) {
                keep_branch();
                xint = fma(half_h, x_prev + x, xint);
            }
        }
        if (read_full) {
            if (lane < n)
                p.out[((size_t)inst * d.n_out + i) * n + lane] = x;
        } else if (!read_int) {
            const double v0 = observe(0, x);
            if (lane == 0)
                p.out[((size_t)inst * d.n_out + i) * d.n_obs] = v0;
            if (d.n_obs > 1) {
                const double v1 = observe(1, x);
                if (lane == 0)
                    p.out[((size_t)inst * d.n_out + i) * d.n_obs + 1] = v1;
            }
        }
    }
    if (read_int) {
        const double v0 = observe(0, xint);
        if (lane == 0)
            p.out[(size_t)inst * d.n_obs] = v0;
        if (d.n_obs > 1) {
            const double v1 = observe(1, xint);
            if (lane == 0)
                p.out[(size_t)inst * d.n_obs + 1] = v1;
        }
    }
    if (!STORE && p.state_out && ckpt_every == 0 && lane < n)
        p.state_out[(size_t)inst * n + lane] = x;
}

// CHAIN > 0: the kernel variant launched for Taylor degree CHAIN, with the
// chain unrolled for exactly that degree.  CHAIN = 0: the chain with
// per-term exits for degrees up to kUnrolledMax (SLOTS <= 16), else the
// runtime loop.  Separate kernels keep each chain's register allocation
// apart (both in one kernel ran every degree 1-2 % slower).
template <int SLOTS, int NPAIR, int CHAIN, int MG, bool STORE = false,
          bool UNIT0 = false>
__device__ __forceinline__ void propagate_instance_any(const Params &p, const double *vals, double *xbuf, int M, int inst)
{
    if constexpr (CHAIN > 0) {
        propagate_instance<SLOTS, NPAIR, CHAIN, MG, STORE, UNIT0>(
            p, vals, xbuf, M, inst);
    } else {
        if constexpr (SLOTS <= 16) {
            if (p.d.taylor >= 1 && p.d.taylor <= kUnrolledMax) {
                propagate_instance<SLOTS, NPAIR, -kUnrolledMax>(p, vals, xbuf, M, inst);
                return;
            }
        }
        propagate_instance<SLOTS, NPAIR, 0>(p, vals, xbuf, M, inst);
    }
}

// The instance's layout variant: slot pairs read whole (every one, the
// first, or none; other counts take the largest variant below them, which
// reads the same values).
template <int SLOTS, int CHAIN, int MG, bool STORE, bool UNIT0>
__device__ __forceinline__ void dispatch_layout(const Params &p,
                                                const double *lds,
                                                double *xbuf, int M, int inst,
                                                int np)
{
    if constexpr (SLOTS >= 4) {
        constexpr int kAll = (SLOTS - 2) / 2;
        if constexpr (kAll > 1 && SLOTS <= 16) {
            if (np >= kAll) {
                propagate_instance_any<SLOTS, kAll, CHAIN, MG, STORE, UNIT0>(
                    p, lds, xbuf, M, inst);
                return;
            }
        }
        if (np >= 1) {
            propagate_instance_any<SLOTS, 1, CHAIN, MG, STORE, UNIT0>(
                p, lds, xbuf, M, inst);
            return;
        }
    }
    propagate_instance_any<SLOTS, 0, CHAIN, MG, STORE, UNIT0>(p, lds, xbuf,
                                                              M, inst);
}

template <int SLOTS, int CHAIN, int MG, bool STORE = false>
// At most 128 VGPRs: four wavefronts per SIMD (4096 instances fill the
// device once).
__global__ __launch_bounds__(kWave *kWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(4))) void propagate_kernel(
    Params p)
{
    // LDS: generator values staged as [m][slot / 2][lane][slot % 2] (one
    // ds_read_b128 returns two slots of a row), one 64-double exchange
    // buffer per wavefront, then the kObsRows readout rows.
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    extern __shared__ double lds[];
    const ssmx_propagate_desc &d = p.d;
    const int lane = threadIdx.x & (kWave - 1);
    const bool rotating = (d.flags & SSMX_ROTATING) != 0;
    const int M = (rotating ? d.n_node : 1) * d.n_gen;  // matrices
    {
        const int nval = M * SLOTP * kWave;
        for (int e = threadIdx.x; e < nval; e += blockDim.x) {
            const int half = e & 1;
            const int row = (e >> 1) % kWave;
            const int pair = (e / (2 * kWave)) % (SLOTP / 2);
            const int m = e / (kWave * SLOTP);
            const int slot = 2 * pair + half;
            lds[e] = slot < d.slots
                         ? p.ell_val[((size_t)m * d.slots + slot) * kWave +
                                     row]
                         : 0.0;
        }
        if (!(d.flags & SSMX_READ_FULL)) {
            double *obs_rows = lds + nval + kWavesPerBlock * kWave;
            for (int e = threadIdx.x; e < d.n_obs * kWave; e += blockDim.x)
                obs_rows[e] = p.obs[e];
        }
        __syncthreads();
    }

    const int inst = __builtin_amdgcn_readfirstlane(
        (int)blockIdx.x * kWavesPerBlock + (int)(threadIdx.x / kWave));
    if (inst >= d.n_inst)
        return;  // whole wavefront; no barrier follows
    double *xbuf = lds + M * SLOTP * kWave + (threadIdx.x / kWave) * kWave;

    // slot pairs read whole: every one (rows laid out by pairs), the first
    // (the other pair of a row's quad) or none; other counts take the
    // largest variant below them, which reads the same values
    int np;
    {
        int col[SLOTS];
        load_columns<SLOTS>(p, lane, col);
        np = fetched_pairs(col, lane, d.partner != 0);
    }
    if constexpr ((MG == 2 || MG == -2) && CHAIN > 0) {
        // matrix 0 with coefficient 1 at every node: the instance's
        // waveform gives generator 0 the base 1 and no segments (the
        // folded constants); wave-uniform, read once
        if (!rotating && (d.flags & SSMX_COEF_WAVE) && d.n_exp == 2 &&
            d.substeps == 1 && p.waves[inst].seg_count[0] == 0 &&
            p.waves[inst].base[0] == 1.0 &&
            unit_weight_sum(p.w, 0, d.n_node) ==
                unit_weight_sum(p.w, 1, d.n_node)) {
            dispatch_layout<SLOTS, CHAIN, MG, STORE, true>(p, lds, xbuf, M,
                                                           inst, np);
            return;
        }
    }
    dispatch_layout<SLOTS, CHAIN, MG, STORE, false>(p, lds, xbuf, M, inst,
                                                    np);
}

template <int SLOTS, int CHAIN, int MG, bool STORE = false>
int launch_chain(const Params &p, hipStream_t stream)
{
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    const int M = ((p.d.flags & SSMX_ROTATING) ? p.d.n_node : 1) * p.d.n_gen;
    const size_t lds = ((size_t)M * SLOTP * kWave +
                        (size_t)(kWavesPerBlock + kObsRows) * kWave) *
                       sizeof(double);
    if (lds > 160 * 1024)
        return set_error(SSMX_EINVAL,
                         "generator values need %zu bytes of LDS (> 160 KiB)",
                         lds);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void *>(
                &propagate_kernel<SLOTS, CHAIN, MG, STORE>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return set_error(SSMX_ELAUNCH, "hipFuncSetAttribute: %s",
                             hipGetErrorString(e));
    }
    const int blocks = (p.d.n_inst + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL((propagate_kernel<SLOTS, CHAIN, MG, STORE>),
                       dim3(blocks),
                       dim3(kWave * kWavesPerBlock), lds, stream, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(SSMX_ELAUNCH, "propagate launch: %s",
                         hipGetErrorString(e));
    return SSMX_OK;
}

// The exact-degree kernel for the common degree, else the general one.
template <int SLOTS>
int launch(const Params &p, hipStream_t stream)
{
    const int M = ((p.d.flags & SSMX_ROTATING) ? p.d.n_node : 1) * p.d.n_gen;
    if (p.d.flags & SSMX_STORE_INPUTS) {
        // the adjoint's forward: degree 11, at most 8 slots (validated)
        if constexpr (SLOTS <= 8) {
            if (M == 2)
                return launch_chain<SLOTS, kUnrolledTaylor, 2, true>(p,
                                                                    stream);
            return launch_chain<SLOTS, kUnrolledTaylor, -2, true>(p, stream);
        }
        return set_error(SSMX_EINVAL, "SSMX_STORE_INPUTS needs slots <= 8");
    }
    if constexpr (SLOTS <= 8) {
        // two generator rows in registers: no LDS reads in the assembly
        if (p.d.taylor == kUnrolledTaylor && M == 2)
            return launch_chain<SLOTS, kUnrolledTaylor, 2>(p, stream);
        // more generators: the first two rows in registers
        if (p.d.taylor == kUnrolledTaylor && M > 2)
            return launch_chain<SLOTS, kUnrolledTaylor, -2>(p, stream);
    }
    if constexpr (SLOTS <= 16) {
        if (p.d.taylor == kUnrolledTaylor)
            return launch_chain<SLOTS, kUnrolledTaylor, 0>(p, stream);
    }
    return launch_chain<SLOTS, 0, 0>(p, stream);
}

int validate(const ssmx_propagate_desc *d)
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
    if (d->slots < 1 || d->slots > SSMX_MAX_N)
        return set_error(SSMX_EINVAL, "slots=%d outside 1..64", d->slots);
    if (d->n_exp < 1 || d->n_exp > SSMX_MAX_EXP)
        return set_error(SSMX_EINVAL, "n_exp=%d outside 1..%d", d->n_exp,
                         SSMX_MAX_EXP);
    if (d->n_node < 1 || d->n_node > SSMX_MAX_NODE)
        return set_error(SSMX_EINVAL, "n_node=%d outside 1..%d", d->n_node,
                         SSMX_MAX_NODE);
    const int M = ((d->flags & SSMX_ROTATING) ? d->n_node : 1) * d->n_gen;
    if (d->n_exp * M > kWave)
        return set_error(SSMX_EINVAL,
                         "n_exp * matrices = %d exceeds one batch of 64",
                         d->n_exp * M);
    if (d->fine < 1 || d->n_out < 0)
        return set_error(SSMX_EINVAL, "fine=%d, n_out=%d", d->fine, d->n_out);
    if (d->taylor < 1 || d->taylor > SSMX_MAX_TAYLOR)
        return set_error(SSMX_EINVAL, "taylor=%d outside 1..%d", d->taylor,
                         SSMX_MAX_TAYLOR);
    if (d->substeps < 1 || (d->substeps & (d->substeps - 1)) != 0 ||
        d->substeps > 1024)
        return set_error(SSMX_EINVAL,
                         "substeps=%d is not a power of two <= 1024",
                         d->substeps);
    if (!(d->flags & SSMX_READ_FULL) && (d->n_obs < 1 || d->n_obs > 2))
        return set_error(SSMX_EINVAL, "n_obs=%d outside 1..2", d->n_obs);
    if ((d->flags & SSMX_READ_FULL) && (d->flags & SSMX_READ_INTEGRAL))
        return set_error(SSMX_EINVAL,
                         "READ_FULL and READ_INTEGRAL are exclusive");
    if (d->flags & SSMX_STORE_INPUTS) {
        const int Mg = ((d->flags & SSMX_ROTATING) ? d->n_node : 1) *
                       d->n_gen;
        if (d->n > SSMX_MAX_N || (d->flags & SSMX_ROTATING) ||
            d->substeps != 1 || d->ckpt_every != 0 ||
            d->taylor != kUnrolledTaylor || d->slots > 8 || Mg < 2)
            return set_error(SSMX_EINVAL,
                             "SSMX_STORE_INPUTS needs n <= 64, the lab "
                             "frame, substeps 1, ckpt_every 0, Taylor degree "
                             "%d, at most 8 slots and 2 or more matrices",
                             kUnrolledTaylor);
    }
    if (!(d->h > 0.0) || !(d->dt > 0.0))
        return set_error(SSMX_EINVAL, "h and dt must be positive");
    if (d->first_out < 0)
        return set_error(SSMX_EINVAL, "first_out=%d < 0", d->first_out);
    if (d->partner != 0 && d->partner != 1)
        return set_error(SSMX_EINVAL, "partner flag must be 0 or 1");
    if (d->ckpt_every < 0)
        return set_error(SSMX_EINVAL, "ckpt_every=%d < 0", d->ckpt_every);
    if (d->n > SSMX_MAX_N) {
        // one row per thread: at most 32 slots, and the staged generator
        // values plus one instance's two 512-row exchange buffers within
        // 160 KiB of LDS
        if (d->slots > 32)
            return set_error(SSMX_EINVAL, "slots=%d > 32 for n > 64",
                             d->slots);
        const long np = (long)(d->n + kWave - 1) / kWave * kWave;
        const long slotp = d->slots + (d->slots & 1);
        const long lds =
            (M * slotp * np + 2 * SSMX_MAX_N_WIDE + np / kWave) * 8;
        if (lds > 160 * 1024)
            return set_error(SSMX_EINVAL,
                             "n=%d with %d matrices x %d slots needs %ld "
                             "bytes of LDS (> 160 KiB)",
                             d->n, M, d->slots, lds);
    }
    return SSMX_OK;
}

}  // namespace

// n > 64: one workgroup per instance (ssmx_wide.hip).
int ssmx_wide_propagate(const ssmx_propagate_desc *desc,
                        const int32_t *ell_col, const double *ell_val,
                        const double *cf_weights, const double *cf_nodes,
                        const void *coef, const ssmx_segment *segments,
                        const double *rho0, const double *eig,
                        const double *obs, const double *obs_offset,
                        double *out, double *state_out, hipStream_t stream);

extern "C" {

int ssmx_version(void) { return 103; }

const char *ssmx_last_error(void) { return ssmx::error_buffer(); }

int ssmx_device_count(void)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess)
        return set_error(SSMX_EDEVICE, "hipGetDeviceCount: %s",
                         hipGetErrorString(e));
    return count;
}

int ssmx_propagate(const ssmx_propagate_desc *desc, const int32_t *ell_col,
                   const double *ell_val, const double *cf_weights,
                   const double *cf_nodes, const void *coef,
                   const ssmx_segment *segments, const double *rho0,
                   const double *eig, const double *obs,
                   const double *obs_offset, double *out, double *state_out,
                   void *stream)
{
    ssmx::error_buffer()[0] = '\0';
    int rc = validate(desc);
    if (rc)
        return rc;
    const bool rotating = (desc->flags & SSMX_ROTATING) != 0;
    const bool wave_mode = (desc->flags & SSMX_COEF_WAVE) != 0;
    if (!ell_col || !ell_val || !cf_weights || !cf_nodes || !coef || !rho0 ||
        !out)
        return set_error(SSMX_EINVAL, "null required pointer");
    if (rotating && !eig)
        return set_error(SSMX_EINVAL, "rotating frame needs eig factors");
    if (wave_mode && !segments)
        return set_error(SSMX_EINVAL, "waveform mode needs segments");
    if (!(desc->flags & SSMX_READ_FULL) && !obs)
        return set_error(SSMX_EINVAL, "obs readout needs obs rows");
    if ((desc->flags & SSMX_STORE_INPUTS) && !state_out)
        return set_error(SSMX_EINVAL, "SSMX_STORE_INPUTS needs state_out");

    Params p;
    p.d = *desc;
    p.ell_col = ell_col;
    p.ell_val = ell_val;
    p.w = cf_weights;
    p.nodes = cf_nodes;
    p.coef = wave_mode ? nullptr : static_cast<const double *>(coef);
    p.waves = wave_mode ? static_cast<const ssmx_wave *>(coef) : nullptr;
    p.segs = segments;
    p.rho0 = rho0;
    p.eig = eig;
    p.obs = obs;
    p.obs_offset = obs_offset;
    p.out = out;
    p.state_out = state_out;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (desc->n > SSMX_MAX_N)
        return ssmx_wide_propagate(desc, ell_col, ell_val, cf_weights,
                                   cf_nodes, coef, segments, rho0, eig, obs,
                                   obs_offset, out, state_out, s);

    switch (desc->slots) {
    case 1: case 2: return launch<2>(p, s);
    case 3: return launch<3>(p, s);
    case 4: return launch<4>(p, s);
    case 5: return launch<5>(p, s);
    case 6: return launch<6>(p, s);
    case 7: return launch<7>(p, s);
    case 8: return launch<8>(p, s);
    case 9: return launch<9>(p, s);
    case 10: return launch<10>(p, s);
    case 11: case 12: return launch<12>(p, s);
    case 13: case 14: return launch<14>(p, s);
    case 15: case 16: return launch<16>(p, s);
    default: break;
    }
    if (desc->slots <= 20) return launch<20>(p, s);
    if (desc->slots <= 24) return launch<24>(p, s);
    if (desc->slots <= 32) return launch<32>(p, s);
    if (desc->slots <= 48) return launch<48>(p, s);
    return launch<64>(p, s);
}

}  // extern "C"
