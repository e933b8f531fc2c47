// ssmx_queue.hip - persistent work-queue propagator for batches of uneven
// protocols (gfx950).
//
// Every instance has its own number of propagator steps and its own readout
// window; the integrated readout F_o = sum over the window of the trapezoid
// (h/2)(obs_o . x_k + obs_o . x_{k+1}) is written per instance.  The grid is
// sized to the wavefronts the device holds at once; each wavefront takes the
// next instance from an atomic counter (in the caller's order, normally
// longest first) until the queue is empty, so the long protocols start first
// and the short ones fill the tail.
//
// The per-step arithmetic is the propagator's (ssmx_propagate.hip): same
// node times, coefficient evaluation, assembly, Taylor action, product order
// and rotating eigen-step, so an instance gives the same states here as in a
// launch of its own.

#include "ssmx_common.h"

#include <mutex>
#include <type_traits>

namespace {

using ssmx::assemble_row;
using ssmx::assemble_row_mix;
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
using ssmx::wave_sum;

constexpr int kMaxWavesPerBlock = 8;

// 1 / q! for the Taylor loop, q = 0..40 (lane q reads entry q).
__constant__ ssmx::InvFactorials kInvQ =
    ssmx::make_inv_factorials();

struct QParams {
    ssmx_propagate_desc d;
    const int32_t *ell_col;
    const double *ell_val;
    const double *w;
    const double *nodes;
    const ssmx_wave *waves;
    const ssmx_segment *segs;
    const double *rho0;
    const double *eig;
    const double *obs;
    const int32_t *order;
    const int64_t *steps;
    const int64_t *read_from;
    int32_t *counter;
    double *out;
};

// Scalar factor of matrix m in exponential j at step `step` of the
// instance whose waveform is `wv` (the propagator's step_factor, wave mode).
__device__ double step_factor_q(const QParams &p, const ssmx_wave *wv,
                                long step, int j, int m, int &cursor)
{
    const ssmx_propagate_desc &d = p.d;
    const int S = d.n_node, K = d.n_gen;
    const long i_loc = step / d.fine;
    const int r = (int)(step - i_loc * d.fine);
    const bool rotating = (d.flags & SSMX_ROTATING) != 0;
    if (rotating) {
        const int s = m / K, g = m % K;
        const double c = wave_coef(
            wv, p.segs, g, node_time(d.t0, d.dt, d.fine, r, i_loc, p.nodes[s]),
            cursor);
        return c * (d.h * p.w[j * S + s]);
    }
    double cw = 0.0;
    for (int s = 0; s < S; ++s) {
        const double c = wave_coef(
            wv, p.segs, m, node_time(d.t0, d.dt, d.fine, r, i_loc, p.nodes[s]),
            cursor);
        cw = fma(p.w[j * S + s], c, cw);
    }
    return d.h * cw;
}

// The queue loop of one wavefront, after the kernel has staged the
// generator values (`vals`) and loaded the lane's ELL columns.  NPAIR: leading
// slot pairs fetched whole (ell_apply).  MG > 0: the first MG generator
// rows are read from LDS once into registers, the others at each exponent.
template <int SLOTS, int NPAIR, int TAY, int MG = 0>
__device__ __forceinline__ void queue_loop(const QParams &p,
                                           const double *vals, double *xbuf,
                                           int M, const int (&col)[SLOTS])
{
    const ssmx_propagate_desc &d = p.d;
    const int lane = threadIdx.x & (kWave - 1);
    const bool rotating = (d.flags & SSMX_ROTATING) != 0;
    const int n = d.n;
    const int J = d.n_exp;
    const bool use_partner = d.partner != 0;
    const double inv_lane = kInvQ.v[lane <= SSMX_MAX_TAYLOR ? lane : 0];
    double g[MG > 0 ? MG : 1][SLOTS];
    if constexpr (MG > 0) {
        constexpr int SLOTP = SLOTS + (SLOTS & 1);
#pragma unroll
        for (int m = 0; m < MG; ++m)
#pragma unroll
            for (int sl = 0; sl < SLOTS; ++sl)
                g[m][sl] = vals[((m * (SLOTP / 2) + sl / 2) * kWave + lane) *
                                    2 +
                                (sl & 1)];
    }
    double eig_own = 0.0, eig_part = 0.0;
    if (rotating) {
        eig_own = p.eig[lane];
        eig_part = p.eig[kWave + lane];
    }
    const double obs0 = p.obs[lane];
    const double obs1 = d.n_obs > 1 ? p.obs[kWave + lane] : 0.0;
    const double inv_sub = 1.0 / (double)d.substeps;
    const double half_h = 0.5 * d.h;

    // Matrix 0 with coefficient 1 at every node (the folded constant
    // generators: the protocol's waveform gives generator 0 the base 1 and
    // no segments): its share of both exponents, fma(f0, G0, 0), is formed
    // once per wavefront when the two weight sums are equal (CF4:2), as in
    // ssmx_propagate.  f0 is the step factor's value, so the rows are the
    // same bits.
    constexpr bool kUnitPath = MG == 2 && TAY > 0;
    double cst[SLOTS];
    bool unit_ok = false;
    if constexpr (kUnitPath) {
        if (!rotating && J == 2 && d.substeps == 1 &&
            unit_weight_sum(p.w, 0, d.n_node) ==
                unit_weight_sum(p.w, 1, d.n_node)) {
            unit_ok = true;
            const double f0 = d.h * unit_weight_sum(p.w, 0, d.n_node);
#pragma unroll
            for (int sl = 0; sl < SLOTS; ++sl)
                cst[sl] = fma(f0, g[0][sl], 0.0);
        }
    }

    // One protocol.  UNIT: the constant-row path (one substep, factors of
    // matrices 1..M-1 only in the coefficient batch, each exponent's row
    // assembled inside its first product); otherwise the general path.
    auto run = [&](auto unit_tag, int inst, const ssmx_wave *wv) {
        constexpr bool UNIT = decltype(unit_tag)::value;
        const int Mb = UNIT ? M - 1 : M;
        const int JM = J * Mb;
        const int B = kWave / JM;
        const int my_b = lane / JM;
        const int my_j = (lane % JM) / Mb;
        const int my_m = UNIT ? 1 + lane % Mb : lane % M;
        const long total = p.steps[inst];
        const long k_from = p.read_from[inst];
        double x = lane < n ? p.rho0[(size_t)inst * n + lane] : 0.0;
        double xint = 0.0;
        int cursor = 0;
        double factor = 0.0;
        int b = 0;
        for (long step = 0; step < total; ++step, ++b) {
            if (b == B)
                b = 0;
            if (b == 0) {
                const long mine = step + my_b;
                factor = (my_b < B && mine < total)
                             ? step_factor_q(p, wv, mine, my_j, my_m, cursor)
                             : 0.0;
            }
            const double x_prev = x;
            if constexpr (UNIT) {
#pragma unroll
                for (int jj = 1; jj >= 0; --jj) {
                    double a[SLOTS];
                    auto assemble = [&]() {
                        const int base = (b * 2 + jj) * Mb;
                        const double f = read_lane(factor, base);
#pragma unroll
                        for (int sl = 0; sl < SLOTS; ++sl)
                            a[sl] = fma(f, g[1][sl], cst[sl]);
                        // matrices 2..M-1 from the staged values, in
                        // assemble_row_mix's order; a matrix whose factor
                        // is 0 (a generator switched off in this step, such
                        // as the readout's excitation before its window)
                        // would add exact zeros (the row is never -0)
                        constexpr int SLOTP = SLOTS + (SLOTS & 1);
                        for (int m = 2; m < M; ++m) {
                            const double fm = read_lane(factor, base + m - 1);
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
                    };
                    x = taylor_action<SLOTS, NPAIR, TAY>(
                        a, col, x, xbuf, lane, use_partner, d.taylor,
                        inv_lane, assemble);
                }
            } else {
                for (int j = J - 1; j >= 0; --j) {
                    double a[SLOTS];
                    if constexpr (MG > 0)
                        assemble_row_mix<SLOTS, MG>(a, g, vals, factor,
                                                    (b * J + j) * M, M, lane);
                    else
                        assemble_row(a, vals, factor, (b * J + j) * M, M,
                                     lane);
                    if (d.substeps > 1) {
                        keep_branch();
#pragma unroll
                        for (int sl = 0; sl < SLOTS; ++sl)
                            a[sl] *= inv_sub;
                    }
                    // the adjoint's recomputation repeats these roundings
                    for (int sub = 0; sub < d.substeps; ++sub)
                        x = taylor_action<SLOTS, NPAIR, TAY>(
                            a, col, x, xbuf, lane, use_partner, d.taylor,
                            inv_lane);
                }
            }
            if (rotating) {
                const double xp = swap_pair(x);
                x = x + fma(eig_own, x, eig_part * xp);
            }
            if (step >= k_from)
                xint = fma(half_h, x_prev + x, xint);
        }
        const double v0 = wave_sum(obs0 * xint);
        if (lane == 0)
            p.out[(size_t)inst * d.n_obs] = v0;
        if (d.n_obs > 1) {
            const double v1 = wave_sum(obs1 * xint);
            if (lane == 0)
                p.out[(size_t)inst * d.n_obs + 1] = v1;
        }
    };

    // Every pass takes one queue entry; the counter only grows, so every
    // wavefront leaves the loop once the n_inst entries are handed out.
    for (;;) {
        int q = 0;
        if (lane == 0)
            q = atomicAdd(p.counter, 1);
        q = __builtin_amdgcn_readfirstlane(q);
        if (q >= d.n_inst)
            break;
        const int inst = p.order[q];
        const ssmx_wave *wv = p.waves + inst;
        if constexpr (kUnitPath) {
            if (unit_ok && wv->seg_count[0] == 0 && wv->base[0] == 1.0) {
                run(std::true_type{}, inst, wv);
                continue;
            }
        }
        run(std::false_type{}, inst, wv);
    }
}

// CHAIN > 0: the kernel variant launched for Taylor degree CHAIN, with the
// chain unrolled for exactly that degree.  CHAIN = 0: the chain with
// per-term exits for degrees up to kUnrolledMax (SLOTS <= 16), else the
// runtime loop.  Separate kernels keep each chain's register allocation
// apart (both in one kernel ran every degree 1-2 % slower).
template <int SLOTS, int NPAIR, int CHAIN, int MG>
__device__ __forceinline__ void queue_loop_any(const QParams &p, const double *vals, double *xbuf, int M, const int (&col)[SLOTS])
{
    if constexpr (CHAIN > 0) {
        queue_loop<SLOTS, NPAIR, CHAIN, MG>(p, vals, xbuf, M, col);
    } else {
        if constexpr (SLOTS <= 16) {
            if (p.d.taylor >= 1 && p.d.taylor <= kUnrolledMax) {
                queue_loop<SLOTS, NPAIR, -kUnrolledMax>(p, vals, xbuf, M, col);
                return;
            }
        }
        queue_loop<SLOTS, NPAIR, 0>(p, vals, xbuf, M, col);
    }
}

template <int SLOTS, int CHAIN, int MG>
// At most 128 VGPRs: the grid is sized for four wavefronts per SIMD.
__global__ __launch_bounds__(kWave *kMaxWavesPerBlock)
__attribute__((amdgpu_waves_per_eu(4))) void queue_kernel(
    QParams p)
{
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    extern __shared__ double lds[];
    const ssmx_propagate_desc &d = p.d;
    const int lane = threadIdx.x & (kWave - 1);
    const bool rotating = (d.flags & SSMX_ROTATING) != 0;
    const int M = (rotating ? d.n_node : 1) * d.n_gen;
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
        __syncthreads();
    }
    double *xbuf = lds + M * SLOTP * kWave + (threadIdx.x / kWave) * kWave;

    const bool use_partner = d.partner != 0;
    int col[SLOTS];
    col[0] = lane;
#pragma unroll
    for (int s = 1; s < SLOTS; ++s)
        col[s] = s < d.slots ? p.ell_col[s * kWave + lane]
                             : pad_column(s, lane);
    // slot pairs read whole: every one (rows laid out by pairs), the first
    // (the other pair of a row's quad) or none; other counts take the
    // largest variant below them, which reads the same values
    const int np = fetched_pairs(col, lane, use_partner);
    if constexpr (SLOTS >= 4) {
        constexpr int kAll = (SLOTS - 2) / 2;
        if constexpr (kAll > 1 && SLOTS <= 16) {
            if (np >= kAll) {
                queue_loop_any<SLOTS, kAll, CHAIN, MG>(p, lds, xbuf, M, col);
                return;
            }
        }
        if (np >= 1) {
            queue_loop_any<SLOTS, 1, CHAIN, MG>(p, lds, xbuf, M, col);
            return;
        }
    }
    queue_loop_any<SLOTS, 0, CHAIN, MG>(p, lds, xbuf, M, col);
}

template <int SLOTS, int CHAIN, int MG>
int launch_chain(const QParams &p, int n_waves, int waves_per_block,
                 hipStream_t stream)
{
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    const int M = ((p.d.flags & SSMX_ROTATING) ? p.d.n_node : 1) * p.d.n_gen;
    const size_t lds = ((size_t)M * SLOTP * kWave +
                        (size_t)waves_per_block * kWave) *
                       sizeof(double);
    if (lds > 160 * 1024)
        return set_error(SSMX_EINVAL,
                         "generator values need %zu bytes of LDS (> 160 KiB)",
                         lds);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(
            reinterpret_cast<const void *>(&queue_kernel<SLOTS, CHAIN, MG>),
            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return set_error(SSMX_ELAUNCH, "hipFuncSetAttribute: %s",
                             hipGetErrorString(e));
    }
    const int blocks = (n_waves + waves_per_block - 1) / waves_per_block;
    hipLaunchKernelGGL((queue_kernel<SLOTS, CHAIN, MG>), dim3(blocks),
                       dim3(kWave * waves_per_block), lds, stream, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(SSMX_ELAUNCH, "queue launch: %s",
                         hipGetErrorString(e));
    return SSMX_OK;
}

// The exact-degree kernel for the common degree, else the general one.
template <int SLOTS>
int launch(const QParams &p, int n_waves, int waves_per_block,
           hipStream_t stream)
{
    const int M = ((p.d.flags & SSMX_ROTATING) ? p.d.n_node : 1) * p.d.n_gen;
    if constexpr (SLOTS <= 8) {
        // two generator rows in registers, the others from LDS
        if (p.d.taylor == kUnrolledTaylor && M >= 2)
            return launch_chain<SLOTS, kUnrolledTaylor, 2>(
                p, n_waves, waves_per_block, stream);
    }
    if constexpr (SLOTS <= 16) {
        if (p.d.taylor == kUnrolledTaylor)
            return launch_chain<SLOTS, kUnrolledTaylor, 0>(
                p, n_waves, waves_per_block, stream);
    }
    return launch_chain<SLOTS, 0, 0>(p, n_waves, waves_per_block, stream);
}

int validate(const ssmx_propagate_desc *d, int n_waves, int waves_per_block)
{
    if (!d)
        return set_error(SSMX_EINVAL, "null descriptor");
    if (d->n < 1 || d->n > SSMX_MAX_N)
        return set_error(SSMX_EINVAL, "n=%d outside 1..%d", d->n, SSMX_MAX_N);
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
    if (d->fine < 1)
        return set_error(SSMX_EINVAL, "fine=%d < 1", d->fine);
    if (d->taylor < 1 || d->taylor > SSMX_MAX_TAYLOR)
        return set_error(SSMX_EINVAL, "taylor=%d outside 1..%d", d->taylor,
                         SSMX_MAX_TAYLOR);
    if (d->substeps < 1 || (d->substeps & (d->substeps - 1)) != 0 ||
        d->substeps > 1024)
        return set_error(SSMX_EINVAL,
                         "substeps=%d is not a power of two <= 1024",
                         d->substeps);
    if (d->n_obs < 1 || d->n_obs > 2)
        return set_error(SSMX_EINVAL, "n_obs=%d outside 1..2", d->n_obs);
    if ((d->flags & ~(SSMX_ROTATING | SSMX_COEF_WAVE | SSMX_READ_INTEGRAL)) ||
        !(d->flags & SSMX_COEF_WAVE) || !(d->flags & SSMX_READ_INTEGRAL))
        return set_error(SSMX_EINVAL,
                         "queue mode needs SSMX_COEF_WAVE | "
                         "SSMX_READ_INTEGRAL (optionally SSMX_ROTATING)");
    if (!(d->h > 0.0) || !(d->dt > 0.0))
        return set_error(SSMX_EINVAL, "h and dt must be positive");
    if (d->partner != 0 && d->partner != 1)
        return set_error(SSMX_EINVAL, "partner flag must be 0 or 1");
    if (n_waves < 1)
        return set_error(SSMX_EINVAL, "n_waves=%d < 1", n_waves);
    if (waves_per_block < 1 || waves_per_block > kMaxWavesPerBlock)
        return set_error(SSMX_EINVAL, "waves_per_block=%d outside 1..%d",
                         waves_per_block, kMaxWavesPerBlock);
    return SSMX_OK;
}

int queue_launch(const ssmx_propagate_desc *desc, const int32_t *ell_col,
                 const double *ell_val, const double *cf_weights,
                 const double *cf_nodes, const ssmx_wave *waves,
                 const ssmx_segment *segments, const double *rho0,
                 const double *eig, const double *obs, const int32_t *order,
                 const int64_t *steps, const int64_t *read_from,
                 int32_t *counter, int32_t n_waves, int32_t waves_per_block,
                 double *out, hipStream_t s)
{
    QParams p;
    p.d = *desc;
    p.ell_col = ell_col;
    p.ell_val = ell_val;
    p.w = cf_weights;
    p.nodes = cf_nodes;
    p.waves = waves;
    p.segs = segments;
    p.rho0 = rho0;
    p.eig = eig;
    p.obs = obs;
    p.order = order;
    p.steps = steps;
    p.read_from = read_from;
    p.counter = counter;
    p.out = out;
    switch (desc->slots) {
    case 1: case 2: return launch<2>(p, n_waves, waves_per_block, s);
    case 3: return launch<3>(p, n_waves, waves_per_block, s);
    case 4: return launch<4>(p, n_waves, waves_per_block, s);
    case 5: return launch<5>(p, n_waves, waves_per_block, s);
    case 6: return launch<6>(p, n_waves, waves_per_block, s);
    case 7: return launch<7>(p, n_waves, waves_per_block, s);
    case 8: return launch<8>(p, n_waves, waves_per_block, s);
    case 9: case 10: return launch<10>(p, n_waves, waves_per_block, s);
    case 11: case 12: return launch<12>(p, n_waves, waves_per_block, s);
    default: break;
    }
    if (desc->slots <= 16) return launch<16>(p, n_waves, waves_per_block, s);
    if (desc->slots <= 24) return launch<24>(p, n_waves, waves_per_block, s);
    if (desc->slots <= 32) return launch<32>(p, n_waves, waves_per_block, s);
    return launch<64>(p, n_waves, waves_per_block, s);
}

int check_pointers(const ssmx_propagate_desc *desc, const int32_t *ell_col,
                   const double *ell_val, const double *cf_weights,
                   const double *cf_nodes, const ssmx_wave *waves,
                   const ssmx_segment *segments, const double *rho0,
                   const double *eig, const double *obs,
                   const int32_t *order, const int64_t *steps,
                   const int64_t *read_from, const double *out)
{
    if (!ell_col || !ell_val || !cf_weights || !cf_nodes || !waves ||
        !segments || !rho0 || !obs || !order || !steps || !read_from || !out)
        return set_error(SSMX_EINVAL, "null required pointer");
    if ((desc->flags & SSMX_ROTATING) && !eig)
        return set_error(SSMX_EINVAL, "rotating frame needs eig factors");
    return SSMX_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free queue launches.
//
// `scratch` holds the work-queue counter of ssmx_propagate_queue_noalloc
// (element 0; the rest pads the array to 256 bytes).  Every device the
// library runs on keeps its own copy for the life of the process.  In host
// code the name `scratch` is only a host-side stand-in: the device address
// comes from hipGetSymbolAddress, looked up once per device (with that
// device current) and kept in g_scratch[device].
// ---------------------------------------------------------------------------
constexpr int kScratchInts = 64;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) int32_t scratch[kScratchInts];

std::mutex g_scratch_mutex;
int32_t *g_scratch[kMaxDevices] = {};

__global__ void clear_queue_counter(int32_t *counter)
{
    if (threadIdx.x == 0)
        counter[0] = 0;
}

// Device address of this device's `scratch` (nullptr and an error message
// on failure).
int32_t *scratch_counter(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(SSMX_ELAUNCH, "hipGetSymbolAddress(scratch): %s",
                      hipGetErrorString(e));
            return nullptr;
        }
        g_scratch[device] = static_cast<int32_t *>(ptr);
    }
    return g_scratch[device];
}

}  // namespace

extern "C" {

int ssmx_propagate_queue(const ssmx_propagate_desc *desc,
                         const int32_t *ell_col, const double *ell_val,
                         const double *cf_weights, const double *cf_nodes,
                         const ssmx_wave *waves,
                         const ssmx_segment *segments, const double *rho0,
                         const double *eig, const double *obs,
                         const int32_t *order, const int64_t *steps,
                         const int64_t *read_from, int32_t *counter,
                         int32_t n_waves, int32_t waves_per_block,
                         double *out, void *stream)
{
    ssmx::error_buffer()[0] = '\0';
    int rc = validate(desc, n_waves, waves_per_block);
    if (rc)
        return rc;
    rc = check_pointers(desc, ell_col, ell_val, cf_weights, cf_nodes, waves,
                        segments, rho0, eig, obs, order, steps, read_from,
                        out);
    if (rc)
        return rc;
    if (!counter)
        return set_error(SSMX_EINVAL, "null required pointer");
    return queue_launch(desc, ell_col, ell_val, cf_weights, cf_nodes, waves,
                        segments, rho0, eig, obs, order, steps, read_from,
                        counter, n_waves, waves_per_block, out,
                        static_cast<hipStream_t>(stream));
}

int ssmx_propagate_queue_noalloc(const ssmx_propagate_desc *desc,
                                 const int32_t *ell_col,
                                 const double *ell_val,
                                 const double *cf_weights,
                                 const double *cf_nodes,
                                 const ssmx_wave *waves,
                                 const ssmx_segment *segments,
                                 const double *rho0, const double *eig,
                                 const double *obs, const int32_t *order,
                                 const int64_t *steps,
                                 const int64_t *read_from, int32_t n_waves,
                                 int32_t waves_per_block, double *out,
                                 void *stream)
{
    ssmx::error_buffer()[0] = '\0';
    int rc = validate(desc, n_waves, waves_per_block);
    if (rc)
        return rc;
    rc = check_pointers(desc, ell_col, ell_val, cf_weights, cf_nodes, waves,
                        segments, rho0, eig, obs, order, steps, read_from,
                        out);
    if (rc)
        return rc;
    // The counter ends at n_inst + n_waves; the scratch holds one int32.
    if (static_cast<int64_t>(desc->n_inst) + n_waves >
        static_cast<int64_t>(SSMX_NOALLOC_MAX_QUEUE))
        return set_error(SSMX_EINVAL,
                         "n_inst + n_waves = %lld exceeds the scratch "
                         "counter's limit %d",
                         static_cast<long long>(desc->n_inst) + n_waves,
                         SSMX_NOALLOC_MAX_QUEUE);
    hipStream_t s = static_cast<hipStream_t>(stream);
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess)
        return set_error(SSMX_EDEVICE, "hipGetDevice: %s",
                         hipGetErrorString(e));
    if (device < 0 || device >= kMaxDevices)
        return set_error(SSMX_EDEVICE, "device %d outside 0..%d", device,
                         kMaxDevices - 1);
    if (s) {
        int stream_device = -1;
        e = hipStreamGetDevice(s, &stream_device);
        if (e != hipSuccess)
            return set_error(SSMX_EINVAL, "hipStreamGetDevice: %s",
                             hipGetErrorString(e));
        if (stream_device != device)
            return set_error(SSMX_EINVAL,
                             "stream belongs to device %d, current device "
                             "is %d", stream_device, device);
    }
    int32_t *counter = scratch_counter(device);
    if (!counter)
        return SSMX_ELAUNCH;
    // a fill kernel, not a memset, so that a captured graph replays it
    hipLaunchKernelGGL(clear_queue_counter, dim3(1), dim3(64), 0, s,
                       counter);
    e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(SSMX_ELAUNCH, "counter clear: %s",
                         hipGetErrorString(e));
    return queue_launch(desc, ell_col, ell_val, cf_weights, cf_nodes, waves,
                        segments, rho0, eig, obs, order, steps, read_from,
                        counter, n_waves, waves_per_block, out, s);
}

}  // extern "C"
