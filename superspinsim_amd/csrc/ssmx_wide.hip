// ssmx_wide.hip - propagator for superoperators of dimension up to 512
// (gfx950), e.g. the 21-level NV + 14N hyperfine model (n = 441, or 171
// with the compressed vectorisation).
//
// W = ceil(n / 64) wavefronts advance one instance: thread r of the
// instance owns row r of the superoperator and component r of the state.
// A workgroup holds IPB <= floor(16 / W) instances, which share one LDS
// copy of the generator values (every instance uses the same matrices;
// only the coefficients differ).  The launch puts one instance in a
// workgroup when two such workgroups fit on a CU (n = 171), else the count
// that keeps the most instances resident (2 at n = 441).  The
// Taylor action exp(A) x is evaluated as in ssmx_propagate.hip, with the
// exchange of each term across the instance's wavefronts:
//  - slot 0 (diagonal) needs no exchange, slot 1 (partner r ^ 1, same
//    wavefront) is a DPP quad permutation;
//  - other slots are read from an LDS buffer of the term; two buffers
//    alternate, so one workgroup barrier per term orders every read of a
//    term before the next write of the same buffer.  The buffers are a
//    fixed 4 KiB apart, so the gather addresses are computed once and the
//    buffer is selected by an immediate offset.
// A_j's row is assembled once per exponential from the staged values with
// the 64-row kernel's roundings, so both kernels give the same states for
// the same problem.
// Each wavefront runs an exponential specialised to the slots its widest row
// uses (the host orders row pairs by width, so most wavefronts need far
// fewer than the widest row of the matrix).  Every wavefront of a workgroup
// meets every other at each product's barrier, so the product's instruction
// count, times the wavefronts per SIMD, sets the pace: no address
// arithmetic and no work on unused slots in the product.
//
// Dense FP64 MFMA is not used: at n = 441 the union of the generators has
// 3,677 non-zeros (at most 14 per row), so a dense product would do about
// 50 times the arithmetic of the sparse action for the same result, and
// MI355X's FP64 matrix and vector peaks are equal.

#include "ssmx_wide_exp.h"

namespace {

// workgroups of at most this many wavefronts get 256 VGPRs per lane
constexpr int kSmallBlockWaves = 4;

using ssmx::node_time;
using ssmx::set_error;
using ssmx::wave_coef;
using ssmx::wave_max;
using ssmx::wave_sum;

struct WParams {
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

// The propagator's step_factor (ssmx_propagate.hip), for one instance.
__device__ double wide_factor(const WParams &p, const double *coef_i,
                              const ssmx_wave *wv, long step, int j, int m,
                              int &cursor)
{
    const ssmx_propagate_desc &d = p.d;
    const int S = d.n_node, K = d.n_gen;
    const long i_loc = step / d.fine;
    const int r = (int)(step - i_loc * d.fine);
    const long ig = (long)d.first_out + i_loc;
    const bool rotating = (d.flags & SSMX_ROTATING) != 0;
    const double *row = nullptr;
    if (!wv)
        row = coef_i + ((size_t)(i_loc * d.fine + r) * S) * K;
    if (rotating) {
        const int s = m / K, g = m % K;
        const double c =
            wv ? wave_coef(wv, p.segs, g,
                           node_time(d.t0, d.dt, d.fine, r, ig, p.nodes[s]),
                           cursor)
               : row[s * K + g];
        return c * (d.h * p.w[j * S + s]);
    }
    double cw = 0.0;
    for (int s = 0; s < S; ++s) {
        const double c =
            wv ? wave_coef(wv, p.segs, m,
                           node_time(d.t0, d.dt, d.fine, r, ig, p.nodes[s]),
                           cursor)
               : row[s * K + m];
        cw = fma(p.w[j * S + s], c, cw);
    }
    return d.h * cw;
}

// WPB: most wavefronts per workgroup.  At the full kMaxBlockWaves four
// wavefronts share a SIMD, so each gets 128 VGPRs; workgroups of at most
// kSmallBlockWaves (one instance, n <= 256) keep two wavefronts per SIMD
// and 256 VGPRs, where the 14- and 12-slot rows no longer spill.
template <int SLOTS, int WPB = kMaxBlockWaves>
__global__ __launch_bounds__(kWave *WPB)
__attribute__((amdgpu_waves_per_eu(WPB <= kSmallBlockWaves ? 2 : 4)))
void wide_kernel(
    WParams p, int W, int IPB)
{
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    extern __shared__ double lds[];
    const ssmx_propagate_desc &d = p.d;
    const int NP = W * kWave;                 // rows of one instance
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int ii = tid / NP;                  // instance slot in the block
    const int row = tid - ii * NP;
    const int wave = tid / kWave;
    const int inst = blockIdx.x * IPB + ii;
    const bool active = inst < d.n_inst;

    const bool rotating = (d.flags & SSMX_ROTATING) != 0;
    const bool wave_mode = (d.flags & SSMX_COEF_WAVE) != 0;
    const bool read_full = (d.flags & SSMX_READ_FULL) != 0;
    const bool read_int = (d.flags & SSMX_READ_INTEGRAL) != 0;
    const bool use_partner = d.partner != 0;
    const int M = (rotating ? d.n_node : 1) * d.n_gen;
    const int n = d.n;
    const int J = d.n_exp;

    // LDS: staged values [m][slot / 2][row][slot % 2], then per instance
    // slot two exchange buffers kBuf doubles apart, then one partial sum
    // per wavefront
    double *lv = lds;
    double *xb0 = lv + (size_t)M * SLOTP * NP + (size_t)ii * 2 * kBuf;
    double *red = lv + (size_t)M * SLOTP * NP + (size_t)IPB * 2 * kBuf;
    {
        const int nval = M * SLOTP * NP;
        for (int e = tid; e < nval; e += blockDim.x) {
            const int half = e & 1;
            const int r = (e >> 1) % NP;
            const int pair = (e / (2 * NP)) % (SLOTP / 2);
            const int m = e / (NP * SLOTP);
            const int slot = 2 * pair + half;
            lv[e] = slot < d.slots
                        ? p.ell_val[((size_t)m * d.slots + slot) * NP + r]
                        : 0.0;
        }
        __syncthreads();
    }

    // slots this wavefront reads: its widest row (slot 0's column entry
    // holds the row's used-slot count)
    const int ws = wave_max(p.ell_col[row]);
    // gather addresses (buffer 0); column indices are taken modulo the
    // buffer size, so no table can address outside the instance's buffers
    const double *gp[SLOTS];
    gp[0] = xb0 + row;
#pragma unroll
    for (int sl = 1; sl < SLOTS; ++sl)
        gp[sl] = xb0 + ((sl < d.slots ? p.ell_col[sl * NP + row] : row) &
                        (kBuf - 1));
    double *own = xb0 + row;
    const double inv_lane = kInvW.v[lane <= SSMX_MAX_TAYLOR ? lane : 0];

    const int inst_c = active ? inst : 0;     // idle slots read instance 0
    const double *coef_i =
        wave_mode ? nullptr : p.coef + (size_t)inst_c * d.coef_inst_stride;
    const ssmx_wave *wv = wave_mode ? p.waves + inst_c : nullptr;
    double x = (active && row < n) ? p.rho0[(size_t)inst * n + row] : 0.0;
    double eig_own = 0.0, eig_part = 0.0;
    if (rotating) {
        eig_own = p.eig[row];
        eig_part = p.eig[NP + row];
    }
    double obs0 = 0.0, obs1 = 0.0, off0 = 0.0, off1 = 0.0;
    if (!read_full) {
        obs0 = p.obs[row];
        off0 = (p.obs_offset && active)
                   ? p.obs_offset[(size_t)inst * d.n_obs] : 0.0;
        if (d.n_obs > 1) {
            obs1 = p.obs[NP + row];
            off1 = (p.obs_offset && active)
                       ? p.obs_offset[(size_t)inst * d.n_obs + 1] : 0.0;
        }
    }

    // coefficient batches (every wavefront computes its instance's batch)
    const int JM = J * M;
    const int B = kWave / JM;
    const long total = (long)d.n_out * d.fine;
    const int my_b = lane / JM;
    const int my_j = (lane % JM) / M;
    const int my_m = lane % M;
    int cursor = 0;
    double factor = 0.0;
    const double inv_sub = 1.0 / (double)d.substeps;
    const double half_h = 0.5 * d.h;
    double xint = 0.0;
    int parity = 0;
    ExpArgs e;
    e.lv = lv;
    e.NP = NP;
    e.M = M;
    e.base = 0;
    e.substeps = d.substeps;
    e.taylor = d.taylor;
    e.inv_sub = inv_sub;
    e.partner = use_partner;

    // Sum of v over the instance's W wavefronts (every thread gets it).
    auto inst_sum = [&](double v) -> double {
        v = wave_sum(v);
        if (lane == 0)
            red[wave] = v;
        __syncthreads();
        double s = 0.0;
        for (int k = 0; k < W; ++k)
            s += red[ii * W + k];
        __syncthreads();
        return s;
    };

    const int ckpt_every = p.state_out ? d.ckpt_every : 0;
    const long n_ckpt = ckpt_every > 0 ? (total + ckpt_every - 1) / ckpt_every
                                       : 0;
    int ckpt_count = 0;
    long ckpt_index = 0;
    const bool own_row = active && row < n;

    long step = 0;
    int b = 0;
    for (int i = 0; i < d.n_out; ++i) {
        for (int r = 0; r < d.fine; ++r, ++step, ++b) {
            if (b == B)
                b = 0;
            if (ckpt_every > 0) {
                if (ckpt_count == 0 && own_row)
                    p.state_out[((size_t)inst * n_ckpt + ckpt_index) * n +
                                row] = x;
                if (ckpt_count == 0)
                    ++ckpt_index;
                if (++ckpt_count == ckpt_every)
                    ckpt_count = 0;
            }
            if (b == 0) {
                const long mine = step + my_b;
                factor = (my_b < B && mine < total)
                             ? wide_factor(p, coef_i, wv, mine, my_j, my_m,
                                           cursor)
                             : 0.0;
            }
            const double x_prev = x;
            for (int j = J - 1; j >= 0; --j) {
                e.base = (b * J + j) * M;
                x = wide_exp_dispatch<SLOTS, true>(ws, x, e, row, factor,
                                                   inv_lane, own, gp, parity);
            }
            if (rotating) {
                const double xp = swap_pair(x);
                x = x + fma(eig_own, x, eig_part * xp);
            }
            if (read_int)
                xint = fma(half_h, x_prev + x, xint);
        }
        if (read_full) {
            if (own_row)
                p.out[((size_t)inst * d.n_out + i) * n + row] = x;
        } else if (!read_int) {
            const double v0 = inst_sum(obs0 * x);
            if (row == 0 && active)
                p.out[((size_t)inst * d.n_out + i) * d.n_obs] = v0 + off0;
            if (d.n_obs > 1) {
                const double v1 = inst_sum(obs1 * x);
                if (row == 0 && active)
                    p.out[((size_t)inst * d.n_out + i) * d.n_obs + 1] =
                        v1 + off1;
            }
        }
    }
    if (read_int) {
        const double v0 = inst_sum(obs0 * xint);
        if (row == 0 && active)
            p.out[(size_t)inst * d.n_obs] = v0 + off0;
        if (d.n_obs > 1) {
            const double v1 = inst_sum(obs1 * xint);
            if (row == 0 && active)
                p.out[(size_t)inst * d.n_obs + 1] = v1 + off1;
        }
    }
    if (p.state_out && ckpt_every == 0 && own_row)
        p.state_out[(size_t)inst * n + row] = x;
}

template <int SLOTS>
int launch(const WParams &p, hipStream_t stream)
{
    constexpr int SLOTP = SLOTS + (SLOTS & 1);
    const int M = ((p.d.flags & SSMX_ROTATING) ? p.d.n_node : 1) * p.d.n_gen;
    const int W = (p.d.n + kWave - 1) / kWave;
    const int NP = W * kWave;
    // instances per workgroup: at most 16 wavefronts, and the staged values
    // plus two exchange buffers per instance within 160 KiB of LDS
    const size_t staged = (size_t)M * SLOTP * NP * sizeof(double);
    auto lds_of = [&](int ipb) {
        return staged + (size_t)ipb * (2 * kBuf + W) * sizeof(double);
    };
    int ipb_max = kMaxBlockWaves / W;
    if (ipb_max > p.d.n_inst)
        ipb_max = p.d.n_inst;
    while (ipb_max > 1 && lds_of(ipb_max) > 160 * 1024)
        --ipb_max;
    if (lds_of(ipb_max) > 160 * 1024)
        return set_error(SSMX_EINVAL,
                         "n=%d: generator values need %zu bytes of LDS "
                         "(> 160 KiB)",
                         p.d.n, lds_of(ipb_max));
    // n <= 256: one instance per workgroup of at most kSmallBlockWaves
    // wavefronts with 256 VGPRs, when two or more such workgroups are
    // resident per CU (the same operations, so the same results)
    if (W <= kSmallBlockWaves) {
        const void *small = reinterpret_cast<const void *>(
            &wide_kernel<SLOTS, kSmallBlockWaves>);
        bool ok = true;
        if (lds_of(1) > 64 * 1024)
            ok = hipFuncSetAttribute(
                     small, hipFuncAttributeMaxDynamicSharedMemorySize,
                     (int)lds_of(1)) == hipSuccess;
        int nb = 0;
        if (ok)
            ok = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                     &nb, small, NP, lds_of(1)) == hipSuccess;
        if (!ok)
            (void)hipGetLastError();    // fall back to the full variant
        if (ok && nb >= 2) {
            hipLaunchKernelGGL((wide_kernel<SLOTS, kSmallBlockWaves>),
                               dim3(p.d.n_inst), dim3(NP), lds_of(1), stream,
                               p, W, 1);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess)
                return set_error(SSMX_ELAUNCH, "wide propagate launch: %s",
                                 hipGetErrorString(e));
            return SSMX_OK;
        }
    }
    const void *kern = reinterpret_cast<const void *>(&wide_kernel<SLOTS>);
    if (lds_of(ipb_max) > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(
            kern, hipFuncAttributeMaxDynamicSharedMemorySize,
            (int)lds_of(ipb_max));
        if (e != hipSuccess)
            return set_error(SSMX_ELAUNCH, "hipFuncSetAttribute: %s",
                             hipGetErrorString(e));
    }
    // One instance per workgroup when two or more such workgroups fit on a
    // CU: fewer wavefronts meet at each barrier.  On the n = 171 model this
    // ran 512-4096 points 1.2-1.8x faster than 5 per workgroup (and
    // within 1 % of it at 1024).  Otherwise (n = 441: the staged values
    // allow one workgroup per CU) the count that keeps the most instances
    // resident per CU.
    auto resident = [&](int c) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &nb, kern, c * NP, lds_of(c)) != hipSuccess) {
            (void)hipGetLastError();    // not a launch error: count 0
            return 0;
        }
        return nb;
    };
    int ipb = 1;
    if (resident(1) < 2) {
        int most = -1;
        for (int c = 1; c <= ipb_max; ++c) {
            const int r = resident(c) * c;
            if (r > most) {
                most = r;
                ipb = c;
            }
        }
    }
    const size_t lds = lds_of(ipb);
    const int blocks = (p.d.n_inst + ipb - 1) / ipb;
    hipLaunchKernelGGL((wide_kernel<SLOTS>), dim3(blocks),
                       dim3(ipb * NP), lds, stream, p, W, ipb);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(SSMX_ELAUNCH, "wide propagate launch: %s",
                         hipGetErrorString(e));
    return SSMX_OK;
}

int dispatch_slots(const WParams &p, hipStream_t s)
{
    const int sl = p.d.slots;
    if (sl <= 8) return launch<8>(p, s);
    if (sl <= 12) return launch<12>(p, s);
    if (sl <= 14) return launch<14>(p, s);
    if (sl <= 16) return launch<16>(p, s);
    if (sl <= 24) return launch<24>(p, s);
    return launch<32>(p, s);
}

}  // namespace

// Entry from ssmx_propagate (ssmx_propagate.hip) for n > 64; the descriptor
// and pointers are validated there.
int ssmx_wide_propagate(const ssmx_propagate_desc *desc,
                        const int32_t *ell_col, const double *ell_val,
                        const double *cf_weights, const double *cf_nodes,
                        const void *coef, const ssmx_segment *segments,
                        const double *rho0, const double *eig,
                        const double *obs, const double *obs_offset,
                        double *out, double *state_out, hipStream_t stream)
{
    const bool wave_mode = (desc->flags & SSMX_COEF_WAVE) != 0;
    WParams p;
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
    return dispatch_slots(p, stream);
}
