// ssmx_chunks.hip - chaining per-chunk propagators of a time-chunked run.
//
// A long run of a small system can be cut into P time chunks of L outputs.
// Each chunk's propagator history U_p(t_k) (the operator from the chunk's
// start to its output k) is computed in parallel by propagating the n basis
// vectors through the chunk (ssmx_propagate, one member per basis vector).
// This kernel then does what the reference's serial prefix product and
// application to rho0 do (generate_simulator.py:1204-1257, 1308-1321,
// 1325-1390): x_0 = rho0, the state at output k of chunk p is
// U_p(t_k) x_p, and x_{p+1} = U_p(t_{L-1}) x_p.
//
// One wavefront per instance.  Only the chain of chunk start states is
// sequential; the outputs inside a chunk are independent given x_p and are
// computed a tile at a time, one (output, row) value per lane.
#include "ssmx_common.h"

using ssmx::kWave;
using ssmx::set_error;

namespace {

struct CParams {
    ssmx_chunk_desc d;
    const double *ops;
    const double *rho0;
    const double *obs;
    double *out;
};

// Outputs of one chunk handled per tile: kTile values y_k[r] (k, r) in the
// tile, at most four per lane.
constexpr int kTile = 4 * kWave;

__global__ __launch_bounds__(kWave) void chunk_combine_kernel(CParams p)
{
    __shared__ double xs[SSMX_MAX_N];      // chunk start state x_p
    __shared__ double ys[kTile];           // the tile's states y_k[r]
    const ssmx_chunk_desc &d = p.d;
    const int lane = threadIdx.x;
    const int inst = blockIdx.x;
    const int n = d.n;
    const int L = d.chunk_out;
    const int pack = d.pack;
    const int per_tile = kTile / n;        // outputs per tile
    if (lane < n)
        xs[lane] = p.rho0[(size_t)inst * n + lane];
    __syncthreads();
    for (int c_p = 0; c_p < d.n_chunks; ++c_p) {
        const long first = (long)c_p * L;
        const int outs = (int)(d.n_out - first < L ? d.n_out - first : L);
        // xs is read by every tile of this chunk and rewritten only after
        // the last tile's barrier
        for (int k0 = 0; k0 < outs; k0 += per_tile) {
            const int tile = outs - k0 < per_tile ? outs - k0 : per_tile;
            // y_k[r] = sum_c U_p(t_k)[r][c] x_p[c] for every (k, r) of the
            // tile; member (inst, c_p, c) is block member % pack of system
            // member / pack, row r at output k at
            // ops[(system * L + k) * pack * n + block * n + r]
            for (int e = lane; e < tile * n; e += kWave) {
                const int k = k0 + e / n;
                const int r = e - (e / n) * n;
                double y = 0.0;
                for (int c = 0; c < n; ++c) {
                    const long member =
                        ((long)inst * d.n_chunks + c_p) * n + c;
                    const long system = member / pack;
                    const int block = (int)(member - system * pack);
                    y = fma(p.ops[((size_t)system * L + k) * pack * n +
                                  (size_t)block * n + r],
                            xs[c], y);
                }
                ys[e] = y;
                if (d.n_obs == 0)
                    p.out[((size_t)inst * d.n_out + first + k) * n + r] = y;
            }
            __syncthreads();
            if (d.n_obs > 0) {
                for (int e = lane; e < tile * d.n_obs; e += kWave) {
                    const int k = e / d.n_obs;
                    const int o = e - k * d.n_obs;
                    double v = 0.0;
                    for (int r = 0; r < n; ++r)
                        v = fma(p.obs[o * n + r], ys[k * n + r], v);
                    p.out[((size_t)inst * d.n_out + first + k0 + k) *
                              d.n_obs + o] = v;
                }
            }
            // the chunk's last output is the next chunk's start state
            if (k0 + tile == outs && lane < n)
                xs[lane] = ys[(tile - 1) * n + lane];
            __syncthreads();
        }
    }
}

}  // namespace

extern "C" int ssmx_chunk_combine(const ssmx_chunk_desc *desc,
                                  const double *ops, const double *rho0,
                                  const double *obs, double *out, void *stream)
{
    ssmx::error_buffer()[0] = '\0';
    if (!desc)
        return set_error(SSMX_EINVAL, "null descriptor");
    const ssmx_chunk_desc &d = *desc;
    if (d.n < 1 || d.n > SSMX_MAX_N)
        return set_error(SSMX_EINVAL, "n=%d outside 1..%d", d.n, SSMX_MAX_N);
    if (d.n_inst < 1 || d.n_chunks < 1 || d.chunk_out < 1)
        return set_error(SSMX_EINVAL,
                         "n_inst=%d, n_chunks=%d, chunk_out=%d must be >= 1",
                         d.n_inst, d.n_chunks, d.chunk_out);
    if (d.n_out < 1 || d.n_out > (long)d.n_chunks * d.chunk_out ||
        d.n_out <= (long)(d.n_chunks - 1) * d.chunk_out)
        return set_error(SSMX_EINVAL,
                         "n_out=%d does not end in the last of %d chunks of "
                         "%d outputs", d.n_out, d.n_chunks, d.chunk_out);
    if (d.pack < 1 || (long)d.pack * d.n > SSMX_MAX_N)
        return set_error(SSMX_EINVAL, "pack=%d x n=%d exceeds %d rows",
                         d.pack, d.n, SSMX_MAX_N);
    if (d.n_obs < 0 || d.n_obs > SSMX_MAX_OBS)
        return set_error(SSMX_EINVAL, "n_obs=%d outside 0..%d", d.n_obs,
                         SSMX_MAX_OBS);
    if (d.reserved != 0)
        return set_error(SSMX_EINVAL, "reserved field must be 0");
    if (!ops || !rho0 || !out || (d.n_obs > 0 && !obs))
        return set_error(SSMX_EINVAL, "null required pointer");
    CParams p;
    p.d = d;
    p.ops = ops;
    p.rho0 = rho0;
    p.obs = obs;
    p.out = out;
    hipLaunchKernelGGL(chunk_combine_kernel, dim3(d.n_inst), dim3(kWave), 0,
                       static_cast<hipStream_t>(stream), p);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(SSMX_ELAUNCH, "chunk combine launch: %s",
                         hipGetErrorString(e));
    return SSMX_OK;
}
