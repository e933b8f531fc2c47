/*
 * ssmx.h - C ABI of the MI355X (gfx950) propagator for superspinsim-style
 * open spin systems.
 *
 * The library replaces the numba.cuda kernel sequence that
 * superspinsim/generate_simulator.py launches inside simulate()
 * (generate_simulator.py:1945-2116: node times, sampling, coefficient
 * combination, differential, scaling, 2R squarings, CF product, rotating
 * eigen-step, prefix product, application to rho0).  The Python factory
 * superspinsim_amd.generate_simulator keeps the reference signature and
 * calls these entry points through ctypes.
 *
 * Conventions
 *  - Every pointer argument is a device pointer (HBM), owned by the caller.
 *    The library never allocates or frees device memory.
 *  - Calls are asynchronous on the given stream (NULL = default stream).
 *  - Return 0 on success or a negative SSMX_E* code; ssmx_last_error()
 *    returns a message for the last failure of the calling thread.
 *  - All arithmetic is IEEE binary64.
 *
 * Data layout of a propagation problem (one launch advances n_inst
 * independent instances over n_out * fine propagator steps each):
 *  - Superoperators are n x n real matrices (n <= 64; ssmx_propagate also
 *    n <= 512), stored in a shared ELL pattern with row stride
 *    NP = 64 * ceil(n / 64) (64 for n <= 64): column index
 *    ell_col[slot * NP + row] (int32).  Slot 0 of row r is always column r
 *    (the diagonal, value possibly 0) and ell_col[row] (slot 0) instead
 *    holds the number of used slots of that row; unused slots point at
 *    column r with value 0; values
 *    ell_val[((node * n_gen + g) * slots + slot) * NP + row]; node ranges
 *    over n_node for the rotating frame (node-wise interaction-picture
 *    generators) and is 0 otherwise.  Readout rows (obs) and eigen-step
 *    factors (eig) use the same stride NP.  For n > 64 the pattern may use
 *    at most 32 slots, and the staged values (matrices x slots x NP
 *    doubles) plus 8 KiB of exchange buffers must fit 160 KiB of LDS; the
 *    used-slot counts let each wavefront skip slots none of its rows use.
 *    For n <= 64 with partner = 1, a slot pair (2 + 2k, 3 + 2k) whose two
 *    columns are 2p and 2p + 1 (an aligned pair) in every row of the
 *    64-row table is read whole instead of by two gathers: when only the
 *    first slot pair holds one (it is then the other pair of the row's
 *    quad, (r & ~1) ^ 2 and its odd neighbour) by DPP, else by one 16-byte
 *    LDS read each.  The kernels detect such leading slot pairs from
 *    ell_col; any column order gives the same result, and the layout only
 *    changes speed.  Rows >= n take part in the detection, so their slots
 *    must follow the layout too (values 0).
 *  - One propagator step of size h applies exp(A_{J-1}) first and exp(A_0)
 *    last, with A_j = h sum_g (sum_s w[j,s] c[s,g]) G_g, or in the rotating
 *    frame A_j = sum_s sum_g (G~[s,g] c[s,g]) (h w[j,s]) followed by the
 *    diagonal block step exp(L0 h).  These are the reference's
 *    _combine_coefficients/_calculate_differential(_rotating)
 *    (generate_simulator.py:278-532) and _quadrature_combine_run order
 *    (:1259-1294).
 *  - The state vector (not an operator prefix product) is propagated; the
 *    two are equal in exact arithmetic.
 */
#ifndef SSMX_H
#define SSMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define SSMX_OK 0
#define SSMX_EINVAL (-1)   /* bad sizes or flags */
#define SSMX_ELAUNCH (-2)  /* HIP launch or attribute failure */
#define SSMX_EDEVICE (-3)  /* no usable gfx950 device */

/* flags of ssmx_propagate_desc.flags */
#define SSMX_ROTATING 1      /* node-wise generators + diagonal eigen-step */
#define SSMX_COEF_WAVE 2     /* coefficients from waveform descriptors */
#define SSMX_READ_FULL 4     /* readout: whole state vector (n values) */
#define SSMX_READ_INTEGRAL 8 /* readout: trapezoid integral of obs rows */
#define SSMX_STORE_INPUTS 16 /* state_out: every exponential's input, for
                                ssmx_adjoint (see ssmx_propagate) */

/* limits */
#define SSMX_MAX_N 64       /* one wavefront per instance */
#define SSMX_MAX_N_WIDE 512 /* ssmx_propagate only: one workgroup of
                               ceil(n / 64) wavefronts per instance */
#define SSMX_MAX_GEN 8
#define SSMX_MAX_EXP 6
#define SSMX_MAX_NODE 3
#define SSMX_MAX_OBS 2
#define SSMX_MAX_TAYLOR 40

typedef struct ssmx_propagate_desc {
    int32_t n;        /* superoperator dimension, 1..64 (ssmx_propagate:
                         1..512) */
    int32_t n_inst;   /* independent instances (one wavefront each) */
    int32_t n_gen;    /* K generators in the ELL pattern, 1..8 */
    int32_t slots;    /* ELL slots per row, 1..64 */
    int32_t n_exp;    /* J exponentials per step, 1..6 */
    int32_t n_node;   /* S quadrature nodes, 1..3 */
    int32_t fine;     /* F propagator steps per output sample */
    int32_t n_out;    /* output samples per instance */
    int32_t taylor;   /* Taylor degree m of the action exp(A) v, 1..40 */
    int32_t substeps; /* s: exp(A) v = (exp(A/s))^s v, power of two */
    int32_t n_obs;    /* readout rows for obs readouts, 1..2 */
    int32_t flags;    /* SSMX_* flags */
    int32_t first_out; /* global index of the first output interval of this
                          launch (node times count intervals from t0) */
    int32_t partner;  /* 1: ELL slot 1 of row r is column r ^ 1 for every
                         row (exchanged by DPP instead of ds_bpermute) */
    int32_t ckpt_every; /* K > 0: state_out receives checkpoints (the state
                           before steps 0, K, 2K, ...); 0: the final state */
    int32_t reserved;   /* 0 */
    int64_t coef_inst_stride; /* table mode: doubles between instances;
                                 0 = one table shared by all instances */
    double h;         /* propagator step dt / F (s) */
    double t0;        /* start time (s) */
    double dt;        /* output step (s) */
} ssmx_propagate_desc;

/*
 * Waveform descriptor (SSMX_COEF_WAVE).  Coefficient g at node time t is
 *   c_g(t) = base[g] + sum over segments k of generator g with
 *            t_start <= t < t_end of  amp * cos(omega * t + phase)
 * Segments of one generator are sorted by t_start and do not overlap.
 * Node times are t = ((t0 + (r dt) / F) + dt i) + (dt c_s) / F exactly as
 * generate_simulator.py:161-167, 1956-1961 evaluates them.
 */
typedef struct ssmx_segment {
    double t_start, t_end, amp, omega, phase;
} ssmx_segment;

typedef struct ssmx_wave {
    double base[SSMX_MAX_GEN];
    int32_t seg_begin[SSMX_MAX_GEN]; /* index into the segment array */
    int32_t seg_count[SSMX_MAX_GEN];
} ssmx_wave;

/*
 * Propagate n_inst instances (one wavefront each).
 *
 * ell_col    int32 [slots * NP]
 * ell_val    double [(n_node or 1) * n_gen * slots * NP]
 * cf_weights double [n_exp * n_node]     (row j: weights of exponential j)
 * cf_nodes   double [n_node]             (Gauss-Legendre nodes on [0, 1])
 * coef       table mode: double [n_inst? * n_out * fine * n_node * n_gen]
 *            ordered (instance, i, r, s, g); wave mode: ssmx_wave [n_inst]
 * segments   wave mode: ssmx_segment [...]; table mode: unused (NULL)
 * rho0       double [n_inst * n]
 * eig        rotating: double [2 * NP] per-row (own, partner) factors of
 *            exp(L0 h) - I; partner of row r is row r ^ 1
 * obs        double [n_obs * NP] readout rows (obs readouts, n_obs <= 2)
 * obs_offset double [n_inst * n_obs] added to each obs readout (or NULL)
 * out        READ_FULL: double [n_inst * n_out * n]
 *            obs: double [n_inst * n_out * n_obs]
 *            READ_INTEGRAL: double [n_inst * n_obs]
 * state_out  ckpt_every = 0: double [n_inst * n] final state (or NULL);
 *            ckpt_every = K: double [n_inst * ceil(n_out * fine / K) * n],
 *            the state before every K-th step (input of ssmx_adjoint);
 *            SSMX_STORE_INPUTS: double [n_inst][n_out * fine][n_exp][64],
 *            the input of every exponential (rows >= n hold 0), for
 *            ssmx_adjoint's stored-input mode; needs n <= 64, the lab
 *            frame, substeps 1, ckpt_every 0, taylor 11, slots <= 8 and at
 *            least two matrices
 *
 * Replaces: generate_simulator.py:1945-2116 (simulate's kernel sequence).
 */
int ssmx_propagate(const ssmx_propagate_desc *desc, const int32_t *ell_col,
                   const double *ell_val, const double *cf_weights,
                   const double *cf_nodes, const void *coef,
                   const ssmx_segment *segments, const double *rho0,
                   const double *eig, const double *obs,
                   const double *obs_offset, double *out, double *state_out,
                   void *stream);

/*
 * The reference's non-residual exponential (use_residual=False).
 *
 * Each exponential is (I + A / scale)^(2^squarings) with scale = 4^R and
 * squarings = 2R, formed densely: A_j is assembled as ssmx_propagate does,
 * divided by `scale`, the identity added, then squared with plain products
 * (k ascending in every dot product, fused multiply-adds).  In FP64 the
 * diagonal 1 + a / 4^R rounds to 1 for small |a|, so damping terms are
 * lost exactly as the reference loses them.  The state is propagated
 * through the resulting matrices (exponential J-1 first, then in the
 * rotating frame the eigen-step exp(L0 h)) and written at every output.
 *
 * gens       non-rotating: double [n_gen][n][n]; rotating: [n_node][n_gen]
 *            [n][n] (node-wise interaction-picture generators), row-major
 * cf_weights double [n_exp * n_node]
 * coef       double [n_inst? * n_out * fine * n_node * n_gen] ordered
 *            (instance, i, r, s, g) as the ssmx_propagate table mode
 * rho0       double [n_inst * n]
 * eig        rotating: double [2 n_doubles + n_singles]: (cos, sin) factors
 *            (d0, d1) of each 2x2 block acting on rows (2y, 2y + 1) as
 *            [[d0, -d1], [d1, d0]], then the factor of each single row
 *            2 n_doubles + y; NULL otherwise
 * out        double [n_inst * n_out * n] states at every output
 *
 * Replaces: generate_simulator.py:534-541 (scaling, + I) and :706-724,
 * :882-897 (squarings) with use_residual=False, within the kernel sequence
 * :1945-2116.
 */
typedef struct ssmx_dense_desc {
    int32_t n;          /* superoperator dimension, 1..64 */
    int32_t n_inst;     /* one workgroup each */
    int32_t n_gen;      /* G, 1..8 */
    int32_t n_exp;      /* J, 1..6 */
    int32_t n_node;     /* S, 1..3 */
    int32_t fine;       /* F propagator steps per output */
    int32_t n_out;      /* outputs per instance */
    int32_t squarings;  /* 2R, 0..1024 */
    int32_t flags;      /* 0 or SSMX_ROTATING */
    int32_t n_doubles;  /* rotating: 2x2 eigen-step blocks */
    int32_t n_singles;  /* rotating: 1x1 eigen-step blocks */
    int32_t reserved;   /* 0 */
    int64_t coef_inst_stride; /* doubles between instances; 0 = shared */
    double h;           /* propagator step dt / F (s) */
    double scale;       /* 4^R */
} ssmx_dense_desc;

int ssmx_propagate_dense(const ssmx_dense_desc *desc, const double *gens,
                         const double *cf_weights, const double *coef,
                         const double *rho0, const double *eig, double *out,
                         void *stream);

/*
 * Chain the per-chunk propagators of a time-chunked run.
 *
 * A run of N outputs is cut into n_chunks chunks of chunk_out outputs (the
 * last may be shorter).  ops holds, for every instance, chunk p and basis
 * column c, the full-state history of ssmx_propagate started from the
 * unit vector e_c over chunk p (time-shifted to the chunk's start): the
 * propagator U_p(t_k) column by column.  With x_0 = rho0 the kernel writes
 * the state U_p(t_k) x_p at every global output p chunk_out + k (or its
 * n_obs readout rows) and carries x_{p+1} = U_p(t_last) x_p to the next
 * chunk.  This is the reference's serial prefix product and application
 * to rho0 (generate_simulator.py:1204-1257, 1308-1321, 1325-1390) over
 * chunks instead of single outputs.
 *
 * Member (inst, p, c) = (inst * n_chunks + p) * n + c sits in system
 * member / pack at block member % pack of a block-diagonal ssmx_propagate
 * run of pack * n rows:
 *   ops   double [systems][chunk_out][pack * n], systems =
 *         ceil(n_inst * n_chunks * n / pack), standard basis
 *   rho0  double [n_inst][n]
 *   obs   double [n_obs][n] (n_obs 1..2), or NULL with n_obs = 0
 *   out   n_obs = 0: double [n_inst][n_out][n]; else [n_inst][n_out][n_obs]
 */
typedef struct ssmx_chunk_desc {
    int32_t n;          /* state dimension, 1..64 */
    int32_t n_inst;
    int32_t n_chunks;   /* P */
    int32_t chunk_out;  /* L outputs per chunk */
    int32_t n_out;      /* N, (P - 1) L < N <= P L */
    int32_t pack;       /* members per system, pack * n <= 64 */
    int32_t n_obs;      /* 0 (full states) or 1..2 readout rows */
    int32_t reserved;   /* 0 */
} ssmx_chunk_desc;

int ssmx_chunk_combine(const ssmx_chunk_desc *desc, const double *ops,
                       const double *rho0, const double *obs, double *out,
                       void *stream);

/*
 * Persistent work-queue propagation of protocols with uneven lengths.
 *
 * Instance i runs steps[i] propagator steps (node times as ssmx_propagate
 * with first_out = 0; n_out is unused) and writes the trapezoid integral of
 * its readout rows over steps read_from[i] .. steps[i]:
 *   out[i][o] = sum_{k = read_from[i]}^{steps[i]-1} (h/2) obs_o.(x_k + x_k+1)
 * desc->flags must hold SSMX_COEF_WAVE | SSMX_READ_INTEGRAL (and may hold
 * SSMX_ROTATING).  The grid holds n_waves wavefronts in blocks of
 * waves_per_block (1..8); each wavefront takes queue entries q = 0, 1, ...
 * from *counter with an atomic increment and runs instance order[q] to the
 * end, so order the entries longest first.  *counter must be 0 at launch
 * (clear it with a kernel on the same stream, not a memset node, when the
 * launch is captured in a graph); it ends at n_inst + n_waves.
 *
 * waves      ssmx_wave [n_inst]; segments as ssmx_propagate
 * rho0       double [n_inst * n]
 * obs        double [n_obs * 64]
 * order      int32 [n_inst], a permutation of 0 .. n_inst - 1
 * steps      int64 [n_inst]; read_from int64 [n_inst] (0 <= read_from)
 * counter    int32 [1] device
 * out        double [n_inst * n_obs]
 *
 * No counterpart in the reference (one simulate call per protocol); each
 * instance runs the path ssmx_propagate replaces
 * (generate_simulator.py:1945-2116).
 */
int ssmx_propagate_queue(const ssmx_propagate_desc *desc,
                         const int32_t *ell_col, const double *ell_val,
                         const double *cf_weights, const double *cf_nodes,
                         const ssmx_wave *waves,
                         const ssmx_segment *segments, const double *rho0,
                         const double *eig, const double *obs,
                         const int32_t *order, const int64_t *steps,
                         const int64_t *read_from, int32_t *counter,
                         int32_t n_waves, int32_t waves_per_block,
                         double *out, void *stream);

/*
 * ssmx_propagate_queue without the counter argument, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.
 *
 * The queue counter lives in a file-scope device array of the library
 * (256 bytes per device, held until the process ends), cleared by a fill
 * kernel on `stream` before the launch; nothing is allocated or freed.
 * Every call on one device shares that counter: calls that overlap on one
 * device (concurrent streams, threads) race and must be serialised by the
 * caller.  `stream` must belong to the current device (NULL = its default
 * stream).  Arguments and results are those of ssmx_propagate_queue;
 * n_inst + n_waves must not exceed SSMX_NOALLOC_MAX_QUEUE.
 */
#define SSMX_NOALLOC_MAX_QUEUE 0x7fffffff

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
                                 void *stream);

/*
 * Adjoint (reverse-mode gradient) of the trapezoid-integrated readout
 *   F_o = sum_k w_k obs_o . x_k,  w_0 = w_T = h / 2, other w_k = h,
 * over the same propagation as ssmx_propagate with SSMX_READ_INTEGRAL
 * (non-rotating frame, substeps = 1).  Given seeds g[inst][o] = dLoss/dF_o
 * it returns, per instance,
 *   grad_theta[inst][t] = sum_{k,j} f_{kj,theta_gen} <Abar_kj, D_t>
 *   grad_seg[rec]      += sum_{k,j,s} h w[j,s] dc(t_s)/d amp_rec
 *                                     <Abar_kj, G_grad_gen>
 *   lambda_out[inst]    = dLoss/d rho0 (the adjoint state at step 0)
 * where Abar_kj is the gradient of the loss with respect to the exponent
 * A_kj of exponential j of step k, f_{kj,m} is the scalar factor of matrix
 * m in A_kj, and D_t = dG_theta_gen / d theta_t are derivative tables.
 * Abar is the adjoint of the Taylor action of the block-triangular
 * augmented exponential exp([[A, E], [0, A]]) (its upper block is the
 * Frechet derivative of exp at A in direction E), truncated at degree
 * taylor_adj: Abar = sum_{i+j<taylor_adj} c_ij u_i w_j^T with
 * u_i = (A^T)^i mu / i!, w_j = A^j y / j!, c_ij = i! j! / (i+j+1)!.
 *
 * The forward states are recomputed from checkpoints every ckpt_every steps
 * (state_out of a ssmx_propagate call with the same ckpt_every); one
 * wavefront recomputes segment p - 1 while a second one runs the adjoint of
 * segment p, exchanging states through `workspace`.
 *
 * n > 64 (up to 512; one workgroup of ceil(n / 64) wavefronts per
 * instance, recomputation then reverse sweep per segment): tables use the
 * row stride NP = 64 ceil(n / 64) throughout, and
 *  - ell_col/ell_val: the matrices' pattern extended by the derivative
 *    tables' non-zeros (`slots` in all; the matrices' own columns first,
 *    values 0 in the derivative-only slots); ell_col[row] holds the used
 *    slots in bits 0-7 and the matrices' used slots in bits 8-15;
 *    `reserved` = the matrices' slot count (<= 16; 0 = slots);
 *  - ell_col_t: the transposed pattern of the matrices (<= 16 slots);
 *    ell_val_t is not read (the values of A^T are taken from ell_val);
 *  - dval_t: D_t[r][c] at the slots of ell_col, [n_theta][slots][NP];
 *  - obs rows [n_obs][NP]; grad_gen must be -1, partner 1;
 *  - workspace: n_inst * ckpt_every * n_exp * NP doubles.
 *
 * No counterpart in the reference (it has no gradients); it differentiates
 * the path ssmx_propagate replaces (generate_simulator.py:1945-2116).
 */
#define SSMX_ADJ_MAX_TAYLOR 10 /* limit of taylor_adj */

typedef struct ssmx_adjoint_desc {
    int32_t n;          /* as ssmx_propagate_desc */
    int32_t n_inst;
    int32_t n_gen;      /* matrices M in both ELL stacks, 1..8 */
    int32_t slots;      /* forward ELL slots (pattern of the M matrices) */
    int32_t slots_t;    /* transposed ELL slots: pattern of the transposed
                           matrices and derivative tables */
    int32_t n_exp;
    int32_t n_node;
    int32_t fine;
    int32_t n_out;
    int32_t taylor;     /* degree of the forward (recomputation) */
    int32_t taylor_adj; /* planned degree of the adjoint action, 1..10;
                           the kernel evaluates degree 10 (a lower plan is
                           raised, which only adds accuracy) */
    int32_t n_obs;      /* 1..2 */
    int32_t flags;      /* 0 or SSMX_COEF_WAVE */
    int32_t first_out;
    int32_t partner;    /* slot 1 is column r ^ 1 in both ELL stacks */
    int32_t ckpt_every; /* K, as the forward call that wrote ckpt */
    int32_t n_theta;    /* derivative tables, 0..16 */
    int32_t theta_gen;  /* matrix the derivative tables belong to (or -1) */
    int32_t grad_gen;   /* matrix whose waveform segments get amplitude
                           gradients (or -1; needs SSMX_COEF_WAVE) */
    int32_t reserved;   /* 0; n > 64: the matrices' slot count */
    int64_t coef_inst_stride;
    double h;
    double t0;
    double dt;
} ssmx_adjoint_desc;

#define SSMX_ADJ_MAX_THETA 16

/* Bytes of device workspace ssmx_adjoint needs: n_inst * 2 * ckpt_every *
 * n_exp * 64 doubles (n > 64: n_inst * ckpt_every * n_exp * NP). */
int ssmx_adjoint_workspace_bytes(const ssmx_adjoint_desc *desc,
                                 uint64_t *bytes);

/*
 * ell_col, ell_val   forward ELL of the M matrices (ssmx_propagate layout)
 * ell_col_t          int32 [slots_t * 64]: ELL of the transposed union
 *                    pattern; entry (slot, c) names row r of the original
 *                    matrices (slot 0 = diagonal, holds the used count)
 * ell_val_t          double [M * slots_t * 64]: G_m[r][c] at (slot, c)
 * dval_t             double [n_theta * slots_t * 64]: D_t[r][c] at
 *                    (slot, c) (NULL when n_theta = 0)
 * cf_weights, cf_nodes, coef, segments   as ssmx_propagate
 * ckpt               double [n_inst * ceil(n_out * fine / K) * n]
 * obs                double [n_obs * 64] readout rows
 * seed               double [n_inst * n_obs] dLoss/dF
 * workspace          ssmx_adjoint_workspace_bytes() bytes
 * grad_theta         double [n_inst * n_theta] (or NULL)
 * grad_seg           double [number of segment records], accumulated
 *                    with += (zero it first; or NULL)
 * lambda_out         double [n_inst * n] (or NULL)
 */
int ssmx_adjoint(const ssmx_adjoint_desc *desc, const int32_t *ell_col,
                 const double *ell_val, const int32_t *ell_col_t,
                 const double *ell_val_t, const double *dval_t,
                 const double *cf_weights, const double *cf_nodes,
                 const void *coef, const ssmx_segment *segments,
                 const double *ckpt, const double *obs, const double *seed,
                 double *workspace, double *grad_theta, double *grad_seg,
                 double *lambda_out, void *stream);

/* Message of the last failure on this thread ("" if none). */
const char *ssmx_last_error(void);

/* ABI version (major * 100 + minor). */
int ssmx_version(void);

/* Number of HIP devices visible, or a negative code. */
int ssmx_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* SSMX_H */
