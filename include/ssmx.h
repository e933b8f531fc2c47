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
 *  - Superoperators are n x n real matrices (n <= 64), stored in a shared
 *    ELL pattern: column index ell_col[slot * 64 + row] (int32).  Slot 0
 *    of row r is always column r (the diagonal, value possibly 0) and
 *    ell_col[row] (slot 0) instead holds the number of used slots of that
 *    row; unused slots point at column r with value 0; values
 *    ell_val[((node * n_gen + g) * slots + slot) * 64 + row]; node ranges
 *    over n_node for the rotating frame (node-wise interaction-picture
 *    generators) and is 0 otherwise.
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

/* limits */
#define SSMX_MAX_N 64
#define SSMX_MAX_GEN 8
#define SSMX_MAX_EXP 6
#define SSMX_MAX_NODE 3
#define SSMX_MAX_OBS 2
#define SSMX_MAX_TAYLOR 40

typedef struct ssmx_propagate_desc {
    int32_t n;        /* superoperator dimension, 1..64 */
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
 * ell_col    int32 [slots * 64]
 * ell_val    double [(n_node or 1) * n_gen * slots * 64]
 * cf_weights double [n_exp * n_node]     (row j: weights of exponential j)
 * cf_nodes   double [n_node]             (Gauss-Legendre nodes on [0, 1])
 * coef       table mode: double [n_inst? * n_out * fine * n_node * n_gen]
 *            ordered (instance, i, r, s, g); wave mode: ssmx_wave [n_inst]
 * segments   wave mode: ssmx_segment [...]; table mode: unused (NULL)
 * rho0       double [n_inst * n]
 * eig        rotating: double [2 * 64] per-row (own, partner) factors of
 *            exp(L0 h) - I; partner of row r is row r ^ 1
 * obs        double [n_obs * 64] readout rows (obs readouts, n_obs <= 2)
 * obs_offset double [n_inst * n_obs] added to each obs readout (or NULL)
 * out        READ_FULL: double [n_inst * n_out * n]
 *            obs: double [n_inst * n_out * n_obs]
 *            READ_INTEGRAL: double [n_inst * n_obs]
 * state_out  double [n_inst * n] final state (or NULL)
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
