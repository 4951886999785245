/* moloch_lac.h -- learned acoustic corrector of the MOLOCH timestep.
 *
 * No counterpart in the reference: its acoustic loop always runs nsound
 * substeps of dtstepa/nsound (nompi/mod_moloch.F90:326-329, :483-851).  This
 * API trains, on the device, a column-stencil network that replaces the
 * first nsound-1 substeps of every acoustic phase by one substep of
 * (nsound-1)*dts plus a learned increment, after which the last ordinary
 * substep runs unchanged; a periodic check against the ordinary loop falls
 * back to it when the increment stops helping.
 *
 * The network (nin -> nhid -> nhid -> nout, ReLU) runs on the matrix cores
 * with BF16 operands and FP32 accumulation; nin = 9 * roundup(6 kz, 16)
 * (u, v, w, Exner, theta, qv on the 3 x 3 stencil of columns; 2160 at
 * kz = 40) and nout = 4 kz (increments of u, v, w and Exner).
 */
#ifndef MOLOCH_LAC_H
#define MOLOCH_LAC_H

#include "moloch_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- network */
typedef struct mlc_mlp mlc_mlp;

/* A network nin -> nhid -> nhid -> nout with He-uniform weights from seed,
 * zero biases and its own stream on device `device`. */
int mlc_mlp_create(int nin, int nhid, int nout, unsigned long long seed, int device, mlc_mlp **out);
int mlc_mlp_destroy(mlc_mlp *p);
/* Copy the FP32 weights w [in][out] and biases b [out] of layer 0..2 from
 * (set = 1) or to (set = 0) host memory; either pointer may be NULL. */
int mlc_mlp_weights(mlc_mlp *p, int layer, float *w, float *b, int set);
/* Gradients of layer 0..2 from the last mlc_mlp_grad. */
int mlc_mlp_gradients(mlc_mlp *p, int layer, float *gw, float *gb);
/* Y [rows x nout] = network(X [rows x nin]); host arrays, X rounded to BF16. */
int mlc_mlp_forward(mlc_mlp *p, const float *X, int rows, float *Y);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Y [rows x nout] = network(X [rows x nin]) like
 * mlc_mlp_forward, the same kernels and bits, but X and Y are DEVICE arrays
 * (fp32) and the bf16 inputs and hidden activations live in a fixed 64 MiB
 * scratch of the library on the network's device, so the call allocates
 * nothing; rows beyond what fits (rows * (2 nin + 4 nhid) bytes) are refused
 * with MLC_EINVAL.  Calls that overlap on one device share that scratch and
 * race; the call returns after the network's stream has finished. */
int mlc_mlp_forward_noalloc(mlc_mlp *p, const float *X, int rows, float *Y);
/* Mean-square-error loss against T [rows x nout] and its gradients. */
int mlc_mlp_grad(mlc_mlp *p, const float *X, const float *T, int rows, double *loss);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the last gradients. */
int mlc_mlp_adam(mlc_mlp *p, double lr);
/* C [M x N] = A . B on the MFMA GEMM with host fp32 operands rounded to BF16;
 * element (m,k) of A at A[m*sam + k*sak] (asize elements in all), (k,n) of B
 * at B[k*sbk + n*sbn]. */
int mlc_gemm_bf16_host(const float *A, long sam, long sak, long asize, const float *B, long sbk, long sbn,
                       long bsize, int M, int N, int K, float *C);

/* -------------------------------------------------- corrector of a handle */
#define MLC_LAC_OFF 0     /* ordinary acoustic loop (the reference's)         */
#define MLC_LAC_ON 1      /* long substep + learned increment + last substep  */
#define MLC_LAC_REDUCED 2 /* long substep + last substep, no increment        */

typedef struct mlc_lac_config {
  int nhid;              /* hidden width (512)                                 */
  int samples_per_phase; /* columns sampled per acoustic phase by collect      */
  int capacity;          /* rows kept in the training set (a ring)             */
  int chunk_rows;        /* rows per inference batch (a multiple of 256)       */
  int check_every;       /* check every this many MLC_LAC_ON steps (0: never)  */
  double check_tol;      /* fall back when a field's error ratio exceeds it    */
                         /* or is NaN                                          */
  unsigned long long seed;
  int engine;            /* inference: 1 fused MFMA kernels reading the fields,
                          * 0 materialised features + the generic GEMM       */
  /* the error budget of a check: fall back also when a field's relative L2
   * error with the increment (rel below) exceeds rel_tol or is NaN; 0: no
   * budget, only check_tol                                                 */
  double rel_tol;
  /* after a failed check: retrain_steps ordinary steps that collect fresh
   * pairs (the training set restarts), then retrain_iters Adam steps of
   * retrain_batch rows at retrain_lr with the normalisation recomputed,
   * then the corrector resumes with a check; at most max_retrains times in
   * a row without a passed check, after which it stays off until trained
   * again.  retrain_steps = 0: no retraining (stay off)                    */
  int retrain_steps;
  int retrain_iters;
  int retrain_batch;
  double retrain_lr;
  int max_retrains;
} mlc_lac_config;

typedef struct mlc_lac_stats {
  long long samples;        /* rows in the training set                       */
  long long learned_phases; /* acoustic phases run with the increment         */
  long long plain_phases;   /* acoustic phases run with the ordinary loop     */
  long long checks;         /* checks run                                     */
  int fallback;             /* 1 once a check failed (until the next train)   */
  double last_loss;         /* last training loss (normalised targets)        */
  /* last check, per field u, v, w (levels 2..kz), Exner: ratio of the error
   * with the increment to the error without it, and the relative L2 error
   * with the increment, both against the ordinary substeps */
  double ratio[4];
  double rel[4];
  double worst_ratio;       /* largest ratio over the phases of the last check */
  double worst_rel;         /* largest rel over the phases of the last check   */
  long long retrains;       /* retrainings after failed checks                 */
  long long collect_phases; /* acoustic phases run to collect retraining pairs */
  long long failed_checks;  /* checks that failed                              */
} mlc_lac_stats;

int mlc_lac_default_config(mlc_lac_config *c);
/* Attach a corrector to the handle (after mlc_init).  Needs nsound >= 2. */
int mlc_lac_create(mlc_handle *h, const mlc_lac_config *c);
/* Run nsteps ordinary steps (the state evolves exactly as under mlc_step
 * with MLC_LAC_OFF) while adding (input, target) pairs of sampled columns
 * from every acoustic phase to the training set. */
int mlc_lac_collect(mlc_handle *h, int nsteps);
/* iters Adam steps on random batches of the training set; the first call
 * fixes the input and target normalisation (a retraining after a failed
 * check recomputes it).  *loss (may be NULL) receives the last batch's
 * loss.  With iters > 0 it clears the fallback flag. */
int mlc_lac_train(mlc_handle *h, int iters, int batch, double lr, double *loss);
/* Acoustic loop used by mlc_step: MLC_LAC_OFF, MLC_LAC_ON, MLC_LAC_REDUCED. */
int mlc_lac_mode(mlc_handle *h, int mode);
int mlc_lac_get_stats(mlc_handle *h, mlc_lac_stats *s);
/* Copy rows first .. first+n-1 of the training set: raw inputs X [n x nin]
 * (feature order of the network), targets T [n x nout] (ordinary substeps
 * minus the long substep) and the sampled columns (i-1)*jx + (j-1); any
 * pointer may be NULL. */
int mlc_lac_dataset(mlc_handle *h, long first, int n, float *X, float *T, int *cols);
/* The corrector's network (owned by the handle; for weight access). */
mlc_mlp *mlc_lac_network(mlc_handle *h);

#ifdef __cplusplus
}
#endif
#endif
