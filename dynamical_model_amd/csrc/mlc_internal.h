// mlc_internal.h -- helpers shared by the translation units of the library.
#pragma once

// Record an error message for mlc_last_error() and return the code.
int mlc_fail(int code, const char *fmt, ...);

// Host copies of reference functions used during set-up.
double gzita_host(double zita, double ztop, double a0);  // mod_zita.F90:110-120
double pfwsat_host(double t, double p);                   // nompi/mod_moloch.F90:119-181

// ------------------------------------------------------------ the MLP engine
#include <hip/hip_runtime.h>
#include <cstdint>
#include <vector>

// Host part of a case (levels, 2-D geometry, surface pressure) in the
// canonical layout (-1:jx+2, -1:iy+2), and the state of the case's
// splitmix64 generator after the terrain draws.
struct mlc_case;
struct CaseHost {
  std::vector<double> zita, zitah, ht, hx, hy, msf, ulat, vlat, ps;
  uint64_t rng = 0;
};
int case_host_part(const mlc_case *c, CaseHost &o);

struct mlc_mlp {
  int nin, nhid, nout;
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = true;  // false when the stream belongs to a model handle
  // fp32 master weights (row-major [in][out]) and biases, bf16 copies
  float *W[3] = {}, *bvec[3] = {};
  __bf16 *Wb[3] = {};
  float *gW[3] = {}, *gb[3] = {};
  float *mW[3] = {}, *vW[3] = {}, *mb[3] = {}, *vb[3] = {};
  long step = 0;
  // activations of a batch of up to cap rows
  int cap = 0;
  __bf16 *H1 = nullptr, *H2 = nullptr, *dH1 = nullptr, *dH2 = nullptr, *dYb = nullptr;
  float *Y = nullptr, *dYf = nullptr, *dH1f = nullptr, *dH2f = nullptr;
  double *partial = nullptr;
  // split-k partial sums of the gradient products (ws_cap floats)
  float *ws = nullptr;
  long ws_cap = 0;
  // weights packed as MFMA B fragments for the fused inference kernels:
  // fragment (ks, nt) holds 64 lanes x 8 bf16, lane = n%32 + 32*((k%16)/8),
  // element e = k%8, for k-step ks = k/16 and n-tile nt = n/32; the n extent
  // is padded to a multiple of 256 and k to a multiple of 16 with zeros.
  __bf16 *Wp[3] = {};
  int wp_nt[3] = {}, wp_ks[3] = {};
  bool pack_dirty = true;
  std::vector<void *> allocs;
  int in_dim(int l) const { return l == 0 ? nin : nhid; }
  int out_dim(int l) const { return l == 2 ? nout : nhid; }
};

// A network on device `device`; with stream == nullptr it creates and owns a
// stream, otherwise it queues its work on `stream`.
int mlp_new(int nin, int nhid, int nout, unsigned long long seed, int device, hipStream_t stream, mlc_mlp **out);
// Release the device memory (and an owned stream) of a network.
void mlp_free(mlc_mlp *p);
// Forward of rows X [rows x nin] (bf16, row stride ldx) into Yout (fp32,
// row stride ldy; the engine's own buffer when Yout is null).
int mlp_forward_dev(mlc_mlp *p, const __bf16 *X, long ldx, int rows, float *Yout, long ldy);
// Loss (mean square error) and gradients of one batch with fp32 targets T.
int mlp_backward_dev(mlc_mlp *p, const __bf16 *X, long ldx, const float *T, int rows, double *loss);
// One Adam step with the gradients of the last backward.
int mlp_adam_dev(mlc_mlp *p, double lr);
// Make room for activations of `rows` rows (H1, H2, ...).
int mlp_reserve_rows(mlc_mlp *p, int rows);
// Refresh the packed weight fragments if the weights changed.
int mlp_pack_dev(mlc_mlp *p);
