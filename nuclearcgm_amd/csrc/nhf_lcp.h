// nhf_lcp.h -- learned 3D convolutional preconditioner of the GCG W-block.
//
// The net maps a residual column r of (H + sigma) w = (lambda + sigma) x to
// an approximate correction c ~ (H + sigma)^-1 r.  It works in lattice units:
// with s = rms(r) over the column and the local scales
//   kappa(p) = h^2 (U(p) + sigma) / f(p),   hf(p) = h^2 / f(p),
// the input per lattice point is (r/s [4 real spinor channels], kappa) and the
// output is c = s hf(p) net(...).  Since (H + sigma) ~ (f/h^2)(-lap_lattice) +
// (U + sigma), the same weights serve lattices of different spacing and
// nuclei of different size.
//
// Architecture: four 3x3x3 convolutions, 5 -> 16 -> 16 -> 16 -> 4 channels,
// zero padding at the box faces, LeakyReLU(0.1) after the first three.
// Parameters (FP32 master copy), per layer l: W_l[tap][ci][co], then b_l[co];
// tap = (dz + 1) 9 + (dy + 1) 3 + (dx + 1).
#pragma once
#include <mutex>
#include <vector>

#include "nhf_internal.h"

namespace nhf {

constexpr int LCP_LAYERS = 4;
constexpr int LCP_HID = 16;
constexpr int LCP_TAPS = 27;
constexpr int LCP_IN_STRIDE = 8;  // stored input channels (5 used)
constexpr float LCP_SLOPE = 0.1f;
constexpr int LCP_CI[LCP_LAYERS] = {5, 16, 16, 16};
constexpr int LCP_CO[LCP_LAYERS] = {16, 16, 16, 4};

constexpr int lcp_layer_size(int l) { return LCP_TAPS * LCP_CI[l] * LCP_CO[l] + LCP_CO[l]; }
constexpr int lcp_layer_offset(int l) { return l == 0 ? 0 : lcp_layer_offset(l - 1) + lcp_layer_size(l - 1); }
constexpr int LCP_NPARAM = lcp_layer_offset(LCP_LAYERS);

struct LcpPair {
  int n;
  float* data;  // device: n^3 x 8 input channels, then n^3 x 4 target channels
  double tn2;   // |A target|^2 (A: the loss operator)
};

struct Lcp {
  int device = 0;
  bool ready = false;  // parameters initialised or set
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr, *part = nullptr;
  double* loss = nullptr;  // device accumulators
  void* frags = nullptr;   // BF16 MFMA weight fragments of the inference kernels
  int adam_t = 0;
  // training pairs and their collection
  std::vector<LcpPair> pairs;
  int collect_left = 0, collect_every = 1;
  std::mutex mu;  // the two species threads add pairs concurrently
  // training / test scratch (batch x n^3 points)
  size_t scratch_pts = 0;
  float *x0 = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr, *y = nullptr, *t = nullptr, *gy = nullptr,
        *g1 = nullptr, *g2 = nullptr;
  ~Lcp();
  int ensure_scratch(size_t pts);
};

int lcp_init(Lcp& L, unsigned long long seed);
// Per-point scales of one species: kap = h^2 (U + shift) / f, hf = h^2 / f.
template <typename R>
void launch_lcp_planes(hipStream_t st, const R* fld, size_t N3, double h, double shift, float* kap, float* hf);
// z_c = M(r_c) for the nb columns clist[0..nb) (all if clist is null);
// rn2[c * rn2_stride] = |r_c|^2.  h0, h1: scratch of nb n^3 x 16 floats each.  BF16
// MFMA kernels with FP32 accumulation; the fragments must be current
// (lcp_make_frags after every change of the parameters).
template <typename R>
void lcp_apply_fp32(const Lcp& L, hipStream_t st, const Sp<R>* r, Sp<R>* z, size_t ld, int nb, const int* clist,
                    const double* rn2, const float* kap, const float* hf, int n, float* h0, float* h1);
int lcp_make_frags(Lcp& L, hipStream_t st);
template <typename R>
void lcp_apply(const Lcp& L, hipStream_t st, const Sp<R>* r, Sp<R>* z, size_t ld, int nb, const int* clist,
               const double* rn2, const float* kap, const float* hf, int n, float* h0, float* h1, int rn2_stride = 1);
// Training pairs from a finished plain W-block: r_c = T_c - ls_c X_c with
// T = (H + shift) X, correction X_c - W_c.  Host-synchronous.
template <typename R>
int lcp_add_pairs(Lcp& L, hipStream_t st, const Sp<R>* T, const Sp<R>* X, const Sp<R>* W, const double* ls,
                  int ncols, size_t ld, const float* kap, const float* hf, int n, double* scratch2);
// Forward (FP32) of nb samples already in L.x0 (nb n^3 x 8); output in L.y.
void lcp_forward_batch(Lcp& L, hipStream_t st, int n, int nb);
// Loss sum_b wts[b] |A (y_b - t_b)|^2 (targets in L.t; A = -lap4 + kappa
// in lattice units) and its parameter gradient into L.g; the loss is added
// to *L.loss.
void lcp_loss_grad_batch(Lcp& L, hipStream_t st, int n, int nb, const double* wts);
void lcp_adam(Lcp& L, hipStream_t st, float lr);
// |A t_b|^2 of the nb targets in L.t (inputs in L.x0), to the host
int lcp_target_norms(Lcp& L, hipStream_t st, int n, int nb, double* out);
// Adam on random batches of the stored pairs (relative squared error of the
// correction); mean loss of the first and last up to 10 steps.
// learned W-block: preconditioned minimal-residual steps (see
// nhf_lcp.hip); LW_NS scalars per iteration and column
constexpr int LW_NS = 4;
template <typename R>
void launch_lw_init(hipStream_t st, const Sp<R>* X, const Sp<R>* T, const double* ls, Sp<R>* W, Sp<R>* Rr, size_t N3,
                    size_t ld, int m, double* sc);
template <typename R>
void launch_lw_dots(hipStream_t st, const Sp<R>* Q, const Sp<R>* Rr, size_t N3, size_t ld, int m, double* sc);
template <typename R>
void launch_lw_update(hipStream_t st, const Sp<R>* Z, const Sp<R>* Q, Sp<R>* W, Sp<R>* Rr, size_t N3, size_t ld, int m,
                      const double* sc, double* sc_next);
int lcp_train(Lcp& L, hipStream_t st, int steps, int batch, double lr, unsigned long long seed, double* loss_first,
              double* loss_last);

}  // namespace nhf
