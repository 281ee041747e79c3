// Training of the learned LAD coefficient on the GPU (ajet_trainer_*, include/ajet_capi.h) and
// the packing of an ajet_mlp into the device layouts the kernels read.  The network has no
// counterpart in the reference; its targets are the beta* and kappa* of
// ComputeArtTransCoeffs (ModShockCapturing.f90:5-106).
//
// One Adam step: minibatch rows -> three GEMMs with the bias + GELU epilogue (activations kept
// as B x 128 and 128 x B in BF16, pre-activations in FP32) -> output GEMM -> dy -> for each
// layer the split-K weight-gradient GEMM, the bias column sums and the GEMM back through W^T
// with the GELU' epilogue -> Adam on every layer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ajet_internal.h"
#include "ajet_kernels.h"

using namespace ajet;

namespace ajet {

uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u = u + 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

// W: (3 x 128 + 16) x 128 BF16 bits (W1 with K padded to 128, W2, W3, W4 in rows 384, 385 and
// zero rows up to 399); bias: b1, b2, b3, b4 (386 floats).
void mlp_pack(const ajet_mlp& m, std::vector<uint16_t>& W, std::vector<float>& bias) {
  W.assign((size_t)(3 * MLP_H + 16) * MLP_K, 0);
  bias.assign(3 * MLP_H + MLP_NOUT, 0.0f);
  for (int r = 0; r < MLP_H; ++r) {
    for (int k = 0; k < MLP_NIN; ++k) W[(size_t)r * MLP_K + k] = bf16_bits(m.w1[r][k]);
    for (int k = 0; k < MLP_H; ++k) {
      W[(size_t)(MLP_H + r) * MLP_K + k] = bf16_bits(m.w2[r][k]);
      W[(size_t)(2 * MLP_H + r) * MLP_K + k] = bf16_bits(m.w3[r][k]);
    }
    bias[r] = m.b1[r];
    bias[MLP_H + r] = m.b2[r];
    bias[2 * MLP_H + r] = m.b3[r];
  }
  for (int r = 0; r < MLP_NOUT; ++r) {
    for (int k = 0; k < MLP_H; ++k) W[(size_t)(3 * MLP_H + r) * MLP_K + k] = bf16_bits(m.w4[r][k]);
    bias[3 * MLP_H + r] = m.b4[r];
  }
}

}  // namespace ajet

namespace {

constexpr int NL = 4;
constexpr int NOUTS[NL] = {MLP_H, MLP_H, MLP_H, MLP_NOUT_PAD};  // rows of each layer's W

}  // namespace

struct ajet_trainer {
  int batch = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  float *p[NL] = {}, *m[NL] = {}, *v[NL] = {};  // FP32 masters [W (N x 128), b (N)] and Adam moments
  __bf16 *W[NL] = {}, *WT[NL] = {};
  float* bias[NL] = {};
  __bf16 *a[NL] = {}, *aT[NL] = {};  // a0 .. a3: batch x 128 and 128 x batch
  float* z[NL] = {};                 // z1 .. z3 (index 1 .. 3): batch x 128
  float* y = nullptr;                // batch x 32
  __bf16 *dz[NL + 1] = {}, *dzT[NL + 1] = {};  // dz1 .. dz4
  float *part[NL] = {}, *db[NL] = {}, *grad[NL] = {};
  float* tb = nullptr;
  long long* idx = nullptr;
  float *X = nullptr, *T = nullptr;
  long long nsamples = 0;
  double* loss = nullptr;  // per step
  int loss_cap = 0;
  float* xin = nullptr;    // test hooks: rows as given
  float* tin = nullptr;
  long long tstep = 0;
  ajet_mlp meta{};
  std::vector<void*> bufs;
  std::string err;
};

static int tfail(ajet_trainer* t, int code, const std::string& msg) {
  if (t) t->err = msg;
  return code;
}

#define THIP(t, expr)                                                           \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) return tfail((t), (int)e_, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
static hipError_t talloc(ajet_trainer* t, T** ptr, size_t n) {
  hipError_t e = hipMalloc((void**)ptr, sizeof(T) * (n ? n : 1));
  if (e == hipSuccess) t->bufs.push_back(*ptr);
  return e;
}

// Split-K of the weight-gradient GEMM over n rows: slices of 256 rows when n allows it.
static int splits_for(int n) { return (n % 256 == 0 && n >= 256) ? n / 256 : 1; }

// Forward over n rows already in a[0] / aT[0].
static hipError_t forward(ajet_trainer* t, int n) {
  hipError_t e;
  for (int l = 0; l < NL; ++l) {
    MlpGemm g{};
    g.A = t->a[l];
    g.lda = MLP_K;
    g.B = t->W[l];
    g.ldb = MLP_K;
    g.M = n;
    g.N = NOUTS[l];
    g.K = MLP_K;
    g.splits = 1;
    g.bias = t->bias[l];
    if (l < NL - 1) {
      g.Z = t->z[l + 1];
      g.O = t->a[l + 1];
      g.OT = t->aT[l + 1];
      e = launch_mlp_gemm(MLP_EPI_FWD_GELU, g, t->stream);
    } else {
      g.Z = t->y;
      e = launch_mlp_gemm(MLP_EPI_FWD_LAST, g, t->stream);
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Loss and backward over n rows with targets tb; leaves the weight-gradient partials in part
// and the bias gradients in db.
static hipError_t backward(ajet_trainer* t, int n, double* loss) {
  hipError_t e = launch_mlp_loss(t->y, t->tb, n, t->dz[4], t->dzT[4], loss, t->stream);
  if (e != hipSuccess) return e;
  const int S = splits_for(n);
  for (int l = NL - 1; l >= 0; --l) {
    const int N = NOUTS[l];
    MlpGemm g{};
    g.A = t->dzT[l + 1];  // N x n
    g.lda = n;
    g.B = t->aT[l];       // 128 x n
    g.ldb = n;
    g.M = N;
    g.N = MLP_K;
    g.K = n / S;
    g.splits = S;
    g.Z = t->part[l];
    if ((e = launch_mlp_gemm(MLP_EPI_PARTIAL, g, t->stream)) != hipSuccess) return e;
    if ((e = launch_mlp_bias_grad(t->dzT[l + 1], n, N, t->db[l], t->stream)) != hipSuccess) return e;
    if (l == 0) break;
    MlpGemm h{};
    h.A = t->dz[l + 1];   // n x N
    h.lda = N;
    h.B = t->WT[l];       // 128 x N
    h.ldb = N;
    h.M = n;
    h.N = MLP_K;
    h.K = N;
    h.splits = 1;
    h.Zin = t->z[l];
    h.O = t->dz[l];
    h.OT = t->dzT[l];
    if ((e = launch_mlp_gemm(MLP_EPI_BWD_GELU, h, t->stream)) != hipSuccess) return e;
  }
  return hipSuccess;
}

static MlpAdam adam_args(ajet_trainer* t, int l, int n, float lr, int grad_only) {
  MlpAdam a{};
  a.p = t->p[l];
  a.m = t->m[l];
  a.v = t->v[l];
  a.partial = t->part[l];
  a.db = t->db[l];
  a.splits = splits_for(n);
  a.N = NOUTS[l];
  a.K = MLP_K;
  a.W = t->W[l];
  a.WT = t->WT[l];
  a.b = t->bias[l];
  const double beta1 = 0.9, beta2 = 0.999;
  const double bc1 = 1.0 - std::pow(beta1, (double)t->tstep), bc2 = 1.0 - std::pow(beta2, (double)t->tstep);
  a.beta2 = (float)beta2;
  a.one_m_beta1 = (float)(1.0 - beta1);
  a.one_m_beta2 = (float)(1.0 - beta2);
  a.step = grad_only ? 0.0f : (float)(lr / bc1);
  a.bc2_sqrt = grad_only ? 1.0f : (float)std::sqrt(bc2);
  a.eps = 1e-8f;
  a.grad_only = grad_only;
  a.grad = t->grad[l];
  return a;
}

// Upload the masters of a model (zero padding in W1's columns >= 104 and W4's rows >= 2) and
// write the BF16 copies in both layouts.
static int set_model(ajet_trainer* t, const ajet_mlp& mdl) {
  for (int l = 0; l < NL; ++l) {
    const int N = NOUTS[l];
    std::vector<float> h((size_t)N * MLP_K + N, 0.0f);
    for (int r = 0; r < N; ++r) {
      for (int k = 0; k < MLP_K; ++k) {
        float w = 0.0f;
        if (l == 0 && k < MLP_NIN) w = mdl.w1[r][k];
        if (l == 1) w = mdl.w2[r][k];
        if (l == 2) w = mdl.w3[r][k];
        if (l == 3 && r < MLP_NOUT) w = mdl.w4[r][k];
        h[(size_t)r * MLP_K + k] = w;
      }
      float b = 0.0f;
      if (l == 0) b = mdl.b1[r];
      if (l == 1) b = mdl.b2[r];
      if (l == 2) b = mdl.b3[r];
      if (l == 3 && r < MLP_NOUT) b = mdl.b4[r];
      h[(size_t)N * MLP_K + r] = b;
    }
    std::vector<uint16_t> w16((size_t)N * MLP_K), wt16((size_t)N * MLP_K);
    for (int r = 0; r < N; ++r)
      for (int k = 0; k < MLP_K; ++k) {
        const uint16_t b = bf16_bits(h[(size_t)r * MLP_K + k]);
        w16[(size_t)r * MLP_K + k] = b;
        wt16[(size_t)k * N + r] = b;
      }
    THIP(t, hipMemcpy(t->p[l], h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    THIP(t, hipMemcpy(t->bias[l], h.data() + (size_t)N * MLP_K, sizeof(float) * N, hipMemcpyHostToDevice));
    THIP(t, hipMemcpy(t->W[l], w16.data(), 2 * w16.size(), hipMemcpyHostToDevice));
    THIP(t, hipMemcpy(t->WT[l], wt16.data(), 2 * wt16.size(), hipMemcpyHostToDevice));
    THIP(t, hipMemset(t->m[l], 0, sizeof(float) * h.size()));
    THIP(t, hipMemset(t->v[l], 0, sizeof(float) * h.size()));
  }
  t->tstep = 0;
  return AJET_OK;
}

// Unpack layer arrays [W (N x 128), b (N)] into the w and b members of an ajet_mlp.
static void unpack(const std::vector<float> (&h)[NL], ajet_mlp& out) {
  for (int r = 0; r < MLP_H; ++r) {
    for (int k = 0; k < MLP_NIN; ++k) out.w1[r][k] = h[0][(size_t)r * MLP_K + k];
    for (int k = 0; k < MLP_H; ++k) {
      out.w2[r][k] = h[1][(size_t)r * MLP_K + k];
      out.w3[r][k] = h[2][(size_t)r * MLP_K + k];
    }
    out.b1[r] = h[0][(size_t)MLP_H * MLP_K + r];
    out.b2[r] = h[1][(size_t)MLP_H * MLP_K + r];
    out.b3[r] = h[2][(size_t)MLP_H * MLP_K + r];
  }
  for (int r = 0; r < MLP_NOUT; ++r) {
    for (int k = 0; k < MLP_H; ++k) out.w4[r][k] = h[3][(size_t)r * MLP_K + k];
    out.b4[r] = h[3][(size_t)MLP_NOUT_PAD * MLP_K + r];
  }
}

extern "C" {

int ajet_trainer_create(int32_t batch, const ajet_mlp* init, ajet_trainer** out) {
  if (!out) return AJET_EINVAL;
  *out = nullptr;
  if (!init || batch < 32 || batch % 32 != 0 || batch > (1 << 20)) return AJET_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return AJET_ENODEV;
  ajet_trainer* t = new ajet_trainer();
  *out = t;
  t->batch = batch;
  t->meta = *init;
  THIP(t, hipGetDevice(&t->device));
  THIP(t, hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
  const size_t B = (size_t)batch;
  for (int l = 0; l < NL; ++l) {
    const size_t np = (size_t)NOUTS[l] * MLP_K + NOUTS[l];
    THIP(t, talloc(t, &t->p[l], np));
    THIP(t, talloc(t, &t->m[l], np));
    THIP(t, talloc(t, &t->v[l], np));
    THIP(t, talloc(t, &t->grad[l], np));
    THIP(t, talloc(t, &t->W[l], (size_t)NOUTS[l] * MLP_K));
    THIP(t, talloc(t, &t->WT[l], (size_t)NOUTS[l] * MLP_K));
    THIP(t, talloc(t, &t->bias[l], (size_t)NOUTS[l]));
    THIP(t, talloc(t, &t->a[l], B * MLP_K));
    THIP(t, talloc(t, &t->aT[l], B * MLP_K));
    THIP(t, talloc(t, &t->part[l], (size_t)splits_for(batch) * NOUTS[l] * MLP_K));
    THIP(t, talloc(t, &t->db[l], (size_t)NOUTS[l]));
    if (l >= 1) THIP(t, talloc(t, &t->z[l], B * MLP_H));
  }
  for (int l = 1; l <= NL; ++l) {
    const size_t w = l == NL ? MLP_NOUT_PAD : MLP_H;
    THIP(t, talloc(t, &t->dz[l], B * w));
    THIP(t, talloc(t, &t->dzT[l], B * w));
  }
  THIP(t, talloc(t, &t->y, B * MLP_NOUT_PAD));
  THIP(t, talloc(t, &t->tb, B * 2));
  THIP(t, talloc(t, &t->idx, B));
  THIP(t, talloc(t, &t->xin, B * MLP_NIN));
  THIP(t, talloc(t, &t->tin, B * 2));
  return set_model(t, *init);
}

int ajet_trainer_destroy(ajet_trainer* t) {
  if (!t) return AJET_EINVAL;
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  for (void* b : t->bufs) (void)hipFree(b);
  if (t->X) (void)hipFree(t->X);
  if (t->T) (void)hipFree(t->T);
  if (t->loss) (void)hipFree(t->loss);
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
  return AJET_OK;
}

const char* ajet_trainer_last_error(const ajet_trainer* t) { return t ? t->err.c_str() : "null trainer"; }

int ajet_trainer_load(ajet_trainer* t, const float* x, const float* tg, int64_t n) {
  if (!t || !x || !tg || n < 1) return AJET_EINVAL;
  if (t->X) (void)hipFree(t->X);
  if (t->T) (void)hipFree(t->T);
  t->X = t->T = nullptr;
  THIP(t, hipMalloc((void**)&t->X, sizeof(float) * (size_t)n * MLP_NIN));
  THIP(t, hipMalloc((void**)&t->T, sizeof(float) * (size_t)n * 2));
  THIP(t, hipMemcpy(t->X, x, sizeof(float) * (size_t)n * MLP_NIN, hipMemcpyHostToDevice));
  THIP(t, hipMemcpy(t->T, tg, sizeof(float) * (size_t)n * 2, hipMemcpyHostToDevice));
  t->nsamples = n;
  return AJET_OK;
}

int ajet_trainer_run(ajet_trainer* t, int32_t steps, float lr, uint64_t seed, double* losses) {
  if (!t || steps < 0) return AJET_EINVAL;
  if (!t->X) return tfail(t, AJET_EINVAL, "no samples loaded");
  if (steps == 0) return AJET_OK;
  if (steps > t->loss_cap) {
    if (t->loss) (void)hipFree(t->loss);
    t->loss = nullptr;
    THIP(t, hipMalloc((void**)&t->loss, sizeof(double) * (size_t)steps));
    t->loss_cap = steps;
  }
  THIP(t, hipMemsetAsync(t->loss, 0, sizeof(double) * (size_t)steps, t->stream));
  const int B = t->batch;
  // exactly one batch of samples: every step takes it whole (full-batch), else a minibatch
  // drawn with replacement
  const bool full = t->nsamples == B;
  for (int s = 0; s < steps; ++s) {
    if (!full) THIP(t, launch_mlp_sample(seed, t->tstep, t->nsamples, B, t->idx, t->stream));
    THIP(t, launch_mlp_rows(t->X, t->T, full ? nullptr : t->idx, B, t->a[0], t->aT[0], t->tb, t->stream));
    THIP(t, forward(t, B));
    THIP(t, backward(t, B, t->loss + s));
    ++t->tstep;
    for (int l = 0; l < NL; ++l) THIP(t, launch_mlp_adam(adam_args(t, l, B, lr, 0), t->stream));
  }
  if (losses) THIP(t, hipMemcpyAsync(losses, t->loss, sizeof(double) * (size_t)steps, hipMemcpyDeviceToHost, t->stream));
  THIP(t, hipStreamSynchronize(t->stream));
  return AJET_OK;
}

static int load_rows(ajet_trainer* t, const float* x, const float* tg, int32_t n) {
  if (n < 32 || n % 32 != 0 || n > t->batch) return tfail(t, AJET_EINVAL, "n must be a multiple of 32 up to the batch");
  THIP(t, hipMemcpyAsync(t->xin, x, sizeof(float) * (size_t)n * MLP_NIN, hipMemcpyHostToDevice, t->stream));
  if (tg) THIP(t, hipMemcpyAsync(t->tin, tg, sizeof(float) * (size_t)n * 2, hipMemcpyHostToDevice, t->stream));
  THIP(t, launch_mlp_rows(t->xin, t->tin, nullptr, n, t->a[0], t->aT[0], tg ? t->tb : nullptr, t->stream));
  return AJET_OK;
}

int ajet_trainer_forward(ajet_trainer* t, const float* x, int32_t n, float* y) {
  if (!t || !x || !y) return AJET_EINVAL;
  int rc = load_rows(t, x, nullptr, n);
  if (rc != AJET_OK) return rc;
  THIP(t, forward(t, n));
  std::vector<float> h((size_t)n * MLP_NOUT_PAD);
  THIP(t, hipMemcpyAsync(h.data(), t->y, sizeof(float) * h.size(), hipMemcpyDeviceToHost, t->stream));
  THIP(t, hipStreamSynchronize(t->stream));
  for (int b = 0; b < n; ++b)
    for (int c = 0; c < MLP_NOUT; ++c) y[2 * b + c] = h[(size_t)b * MLP_NOUT_PAD + c];
  return AJET_OK;
}

int ajet_trainer_grads(ajet_trainer* t, const float* x, const float* tg, int32_t n, ajet_mlp* grads, double* loss) {
  if (!t || !x || !tg || !grads) return AJET_EINVAL;
  int rc = load_rows(t, x, tg, n);
  if (rc != AJET_OK) return rc;
  if (t->loss_cap < 1) {
    THIP(t, hipMalloc((void**)&t->loss, sizeof(double)));
    t->loss_cap = 1;
  }
  THIP(t, hipMemsetAsync(t->loss, 0, sizeof(double), t->stream));
  THIP(t, forward(t, n));
  THIP(t, backward(t, n, t->loss));
  std::vector<float> h[NL];
  for (int l = 0; l < NL; ++l) {
    THIP(t, launch_mlp_adam(adam_args(t, l, n, 0.0f, 1), t->stream));
    h[l].resize((size_t)NOUTS[l] * MLP_K + NOUTS[l]);
    THIP(t, hipMemcpyAsync(h[l].data(), t->grad[l], sizeof(float) * h[l].size(), hipMemcpyDeviceToHost, t->stream));
  }
  double lv = 0.0;
  THIP(t, hipMemcpyAsync(&lv, t->loss, sizeof(double), hipMemcpyDeviceToHost, t->stream));
  THIP(t, hipStreamSynchronize(t->stream));
  *grads = t->meta;
  unpack(h, *grads);
  if (loss) *loss = lv;
  return AJET_OK;
}

int ajet_trainer_get(ajet_trainer* t, ajet_mlp* model) {
  if (!t || !model) return AJET_EINVAL;
  std::vector<float> h[NL];
  for (int l = 0; l < NL; ++l) {
    h[l].resize((size_t)NOUTS[l] * MLP_K + NOUTS[l]);
    THIP(t, hipMemcpyAsync(h[l].data(), t->p[l], sizeof(float) * h[l].size(), hipMemcpyDeviceToHost, t->stream));
  }
  THIP(t, hipStreamSynchronize(t->stream));
  *model = t->meta;
  unpack(h, *model);
  return AJET_OK;
}

}  // extern "C"
