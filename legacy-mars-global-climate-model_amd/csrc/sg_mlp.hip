// sg_mlp.hip -- the learned IR surrogate (include/mgcm_sg.h): an MLP whose
// matrix products run on the gfx950 matrix cores (BF16 operands, FP32
// accumulation), with the bias, GELU and GELU-derivative work fused into the
// products' epilogues, and the loss, bias gradients, Adam and feature assembly
// as small element-wise kernels.  Training and inference use the same kernels.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <new>
#include <vector>

#include "../../include/mgcm_rt.h"
#include "../../include/mgcm_sg.h"

namespace sgk {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int NIN = MGCM_SG_NIN, KIN = MGCM_SG_KIN, NH = MGCM_SG_NHID, NOUT = MGCM_SG_NOUT;
constexpr int NOUT_PAD = 64;  // the output layer's N padded to one tile for the backward product
// parameter offsets (W[k][n] then b[n] per layer)
constexpr long OW1 = 0, OB1 = OW1 + (long)KIN * NH, OW2 = OB1 + NH, OB2 = OW2 + (long)NH * NH, OW3 = OB2 + NH,
               OB3 = OW3 + (long)NH * NH, OW4 = OB3 + NH, OB4 = OW4 + (long)NH * NOUT, NPARAM = OB4 + NOUT;
static_assert(NPARAM == MGCM_SG_NPARAM, "parameter layout");

// ---------------------------------------------------------------------------
// Matrix product C(m,n) = sum_k A(m,k) B(k,n) on the matrix cores.
//
// Block tile 64 x 64, K step 32, 4 waves each owning a 32 x 32 quarter as
// 2 x 2 tiles of v_mfma_f32_16x16x32_bf16.  A(m,k) lives at A[m*sam + k*sak]
// and B(k,n) at B[k*sbk + n*sbn]; AK / BK say which of the two is the unit
// stride (k: 16-byte vector loads straight into LDS; m or n: the tile is
// transposed on its way into LDS).  LDS holds both tiles k-contiguous so that
// each lane's 8-element MFMA operand is one 16-byte LDS read: lane l holds
// row (l % 16) and k = 8 (l / 16) .. +7 of its 16 x 32 operand, and receives
// rows 4 (l / 16) .. +3 of column l % 16 of the 16 x 16 result.
// K must be a multiple of 32; M and N are bounds-checked.

enum { EPI_F32 = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_DGELU = 3 };

struct Epi {
  float *c;            // FP32 result (EPI_F32, EPI_BIAS) or pre-activation Z (EPI_BIAS_GELU)
  long ldc;
  const float *bias;   // [N] (EPI_BIAS, EPI_BIAS_GELU)
  bf16 *h;             // BF16 GELU(Z) (EPI_BIAS_GELU) or BF16 dZ (EPI_DGELU)
  long ldh;
  const float *z;      // pre-activation for the GELU derivative (EPI_DGELU)
  long ldz;
};

constexpr int BM = 64, BN = 64, BKS = 32, LDK = BKS + 8;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

template <bool AK, bool BK, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const bf16 *__restrict__ A, long sam, long sak,
                                              const bf16 *__restrict__ B, long sbk, long sbn, Epi e) {
  __shared__ __attribute__((aligned(16))) bf16 As[BM * LDK];
  __shared__ __attribute__((aligned(16))) bf16 Bs[BN * LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const bf16 zero = (bf16)0.0f;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < K; k0 += BKS) {
    if (AK) {  // 64 rows x 4 chunks of 8 k
      const int r = tid >> 2, kc = (tid & 3) * 8, m = m0 + r;
      bf16x8 v;
      if (m < M) {
        v = *(const bf16x8 *)(A + (long)m * sam + k0 + kc);
      } else {
        for (int i = 0; i < 8; i++) v[i] = zero;
      }
      *(bf16x8 *)(As + r * LDK + kc) = v;
    } else {  // 32 k x 8 chunks of 8 m
      const int kr = tid >> 3, mc = (tid & 7) * 8;
      const bf16 *src = A + (long)(k0 + kr) * sak;
      for (int i = 0; i < 8; i++) {
        const int m = m0 + mc + i;
        As[(mc + i) * LDK + kr] = (m < M) ? src[m] : zero;
      }
    }
    if (BK) {
      const int r = tid >> 2, kc = (tid & 3) * 8, n = n0 + r;
      bf16x8 v;
      if (n < N) {
        v = *(const bf16x8 *)(B + (long)n * sbn + k0 + kc);
      } else {
        for (int i = 0; i < 8; i++) v[i] = zero;
      }
      *(bf16x8 *)(Bs + r * LDK + kc) = v;
    } else {
      const int kr = tid >> 3, nc = (tid & 7) * 8;
      const bf16 *src = B + (long)(k0 + kr) * sbk;
      for (int i = 0; i < 8; i++) {
        const int n = n0 + nc + i;
        Bs[(nc + i) * LDK + kr] = (n < N) ? src[n] : zero;
      }
    }
    __syncthreads();
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; i++) a[i] = *(const bf16x8 *)(As + (wm * 32 + i * 16 + (lane & 15)) * LDK + 8 * (lane >> 4));
    for (int j = 0; j < 2; j++) b[j] = *(const bf16x8 *)(Bs + (wn * 32 + j * 16 + (lane & 15)) * LDK + 8 * (lane >> 4));
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (m >= M || n >= N) continue;
        const float v = acc[i][j][r];
        if (EPI == EPI_F32) {
          e.c[(long)m * e.ldc + n] = v;
        } else if (EPI == EPI_BIAS) {
          e.c[(long)m * e.ldc + n] = v + e.bias[n];
        } else if (EPI == EPI_BIAS_GELU) {
          const float z = v + e.bias[n];
          e.c[(long)m * e.ldc + n] = z;
          e.h[(long)m * e.ldh + n] = (bf16)gelu(z);
        } else {
          e.h[(long)m * e.ldh + n] = (bf16)(v * gelu_d(e.z[(long)m * e.ldz + n]));
        }
      }
}

template <bool AK, bool BK, int EPI>
static hipError_t gemm(int M, int N, int K, const bf16 *A, long sam, long sak, const bf16 *B, long sbk, long sbn,
                       const Epi &e, hipStream_t st) {
  const dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  hipLaunchKernelGGL((k_gemm<AK, BK, EPI>), grid, dim3(256), 0, st, M, N, K, A, sam, sak, B, sbk, sbn, e);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Element-wise kernels.

__global__ void k_to_bf16_pad(int B, const float *__restrict__ x, bf16 *__restrict__ x0) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)B * KIN) return;
  const long b = i / KIN;
  const int k = (int)(i % KIN);
  x0[i] = (bf16)(k < NIN ? x[b * NIN + k] : 0.0f);
}

__global__ void k_fill_f32(float *p, long n, float v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// MSE over the NOUT outputs: dY = 2 (y - t) / (B NOUT) in BF16, padded to
// NOUT_PAD columns with zeros; the loss is added to *loss by block partials.
__global__ void k_loss(int B, const float *__restrict__ y, const float *__restrict__ t, bf16 *__restrict__ dy,
                       float *loss) {
  __shared__ float part[256];
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float sq = 0.f;
  if (i < (long)B * NOUT_PAD) {
    const long b = i / NOUT_PAD;
    const int o = (int)(i % NOUT_PAD);
    float g = 0.f;
    if (o < NOUT) {
      const float d = y[b * NOUT + o] - t[b * NOUT + o];
      sq = d * d;
      g = 2.0f * d / ((float)B * NOUT);
    }
    dy[i] = (bf16)g;
  }
  part[threadIdx.x] = sq;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0] / ((float)B * NOUT));
}

// db[n] = sum over rows of dZ[b][n], deterministic and spread over the CUs:
// stage 1 -- block (column chunk, row slice p) sums rows p*RS.. of 64 columns
// in 4 row groups combined in a fixed order into part[p][n]; stage 2 -- one
// thread per column adds the slices in order.
constexpr int COLSUM_ROWS = 64;  // rows per slice
__global__ void k_colsum_part(int B, int N, const bf16 *__restrict__ dz, long ld, float *__restrict__ part) {
  __shared__ float acc[4][64];
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + c, p = blockIdx.y;
  const int b1 = min(B, (p + 1) * COLSUM_ROWS);
  float s = 0.f;
  if (n < N)
    for (int b = p * COLSUM_ROWS + g; b < b1; b += 4) s += (float)dz[(long)b * ld + n];
  acc[g][c] = s;
  __syncthreads();
  if (g == 0 && n < N) part[(long)p * N + n] = ((acc[0][c] + acc[1][c]) + acc[2][c]) + acc[3][c];
}
__global__ void k_colsum_fin(int N, int P, const float *__restrict__ part, float *__restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int p = 0; p < P; p++) s += part[(long)p * N + n];
  db[n] = s;
}
static void colsum(int B, int N, const bf16 *dz, long ld, float *part, float *db, hipStream_t st) {
  const int P = (B + COLSUM_ROWS - 1) / COLSUM_ROWS;
  hipLaunchKernelGGL(k_colsum_part, dim3((N + 63) / 64, P), dim3(256), 0, st, B, N, dz, ld, part);
  hipLaunchKernelGGL(k_colsum_fin, dim3((N + 255) / 256), dim3(256), 0, st, N, P, part, db);
}

__global__ void k_adam(long n, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// WT[n][k] = bf16(W[k][n]) for n < N, zero rows for N <= n < NPAD.
__global__ void k_pack(int K, int N, int NPAD, const float *__restrict__ W, bf16 *__restrict__ WT) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)NPAD * K) return;
  const int n = (int)(i / K), k = (int)(i % K);
  WT[i] = (bf16)(n < N ? W[(long)k * N + n] : 0.0f);
}

// Standardised features of one column per thread.
__global__ void k_features(int ncol, int nlay, const double *__restrict__ plev, const double *__restrict__ tlev,
                           const double *__restrict__ pmid, const double *__restrict__ tmid,
                           const double *__restrict__ qh2o, const double *__restrict__ tauref,
                           const double *__restrict__ albi, const float *__restrict__ mean,
                           const float *__restrict__ inv_std, float *__restrict__ x) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncol) return;
  const int nlev = 2 * nlay + 3;
  const double *pm = pmid + (long)c * nlev, *tm = tmid + (long)c * nlev, *q = qh2o + (long)c * nlev;
  const double *tr = tauref + (long)c * (nlev + 1);
  float *xc = x + (long)c * NIN;
  for (int l = 1; l <= MGCM_SG_NLAY; l++) {
    const int L = l + 1, km = 2 * L - 1;  // 0-based index of the mid-level K = 2L
    const float f[4] = {(float)tm[km], (float)log10(pm[km]), (float)log10(fmax(tr[2 * L] + tr[2 * L + 1], 1e-12)),
                        (float)log10(fmax(q[km], 1e-30))};
    for (int j = 0; j < 4; j++) {
      const int i = j * MGCM_SG_NLAY + (l - 1);
      xc[i] = (f[j] - mean[i]) * inv_std[i];
    }
  }
  const float s[3] = {(float)tlev[(long)c * nlev + nlev - 1], (float)(1.0 - albi[c]),
                      (float)plev[(long)c * nlev + nlev - 1]};
  for (int j = 0; j < 3; j++) {
    const int i = 4 * MGCM_SG_NLAY + j;
    xc[i] = (s[j] - mean[i]) * inv_std[i];
  }
}

static unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace sgk

using namespace sgk;

struct mgcm_sg {
  int maxb;
  int step;
  float *p, *g, *m, *v;        // FP32 master parameters, gradients, Adam moments
  bf16 *wt[4];                 // BF16 W^T per layer: [NH][KIN], [NH][NH], [NH][NH], [NOUT_PAD][NH]
  bf16 *x0, *h[3], *dy, *dz[2];
  float *part;  // bias-gradient partial sums, [ceil(B/COLSUM_ROWS)][NH]
  float *z[3], *y;
};

namespace {

const int KL[4] = {KIN, NH, NH, NH}, NL[4] = {NH, NH, NH, NOUT}, NP[4] = {NH, NH, NH, NOUT_PAD};
const long OW[4] = {OW1, OW2, OW3, OW4}, OB[4] = {OB1, OB2, OB3, OB4};

hipError_t pack_all(mgcm_sg *s, hipStream_t st) {
  for (int l = 0; l < 4; l++) {
    hipLaunchKernelGGL(k_pack, dim3(nblk((long)NP[l] * KL[l], 256)), dim3(256), 0, st, KL[l], NL[l], NP[l],
                       s->p + OW[l], s->wt[l]);
  }
  return hipGetLastError();
}

hipError_t forward(mgcm_sg *s, int B, const float *x, float *y, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf16_pad, dim3(nblk((long)B * KIN, 256)), dim3(256), 0, st, B, x, s->x0);
  const bf16 *in = s->x0;
  long ldin = KIN;
  for (int l = 0; l < 3; l++) {
    Epi e{s->z[l], NH, s->p + OB[l], s->h[l], NH, nullptr, 0};
    hipError_t r = gemm<true, true, EPI_BIAS_GELU>(B, NH, KL[l], in, ldin, 1, s->wt[l], 1, KL[l], e, st);
    if (r != hipSuccess) return r;
    in = s->h[l];
    ldin = NH;
  }
  Epi e{y, NOUT, s->p + OB4, nullptr, 0, nullptr, 0};
  return gemm<true, true, EPI_BIAS>(B, NOUT, NH, s->h[2], NH, 1, s->wt[3], 1, NH, e, st);
}

// Gradients of the MSE into s->g (after forward into s->y).
hipError_t backward(mgcm_sg *s, int B, const float *t, float *loss, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, st, loss, 1L, 0.0f);
  hipLaunchKernelGGL(k_loss, dim3(nblk((long)B * NOUT_PAD, 256)), dim3(256), 0, st, B, s->y, t, s->dy, loss);
  // output layer: dW4 = H3^T dY, db4, dZ3 = (dY W4^T) * GELU'(Z3)
  {
    Epi e{s->g + OW4, NOUT, nullptr, nullptr, 0, nullptr, 0};
    hipError_t r = gemm<false, false, EPI_F32>(NH, NOUT, B, s->h[2], 1, NH, s->dy, NOUT_PAD, 1, e, st);
    if (r != hipSuccess) return r;
    colsum(B, NOUT, s->dy, (long)NOUT_PAD, s->part, s->g + OB4, st);
    Epi d{nullptr, 0, nullptr, s->dz[0], NH, s->z[2], NH};
    r = gemm<true, false, EPI_DGELU>(B, NH, NOUT_PAD, s->dy, NOUT_PAD, 1, s->wt[3], NH, 1, d, st);
    if (r != hipSuccess) return r;
  }
  // hidden layers 3, 2, 1: dz[cur] holds dZ of the layer
  int cur = 0;
  for (int l = 2; l >= 0; l--) {
    const bf16 *hin = (l == 0) ? s->x0 : s->h[l - 1];
    const int kin = KL[l];
    Epi e{s->g + OW[l], NH, nullptr, nullptr, 0, nullptr, 0};
    hipError_t r = gemm<false, false, EPI_F32>(kin, NH, B, hin, 1, kin, s->dz[cur], NH, 1, e, st);
    if (r != hipSuccess) return r;
    colsum(B, NH, s->dz[cur], (long)NH, s->part, s->g + OB[l], st);
    if (l > 0) {
      Epi d{nullptr, 0, nullptr, s->dz[1 - cur], NH, s->z[l - 1], NH};
      r = gemm<true, false, EPI_DGELU>(B, NH, NH, s->dz[cur], NH, 1, s->wt[l], KL[l], 1, d, st);
      if (r != hipSuccess) return r;
      cur = 1 - cur;
    }
  }
  return hipGetLastError();
}

// splitmix64 + Box-Muller
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() { return sqrt(-2.0 * log(uniform())) * cos(6.283185307179586 * uniform()); }
};

}  // namespace

#define SG_TRY(x)                                    \
  do {                                               \
    if ((x) != hipSuccess) return MGCM_RT_ERR_DEVICE; \
  } while (0)

extern "C" int mgcm_sg_destroy(mgcm_sg *s) {
  if (!s) return MGCM_RT_ERR_ARG;
  void *bufs[] = {s->p, s->g, s->m, s->v, s->wt[0], s->wt[1], s->wt[2], s->wt[3], s->x0, s->h[0], s->h[1],
                  s->h[2], s->dy, s->dz[0], s->dz[1], s->z[0], s->z[1], s->z[2], s->y, s->part};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  delete s;
  return MGCM_RT_OK;
}

extern "C" int mgcm_sg_set_params(mgcm_sg *s, const float *params) {
  if (!s || !params) return MGCM_RT_ERR_ARG;
  std::vector<float> h(params, params + NPARAM);
  for (int k = NIN; k < KIN; k++)  // padding rows of W1 stay zero
    for (int n = 0; n < NH; n++) h[OW1 + (long)k * NH + n] = 0.f;
  SG_TRY(hipMemcpy(s->p, h.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
  SG_TRY(hipMemset(s->m, 0, NPARAM * sizeof(float)));
  SG_TRY(hipMemset(s->v, 0, NPARAM * sizeof(float)));
  s->step = 0;
  SG_TRY(pack_all(s, nullptr));
  SG_TRY(hipDeviceSynchronize());
  return MGCM_RT_OK;
}

extern "C" int mgcm_sg_get_params(mgcm_sg *s, float *params) {
  if (!s || !params) return MGCM_RT_ERR_ARG;
  SG_TRY(hipDeviceSynchronize());
  SG_TRY(hipMemcpy(params, s->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return MGCM_RT_OK;
}

extern "C" int mgcm_sg_create(int max_batch, unsigned long long seed, mgcm_sg **out) {
  if (!out || max_batch < 1) return MGCM_RT_ERR_ARG;
  *out = nullptr;
  mgcm_sg *s = new (std::nothrow) mgcm_sg();
  if (!s) return MGCM_RT_ERR_NOMEM;
  memset(s, 0, sizeof(*s));
  s->maxb = max_batch;
  const long B = max_batch;
  bool ok = hipMalloc(&s->p, NPARAM * 4) == hipSuccess && hipMalloc(&s->g, NPARAM * 4) == hipSuccess &&
            hipMalloc(&s->m, NPARAM * 4) == hipSuccess && hipMalloc(&s->v, NPARAM * 4) == hipSuccess;
  for (int l = 0; l < 4 && ok; l++) ok = hipMalloc(&s->wt[l], (long)NP[l] * KL[l] * 2) == hipSuccess;
  ok = ok && hipMalloc(&s->x0, B * KIN * 2) == hipSuccess && hipMalloc(&s->dy, B * NOUT_PAD * 2) == hipSuccess &&
       hipMalloc(&s->y, B * NOUT * 4) == hipSuccess;
  for (int l = 0; l < 3 && ok; l++)
    ok = hipMalloc(&s->h[l], B * NH * 2) == hipSuccess && hipMalloc(&s->z[l], B * NH * 4) == hipSuccess;
  for (int l = 0; l < 2 && ok; l++) ok = hipMalloc(&s->dz[l], B * NH * 2) == hipSuccess;
  ok = ok && hipMalloc(&s->part, (long)((B + COLSUM_ROWS - 1) / COLSUM_ROWS) * NH * 4) == hipSuccess;
  if (!ok) {
    mgcm_sg_destroy(s);
    return MGCM_RT_ERR_NOMEM;
  }
  // He-normal hidden weights (fan-in), 1/sqrt(fan-in) output weights, zero biases
  std::vector<float> h(NPARAM, 0.f);
  Rng rng{seed};
  const int fan[4] = {NIN, NH, NH, NH};
  for (int l = 0; l < 4; l++) {
    const double sd = (l < 3 ? sqrt(2.0) : 1.0) / sqrt((double)fan[l]);
    for (int k = 0; k < fan[l]; k++)
      for (int n = 0; n < NL[l]; n++) h[OW[l] + (long)k * NL[l] + n] = (float)(sd * rng.normal());
  }
  const int st = mgcm_sg_set_params(s, h.data());
  if (st != MGCM_RT_OK) {
    mgcm_sg_destroy(s);
    return st;
  }
  *out = s;
  return MGCM_RT_OK;
}

extern "C" int mgcm_sg_forward(mgcm_sg *s, int batch, const float *x, float *y, void *stream) {
  if (!s || !x || !y || batch < 0 || batch > s->maxb) return MGCM_RT_ERR_ARG;
  if (batch == 0) return MGCM_RT_OK;
  return forward(s, batch, x, y, (hipStream_t)stream) == hipSuccess ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
}

extern "C" int mgcm_sg_train_step(mgcm_sg *s, int batch, const float *x, const float *t, float lr, float *loss,
                                  void *stream) {
  if (!s || !x || !t || !loss || batch < 32 || batch > s->maxb || batch % 32) return MGCM_RT_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (forward(s, batch, x, s->y, st) != hipSuccess) return MGCM_RT_ERR_DEVICE;
  if (backward(s, batch, t, loss, st) != hipSuccess) return MGCM_RT_ERR_DEVICE;
  s->step++;
  const float c1 = 1.0f - powf(0.9f, (float)s->step), c2 = 1.0f - powf(0.999f, (float)s->step);
  hipLaunchKernelGGL(k_adam, dim3(nblk(NPARAM, 256)), dim3(256), 0, st, (long)NPARAM, s->p, s->g, s->m, s->v, lr,
                     c1, c2);
  return pack_all(s, st) == hipSuccess ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
}

extern "C" int mgcm_sg_gradients(mgcm_sg *s, int batch, const float *x, const float *t, float *grads) {
  if (!s || !x || !t || !grads || batch < 32 || batch > s->maxb || batch % 32) return MGCM_RT_ERR_ARG;
  float *loss = nullptr;
  SG_TRY(hipMalloc(&loss, sizeof(float)));
  const bool ok = forward(s, batch, x, s->y, nullptr) == hipSuccess &&
                  backward(s, batch, t, loss, nullptr) == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
                  hipMemcpy(grads, s->g, NPARAM * 4, hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(loss);
  return ok ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
}

extern "C" int mgcm_sg_features(int ncol, int nlay, const double *plev, const double *tlev, const double *pmid,
                                const double *tmid, const double *qh2o, const double *tauref, const double *albi,
                                const float *mean, const float *inv_std, float *x, void *stream) {
  if (ncol < 0 || nlay != MGCM_SG_NLAY) return ncol < 0 ? MGCM_RT_ERR_ARG : MGCM_RT_ERR_NLAY;
  if (ncol == 0) return MGCM_RT_OK;
  if (!plev || !tlev || !pmid || !tmid || !qh2o || !tauref || !albi || !mean || !inv_std || !x)
    return MGCM_RT_ERR_ARG;
  hipLaunchKernelGGL(k_features, dim3(nblk(ncol, 128)), dim3(128), 0, (hipStream_t)stream, ncol, nlay, plev, tlev,
                     pmid, tmid, qh2o, tauref, albi, mean, inv_std, x);
  return hipGetLastError() == hipSuccess ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
}

// The surrogate's check, one thread per column: both networks' outputs
// de-standardised, h_k = scale * sinh(y_k * tstd + tmean) (the inverse of the
// asinh target transform, FP64), their largest disagreement over the outputs,
// and whether every feature lies in the trusted range [xlo, xhi].
__global__ void k_check(int ncol, const float *__restrict__ y0, const float *__restrict__ y1,
                        const float *__restrict__ x, const double *__restrict__ tmean,
                        const double *__restrict__ tstd, const float *__restrict__ xlo,
                        const float *__restrict__ xhi, double scale, double *__restrict__ h,
                        double *__restrict__ dis, int *__restrict__ inr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncol) return;
  double d = 0.0;
  for (int j = 0; j < NOUT; j++) {
    const double a = scale * sinh((double)y0[(long)i * NOUT + j] * tstd[j] + tmean[j]);
    const double b = y1 ? scale * sinh((double)y1[(long)i * NOUT + j] * tstd[j] + tmean[j]) : a;
    h[(long)i * NOUT + j] = a;
    d = fmax(d, fabs(a - b));
  }
  dis[i] = d;
  int ok = 1;
  for (int k = 0; k < NIN; k++) {
    const float v = x[(long)i * NIN + k];
    ok &= (v >= xlo[k]) & (v <= xhi[k]);
  }
  inr[i] = ok;
}

extern "C" int mgcm_sg_check(int ncol, const float *y0, const float *y1, const float *x, const double *tmean,
                             const double *tstd, const float *xlo, const float *xhi, double scale, double *h,
                             double *dis, int *inr, void *stream) {
  if (ncol < 0) return MGCM_RT_ERR_ARG;
  if (ncol == 0) return MGCM_RT_OK;
  if (!y0 || !x || !tmean || !tstd || !xlo || !xhi || !h || !dis || !inr) return MGCM_RT_ERR_ARG;
  hipLaunchKernelGGL(k_check, dim3(nblk(ncol, 128)), dim3(128), 0, (hipStream_t)stream, ncol, y0, y1, x, tmean,
                     tstd, xlo, xhi, scale, h, dis, inr);
  return hipGetLastError() == hipSuccess ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
}

extern "C" int mgcm_sg_gemm_check(int m, int n, int k, const void *a, const void *b, float *c) {
  if (!a || !b || !c || m < 1 || n < 1 || k < 32 || k % 32) return MGCM_RT_ERR_ARG;
  // A row-major [m][k] (k contiguous), B row-major [k][n] (n contiguous)
  Epi e{c, n, nullptr, nullptr, 0, nullptr, 0};
  if (gemm<true, false, EPI_F32>(m, n, k, (const bf16 *)a, k, 1, (const bf16 *)b, n, 1, e, nullptr) != hipSuccess)
    return MGCM_RT_ERR_DEVICE;
  return hipDeviceSynchronize() == hipSuccess ? MGCM_RT_OK : MGCM_RT_ERR_DEVICE;
}
