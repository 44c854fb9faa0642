// Learned warm start for the per-leaf ionisation / temperature solve.
//
// A multilayer perceptron 8 -> 64 -> 64 -> 64 -> 4 with GELU activations maps
// the tallied rates of a leaf to a first guess of its balance state:
//   inputs  (log10 of) n_H, Gamma_HI, Gamma_HeI, Gamma_HeII, H_HI, H_HeI,
//           H_HeII and the current n_e, each standardised by (mu, sd);
//   outputs (standardised) log10 x_HI, log10 x_HeI, log10 x_HeII, log10 T_e.
// The safeguarded Newton solve of the thermal update starts from it and judges
// every leaf (mc_kernels.hip); nothing here replaces the solve.  The reference
// has no counterpart: this is the build's learned component.
//
// Training runs in FP32 with hand-written forward / backward kernels and
// Adam; inference runs the same network with BF16 weights and activations on
// the matrix cores (v_mfma_f32_32x32x8_bf16, FP32 accumulate), one wavefront
// per tile of 32 leaves, weights staged in LDS.
//
// Parameter blob (FP32, MC_MLP_NPARAM values, the layout of mc_mlp_set/get):
//   W1[8][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64][4] b4[4]
//   in_mu[8] in_sd[8] out_mu[4] out_sd[4]      (W[k][n]: input k -> unit n)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "mc_internal.h"

namespace {

constexpr int NI = MC_MLP_IN, NH = MC_MLP_HID, NO = MC_MLP_OUT;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NI * NH;
constexpr int OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NH;
constexpr int OFF_W4 = OFF_B3 + NH, OFF_B4 = OFF_W4 + NH * NO;
constexpr int OFF_INMU = OFF_B4 + NO, OFF_INSD = OFF_INMU + NI;
constexpr int OFF_OUTMU = OFF_INSD + NI, OFF_OUTSD = OFF_OUTMU + NO;
constexpr int NTRAIN = OFF_INMU;                 // trainable parameters
static_assert(OFF_OUTSD + NO == MC_MLP_NPARAM, "parameter layout");

constexpr int kB = 256;

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float z)
{
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// ---- training path (FP32) ---------------------------------------------------
// standardise raw features / targets of the selected rows
__global__ void k_prep(const float *X, const float *Y, const int32_t *idx, int B, const float *P,
                       float *Xn, float *Yn)
{
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= B) return;
    const int r = idx ? idx[i] : i;
    for (int k = 0; k < NI; ++k)
        Xn[i * NI + k] = (X[(size_t)r * NI + k] - P[OFF_INMU + k]) / P[OFF_INSD + k];
    if (Y)
        for (int k = 0; k < NO; ++k)
            Yn[i * NO + k] = (Y[(size_t)r * NO + k] - P[OFF_OUTMU + k]) / P[OFF_OUTSD + k];
}

// Z = A W + b, H = gelu(Z) (act) or Z
__global__ void k_fwd(const float *A, const float *W, const float *b, float *Z, float *H, int B,
                      int K, int N, int act)
{
    const int t = blockIdx.x * kB + threadIdx.x;
    if (t >= B * N) return;
    const int r = t / N, n = t - r * N;
    float z = b[n];
    for (int k = 0; k < K; ++k) z += A[r * K + k] * W[k * N + n];
    Z[t] = z;
    H[t] = act ? gelu(z) : z;
}

// mean squared error over B x NO standardised outputs; dZ = dL/dZ
__global__ void k_loss(const float *Z, const float *Y, float *dZ, int B, float *part)
{
    __shared__ float sh[kB];
    const int t = blockIdx.x * kB + threadIdx.x;
    float e2 = 0.0f;
    if (t < B * NO) {
        const float e = Z[t] - Y[t];
        dZ[t] = 2.0f * e / (float)(B * NO);
        e2 = e * e;
    }
    sh[threadIdx.x] = e2;
    __syncthreads();
    for (int s = kB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// dW[k][n] = sum_r A[r][k] dZ[r][n]; db[n] = sum_r dZ[r][n] (row k == K):
// partial sums over kWgradParts row ranges (blockIdx.y), then a fixed-order
// sum of the partials
constexpr int kWgradParts = 32;

__global__ void k_wgrad_part(const float *A, const float *dZ, float *part, int B, int K, int N)
{
    const int t = blockIdx.x * kB + threadIdx.x;
    const int M = (K + 1) * N;
    if (t >= M) return;
    const int k = t / N, n = t - k * N;
    const int r0 = (int)((long long)B * blockIdx.y / kWgradParts);
    const int r1 = (int)((long long)B * (blockIdx.y + 1) / kWgradParts);
    float s = 0.0f;
    if (k < K)
        for (int r = r0; r < r1; ++r) s += A[r * K + k] * dZ[r * N + n];
    else
        for (int r = r0; r < r1; ++r) s += dZ[r * N + n];
    part[(size_t)blockIdx.y * M + t] = s;
}

__global__ void k_wgrad_sum(const float *part, float *dW, float *db, int K, int N)
{
    const int t = blockIdx.x * kB + threadIdx.x;
    const int M = (K + 1) * N;
    if (t >= M) return;
    float s = 0.0f;
    for (int p = 0; p < kWgradParts; ++p) s += part[(size_t)p * M + t];
    const int k = t / N, n = t - k * N;
    if (k < K) dW[k * N + n] = s; else db[n] = s;
}

// dZprev[r][k] = (sum_n dZ[r][n] W[k][n]) gelu'(Zprev[r][k])
__global__ void k_dgrad(const float *dZ, const float *W, const float *Zp, float *dZp, int B, int K,
                        int N)
{
    const int t = blockIdx.x * kB + threadIdx.x;
    if (t >= B * K) return;
    const int r = t / K, k = t - r * K;
    float s = 0.0f;
    for (int n = 0; n < N; ++n) s += dZ[r * N + n] * W[k * N + n];
    dZp[t] = s * gelu_grad(Zp[t]);
}

__global__ void k_adam(float *P, const float *G, float *m, float *v, int n, float lr, float b1,
                       float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g, vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---- inference path (BF16 matrix cores) ---------------------------------------
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);                   // round to nearest even
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned int)h << 16); }

// LDS image of the BF16 network: weights stored k-major per output column
// block so that one lane's 4 consecutive k values are one 8-byte load.
struct InferLds {
    unsigned short w1[NH][8];          // [n][k]
    unsigned short w2[NH][NH];         // [n][k]
    unsigned short w3[NH][NH];
    unsigned short w4[32][NH];         // [n][k], n >= NO zero
    float b1[NH], b2[NH], b3[NH], b4[32];
    float inmu[NI], insd[NI], outmu[NO], outsd[NO];
    unsigned short act[4][32][NH + 4]; // per-wave activation tile [leaf][unit] (+pad)
};

// C += A(32 x 8k) * B(8k x 32): lane (i = l%32, kb = l/32) supplies
// A[i][4kb..4kb+3] and B[4kb..4kb+3][i]; D holds column l%32, rows
// 8*(r/4) + 4*(l/32) + r%4 for r = 0..15.
__device__ __forceinline__ v16f mfma_k8(v4s a, v4s b, v16f c)
{
    return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4s load4(const unsigned short *p)
{
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    v4s r;
    r[0] = (short)(u.x & 0xFFFF); r[1] = (short)(u.x >> 16);
    r[2] = (short)(u.y & 0xFFFF); r[3] = (short)(u.y >> 16);
    return r;
}

// hidden layer on the wave's activation tile: act <- gelu(act W + b), K = NH
__device__ __forceinline__ void hidden_layer(InferLds &L, int wv, int lane,
                                             const unsigned short (*w)[NH], const float *b)
{
    const int i = lane & 31, kb = lane >> 5;
    v16f acc[2];
    for (int t = 0; t < 2; ++t) acc[t] = (v16f){0};
    for (int s = 0; s < NH / 8; ++s) {
        const v4s a = load4(&L.act[wv][i][8 * s + 4 * kb]);
        for (int t = 0; t < 2; ++t) acc[t] = mfma_k8(a, load4(&w[32 * t + i][8 * s + 4 * kb]), acc[t]);
    }
    __syncthreads();
    for (int t = 0; t < 2; ++t)
        for (int r = 0; r < 16; ++r) {
            const int row = 8 * (r >> 2) + 4 * kb + (r & 3), col = 32 * t + i;
            L.act[wv][row][col] = f2bf(gelu(acc[t][r] + b[col]));
        }
    __syncthreads();
}

}  // namespace

// features of n leaves (host or device layout n x NI) -> standardised log10
// predictions, de-standardised: out[n][4] = log10 xHI, xHeI, xHeII, T
__global__ __launch_bounds__(256) void k_mlp_infer(const float *P, const float *X, int n, float *out)
{
    __shared__ InferLds L;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    for (int e = tid; e < NH * 8; e += 256) {
        const int nn = e / 8, k = e % 8;
        L.w1[nn][k] = f2bf(P[OFF_W1 + k * NH + nn]);
    }
    for (int e = tid; e < NH * NH; e += 256) {
        const int nn = e / NH, k = e % NH;
        L.w2[nn][k] = f2bf(P[OFF_W2 + k * NH + nn]);
        L.w3[nn][k] = f2bf(P[OFF_W3 + k * NH + nn]);
    }
    for (int e = tid; e < 32 * NH; e += 256) {
        const int nn = e / NH, k = e % NH;
        L.w4[nn][k] = nn < NO ? f2bf(P[OFF_W4 + k * NO + nn]) : (unsigned short)0;
    }
    for (int e = tid; e < NH; e += 256) {
        L.b1[e] = P[OFF_B1 + e]; L.b2[e] = P[OFF_B2 + e]; L.b3[e] = P[OFF_B3 + e];
    }
    if (tid < 32) L.b4[tid] = tid < NO ? P[OFF_B4 + tid] : 0.0f;
    if (tid < NI) { L.inmu[tid] = P[OFF_INMU + tid]; L.insd[tid] = P[OFF_INSD + tid]; }
    if (tid < NO) { L.outmu[tid] = P[OFF_OUTMU + tid]; L.outsd[tid] = P[OFF_OUTSD + tid]; }
    __syncthreads();
    const int i = lane & 31, kb = lane >> 5;
    const int ntiles = (n + 31) / 32;
    // every wave of the block runs the same number of tile iterations so the
    // block-wide barriers inside the layers stay matched
    const int per_block = (ntiles + gridDim.x - 1) / gridDim.x;
    const int rounds = (per_block + 3) / 4;
    for (int rd = 0; rd < rounds; ++rd) {
        const int tile = (blockIdx.x * rounds + rd) * 4 + wv;
        const int leaf = tile * 32 + i;
        const bool live = tile < ntiles && leaf < n;
        // layer 1 input: lane supplies features 4kb..4kb+3 of its leaf
        v4s a;
        for (int q = 0; q < 4; ++q) {
            const int k = 4 * kb + q;
            const float x = live ? (X[(size_t)leaf * NI + k] - L.inmu[k]) / L.insd[k] : 0.0f;
            a[q] = (short)f2bf(x);
        }
        v16f acc[2];
        for (int t = 0; t < 2; ++t) acc[t] = mfma_k8(a, load4(&L.w1[32 * t + i][4 * kb]), (v16f){0});
        for (int t = 0; t < 2; ++t)
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * kb + (r & 3), col = 32 * t + i;
                L.act[wv][row][col] = f2bf(gelu(acc[t][r] + L.b1[col]));
            }
        __syncthreads();
        hidden_layer(L, wv, lane, L.w2, L.b2);
        hidden_layer(L, wv, lane, L.w3, L.b3);
        v16f o = (v16f){0};
        for (int s = 0; s < NH / 8; ++s)
            o = mfma_k8(load4(&L.act[wv][i][8 * s + 4 * kb]), load4(&L.w4[i][8 * s + 4 * kb]), o);
        if (i < NO)
            for (int r = 0; r < 16; ++r) {
                const int row = 8 * (r >> 2) + 4 * kb + (r & 3);
                const int lf = tile * 32 + row;
                if (tile < ntiles && lf < n)
                    out[(size_t)lf * NO + i] = (o[r] + L.b4[i]) * L.outsd[i] + L.outmu[i];
            }
        __syncthreads();
    }
}

// per-leaf features from the device state and rates
__global__ __launch_bounds__(256) void k_mlp_features(DevState s, int n, float *X)
{
    const int il = blockIdx.x * 256 + threadIdx.x;
    if (il >= n) return;
    const size_t N = n;
    float *x = X + (size_t)il * NI;
    x[0] = log10f((float)fmax(s.nH[il], 1e-30));
    for (int k = 0; k < 3; ++k) x[1 + k] = log10f((float)fmax(s.rates[k * N + il], 1e-30));
    for (int k = 0; k < 3; ++k) x[4 + k] = log10f((float)fmax(s.rates[(3 + k) * N + il], 1e-40));
    x[7] = log10f((float)fmax(s.ne[il], 1e-30));
}

// ---- host side ------------------------------------------------------------------
namespace {

struct Grid1 {
    unsigned g;
    explicit Grid1(long long n) : g((unsigned)((n + kB - 1) / kB)) {}
};

int ensure_mlp(mc_ctx *c)
{
    if (c->mlp.P) return 0;
    MC_HIP(hipMalloc(&c->mlp.P, sizeof(float) * MC_MLP_NPARAM));
    MC_HIP(hipMemset(c->mlp.P, 0, sizeof(float) * MC_MLP_NPARAM));
    return 0;
}

// device scratch of one training / gradient call (freed by the caller)
struct TrainBufs {
    float *Xn = nullptr, *Yn = nullptr, *Z[4] = {}, *H[4] = {}, *dZ[4] = {}, *G = nullptr,
          *part = nullptr, *wpart = nullptr;
    int B = 0;
    int alloc(int b)
    {
        B = b;
        const int w[4] = {NH, NH, NH, NO};
        MC_HIP(hipMalloc(&Xn, sizeof(float) * b * NI));
        MC_HIP(hipMalloc(&Yn, sizeof(float) * b * NO));
        for (int l = 0; l < 4; ++l) {
            MC_HIP(hipMalloc(&Z[l], sizeof(float) * b * w[l]));
            MC_HIP(hipMalloc(&H[l], sizeof(float) * b * w[l]));
            MC_HIP(hipMalloc(&dZ[l], sizeof(float) * b * w[l]));
        }
        MC_HIP(hipMalloc(&G, sizeof(float) * MC_MLP_NPARAM));
        MC_HIP(hipMalloc(&part, sizeof(float) * ((b * NO + kB - 1) / kB)));
        MC_HIP(hipMalloc(&wpart, sizeof(float) * kWgradParts * (NH + 1) * NH));
        return 0;
    }
    ~TrainBufs()
    {
        for (void *p : {(void *)Xn, (void *)Yn, (void *)G, (void *)part, (void *)wpart})
            if (p) (void)hipFree(p);
        for (int l = 0; l < 4; ++l)
            for (void *p : {(void *)Z[l], (void *)H[l], (void *)dZ[l]}) if (p) (void)hipFree(p);
    }
};

// forward + backward of one standardised batch (Xn, Yn filled); gradient into
// T.G (trainable part), returns the loss
int fwd_bwd(mc_ctx *c, TrainBufs &T, hipStream_t st, float *loss)
{
    const float *P = c->mlp.P;
    const int B = T.B;
    const int K[4] = {NI, NH, NH, NH}, N[4] = {NH, NH, NH, NO};
    const int offW[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4}, offB[4] = {OFF_B1, OFF_B2, OFF_B3, OFF_B4};
    const float *A = T.Xn;
    for (int l = 0; l < 4; ++l) {
        k_fwd<<<Grid1((long long)B * N[l]).g, kB, 0, st>>>(A, P + offW[l], P + offB[l], T.Z[l],
                                                           T.H[l], B, K[l], N[l], l < 3);
        A = T.H[l];
    }
    const int nblk = (B * NO + kB - 1) / kB;
    k_loss<<<nblk, kB, 0, st>>>(T.Z[3], T.Yn, T.dZ[3], B, T.part);
    MC_HIP(hipMemsetAsync(T.G, 0, sizeof(float) * MC_MLP_NPARAM, st));
    for (int l = 3; l >= 0; --l) {
        const float *Ain = l == 0 ? T.Xn : T.H[l - 1];
        const unsigned gw = Grid1((long long)(K[l] + 1) * N[l]).g;
        k_wgrad_part<<<dim3(gw, kWgradParts), kB, 0, st>>>(Ain, T.dZ[l], T.wpart, B, K[l], N[l]);
        k_wgrad_sum<<<gw, kB, 0, st>>>(T.wpart, T.G + offW[l], T.G + offB[l], K[l], N[l]);
        if (l > 0)
            k_dgrad<<<Grid1((long long)B * K[l]).g, kB, 0, st>>>(T.dZ[l], P + offW[l], T.Z[l - 1],
                                                                 T.dZ[l - 1], B, K[l], N[l]);
    }
    MC_HIP(hipGetLastError());
    std::vector<float> h(nblk);
    MC_HIP(hipMemcpyAsync(h.data(), T.part, sizeof(float) * nblk, hipMemcpyDeviceToHost, st));
    MC_HIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (float v : h) s += v;
    *loss = (float)(s / (double)(B * NO));
    return 0;
}

}  // namespace

#define CHECK_CTX_MLP(c)                                   \
    do {                                                   \
        if (!(c)) {                                        \
            mc_set_error("null context");                  \
            return -1;                                     \
        }                                                  \
    } while (0)

extern "C" int mc_mlp_set(mc_ctx *c, const float *params, int nparam)
{
    CHECK_CTX_MLP(c);
    if (!params || nparam != MC_MLP_NPARAM) {
        mc_set_error("mc_mlp_set: need MC_MLP_NPARAM parameters");
        return -1;
    }
    for (int k = 0; k < NI; ++k)
        if (!(params[OFF_INSD + k] > 0.0f)) { mc_set_error("mc_mlp_set: in_sd must be > 0"); return -1; }
    for (int k = 0; k < NO; ++k)
        if (!(params[OFF_OUTSD + k] > 0.0f)) { mc_set_error("mc_mlp_set: out_sd must be > 0"); return -1; }
    if (ensure_mlp(c)) return -100;
    MC_HIP(hipMemcpy(c->mlp.P, params, sizeof(float) * MC_MLP_NPARAM, hipMemcpyHostToDevice));
    c->mlp.ready = true;
    return 0;
}

extern "C" int mc_mlp_get(mc_ctx *c, float *params)
{
    CHECK_CTX_MLP(c);
    if (!c->mlp.ready || !params) { mc_set_error("mc_mlp_get: no network set"); return -1; }
    MC_HIP(hipMemcpy(params, c->mlp.P, sizeof(float) * MC_MLP_NPARAM, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int mc_mlp_grad(mc_ctx *c, const float *X, const float *Y, int n, float *loss,
                           float *grad)
{
    CHECK_CTX_MLP(c);
    if (!c->mlp.ready || !X || !Y || n < 1 || n > (1 << 20) || !grad) {
        mc_set_error("mc_mlp_grad: need a network, 1 <= n <= 2^20 rows, X, Y, grad");
        return -1;
    }
    hipStream_t st = c->stream;
    TrainBufs T;
    if (T.alloc(n)) return -100;
    float *dX = nullptr, *dY = nullptr;
    MC_HIP(hipMalloc(&dX, sizeof(float) * (size_t)n * NI));
    MC_HIP(hipMalloc(&dY, sizeof(float) * (size_t)n * NO));
    int rc = 0;
    if (hipMemcpy(dX, X, sizeof(float) * (size_t)n * NI, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dY, Y, sizeof(float) * (size_t)n * NO, hipMemcpyHostToDevice) != hipSuccess) {
        mc_set_error("mc_mlp_grad: copy failed");
        rc = -100;
    }
    if (!rc) {
        k_prep<<<Grid1(n).g, kB, 0, st>>>(dX, dY, nullptr, n, c->mlp.P, T.Xn, T.Yn);
        float l = 0.0f;
        rc = fwd_bwd(c, T, st, &l);
        if (!rc) {
            if (loss) *loss = l;
            std::vector<float> g(MC_MLP_NPARAM, 0.0f);
            if (hipMemcpy(g.data(), T.G, sizeof(float) * NTRAIN, hipMemcpyDeviceToHost) != hipSuccess) {
                mc_set_error("mc_mlp_grad: copy failed");
                rc = -100;
            } else {
                std::memcpy(grad, g.data(), sizeof(float) * MC_MLP_NPARAM);
            }
        }
    }
    (void)hipFree(dX);
    (void)hipFree(dY);
    return rc;
}

extern "C" int mc_mlp_train(mc_ctx *c, const float *X, const float *Y, int64_t n, int steps,
                            int batch, float lr, uint32_t seed, float *loss_hist)
{
    CHECK_CTX_MLP(c);
    if (!c->mlp.ready || !X || !Y || n < 1 || steps < 0 || batch < 1 || batch > (1 << 16) ||
        !(lr > 0.0f)) {
        mc_set_error("mc_mlp_train: need a network (mc_mlp_set), data, steps >= 0, "
                     "1 <= batch <= 65536, lr > 0");
        return -1;
    }
    hipStream_t st = c->stream;
    const int B = (int)std::min<int64_t>(batch, n);
    TrainBufs T;
    if (T.alloc(B)) return -100;
    float *dX = nullptr, *dY = nullptr, *m = nullptr, *v = nullptr;
    int32_t *didx = nullptr;
    int rc = 0;
    if (hipMalloc(&dX, sizeof(float) * (size_t)n * NI) != hipSuccess ||
        hipMalloc(&dY, sizeof(float) * (size_t)n * NO) != hipSuccess ||
        hipMalloc(&m, sizeof(float) * NTRAIN) != hipSuccess ||
        hipMalloc(&v, sizeof(float) * NTRAIN) != hipSuccess ||
        hipMalloc(&didx, sizeof(int32_t) * B) != hipSuccess) {
        mc_set_error("mc_mlp_train: device allocation failed");
        rc = -100;
    }
    if (!rc && (hipMemcpy(dX, X, sizeof(float) * (size_t)n * NI, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dY, Y, sizeof(float) * (size_t)n * NO, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(m, 0, sizeof(float) * NTRAIN) != hipSuccess ||
                hipMemset(v, 0, sizeof(float) * NTRAIN) != hipSuccess)) {
        mc_set_error("mc_mlp_train: upload failed");
        rc = -100;
    }
    // minibatches without replacement from a fresh permutation per epoch
    std::mt19937_64 rng(seed);
    std::vector<int32_t> perm((size_t)n);
    for (int64_t i = 0; i < n; ++i) perm[i] = (int32_t)i;
    int64_t pos = n;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < steps && !rc; ++it) {
        if (pos + B > n) {
            std::shuffle(perm.begin(), perm.end(), rng);
            pos = 0;
        }
        if (hipMemcpyAsync(didx, perm.data() + pos, sizeof(int32_t) * B, hipMemcpyHostToDevice, st) !=
            hipSuccess) {
            mc_set_error("mc_mlp_train: index upload failed");
            rc = -100;
            break;
        }
        pos += B;
        k_prep<<<Grid1(B).g, kB, 0, st>>>(dX, dY, didx, B, c->mlp.P, T.Xn, T.Yn);
        float l = 0.0f;
        rc = fwd_bwd(c, T, st, &l);
        if (rc) break;
        if (loss_hist) loss_hist[it] = l;
        const float c1 = 1.0f - std::pow(b1, (float)(it + 1)), c2 = 1.0f - std::pow(b2, (float)(it + 1));
        k_adam<<<Grid1(NTRAIN).g, kB, 0, st>>>(c->mlp.P, T.G, m, v, NTRAIN, lr, b1, b2, eps, c1, c2);
        if (hipGetLastError() != hipSuccess) {
            mc_set_error("mc_mlp_train: kernel launch failed");
            rc = -100;
        }
    }
    if (!rc && hipStreamSynchronize(st) != hipSuccess) {
        mc_set_error("mc_mlp_train: stream error");
        rc = -100;
    }
    for (void *p : {(void *)dX, (void *)dY, (void *)m, (void *)v, (void *)didx}) if (p) (void)hipFree(p);
    return rc;
}

int mc_run_mlp_infer(mc_ctx *c, const float *dX, int n, float *dOut)
{
    if (n < 1) return 0;
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    const int ntiles = (n + 31) / 32;
    const int nblk = std::max(1, std::min((ntiles + 3) / 4, 4 * std::max(ncu, 1)));
    k_mlp_infer<<<nblk, 256, 0, c->stream>>>(c->mlp.P, dX, n, dOut);
    MC_HIP(hipGetLastError());
    return 0;
}

int mc_run_mlp_features(mc_ctx *c, float *dX)
{
    k_mlp_features<<<(c->nleaf + 255) / 256, 256, 0, c->stream>>>(c->st, c->nleaf, dX);
    MC_HIP(hipGetLastError());
    return 0;
}

extern "C" int mc_mlp_infer(mc_ctx *c, const float *X, int n, float *out)
{
    CHECK_CTX_MLP(c);
    if (!c->mlp.ready || !X || !out || n < 1 || n > (1 << 24)) {
        mc_set_error("mc_mlp_infer: need a network, 1 <= n <= 2^24 rows, X, out");
        return -1;
    }
    float *dX = nullptr, *dO = nullptr;
    MC_HIP(hipMalloc(&dX, sizeof(float) * (size_t)n * NI));
    if (hipMalloc(&dO, sizeof(float) * (size_t)n * NO) != hipSuccess) {
        (void)hipFree(dX);
        mc_set_error("mc_mlp_infer: allocation failed");
        return -100;
    }
    int rc = 0;
    if (hipMemcpy(dX, X, sizeof(float) * (size_t)n * NI, hipMemcpyHostToDevice) != hipSuccess) rc = -100;
    if (!rc) rc = mc_run_mlp_infer(c, dX, n, dO);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = -100;
    if (!rc && hipMemcpy(out, dO, sizeof(float) * (size_t)n * NO, hipMemcpyDeviceToHost) != hipSuccess)
        rc = -100;
    if (rc) mc_set_error("mc_mlp_infer: device error");
    (void)hipFree(dX);
    (void)hipFree(dO);
    return rc;
}

extern "C" int mc_mlp_features(mc_ctx *c, float *X)
{
    CHECK_CTX_MLP(c);
    if (!X || !c->st.nH) { mc_set_error("mc_mlp_features: no grid/state or null output"); return -1; }
    float *dX = nullptr;
    MC_HIP(hipMalloc(&dX, sizeof(float) * (size_t)c->nleaf * NI));
    int rc = mc_run_mlp_features(c, dX);
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = -100;
    if (!rc && hipMemcpy(X, dX, sizeof(float) * (size_t)c->nleaf * NI, hipMemcpyDeviceToHost) != hipSuccess)
        rc = -100;
    if (rc) mc_set_error("mc_mlp_features: device error");
    (void)hipFree(dX);
    return rc;
}
