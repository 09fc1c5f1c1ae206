// Multilayer perceptron of the learned GMRES warm start: forward, backward
// and Adam training on the GPU with BF16 matrix-core products and FP32
// accumulation.  No counterpart in the reference (its GMRES always starts
// from zeros, ErosionPackage/src/DensityStress.jl:45).
//
// Network: nlayers hidden layers of nhid units with tanh, then a linear
// output layer: h_0 = x, h_{l+1} = tanh(W_l h_l + b_l), y = W_L h_L + b_L.
// Parameters are FP32 master copies in one flat array, layer by layer,
// W_l (rows = outputs, row-major) then b_l.  Every product rounds both
// operands to BF16 (round to nearest even) and accumulates in FP32 with
// v_mfma_f32_16x16x16_bf16; bias, tanh, loss and Adam run in FP32.
//
// Loss: mean squared error over the batch and the outputs,
// L = sum (y - t)^2 / (B nout).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "common.hpp"

namespace stokes {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kT = 16;  // MFMA tile (16 x 16 x 16)

// FP32 -> BF16 bits, round to nearest even (inputs are finite)
__device__ __forceinline__ short bf16_bits(float f) {
    const unsigned u = __float_as_uint(f);
    return static_cast<short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// C[i][j] = sum_k A(i,k) B(k,j) (+ bias[j], tanh), FP32 out, BF16 products.
// A(i,k) = A[i lai + k lak], B(k,j) = B[k lbk + j lbj], C[i ldc + j].
// One wave per 16 x 16 tile of C; 4 waves per block (2 x 2 tiles).  Lane l
// feeds A[l%16][4(l/16)..+3] and B[4(l/16)..+3][l%16] and receives
// C[4(l/16)..+3][l%16].  epi: 0 store, 1 + bias, 2 + bias then tanh,
// 3 accumulate (C += product).
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const float* __restrict__ A,
                                                   long lai, long lak, const float* __restrict__ B,
                                                   long lbk, long lbj, float* __restrict__ C,
                                                   long ldc, const float* __restrict__ bias,
                                                   int epi) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int i0 = (blockIdx.y * 2 + (wv >> 1)) * kT;
    const int j0 = (blockIdx.x * 2 + (wv & 1)) * kT;
    if (i0 >= M || j0 >= N) return;  // wave-uniform
    const int r = lane % kT, q = lane / kT;
    const int ia = i0 + r, jb = j0 + r;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += kT) {
        v4s a, b;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int k = k0 + 4 * q + t;
            const float av = (ia < M && k < K) ? A[ia * lai + k * lak] : 0.f;
            const float bv = (jb < N && k < K) ? B[k * lbk + jb * lbj] : 0.f;
            a[t] = bf16_bits(av);
            b[t] = bf16_bits(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int j = j0 + r;
    if (j >= N) return;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int i = i0 + 4 * q + t;
        if (i >= M) continue;
        float v = acc[t];
        if (epi == 1 || epi == 2) v += bias[j];
        if (epi == 2) v = tanhf(v);
        float* c = C + static_cast<size_t>(i) * ldc + j;
        *c = epi == 3 ? *c + v : v;
    }
}

// g[b][j] *= 1 - h[b][j]^2 (tanh derivative from the layer output)
__global__ void k_tanh_back(float* __restrict__ g, const float* __restrict__ h, size_t n) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n) g[e] *= 1.f - h[e] * h[e];
}

// db[j] = sum_b g[b][j]
__global__ void k_colsum(const float* __restrict__ g, int B, int N, float* __restrict__ db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float s = 0.f;
    for (int b = 0; b < B; ++b) s += g[static_cast<size_t>(b) * N + j];
    db[j] = s;
}

// g = 2 (y - t) / (B nout); per-block partial sums of (y - t)^2
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, size_t n,
                      float scale, float* __restrict__ g, double* __restrict__ part) {
    __shared__ double sh[256];
    const size_t e = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    double d2 = 0.0;
    if (e < n) {
        const float d = y[e] - t[e];
        g[e] = 2.f * d * scale;
        d2 = static_cast<double>(d) * d;
    }
    sh[threadIdx.x] = d2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// rows of src selected by idx
__global__ void k_gather_rows(const float* __restrict__ src, const int* __restrict__ idx, int B,
                              int W, float* __restrict__ dst) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<size_t>(B) * W) return;
    const int b = static_cast<int>(e / W), c = static_cast<int>(e - static_cast<size_t>(b) * W);
    dst[e] = src[static_cast<size_t>(idx[b]) * W + c];
}

// Adam (Kingma and Ba 2015) with bias correction
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    w[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline int cdiv_int(int a, int b) { return (a + b - 1) / b; }

struct Layer {
    int nin, nout;
    size_t w, b;  // offsets of W and b in the flat parameter array
};

std::vector<Layer> layout(int nin, int nhid, int nlayers, int nout, size_t* total) {
    std::vector<Layer> L;
    size_t off = 0;
    int prev = nin;
    for (int l = 0; l <= nlayers; ++l) {
        const int o = l < nlayers ? nhid : nout;
        Layer y{prev, o, off, off + static_cast<size_t>(o) * prev};
        off = y.b + o;
        L.push_back(y);
        prev = o;
    }
    *total = off;
    return L;
}

void gemm(hipStream_t s, int M, int N, int K, const float* A, long lai, long lak, const float* B,
          long lbk, long lbj, float* C, long ldc, const float* bias, int epi) {
    const dim3 grid(cdiv_int(N, 2 * kT), cdiv_int(M, 2 * kT));
    k_gemm_bf16<<<grid, 256, 0, s>>>(M, N, K, A, lai, lak, B, lbk, lbj, C, ldc, bias, epi);
    STOKES_LAUNCH_CHECK();
}

// Device workspace of one network and batch size.
struct Net {
    int nin, nhid, nlayers, nout, B;
    std::vector<Layer> L;
    size_t nparam = 0;
    DevBuf<float> w, g, m, v;       // parameters, gradients, Adam moments
    std::vector<DevBuf<float>> h;   // activations h_0 .. h_{L+1} (h_{L+1} = y)
    DevBuf<float> dh, dh2, tgt;     // backward scratch, targets
    DevBuf<double> part;

    Net(int nin_, int nhid_, int nlayers_, int nout_, int B_)
        : nin(nin_), nhid(nhid_), nlayers(nlayers_), nout(nout_), B(B_), h(nlayers_ + 2) {
        L = layout(nin, nhid, nlayers, nout, &nparam);
        w.ensure(nparam);
        g.ensure(nparam);
        h[0].ensure(static_cast<size_t>(B) * nin);
        for (int l = 0; l <= nlayers; ++l)
            h[l + 1].ensure(static_cast<size_t>(B) * L[l].nout);
        const size_t wide = static_cast<size_t>(B) * std::max(nhid, nout);
        dh.ensure(wide);
        dh2.ensure(wide);
        tgt.ensure(static_cast<size_t>(B) * nout);
        part.ensure(static_cast<size_t>(cdiv_int(B * nout, 256)) + 1);
    }

    void forward(hipStream_t s, int nb) {
        for (int l = 0; l <= nlayers; ++l) {
            const Layer& y = L[l];
            // h_{l+1}[b][o] = sum_k h_l[b][k] W[o][k] + b[o]
            gemm(s, nb, y.nout, y.nin, h[l].p, y.nin, 1, w.p + y.w, 1, y.nin, h[l + 1].p, y.nout,
                 w.p + y.b, l < nlayers ? 2 : 1);
        }
    }

    // gradients of the mean squared error into g; returns the loss
    double backward(hipStream_t s, int nb) {
        const size_t ny = static_cast<size_t>(nb) * nout;
        const int nbk = cdiv_int(static_cast<int>(ny), 256);
        k_mse<<<nbk, 256, 0, s>>>(h[nlayers + 1].p, tgt.p, ny,
                                  1.f / static_cast<float>(ny), dh.p, part.p);
        STOKES_LAUNCH_CHECK();
        float* gcur = dh.p;  // gradient w.r.t. the pre-activation of layer l
        float* gnext = dh2.p;
        for (int l = nlayers; l >= 0; --l) {
            const Layer& y = L[l];
            // dW[o][k] = sum_b gcur[b][o] h_l[b][k]
            gemm(s, y.nout, y.nin, nb, gcur, 1, y.nout, h[l].p, y.nin, 1, g.p + y.w, y.nin,
                 nullptr, 0);
            k_colsum<<<cdiv_int(y.nout, 256), 256, 0, s>>>(gcur, nb, y.nout, g.p + y.b);
            STOKES_LAUNCH_CHECK();
            if (l == 0) break;
            // dh_l[b][k] = sum_o gcur[b][o] W[o][k], then times tanh'
            gemm(s, nb, y.nin, y.nout, gcur, y.nout, 1, w.p + y.w, y.nin, 1, gnext, y.nin,
                 nullptr, 0);
            const size_t nh = static_cast<size_t>(nb) * y.nin;
            k_tanh_back<<<static_cast<int>((nh + 255) / 256), 256, 0, s>>>(gnext, h[l].p, nh);
            STOKES_LAUNCH_CHECK();
            std::swap(gcur, gnext);
        }
        std::vector<double> hp(static_cast<size_t>(nbk));
        STOKES_HIP(hipMemcpyAsync(hp.data(), part.p, hp.size() * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
        STOKES_HIP(hipStreamSynchronize(s));
        double loss = 0.0;
        for (double v : hp) loss += v;
        return loss / static_cast<double>(ny);
    }
};

void check_shape(int nin, int nhid, int nlayers, int nout) {
    if (nin < 1 || nhid < 1 || nlayers < 1 || nout < 1 || nin > 65536 || nhid > 65536 ||
        nout > 65536 || nlayers > 64)
        throw Error("invalid network shape");
}

}  // namespace

namespace {

// ----------------------------------------------- allocation-free forward
// Hidden activations of mlp_forward_noalloc live in this file-scope array
// (8 MiB on every GPU the library runs on, held until the process ends):
// two halves of kScratchFloats / 2 floats, layer l writing one half and
// reading the other, rows of samples processed in chunks that fit a half.
constexpr size_t kScratchFloats = size_t{1} << 21;
static __device__ __align__(256) float scratch[kScratchFloats];

std::mutex g_scratch_mu;
std::map<int, float*> g_scratch_ptr;  // device -> address of `scratch` there

// The device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array).
float* scratch_on_current_device() {
    int dev = 0;
    STOKES_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void* p = nullptr;
    STOKES_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error("no device address for the forward pass's scratch array");
    g_scratch_ptr.emplace(dev, static_cast<float*>(p));
    return static_cast<float*>(p);
}

}  // namespace

size_t mlp_param_count(int nin, int nhid, int nlayers, int nout) {
    check_shape(nin, nhid, nlayers, nout);
    size_t total = 0;
    layout(nin, nhid, nlayers, nout, &total);
    return total;
}

void mlp_forward(Context& c, int nin, int nhid, int nlayers, int nout, const float* params,
                 int nsamples, const float* x, float* y) {
    check_shape(nin, nhid, nlayers, nout);
    if (nsamples < 0) throw Error("negative sample count");
    if (nsamples == 0) return;
    const int B = std::min(nsamples, 16384);
    Net net(nin, nhid, nlayers, nout, B);
    hipStream_t s = c.stream;
    STOKES_HIP(hipMemcpyAsync(net.w.p, params, net.nparam * sizeof(float), hipMemcpyHostToDevice, s));
    for (int b0 = 0; b0 < nsamples; b0 += B) {
        const int nb = std::min(B, nsamples - b0);
        STOKES_HIP(hipMemcpyAsync(net.h[0].p, x + static_cast<size_t>(b0) * nin,
                                  static_cast<size_t>(nb) * nin * sizeof(float),
                                  hipMemcpyHostToDevice, s));
        net.forward(s, nb);
        STOKES_HIP(hipMemcpyAsync(y + static_cast<size_t>(b0) * nout, net.h[nlayers + 1].p,
                                  static_cast<size_t>(nb) * nout * sizeof(float),
                                  hipMemcpyDeviceToHost, s));
    }
    STOKES_HIP(hipStreamSynchronize(s));
}

void mlp_forward_noalloc(int nin, int nhid, int nlayers, int nout, const float* params_dev,
                         int nsamples, const float* x_dev, float* y_dev, hipStream_t s) {
    check_shape(nin, nhid, nlayers, nout);
    if (nsamples < 0) throw Error("negative sample count");
    const size_t half = kScratchFloats / 2;
    // a chunk holds at least one 32-row tile pair of hidden activations
    if (static_cast<size_t>(nhid) * 2 * kT > half)
        throw Error("hidden width " + std::to_string(nhid) + " exceeds the forward scratch (" +
                    std::to_string(half / (2 * kT)) + " units at most)");
    float* base = scratch_on_current_device();
    if (nsamples == 0) return;
    size_t total = 0;
    const std::vector<Layer> L = layout(nin, nhid, nlayers, nout, &total);
    // rows per chunk: as many as a half holds, a multiple of the 32-row tile
    const int rows = static_cast<int>(std::min<size_t>(half / nhid / (2 * kT) * (2 * kT),
                                                       static_cast<size_t>(1) << 20));
    if (rows < 2 * kT) throw Error("forward scratch holds no row chunk");
    float* buf[2] = {base, base + half};
    for (int b0 = 0; b0 < nsamples; b0 += rows) {
        const int nb = std::min(rows, nsamples - b0);
        const float* in = x_dev + static_cast<size_t>(b0) * nin;
        for (int l = 0; l <= nlayers; ++l) {
            const Layer& y = L[static_cast<size_t>(l)];
            const bool last = l == nlayers;
            float* out = last ? y_dev + static_cast<size_t>(b0) * nout : buf[l & 1];
            gemm(s, nb, y.nout, y.nin, in, y.nin, 1, params_dev + y.w, 1, y.nin, out, y.nout,
                 params_dev + y.b, last ? 1 : 2);
            in = out;
        }
    }
}

double mlp_gradients(Context& c, int nin, int nhid, int nlayers, int nout, const float* params,
                     int nsamples, const float* x, const float* t, float* grads) {
    check_shape(nin, nhid, nlayers, nout);
    if (nsamples < 1) throw Error("gradients need at least one sample");
    Net net(nin, nhid, nlayers, nout, nsamples);
    hipStream_t s = c.stream;
    STOKES_HIP(hipMemcpyAsync(net.w.p, params, net.nparam * sizeof(float), hipMemcpyHostToDevice, s));
    STOKES_HIP(hipMemcpyAsync(net.h[0].p, x, static_cast<size_t>(nsamples) * nin * sizeof(float),
                              hipMemcpyHostToDevice, s));
    STOKES_HIP(hipMemcpyAsync(net.tgt.p, t, static_cast<size_t>(nsamples) * nout * sizeof(float),
                              hipMemcpyHostToDevice, s));
    net.forward(s, nsamples);
    const double loss = net.backward(s, nsamples);
    STOKES_HIP(hipMemcpyAsync(grads, net.g.p, net.nparam * sizeof(float), hipMemcpyDeviceToHost, s));
    STOKES_HIP(hipStreamSynchronize(s));
    return loss;
}

void mlp_train(Context& c, int nin, int nhid, int nlayers, int nout, float* params, int nsamples,
               const float* x, const float* t, int epochs, int batch, float lr, uint64_t seed,
               double* epoch_loss) {
    check_shape(nin, nhid, nlayers, nout);
    if (nsamples < 1 || epochs < 1 || batch < 1 || !(lr > 0.f)) throw Error("invalid training setup");
    batch = std::min(batch, nsamples);
    Net net(nin, nhid, nlayers, nout, batch);
    hipStream_t s = c.stream;
    net.m.ensure(net.nparam);
    net.v.ensure(net.nparam);
    STOKES_HIP(hipMemsetAsync(net.m.p, 0, net.nparam * sizeof(float), s));
    STOKES_HIP(hipMemsetAsync(net.v.p, 0, net.nparam * sizeof(float), s));
    STOKES_HIP(hipMemcpyAsync(net.w.p, params, net.nparam * sizeof(float), hipMemcpyHostToDevice, s));
    DevBuf<float> X, T;
    DevBuf<int> idx;
    X.ensure(static_cast<size_t>(nsamples) * nin);
    T.ensure(static_cast<size_t>(nsamples) * nout);
    idx.ensure(static_cast<size_t>(nsamples));
    STOKES_HIP(hipMemcpyAsync(X.p, x, static_cast<size_t>(nsamples) * nin * sizeof(float),
                              hipMemcpyHostToDevice, s));
    STOKES_HIP(hipMemcpyAsync(T.p, t, static_cast<size_t>(nsamples) * nout * sizeof(float),
                              hipMemcpyHostToDevice, s));
    std::mt19937_64 rng(seed);
    std::vector<int> perm(static_cast<size_t>(nsamples));
    for (int i = 0; i < nsamples; ++i) perm[static_cast<size_t>(i)] = i;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    long long step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        STOKES_HIP(hipMemcpyAsync(idx.p, perm.data(), perm.size() * sizeof(int),
                                  hipMemcpyHostToDevice, s));
        double loss_sum = 0.0;
        int nbatches = 0;
        for (int b0 = 0; b0 + batch <= nsamples; b0 += batch) {
            const size_t nx = static_cast<size_t>(batch) * nin, nt = static_cast<size_t>(batch) * nout;
            k_gather_rows<<<static_cast<int>((nx + 255) / 256), 256, 0, s>>>(X.p, idx.p + b0, batch,
                                                                             nin, net.h[0].p);
            k_gather_rows<<<static_cast<int>((nt + 255) / 256), 256, 0, s>>>(T.p, idx.p + b0, batch,
                                                                             nout, net.tgt.p);
            STOKES_LAUNCH_CHECK();
            net.forward(s, batch);
            loss_sum += net.backward(s, batch);
            ++step;
            const float c1 = 1.f - std::pow(b1, static_cast<float>(step));
            const float c2 = 1.f - std::pow(b2, static_cast<float>(step));
            k_adam<<<static_cast<int>((net.nparam + 255) / 256), 256, 0, s>>>(
                net.w.p, net.g.p, net.m.p, net.v.p, net.nparam, lr, b1, b2, eps, c1, c2);
            STOKES_LAUNCH_CHECK();
            ++nbatches;
        }
        if (epoch_loss) epoch_loss[ep] = nbatches ? loss_sum / nbatches : 0.0;
    }
    STOKES_HIP(hipMemcpyAsync(params, net.w.p, net.nparam * sizeof(float), hipMemcpyDeviceToHost, s));
    STOKES_HIP(hipStreamSynchronize(s));
}

}  // namespace stokes
