// ft_learned.hip -- learned k-th-neighbour radius (no reference counterpart).
//
// A 12 -> 64 -> 64 -> 64 -> 1 ReLU MLP predicts log d2_k of a self query from
// local tree density.  Its prediction, plus a calibrated margin, becomes the
// query's seed bound for the packet traversal; a query that finds fewer than k
// neighbours inside its seed is re-run unbounded, so results never change.
//
// Features (fp32, 16 slots, the last four zero):
//   0-3   log box volume of the ancestors 3, 6, 9, 12 levels above the leaf
//   4-7   log leaf count of the same ancestors
//   8     log k
//   9-11  log d2 to the leaves at ranks r-1, r+1 and the larger of r +- k/2
// Volumes and d2 are in coordinate units; eps = 1 for integer coordinates and
// 1e-12 * (root extent)^2 for FP coordinates keeps the logs finite.
//
// Layers 1-3 run on MFMA (v_mfma_f32_16x16x16_bf16): bf16 weights and bf16
// activations, fp32 accumulation; the 64 -> 1 output is an fp32 dot product.
// Training: mean-squared error on the standardised target, gradients from the
// hand-written backward (also MFMA), Adam on fp32 master weights.
#include "ft_internal.hpp"

#include <cmath>
#include <vector>

namespace ft {

int knn_entry(hipStream_t s, const ft_tree *tree, const void *qx, const void *qy, const void *qz,
              const int32_t *leaf_ids, int nq, int k, int excl, uint64_t *res, void *dists,
              const ft_query_opts *opts);

// ---- parameter layout (fp32) ------------------------------------------------
constexpr int NF = 16;   // padded feature count
constexpr int NH = 64;   // hidden width
constexpr int P_W1 = 0;                  // [16][64]  (in, out)
constexpr int P_B1 = P_W1 + NF * NH;     // [64]
constexpr int P_W2 = P_B1 + NH;          // [64][64]
constexpr int P_B2 = P_W2 + NH * NH;     // [64]
constexpr int P_W3 = P_B2 + NH;          // [64][64]
constexpr int P_B3 = P_W3 + NH * NH;     // [64]
constexpr int P_W4 = P_B3 + NH;          // [64]
constexpr int P_B4 = P_W4 + NH;          // [1] (+15 unused)
constexpr int NPARAM = P_B4 + 16;        // 9488
// Normalisation block: feature mean[16], feature 1/std[16], target mean, target std.
constexpr int NNORM = 2 * NF + 2;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __host__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite values)
    return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
// The same rounding by the hardware conversion (v_cvt_pk_bf16_f32, round to
// nearest even; equal to to_bf16 for every finite value), for the inference
// loop's activations.
__device__ inline short bf16_hw(float f) { return __builtin_bit_cast(short, (__bf16)f); }

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// ===========================================================================
// Features
// ===========================================================================
template <typename T>
__global__ __launch_bounds__(256) void features_kernel(DTree<T> t, const int32_t *__restrict__ leaf_ids,
                                                       int nq, int k, float *__restrict__ feat) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const int n = t.n;
    const int off = n - 1;
    const int r = leaf_ids ? leaf_ids[q] : q;
    double eps = 1.0;
    if constexpr (CoordTraits<T>::fp) {
        const double ex = (double)t.mxx[0] - (double)t.mnx[0];
        const double ey = (double)t.mxy[0] - (double)t.mny[0];
        const double ez = (double)t.mxz[0] - (double)t.mnz[0];
        const double e = fmax(ex, fmax(ey, ez));
        eps = e > 0 ? 1e-12 * e * e : 1e-300;
    }
    const double vol_eps = CoordTraits<T>::fp ? cbrt(eps) : 1.0;
    float f[NF];
    int node = r + off;
    int anc = node;
    int level = 0;
    for (int a = 0; a < 4; ++a) {
        const int want = 3 * (a + 1);
        while (level < want && n > 1) {
            const int p = t.parent[anc];
            if (p < 0) break;
            anc = p;
            ++level;
        }
        double cnt = 1.0, vol;
        const double ex = (double)(t.mxx[anc] - t.mnx[anc]) + vol_eps;
        const double ey = (double)(t.mxy[anc] - t.mny[anc]) + vol_eps;
        const double ez = (double)(t.mxz[anc] - t.mnz[anc]) + vol_eps;
        vol = ex * ey * ez;
        if (anc < off) cnt = (double)(t.rlast[anc] - t.rfirst[anc] + 1);
        f[a] = (float)log(vol);
        f[4 + a] = (float)log(cnt);
    }
    f[8] = (float)log((double)(k > 0 ? k : 1));
    const T px = t.mnx[node], py = t.mny[node], pz = t.mnz[node];
    auto d2to = [&](int rr) -> double {
        rr = rr < 0 ? 0 : (rr >= n ? n - 1 : rr);
        const int lf = rr + off;
        return (double)leaf_d2<T>(t.m, px, py, pz, t.mnx[lf], t.mny[lf], t.mnz[lf]);
    };
    const int h = k / 2 > 0 ? k / 2 : 1;
    f[9] = (float)log(d2to(r > 0 ? r - 1 : r + 1) + eps);
    f[10] = (float)log(d2to(r + 1 < n ? r + 1 : r - 1) + eps);
    f[11] = (float)log(fmax(d2to(r - h), d2to(r + h)) + eps);
    for (int j = 12; j < NF; ++j) f[j] = 0.f;
    float4 *o = (float4 *)(feat + (size_t)q * NF);
    o[0] = make_float4(f[0], f[1], f[2], f[3]);
    o[1] = make_float4(f[4], f[5], f[6], f[7]);
    o[2] = make_float4(f[8], f[9], f[10], f[11]);
    o[3] = make_float4(f[12], f[13], f[14], f[15]);
}

// ===========================================================================
// MLP forward / backward on MFMA.  MFMA 16x16x16 operand layouts (lane l):
//   A (M x K): A[l % 16][4*(l/16) + i]   B (K x N): B[4*(l/16) + i][l % 16]
//   D (M x N): D[4*(l/16) + i][l % 16]   for i = 0..3
// A tile is 16 samples.  Per block: weights staged in LDS as bf16 in both
// orientations (W[in][out] for dz W^T, Wt[out][in] for x W).  Per wave: the
// normalised input and the activations as bf16 row-major [16][64] (next
// layer's A operand) and transposed [64][16] (A operand of dW = a^T dz).
// ===========================================================================
struct MlpSmem {
    uint16_t w1[NF * NH], w1t[NH * NF];
    uint16_t w2[NH * NH], w2t[NH * NH];
    uint16_t w3[NH * NH], w3t[NH * NH];
    float b1[NH], b2[NH], b3[NH], w4[NH], b4;
    float fmean[NF], finv[NF], ymean, ystd;
};
struct MlpWave {
    uint16_t x[16 * NF], xt[NF * 16];
    uint16_t a1[16 * NH], a1t[NH * 16];
    uint16_t a2[16 * NH], a2t[NH * 16];
    uint16_t dz[16 * NH];
};

__device__ inline void stage_weights(MlpSmem &S, const float *__restrict__ prm,
                                     const float *__restrict__ norm) {
    for (int i = threadIdx.x; i < NF * NH; i += blockDim.x) {
        const int in = i / NH, out = i % NH;
        const uint16_t b = to_bf16(prm[P_W1 + i]);
        S.w1[i] = b;
        S.w1t[out * NF + in] = b;
    }
    for (int i = threadIdx.x; i < NH * NH; i += blockDim.x) {
        const int in = i / NH, out = i % NH;
        const uint16_t b2 = to_bf16(prm[P_W2 + i]);
        const uint16_t b3 = to_bf16(prm[P_W3 + i]);
        S.w2[i] = b2;
        S.w2t[out * NH + in] = b2;
        S.w3[i] = b3;
        S.w3t[out * NH + in] = b3;
    }
    for (int i = threadIdx.x; i < NH; i += blockDim.x) {
        S.b1[i] = prm[P_B1 + i];
        S.b2[i] = prm[P_B2 + i];
        S.b3[i] = prm[P_B3 + i];
        S.w4[i] = prm[P_W4 + i];
    }
    for (int i = threadIdx.x; i < NF; i += blockDim.x) {
        S.fmean[i] = norm[i];
        S.finv[i] = norm[NF + i];
    }
    if (threadIdx.x == 0) {
        S.b4 = prm[P_B4];
        S.ymean = norm[2 * NF];
        S.ystd = norm[2 * NF + 1];
    }
}

__device__ inline bf16x4 ld4(const uint16_t *p) {
    bf16x4 v;
    const uint2 u = *(const uint2 *)p;
    v[0] = (short)(u.x & 0xFFFF);
    v[1] = (short)(u.x >> 16);
    v[2] = (short)(u.y & 0xFFFF);
    v[3] = (short)(u.y >> 16);
    return v;
}
__device__ inline bf16x4 pack4(f32x4 f) {
    bf16x4 v;
    v[0] = (short)to_bf16(f[0]);
    v[1] = (short)to_bf16(f[1]);
    v[2] = (short)to_bf16(f[2]);
    v[3] = (short)to_bf16(f[3]);
    return v;
}

// Forward of one 16-sample tile.  Leaves z1..z3 (pre-activations, D layout,
// [nt][i]) in registers, activations in W, and returns the normalised output
// for samples 4*(lane/16) + i.
struct Fwd {
    f32x4 z1[4], z2[4], z3[4];
    f32x4 out;
};

__device__ inline void mlp_tile_forward(const MlpSmem &S, MlpWave &W, const float *__restrict__ feat,
                                        int s0, int n, Fwd &F) {
    const int l = threadIdx.x & 63;
    const int lr = l & 15, lg = l >> 4;
    // normalised inputs: A layout x[l%16][4*(l/16)+i]
    {
        const int s = s0 + lr;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (s < n) v = ((const float4 *)(feat + (size_t)s * NF))[lg];
        const float vi[4] = {v.x, v.y, v.z, v.w};
        for (int i = 0; i < 4; ++i) {
            const int f = 4 * lg + i;
            const uint16_t b = to_bf16((vi[i] - S.fmean[f]) * S.finv[f]);
            W.x[lr * NF + f] = b;
            W.xt[f * 16 + lr] = b;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const bf16x4 ax = ld4(&W.x[lr * NF + 4 * lg]);  // this lane's own four inputs
    // layer 1: K = 16
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(ax, ld4(&S.w1t[(nt * 16 + lr) * NF + 4 * lg]), acc);
        const float b = S.b1[nt * 16 + lr];
        for (int i = 0; i < 4; ++i) {
            acc[i] += b;
            const float a = acc[i] > 0.f ? acc[i] : 0.f;
            const uint16_t h = to_bf16(a);
            W.a1[(4 * lg + i) * NH + nt * 16 + lr] = h;
            W.a1t[(nt * 16 + lr) * 16 + 4 * lg + i] = h;
        }
        F.z1[nt] = acc;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // layer 2: K = 64 in 4 chunks
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < 4; ++kc)
            acc = mfma16(ld4(&W.a1[lr * NH + kc * 16 + 4 * lg]),
                         ld4(&S.w2t[(nt * 16 + lr) * NH + kc * 16 + 4 * lg]), acc);
        const float b = S.b2[nt * 16 + lr];
        for (int i = 0; i < 4; ++i) {
            acc[i] += b;
            const float a = acc[i] > 0.f ? acc[i] : 0.f;
            const uint16_t h = to_bf16(a);
            W.a2[(4 * lg + i) * NH + nt * 16 + lr] = h;
            W.a2t[(nt * 16 + lr) * 16 + 4 * lg + i] = h;
        }
        F.z2[nt] = acc;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // layer 3 + output dot product (fp32)
    f32x4 part = {0.f, 0.f, 0.f, 0.f};
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < 4; ++kc)
            acc = mfma16(ld4(&W.a2[lr * NH + kc * 16 + 4 * lg]),
                         ld4(&S.w3t[(nt * 16 + lr) * NH + kc * 16 + 4 * lg]), acc);
        const float b = S.b3[nt * 16 + lr];
        const float w = S.w4[nt * 16 + lr];
        for (int i = 0; i < 4; ++i) {
            acc[i] += b;
            part[i] += (acc[i] > 0.f ? acc[i] : 0.f) * w;
        }
        F.z3[nt] = acc;
    }
    // reduce over the 16 lanes of a group (same samples, different units)
    for (int i = 0; i < 4; ++i) {
        float v = part[i];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        F.out[i] = v + S.b4;
    }
}

// ===========================================================================
// Inference (predictions and seeds).  Activations are kept transposed:
// Z^T = W^T H^T, with the weight tile as the A operand and the activation tile
// as B, so a layer's MFMA output (D: unit 4*(l/16)+i of the tile, sample
// l%16) is already the B operand of the next layer and never goes through
// LDS.  Each wave holds every weight tile in registers (A operands), the
// biases as accumulator initial values, and streams 16-sample tiles.
// Same roundings as the training forward: bf16 weights, bf16 normalised
// inputs and hidden activations, fp32 accumulation and fp32 output layer.
// Seeds (dist_t) = exp(ymean + ystd * out + margin) - eps; `pred` receives the
// raw prediction (log d2) when given.
// ===========================================================================
template <typename D>
__global__ __launch_bounds__(256) void mlp_infer_kernel(const float *__restrict__ prm,
                                                        const float *__restrict__ norm,
                                                        const float *__restrict__ feat, int n,
                                                        float *__restrict__ pred,
                                                        D *__restrict__ seed, double margin,
                                                        double eps) {
    const int l = threadIdx.x & 63, lr = l & 15, lg = l >> 4;
    auto wt = [&](int base, int kc, int nt) -> bf16x4 {  // W^T tile: W[kc*16+4lg+i][nt*16+lr]
        bf16x4 v;
        for (int i = 0; i < 4; ++i)
            v[i] = (short)to_bf16(prm[base + (kc * 16 + 4 * lg + i) * NH + nt * 16 + lr]);
        return v;
    };
    auto col = [&](int base, int nt) -> f32x4 {  // per D row: unit nt*16+4lg+i
        f32x4 v;
        for (int i = 0; i < 4; ++i) v[i] = prm[base + nt * 16 + 4 * lg + i];
        return v;
    };
    bf16x4 a1[4], a2[4][4], a3[4][4];
    f32x4 b1[4], b2[4], b3[4], w4[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        a1[nt] = wt(P_W1, 0, nt);
#pragma unroll
        for (int kc = 0; kc < 4; ++kc) {
            a2[nt][kc] = wt(P_W2, kc, nt);
            a3[nt][kc] = wt(P_W3, kc, nt);
        }
        b1[nt] = col(P_B1, nt);
        b2[nt] = col(P_B2, nt);
        b3[nt] = col(P_B3, nt);
        w4[nt] = col(P_W4, nt);
    }
    const float b4 = prm[P_B4];
    float fm[4], fi[4];
    for (int i = 0; i < 4; ++i) {
        fm[i] = norm[4 * lg + i];
        fi[i] = norm[NF + 4 * lg + i];
    }
    const float ymean = norm[2 * NF], ystd = norm[2 * NF + 1];
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * 4;
    // The next tile's features are loaded while this tile is computed.
    auto load_tile = [&](int t0) -> float4 {
        const int s = t0 + lr;
        return s < n ? ((const float4 *)(feat + (size_t)s * NF))[lg]
                     : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    float4 vnext = load_tile(wave * 16);
    for (int s0 = wave * 16; s0 < n; s0 += nwaves * 16) {
        const int s = s0 + lr;
        const float4 v = vnext;
        vnext = load_tile(s0 + nwaves * 16);
        const float vi[4] = {v.x, v.y, v.z, v.w};
        bf16x4 x;
        for (int i = 0; i < 4; ++i) x[i] = bf16_hw((vi[i] - fm[i]) * fi[i]);
        bf16x4 h1[4], h2[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const f32x4 z = mfma16(a1[nt], x, b1[nt]);
            for (int i = 0; i < 4; ++i) h1[nt][i] = bf16_hw(z[i] > 0.f ? z[i] : 0.f);
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 z = b2[nt];
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) z = mfma16(a2[nt][kc], h1[kc], z);
            for (int i = 0; i < 4; ++i) h2[nt][i] = bf16_hw(z[i] > 0.f ? z[i] : 0.f);
        }
        float part = 0.f;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 z = b3[nt];
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) z = mfma16(a3[nt][kc], h2[kc], z);
            for (int i = 0; i < 4; ++i) part += (z[i] > 0.f ? z[i] : 0.f) * w4[nt][i];
        }
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (lg == 0 && s < n) {
            const float y = ymean + ystd * (part + b4);
            if (pred) pred[s] = y;
            if (seed) {
                double sv = exp((double)y + margin) - eps;
                if (sv < 0) sv = 0;
                if constexpr (std::is_floating_point<D>::value) {
                    seed[s] = sv;
                } else {
                    seed[s] = sv >= 9.0e18 ? (D)9000000000000000000ull : (D)sv;
                }
            }
        }
    }
}

// Gradient of L = (1/total) * sum (out - t)^2 over the samples of this call,
// t = (y - ymean) / ystd, accumulated into grad[NPARAM] (fp32 atomics) and
// loss[0].
__global__ __launch_bounds__(256) void mlp_grad_kernel(const float *__restrict__ prm,
                                                       const float *__restrict__ norm,
                                                       const float *__restrict__ feat,
                                                       const float *__restrict__ target, int n,
                                                       float inv_total, float *__restrict__ grad,
                                                       float *__restrict__ loss) {
    __shared__ MlpSmem S;
    __shared__ MlpWave Wv[4];
    stage_weights(S, prm, norm);
    __syncthreads();
    MlpWave &W = Wv[threadIdx.x >> 6];
    const int l = threadIdx.x & 63, lg = l >> 4, lr = l & 15;
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int nwaves = gridDim.x * 4;
    f32x4 gW1[4], gW2[4][4], gW3[4][4];
    float gb1[4], gb2[4], gb3[4], gw4[4], gb4 = 0.f, lsum = 0.f;
    for (int a = 0; a < 4; ++a) {
        gW1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
        gb1[a] = gb2[a] = gb3[a] = gw4[a] = 0.f;
        for (int b = 0; b < 4; ++b) {
            gW2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            gW3[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    }
    for (int s0 = wave * 16; s0 < n; s0 += nwaves * 16) {
        Fwd F;
        mlp_tile_forward(S, W, feat, s0, n, F);
        // dout for samples 4*lg + i
        f32x4 dout;
        for (int i = 0; i < 4; ++i) {
            const int s = s0 + 4 * lg + i;
            float d = 0.f;
            if (s < n) {
                const float t = (target[s] - S.ymean) / S.ystd;
                const float e = F.out[i] - t;
                if (lr == 0) lsum += e * e;
                d = 2.f * e * inv_total;
            }
            dout[i] = d;
        }
        // layer 4 and dz3
        f32x4 dz3[4];
        for (int nt = 0; nt < 4; ++nt) {
            const float w = S.w4[nt * 16 + lr];
            for (int i = 0; i < 4; ++i) {
                const float z = F.z3[nt][i];
                const float a = z > 0.f ? z : 0.f;
                gw4[nt] += dout[i] * a;
                dz3[nt][i] = z > 0.f ? dout[i] * w : 0.f;
            }
        }
        gb4 += dout[0] + dout[1] + dout[2] + dout[3];
        // dW3 += a2^T dz3 ; db3
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x4 bz = pack4(dz3[nt]);
            for (int mt = 0; mt < 4; ++mt)
                gW3[mt][nt] = mfma16(ld4(&W.a2t[(mt * 16 + lr) * 16 + 4 * lg]), bz, gW3[mt][nt]);
            gb3[nt] += dz3[nt][0] + dz3[nt][1] + dz3[nt][2] + dz3[nt][3];
            for (int i = 0; i < 4; ++i) W.dz[(4 * lg + i) * NH + nt * 16 + lr] = to_bf16(dz3[nt][i]);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        // da2 = dz3 W3^T ; dz2
        f32x4 dz2[4];
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kc = 0; kc < 4; ++kc)
                acc = mfma16(ld4(&W.dz[lr * NH + kc * 16 + 4 * lg]),
                             ld4(&S.w3[(nt * 16 + lr) * NH + kc * 16 + 4 * lg]), acc);
            for (int i = 0; i < 4; ++i) dz2[nt][i] = F.z2[nt][i] > 0.f ? acc[i] : 0.f;
        }
        __builtin_amdgcn_wave_barrier();
        // dW2 += a1^T dz2 ; db2
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x4 bz = pack4(dz2[nt]);
            for (int mt = 0; mt < 4; ++mt)
                gW2[mt][nt] = mfma16(ld4(&W.a1t[(mt * 16 + lr) * 16 + 4 * lg]), bz, gW2[mt][nt]);
            gb2[nt] += dz2[nt][0] + dz2[nt][1] + dz2[nt][2] + dz2[nt][3];
            for (int i = 0; i < 4; ++i) W.dz[(4 * lg + i) * NH + nt * 16 + lr] = to_bf16(dz2[nt][i]);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        // da1 = dz2 W2^T ; dz1 ; dW1 += x^T dz1 ; db1
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kc = 0; kc < 4; ++kc)
                acc = mfma16(ld4(&W.dz[lr * NH + kc * 16 + 4 * lg]),
                             ld4(&S.w2[(nt * 16 + lr) * NH + kc * 16 + 4 * lg]), acc);
            f32x4 dz1;
            for (int i = 0; i < 4; ++i) dz1[i] = F.z1[nt][i] > 0.f ? acc[i] : 0.f;
            gW1[nt] = mfma16(ld4(&W.xt[lr * 16 + 4 * lg]), pack4(dz1), gW1[nt]);
            gb1[nt] += dz1[0] + dz1[1] + dz1[2] + dz1[3];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    // Bias and w4 sums: reduce over the four lane groups (same unit l%16).
    auto grp = [](float v) {
        v += __shfl_xor(v, 16, 64);
        v += __shfl_xor(v, 32, 64);
        return v;
    };
    for (int nt = 0; nt < 4; ++nt) {
        const float a = grp(gb1[nt]), b = grp(gb2[nt]), c = grp(gb3[nt]), d = grp(gw4[nt]);
        if (lg == 0) {
            atomicAdd(&grad[P_B1 + nt * 16 + lr], a);
            atomicAdd(&grad[P_B2 + nt * 16 + lr], b);
            atomicAdd(&grad[P_B3 + nt * 16 + lr], c);
            atomicAdd(&grad[P_W4 + nt * 16 + lr], d);
        }
    }
    float gb4s = gb4;
    gb4s += __shfl_xor(gb4s, 16, 64);
    gb4s += __shfl_xor(gb4s, 32, 64);
    if (l == 0) atomicAdd(&grad[P_B4], gb4s);
    // Weight tiles: lane holds dW[mt*16 + 4*lg + i][nt*16 + lr].
    for (int nt = 0; nt < 4; ++nt) {
        for (int i = 0; i < 4; ++i) {
            const int in = 4 * lg + i, out = nt * 16 + lr;
            atomicAdd(&grad[P_W1 + in * NH + out], gW1[nt][i]);
            for (int mt = 0; mt < 4; ++mt) {
                const int in2 = mt * 16 + 4 * lg + i;
                atomicAdd(&grad[P_W2 + in2 * NH + out], gW2[mt][nt][i]);
                atomicAdd(&grad[P_W3 + in2 * NH + out], gW3[mt][nt][i]);
            }
        }
    }
    float ls = lsum;
    ls += __shfl_xor(ls, 16, 64);
    ls += __shfl_xor(ls, 32, 64);
    if (l == 0) atomicAdd(loss, ls * inv_total);
}

// Adam on the fp32 master weights.  The padded input rows of W1 stay zero.
__global__ void adam_kernel(float *__restrict__ prm, const float *__restrict__ grad,
                            float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    if (i >= P_W1 + 12 * NH && i < P_W1 + NF * NH) return;
    if (i > P_B4) return;
    const float g = grad[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    prm[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void zero_f32_kernel(float *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

}  // namespace ft

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
struct ft_mlp {
    float *prm = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr, *norm = nullptr,
          *loss = nullptr;
    void *block = nullptr;
    double margin = 0.0;
    long step = 0;
};

using namespace ft;

extern "C" {

int ft_mlp_num_params(void) { return NPARAM; }

int ft_mlp_create(ft_mlp **out) {
    if (!out) return fail(FT_EINVAL, "null model pointer");
    ft_mlp *mdl = new ft_mlp();
    const size_t bytes = (size_t)(4 * NPARAM + NNORM + 16) * 4;
    if (hipMalloc(&mdl->block, bytes) != hipSuccess) {
        delete mdl;
        return fail(FT_ENOMEM, "model allocation failed");
    }
    FT_HIP(hipMemset(mdl->block, 0, bytes));
    FT_HIP(hipDeviceSynchronize());  // the null-stream clear, before any stream uses the block
    float *b = (float *)mdl->block;
    mdl->prm = b;
    mdl->grad = b + NPARAM;
    mdl->m = b + 2 * NPARAM;
    mdl->v = b + 3 * NPARAM;
    mdl->norm = b + 4 * NPARAM;
    mdl->loss = b + 4 * NPARAM + NNORM;
    *out = mdl;
    return FT_OK;
}

int ft_mlp_destroy(ft_mlp *mdl) {
    if (!mdl) return FT_OK;
    if (mdl->block) (void)hipFree(mdl->block);
    delete mdl;
    return FT_OK;
}

// params: NPARAM floats; norm: NNORM floats (feature mean[16], 1/std[16], target mean, std).
int ft_mlp_set(ft_mlp *mdl, const float *params, const float *norm, double margin) {
    if (!mdl) return fail(FT_EINVAL, "null model");
    if (params) FT_HIP(hipMemcpy(mdl->prm, params, NPARAM * 4, hipMemcpyHostToDevice));
    if (norm) FT_HIP(hipMemcpy(mdl->norm, norm, NNORM * 4, hipMemcpyHostToDevice));
    mdl->margin = margin;
    if (params) {
        FT_HIP(hipMemset(mdl->m, 0, NPARAM * 4));
        FT_HIP(hipMemset(mdl->v, 0, NPARAM * 4));
        mdl->step = 0;
    }
    if (params || norm) {
        FT_HIP(hipDeviceSynchronize());  // null-stream copies and clears, before stream work
    }
    return FT_OK;
}

int ft_mlp_get(ft_mlp *mdl, float *params, float *norm, double *margin) {
    if (!mdl) return fail(FT_EINVAL, "null model");
    if (params) FT_HIP(hipMemcpy(params, mdl->prm, NPARAM * 4, hipMemcpyDeviceToHost));
    if (norm) FT_HIP(hipMemcpy(norm, mdl->norm, NNORM * 4, hipMemcpyDeviceToHost));
    if (margin) *margin = mdl->margin;
    return FT_OK;
}

int ft_knn_features(ft_stream s, const ft_tree *tree, const int32_t *leaf_ids, int nq, int k,
                    float *features) {
    if (!tree) return fail(FT_EINVAL, "null tree");
    if (nq <= 0 || tree->num_leaves == 0) return FT_OK;
    const hipStream_t hs = as_stream(s);
    FT_DISPATCH_COORD(tree->coord_type, {
        const DTree<T> dt = make_dtree<T>(*tree);
        hipLaunchKernelGGL(features_kernel<T>, dim3((nq + 255) / 256), dim3(256), 0, hs, dt,
                           leaf_ids, nq, k, features);
        FT_CHECK_LAUNCH();
    });
    return FT_OK;
}

// Predicted log d2_k (float[n]) for feature rows (device memory).
int ft_mlp_forward(ft_stream s, ft_mlp *mdl, const float *features, int n, float *pred) {
    if (!mdl) return fail(FT_EINVAL, "null model");
    if (n <= 0) return FT_OK;
    const int blocks = std::min((n + 63) / 64, 2048);
    hipLaunchKernelGGL(mlp_infer_kernel<double>, dim3(blocks), dim3(256), 0, as_stream(s),
                       (const float *)mdl->prm, (const float *)mdl->norm, features, n, pred,
                       (double *)nullptr, 0.0, 0.0);
    FT_CHECK_LAUNCH();
    return FT_OK;
}

// Gradient (fp32, NPARAM) of the mean squared error of the standardised target
// over n samples; loss receives the loss value.  Device pointers.
int ft_mlp_grad(ft_stream s, ft_mlp *mdl, const float *features, const float *target, int n,
                float *grad, float *loss) {
    if (!mdl) return fail(FT_EINVAL, "null model");
    const hipStream_t hs = as_stream(s);
    hipLaunchKernelGGL(zero_f32_kernel, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, hs, grad,
                       NPARAM);
    hipLaunchKernelGGL(zero_f32_kernel, dim3(1), dim3(64), 0, hs, loss, 1);
    if (n > 0) {
        const int blocks = std::min((n + 63) / 64, 1024);
        hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(256), 0, hs,
                           (const float *)mdl->prm, (const float *)mdl->norm, features, target,
                           n, 1.0f / (float)n, grad, loss);
    }
    FT_CHECK_LAUNCH();
    return FT_OK;
}

// One Adam step on n samples.
int ft_mlp_train_step(ft_stream s, ft_mlp *mdl, const float *features, const float *target, int n,
                      float lr) {
    FT_TRY(ft_mlp_grad(s, mdl, features, target, n, mdl->grad, mdl->loss));
    mdl->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)mdl->step));
    const float c2 = 1.0f / (1.0f - std::pow(b2, (float)mdl->step));
    hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, as_stream(s),
                       mdl->prm, (const float *)mdl->grad, mdl->m, mdl->v, lr, b1, b2, 1e-8f, c1,
                       c2);
    FT_CHECK_LAUNCH();
    return FT_OK;
}

int ft_mlp_last_loss(ft_stream s, ft_mlp *mdl, float *loss) {
    if (!mdl) return fail(FT_EINVAL, "null model");
    FT_HIP(hipMemcpyAsync(loss, mdl->loss, 4, hipMemcpyDeviceToHost, as_stream(s)));
    FT_HIP(hipStreamSynchronize(as_stream(s)));
    return FT_OK;
}

// Seed bounds (dist_t[nq], device) for self queries: features + forward +
// exp(prediction + margin).  `scratch_features` holds nq * 16 floats.
int ft_mlp_seed(ft_stream s, ft_mlp *mdl, const ft_tree *tree, const int32_t *leaf_ids, int nq,
                int k, float *scratch_features, void *seed) {
    if (!mdl || !tree) return fail(FT_EINVAL, "null argument");
    if (nq <= 0 || tree->num_leaves == 0) return FT_OK;
    FT_TRY(ft_knn_features(s, tree, leaf_ids, nq, k, scratch_features));
    const hipStream_t hs = as_stream(s);
    const int blocks = std::min((nq + 63) / 64, 2048);
    if (tree->coord_type == FT_COORD_F32 || tree->coord_type == FT_COORD_F64) {
        double eps = 0.0;  // the features' eps is relative; seeds need no offset for FP
        hipLaunchKernelGGL(mlp_infer_kernel<double>, dim3(blocks), dim3(256), 0, hs,
                           (const float *)mdl->prm, (const float *)mdl->norm,
                           (const float *)scratch_features, nq, (float *)nullptr, (double *)seed,
                           mdl->margin, eps);
    } else {
        hipLaunchKernelGGL(mlp_infer_kernel<uint64_t>, dim3(blocks), dim3(256), 0, hs,
                           (const float *)mdl->prm, (const float *)mdl->norm,
                           (const float *)scratch_features, nq, (float *)nullptr,
                           (uint64_t *)seed, mdl->margin, 1.0);
    }
    FT_CHECK_LAUNCH();
    return FT_OK;
}

}  // extern "C"
