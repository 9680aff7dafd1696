// Learned preconditioner network: a 4-level U-Net of 3x3 convolutions run
// and trained with BF16 MFMA (gfx950's v_mfma_f32_16x16x32_bf16, FP32
// accumulate).
// It has no counterpart in the reference.
//
// Network (level 0 = full image, H and W multiples of 8, NHWC activations):
//   e0 = relu(c0(x))             4 -> 32      e1 = relu(c1(pool e0))  32 -> 32
//   e2 = relu(c2(pool e1))      32 -> 64      e3 = relu(c3(pool e2))  64 -> 64
//   d2 = relu(c4(e2 + up e3))   64 -> 32      d1 = relu(c5(e1 + up d2)) 32 -> 32
//   y  =      c6(e0 + up d1)    32 -> 3
// pool = 2x2 mean, up = nearest x2.  Weights W[tap][ci][co] (tap = 3*(dy+1) +
// (dx+1)) are kept in FP32 with Adam moments; two BF16 copies are packed
// after every update: Wf[co][tap*Ci + ci] for the forward convolution and
// Wb[ci][tap'*Co + co] = W[8-tap'][ci][co] for the input gradient (a
// convolution with flipped taps), both as [N rows][K] so one MFMA B fragment
// is one 8-byte load.
//
// Convolution as implicit GEMM: a wavefront owns 64 consecutive pixels of a
// row (M, four 16-pixel tiles) and all output channels (N = 16*NT); K =
// 9*Cin runs in chunks of 32, each lane reading two groups of 4 consecutive
// channels (8 bytes each, NHWC, from the workgroup's LDS tile; each group of
// one tap, as Cin is a multiple of 4).  Lane l holds
// A[l%16][8(l/16)..+7], B[8(l/16)..+7][l%16] and receives
// D[4(l/16)+i][l%16].  The weight gradient is the GEMM over pixels
// dW[k][co] = sum_p X[p + off(k)][ci(k)] dZ[p][co], split over pixel chunks
// whose partial sums are added in a fixed order (no atomics).
#include "mdg_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;

namespace {

constexpr int NL = 7;
constexpr int L_CIN[NL] = {4, 32, 32, 64, 64, 32, 32};
constexpr int L_COUT[NL] = {32, 32, 64, 64, 32, 32, 3};
constexpr int L_LEVEL[NL] = {0, 1, 2, 3, 2, 1, 0};
constexpr int WG_CHUNK = 1024;   // pixels per weight-gradient partial

__host__ __device__ inline int rup(int a, int m) { return (a + m - 1) / m * m; }

__device__ __forceinline__ bf16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);           // round to nearest even
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
// two groups of 4 bf16 bit patterns as one 8-element MFMA operand
__device__ __forceinline__ bf16x8 pack8(bf16x4 lo, bf16x4 hi) {
    typedef short s16x8 __attribute__((ext_vector_type(8)));
    const s16x8 v = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    return __builtin_bit_cast(bf16x8, v);
}

// y = conv3x3(x) + bias (optional ReLU).  x: H*W*Cs_in (bf16 or fp32),
// Wt: (16*NT) x Kp bf16, out: H*W*Cs_out (bf16 or fp32), channels < Cout.
// A workgroup owns CT_H rows of CT_W pixels: it stages the rows and columns
// the 3 x 3 taps reach (one pixel of halo, zero outside the image), all
// input channels, once into LDS as bf16 (pixel stride Cs_in + 4 elements,
// so the 16 pixels of an MFMA fragment fall on distinct banks); each wave
// then computes one row as 4 M tiles of 16 pixels, every B fragment (one
// 16-byte load per lane) feeding 4 MFMAs.
#define CT_H 4
#define CT_W 64
template <int NT, bool IN_F32, bool RELU, bool OUT_F32>
__global__ void __launch_bounds__(256)
k_conv3x3(int H, int W, int Cs_in, int Kp, int Cout, int Cs_out, const void *__restrict__ in,
          const bf16_t *__restrict__ Wt, const float *__restrict__ bias, void *__restrict__ out) {
    extern __shared__ bf16_t tile[];
    const int CS = Cs_in + 4, TW = CT_W + 2;
    const int x0 = blockIdx.x * CT_W, y0 = blockIdx.y * CT_H;
    const int t = threadIdx.x;
    const int cq = Cs_in >> 2;                          // 4-channel groups per pixel
    const int nel = (CT_H + 2) * TW * cq;
    for (int e = t; e < nel; e += 256) {
        const int q = e % cq, pix = e / cq, lx = pix % TW, ly = pix / TW;
        const int gx = x0 + lx - 1, gy = y0 + ly - 1;
        bf16x4 v = {0, 0, 0, 0};
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const long off = ((long)gy * W + gx) * Cs_in + 4 * q;
            if (IN_F32) {
                const float4 f = *(const float4 *)((const float *)in + off);
                v = bf16x4{(short)f2bf(f.x), (short)f2bf(f.y), (short)f2bf(f.z), (short)f2bf(f.w)};
            } else {
                v = *(const bf16x4 *)((const bf16_t *)in + off);
            }
        }
        *(bf16x4 *)(tile + (ly * TW + lx) * CS + 4 * q) = v;
    }
    __syncthreads();
    const int lane = t & 63, wave = t >> 6, m = lane & 15, g = lane >> 4;
    const int K = 9 * Cs_in;
    f32x4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[mt][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < Kp; kc += 32) {
        bf16x8 bw[NT];
#pragma unroll
        for (int n = 0; n < NT; n++) bw[n] = *(const bf16x8 *)(Wt + (long)(n * 16 + m) * Kp + kc + 8 * g);
        int base[2];
        bool kv[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int k = kc + 8 * g + 4 * h;
            kv[h] = k < K;
            const int tap = kv[h] ? k / Cs_in : 0, ci = kv[h] ? k - tap * Cs_in : 0;
            // tap (dy, dx) of the wave's row and of fragment pixel m, tile coordinates
            base[h] = ((wave + tap / 3) * TW + m + tap % 3) * CS + ci;
        }
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            bf16x4 ah[2];
#pragma unroll
            for (int h = 0; h < 2; h++)
                ah[h] = kv[h] ? *(const bf16x4 *)(tile + base[h] + 16 * mt * CS) : bf16x4{0, 0, 0, 0};
            const bf16x8 a = pack8(ah[0], ah[1]);
#pragma unroll
            for (int n = 0; n < NT; n++) acc[mt][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[n], acc[mt][n], 0, 0, 0);
        }
    }
    const int y = y0 + wave;
    if (y >= H) return;
#pragma unroll
    for (int n = 0; n < NT; n++) {
        const int co = n * 16 + m;
        if (co >= Cout) continue;
        const float bn = bias ? bias[co] : 0.f;
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int x = x0 + 16 * mt + 4 * g + i;
                if (x >= W) continue;
                const long q = (long)y * W + x;
                float v = acc[mt][n][i] + bn;
                if (RELU) v = fmaxf(v, 0.f);
                if (OUT_F32) ((float *)out)[q * Cs_out + co] = v;
                else ((bf16_t *)out)[q * Cs_out + co] = f2bf(v);
            }
    }
}

// Partial weight gradients of one pixel chunk: part[chunk][k][co16]
// (k < Kp, co16 < 16*NTn) for one (k tile, n tile) per wavefront.
__global__ void __launch_bounds__(64)
k_conv_wgrad(int H, int W, int Cs_in, int Kp, int Cout, int Cs_g, int NTn,
             const bf16_t *__restrict__ X, const float *__restrict__ G, float *__restrict__ part) {
    const int lane = threadIdx.x, m = lane & 15, g = lane >> 4;
    const int kt = blockIdx.y / NTn, nt = blockIdx.y - kt * NTn;
    const long npix = (long)H * W;
    const long q0 = (long)blockIdx.x * WG_CHUNK;
    const long q1 = (q0 + WG_CHUNK < npix) ? q0 + WG_CHUNK : npix;
    const int k = kt * 16 + m, K = 9 * Cs_in;
    const bool kv = k < K;
    const int tap = kv ? k / Cs_in : 0, ci = kv ? k - tap * Cs_in : 0;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int n = nt * 16 + m;
    const bool nv = n < Cout;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    // the lane's 8 consecutive pixels of each 32-pixel step start at
    // qb + 8g; their (y, x) advance by 32 pixels a step (one division at
    // the start, wrap-arounds after)
    long qs = q0 + 8 * g;
    int ys = (int)(qs / W), xs = (int)(qs - (long)ys * W);
    for (long qb = q0; qb < q1; qb += 32) {
        bf16x4 a4[2], b4[2];
        int y = ys, x = xs;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const long q = qb + 8 * g + j;
            bf16_t av = 0, bv = 0;
            if (q < q1) {
                if (kv) {
                    const int yy = y + dy, xx = x + dx;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) av = X[((long)yy * W + xx) * Cs_in + ci];
                }
                if (nv) bv = f2bf(G[q * Cs_g + n]);
            }
            a4[j >> 2][j & 3] = (short)av;
            b4[j >> 2][j & 3] = (short)bv;
            if (++x == W) { x = 0; y++; }
        }
        xs += 32;
        while (xs >= W) { xs -= W; ys++; }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pack8(a4[0], a4[1]), pack8(b4[0], b4[1]), acc, 0, 0, 0);
    }
    const int ld = 16 * NTn;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int kk = kt * 16 + 4 * g + i;
        part[((long)blockIdx.x * Kp + kk) * ld + n] = acc[i];
    }
}

// dW[tap][ci][co] (+)= sum over chunks.  A 256-thread block owns 16
// consecutive weights; 16 threads per weight sum every 16th chunk and the
// 16 partial sums are combined in LDS in a fixed order.
__global__ void __launch_bounds__(256) k_wgrad_reduce(int nchunk, int Kp, int ld, int Cin, int Cs_in, int Cout,
                                                      const float *__restrict__ part, float *__restrict__ dW,
                                                      int accumulate) {
    __shared__ float sh[16][17];
    const int w = threadIdx.x & 15, j = threadIdx.x >> 4;
    const int i = blockIdx.x * 16 + w, nw = 9 * Cin * Cout;
    float s = 0.f;
    if (i < nw) {
        const int co = i % Cout, ci = (i / Cout) % Cin, tap = i / (Cout * Cin);
        const int k = tap * Cs_in + ci;
        for (int c = j; c < nchunk; c += 16) s += part[((long)c * Kp + k) * ld + co];
    }
    sh[j][w] = s;
    __syncthreads();
    if (j == 0 && i < nw) {
        float r = 0.f;
        for (int q = 0; q < 16; q++) r += sh[q][w];
        dW[i] = accumulate ? dW[i] + r : r;
    }
}
// Bias gradient: per-chunk channel sums, then the sum over chunks.  A
// 256-thread block splits its pixels (or chunks) over 256 / Cout groups of
// Cout channel threads and combines the groups in LDS in a fixed order.
__global__ void __launch_bounds__(256) k_bias_part(long npix, int Cout, int Cs_g, const float *__restrict__ G,
                                                   float *__restrict__ part) {
    __shared__ float sh[256];
    const int t = threadIdx.x, ng = 256 / Cout, co = t % Cout, pg = t / Cout;
    const long q0 = (long)blockIdx.x * WG_CHUNK;
    const long q1 = (q0 + WG_CHUNK < npix) ? q0 + WG_CHUNK : npix;
    float s = 0.f;
    if (pg < ng)
        for (long q = q0 + pg; q < q1; q += ng) s += G[q * Cs_g + co];
    sh[t] = s;
    __syncthreads();
    if (t < Cout) {
        float r = 0.f;
        for (int k = 0; k < ng; k++) r += sh[k * Cout + t];
        part[(long)blockIdx.x * 64 + t] = r;
    }
}
__global__ void __launch_bounds__(256) k_bias_reduce(int nchunk, int Cout, const float *__restrict__ part,
                                                     float *__restrict__ db, int accumulate) {
    __shared__ float sh[256];
    const int t = threadIdx.x, ng = 256 / Cout, co = t % Cout, pg = t / Cout;
    float s = 0.f;
    if (pg < ng)
        for (int c = pg; c < nchunk; c += ng) s += part[(long)c * 64 + co];
    sh[t] = s;
    __syncthreads();
    if (t < Cout) {
        float r = 0.f;
        for (int k = 0; k < ng; k++) r += sh[k * Cout + t];
        db[t] = accumulate ? db[t] + r : r;
    }
}

// 2x2 mean pool (fine H x W -> H/2 x W/2), bf16 in/out, fp32 arithmetic.
__global__ void k_pool(int Hc, int Wc, int C, int Cs, const bf16_t *__restrict__ in, bf16_t *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Hc * Wc * C) return;
    const int c = (int)(i % C);
    const long pc = i / C;
    const int y = (int)(pc / Wc), x = (int)(pc - (long)y * Wc);
    const int W = 2 * Wc;
    const long f = ((long)(2 * y) * W + 2 * x) * Cs + c;
    const float s = (bf2f(in[f]) + bf2f(in[f + Cs])) + (bf2f(in[f + (long)W * Cs]) + bf2f(in[f + (long)W * Cs + Cs]));
    out[pc * Cs + c] = f2bf(s * 0.25f);
}
// s = bf16(fine + up(coarse)) on the fine grid.
__global__ void k_up_add(int H, int W, int C, int Cs, const bf16_t *__restrict__ fine,
                         const bf16_t *__restrict__ coarse, bf16_t *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W * C) return;
    const int c = (int)(i % C);
    const long p = i / C;
    const int y = (int)(p / W), x = (int)(p - (long)y * W);
    const long pc = (long)(y >> 1) * (W >> 1) + (x >> 1);
    out[p * Cs + c] = f2bf(bf2f(fine[p * Cs + c]) + bf2f(coarse[pc * Cs + c]));
}
// out(coarse) = (sum of the 4 fine gradients) * (act > 0)
__global__ void k_up_bwd_mask(int Hc, int Wc, int C, int Cs, const float *__restrict__ g,
                              const bf16_t *__restrict__ act, float *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Hc * Wc * C) return;
    const int c = (int)(i % C);
    const long pc = i / C;
    const int y = (int)(pc / Wc), x = (int)(pc - (long)y * Wc);
    const int W = 2 * Wc;
    const long f = ((long)(2 * y) * W + 2 * x) * Cs + c;
    const float s = (g[f] + g[f + Cs]) + (g[f + (long)W * Cs] + g[f + (long)W * Cs + Cs]);
    out[pc * Cs + c] = bf2f(act[pc * Cs + c]) > 0.f ? s : 0.f;
}
// fine += coarse / 4 replicated (pool backward)
__global__ void k_pool_bwd_acc(int H, int W, int C, int Cs, const float *__restrict__ coarse, float *__restrict__ fine) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W * C) return;
    const int c = (int)(i % C);
    const long p = i / C;
    const int y = (int)(p / W), x = (int)(p - (long)y * W);
    const long pc = (long)(y >> 1) * (W >> 1) + (x >> 1);
    fine[p * Cs + c] += coarse[pc * Cs + c] * 0.25f;
}
__global__ void k_mask_mul(long n, int C, int Cs, float *__restrict__ g, const bf16_t *__restrict__ act) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * C) return;
    const long o = (i / C) * Cs + i % C;
    if (!(bf2f(act[o]) > 0.f)) g[o] = 0.f;
}
// 0.5 * sum(mask (y - t)^2) / sum(mask) partials; dy = mask (y - t) / sum(mask)
__global__ void __launch_bounds__(256)
k_loss(long n4, const float *__restrict__ y, const float *__restrict__ t, const float *__restrict__ mask,
       float inv_cnt, float *__restrict__ dy, double *__restrict__ part) {
    __shared__ double s[256];
    double acc = 0.0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n4; i += (long)gridDim.x * 256) {
        const float d = (i % 4 < 3) ? mask[i] * (y[i] - t[i]) : 0.f;
        dy[i] = d * inv_cnt;
        acc += 0.5 * (double)d * (double)d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void k_sum_partials(int n, const double *__restrict__ part, double scale, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += part[i];
    *out = s * scale;
}

// Adam on one layer's FP32 weights and bias, then repack the BF16 copies.
__global__ void k_adam(int Cin, int Cout, int Cs_in, int Cs_out, int Kpf, int Kpb, float lr, float c1,
                       float c2, float *__restrict__ Wm, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ gW, float *__restrict__ b, float *__restrict__ mb,
                       float *__restrict__ vb, const float *__restrict__ gb, bf16_t *__restrict__ Wf,
                       bf16_t *__restrict__ Wb, int step) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nw = 9 * Cin * Cout;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    if (i < nw) {
        if (step) {
            const float g = gW[i];
            const float mi = b1 * m[i] + (1.f - b1) * g;
            const float vi = b2 * v[i] + (1.f - b2) * g * g;
            m[i] = mi;
            v[i] = vi;
            Wm[i] = Wm[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        }
        const int co = i % Cout, ci = (i / Cout) % Cin, tap = i / (Cout * Cin);
        const bf16_t w = f2bf(Wm[i]);
        Wf[(long)co * Kpf + tap * Cs_in + ci] = w;
        Wb[(long)ci * Kpb + (8 - tap) * Cs_out + co] = w;
    } else if (i < nw + Cout && step) {
        const int j = i - nw;
        const float g = gb[j];
        const float mi = b1 * mb[j] + (1.f - b1) * g;
        const float vi = b2 * vb[j] + (1.f - b2) * g * g;
        mb[j] = mi;
        vb[j] = vi;
        b[j] = b[j] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

struct pc_layer {
    int cin, cout, csi, cso, kpf, kpb, ntf, ntb, level;
    long woff, boff;   // offsets in the flat parameter vector
};

}  // namespace

// Parameters, optimiser state and activation buffers of the network.
#define OUTER_PEND_MAX 256   // outer pairs kept per solve

struct pcnet_state {
    pc_layer L[NL];
    long nparam = 0;
    dbuf<float> P, M, V, G;        // flat params / Adam moments / gradients
    dbuf<bf16_t> Wf, Wb;           // packed bf16 weights (all layers)
    long wf_off[NL], wb_off[NL];
    int step = 0;
    int H = 0, W = 0;              // size the activations are allocated for
    dbuf<bf16_t> act;              // x, e0, p0, e1, p1, e2, p2, e3, s2, d2, s1, d1, s0
    dbuf<float> y, gbuf, tbuf, mbuf, part;
    dbuf<double> lpart, loss, guess_scale;
    long aoff[13];
    long goff[16];
    // captured training pairs (input image, target, mask; H*W*4 floats each)
    int cap_on = 0, cap_max = 0, cap_n = 0, cap_next = 0, cap_pending = 0, cap_H = 0, cap_W = 0;
    dbuf<float> cap_x, cap_t, cap_m;
    dbuf<double> cap_scale;
    // outer-level capture (kind 1): iterates x_k of the running solve, kept
    // until the converged x* gives the targets x* - x_k
    int cap_kind = 0;
    int pend_n = 0;
    int pend_slot[OUTER_PEND_MAX];
    dbuf<double> pend_x, oscale;
    // per outer pair, for the residual-loss training: R_k, eta_n, eta_s, gamma
    dbuf<double> cap_R, cap_en, cap_es, cap_gamma;
    // training work vectors (grid layout)
    dbuf<double> t_zu, t_zp, t_ku, t_kp, t_r, t_g, t_part;
};

namespace {

enum { A_X, A_E0, A_P0, A_E1, A_P1, A_E2, A_P2, A_E3, A_S2, A_D2, A_S1, A_D1, A_S0 };

void layout_layers(pcnet_state *s) {
    long off = 0, of = 0, ob = 0;
    for (int k = 0; k < NL; k++) {
        pc_layer &l = s->L[k];
        l.cin = L_CIN[k];
        l.cout = L_COUT[k];
        l.csi = rup(l.cin, 4);
        l.cso = rup(l.cout, 4);
        l.kpf = rup(9 * l.csi, 32);
        l.kpb = rup(9 * l.cso, 32);
        l.ntf = rup(l.cout, 16) / 16;
        l.ntb = rup(l.cin, 16) / 16;
        l.level = L_LEVEL[k];
        l.woff = off;
        off += 9L * l.cin * l.cout;
        l.boff = off;
        off += l.cout;
        s->wf_off[k] = of;
        of += (long)l.ntf * 16 * l.kpf;
        s->wb_off[k] = ob;
        ob += (long)l.ntb * 16 * l.kpb;
    }
    s->nparam = off;
}

long wf_total(const pcnet_state *s) { return s->wf_off[NL - 1] + (long)s->L[NL - 1].ntf * 16 * s->L[NL - 1].kpf; }
long wb_total(const pcnet_state *s) { return s->wb_off[NL - 1] + (long)s->L[NL - 1].ntb * 16 * s->L[NL - 1].kpb; }

pcnet_state *net_of(mdg_ctx *c) {
    if (!c->pcnet) {
        c->pcnet = new pcnet_state;
        layout_layers(c->pcnet);
    }
    return c->pcnet;
}

int nblk256(long n) { return (int)((n + 255) / 256); }

// Size (in elements) and channel stride of each activation tensor.
struct tdesc { int lev, C; };
const tdesc ADESC[13] = {{0, 4}, {0, 32}, {1, 32}, {1, 32}, {2, 32}, {2, 64}, {3, 64}, {3, 64},
                         {2, 64}, {2, 32}, {1, 32}, {1, 32}, {0, 32}};

int ensure_buffers(mdg_ctx *c, pcnet_state *s, int H, int W) {
    if (H == s->H && W == s->W) return MDG_OK;
    long off = 0;
    for (int a = 0; a < 13; a++) {
        const long npix = (long)(H >> ADESC[a].lev) * (W >> ADESC[a].lev);
        s->aoff[a] = off;
        off += npix * rup(ADESC[a].C, 4);
    }
    HIPCHK(c, s->act.reserve(off));
    HIPCHK(c, hipMemsetAsync(s->act.p, 0, off * sizeof(bf16_t), c->stream));
    // gradient buffers: 0 dy (lev0, 4), 1 ds0/de0 (lev0, 32), 2 dz (lev0 32), 3 de1 (lev1 32),
    // 4 dz1 (lev1 32), 5 de2 (lev2 64), 6 dz2 (lev2 64), 7 dz3 (lev3 64), 8 dp (lev0..2 scratch, 64)
    const int glev[9] = {0, 0, 0, 1, 1, 2, 2, 3, 0}, gc[9] = {4, 32, 32, 32, 32, 64, 64, 64, 64};
    long go = 0;
    for (int i = 0; i < 9; i++) {
        s->goff[i] = go;
        go += (long)(H >> glev[i]) * (W >> glev[i]) * gc[i];
    }
    HIPCHK(c, s->gbuf.reserve(go));
    HIPCHK(c, hipMemsetAsync(s->gbuf.p, 0, go * sizeof(float), c->stream));
    const long n0 = (long)H * W;
    HIPCHK(c, s->y.reserve(n0 * 4));
    HIPCHK(c, hipMemsetAsync(s->y.p, 0, n0 * 4 * sizeof(float), c->stream));
    HIPCHK(c, s->tbuf.reserve(n0 * 4));
    HIPCHK(c, s->mbuf.reserve(n0 * 4));
    long maxpart = 0;
    for (int k = 0; k < NL; k++) {
        const pc_layer &l = s->L[k];
        const long np = (long)(H >> l.level) * (W >> l.level);
        const long nch = (np + WG_CHUNK - 1) / WG_CHUNK;
        const long need = nch * l.kpf * 16 * l.ntf;
        if (need > maxpart) maxpart = need;
        if (nch * 64 > maxpart) maxpart = nch * 64;
    }
    HIPCHK(c, s->part.reserve(maxpart));
    HIPCHK(c, s->lpart.reserve(1024));
    HIPCHK(c, s->loss.reserve(1));
    s->H = H;
    s->W = W;
    return MDG_OK;
}

bf16_t *A(pcnet_state *s, int a) { return s->act.p + s->aoff[a]; }
float *Gb(pcnet_state *s, int i) { return s->gbuf.p + s->goff[i]; }

template <int NT, bool IN_F32, bool RELU, bool OUT_F32>
void conv_launch(mdg_ctx *c, int H, int W, int Cs_in, int Kp, int Cout, int Cs_out, const void *in,
                 const bf16_t *Wt, const float *bias, void *out) {
    const dim3 grid((W + CT_W - 1) / CT_W, (H + CT_H - 1) / CT_H);
    const size_t lds = (size_t)(CT_H + 2) * (CT_W + 2) * (Cs_in + 4) * sizeof(bf16_t);
    hipLaunchKernelGGL((k_conv3x3<NT, IN_F32, RELU, OUT_F32>), grid, dim3(256), lds, c->stream, H, W,
                       Cs_in, Kp, Cout, Cs_out, in, Wt, bias, out);
}

// forward convolution of layer k: bf16 input, bf16 (ReLU) or fp32 output
void conv_fwd(mdg_ctx *c, pcnet_state *s, int k, const bf16_t *in, void *out) {
    const pc_layer &l = s->L[k];
    const int H = s->H >> l.level, W = s->W >> l.level;
    const bf16_t *Wt = s->Wf.p + s->wf_off[k];
    const float *b = s->P.p + l.boff;
    if (k == NL - 1) { conv_launch<1, false, false, true>(c, H, W, l.csi, l.kpf, l.cout, l.cso, in, Wt, b, out); return; }
    switch (l.ntf) {
    case 2: conv_launch<2, false, true, false>(c, H, W, l.csi, l.kpf, l.cout, l.cso, in, Wt, b, out); break;
    default: conv_launch<4, false, true, false>(c, H, W, l.csi, l.kpf, l.cout, l.cso, in, Wt, b, out); break;
    }
}

// input gradient of layer k: dX (fp32, Cs_in) = conv(dZ (fp32, Cs_out), Wb)
void conv_bwd_data(mdg_ctx *c, pcnet_state *s, int k, const float *dz, float *dx) {
    const pc_layer &l = s->L[k];
    const int H = s->H >> l.level, W = s->W >> l.level;
    const bf16_t *Wt = s->Wb.p + s->wb_off[k];
    switch (l.ntb) {
    case 1: conv_launch<1, true, false, true>(c, H, W, l.cso, l.kpb, l.cin, l.csi, dz, Wt, nullptr, dx); break;
    case 2: conv_launch<2, true, false, true>(c, H, W, l.cso, l.kpb, l.cin, l.csi, dz, Wt, nullptr, dx); break;
    default: conv_launch<4, true, false, true>(c, H, W, l.cso, l.kpb, l.cin, l.csi, dz, Wt, nullptr, dx); break;
    }
}

void conv_bwd_weights(mdg_ctx *c, pcnet_state *s, int k, const bf16_t *x, const float *dz, int accumulate) {
    const pc_layer &l = s->L[k];
    const int H = s->H >> l.level, W = s->W >> l.level;
    const long npix = (long)H * W;
    const int nch = (int)((npix + WG_CHUNK - 1) / WG_CHUNK);
    hipLaunchKernelGGL(k_conv_wgrad, dim3(nch, (l.kpf / 16) * l.ntf), dim3(64), 0, c->stream, H, W, l.csi,
                       l.kpf, l.cout, l.cso, l.ntf, x, dz, s->part.p);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((int)((9L * l.cin * l.cout + 15) / 16)), dim3(256), 0, c->stream, nch,
                       l.kpf, 16 * l.ntf, l.cin, l.csi, l.cout, (const float *)s->part.p,
                       s->G.p + l.woff, accumulate);
    hipLaunchKernelGGL(k_bias_part, dim3(nch), dim3(256), 0, c->stream, npix, l.cout, l.cso, dz, s->part.p);
    hipLaunchKernelGGL(k_bias_reduce, dim3(1), dim3(256), 0, c->stream, nch, l.cout, (const float *)s->part.p,
                       s->G.p + l.boff, accumulate);
}

void pool_launch(mdg_ctx *c, int Hf, int Wf, int C, const bf16_t *in, bf16_t *out) {
    const int Hc = Hf / 2, Wc = Wf / 2;
    hipLaunchKernelGGL(k_pool, dim3(nblk256((long)Hc * Wc * C)), dim3(256), 0, c->stream, Hc, Wc, C, rup(C, 4), in, out);
}
void up_add_launch(mdg_ctx *c, int H, int W, int C, const bf16_t *fine, const bf16_t *coarse, bf16_t *out) {
    hipLaunchKernelGGL(k_up_add, dim3(nblk256((long)H * W * C)), dim3(256), 0, c->stream, H, W, C, rup(C, 4), fine, coarse, out);
}

int forward(mdg_ctx *c, pcnet_state *s) {
    const int H = s->H, W = s->W;
    conv_fwd(c, s, 0, A(s, A_X), A(s, A_E0));
    pool_launch(c, H, W, 32, A(s, A_E0), A(s, A_P0));
    conv_fwd(c, s, 1, A(s, A_P0), A(s, A_E1));
    pool_launch(c, H / 2, W / 2, 32, A(s, A_E1), A(s, A_P1));
    conv_fwd(c, s, 2, A(s, A_P1), A(s, A_E2));
    pool_launch(c, H / 4, W / 4, 64, A(s, A_E2), A(s, A_P2));
    conv_fwd(c, s, 3, A(s, A_P2), A(s, A_E3));
    up_add_launch(c, H / 4, W / 4, 64, A(s, A_E2), A(s, A_E3), A(s, A_S2));
    conv_fwd(c, s, 4, A(s, A_S2), A(s, A_D2));
    up_add_launch(c, H / 2, W / 2, 32, A(s, A_E1), A(s, A_D2), A(s, A_S1));
    conv_fwd(c, s, 5, A(s, A_S1), A(s, A_D1));
    up_add_launch(c, H, W, 32, A(s, A_E0), A(s, A_D1), A(s, A_S0));
    conv_fwd(c, s, 6, A(s, A_S0), s->y.p);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

int backward(mdg_ctx *c, pcnet_state *s, int accumulate) {
    const int H = s->H, W = s->W;
    float *dy = Gb(s, 0), *de0 = Gb(s, 1), *dz0 = Gb(s, 2), *de1 = Gb(s, 3), *dz1 = Gb(s, 4);
    float *de2 = Gb(s, 5), *dz2 = Gb(s, 6), *dz3 = Gb(s, 7), *dp = Gb(s, 8);
    // layer 6 (level 0): ds0 -> de0
    conv_bwd_weights(c, s, 6, A(s, A_S0), dy, accumulate);
    conv_bwd_data(c, s, 6, dy, de0);
    // layer 5 (level 1): dz5 = up_bwd(ds0) * (d1 > 0)
    hipLaunchKernelGGL(k_up_bwd_mask, dim3(nblk256((long)(H / 2) * (W / 2) * 32)), dim3(256), 0, c->stream, H / 2, W / 2, 32, 32, (const float *)de0, (const bf16_t *)A(s, A_D1), dz1);
    conv_bwd_weights(c, s, 5, A(s, A_S1), dz1, accumulate);
    conv_bwd_data(c, s, 5, dz1, de1);                      // ds1 -> de1
    // layer 4 (level 2)
    hipLaunchKernelGGL(k_up_bwd_mask, dim3(nblk256((long)(H / 4) * (W / 4) * 32)), dim3(256), 0, c->stream, H / 4, W / 4, 32, 32, (const float *)de1, (const bf16_t *)A(s, A_D2), dz2);
    conv_bwd_weights(c, s, 4, A(s, A_S2), dz2, accumulate);
    conv_bwd_data(c, s, 4, dz2, de2);                      // ds2 -> de2
    // layer 3 (level 3)
    hipLaunchKernelGGL(k_up_bwd_mask, dim3(nblk256((long)(H / 8) * (W / 8) * 64)), dim3(256), 0, c->stream, H / 8, W / 8, 64, 64, (const float *)de2, (const bf16_t *)A(s, A_E3), dz3);
    conv_bwd_weights(c, s, 3, A(s, A_P2), dz3, accumulate);
    conv_bwd_data(c, s, 3, dz3, dp);                       // dp2 (level 3, 64)
    hipLaunchKernelGGL(k_pool_bwd_acc, dim3(nblk256((long)(H / 4) * (W / 4) * 64)), dim3(256), 0, c->stream, H / 4, W / 4, 64, 64, (const float *)dp, de2);
    // layer 2 (level 2)
    hipLaunchKernelGGL(k_mask_mul, dim3(nblk256((long)(H / 4) * (W / 4) * 64)), dim3(256), 0, c->stream, (long)(H / 4) * (W / 4), 64, 64, de2, (const bf16_t *)A(s, A_E2));
    conv_bwd_weights(c, s, 2, A(s, A_P1), de2, accumulate);
    conv_bwd_data(c, s, 2, de2, dp);                       // dp1 (level 2, 32)
    hipLaunchKernelGGL(k_pool_bwd_acc, dim3(nblk256((long)(H / 2) * (W / 2) * 32)), dim3(256), 0, c->stream, H / 2, W / 2, 32, 32, (const float *)dp, de1);
    // layer 1 (level 1)
    hipLaunchKernelGGL(k_mask_mul, dim3(nblk256((long)(H / 2) * (W / 2) * 32)), dim3(256), 0, c->stream, (long)(H / 2) * (W / 2), 32, 32, de1, (const bf16_t *)A(s, A_E1));
    conv_bwd_weights(c, s, 1, A(s, A_P0), de1, accumulate);
    conv_bwd_data(c, s, 1, de1, dp);                       // dp0 (level 1, 32)
    hipLaunchKernelGGL(k_pool_bwd_acc, dim3(nblk256((long)H * W * 32)), dim3(256), 0, c->stream, H, W, 32, 32, (const float *)dp, de0);
    // layer 0 (level 0)
    hipLaunchKernelGGL(k_mask_mul, dim3(nblk256((long)H * W * 32)), dim3(256), 0, c->stream, (long)H * W, 32, 32, de0, (const bf16_t *)A(s, A_E0));
    conv_bwd_weights(c, s, 0, A(s, A_X), de0, accumulate);
    (void)dz0;
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

int repack(mdg_ctx *c, pcnet_state *s, int step, float lr) {
    const float c1 = 1.f - powf(0.9f, (float)s->step), c2 = 1.f - powf(0.999f, (float)s->step);
    for (int k = 0; k < NL; k++) {
        const pc_layer &l = s->L[k];
        const long n = 9L * l.cin * l.cout + l.cout;
        hipLaunchKernelGGL(k_adam, dim3(nblk256(n)), dim3(256), 0, c->stream, l.cin, l.cout, l.csi, l.cso, l.kpf,
                           l.kpb, lr, c1, c2, s->P.p + l.woff, s->M.p + l.woff, s->V.p + l.woff,
                           (const float *)s->G.p + l.woff, s->P.p + l.boff, s->M.p + l.boff, s->V.p + l.boff,
                           (const float *)s->G.p + l.boff, s->Wf.p + s->wf_off[k], s->Wb.p + s->wb_off[k], step);
    }
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

// Host image (fp32, H*W*4) into the bf16 input tensor.
__global__ void k_to_bf16(long n, const float *__restrict__ x, bf16_t *__restrict__ y) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

int check_dims(mdg_ctx *c, int H, int W) {
    if (H < 8 || W < 8 || H % 8 || W % 8 || (long)H * W > (1L << 26))
        return mdg_fail(c, MDG_ERR_ARG, "image %d x %d: H and W must be multiples of 8 (>= 8)", H, W);
    return MDG_OK;
}

int stage_image(mdg_ctx *c, pcnet_state *s, const float *x) {
    const long n = (long)s->H * s->W * 4;
    const float *xd = x;
    if (!mdg_is_device_ptr(x)) {
        HIPCHK(c, hipMemcpyAsync(s->tbuf.p, x, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
        xd = s->tbuf.p;
    }
    hipLaunchKernelGGL(k_to_bf16, dim3(nblk256(n)), dim3(256), 0, c->stream, n, xd, A(s, A_X));
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

}  // namespace

void mdg_pcnet_release(mdg_ctx *c) {
    pcnet_state *s = c->pcnet;
    if (!s) return;
    s->P.release(); s->M.release(); s->V.release(); s->G.release(); s->Wf.release(); s->Wb.release();
    s->act.release(); s->y.release(); s->gbuf.release(); s->tbuf.release(); s->mbuf.release();
    s->part.release(); s->lpart.release(); s->loss.release(); s->guess_scale.release();
    s->cap_x.release(); s->cap_t.release(); s->cap_m.release(); s->cap_scale.release();
    s->pend_x.release(); s->oscale.release();
    s->cap_R.release(); s->cap_en.release(); s->cap_es.release(); s->cap_gamma.release();
    s->t_zu.release(); s->t_zp.release(); s->t_ku.release(); s->t_kp.release(); s->t_r.release(); s->t_g.release();
    s->t_part.release();
    delete s;
    c->pcnet = nullptr;
}

// internal entry points used by the solver
int mdg_pcnet_ready(mdg_ctx *c) { return c->pcnet && c->pcnet->P.p ? 1 : 0; }
int mdg_pcnet_forward_dev(mdg_ctx *c, int H, int W, const bf16_t *xin_dev, float **y_dev) {
    pcnet_state *s = net_of(c);
    int rc;
    if ((rc = check_dims(c, H, W)) || (rc = ensure_buffers(c, s, H, W))) return rc;
    if (xin_dev)
        HIPCHK(c, hipMemcpyAsync(A(s, A_X), xin_dev, (long)H * W * 4 * sizeof(bf16_t), hipMemcpyDeviceToDevice, c->stream));
    if ((rc = forward(c, s))) return rc;
    *y_dev = s->y.p;
    return MDG_OK;
}
unsigned short *mdg_pcnet_input_dev(mdg_ctx *c, int H, int W) {
    pcnet_state *s = net_of(c);
    if (check_dims(c, H, W) || ensure_buffers(c, s, H, W)) return nullptr;
    return A(s, A_X);
}

extern "C" int mdg_pcnet_param_count(mdg_ctx *c, long long *n) {
    if (!c || !n) return MDG_ERR_ARG;
    *n = net_of(c)->nparam;
    return MDG_OK;
}

extern "C" int mdg_pcnet_set_params(mdg_ctx *c, const float *params, long long n) {
    if (!c || !params) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    pcnet_state *s = net_of(c);
    if (n != s->nparam) return mdg_fail(c, MDG_ERR_ARG, "expected %ld parameters, got %lld", s->nparam, n);
    HIPCHK(c, s->P.reserve(n));
    HIPCHK(c, s->M.reserve(n));
    HIPCHK(c, s->V.reserve(n));
    HIPCHK(c, s->G.reserve(n));
    HIPCHK(c, s->Wf.reserve(wf_total(s)));
    HIPCHK(c, s->Wb.reserve(wb_total(s)));
    HIPCHK(c, hipMemcpyAsync(s->P.p, params, n * sizeof(float), hipMemcpyDefault, c->stream));
    HIPCHK(c, hipMemsetAsync(s->M.p, 0, n * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(s->V.p, 0, n * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(s->G.p, 0, n * sizeof(float), c->stream));
    HIPCHK(c, hipMemsetAsync(s->Wf.p, 0, wf_total(s) * sizeof(bf16_t), c->stream));
    HIPCHK(c, hipMemsetAsync(s->Wb.p, 0, wb_total(s) * sizeof(bf16_t), c->stream));
    s->step = 0;
    int rc = repack(c, s, 0, 0.f);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_pcnet_get_params(mdg_ctx *c, float *params, float *grads) {
    if (!c) return MDG_ERR_ARG;
    pcnet_state *s = net_of(c);
    if (!s->P.p) return mdg_fail(c, MDG_ERR_STATE, "network parameters not set");
    if (params) HIPCHK(c, hipMemcpyAsync(params, s->P.p, s->nparam * sizeof(float), hipMemcpyDefault, c->stream));
    if (grads) HIPCHK(c, hipMemcpyAsync(grads, s->G.p, s->nparam * sizeof(float), hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_pcnet_forward(mdg_ctx *c, int H, int W, const float *x, float *y) {
    if (!c || !x || !y) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    pcnet_state *s = net_of(c);
    if (!s->P.p) return mdg_fail(c, MDG_ERR_STATE, "network parameters not set");
    int rc;
    if ((rc = check_dims(c, H, W)) || (rc = ensure_buffers(c, s, H, W)) || (rc = stage_image(c, s, x)) ||
        (rc = forward(c, s)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(y, s->y.p, (long)H * W * 4 * sizeof(float), hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_pcnet_grad(mdg_ctx *c, int H, int W, const float *x, const float *t,
                              const float *mask, int accumulate, double *loss) {
    if (!c || !x || !t || !mask) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    pcnet_state *s = net_of(c);
    if (!s->P.p) return mdg_fail(c, MDG_ERR_STATE, "network parameters not set");
    int rc;
    if ((rc = check_dims(c, H, W)) || (rc = ensure_buffers(c, s, H, W)) || (rc = stage_image(c, s, x)) ||
        (rc = forward(c, s)))
        return rc;
    const long n4 = (long)H * W * 4;
    const float *td = t, *md = mask;
    if (!mdg_is_device_ptr(t)) {
        HIPCHK(c, hipMemcpyAsync(s->tbuf.p, t, n4 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        td = s->tbuf.p;
    }
    if (!mdg_is_device_ptr(mask)) {
        HIPCHK(c, hipMemcpyAsync(s->mbuf.p, mask, n4 * sizeof(float), hipMemcpyHostToDevice, c->stream));
        md = s->mbuf.p;
    }
    // the loss normaliser (number of unmasked output values) is counted on the host side copy
    std::vector<float> hm(n4);
    HIPCHK(c, hipMemcpyAsync(hm.data(), md, n4 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double cnt = 0.0;
    for (long i = 0; i < n4; i++)
        if (i % 4 < 3) cnt += hm[i];
    if (cnt < 1.0) cnt = 1.0;
    const int nb = 256;
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, c->stream, n4, (const float *)s->y.p, td, md,
                       (float)(1.0 / cnt), Gb(s, 0), s->lpart.p);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(1), 0, c->stream, nb, (const double *)s->lpart.p, 1.0 / cnt, s->loss.p);
    HIPCHK(c, hipGetLastError());
    if ((rc = backward(c, s, accumulate))) return rc;
    if (loss) HIPCHK(c, hipMemcpyAsync(loss, s->loss.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_pcnet_adam(mdg_ctx *c, float lr) {
    if (!c) return MDG_ERR_ARG;
    pcnet_state *s = net_of(c);
    if (!s->P.p) return mdg_fail(c, MDG_ERR_STATE, "network parameters not set");
    s->step++;
    int rc = repack(c, s, 1, lr);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

// ------------------------------------------------------------------ solver hooks
#include "pcnet_internal.h"

namespace {

// sum over free velocity faces of (dinv*b)^2 -> part[block]
template <typename V, typename T>
__global__ void __launch_bounds__(256)
k_scaled_sumsq(sgrid g, const V *__restrict__ b, const T *__restrict__ dinv, double *__restrict__ part) {
    __shared__ double sh[256];
    const int nu = nvx(g), n = nu + nvz(g);
    double acc = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const double v = (dinv ? (double)dinv[i] : 1.0) * (double)b[i];
        acc += v * v;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
// scale[0] = sqrt(sum / nfree) (1 when zero); nfree = free velocity faces
__global__ void k_scale_from(int nb, const double *__restrict__ part, double nfree, double *__restrict__ scale) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += part[i];
    s = sqrt(s / nfree);
    scale[0] = (s > 0.0) ? s : 1.0;
}

// Input image (H x W x 4) of cell (x = k, y = l): log10(eta_n)/3, then
// dinv*b/s on the cell's W face Vx(k, l+1) and S face Vz(k+1, l) where free.
template <typename V, typename T, bool F32>
__global__ void k_pack_input(sgrid g, int H, int W, const V *__restrict__ b, const T *__restrict__ dinv,
                             const T *__restrict__ en, const double *__restrict__ scale, void *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    const int Ncx = g.Nx - 1, Ncz = g.Nz - 1, nu = nvx(g);
    float v0 = 0.f, v1 = 0.f, v2 = 0.f;
    if (x < Ncx && y < Ncz) {
        const double is = 1.0 / scale[0];
        v0 = (float)(log10((double)en[x + y * Ncx]) / 3.0);
        const int iu = x + (y + 1) * g.Nx;
        if (vx_free(g, x, y + 1)) v1 = (float)((double)dinv[iu] * (double)b[iu] * is);
        const int iv = nu + (x + 1) + y * (g.Nx + 1);
        if (vz_free(g, x + 1, y)) v2 = (float)((double)dinv[iv] * (double)b[iv] * is);
    }
    if (F32) {
        float *o = (float *)out + i * 4;
        o[0] = v0; o[1] = v1; o[2] = v2; o[3] = 0.f;
    } else {
        bf16_t *o = (bf16_t *)out + i * 4;
        o[0] = f2bf(v0); o[1] = f2bf(v1); o[2] = f2bf(v2); o[3] = 0;
    }
}
// x0 on the free faces from the network output (zero elsewhere)
template <typename T>
__global__ void k_unpack_guess(sgrid g, int H, int W, const float *__restrict__ y, const double *__restrict__ scale,
                               T *__restrict__ x0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), n = nu + nvz(g);
    if (i >= n) return;
    const double s = scale[0];
    T v = T(0);
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;    // W face of cell (k, l-1)
        if (vx_free(g, k, l) && l - 1 < H && k < W) v = (T)(s * y[((long)(l - 1) * W + k) * 4 + 0]);
    } else {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);   // S face of cell (k-1, l)
        if (vz_free(g, k, l) && l < H && k - 1 < W) v = (T)(s * y[((long)l * W + (k - 1)) * 4 + 1]);
    }
    x0[i] = v;
}
// Target image and mask of a captured pair: x/s on the W and S faces.
template <typename T>
__global__ void k_pack_target(sgrid g, int H, int W, const T *__restrict__ xs, const double *__restrict__ scale,
                              float *__restrict__ t, float *__restrict__ m) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    const int Ncx = g.Nx - 1, Ncz = g.Nz - 1, nu = nvx(g);
    float t0 = 0.f, t1 = 0.f, m0 = 0.f, m1 = 0.f;
    if (x < Ncx && y < Ncz) {
        const double is = 1.0 / scale[0];
        const int iu = x + (y + 1) * g.Nx;
        if (vx_free(g, x, y + 1)) { t0 = (float)((double)xs[iu] * is); m0 = 1.f; }
        const int iv = nu + (x + 1) + y * (g.Nx + 1);
        if (vz_free(g, x + 1, y)) { t1 = (float)((double)xs[iv] * is); m1 = 1.f; }
    }
    float *to = t + i * 4, *mo = m + i * 4;
    to[0] = t0; to[1] = t1; to[2] = 0.f; to[3] = 0.f;
    mo[0] = m0; mo[1] = m1; mo[2] = 0.f; mo[3] = 0.f;
}

int image_dims(const sgrid &g, int *H, int *W) {
    *W = rup(g.Nx - 1, 8);
    *H = rup(g.Nz - 1, 8);
    return MDG_OK;
}

double free_faces(const sgrid &g) {
    return (double)(g.Nx - 2) * (g.Nz - 1) + (double)(g.Nx - 1) * (g.Nz - 2);
}

template <typename V, typename T>
int compute_scale(mdg_ctx *c, pcnet_state *s, const sgrid &g, const V *b, const T *dinv, double *scale) {
    const int nb = 256;
    HIPCHK(c, s->lpart.reserve(1024));
    hipLaunchKernelGGL((k_scaled_sumsq<V, T>), dim3(nb), dim3(256), 0, c->stream, g, b, dinv, s->lpart.p);
    hipLaunchKernelGGL(k_scale_from, dim3(1), dim3(1), 0, c->stream, nb, (const double *)s->lpart.p, free_faces(g), scale);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

}  // namespace

template <typename V, typename T>
int mdg_pcnet_guess(mdg_ctx *c, const sgrid &g, const V *b, const T *dinv, const T *en, V *x0) {
    pcnet_state *s = net_of(c);
    int H, W, rc;
    image_dims(g, &H, &W);
    if ((rc = check_dims(c, H, W)) || (rc = ensure_buffers(c, s, H, W))) return rc;
    HIPCHK(c, s->guess_scale.reserve(1));
    if ((rc = compute_scale<V, T>(c, s, g, b, dinv, s->guess_scale.p))) return rc;
    hipLaunchKernelGGL((k_pack_input<V, T, false>), dim3(nblk256((long)H * W)), dim3(256), 0, c->stream, g, H, W, b, dinv,
                       en, (const double *)s->guess_scale.p, (void *)A(s, A_X));
    if ((rc = forward(c, s))) return rc;
    hipLaunchKernelGGL(k_unpack_guess<V>, dim3(nblk256(nvx(g) + nvz(g))), dim3(256), 0, c->stream, g, H, W,
                       (const float *)s->y.p, (const double *)s->guess_scale.p, x0);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

template <typename V, typename T>
int mdg_pcnet_capture_in(mdg_ctx *c, const sgrid &g, const V *b, const T *dinv, const T *en) {
    pcnet_state *s = c->pcnet;
    if (!s || !s->cap_on || s->cap_kind != 0 || s->cap_pending) return MDG_OK;
    int H, W;
    image_dims(g, &H, &W);
    if (s->cap_n > 0 && (H != s->cap_H || W != s->cap_W))
        return mdg_fail(c, MDG_ERR_ARG, "captured pairs must share one grid size (%dx%d vs %dx%d)", H, W, s->cap_H, s->cap_W);
    const long img = (long)H * W * 4;
    if (s->cap_n == 0) {
        s->cap_H = H;
        s->cap_W = W;
        HIPCHK(c, s->cap_x.reserve(img * s->cap_max));
        HIPCHK(c, s->cap_t.reserve(img * s->cap_max));
        HIPCHK(c, s->cap_m.reserve(img * s->cap_max));
        HIPCHK(c, s->cap_scale.reserve(s->cap_max));
    }
    const int slot = s->cap_next;
    int rc;
    if ((rc = compute_scale<V, T>(c, s, g, b, dinv, s->cap_scale.p + slot))) return rc;
    hipLaunchKernelGGL((k_pack_input<V, T, true>), dim3(nblk256((long)H * W)), dim3(256), 0, c->stream, g, H, W, b, dinv,
                       en, (const double *)(s->cap_scale.p + slot), (void *)(s->cap_x.p + img * slot));
    HIPCHK(c, hipGetLastError());
    s->cap_pending = 1;
    return MDG_OK;
}

template <typename T>
int mdg_pcnet_capture_out(mdg_ctx *c, const sgrid &g, const T *x) {
    pcnet_state *s = c->pcnet;
    if (!s || !s->cap_on || s->cap_kind != 0 || !s->cap_pending) return MDG_OK;
    const int H = s->cap_H, W = s->cap_W, slot = s->cap_next;
    const long img = (long)H * W * 4;
    // the target is the solution's direction: x scaled to unit RMS over the
    // free faces (the solver rescales the network's guess optimally)
    HIPCHK(c, s->guess_scale.reserve(1));
    int rc = compute_scale<T, T>(c, s, g, x, (const T *)nullptr, s->guess_scale.p);
    if (rc) return rc;
    hipLaunchKernelGGL(k_pack_target<T>, dim3(nblk256((long)H * W)), dim3(256), 0, c->stream, g, H, W, x,
                       (const double *)s->guess_scale.p, s->cap_t.p + img * slot, s->cap_m.p + img * slot);
    HIPCHK(c, hipGetLastError());
    s->cap_pending = 0;
    s->cap_next = (slot + 1) % s->cap_max;
    if (s->cap_n < s->cap_max) s->cap_n++;
    return MDG_OK;
}

template int mdg_pcnet_guess<double, double>(mdg_ctx *, const sgrid &, const double *, const double *, const double *,
                                             double *);
template int mdg_pcnet_guess<double, float>(mdg_ctx *, const sgrid &, const double *, const float *, const float *,
                                            double *);
template int mdg_pcnet_capture_in<double, double>(mdg_ctx *, const sgrid &, const double *, const double *,
                                                  const double *);
template int mdg_pcnet_capture_in<double, float>(mdg_ctx *, const sgrid &, const double *, const float *, const float *);
template int mdg_pcnet_capture_out<double>(mdg_ctx *, const sgrid &, const double *);


// ------------------------------------------------- outer (FGCR) directions
namespace {

// Sums of squares of the velocity part (free faces, optionally times dinv)
// and of the pressure part (cells, optionally times pfac * eta_n) of a
// grid-layout vector: part[b] and part[nb + b].
template <typename T>
__global__ void __launch_bounds__(256)
k_sumsq2(sgrid g, const double *__restrict__ v, const T *__restrict__ dinv, const double *__restrict__ en,
         double pfac, double *__restrict__ part) {
    __shared__ double sh[2][256];
    const int nu = nvx(g), nvel = nu + nvz(g), n = nvel + ncell(g);
    double au = 0.0, ap = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (i < nvel) {
            const double a = (dinv ? (double)dinv[i] : 1.0) * v[i];
            au += a * a;
        } else {
            const double a = (en ? pfac * en[i - nvel] : 1.0) * v[i];
            ap += a * a;
        }
    }
    sh[0][threadIdx.x] = au;
    sh[1][threadIdx.x] = ap;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sh[0][0];
        part[gridDim.x + blockIdx.x] = sh[1][0];
    }
}
// scale[0] = RMS of the velocity part over the free faces, scale[1] = RMS of
// the pressure part over the cells (1 where zero)
__global__ void k_scale2_from(int nb, const double *__restrict__ part, double nfree, double ncells,
                              double *__restrict__ scale) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nb; i++) { a += part[i]; b += part[nb + i]; }
    a = sqrt(a / nfree);
    b = sqrt(b / ncells);
    scale[0] = (a > 0.0) ? a : 1.0;
    scale[1] = (b > 0.0) ? b : 1.0;
}

// Input image of cell (x, y) from the augmented residual R: log10(eta_n)/3;
// dinv*R_u on the W face and dinv*R_v on the S face over scale[0] (the
// Jacobi-preconditioned momentum residual); pfac*eta_n*R_p over scale[1]
// (the Schur-approximated continuity residual).
template <typename T, bool F32>
__global__ void k_pack_outer(sgrid g, int H, int W, double pfac, const double *__restrict__ R,
                             const T *__restrict__ dinv, const double *__restrict__ en,
                             const double *__restrict__ scale, void *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    const int Ncx = g.Nx - 1, Ncz = g.Nz - 1, nu = nvx(g), nvel = nu + nvz(g);
    float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f;
    if (x < Ncx && y < Ncz) {
        const int cc = x + y * Ncx;
        const double iu = 1.0 / scale[0];
        v0 = (float)(log10(en[cc]) / 3.0);
        const int ju = x + (y + 1) * g.Nx;
        if (vx_free(g, x, y + 1)) v1 = (float)((double)dinv[ju] * R[ju] * iu);
        const int jv = nu + (x + 1) + y * (g.Nx + 1);
        if (vz_free(g, x + 1, y)) v2 = (float)((double)dinv[jv] * R[jv] * iu);
        v3 = (float)(pfac * en[cc] * R[nvel + cc] / scale[1]);
    }
    if (F32) {
        float *o = (float *)out + i * 4;
        o[0] = v0; o[1] = v1; o[2] = v2; o[3] = v3;
    } else {
        bf16_t *o = (bf16_t *)out + i * 4;
        o[0] = f2bf(v0); o[1] = f2bf(v1); o[2] = f2bf(v2); o[3] = f2bf(v3);
    }
}
// Target image and mask of an outer pair from the error e = x* - x_k:
// e_u (W face), e_v (S face) over scale[0], e_p over scale[1].
__global__ void k_pack_outer_target(sgrid g, int H, int W, const double *__restrict__ e,
                                    const double *__restrict__ scale, float *__restrict__ t,
                                    float *__restrict__ m) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    const int Ncx = g.Nx - 1, Ncz = g.Nz - 1, nu = nvx(g), nvel = nu + nvz(g);
    float t0 = 0.f, t1 = 0.f, t2 = 0.f, m0 = 0.f, m1 = 0.f, m2 = 0.f;
    if (x < Ncx && y < Ncz) {
        const int ju = x + (y + 1) * g.Nx;
        if (vx_free(g, x, y + 1)) { t0 = (float)(e[ju] / scale[0]); m0 = 1.f; }
        const int jv = nu + (x + 1) + y * (g.Nx + 1);
        if (vz_free(g, x + 1, y)) { t1 = (float)(e[jv] / scale[0]); m1 = 1.f; }
        t2 = (float)(e[nvel + x + y * Ncx] / scale[1]);
        m2 = 1.f;
    }
    float *to = t + i * 4, *mo = m + i * 4;
    to[0] = t0; to[1] = t1; to[2] = t2; to[3] = 0.f;
    mo[0] = m0; mo[1] = m1; mo[2] = m2; mo[3] = 0.f;
}
// Network output -> two grid-layout directions: zu (velocity channels on
// the free faces, zero pressure) and zp (pressure channel, zero velocity).
__global__ void k_unpack_outer(sgrid g, int H, int W, const float *__restrict__ y, double *__restrict__ zu,
                               double *__restrict__ zp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nvel = nu + nvz(g), n = nvel + ncell(g), Ncx = g.Nx - 1;
    if (i >= n) return;
    double a = 0.0, b = 0.0;
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        if (vx_free(g, k, l) && l - 1 < H && k < W) a = y[((long)(l - 1) * W + k) * 4 + 0];
    } else if (i < nvel) {
        const int j = i - nu, k = j % (g.Nx + 1), l = j / (g.Nx + 1);
        if (vz_free(g, k, l) && l < H && k - 1 < W) a = y[((long)l * W + (k - 1)) * 4 + 1];
    } else {
        const int j = i - nvel, x = j % Ncx, yy = j / Ncx;
        b = y[((long)yy * W + x) * 4 + 2];
    }
    zu[i] = a;
    zp[i] = b;
}
__global__ void k_err(long n, const double *__restrict__ xs, double *__restrict__ xk) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xk[i] = xs[i] - xk[i];
}

template <typename T>
int outer_scales(mdg_ctx *c, pcnet_state *s, const sgrid &g, const double *v, const T *dinv, const double *en,
                 double pfac, double *scale) {
    const int nb = 256;
    HIPCHK(c, s->lpart.reserve(2 * nb));
    hipLaunchKernelGGL(k_sumsq2<T>, dim3(nb), dim3(256), 0, c->stream, g, v, dinv, en, pfac, s->lpart.p);
    hipLaunchKernelGGL(k_scale2_from, dim3(1), dim3(1), 0, c->stream, nb, (const double *)s->lpart.p, free_faces(g),
                       (double)ncell(g), scale);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

}  // namespace

int mdg_pcnet_outer_capturing(mdg_ctx *c) {
    return c->pcnet && c->pcnet->cap_on && c->pcnet->cap_kind == 1;
}

template <typename T>
int mdg_pcnet_outer_dirs(mdg_ctx *c, const sgrid &g, double gamma, const double *R, const T *dinv,
                         const double *en, double *zu, double *zp) {
    pcnet_state *s = net_of(c);
    int H, W, rc;
    image_dims(g, &H, &W);
    if ((rc = check_dims(c, H, W)) || (rc = ensure_buffers(c, s, H, W))) return rc;
    HIPCHK(c, s->oscale.reserve(2));
    const double pfac = (1.0 + gamma) / (g.dx * g.dz);
    if ((rc = outer_scales<T>(c, s, g, R, dinv, en, pfac, s->oscale.p))) return rc;
    hipLaunchKernelGGL((k_pack_outer<T, false>), dim3(nblk256((long)H * W)), dim3(256), 0, c->stream, g, H, W, pfac, R,
                       dinv, en, (const double *)s->oscale.p, (void *)A(s, A_X));
    if ((rc = forward(c, s))) return rc;
    const int n = nvx(g) + nvz(g) + ncell(g);
    hipLaunchKernelGGL(k_unpack_outer, dim3(nblk256(n)), dim3(256), 0, c->stream, g, H, W, (const float *)s->y.p, zu, zp);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

template <typename T>
int mdg_pcnet_outer_capture_in(mdg_ctx *c, const sgrid &g, double gamma, const double *R, const T *dinv,
                               const double *en, const double *x) {
    pcnet_state *s = c->pcnet;
    if (!s || !s->cap_on || s->cap_kind != 1 || s->pend_n >= OUTER_PEND_MAX) return MDG_OK;
    int H, W;
    image_dims(g, &H, &W);
    if (s->cap_n > 0 && (H != s->cap_H || W != s->cap_W))
        return mdg_fail(c, MDG_ERR_ARG, "captured pairs must share one grid size (%dx%d vs %dx%d)", H, W, s->cap_H, s->cap_W);
    const long img = (long)H * W * 4;
    const long n = (long)nvx(g) + nvz(g) + ncell(g);
    if (s->cap_n == 0 && s->pend_n == 0) {
        s->cap_H = H;
        s->cap_W = W;
        HIPCHK(c, s->cap_x.reserve(img * s->cap_max));
        HIPCHK(c, s->cap_t.reserve(img * s->cap_max));
        HIPCHK(c, s->cap_m.reserve(img * s->cap_max));
        HIPCHK(c, s->cap_scale.reserve(2L * s->cap_max));
    }
    HIPCHK(c, s->pend_x.reserve(n * OUTER_PEND_MAX));
    // the ring slot this pair will take once the solve has converged
    const int slot = (s->cap_next + s->pend_n) % s->cap_max;
    if (s->pend_n >= s->cap_max) return MDG_OK;
    const double pfac = (1.0 + gamma) / (g.dx * g.dz);
    int rc;
    if ((rc = outer_scales<T>(c, s, g, R, dinv, en, pfac, s->cap_scale.p + 2L * slot))) return rc;
    hipLaunchKernelGGL((k_pack_outer<T, true>), dim3(nblk256((long)H * W)), dim3(256), 0, c->stream, g, H, W, pfac, R,
                       dinv, en, (const double *)(s->cap_scale.p + 2L * slot), (void *)(s->cap_x.p + img * slot));
    HIPCHK(c, hipMemcpyAsync(s->pend_x.p + n * s->pend_n, x, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    // what the residual-loss training needs to re-apply this pair's operator
    const long nc = ncell(g), nv = (long)g.Nx * g.Nz;
    HIPCHK(c, s->cap_R.reserve(n * s->cap_max));
    HIPCHK(c, s->cap_en.reserve(nc * s->cap_max));
    HIPCHK(c, s->cap_es.reserve(nv * s->cap_max));
    HIPCHK(c, s->cap_gamma.reserve(s->cap_max));
    HIPCHK(c, hipMemcpyAsync(s->cap_R.p + n * slot, R, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s->cap_en.p + nc * slot, c->eta_n.p, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s->cap_es.p + nv * slot, c->eta_s.p, nv * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s->cap_gamma.p + slot, &gamma, sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));   // gamma is a host local
    HIPCHK(c, hipGetLastError());
    s->pend_slot[s->pend_n++] = slot;
    return MDG_OK;
}

int mdg_pcnet_outer_capture_done(mdg_ctx *c, const sgrid &g, const double *xstar, int converged) {
    pcnet_state *s = c->pcnet;
    if (!s || !s->cap_on || s->cap_kind != 1) return MDG_OK;
    if (!converged) {       // no reference solution: drop the pending pairs
        s->pend_n = 0;
        return MDG_OK;
    }
    const long img = (long)s->cap_H * s->cap_W * 4;
    const long n = (long)nvx(g) + nvz(g) + ncell(g);
    HIPCHK(c, s->oscale.reserve(2));
    for (int q = 0; q < s->pend_n; q++) {
        double *e = s->pend_x.p + n * q;
        const int slot = s->pend_slot[q];
        hipLaunchKernelGGL(k_err, dim3(nblk256(n)), dim3(256), 0, c->stream, n, xstar, e);
        int rc = outer_scales<double>(c, s, g, e, (const double *)nullptr, (const double *)nullptr, 1.0, s->oscale.p);
        if (rc) return rc;
        hipLaunchKernelGGL(k_pack_outer_target, dim3(nblk256(img / 4)), dim3(256), 0, c->stream, g, s->cap_H, s->cap_W,
                           (const double *)e, (const double *)s->oscale.p, s->cap_t.p + img * slot,
                           s->cap_m.p + img * slot);
        HIPCHK(c, hipGetLastError());
    }
    s->cap_next = (s->cap_next + s->pend_n) % s->cap_max;
    s->cap_n = std::min(s->cap_max, s->cap_n + s->pend_n);
    s->pend_n = 0;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

template int mdg_pcnet_outer_dirs<double>(mdg_ctx *, const sgrid &, double, const double *, const double *,
                                          const double *, double *, double *);
template int mdg_pcnet_outer_dirs<float>(mdg_ctx *, const sgrid &, double, const double *, const float *,
                                         const double *, double *, double *);
template int mdg_pcnet_outer_capture_in<double>(mdg_ctx *, const sgrid &, double, const double *, const double *,
                                                const double *, const double *);
template int mdg_pcnet_outer_capture_in<float>(mdg_ctx *, const sgrid &, double, const double *, const float *,
                                               const double *, const double *);


// ---------------------------------------------- residual-loss training
namespace {

// the six inner products of the two-direction least-squares fit:
// part[b*6 + j] for (Ku.Ku, Ku.Kp, Kp.Kp, Ku.R, Kp.R, R.R)
__global__ void __launch_bounds__(256) k_dots6(long n, const double *__restrict__ ku, const double *__restrict__ kp,
                                               const double *__restrict__ r, double *__restrict__ part) {
    __shared__ double sh[6][256];
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const double u = ku[i], p = kp[i], q = r[i];
        a[0] += u * u; a[1] += u * p; a[2] += p * p; a[3] += u * q; a[4] += p * q; a[5] += q * q;
    }
    for (int j = 0; j < 6; j++) sh[j][threadIdx.x] = a[j];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int j = 0; j < 6; j++) sh[j][threadIdx.x] += sh[j][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int j = 0; j < 6; j++) part[blockIdx.x * 6 + j] = sh[j][0];
}
// y = x - a u - b v
__global__ void k_res2(long n, const double *__restrict__ x, double a, const double *__restrict__ u, double b,
                       const double *__restrict__ v, double *__restrict__ y) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = x[i] - a * u[i] - b * v[i];
}
// negate the pressure block of a grid-layout vector (in place)
__global__ void k_flip_p(long nvel, long n, double *__restrict__ x) {
    const long i = nvel + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = -x[i];
}
// Output-image gradient from the grid gradients: the transpose of
// k_unpack_outer (channel 0 <- W face of gu, 1 <- S face of gu, 2 <- cell of
// gp), each times its factor.
__global__ void k_pack_grad(sgrid g, int H, int W, const double *__restrict__ gvec, double fu, double fp,
                            float *__restrict__ dy) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)H * W) return;
    const int y = (int)(i / W), x = (int)(i - (long)y * W);
    const int Ncx = g.Nx - 1, Ncz = g.Nz - 1, nu = nvx(g), nvel = nu + nvz(g);
    float d0 = 0.f, d1 = 0.f, d2 = 0.f;
    if (x < Ncx && y < Ncz) {
        const int ju = x + (y + 1) * g.Nx;
        if (vx_free(g, x, y + 1)) d0 = (float)(fu * gvec[ju]);
        const int jv = nu + (x + 1) + y * (g.Nx + 1);
        if (vz_free(g, x + 1, y)) d1 = (float)(fu * gvec[jv]);
        d2 = (float)(fp * gvec[nvel + x + y * Ncx]);
    }
    float *o = dy + i * 4;
    o[0] = d0; o[1] = d1; o[2] = d2; o[3] = 0.f;
}

}  // namespace

// Trains the network for precond = 2 on the captured outer pairs with the
// loss the solver actually sees: the residual left after the best
// combination of the two network directions,
//   L = min_{a,b} ||R - a K z_u - b K z_p||^2 / ||R||^2,
// whose gradient (envelope theorem) is dL/dz_u = -2a K^T r / ||R||^2 and
// dL/dz_p = -2b K^T r / ||R||^2 with r the residual at the optimum;
// K^T = P K P with P negating the pressure block (B = -C^T).
extern "C" int mdg_pcnet_train_residual(mdg_ctx *c, int steps, int batch, float lr, unsigned long long seed,
                                        double *losses) {
    if (!c || steps < 0 || batch < 1) return mdg_fail(c, MDG_ERR_ARG, "bad training arguments");
    pcnet_state *s = net_of(c);
    if (!s->P.p) return mdg_fail(c, MDG_ERR_STATE, "network parameters not set");
    if (s->cap_n < 1 || s->cap_kind != 1 || !s->cap_R.p)
        return mdg_fail(c, MDG_ERR_STATE, "no captured outer pairs (mdg_pcnet_capture_outer)");
    const sgrid g = c->sg;
    int H, W, rc;
    image_dims(g, &H, &W);
    if (H != s->cap_H || W != s->cap_W) return mdg_fail(c, MDG_ERR_ARG, "pairs were captured on another grid");
    if ((rc = check_dims(c, H, W)) || (rc = ensure_buffers(c, s, H, W))) return rc;
    const long n = (long)nvx(g) + nvz(g) + ncell(g), nvel = (long)nvx(g) + nvz(g);
    const long nc = ncell(g), nv = (long)g.Nx * g.Nz, img = (long)H * W * 4;
    HIPCHK(c, s->t_zu.reserve(n)); HIPCHK(c, s->t_zp.reserve(n)); HIPCHK(c, s->t_ku.reserve(n));
    HIPCHK(c, s->t_kp.reserve(n)); HIPCHK(c, s->t_r.reserve(n)); HIPCHK(c, s->t_g.reserve(n));
    const int nbd = 256;
    HIPCHK(c, s->t_part.reserve(6 * nbd));
    HIPCHK(c, c->eta_n.reserve(nc));
    HIPCHK(c, c->eta_s.reserve(nv));
    std::vector<double> gam(s->cap_n), part(6 * nbd);
    HIPCHK(c, hipMemcpy(gam.data(), s->cap_gamma.p, s->cap_n * sizeof(double), hipMemcpyDeviceToHost));
    for (int st = 0; st < steps; st++) {
        double lsum = 0.0;
        for (int j = 0; j < batch; j++) {
            unsigned long long h = seed * 0x9E3779B97F4A7C15ull + (unsigned long long)st * 0xBF58476D1CE4E5B9ull + (unsigned long long)j;
            h ^= h >> 31; h *= 0x94D049BB133111EBull; h ^= h >> 29;
            const int i = (int)(h % (unsigned long long)s->cap_n);
            const double *R = s->cap_R.p + n * i;
            HIPCHK(c, hipMemcpyAsync(c->eta_n.p, s->cap_en.p + nc * i, nc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->eta_s.p, s->cap_es.p + nv * i, nv * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            hipLaunchKernelGGL(k_to_bf16, dim3(nblk256(img)), dim3(256), 0, c->stream, img, (const float *)(s->cap_x.p + img * i), A(s, A_X));
            if ((rc = forward(c, s))) return rc;
            hipLaunchKernelGGL(k_unpack_outer, dim3(nblk256(n)), dim3(256), 0, c->stream, g, H, W, (const float *)s->y.p,
                               s->t_zu.p, s->t_zp.p);
            if ((rc = mdg_apply_Kg_grid(c, gam[i], s->t_zu.p, s->t_ku.p)) || (rc = mdg_apply_Kg_grid(c, gam[i], s->t_zp.p, s->t_kp.p)))
                return rc;
            hipLaunchKernelGGL(k_dots6, dim3(nbd), dim3(256), 0, c->stream, n, (const double *)s->t_ku.p, (const double *)s->t_kp.p, R,
                               s->t_part.p);
            HIPCHK(c, hipMemcpyAsync(part.data(), s->t_part.p, 6 * nbd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            double d[6] = {0, 0, 0, 0, 0, 0};
            for (int b = 0; b < nbd; b++)
                for (int q = 0; q < 6; q++) d[q] += part[b * 6 + q];
            const double det = d[0] * d[2] - d[1] * d[1];
            double a = 0.0, bb = 0.0, loss = 1.0;
            if (d[5] > 0.0 && det > 1e-12 * d[0] * d[2] && det > 0.0) {
                a = (d[3] * d[2] - d[4] * d[1]) / det;
                bb = (d[0] * d[4] - d[1] * d[3]) / det;
                loss = (d[5] - a * d[3] - bb * d[4]) / d[5];
            }
            // r = R - a Ku - b Kp;  g = K^T r = P K P r
            hipLaunchKernelGGL(k_res2, dim3(nblk256(n)), dim3(256), 0, c->stream, n, R, a, (const double *)s->t_ku.p, bb,
                               (const double *)s->t_kp.p, s->t_r.p);
            hipLaunchKernelGGL(k_flip_p, dim3(nblk256(n - nvel)), dim3(256), 0, c->stream, nvel, n, s->t_r.p);
            if ((rc = mdg_apply_Kg_grid(c, gam[i], s->t_r.p, s->t_g.p))) return rc;
            hipLaunchKernelGGL(k_flip_p, dim3(nblk256(n - nvel)), dim3(256), 0, c->stream, nvel, n, s->t_g.p);
            const double inv = d[5] > 0.0 ? 1.0 / d[5] : 0.0;
            hipLaunchKernelGGL(k_pack_grad, dim3(nblk256((long)H * W)), dim3(256), 0, c->stream, g, H, W,
                               (const double *)s->t_g.p, -2.0 * a * inv, -2.0 * bb * inv, Gb(s, 0));
            HIPCHK(c, hipGetLastError());
            if ((rc = backward(c, s, j > 0))) return rc;
            lsum += loss;
        }
        if ((rc = mdg_pcnet_adam(c, lr / (float)batch))) return rc;
        if (losses) losses[st] = lsum / batch;
    }
    return MDG_OK;
}

extern "C" int mdg_pcnet_capture(mdg_ctx *c, int max_pairs) {
    if (!c || max_pairs < 0) return mdg_fail(c, MDG_ERR_ARG, "max_pairs must be >= 0");
    pcnet_state *s = net_of(c);
    s->cap_pending = 0;
    if (max_pairs == 0) {       // stop; the captured pairs stay available
        s->cap_on = 0;
        return MDG_OK;
    }
    if (max_pairs > s->cap_max) {   // grow the ring before the first capture
        s->cap_x.release(); s->cap_t.release(); s->cap_m.release(); s->cap_scale.release();
    }
    s->cap_max = max_pairs;
    s->cap_on = 1;
    s->cap_n = 0;
    s->cap_next = 0;
    s->cap_kind = 0;
    s->pend_n = 0;
    return MDG_OK;
}

extern "C" int mdg_pcnet_capture_outer(mdg_ctx *c, int max_pairs) {
    int rc = mdg_pcnet_capture(c, max_pairs);
    if (rc) return rc;
    if (max_pairs > 0) c->pcnet->cap_kind = 1;
    return MDG_OK;
}

extern "C" int mdg_pcnet_captured(mdg_ctx *c, int *n, int *H, int *W) {
    if (!c) return MDG_ERR_ARG;
    pcnet_state *s = net_of(c);
    if (n) *n = s->cap_n;
    if (H) *H = s->cap_H;
    if (W) *W = s->cap_W;
    return MDG_OK;
}

// Captured pair i as host arrays (H*W*4 each; any may be NULL).
extern "C" int mdg_pcnet_get_pair(mdg_ctx *c, int i, float *x, float *t, float *m) {
    if (!c) return MDG_ERR_ARG;
    pcnet_state *s = net_of(c);
    if (i < 0 || i >= s->cap_n) return mdg_fail(c, MDG_ERR_ARG, "pair %d of %d", i, s->cap_n);
    const long img = (long)s->cap_H * s->cap_W * 4;
    if (x) HIPCHK(c, hipMemcpyAsync(x, s->cap_x.p + img * i, img * 4, hipMemcpyDefault, c->stream));
    if (t) HIPCHK(c, hipMemcpyAsync(t, s->cap_t.p + img * i, img * 4, hipMemcpyDefault, c->stream));
    if (m) HIPCHK(c, hipMemcpyAsync(m, s->cap_m.p + img * i, img * 4, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

// steps Adam updates, each on `batch` captured pairs drawn by a counter
// hash of (seed, step, j); loss per step into losses (may be NULL).
extern "C" int mdg_pcnet_train(mdg_ctx *c, int steps, int batch, float lr, unsigned long long seed,
                               double *losses) {
    if (!c || steps < 0 || batch < 1) return mdg_fail(c, MDG_ERR_ARG, "bad training arguments");
    pcnet_state *s = net_of(c);
    if (!s->P.p) return mdg_fail(c, MDG_ERR_STATE, "network parameters not set");
    if (s->cap_n < 1) return mdg_fail(c, MDG_ERR_STATE, "no captured training pairs");
    const long img = (long)s->cap_H * s->cap_W * 4;
    for (int st = 0; st < steps; st++) {
        double lsum = 0.0;
        for (int j = 0; j < batch; j++) {
            unsigned long long h = seed * 0x9E3779B97F4A7C15ull + (unsigned long long)st * 0xBF58476D1CE4E5B9ull + (unsigned long long)j;
            h ^= h >> 31; h *= 0x94D049BB133111EBull; h ^= h >> 29;
            const int i = (int)(h % (unsigned long long)s->cap_n);
            double l = 0.0;
            int rc = mdg_pcnet_grad(c, s->cap_H, s->cap_W, s->cap_x.p + img * i, s->cap_t.p + img * i,
                                    s->cap_m.p + img * i, j > 0, &l);
            if (rc) return rc;
            lsum += l;
        }
        int rc = mdg_pcnet_adam(c, lr / (float)batch);
        if (rc) return rc;
        if (losses) losses[st] = lsum / batch;
    }
    return MDG_OK;
}
