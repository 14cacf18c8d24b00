// Periodic 3D CNN that predicts the L-FTS pressure field w+ from the exchange
// field w- (the learned initial guess of the saddle-point iteration; it has no
// counterpart in the reference, whose drivers start each saddle point from the
// previous step's w+, examples/fts/lfts.py:701-717).
//
// Network: conv(n_in->C) ReLU, 4 x [conv(C->C) ReLU], conv(C->1); 3x3x3 kernels with
// circular padding, C = 32, n_in = 1 (w- alone) or 3 (w-, its change in the Langevin
// update and the previous w+). Hidden activations are bf16, channels-last [voxel][C];
// every product accumulates in FP32.
//
//   hidden layers    implicit GEMM on v_mfma_f32_32x32x16_bf16: one wavefront owns
//                    32 voxels x 32 output channels and accumulates 27 taps x 2
//                    input-channel chunks of 16
//   first / last     FP32 FMA (n_in input or 1 output channel)
//   backward         data gradients: the same MFMA kernel with flipped, transposed
//                    bf16 weights and the ReLU mask of the layer input; hidden weight
//                    gradients: MFMA GEMMs over voxels, split over voxel chunks and
//                    reduced in a fixed order, with both operands read from
//                    channel-major copies (written by the producing kernels while
//                    training) when nz % 8 == 0; first- and last-layer weight
//                    gradients: GEMMs over voxels on v_mfma_f32_32x32x2_f32 (FP32
//                    operands), one launch for all input channels
//   optimizer        Adam on FP32 master weights, then bf16 copies for the kernels
//
// Tap t = 9 (dx + 1) + 3 (dy + 1) + (dz + 1); the flipped tap is 26 - t.
#include <cmath>
#include <vector>

#include "fts_internal.hpp"
#include "fts_reduce.hpp"

namespace fts {

namespace {

constexpr int C = FTS_CNN_CHANNELS;
constexpr int TAPS = 27;
constexpr int NHID = 4;                                 // hidden C->C layers
constexpr int HSZ = TAPS * C * C + C;
constexpr int WG_CHUNK = 1024;                          // voxels per block in the weight-gradient kernels
constexpr int HID_CHUNK = 4096;                         // voxels per block of the channel-major hidden kernel

// Offsets in the FP32 master vector for n_in input channels:
// W0[n_in][27][C], b0[C], NHID x (W[27][C co][C ci], b[C]), WL[27][C], bL[1]
struct Layout {
    int w0, b0, wh, wl, bl, n;
    explicit Layout(int n_in) {
        w0 = 0;
        b0 = w0 + n_in * TAPS * C;
        wh = b0 + C;
        wl = wh + NHID * HSZ;
        bl = wl + TAPS * C;
        n = bl + 1;
    }
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct Geo {
    int nx, ny, nz, M;
};

// a voxel's coordinates: computed once per voxel, so the 27 taps need no division
struct Vox {
    int x, y, z;
};

__device__ __forceinline__ Vox coords(const Geo& g, int v) {
    const int r = v / g.nz;
    return Vox{r / g.ny, r % g.ny, v % g.nz};
}

// the voxel k further along the flat index (k >= 0), by carries
__device__ __forceinline__ Vox advance(const Geo& g, Vox c, int k) {
    c.z += k;
    while (c.z >= g.nz) {
        c.z -= g.nz;
        if (++c.y == g.ny) {
            c.y = 0;
            ++c.x;
        }
    }
    return c;
}

// periodic neighbour of the voxel at c through tap t
__device__ __forceinline__ int neighbour(const Geo& g, Vox c, int t) {
    int xx = c.x + t / 9 - 1, yy = c.y + (t / 3) % 3 - 1, zz = c.z + t % 3 - 1;
    xx = xx < 0 ? xx + g.nx : (xx >= g.nx ? xx - g.nx : xx);
    yy = yy < 0 ? yy + g.ny : (yy >= g.ny ? yy - g.ny : yy);
    zz = zz < 0 ? zz + g.nz : (zz >= g.nz ? zz - g.nz : zz);
    return (xx * g.ny + yy) * g.nz + zz;
}

// Channel-major copies for the hidden weight gradients, blocked by 8 voxels: channel c of
// voxel v at (v / 8) 8 C + 8 c + v % 8, so that the 8 voxels of a group are 16 contiguous bytes
// per channel and the 32 channels of a group one contiguous 512-byte block. (A plain
// [C][M] layout puts the channels M apart: a wavefront's load then touched 32 pages.)
__device__ __forceinline__ size_t cm_at(int c, int v) {
    return (size_t)(v >> 3) * (8 * C) + 8 * c + (v & 7);
}

// ---------------------------------------------------------------- forward
// x: [n_in][M], W: [n_in][27][C]
__global__ void __launch_bounds__(256) conv_first(Geo g, int n_in, const float* x, const float* W, const float* b,
                                                  __bf16* out, __bf16* out_cm) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.M) return;
    float acc[C];
#pragma unroll
    for (int c = 0; c < C; c++) acc[c] = b[c];
    const Vox cv = coords(g, v);
    for (int ci = 0; ci < n_in; ci++) {
        const float* xc = x + (size_t)ci * g.M;
        const float* Wc = W + (size_t)ci * TAPS * C;
        for (int t = 0; t < TAPS; t++) {
            const float xv = xc[neighbour(g, cv, t)];
#pragma unroll
            for (int c = 0; c < C; c++) acc[c] += Wc[t * C + c] * xv;
        }
    }
    bf16x8* o = reinterpret_cast<bf16x8*>(out + (size_t)v * C);
#pragma unroll
    for (int q = 0; q < C / 8; q++) {
        bf16x8 h;
#pragma unroll
        for (int e = 0; e < 8; e++) h[e] = (__bf16)fmaxf(acc[q * 8 + e], 0.f);
        o[q] = h;
        if (out_cm) {
#pragma unroll
            for (int e = 0; e < 8; e++) out_cm[cm_at(q * 8 + e, v)] = h[e];
        }
    }
}

// Y[v][co] = act(bias[co] + sum_t sum_ci W[t][co][ci] X[nbr(v,t)][ci]) (x [X_in[v][co] > 0] when
// mask is given: the ReLU derivative of a backward pass). M % 32 == 0.
// Bias, ReLU and ReLU mask of the 32 voxels v0.. x 32 output channels of one wavefront's
// accumulator, stored channels-last and, with Ycm, channel-major.
__device__ __forceinline__ void conv_hidden_store(const Geo& g, const f32x16& acc, int v0, int half, int col,
                                                  const float* bias, const __bf16* mask, int relu, __bf16* Y,
                                                  __bf16* Ycm) {
    // D[i][j]: j = lane % 32 (output channel), i = 8 (r / 4) + 4 (lane / 32) + r % 4 (voxel)
    const float bj = bias ? bias[col] : 0.f;
    // the ReLU mask of the backward pass, all 16 entries loaded before they are used
    bool keep[16];
#pragma unroll
    for (int r = 0; r < 16; r++) keep[r] = true;
    if (mask) {
        __bf16 mk[16];
#pragma unroll
        for (int r = 0; r < 16; r++) mk[r] = mask[(size_t)(v0 + 8 * (r >> 2) + 4 * half + (r & 3)) * C + col];
#pragma unroll
        for (int r = 0; r < 16; r++) keep[r] = (float)mk[r] > 0.f;
    }
    __bf16 h[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int v = v0 + 8 * (r >> 2) + 4 * half + (r & 3);
        float val = acc[r] + bj;
        if (relu) val = fmaxf(val, 0.f);
        if (!keep[r]) val = 0.f;
        h[r] = (__bf16)val;
        Y[(size_t)v * C + col] = h[r];
    }
    // channel-major copy: the 4 consecutive voxels r % 4 of a group in one 8-byte store
    if (Ycm) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
            const bf16x4 pk = {h[4 * q], h[4 * q + 1], h[4 * q + 2], h[4 * q + 3]};
            *reinterpret_cast<bf16x4*>(Ycm + cm_at(col, v0 + 8 * q + 4 * half)) = pk;
        }
    }
}

__global__ void __launch_bounds__(256) conv_hidden(Geo g, const __bf16* X, const __bf16* W, const float* bias,
                                                   const __bf16* mask, int relu, __bf16* Y, __bf16* Ycm) {
    const int lane = threadIdx.x & 63;
    const int v0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 32;
    if (v0 >= g.M) return;
    const int half = lane >> 5, col = lane & 31;
    const int vi = v0 + col;   // voxel of this lane's A rows
    const Vox ci = coords(g, vi);
    f32x16 acc = {};
    // the three z taps of one (dx, dy) issue their 12 loads before their 6 products
#pragma unroll 1
    for (int t0 = 0; t0 < TAPS; t0 += 3) {
        bf16x8 a[3][2], b[3][2];
#pragma unroll
        for (int u = 0; u < 3; u++) {
            const __bf16* xa = X + (size_t)neighbour(g, ci, t0 + u) * C + half * 8;
            const __bf16* wb = W + ((size_t)(t0 + u) * C + col) * C + half * 8;
#pragma unroll
            for (int ch = 0; ch < 2; ch++) {
                a[u][ch] = *reinterpret_cast<const bf16x8*>(xa + ch * 16);
                b[u][ch] = *reinterpret_cast<const bf16x8*>(wb + ch * 16);
            }
        }
#pragma unroll
        for (int u = 0; u < 3; u++)
#pragma unroll
            for (int ch = 0; ch < 2; ch++) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[u][ch], b[u][ch], acc, 0, 0, 0);
    }
    conv_hidden_store(g, acc, v0, half, col, bias, mask, relu, Y, Ycm);
}

// The same layer with its operands staged in LDS (nz % 32 == 0, nz <= 128, ny % TILE_Y == 0):
// a workgroup owns TILE_Y y rows of one x plane, every z, as nz/32 x TILE_Y wavefronts of 32
// voxels. It runs the taps in three phases, one per dx: each phase copies the TILE_Y + 2
// periodic lines of that plane around the rows and the 9 taps' weights into LDS, then every
// wavefront takes its A and B operands there. Activation rows are read about 4.5 times instead
// of 27 and the weights once per workgroup instead of once per wavefront. A 64-byte voxel row
// (or weight row) is four 16-byte chunks, stored at chunk k ^ ((i / 4) % 4) for row i, so that
// 16 lanes reading one chunk of 16 consecutive rows hit 64 distinct banks.
constexpr int TILE_Y = 4;
constexpr int TILE_LINES = TILE_Y + 2;          // lines of one dx plane
constexpr int TILE_W_CHUNKS = 9 * C * 4;        // 16-byte chunks of one phase's weights
__host__ __device__ constexpr size_t conv_tile_lds(int nz) {
    return ((size_t)TILE_LINES * nz * 4 + TILE_W_CHUNKS) * 16;
}

__global__ void __launch_bounds__(1024) conv_hidden_tiled(Geo g, const __bf16* X, const __bf16* W,
                                                          const float* bias, const __bf16* mask, int relu, __bf16* Y,
                                                          __bf16* Ycm) {
    extern __shared__ __align__(16) unsigned char conv_lds[];
    const int nz = g.nz;
    u32x4* LA = reinterpret_cast<u32x4*>(conv_lds);   // [line][z][4 chunks]
    u32x4* LW = LA + TILE_LINES * nz * 4;              // [tap][co][4 chunks]
    const int rows = g.ny / TILE_Y;
    const int x = blockIdx.x / rows, y0 = (blockIdx.x - x * rows) * TILE_Y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int half = lane >> 5, col = lane & 31;
    const int wpr = nz >> 5;   // wavefronts per y row
    const int r = wave / wpr, z0 = (wave - r * wpr) * 32;
    const int z = z0 + col;    // voxel of this lane's A rows
    // the halo lines split evenly over the 8 nz threads: 3 chunks each; the weights need up to 5
    constexpr int PA = TILE_LINES * 4 / 8;
    static_assert(TILE_LINES * 4 % 8 == 0, "the halo lines split evenly over the threads");
    constexpr int PW = (TILE_W_CHUNKS + 255) / 256;
    f32x16 acc = {};
#pragma unroll 1
    for (int dxi = 0; dxi < 3; dxi++) {
        if (dxi > 0) __syncthreads();   // every wavefront is done with the previous phase
        u32x4 av[PA], wv[PW];
        int ad[PA];
#pragma unroll
        for (int k = 0; k < PA; k++) {
            const int c = threadIdx.x + k * blockDim.x;
            const int q = c & 3, vz = c >> 2;
            const int zz = vz % nz, ry = vz / nz;
            int xx = x + dxi - 1, yy = y0 + ry - 1;
            xx = xx < 0 ? xx + g.nx : (xx >= g.nx ? xx - g.nx : xx);
            yy = yy < 0 ? yy + g.ny : (yy >= g.ny ? yy - g.ny : yy);
            av[k] = *reinterpret_cast<const u32x4*>(X + ((size_t)(xx * g.ny + yy) * nz + zz) * C + q * 8);
            ad[k] = (ry * nz + zz) * 4 + (q ^ ((zz >> 2) & 3));
        }
        const u32x4* wsrc = reinterpret_cast<const u32x4*>(W) + (size_t)dxi * TILE_W_CHUNKS;
#pragma unroll
        for (int k = 0; k < PW; k++) {
            const int c = threadIdx.x + k * blockDim.x;
            if (c < TILE_W_CHUNKS) wv[k] = wsrc[c];
        }
#pragma unroll
        for (int k = 0; k < PA; k++) LA[ad[k]] = av[k];
#pragma unroll
        for (int k = 0; k < PW; k++) {
            const int c = threadIdx.x + k * blockDim.x;
            const int row = c >> 2;   // tap * C + co
            if (c < TILE_W_CHUNKS) LW[(row << 2) + ((c & 3) ^ ((row >> 2) & 3))] = wv[k];
        }
        __syncthreads();
        // the three z taps of one dy read their operands before their 6 products
#pragma unroll 1
        for (int dyi = 0; dyi < 3; dyi++) {
            const int line = r + dyi;   // y row r + dy + 1 of the plane's lines
            bf16x8 a[3][2], b[3][2];
#pragma unroll
            for (int u = 0; u < 3; u++) {
                int zz = z + u - 1;
                zz = zz < 0 ? zz + nz : (zz >= nz ? zz - nz : zz);
                const int wrow = (3 * dyi + u) * C + col;
#pragma unroll
                for (int ch = 0; ch < 2; ch++) {
                    const int k = half + 2 * ch;
                    a[u][ch] = __builtin_bit_cast(bf16x8, LA[(line * nz + zz) * 4 + (k ^ ((zz >> 2) & 3))]);
                    b[u][ch] = __builtin_bit_cast(bf16x8, LW[(wrow << 2) + (k ^ ((wrow >> 2) & 3))]);
                }
            }
#pragma unroll
            for (int u = 0; u < 3; u++)
#pragma unroll
                for (int ch = 0; ch < 2; ch++)
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[u][ch], b[u][ch], acc, 0, 0, 0);
        }
    }
    conv_hidden_store(g, acc, ((x * g.ny + y0 + r) * nz) + z0, half, col, bias, mask, relu, Y, Ycm);
}

__global__ void __launch_bounds__(256) conv_last(Geo g, const __bf16* X, const float* W, const float* b, float* y) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.M) return;
    float acc = b[0];
    const Vox cv = coords(g, v);
    for (int t = 0; t < TAPS; t++) {
        const bf16x8* xv = reinterpret_cast<const bf16x8*>(X + (size_t)neighbour(g, cv, t) * C);
#pragma unroll
        for (int q = 0; q < C / 8; q++) {
            const bf16x8 h = xv[q];
#pragma unroll
            for (int e = 0; e < 8; e++) acc += W[t * C + q * 8 + e] * (float)h[e];
        }
    }
    y[v] = acc;
}

// ---------------------------------------------------------------- loss
// d = 2 (y - target) / M, partial sums of (y - target)^2
__global__ void __launch_bounds__(256) loss_grad(Geo g, const float* y, const float* target, float* dy,
                                                 double* partials) {
    __shared__ double ws[16];
    double acc = 0.0;
    for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < g.M; v += gridDim.x * blockDim.x) {
        const float d = y[v] - target[v];
        dy[v] = 2.f * d / (float)g.M;
        acc += (double)d * d;
    }
    acc = block_sum(acc, ws);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(256) sum_partials(const double* partials, int n, double* out) {
    __shared__ double ws[16];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partials[i];
    acc = block_sum(acc, ws);
    if (threadIdx.x == 0) out[0] = acc;
}

// ---------------------------------------------------------------- backward
// dX[u][c] = [X[u][c] > 0] sum_t W[t][c] dy[nbr(u, 26 - t)]  (last layer, 1 output channel)
__global__ void __launch_bounds__(256) back_last(Geo g, const float* dy, const float* W, const __bf16* X, __bf16* dX,
                                                 __bf16* dXcm) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= g.M) return;
    float acc[C] = {};
    const Vox cu = coords(g, u);
    for (int t = 0; t < TAPS; t++) {
        const float d = dy[neighbour(g, cu, 26 - t)];
#pragma unroll
        for (int c = 0; c < C; c++) acc[c] += W[t * C + c] * d;
    }
    const bf16x8* xv = reinterpret_cast<const bf16x8*>(X + (size_t)u * C);
    bf16x8* o = reinterpret_cast<bf16x8*>(dX + (size_t)u * C);
#pragma unroll
    for (int q = 0; q < C / 8; q++) {
        const bf16x8 h = xv[q];
        bf16x8 r;
#pragma unroll
        for (int e = 0; e < 8; e++) r[e] = (__bf16)((float)h[e] > 0.f ? acc[q * 8 + e] : 0.f);
        o[q] = r;
        if (dXcm) {
#pragma unroll
            for (int e = 0; e < 8; e++) dXcm[cm_at(q * 8 + e, u)] = r[e];
        }
    }
}

// Hidden weight gradient, per tap and voxel chunk: G[t][co][ci] = sum_v dZ[v][co] X[nbr(v,t)][ci]
// as D = A B with A = dZ^T (32 x voxels), B = X_t (voxels x 32). One wavefront sums WG_CHUNK / 4
// voxels; the block's 4 wavefronts meet in LDS; partial[chunk][t][co][ci].
__global__ void __launch_bounds__(256) wgrad_hidden(Geo g, const __bf16* dZ, const __bf16* X, float* partial) {
    __shared__ float red[4][C * C];
    const int t = blockIdx.y;
    const int chunk = blockIdx.x;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int half = lane >> 5, col = lane & 31;
    const int vbeg = chunk * WG_CHUNK + wave * (WG_CHUNK / 4);
    const int vend = min(vbeg + WG_CHUNK / 4, g.M);
    f32x16 acc = {};
    // coordinates of this lane's first voxel, carried along the loop
    Vox cb = coords(g, min(vbeg + half * 8, g.M - 1));
    for (int v0 = vbeg; v0 < vend; v0 += 16) {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int v = v0 + half * 8 + e;
            const bool ok = v < vend;
            a[e] = ok ? dZ[(size_t)v * C + col] : (__bf16)0.f;
            b[e] = ok ? X[(size_t)neighbour(g, advance(g, cb, e), t) * C + col] : (__bf16)0.f;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        if (v0 + 16 < vend) cb = advance(g, cb, 16);
    }
    // D[i = co][j = ci]: j = lane % 32, i = 8 (r / 4) + 4 (lane / 32) + r % 4
#pragma unroll
    for (int r = 0; r < 16; r++) red[wave][(8 * (r >> 2) + 4 * half + (r & 3)) * C + col] = acc[r];
    __syncthreads();
    float* out = partial + ((size_t)chunk * TAPS + t) * C * C;
    for (int i = threadIdx.x; i < C * C; i += blockDim.x) out[i] = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
}

// The same gradient from channel-major copies dZc, Xc (blocked by 8 voxels, cm_at; nz % 8 == 0). Block
// (p, chunk) owns the taps t = 3 p + dz + 1 (dz = -1, 0, 1) of the line offset
// p = 3 (dx + 1) + (dy + 1); its 4 wavefronts split the chunk's voxels. A wavefront step
// takes 16 voxels: lane (c, h) holds dZc[c][v .. v + 7], v = v0 + 8 h (the A operand, one
// z line), and builds the three B operands Xc[c][nbr(v .. v + 7)] from the aligned
// 8-voxel vectors at z - 8, z and z + 8 of the neighbour line (periodic in z), shifted by
// one element with v_alignbit for dz = -1 and +1: 4 vector loads for 3 MFMAs.
// partial[chunk][t][co][ci].
__global__ void __launch_bounds__(256) wgrad_hidden_cm(Geo g, int n_chunks, const __bf16* dZc, const __bf16* Xc,
                                                       float* partial) {
    // one tap's four wavefront sums at a time: 16 KB of LDS, so that LDS does not limit
    // the workgroups per CU
    __shared__ float red[4][C * C];
    // the 9 (dx, dy) blocks of a voxel chunk read the same dZ and overlapping X lines: they get
    // linear ids 8 apart (b = 8 (9 q + p) + k for chunk 8 q + k), so that they run on one XCD of
    // the 8, which hand out workgroups round-robin, and share its L2
    const int k8 = blockIdx.x & 7, m = blockIdx.x >> 3;
    const int q = m / 9, p = m - 9 * q, chunk = 8 * q + k8;
    if (chunk >= n_chunks) return;   // pad of the last group of 8 chunks
    const int dx = p / 3 - 1, dy = p % 3 - 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int half = lane >> 5, col = lane & 31;
    const int vbeg = chunk * HID_CHUNK + wave * (HID_CHUNK / 4);
    const int vend = min(vbeg + HID_CHUNK / 4, g.M);
    const __bf16* dz_c = dZc + 8 * col;   // + (v / 8) 8 C for the group of voxel v
    const __bf16* x_c = Xc + 8 * col;
    f32x16 acc0 = {}, acc1 = {}, acc2 = {};
    // four 16-voxel steps' loads are in flight before their products: with two or three
    // wavefronts per SIMD the loop is otherwise one memory latency per step
    for (int v00 = vbeg; v00 < vend; v00 += 64) {
        u32x4 a[4], cu[4];
        unsigned lo[4], hi[4];   // the dwords before and after the 8 z values: z - 2, z - 1 and z + 8, z + 9
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            // branch-free: 8 voxels past the end read those at vbeg and get a zero dZ operand
            // where the products are formed (a select here would wait for the load)
            ok[u] = v00 + 16 * u + 8 * half < vend;
            const int v = ok[u] ? v00 + 16 * u + 8 * half : vbeg;
            a[u] = *reinterpret_cast<const u32x4*>(dz_c + (size_t)(v >> 3) * (8 * C));
            const int z = v % g.nz, r = v / g.nz;
            int xx = r / g.ny + dx, yy = r % g.ny + dy;
            xx = xx < 0 ? xx + g.nx : (xx >= g.nx ? xx - g.nx : xx);
            yy = yy < 0 ? yy + g.ny : (yy >= g.ny ? yy - g.ny : yy);
            // the line's 8-voxel groups, periodic in z: z - 2 and z - 1 end the previous group,
            // z + 8 and z + 9 start the next
            const __bf16* line = x_c + (size_t)((xx * g.ny + yy) * g.nz >> 3) * (8 * C);
            const int gz = z >> 3, ng = g.nz >> 3;
            lo[u] = *reinterpret_cast<const unsigned*>(line + (size_t)(gz == 0 ? ng - 1 : gz - 1) * (8 * C) + 6);
            cu[u] = *reinterpret_cast<const u32x4*>(line + (size_t)gz * (8 * C));
            hi[u] = *reinterpret_cast<const unsigned*>(line + (size_t)(gz + 1 == ng ? 0 : gz + 1) * (8 * C));
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            // element e of a vector sits in bits 16 (e % 2) of dword e / 2
            const u32x4 c = cu[u];
            const u32x4 bm = {__builtin_amdgcn_alignbit(c.x, lo[u], 16), __builtin_amdgcn_alignbit(c.y, c.x, 16),
                              __builtin_amdgcn_alignbit(c.z, c.y, 16), __builtin_amdgcn_alignbit(c.w, c.z, 16)};
            const u32x4 bp = {__builtin_amdgcn_alignbit(c.y, c.x, 16), __builtin_amdgcn_alignbit(c.z, c.y, 16),
                              __builtin_amdgcn_alignbit(c.w, c.z, 16), __builtin_amdgcn_alignbit(hi[u], c.w, 16)};
            const bf16x8 av = __builtin_bit_cast(bf16x8, ok[u] ? a[u] : u32x4{0, 0, 0, 0});
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, bm), acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, c), acc1, 0, 0, 0);
            acc2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, __builtin_bit_cast(bf16x8, bp), acc2, 0, 0, 0);
        }
    }
    // D[i = co][j = ci]: j = lane % 32, i = 8 (r / 4) + 4 (lane / 32) + r % 4
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const f32x16& acc = k == 0 ? acc0 : (k == 1 ? acc1 : acc2);
        if (k > 0) __syncthreads();   // the previous tap's sums are read
#pragma unroll
        for (int r = 0; r < 16; r++) red[wave][(8 * (r >> 2) + 4 * half + (r & 3)) * C + col] = acc[r];
        __syncthreads();
        for (int e = threadIdx.x; e < C * C; e += blockDim.x)
            partial[((size_t)chunk * TAPS + 3 * p + k) * C * C + e] =
                ((red[0][e] + red[1][e]) + red[2][e]) + red[3][e];
    }
}

// Weight gradients of the first and last layers, G[t][c] = sum_v H[v][c] s[nbr(v, t')]:
//   first layer, one grid row k per input channel: H = dZ, s = x_k, t' = t;
//   last layer (FLIP): H = X, the last hidden activation, s = dy, t' = 26 - t, since
//   sum_v dy[v] X[v + o][c] = sum_u X[u][c] dy[u - o] on the periodic grid.
// A GEMM over voxels on v_mfma_f32_32x32x2_f32 (FP32 products and sums): per instruction, lane
// (k, j) supplies A[c = j][k] = H[v_k][j] and B[k][t = j] = s[nbr(v_k, t')] for the voxel pair
// (v_0, v_1), so every lane issues one coalesced H load and one gather per voxel pair; taps 27..31
// are zero. Each wavefront takes a quarter of the chunk with two accumulators; the four
// wavefronts' sums meet in LDS in a fixed order; partial[chunk][k][t][c].
template <bool FLIP>
__global__ void __launch_bounds__(256) wgrad_edge(Geo g, const __bf16* H, const float* s, float* partial) {
    __shared__ float red[4][TAPS * C];
    const int k = blockIdx.y;
    s += (size_t)k * g.M;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = lane & 31, kk = lane >> 5;
    const bool tap = j < TAPS;
    const int tt = FLIP ? TAPS - 1 - j : j;
    const int dx = tt / 9 - 1, dy = (tt / 3) % 3 - 1, dz = tt % 3 - 1;
    const int vbeg = blockIdx.x * WG_CHUNK, vend = min(vbeg + WG_CHUNK, g.M);
    const int p0 = vbeg + wave * (WG_CHUNK / 4), p1 = min(p0 + WG_CHUNK / 4, vend);
    f32x16 acc0 = {}, acc1 = {};
    int v = p0 + kk;
    Vox cv = coords(g, v);
    // eight voxel pairs' loads are in flight before their products (with few wavefronts per SIMD
    // the loop is otherwise one memory latency per pair). Branch-free, so that no control-flow
    // join waits for them: a voxel past the end reads voxel p0 and gets a = 0, a padding tap
    // reads entry 0 and gets b = 0.
    for (int p = p0; p < p1; p += 16) {
        float a[8], b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const bool ok = v < p1;
            int xx = cv.x + dx, yy = cv.y + dy, zz = cv.z + dz;
            xx = xx < 0 ? xx + g.nx : (xx >= g.nx ? xx - g.nx : xx);
            yy = yy < 0 ? yy + g.ny : (yy >= g.ny ? yy - g.ny : yy);
            zz = zz < 0 ? zz + g.nz : (zz >= g.nz ? zz - g.nz : zz);
            const int nb = ok && tap ? (xx * g.ny + yy) * g.nz + zz : 0;
            const int vh = ok ? v : p0;
            const float av = (float)H[(size_t)vh * C + j];
            const float bv = s[nb];
            a[u] = ok ? av : 0.f;
            b[u] = tap ? bv : 0.f;
            // the next voxel pair: nz >= 3, so one carry at most
            v += 2;
            cv.z += 2;
            const bool wz = cv.z >= g.nz;
            cv.z = wz ? cv.z - g.nz : cv.z;
            cv.y += wz ? 1 : 0;
            const bool wy = cv.y == g.ny;
            cv.y = wy ? 0 : cv.y;
            cv.x += wy ? 1 : 0;
        }
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u], b[u], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a[u + 1], b[u + 1], acc1, 0, 0, 0);
        }
    }
    // D[c][t]: register r of lane (kk, j) holds row c = 8 (r / 4) + 4 kk + r % 4, column t = j
    if (tap) {
#pragma unroll
        for (int r = 0; r < 16; r++) red[wave][j * C + 8 * (r >> 2) + 4 * kk + (r & 3)] = acc0[r] + acc1[r];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TAPS * C; i += blockDim.x)
        partial[((size_t)blockIdx.x * gridDim.y + k) * TAPS * C + i] =
            ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
}

// Bias gradients: channel sums of dZ (C channels) or of dy (1 channel); partial[chunk][c]
__global__ void __launch_bounds__(256) bias_grad(Geo g, const __bf16* dZ, const float* dy, float* partial) {
    __shared__ float red[256];
    const int vbeg = blockIdx.x * WG_CHUNK, vend = min(vbeg + WG_CHUNK, g.M);
    const int c = threadIdx.x % C, lane_group = threadIdx.x / C;   // 8 groups of 32
    float acc = 0.f;
    if (dZ) {
#pragma unroll 8
        for (int v = vbeg + lane_group; v < vend; v += 8) acc += (float)dZ[(size_t)v * C + c];
    } else if (c == 0) {
#pragma unroll 8
        for (int v = vbeg + lane_group; v < vend; v += 8) acc += dy[v];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < C) {
        float s = 0.f;
        for (int k = 0; k < 8; k++) s += red[k * C + threadIdx.x];
        partial[(size_t)blockIdx.x * C + threadIdx.x] = s;
    }
}

// grad[i] = sum over chunks of partial[chunk * stride + i], i < n (fixed order)
__global__ void __launch_bounds__(256) reduce_chunks(const float* partial, int n_chunks, int stride, int n,
                                                     float* grad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    // unrolled so that several chunks' loads are in flight; the sum keeps its order
#pragma unroll 8
    for (int k = 0; k < n_chunks; k++) s += partial[(size_t)k * stride + i];
    grad[i] = s;
}

// ---------------------------------------------------------------- parameters
__global__ void __launch_bounds__(256) adam_step(float* p, const float* gr, float* m, float* v, int n, float lr,
                                                 float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of the hidden weights: forward [t][co][ci] and the backward form
// [t'][ci][co] with t' = 26 - t
__global__ void __launch_bounds__(256) hidden_weights(const float* wh, __bf16* wf, __bf16* wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NHID * TAPS * C * C) return;
    const int l = i / (TAPS * C * C), r = i % (TAPS * C * C);
    const int t = r / (C * C), co = (r / C) % C, ci = r % C;
    const __bf16 w = (__bf16)wh[l * HSZ + r];
    wf[i] = w;
    wb[(size_t)l * TAPS * C * C + ((size_t)(26 - t) * C + ci) * C + co] = w;
}

__global__ void __launch_bounds__(256) to_float(const double* in, double scale, int n, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)(in[i] * scale);
}

__global__ void __launch_bounds__(256) to_double(const float* in, double scale, int n, double* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (double)in[i] * scale;
}

int blocks(int n, int b = 256) { return (n + b - 1) / b; }

}  // namespace

int cnn_n_params(int n_in) { return Layout(n_in).n; }

Cnn::Cnn(fts_context* ctx, int nx, int ny, int nz, int n_in, const double* in_scales, double out_scale)
    : ctx_(ctx), nx_(nx), ny_(ny), nz_(nz), n_in_(n_in), n_params_(Layout(n_in).n), out_scale_(out_scale) {
    FTS_REQUIRE(nx >= 3 && ny >= 3 && nz >= 3, "the CNN needs at least 3 points per axis");
    FTS_REQUIRE(n_in >= 1 && n_in <= 4, "the CNN takes 1 to 4 input channels");
    M_ = nx * ny * nz;
    FTS_REQUIRE(M_ % 32 == 0, "the CNN needs a grid size that is a multiple of 32");
    for (int c = 0; c < n_in; c++) FTS_REQUIRE(in_scales[c] > 0, "scales must be positive");
    FTS_REQUIRE(out_scale > 0, "scales must be positive");
    in_scales_.assign(in_scales, in_scales + n_in);
    n_chunks_ = (M_ + WG_CHUNK - 1) / WG_CHUNK;
    params_.allocate(sizeof(float) * n_params_);
    grads_.allocate(sizeof(float) * n_params_);
    adam_m_.allocate(sizeof(float) * n_params_);
    adam_v_.allocate(sizeof(float) * n_params_);
    wf_.allocate(sizeof(__bf16) * NHID * TAPS * C * C);
    wb_.allocate(sizeof(__bf16) * NHID * TAPS * C * C);
    x_.allocate(sizeof(float) * (size_t)n_in * M_);
    target_.allocate(sizeof(float) * M_);
    y_.allocate(sizeof(float) * M_);
    dy_.allocate(sizeof(float) * M_);
    acts_.allocate(sizeof(__bf16) * (size_t)(NHID + 1) * M_ * C);
    grad_a_.allocate(sizeof(__bf16) * (size_t)M_ * C);
    grad_b_.allocate(sizeof(__bf16) * (size_t)M_ * C);
    cm_ = nz % 8 == 0;
    tiled_ = nz % 32 == 0 && nz <= 128 && ny % TILE_Y == 0;
    if (tiled_)
        FTS_HIP_CHECK(hipFuncSetAttribute((const void*)&conv_hidden_tiled, hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (int)conv_tile_lds(128)));
    n_hchunks_ = (M_ + HID_CHUNK - 1) / HID_CHUNK;
    if (cm_) {
        acts_cm_.allocate(sizeof(__bf16) * (size_t)NHID * M_ * C);
        gcm_a_.allocate(sizeof(__bf16) * (size_t)M_ * C);
        gcm_b_.allocate(sizeof(__bf16) * (size_t)M_ * C);
    }
    partial_.allocate(sizeof(float) * (size_t)std::max(TAPS * n_chunks_ * C * C, n_chunks_ * n_in_ * TAPS * C));
    loss_partials_.allocate(sizeof(double) * 256);
    loss_.allocate(sizeof(double));
    stage_in_.allocate(sizeof(double) * (size_t)n_in * M_);
    stage_out_.allocate(sizeof(double) * M_);
    std::vector<float> zero(n_params_, 0.f);
    set_params(zero.data());
}

void Cnn::set_params(const float* host) {
    version_++;
    hipStream_t st = ctx_->stream;
    FTS_HIP_CHECK(hipMemcpyAsync(params_.get(), host, sizeof(float) * n_params_, hipMemcpyHostToDevice, st));
    FTS_HIP_CHECK(hipMemsetAsync(adam_m_.get(), 0, sizeof(float) * n_params_, st));
    FTS_HIP_CHECK(hipMemsetAsync(adam_v_.get(), 0, sizeof(float) * n_params_, st));
    step_ = 0;
    refresh_weights();
    FTS_HIP_CHECK(hipStreamSynchronize(st));
}

void Cnn::get_params(float* host) {
    FTS_HIP_CHECK(hipMemcpyAsync(host, params_.get(), sizeof(float) * n_params_, hipMemcpyDeviceToHost, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void Cnn::get_grads(float* host) {
    FTS_HIP_CHECK(hipMemcpyAsync(host, grads_.get(), sizeof(float) * n_params_, hipMemcpyDeviceToHost, ctx_->stream));
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
}

void Cnn::conv_hidden_launch(const void* X, const void* W, const float* bias, const void* mask, int relu, void* Y,
                             void* Ycm) {
    const Geo g{nx_, ny_, nz_, M_};
    const __bf16* x = static_cast<const __bf16*>(X);
    const __bf16* w = static_cast<const __bf16*>(W);
    const __bf16* m = static_cast<const __bf16*>(mask);
    __bf16* y = static_cast<__bf16*>(Y);
    __bf16* ycm = static_cast<__bf16*>(Ycm);
    if (tiled_) {
        // nz / 32 wavefronts per y row, TILE_Y rows per workgroup (256 to 1024 threads)
        const size_t lds = conv_tile_lds(nz_);
        hipLaunchKernelGGL(conv_hidden_tiled, dim3(nx_ * (ny_ / TILE_Y)), dim3(8 * nz_), lds, ctx_->stream, g, x, w,
                           bias, m, relu, y, ycm);
    } else {
        hipLaunchKernelGGL(conv_hidden, dim3(blocks(M_ / 32, 4)), dim3(256), 0, ctx_->stream, g, x, w, bias, m, relu,
                           y, ycm);
    }
}

void Cnn::refresh_weights() {
    hipLaunchKernelGGL(hidden_weights, dim3(blocks(NHID * TAPS * C * C)), dim3(256), 0, ctx_->stream,
                       params_.as<float>() + Layout(n_in_).wh, wf_.as<__bf16>(), wb_.as<__bf16>());
    FTS_HIP_CHECK(hipGetLastError());
}

void Cnn::load_input(const double* w_minus, bool on_device) {
    hipStream_t st = ctx_->stream;
    const double* src = w_minus;
    if (!on_device) {
        FTS_HIP_CHECK(hipMemcpyAsync(stage_in_.get(), w_minus, sizeof(double) * n_in_ * M_, hipMemcpyHostToDevice,
                                     st));
        src = stage_in_.as<double>();
    }
    for (int c = 0; c < n_in_; c++)
        hipLaunchKernelGGL(to_float, dim3(blocks(M_)), dim3(256), 0, st, src + (size_t)c * M_, 1.0 / in_scales_[c], M_,
                           x_.as<float>() + (size_t)c * M_);
    FTS_HIP_CHECK(hipGetLastError());
}

void Cnn::forward(bool train) {
    hipStream_t st = ctx_->stream;
    const Geo g{nx_, ny_, nz_, M_};
    const float* p = params_.as<float>();
    const Layout L(n_in_);
    __bf16* a = acts_.as<__bf16>();
    const size_t AS = (size_t)M_ * C;
    // channel-major copies of the hidden layers' inputs a[0..NHID-1] for the weight gradients
    __bf16* acm = train && cm_ ? acts_cm_.as<__bf16>() : nullptr;
    hipLaunchKernelGGL(conv_first, dim3(blocks(M_)), dim3(256), 0, st, g, n_in_, x_.as<float>(), p + L.w0, p + L.b0,
                       a, acm);
    for (int l = 0; l < NHID; l++)
        conv_hidden_launch(a + l * AS, wf_.as<__bf16>() + (size_t)l * TAPS * C * C, p + L.wh + l * HSZ + TAPS * C * C,
                           nullptr, 1, a + (l + 1) * AS, acm && l + 1 < NHID ? acm + (l + 1) * AS : nullptr);
    hipLaunchKernelGGL(conv_last, dim3(blocks(M_)), dim3(256), 0, st, g, a + NHID * AS, p + L.wl, p + L.bl,
                       y_.as<float>());
    FTS_HIP_CHECK(hipGetLastError());
}

void Cnn::predict(const double* w_minus, double* w_plus, bool on_device) {
    load_input(w_minus, on_device);
    forward();
    hipStream_t st = ctx_->stream;
    if (on_device) {
        hipLaunchKernelGGL(to_double, dim3(blocks(M_)), dim3(256), 0, st, y_.as<float>(), out_scale_, M_, w_plus);
        FTS_HIP_CHECK(hipGetLastError());
    } else {
        hipLaunchKernelGGL(to_double, dim3(blocks(M_)), dim3(256), 0, st, y_.as<float>(), out_scale_, M_,
                           stage_out_.as<double>());
        FTS_HIP_CHECK(hipGetLastError());
        FTS_HIP_CHECK(hipMemcpyAsync(w_plus, stage_out_.get(), sizeof(double) * M_, hipMemcpyDeviceToHost, st));
    }
    FTS_HIP_CHECK(hipStreamSynchronize(st));
}

double Cnn::gradients(const double* w_minus, const double* w_plus_target, bool on_device) {
    hipStream_t st = ctx_->stream;
    const Geo g{nx_, ny_, nz_, M_};
    load_input(w_minus, on_device);
    const double* tsrc = w_plus_target;
    if (!on_device) {
        FTS_HIP_CHECK(hipMemcpyAsync(stage_out_.get(), w_plus_target, sizeof(double) * M_, hipMemcpyHostToDevice, st));
        tsrc = stage_out_.as<double>();
    }
    hipLaunchKernelGGL(to_float, dim3(blocks(M_)), dim3(256), 0, st, tsrc, 1.0 / out_scale_, M_, target_.as<float>());
    forward(true);
    const int lb = std::min(256, blocks(M_));
    hipLaunchKernelGGL(loss_grad, dim3(lb), dim3(256), 0, st, g, y_.as<float>(), target_.as<float>(), dy_.as<float>(),
                       loss_partials_.as<double>());
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(256), 0, st, loss_partials_.as<double>(), lb, loss_.as<double>());

    const float* p = params_.as<float>();
    const Layout L(n_in_);
    float* gr = grads_.as<float>();
    float* part = partial_.as<float>();
    __bf16* a = acts_.as<__bf16>();
    const size_t AS = (size_t)M_ * C;
    // last layer: weights, bias, then dZ of the last hidden activation
    hipLaunchKernelGGL(wgrad_edge<true>, dim3(n_chunks_), dim3(256), 0, st, g, a + NHID * AS, dy_.as<float>(),
                       part);
    hipLaunchKernelGGL(reduce_chunks, dim3(blocks(TAPS * C)), dim3(256), 0, st, part, n_chunks_, TAPS * C, TAPS * C,
                       gr + L.wl);
    hipLaunchKernelGGL(bias_grad, dim3(n_chunks_), dim3(256), 0, st, g, (const __bf16*)nullptr, dy_.as<float>(), part);
    hipLaunchKernelGGL(reduce_chunks, dim3(1), dim3(256), 0, st, part, n_chunks_, C, 1, gr + L.bl);
    __bf16* dz_cm = cm_ ? gcm_a_.as<__bf16>() : nullptr;
    __bf16* other_cm = cm_ ? gcm_b_.as<__bf16>() : nullptr;
    hipLaunchKernelGGL(back_last, dim3(blocks(M_)), dim3(256), 0, st, g, dy_.as<float>(), p + L.wl, a + NHID * AS,
                       grad_a_.as<__bf16>(), dz_cm);
    __bf16* dz = grad_a_.as<__bf16>();
    __bf16* other = grad_b_.as<__bf16>();
    for (int l = NHID - 1; l >= 0; l--) {
        // layer l maps a[l] -> a[l + 1]; dz = dL/d(pre-activation of a[l + 1])
        if (cm_) {
            hipLaunchKernelGGL(wgrad_hidden_cm, dim3(TAPS / 3 * 8 * ((n_hchunks_ + 7) / 8)), dim3(256), 0, st, g,
                               n_hchunks_, dz_cm, acts_cm_.as<const __bf16>() + l * AS, part);
            hipLaunchKernelGGL(reduce_chunks, dim3(blocks(TAPS * C * C)), dim3(256), 0, st, part, n_hchunks_,
                               TAPS * C * C, TAPS * C * C, gr + L.wh + l * HSZ);
        } else {
            hipLaunchKernelGGL(wgrad_hidden, dim3(n_chunks_, TAPS), dim3(256), 0, st, g, dz, a + l * AS, part);
            hipLaunchKernelGGL(reduce_chunks, dim3(blocks(TAPS * C * C)), dim3(256), 0, st, part, n_chunks_,
                               TAPS * C * C, TAPS * C * C, gr + L.wh + l * HSZ);
        }
        hipLaunchKernelGGL(bias_grad, dim3(n_chunks_), dim3(256), 0, st, g, dz, (const float*)nullptr, part);
        hipLaunchKernelGGL(reduce_chunks, dim3(1), dim3(256), 0, st, part, n_chunks_, C, C,
                           gr + L.wh + l * HSZ + TAPS * C * C);
        // the next lower layer's dz, with its channel-major copy unless it is the first layer's
        conv_hidden_launch(dz, wb_.as<__bf16>() + (size_t)l * TAPS * C * C, nullptr, a + l * AS, 0, other,
                           l > 0 ? other_cm : nullptr);
        std::swap(dz, other);
        std::swap(dz_cm, other_cm);
    }
    // first layer
    hipLaunchKernelGGL(wgrad_edge<false>, dim3(n_chunks_, n_in_), dim3(256), 0, st, g, dz, x_.as<float>(), part);
    hipLaunchKernelGGL(reduce_chunks, dim3(blocks(n_in_ * TAPS * C)), dim3(256), 0, st, part, n_chunks_,
                       n_in_ * TAPS * C, n_in_ * TAPS * C, gr + L.w0);
    hipLaunchKernelGGL(bias_grad, dim3(n_chunks_), dim3(256), 0, st, g, dz, (const float*)nullptr, part);
    hipLaunchKernelGGL(reduce_chunks, dim3(1), dim3(256), 0, st, part, n_chunks_, C, C, gr + L.b0);
    FTS_HIP_CHECK(hipGetLastError());
    double loss = 0.0;
    FTS_HIP_CHECK(hipMemcpyAsync(&loss, loss_.get(), sizeof(double), hipMemcpyDeviceToHost, st));
    FTS_HIP_CHECK(hipStreamSynchronize(st));
    return loss / M_;
}

double Cnn::train_step(const double* w_minus, const double* w_plus_target, bool on_device, double lr) {
    const double loss = gradients(w_minus, w_plus_target, on_device);
    step_++;
    version_++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)step_), c2 = 1.f - std::pow(b2, (float)step_);
    hipLaunchKernelGGL(adam_step, dim3(blocks(n_params_)), dim3(256), 0, ctx_->stream, params_.as<float>(),
                       grads_.as<float>(), adam_m_.as<float>(), adam_v_.as<float>(), n_params_, (float)lr, b1, b2,
                       1e-8f, c1, c2);
    FTS_HIP_CHECK(hipGetLastError());
    refresh_weights();
    FTS_HIP_CHECK(hipStreamSynchronize(ctx_->stream));
    return loss;
}

}  // namespace fts
