// Learned directional proposal: network kernels on the matrix cores and the
// dl_hip_guide_* entry points.
//
// Every GEMM of the 96-64-64-64-64-48 network runs as 32x32x16 bfloat16 MFMA
// tiles with float32 accumulation (v_mfma_f32_32x32x16_bf16).  Operand lane
// layout (checked on the device): lane l holds A[l%32][(l/32)*8 + 0..7] and
// B[(l/32)*8 + 0..7][l%32]; accumulator register r of lane l is
// D[8*(r/4) + 4*(l/32) + (r%4)][l%32].
//
//   forward     a wave takes 64 rows (two 32-row blocks); activations live
//               in the wave's LDS rows (bfloat16, in place layer by layer);
//               weights are read from global memory (43 KB, L2-resident) as
//               B operands from W[out][in]
//   backward    dW = g^T act (tiles of dW per wave, batch chunks reduced by
//               float32 atomics), dAct = g W (B operands from W^T), ReLU
//               masks from the stored bfloat16 activations
//   Adam        float32 master weights; the bfloat16 copies (W and W^T) are
//               refreshed in the same kernel
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../render_sizes.h"
#include "backend.h"
#include "guide.h"

namespace dl {

namespace {

// Network evaluation / training forward on float32 inputs x (n x 96).
__global__ __launch_bounds__(kGWaves * 64) void k_mlp_forward(Net net, const float* __restrict__ x, uint32_t n,
                                                             float* __restrict__ out, int keep, uint16_t* act0,
                                                             uint16_t* act1, uint16_t* act2, uint16_t* act3,
                                                             uint16_t* act4) {
    __shared__ uint16_t s_act[kGWaves][kGRowsPerWave * kGRowStride];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint16_t* s = s_act[wave];
    uint16_t* const act[5] = {act0, act1, act2, act3, act4};
    for (uint32_t tile = blockIdx.x * kGWaves + wave; (size_t)tile * kGRowsPerWave < n; tile += gridDim.x * kGWaves) {
        const uint32_t row0 = tile * kGRowsPerWave;
        const uint32_t row = row0 + lane;
        float f[8];
        for (int q = 0; q < kGIn / 8; q++) {
            for (int j = 0; j < 8; j++) f[j] = row < n ? x[(size_t)row * kGIn + q * 8 + j] : 0.0f;
            *reinterpret_cast<bf16x8*>(s + lane * kGRowStride + q * 8) = pack8(f);
        }
        forward_rows(net, s, row0, n, keep != 0, act, [&](int r, int c, float v) {
            if (row0 + r < n) out[(size_t)(row0 + r) * kGOut + c] = v;
        });
    }
}

// Raw outputs -> mixture parameters (n x 40).
__global__ void k_mlp_decode(const float* __restrict__ out, uint32_t n, float* __restrict__ mix) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float o[kGOut], m[kGMix];
        for (int j = 0; j < kGOut; j++) o[j] = out[(size_t)i * kGOut + j];
        guide_decode(o, m);
        for (int j = 0; j < kGMix; j++) mix[(size_t)i * kGMix + j] = m[j];
    }
}

// Sum of a 256-thread block's values added to *dst with one atomic.
__device__ __forceinline__ void block_sum_add(float v, float* dst) {
    __shared__ float s_part[4];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(dst, s_part[0] + s_part[1] + s_part[2] + s_part[3]);
}

__global__ __launch_bounds__(256) void k_sum(const float* __restrict__ c, uint32_t n, float* __restrict__ total) {
    float acc = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += c[i];
    block_sum_add(acc, total);
}

// dLoss/dout for loss = sum(-c log q(d)) / sum(c); loss accumulated in *loss.
__global__ __launch_bounds__(256) void k_loss_grad(const float* __restrict__ out, const float* __restrict__ dirs,
                            const float* __restrict__ c, uint32_t n, const float* __restrict__ total,
                            float* __restrict__ g, float* __restrict__ loss) {
    const float invC = 1.0f / fmaxf(*total, 1e-30f);
    float lacc = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float o[kGOut];
        for (int j = 0; j < kGOut; j++) o[j] = out[(size_t)i * kGOut + j];
        float mix[kGMix];
        guide_decode(o, mix);
        const vec3 d = v3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        float f[kGLobes], dt[kGLobes], q = 0.0f;
        for (int j = 0; j < kGLobes; j++) {
            const vec3 mu = v3(mix[5 * j + 2], mix[5 * j + 3], mix[5 * j + 4]);
            dt[j] = dot(mu, d);
            f[j] = vmf_pdf(mix[5 * j + 1], mu, d);
            q += mix[5 * j] * f[j];
        }
        const float ci = c[i];
        const float qs = fmaxf(q, 1e-30f);
        lacc += -ci * logf(qs) * invC;
        const float gq = -ci / qs * invC;
        float* gi = g + (size_t)i * kGOut;
        for (int j = 0; j < kGLobes; j++) {
            const float w = mix[5 * j], k = mix[5 * j + 1];
            const float wf = w * f[j];
            gi[5 * j] = gq * (wf - w * q);
            const float sg = 1.0f / (1.0f + expf(-o[5 * j + 1]));
            const float em = -expm1f(-2.0f * k);
            const float dlogf = 1.0f / k - 2.0f * expf(-2.0f * k) / em + (dt[j] - 1.0f);
            gi[5 * j + 1] = gq * wf * dlogf * k * (kGKappaHi - kGKappaLo) * sg * (1.0f - sg);
            const float mx = o[5 * j + 2], my = o[5 * j + 3], mz = o[5 * j + 4];
            const float nm = fmaxf(sqrtf(mx * mx + my * my + mz * mz), 1e-6f);
            const float s = gq * wf * k / nm;
            gi[5 * j + 2] = s * (d.x - dt[j] * mix[5 * j + 2]);
            gi[5 * j + 3] = s * (d.y - dt[j] * mix[5 * j + 3]);
            gi[5 * j + 4] = s * (d.z - dt[j] * mix[5 * j + 4]);
        }
        for (int j = kGMix; j < kGOut; j++) gi[j] = 0.0f;
    }
    block_sum_add(lacc, loss);
}

// Weight and bias gradients of one layer without atomics: block (tile o,
// tile i, chunk of kChunk rows) of kDwWaves waves, each wave accumulating
// g^T act over its share of the chunk on the matrix cores; the waves' tiles
// are summed in LDS and the block writes its chunk's partial dW tile (and,
// for i-tile 0, the partial bias sums of its outputs) to part[chunk]; then
// k_dw_reduce adds the chunks to the gradient in a fixed order, so a
// training step is deterministic.
constexpr int kChunk = 1024, kDwWaves = 4;
__global__ __launch_bounds__(kDwWaves * 64) void k_dw(const float* __restrict__ g, int gw,
                                                      const uint16_t* __restrict__ act, int K, uint32_t n,
                                                      float* __restrict__ part) {
    __shared__ float s_acc[kDwWaves][16][64];
    __shared__ float s_b[kDwWaves][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ot = blockIdx.x, it = blockIdx.y;
    const uint32_t chunk = blockIdx.z;
    const uint32_t r0 = chunk * kChunk + wave * (kChunk / kDwWaves);
    const uint32_t r1 = min(r0 + kChunk / kDwWaves, n);
    const int o = ot * 32 + (lane & 31), i = it * 32 + (lane & 31);
    f32x16 acc = {0};
    float bsum = 0.0f;
    for (uint32_t kb = r0; kb < r1; kb += 16) {
        float av[8];
        bf16x8 b;
        for (int j = 0; j < 8; j++) {
            const uint32_t srow = kb + (lane >> 5) * 8 + j;
            const bool ok = srow < r1;
            av[j] = ok && o < gw ? g[(size_t)srow * gw + o] : 0.0f;
            b[j] = ok ? (short)act[(size_t)srow * K + i] : (short)0;
            bsum += av[j];
        }
        acc = mfma(pack8(av), b, acc);
    }
#pragma unroll
    for (int r = 0; r < 16; r++) s_acc[wave][r][lane] = acc[r];
    s_b[wave][lane] = bsum;
    __syncthreads();
    if (wave != 0) return;
    const size_t stride = (size_t)gw * K + gw;
    float* dst = part + (size_t)chunk * stride;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        float v = 0.0f;
        for (int w = 0; w < kDwWaves; w++) v += s_acc[w][r][lane];
        const int orow = ot * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (orow < gw) dst[(size_t)orow * K + i] = v;
    }
    if (it == 0 && lane < 32 && o < gw) {
        float v = 0.0f;
        for (int w = 0; w < kDwWaves; w++) v += s_b[w][lane] + s_b[w][lane + 32];
        dst[(size_t)gw * K + o] = v;
    }
}

// grad[e] += sum over chunks of part[chunk][e], e over the layer's weights
// then biases (their order in the parameter vector).  Block: 64 elements x 4
// groups of chunks (group q sums chunks q, q + 4, ...); the groups are added
// in a fixed order.
__global__ __launch_bounds__(256) void k_dw_reduce(const float* __restrict__ part, uint32_t chunks, uint32_t cnt,
                                                   float* __restrict__ grad) {
    __shared__ float s_sum[4][64];
    const uint32_t e = blockIdx.x * 64 + (threadIdx.x & 63);
    const uint32_t q = threadIdx.x >> 6;
    float v = 0.0f;
    if (e < cnt)
        for (uint32_t c = q; c < chunks; c += 4) v += part[(size_t)c * cnt + e];
    s_sum[q][threadIdx.x & 63] = v;
    __syncthreads();
    if (q == 0 && e < cnt) grad[e] += (s_sum[0][threadIdx.x] + s_sum[1][threadIdx.x]) + (s_sum[2][threadIdx.x] + s_sum[3][threadIdx.x]);
}

// g_prev = (g W) * (act > 0): rows of 64 per wave; K = gw (64 or 48), out
// width 64 (layer input).
__global__ __launch_bounds__(kGWaves * 64) void k_dact(Net net, int l, const float* __restrict__ g, int gw,
                                                      const uint16_t* __restrict__ act, uint32_t n,
                                                      float* __restrict__ gprev) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint16_t* WT = net.wt + l * kGWSlot;  // [in = 64][64]
    for (uint32_t tile = blockIdx.x * kGWaves + wave; (size_t)tile * kGRowsPerWave < n; tile += gridDim.x * kGWaves) {
        const uint32_t row0 = tile * kGRowsPerWave;
        for (int rb = 0; rb < 2; rb++) {
            f32x16 acc0 = {0}, acc1 = {0};
            const uint32_t arow = row0 + rb * 32 + (lane & 31);
            for (int kt = 0; kt < gw / 16; kt++) {
                float av[8];
                for (int j = 0; j < 8; j++)
                    av[j] = arow < n ? g[(size_t)arow * gw + kt * 16 + (lane >> 5) * 8 + j] : 0.0f;
                const bf16x8 a = pack8(av);
                const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(WT + (lane & 31) * 64 + kt * 16 + (lane >> 5) * 8);
                const bf16x8 b1 =
                    *reinterpret_cast<const bf16x8*>(WT + (32 + (lane & 31)) * 64 + kt * 16 + (lane >> 5) * 8);
                acc0 = mfma(a, b0, acc0);
                acc1 = mfma(a, b1, acc1);
            }
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const uint32_t row = row0 + rb * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (row >= n) continue;
                const int c0 = lane & 31, c1 = 32 + (lane & 31);
                const bool m0 = act[(size_t)row * 64 + c0] != 0 && !(act[(size_t)row * 64 + c0] & 0x8000);
                const bool m1 = act[(size_t)row * 64 + c1] != 0 && !(act[(size_t)row * 64 + c1] & 0x8000);
                gprev[(size_t)row * 64 + c0] = m0 ? acc0[r] : 0.0f;
                gprev[(size_t)row * 64 + c1] = m1 ? acc1[r] : 0.0f;
            }
        }
    }
}

// Adam on the float32 masters; refreshes the bfloat16 W, W^T and biases.
__global__ void k_adam(float* __restrict__ theta, float* __restrict__ m, float* __restrict__ v,
                       float* __restrict__ grad, float lr, float c1, float c2, int apply, uint16_t* __restrict__ w,
                       uint16_t* __restrict__ wt, float* __restrict__ bias) {
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < kGParams; idx += gridDim.x * blockDim.x) {
        float th = theta[idx];
        if (apply) {
            const float gr = grad[idx];
            const float mm = 0.9f * m[idx] + 0.1f * gr;
            const float vv = 0.999f * v[idx] + 0.001f * gr * gr;
            m[idx] = mm;
            v[idx] = vv;
            th = th - lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
            theta[idx] = th;
            grad[idx] = 0.0f;
        }
        int l = 4;
        while (l > 0 && idx < glayer_w(l)) l--;
        const int K = glayer_in(l), off = idx - glayer_w(l);
        if (off < glayer_out(l) * K) {
            const int o = off / K, i = off % K;
            w[l * kGWSlot + o * K + i] = f2bf(th);
            wt[l * kGWSlot + i * 64 + o] = f2bf(th);
        } else {
            bias[l * 64 + (off - glayer_out(l) * K)] = th;
        }
    }
}

// Training weight of one render record: c = L_i(d) / p(d), the incident
// radiance estimated from the path's final radiance as (L_final - L_before)
// / T (T: throughput after the vertex, per channel, averaged over the
// channels with T > 0), divided by the density d was sampled from.  Zero
// for an empty slot, a non-finite estimate or a non-finite direction.
__device__ __forceinline__ float record_weight(const float* __restrict__ q, const float4* __restrict__ rad) {
    if (q[kRecValid] == 0.0f) return 0.0f;
    const uint32_t p = __float_as_uint(q[kRecPath]);
    const float4 rf = rad[p];
    const float fin[3] = {rf.x, rf.y, rf.z};
    float sum = 0.0f;
    int cnt = 0;
    for (int k = 0; k < 3; k++) {
        const float t = q[kRecThr + k];
        if (t > 0.0f) {
            sum += (fin[k] - q[kRecRad + k]) / t;
            cnt++;
        }
    }
    const float pdf = q[kRecPdf];
    float w = (cnt > 0 && pdf > 0.0f) ? (sum / (float)cnt) / pdf : 0.0f;
    if (!(w >= 0.0f) || !isfinite(w)) w = 0.0f;
    if (!isfinite(q[kRecDir]) || !isfinite(q[kRecDir + 1]) || !isfinite(q[kRecDir + 2])) w = 0.0f;
    return w;
}

// Batch row r from record q with weight w (features zeroed when w == 0).
__device__ __forceinline__ void write_batch_row(const float* __restrict__ q, float w, uint32_t r, float* __restrict__ x,
                                                float* __restrict__ d, float* __restrict__ c) {
    const vec3 pn = v3(q[kRecPos], q[kRecPos + 1], q[kRecPos + 2]);
    const vec3 wo = v3(q[kRecWo], q[kRecWo + 1], q[kRecWo + 2]);
    const vec3 nn = v3(q[kRecN], q[kRecN + 1], q[kRecN + 2]);
    for (int ch = 0; ch < kGIn / 8; ch++) {
        float f[8];
        if (w > 0.0f)
            guide_feature_chunk(pn, wo, nn, q[kRecRough], ch, f);
        else
            for (int j = 0; j < 8; j++) f[j] = 0.0f;
        for (int j = 0; j < 8; j++) x[(size_t)r * kGIn + ch * 8 + j] = f[j];
    }
    const bool ok = w > 0.0f;
    d[3 * r] = ok ? q[kRecDir] : 0.0f;
    d[3 * r + 1] = ok ? q[kRecDir + 1] : 0.0f;
    d[3 * r + 2] = ok ? q[kRecDir + 2] : 1.0f;
    c[r] = w;
}

// ---- resampling of records in proportion to their weights ----------------
// The batch gradient sum_r c_r grad(-log q(d_r)) / sum_r c_r equals in
// expectation the mean of grad(-log q(d)) over records drawn with
// probability c_r / sum c; drawing them (systematic resampling over the
// weights' prefix sums) gives every batch row weight 1, so one batch spends
// its rows on the records that carry radiance.
constexpr int kScanThreads = 256, kScanItems = 4, kScanTile = kScanThreads * kScanItems;

// Record weights, each capped at `share` of *total when total is given (a
// rare path of enormous weight would otherwise fill whole batches with
// copies of itself).
__global__ void k_record_weights(const float* __restrict__ rec, uint32_t R, const float4* __restrict__ rad,
                                 const float* __restrict__ total, float share, float* __restrict__ w) {
    const float cap = total ? share * *total : INFINITY;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < R; r += gridDim.x * blockDim.x)
        w[r] = fminf(record_weight(rec + (size_t)r * kGRecFloats, rad), cap);
}

// In-place inclusive prefix sums of each kScanTile-element tile of w; the
// tile totals to tile_sum.  Fixed summation order (deterministic).
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(float* __restrict__ w, uint32_t R,
                                                             float* __restrict__ tile_sum) {
    __shared__ float s_wave[kScanThreads / 64];
    const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float v[kScanItems];
    float run = 0.0f;
    for (int k = 0; k < kScanItems; k++) {
        run += base + k < R ? w[base + k] : 0.0f;
        v[k] = run;
    }
    float x = run;
    for (int off = 1; off < 64; off <<= 1) {
        const float y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    float woff = 0.0f;
    for (int j = 0; j < wave; j++) woff += s_wave[j];
    const float excl = woff + (x - run);
    for (int k = 0; k < kScanItems; k++)
        if (base + k < R) w[base + k] = excl + v[k];
    if (threadIdx.x == kScanThreads - 1) tile_sum[blockIdx.x] = woff + x;
}

// Exclusive prefix of the tile totals: tile_off[b] = sum of tiles < b,
// tile_off[nt] = the grand total.  One block; each thread sums a run of
// tiles serially, then one thread prefixes the runs.
__global__ __launch_bounds__(kScanThreads) void k_scan_top(const float* __restrict__ tile_sum, uint32_t nt,
                                                           float* __restrict__ tile_off) {
    __shared__ float s_run[kScanThreads];
    const uint32_t per = (nt + kScanThreads - 1) / kScanThreads;
    const uint32_t b0 = threadIdx.x * per;
    float acc = 0.0f;
    for (uint32_t b = b0; b < b0 + per && b < nt; b++) acc += tile_sum[b];
    s_run[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.0f;
        for (int j = 0; j < kScanThreads; j++) {
            const float a = s_run[j];
            s_run[j] = t;
            t += a;
        }
        tile_off[nt] = t;
    }
    __syncthreads();
    float t = s_run[threadIdx.x];
    for (uint32_t b = b0; b < b0 + per && b < nt; b++) {
        tile_off[b] = t;
        t += tile_sum[b];
    }
}

// Batch row j <- the record whose cumulative-weight interval holds
// (j + u) / B of the total (first r with G(r) > t, G the inclusive prefix
// sum, G(r) = tile_off[r / tile] + w_scanned[r]).
__global__ void k_resample(const float* __restrict__ rec, uint32_t R, const float* __restrict__ cdf,
                           const float* __restrict__ tile_off, uint32_t nt, uint32_t B, float u,
                           float* __restrict__ x, float* __restrict__ d, float* __restrict__ c) {
    const float total = tile_off[nt];
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < B; j += gridDim.x * blockDim.x) {
        if (!(total > 0.0f) || !isfinite(total)) {
            write_batch_row(rec, 0.0f, j, x, d, c);
            continue;
        }
        const float t = ((float)j + u) * (total / (float)B);
        uint32_t lo = 0, hi = R;  // first r in [lo, hi) with G(r) > t, or R
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            const float g = tile_off[mid / kScanTile] + cdf[mid];
            if (g > t)
                hi = mid;
            else
                lo = mid + 1;
        }
        const uint32_t r = lo < R ? lo : R - 1;
        const float g_prev = r == 0 ? 0.0f : tile_off[(r - 1) / kScanTile] + cdf[r - 1];
        const float g_r = tile_off[r / kScanTile] + cdf[r];
        write_batch_row(rec + (size_t)r * kGRecFloats, g_r > g_prev ? 1.0f : 0.0f, j, x, d, c);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// the device bytes the render service reserves for a guided render
static_assert(kGuideBatch == dl_sizes::kGuideBatchRows && kChunk == (int)dl_sizes::kGuideChunkRows &&
                  kScanTile == (int)dl_sizes::kGuideScanTile && kGParams == dl_sizes::kGuideParams &&
                  kGWSlot == dl_sizes::kGuideWSlot,
              "guided render sizes");
static_assert(dl_sizes::kGuideNetBytes == 4ull * kGParams * sizeof(float) + 5 * 64 * sizeof(float) +
                                              2ull * 5 * kGWSlot * sizeof(uint16_t) + 4 * sizeof(float),
              "guide_init allocations");
static_assert(dl_sizes::kGuideTrainBytes ==
                  (uint64_t)kGuideBatch * ((glayer_in(0) + 4 * glayer_in(1)) * sizeof(uint16_t) +
                                           (kGOut + 2 * 64) * sizeof(float) + (kGIn + 3 + 1) * sizeof(float)) +
                      (uint64_t)(kGuideBatch / kChunk) * (64 * 96 + 64) * sizeof(float),
              "ensure_train and training batch allocations");

struct GuideState {
    // training batch built from render records
    uint32_t batch_cap = 0;
    float *bx = nullptr, *bd = nullptr, *bc = nullptr;
    // record weights (prefix-summed in place) and their tile offsets
    uint32_t rw_cap = 0;
    float *rw = nullptr, *tile_sum = nullptr, *tile_off = nullptr;
    float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *bias = nullptr;
    uint16_t *w = nullptr, *wt = nullptr;
    int step = 0;
    float lr = 1e-2f;
    // training buffers
    uint32_t cap = 0;
    uint16_t* act[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    float *out = nullptr, *g0 = nullptr, *g1 = nullptr, *scalars = nullptr;
    float* part = nullptr;  // per-chunk partial weight/bias gradients of one layer
    std::vector<void*> allocs;
};

namespace {

template <typename T>
dl_status galloc(Ctx* c, GuideState* g, T** p, size_t count) {
    void* q = nullptr;
    DL_HIP_CHECK(c, dmalloc(c, &q, std::max<size_t>(1, count) * sizeof(T)));
    g->allocs.push_back(q);
    *p = (T*)q;
    return DL_OK;
}

#define DL_TRY(expr)                \
    do {                            \
        dl_status s_ = (expr);      \
        if (s_ != DL_OK) return s_; \
    } while (0)

Net net_of(const GuideState* g) { return Net{g->w, g->wt, g->bias}; }

// Deterministic initial weights: He-uniform (last layer x0.1), zero biases,
// uniforms from splitmix64 (Steele et al. 2014).
void init_weights(uint64_t seed, std::vector<float>& th) {
    th.assign(kGParams, 0.0f);
    uint64_t x = seed;
    auto next = [&x]() {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)((z >> 40) * (1.0 / 16777216.0));
    };
    for (int l = 0; l < 5; l++) {
        const float a = std::sqrt(6.0f / glayer_in(l)) * (l == 4 ? 0.1f : 1.0f);
        for (int k = 0; k < glayer_out(l) * glayer_in(l); k++) th[glayer_w(l) + k] = (2.0f * next() - 1.0f) * a;
    }
}

dl_status ensure_train(Ctx* c, GuideState* g, uint32_t n) {
    if (g->cap >= n) return DL_OK;
    for (int l = 0; l < 5; l++)
        if (g->act[l]) {
            dfree(c, g->act[l]);
            g->allocs.erase(std::find(g->allocs.begin(), g->allocs.end(), (void*)g->act[l]));
        }
    for (float** p : {&g->out, &g->g0, &g->g1, &g->part})
        if (*p) {
            dfree(c, *p);
            g->allocs.erase(std::find(g->allocs.begin(), g->allocs.end(), (void*)*p));
        }
    g->cap = n;
    for (int l = 0; l < 5; l++) DL_TRY(galloc(c, g, &g->act[l], (size_t)n * glayer_in(l)));
    DL_TRY(galloc(c, g, &g->out, (size_t)n * kGOut));
    DL_TRY(galloc(c, g, &g->g0, (size_t)n * 64));
    DL_TRY(galloc(c, g, &g->g1, (size_t)n * 64));
    // largest layer: 64 x 96 weights + 64 biases per chunk
    DL_TRY(galloc(c, g, &g->part, (size_t)((n + kChunk - 1) / kChunk) * (64 * 96 + 64)));
    return DL_OK;
}

}  // namespace

void free_guide(Ctx* c) {
    GuideState* g = (GuideState*)c->guide;
    if (!g) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void* p : g->allocs) (void)dfree(c, p);
    delete g;
    c->guide = nullptr;
}

GuideState* guide_state(Ctx* c) { return (GuideState*)c->guide; }
const float* guide_bias(GuideState* g) { return g->bias; }

dl_status guide_init(Ctx* c, uint64_t seed, float lr, const float* params) {
    free_guide(c);
    GuideState* g = new GuideState();
    c->guide = g;
    g->lr = lr;
    DL_TRY(galloc(c, g, &g->theta, kGParams));
    DL_TRY(galloc(c, g, &g->m, kGParams));
    DL_TRY(galloc(c, g, &g->v, kGParams));
    DL_TRY(galloc(c, g, &g->grad, kGParams));
    DL_TRY(galloc(c, g, &g->bias, 5 * 64));
    DL_TRY(galloc(c, g, &g->w, 5 * kGWSlot));
    DL_TRY(galloc(c, g, &g->wt, 5 * kGWSlot));
    DL_TRY(galloc(c, g, &g->scalars, 4));
    std::vector<float> th;
    if (params)
        th.assign(params, params + kGParams);
    else
        init_weights(seed, th);
    DL_HIP_CHECK(c, hipMemcpyAsync(g->theta, th.data(), kGParams * sizeof(float), hipMemcpyHostToDevice, c->stream));
    DL_HIP_CHECK(c, hipMemsetAsync(g->m, 0, kGParams * sizeof(float), c->stream));
    DL_HIP_CHECK(c, hipMemsetAsync(g->v, 0, kGParams * sizeof(float), c->stream));
    DL_HIP_CHECK(c, hipMemsetAsync(g->grad, 0, kGParams * sizeof(float), c->stream));
    DL_HIP_CHECK(c, hipMemsetAsync(g->w, 0, 5 * kGWSlot * sizeof(uint16_t), c->stream));
    DL_HIP_CHECK(c, hipMemsetAsync(g->wt, 0, 5 * kGWSlot * sizeof(uint16_t), c->stream));
    DL_HIP_CHECK(c, hipMemsetAsync(g->bias, 0, 5 * 64 * sizeof(float), c->stream));
    k_adam<<<(kGParams + 255) / 256, 256, 0, c->stream>>>(g->theta, g->m, g->v, g->grad, 0.f, 1.f, 1.f, 0, g->w,
                                                          g->wt, g->bias);
    DL_HIP_CHECK(c, hipGetLastError());
    return DL_OK;
}

// Forward pass over n float32 input rows (device), raw outputs to out.
dl_status guide_forward(Ctx* c, const float* x, uint32_t n, float* out, bool keep) {
    GuideState* g = (GuideState*)c->guide;
    if (n == 0) return DL_OK;
    const uint32_t tiles = (n + kGRowsPerWave - 1) / kGRowsPerWave;
    const int grid = (int)std::max<uint32_t>(1, std::min<uint32_t>((tiles + kGWaves - 1) / kGWaves, (uint32_t)c->num_cus * 4));
    k_mlp_forward<<<grid, kGWaves * 64, 0, c->stream>>>(net_of(g), x, n, out, keep ? 1 : 0, g->act[0], g->act[1],
                                                       g->act[2], g->act[3], g->act[4]);
    DL_HIP_CHECK(c, hipGetLastError());
    return DL_OK;
}

// One training step on n records (device arrays): gradient, then Adam.
dl_status guide_train(Ctx* c, const float* x, const float* dirs, const float* cw, uint32_t n, bool apply) {
    GuideState* g = (GuideState*)c->guide;
    DL_TRY(ensure_train(c, g, n));
    DL_HIP_CHECK(c, hipMemsetAsync(g->scalars, 0, 4 * sizeof(float), c->stream));
    DL_TRY(guide_forward(c, x, n, g->out, true));
    const int eg = std::max(1, std::min<int>((int)((n + 255) / 256), c->num_cus * 8));
    k_sum<<<std::min(eg, 64), 256, 0, c->stream>>>(cw, n, g->scalars);
    k_loss_grad<<<eg, 256, 0, c->stream>>>(g->out, dirs, cw, n, g->scalars, g->g0, g->scalars + 1);
    float* gcur = g->g0;
    float* gnext = g->g1;
    int gw = kGOut;
    const uint32_t chunks = (n + kChunk - 1) / kChunk;
    for (int l = 4; l >= 0; l--) {
        const int K = glayer_in(l);
        const uint32_t cnt = (uint32_t)(gw * K + gw);
        k_dw<<<dim3((gw + 31) / 32, K / 32, chunks), kDwWaves * 64, 0, c->stream>>>(gcur, gw, g->act[l], K, n,
                                                                                    g->part);
        k_dw_reduce<<<(cnt + 63) / 64, 256, 0, c->stream>>>(g->part, chunks, cnt, g->grad + glayer_w(l));
        if (l > 0) {
            const uint32_t tiles = (n + kGRowsPerWave - 1) / kGRowsPerWave;
            const int grid = (int)std::max<uint32_t>(
                1, std::min<uint32_t>((tiles + kGWaves - 1) / kGWaves, (uint32_t)c->num_cus * 4));
            k_dact<<<grid, kGWaves * 64, 0, c->stream>>>(net_of(g), l, gcur, gw, g->act[l], n, gnext);
            std::swap(gcur, gnext);
            gw = 64;
        }
    }
    DL_HIP_CHECK(c, hipGetLastError());
    if (apply) {
        g->step++;
        const float c1 = 1.0f - std::pow(0.9f, (float)g->step), c2 = 1.0f - std::pow(0.999f, (float)g->step);
        k_adam<<<(kGParams + 255) / 256, 256, 0, c->stream>>>(g->theta, g->m, g->v, g->grad, g->lr, c1, c2, 1, g->w,
                                                              g->wt, g->bias);
        DL_HIP_CHECK(c, hipGetLastError());
    }
    return DL_OK;
}

int guide_step_count(Ctx* c) { return c->guide ? ((GuideState*)c->guide)->step : 0; }

void guide_net_ptrs(Ctx* c, const uint16_t** w, const uint16_t** wt, const float** bias) {
    GuideState* g = (GuideState*)c->guide;
    *w = g->w;
    *wt = g->wt;
    *bias = g->bias;
}

// `steps` Adam steps on batches of kGuideBatch records resampled from the R
// render records in proportion to their weights; returns the steps taken.
dl_status guide_train_records(Ctx* c, const float* records, uint32_t R, const float4* rad, int steps) {
    GuideState* g = (GuideState*)c->guide;
    if (R == 0 || steps <= 0) return DL_OK;
    const uint32_t B = kGuideBatch;
    if (g->batch_cap < B) {
        for (float** p : {&g->bx, &g->bd, &g->bc})
            if (*p) {
                dfree(c, *p);
                g->allocs.erase(std::find(g->allocs.begin(), g->allocs.end(), (void*)*p));
            }
        g->batch_cap = B;
        DL_TRY(galloc(c, g, &g->bx, (size_t)B * kGIn));
        DL_TRY(galloc(c, g, &g->bd, (size_t)B * 3));
        DL_TRY(galloc(c, g, &g->bc, B));
    }
    const uint32_t nt = (R + kScanTile - 1) / kScanTile;
    if (g->rw_cap < R) {
        for (float** p : {&g->rw, &g->tile_sum, &g->tile_off})
            if (*p) {
                dfree(c, *p);
                g->allocs.erase(std::find(g->allocs.begin(), g->allocs.end(), (void*)*p));
            }
        g->rw_cap = R;
        DL_TRY(galloc(c, g, &g->rw, R));
        DL_TRY(galloc(c, g, &g->tile_sum, nt));
        DL_TRY(galloc(c, g, &g->tile_off, nt + 1));
    }
    const int eg = std::max(1, std::min<int>((int)((R + 255) / 256), c->num_cus * 8));
    // weights, their total, then the capped weights and their prefix sums
    k_record_weights<<<eg, 256, 0, c->stream>>>(records, R, rad, nullptr, 0.0f, g->rw);
    k_scan_tiles<<<nt, kScanThreads, 0, c->stream>>>(g->rw, R, g->tile_sum);
    k_scan_top<<<1, kScanThreads, 0, c->stream>>>(g->tile_sum, nt, g->tile_off);
    k_record_weights<<<eg, 256, 0, c->stream>>>(records, R, rad, g->tile_off + nt, kGuideMaxShare, g->rw);
    k_scan_tiles<<<nt, kScanThreads, 0, c->stream>>>(g->rw, R, g->tile_sum);
    k_scan_top<<<1, kScanThreads, 0, c->stream>>>(g->tile_sum, nt, g->tile_off);
    DL_HIP_CHECK(c, hipGetLastError());
    const int bg = std::max(1, std::min<int>((int)((B + 255) / 256), c->num_cus * 8));
    for (int s = 0; s < steps; s++) {
        // stratum offset of this step: golden-ratio sequence over the steps
        const double u = std::fmod(0.5 + 0.6180339887498949 * (double)(g->step + 1), 1.0);
        k_resample<<<bg, 256, 0, c->stream>>>(records, R, g->rw, g->tile_off, nt, B, (float)u, g->bx, g->bd, g->bc);
        DL_HIP_CHECK(c, hipGetLastError());
        DL_TRY(guide_train(c, g->bx, g->bd, g->bc, B, true));
    }
    return DL_OK;
}

dl_status guide_decode_rows(Ctx* c, const float* out, uint32_t n, float* mix) {
    if (n == 0) return DL_OK;
    k_mlp_decode<<<std::max(1, std::min<int>((int)((n + 255) / 256), c->num_cus * 8)), 256, 0, c->stream>>>(out, n,
                                                                                                        mix);
    DL_HIP_CHECK(c, hipGetLastError());
    return DL_OK;
}

}  // namespace dl

using namespace dl;

extern "C" {

dl_status dl_hip_guide_init(void* h, uint64_t seed, float lr, const float* params) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    (void)hipSetDevice(c->device);
    return guide_init(c, seed, lr, params);
}

dl_status dl_hip_guide_params(void* h, float* params, float* grads) {
    Ctx* c = (Ctx*)h;
    if (!c) return DL_ERR_INVALID;
    GuideState* g = (GuideState*)c->guide;
    if (!g) return set_error(c, DL_ERR_STATE, "guide not initialised");
    (void)hipSetDevice(c->device);
    if (params)
        DL_HIP_CHECK(c, hipMemcpyAsync(params, g->theta, kGParams * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (grads)
        DL_HIP_CHECK(c, hipMemcpyAsync(grads, g->grad, kGParams * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return DL_OK;
}

dl_status dl_hip_guide_eval(void* h, const float* x, uint32_t n, float* out, float* mix) {
    Ctx* c = (Ctx*)h;
    if (!c || !x || !out) return DL_ERR_INVALID;
    if (!c->guide) return set_error(c, DL_ERR_STATE, "guide not initialised");
    (void)hipSetDevice(c->device);
    DL_TRY(guide_forward(c, x, n, out, false));
    if (mix) DL_TRY(guide_decode_rows(c, out, n, mix));
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return DL_OK;
}

dl_status dl_hip_guide_train_step(void* h, const float* x, const float* dirs, const float* weights, uint32_t n,
                                  int apply, float* loss) {
    Ctx* c = (Ctx*)h;
    if (!c || !x || !dirs || !weights) return DL_ERR_INVALID;
    GuideState* g = (GuideState*)c->guide;
    if (!g) return set_error(c, DL_ERR_STATE, "guide not initialised");
    if (n == 0) return set_error(c, DL_ERR_INVALID, "empty batch");
    (void)hipSetDevice(c->device);
    DL_TRY(guide_train(c, x, dirs, weights, n, apply != 0));
    if (loss) DL_HIP_CHECK(c, hipMemcpyAsync(loss, g->scalars + 1, sizeof(float), hipMemcpyDeviceToHost, c->stream));
    DL_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return DL_OK;
}

}  // extern "C"
