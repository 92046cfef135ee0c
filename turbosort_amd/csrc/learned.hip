// Learned partitioner: a sort of 4- or 8-byte keys whose output equals
// turbosort::sort (lib.rs:99-101) -- the unique ascending order of the radix
// keys -- reached by bucketing instead of by eight byte passes.
//
//   1. A stratified sample of the radix keys (0.5 % by default), sorted by
//      the device radix.
//   2. A monotone piecewise-linear CDF (kKnots knots over a logarithmic key
//      coordinate) fitted to the sample's ranks by gradient descent (Adam) in
//      one workgroup: the knot heights are normalised cumulative softplus of
//      the parameters, so every step keeps the model monotone.
//   3. The model, quantised to 32-bit fixed point, gives each key a bucket;
//      one-sweep passes over the bucket id's bytes (two up to 2^16 buckets)
//      group the keys by bucket (keys stay whole; buckets are contiguous key
//      ranges because the model is monotone).
//   4. Bucket bounds from the grouped keys; every bucket is sorted by the
//      batched segmented sort (LDS classes; buckets too long for one
//      workgroup take the device radix sort).
//
// Model error only unbalances buckets; the output is exact for any model.
// Divert's "bucket, then sort each run" (divert.rs:38-90) is the CPU analogue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "device.h"
#include "radix.h"

namespace ts {
namespace {

constexpr int kFitThreads = 1024;

// Stratified sample: stratum j of n / S positions contributes one key at a
// hashed offset, transformed to its radix key.
template <class U>
__global__ void k_learn_sample(const U* __restrict__ data, size_t n, int mode, unsigned long long* __restrict__ out,
                               size_t S) {
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  // Stratum j is [j n / S, (j + 1) n / S) in double precision (no 64-bit
  // division, no overflow of j * n); the pick inside it is the high word of
  // a hash times the stratum's length, and is clamped into the keys.
  const double step = double(n) / double(S);
  for (size_t j = g; j < S; j += gs) {
    const size_t lo = min(size_t(double(j) * step), n - 1);
    const size_t hi = min(size_t(double(j + 1) * step), n);
    const unsigned long long span = hi > lo ? (unsigned long long)(hi - lo) : 1ull;
    const size_t i = min(n - 1, lo + size_t(__umul64hi(mix64(0x5EEDull + j), span)));
    out[j] = (unsigned long long)to_key<U>(data[i], mode);
  }
}

// Model coordinate range from the sorted sample, and the training points:
// T evenly spaced sample keys, their coordinate u (in knot units) and their
// CDF target (rank + 1/2) / S.
__global__ void k_learn_points(const unsigned long long* __restrict__ sorted, size_t S, LearnedModel* model,
                               float* __restrict__ u, float* __restrict__ t, uint32_t T, uint32_t bucket_shift,
                               int linear) {
  __shared__ LearnedModel m;
  if (threadIdx.x == 0) {
    m.kmin = sorted[0];
    m.linear = uint32_t(linear);
    const unsigned long long span = sorted[S - 1] - sorted[0];
    const int bits = span ? 64 - __clzll((long long)span) : 0;
    m.shift = linear && bits > 32 ? uint32_t(bits - 32) : 0u;
    m.xhi = learned_x(sorted[S - 1], m);
    learned_scale(m.xhi, &m.cmul, &m.cshift);
    m.bucket_shift = bucket_shift;
    if (blockIdx.x == 0) {
      model->kmin = m.kmin;
      model->linear = m.linear;
      model->shift = m.shift;
      model->xhi = m.xhi;
      model->cmul = m.cmul;
      model->cshift = m.cshift;
      model->bucket_shift = bucket_shift;
    }
  }
  __syncthreads();
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t j = g; j < T; j += gs) {
    const size_t r = j * S / T;
    u[j] = float(learned_coord(sorted[r], m)) * (1.0f / float(1u << kKnotFracBits));
    t[j] = (float(r) + 0.5f) / float(S);
  }
}

__device__ __forceinline__ float softplus(float x) { return x > 20.f ? x : log1pf(expf(x)); }
__device__ __forceinline__ float sigmoid(float x) { return 1.f / (1.f + expf(-x)); }

// Inclusive prefix sum over the workgroup's kFitThreads lanes.
__device__ __forceinline__ float block_inclusive(float v, float* wsum) {
  const unsigned w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  float inc = wave_inclusive_add(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  float off = 0.f;
  for (unsigned i = 0; i < w; ++i) off += wsum[i];
  __syncthreads();
  return off + inc;
}

// Fits the knot parameters theta (kKnots, one per interval) to the points by
// minimising L = mean((F(u) - t)^2), where y_0 = 0, y_k = c_{k-1} / C,
// c_k = sum_{j<=k} softplus(theta_j), C = c_{kKnots-1}, and
// F(u) = y_i + (y_{i+1} - y_i) (u - i) with i = floor(u).
//
// dL/dy_k accumulates 2 r (1 - f) / T at the left knot and 2 r f / T at the
// right one (r = F - t, f = u - i); through the normalisation
// dL/ds_j = (sum_{k>j} g_k - sum_k g_k y_k) / C, and dL/dtheta_j =
// dL/ds_j * sigmoid(theta_j).  Adam updates theta.  grad_out receives the
// gradient at the initial theta, loss_out[0] / [1] the loss before and after
// the iters steps.  One workgroup; each lane takes runs of consecutive points
// (sorted by u), so it adds its partial gradient to LDS once per interval it
// meets.  REG: up to kFitThreads * kFitPpl points, read once into registers.
//
// With knot_first (the index of the first sample key at or past each knot,
// over a sorted sample of S keys) the parameters start from the sample's
// interval masses -- softplus(theta_k) proportional to the keys in interval
// k, plus a small floor, so the knots interpolate the sample's CDF -- and the
// optimiser refines that; loss_out[2] receives the loss of uniform knots.
// Otherwise they start from theta_io.
constexpr int kFitPpl = 16;

template <bool REG>
__global__ __launch_bounds__(kFitThreads) void k_pwl_fit(const float* __restrict__ u, const float* __restrict__ t,
                                                         uint32_t T, float* __restrict__ theta_io,
                                                         float* __restrict__ grad_out, float* __restrict__ loss_out,
                                                         int iters, float lr,
                                                         const unsigned long long* __restrict__ knot_first,
                                                         unsigned long long S) {
  static_assert(kKnots == kFitThreads, "one lane per interval");
  constexpr int kStage = 8;  // points per lane per staging chunk
  __shared__ float y[kKnots + 1];
  __shared__ float g[kKnots + 1];
  __shared__ float ws_a[kFitThreads / 64], ws_b[kFitThreads / 64], ws_c[kFitThreads / 64];
  __shared__ float loss_acc;
  __shared__ float stage_u[REG ? kStage * kFitThreads : 1], stage_t[REG ? kStage * kFitThreads : 1];
  const unsigned tid = threadIdx.x, w = tid >> 6;

  // This lane's points: REG, runs of kStage consecutive points per chunk of
  // kStage * kFitThreads; otherwise one run read from global memory.
  float ru[REG ? kFitPpl : 1], rt[REG ? kFitPpl : 1];
  const uint32_t per = (T + kFitThreads - 1) / kFitThreads;
  const uint32_t p0 = min(T, tid * per), p1 = min(T, p0 + per);
  if (REG) {
#pragma unroll
    for (int c = 0; c < kFitPpl / kStage; ++c) {
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        const uint32_t i = uint32_t(c * kStage + k) * kFitThreads + tid;
        stage_u[k * kFitThreads + tid] = i < T ? u[i] : -1.f;
        stage_t[k * kFitThreads + tid] = i < T ? t[i] : 0.f;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kStage; ++k) {
        ru[c * kStage + k] = stage_u[tid * kStage + k];
        rt[c * kStage + k] = stage_t[tid * kStage + k];
      }
      __syncthreads();
    }
  }
  // Calls f(u, t) for each of this lane's points in order.
  auto for_points = [&](auto&& f) {
    if constexpr (REG) {
#pragma unroll
      for (int k = 0; k < kFitPpl; ++k)
        if (ru[k] >= 0.f) f(ru[k], rt[k]);
    } else {
      for (uint32_t q = p0; q < p1; q += 8) {
        float uu[8], tt[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const uint32_t p = min(q + k, p1 - 1);
          uu[k] = u[p];
          tt[k] = t[p];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (q + k < p1) f(uu[k], tt[k]);
      }
    }
  };

  float theta;
  if (knot_first) {
    if (tid == 0) loss_acc = 0.f;
    __syncthreads();
    float lu = 0.f;
    for_points([&](float up, float tp) {
      const float r = up * (1.f / float(kKnots)) - tp;
      lu += r * r;
    });
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) lu += __shfl_xor(lu, d, 64);
    if ((tid & 63) == 0) atomicAdd(&loss_acc, lu);
    const float mean = float(S) / float(kKnots);
    const float sk = float(knot_first[tid + 1] - knot_first[tid]) / mean + 0.002f;
    theta = sk > 20.f ? sk : logf(expm1f(sk));
    __syncthreads();
    if (tid == 0) loss_out[2] = loss_acc / float(T);
    __syncthreads();
  } else {
    theta = theta_io[tid];
  }

  float m1 = 0.f, m2 = 0.f;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float b1t = 1.f, b2t = 1.f;
  for (int it = 0;; ++it) {
    // Knots from the parameters: one workgroup scan of softplus(theta).
    const float c_w = wave_inclusive_add(softplus(theta));
    if ((tid & 63) == 63) ws_a[w] = c_w;
    g[tid] = 0.f;
    if (tid == 0) {
      g[kKnots] = 0.f;
      loss_acc = 0.f;
    }
    __syncthreads();
    float off = 0.f, C = 0.f;
#pragma unroll
    for (int x = 0; x < kFitThreads / 64; ++x) {
      const float v = ws_a[x];
      off += x < int(w) ? v : 0.f;
      C += v;
    }
    y[tid + 1] = (off + c_w) / C;
    if (tid == 0) y[0] = 0.f;
    __syncthreads();

    // Forward and backward over this lane's points.  A lane adds each run of
    // points in one interval to LDS once; in REG mode the last run of each
    // staging chunk is first summed across the wave's lanes that end in the
    // same interval (the lanes hold consecutive points, so such lanes are
    // adjacent), so a dense interval costs one atomic per wave, not per lane.
    float loss = 0.f, gl = 0.f, gr = 0.f;
    int cur = -1;
    auto point = [&](float up, float tp) {
      const int i = min(kKnots - 1, int(up));
      const float f = up - float(i);
      if (i != cur) {
        if (cur >= 0) {
          atomicAdd(&g[cur], gl);
          atomicAdd(&g[cur + 1], gr);
        }
        cur = i;
        gl = gr = 0.f;
      }
      const float y0 = y[i], y1 = y[i + 1];
      const float r = y0 + (y1 - y0) * f - tp;
      loss += r * r;
      gl += 2.f * r * (1.f - f);
      gr += 2.f * r * f;
    };
    if constexpr (REG) {
      const unsigned lane = tid & 63;
#pragma unroll
      for (int c = 0; c < kFitPpl / kStage; ++c) {
#pragma unroll
        for (int k = 0; k < kStage; ++k)
          if (ru[c * kStage + k] >= 0.f) point(ru[c * kStage + k], rt[c * kStage + k]);
        // Segmented sum over lanes by cur (non-decreasing across the wave).
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
          const int oc = __shfl_up(cur, d, 64);
          const float ol = __shfl_up(gl, d, 64), orr = __shfl_up(gr, d, 64);
          if (lane >= unsigned(d) && oc == cur) {
            gl += ol;
            gr += orr;
          }
        }
        const int next = __shfl_down(cur, 1, 64);
        if (cur >= 0 && (lane == 63 || next != cur)) {
          atomicAdd(&g[cur], gl);
          atomicAdd(&g[cur + 1], gr);
        }
        cur = -1;
        gl = gr = 0.f;
      }
    } else {
      for_points(point);
      if (cur >= 0) {
        atomicAdd(&g[cur], gl);
        atomicAdd(&g[cur + 1], gr);
      }
    }
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) loss += __shfl_xor(loss, d, 64);
    if ((tid & 63) == 0) atomicAdd(&loss_acc, loss);
    __syncthreads();
    const float invT = 1.f / float(T);
    const float L = loss_acc * invT;
    if (it == 0 && tid == 0) loss_out[0] = L;
    if (it == iters && tid == 0) loss_out[1] = L;
    // Gradient through the normalisation: one scan gives the prefix of g over
    // the knots k <= tid + 1, its total and the total of g_k y_k.
    const float gk = g[tid + 1] * invT;  // knot tid + 1 is the right end of interval tid
    const float incl_w = wave_inclusive_add(gk);
    float dot_w = gk * y[tid + 1];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) dot_w += __shfl_xor(dot_w, d, 64);
    if ((tid & 63) == 63) {
      ws_b[w] = incl_w;
      ws_c[w] = dot_w;
    }
    __syncthreads();
    float goff = 0.f, total = 0.f, dot = 0.f;
#pragma unroll
    for (int x = 0; x < kFitThreads / 64; ++x) {
      const float v = ws_b[x];
      goff += x < int(w) ? v : 0.f;
      total += v;
      dot += ws_c[x];
    }
    const float suffix = total - (goff + incl_w) + gk;  // sum over knots k >= tid + 1, i.e. k > tid
    const float gs = (suffix - dot) / C;
    const float gt = gs * sigmoid(theta);
    if (it == 0) grad_out[tid] = gt;
    if (it == iters) break;
    // Adam.
    b1t *= b1;
    b2t *= b2;
    m1 = b1 * m1 + (1.f - b1) * gt;
    m2 = b2 * m2 + (1.f - b2) * gt * gt;
    theta -= lr * (m1 / (1.f - b1t)) / (sqrtf(m2 / (1.f - b2t)) + eps);
    __syncthreads();
  }
  theta_io[tid] = theta;
}

// The first sample index at or past each knot (k = 0 .. kKnots), by binary
// search over the sorted sample's model coordinates.
__global__ void k_knot_first(const unsigned long long* __restrict__ sorted, size_t S, const LearnedModel* model,
                             unsigned long long* __restrict__ first) {
  const unsigned k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > unsigned(kKnots)) return;
  const LearnedModel& m = *model;
  const uint32_t bound = uint32_t(k) << kKnotFracBits;
  size_t lo = 0, hi = S;  // first index with coord >= bound
  while (lo < hi) {
    const size_t mid = (lo + hi) >> 1;
    if (learned_coord(sorted[mid], m) >= bound)
      hi = mid;
    else
      lo = mid + 1;
  }
  first[k] = k == unsigned(kKnots) ? S : lo;
}

// Quantises the knots: Y_k = y_k * 2^32, non-decreasing, Y_0 = 0 and the last
// knot just below 2^32.
__global__ __launch_bounds__(kFitThreads) void k_pwl_quantize(const float* __restrict__ theta, LearnedModel* model) {
  __shared__ float wsum[kFitThreads / 64];
  __shared__ double cs[kKnots];
  const unsigned tid = threadIdx.x;
  const float c = block_inclusive(softplus(theta[tid]), wsum);
  cs[tid] = c;
  __syncthreads();
  const double C = cs[kKnots - 1];
  double y = cs[tid] / C * 4294967296.0;
  y = y < 0.0 ? 0.0 : (y > 4294967295.0 ? 4294967295.0 : y);
  if (tid == kKnots - 1) y = 4294967295.0;
  model->knots[tid + 1] = uint32_t(y);
  if (tid == 0) model->knots[0] = 0;
  __syncthreads();
  // Monotone after rounding: cumulative maximum (cs is non-decreasing, so
  // this only guards against double rounding).
  if (tid == 0) {
    uint32_t run = 0;
    for (int k = 0; k <= kKnots; ++k) {
      run = model->knots[k] > run ? model->knots[k] : run;
      model->knots[k] = run;
    }
  }
}

// Uniform knots: the untrained model (equal-width buckets over the sample's
// coordinate range).
__global__ void k_pwl_uniform(LearnedModel* model) {
  for (unsigned k = threadIdx.x; k <= unsigned(kKnots); k += blockDim.x)
    model->knots[k] = k == unsigned(kKnots) ? 0xFFFFFFFFu : uint32_t((unsigned long long)k << 32 >> 10);
}

// Bucket bounds of keys grouped by bucket: offsets[b] = first position whose
// bucket is >= b, offsets[B] = n.  Each lane takes kBoundRun consecutive keys
// and evaluates the model once per key (plus once for the key before its run).
constexpr int kBoundRun = 8;
// Binary searches (k_learn_bounds_search) replace the scan when there are at
// least this many keys per bucket.
constexpr size_t kBoundSearchRatio = 64;

template <class U>
__global__ void k_learn_bounds(const U* __restrict__ data, size_t n, int mode, const LearnedModel* __restrict__ model,
                               unsigned long long* __restrict__ offsets, uint32_t B) {
  __shared__ unsigned long long knots[kKnots];
  for (unsigned i = threadIdx.x; i < unsigned(kKnots); i += blockDim.x) knots[i] = knot_pair(model->knots, int(i));
  __syncthreads();
  const LearnedCoord m = coord_of(*model);
  auto bucket = [&](U x) {
    return learned_cdf_pairs(learned_coord((unsigned long long)to_key<U>(x, mode), m), knots) >> m.bucket_shift;
  };
  const size_t runs = (n + kBoundRun - 1) / kBoundRun;
  const size_t g = size_t(blockIdx.x) * blockDim.x + threadIdx.x, gs = size_t(gridDim.x) * blockDim.x;
  for (size_t r = g; r < runs; r += gs) {
    const size_t i0 = r * kBoundRun;
    U x[kBoundRun];
#pragma unroll
    for (int k = 0; k < kBoundRun; ++k) x[k] = data[min(i0 + k, n - 1)];
    uint32_t prev = i0 == 0 ? 0u : bucket(data[i0 - 1]) + 1;  // first bucket that may start at i0
#pragma unroll
    for (int k = 0; k < kBoundRun; ++k) {
      if (i0 + k >= n) break;
      const uint32_t here = bucket(x[k]);
      for (uint32_t b = prev; b <= here; ++b) offsets[b] = i0 + k;
      prev = here + 1;
    }
    if (i0 + kBoundRun >= n)  // this run holds the last key
      for (uint32_t b = prev; b <= B; ++b) offsets[b] = n;
  }
}

// The same bounds when buckets are few against keys: lane b finds the first
// position whose bucket is >= b by binary search over the keys (in bucket
// order), about log2(n) dependent reads per bucket instead of a read of
// every key.  The searches' upper levels share their positions and hit L2.
template <class U>
__global__ __launch_bounds__(256) void k_learn_bounds_search(const U* __restrict__ data, size_t n, int mode,
                                                             const LearnedModel* __restrict__ model,
                                                             unsigned long long* __restrict__ offsets, uint32_t B) {
  __shared__ unsigned long long knots[kKnots];
  for (unsigned i = threadIdx.x; i < unsigned(kKnots); i += blockDim.x) knots[i] = knot_pair(model->knots, int(i));
  __syncthreads();
  const LearnedCoord m = coord_of(*model);
  const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > B) return;
  size_t lo = 0, hi = n;  // the answer lies in [lo, hi]
  if (b == B) lo = n;
  if (b == 0) hi = 0;
  while (lo < hi) {
    const size_t mid = lo + ((hi - lo) >> 1);
    const uint32_t k = learned_cdf_pairs(learned_coord((unsigned long long)to_key<U>(data[mid], mode), m), knots) >>
                       m.bucket_shift;
    if (k < b)
      lo = mid + 1;
    else
      hi = mid;
  }
  offsets[b] = lo;
}

// Largest bucket: stats[0] = max over b of offsets[b+1] - offsets[b].
__global__ __launch_bounds__(1024) void k_bucket_max(const unsigned long long* __restrict__ offsets, uint32_t B,
                                                     unsigned long long* stats) {
  __shared__ unsigned long long wmax[16];
  unsigned long long m = 0;
  for (uint32_t b = threadIdx.x; b < B; b += blockDim.x) {
    const unsigned long long len = offsets[b + 1] - offsets[b];
    m = len > m ? len : m;
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long o = __shfl_xor(m, d, 64);
    m = o > m ? o : m;
  }
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 16; ++i) m = wmax[i] > m ? wmax[i] : m;
    stats[0] = m;
  }
}

template <class U> hipError_t run_learned(const LearnedJob& j, hipStream_t s, KernelClock* clk) {
  const unsigned sg = unsigned(std::min<size_t>((j.sample + 255) / 256, 4096));
  if (clk) clk->begin(s, "k_learn_sample", double(j.sample) * 16);
  hipLaunchKernelGGL((k_learn_sample<U>), dim3(sg), dim3(256), 0, s, static_cast<const U*>(j.data), j.n, j.mode,
                     j.sample_a, j.sample);
  if (clk) clk->end(s);
  RadixJob r;
  r.key_bytes = 8;
  r.mode = kIdentity;
  r.n = j.sample;
  r.keys_a = j.sample_a;
  r.keys_b = j.sample_b;
  r.meta = j.meta;
  r.status = j.status;
  r.status_cap = j.status_cap;
  r.error = j.error;
  r.num_cus = j.num_cus;
  r.atomic_rank = j.atomic_rank;
  hipError_t e = j.sample > kSmallMax ? radix_sort(r, s, clk)
                                      : small_sort(8, kIdentity, j.sample_a, j.sample, nullptr, 0, 0, nullptr, s, clk);
  if (e != hipSuccess) return e;
  const unsigned pg = unsigned(std::min<size_t>((j.points + 255) / 256, 1024));
  // Float keys: their bit patterns already follow the logarithm of the value.
  hipLaunchKernelGGL(k_learn_points, dim3(pg), dim3(256), 0, s, j.sample_a, j.sample, j.model, j.u, j.t, j.points,
                     j.bucket_shift, int(j.mode == kFloat));
  if (j.train) {
    hipLaunchKernelGGL(k_knot_first, dim3((kKnots + 256) / 256), dim3(256), 0, s, j.sample_a, j.sample, j.model,
                       j.knot_first);
    if (clk) clk->begin(s, "k_pwl_fit", 0);
    if (j.points <= uint32_t(kFitThreads) * kFitPpl)
      hipLaunchKernelGGL(k_pwl_fit<true>, dim3(1), dim3(kFitThreads), 0, s, j.u, j.t, j.points, j.theta, j.grad,
                         j.loss, j.iters, j.lr, j.knot_first, (unsigned long long)j.sample);
    else
      hipLaunchKernelGGL(k_pwl_fit<false>, dim3(1), dim3(kFitThreads), 0, s, j.u, j.t, j.points, j.theta, j.grad,
                         j.loss, j.iters, j.lr, j.knot_first, (unsigned long long)j.sample);
    if (clk) clk->end(s);
    hipLaunchKernelGGL(k_pwl_quantize, dim3(1), dim3(kFitThreads), 0, s, j.theta, j.model);
  } else {
    hipLaunchKernelGGL(k_pwl_uniform, dim3(1), dim3(256), 0, s, j.model);
  }
  // Partition by bucket.
  RadixJob p;
  p.key_bytes = int(sizeof(U));
  p.mode = j.mode;
  p.n = j.n;
  p.keys_a = j.data;
  p.keys_b = j.scratch;
  p.meta = j.meta;
  p.status = j.status;
  p.status_cap = j.status_cap;
  p.error = j.error;
  p.num_cus = j.num_cus;
  p.atomic_rank = j.atomic_rank;
  p.model = j.model;
  p.fingerprint = j.fingerprint;
  e = radix_sort(p, s, clk);
  if (e != hipSuccess) return e;
  if (size_t(j.buckets) * kBoundSearchRatio <= j.n) {
    if (clk) clk->begin(s, "k_learn_bounds_search", 0);
    hipLaunchKernelGGL((k_learn_bounds_search<U>), dim3(j.buckets / 256 + 1), dim3(256), 0, s,
                       static_cast<const U*>(j.data), j.n, j.mode, j.model, j.offsets, j.buckets);
  } else {
    const unsigned bg = unsigned(std::min<size_t>((j.n / kBoundRun + 256) / 256, size_t(j.num_cus) * 16));
    if (clk) clk->begin(s, "k_learn_bounds", double(j.n) * sizeof(U));
    hipLaunchKernelGGL((k_learn_bounds<U>), dim3(bg), dim3(256), 0, s, static_cast<const U*>(j.data), j.n, j.mode,
                       j.model, j.offsets, j.buckets);
  }
  if (clk) clk->end(s);
  hipLaunchKernelGGL(k_bucket_max, dim3(1), dim3(1024), 0, s, j.offsets, j.buckets, j.stats);
  return hipGetLastError();
}

}  // namespace

hipError_t learned_partition(const LearnedJob& j, hipStream_t s, KernelClock* clk) {
  switch (j.key_bytes) {
    case 4: return run_learned<uint32_t>(j, s, clk);
    case 8: return run_learned<uint64_t>(j, s, clk);
    default: return hipErrorInvalidValue;
  }
}

hipError_t pwl_fit(const float* u, const float* t, uint32_t T, float* theta, float* grad, float* loss, int iters,
                   float lr, hipStream_t s) {
  if (T <= uint32_t(kFitThreads) * kFitPpl)
    hipLaunchKernelGGL(k_pwl_fit<true>, dim3(1), dim3(kFitThreads), 0, s, u, t, T, theta, grad, loss, iters, lr,
                       nullptr, 0ull);
  else
    hipLaunchKernelGGL(k_pwl_fit<false>, dim3(1), dim3(kFitThreads), 0, s, u, t, T, theta, grad, loss, iters, lr,
                       nullptr, 0ull);
  return hipGetLastError();
}

}  // namespace ts
