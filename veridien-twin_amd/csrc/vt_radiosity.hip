// Node radiosity on gfx950.
//  * The reference's sparse gather Jacobi (gpu/wgsl/kernels.wgsl.js:416-461): one wave per active node for the
//    gather, a group of 8 / 16 / 32 lanes (one per smoothing neighbour) for the update.
//  * The dense form x <- e + F x of the precision configuration (n = 32768 nodes, 48 right-hand sides):
//      FP32 / FP64 sweeps on the vector ALUs (128-row x 48-RHS block tiles, split over K, x staged in LDS);
//      the BF16 sweep on MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation), F stored once as BF16 so a
//      sweep streams 2 B per matrix entry; the form-factor build traces one visibility ray per node pair.
#include "vt_internal.h"
#include "vt_nodes.h"

namespace vt {

// ------------------------------------------------------------------ sparse gather Jacobi
// One wave per active node, lane j taking gather entries j, j + 64, ...: the wave reads the node's table row as
// one contiguous stretch and has its K dependent fetches (hit normal, material, state) in flight at once; the
// lanes' sums meet in a fixed butterfly, so a step is deterministic.
// the gather entry's static weight: the node's albedo toward the gather direction times the hit node's albedo
// toward the node (both from the tables), the factor that multiplies the hit's radiance
__device__ __forceinline__ v3 gather_weight(const GatherDev &G, uint32_t i, uint32_t j, v3 n, uint32_t mi, uint32_t h) {
  const v3 dir = gather_dir(i, j, G.gather_k, n);
  const float muIn = sqrtf(fmaxf(0.0f, 1.0f - ((float)j + 0.5f) / (float)G.gather_k));
  const float muOut = fminf(fmaxf(-dot(dir, normalize(ld3(G.normals[h]))), 0.0f), 1.0f);
  return albedo_at(G.albedo, G.bins, mi, muIn) * albedo_at(G.albedo, G.bins, G.node_mat[h], muOut);
}

// WEIGHTS: each entry's hit node and weight come from the 16-B records k_jacobi_weights wrote once per node cache
// (the geometry and the albedo tables are fixed), in active-node order: a step streams them instead of reading the
// table row and recomputing the direction and two lookups
template <bool WEIGHTS>
__global__ __launch_bounds__(256) void k_jacobi_gather(GatherDev G) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t k = blockIdx.x * 4u + (threadIdx.x >> 6);
  if (k >= G.n_active) return;     // uniform over the wave
  const uint32_t i = G.active[k];
  const v3 n = normalize(ld3(G.normals[i]));
  const uint32_t mi = G.node_mat[i];
  v3 acc = mk(0.f, 0.f, 0.f);
  for (uint32_t j = lane; j < G.gather_k; j += 64u) {
    float4 rec;
    if (WEIGHTS) rec = G.gw[(size_t)k * G.gather_k + j];
    const uint32_t h = WEIGHTS ? __float_as_uint(rec.w) : G.gtable[(size_t)i * G.gather_k + j];
    if (h == MISS) continue;
    const v3 Ls = ld3(G.node_state[4 * (size_t)h + G.slot_in]);
    const v3 w = WEIGHTS ? ld3(rec) : gather_weight(G, i, j, n, mi, h);
    acc = acc + w * Ls;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    acc.x += __shfl_xor(acc.x, off);
    acc.y += __shfl_xor(acc.y, off);
    acc.z += __shfl_xor(acc.z, off);
  }
  if (lane) return;
  const float4 hh = G.hh[mi];
  const float s = PI_F / (float)G.gather_k;
  G.aux[k] = make_float4(acc.x * s / fmaxf(hh.x, 1e-4f), acc.y * s / fmaxf(hh.y, 1e-4f), acc.z * s / fmaxf(hh.z, 1e-4f), 0.0f);
}

__global__ __launch_bounds__(256) void k_jacobi_weights(GatherDev G) {
  const uint64_t total = (uint64_t)G.n_active * G.gather_k;
  for (uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x; r < total; r += (uint64_t)gridDim.x * 256) {
    const uint32_t k = (uint32_t)(r / G.gather_k), j = (uint32_t)(r - (uint64_t)k * G.gather_k);
    const uint32_t i = G.active[k];
    const uint32_t h = G.gtable[(size_t)i * G.gather_k + j];
    v3 w = mk(0.f, 0.f, 0.f);
    if (h != MISS) w = gather_weight(G, i, j, normalize(ld3(G.normals[i])), G.node_mat[i], h);
    G.gw[r] = make_float4(w.x, w.y, w.z, __uint_as_float(h));
  }
}

// One group of GR lanes per active node (GR >= nbr_k), lane j taking neighbour j: a row ends at its first
// non-positive weight (found by a ballot over the group), the weighted sums meet in a fixed butterfly and the
// group's first lane writes the node.
template <uint32_t GR>
__global__ __launch_bounds__(256) void k_jacobi_update(GatherDev G) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  const uint32_t k = t / GR, j = t % GR;
  const bool node = k < G.n_active;      // uniform over the group: GR divides 64
  float w = 0.0f;
  v3 a = mk(0.f, 0.f, 0.f);
  if (node && j < G.nbr_k) {
    w = G.nbr_w[(size_t)k * G.nbr_k + j];
    if (w > 0.0f) a = ld3(G.aux[G.nbr[(size_t)k * G.nbr_k + j]]);
  }
  // the row's length: the group's first lane with a non-positive weight (or past nbr_k)
  const uint64_t stop = __ballot(!(w > 0.0f));
  const uint32_t gbase = (threadIdx.x & 63u) & ~(GR - 1u);
  const uint64_t gm = (stop >> gbase) & (GR == 64 ? ~0ull : ((1ull << GR) - 1ull));
  const uint32_t len = gm ? (uint32_t)__builtin_ctzll(gm) : GR;
  if (j >= len) {
    w = 0.0f;
    a = mk(0.f, 0.f, 0.f);
  }
  v3 Ef = a * w;
  float ws = w;
#pragma unroll
  for (uint32_t off = GR / 2; off >= 1; off >>= 1) {
    Ef.x += __shfl_xor(Ef.x, off, 64);
    Ef.y += __shfl_xor(Ef.y, off, 64);
    Ef.z += __shfl_xor(Ef.z, off, 64);
    ws += __shfl_xor(ws, off, 64);
  }
  if (!node || j) return;
  const uint32_t i = G.active[k];
  if (ws > 0.0f) Ef = Ef * (1.0f / ws);
  else Ef = ld3(G.aux[k]);
  const float4 st = G.node_state[4 * (size_t)i];
  const float4 B = G.node_state[4 * (size_t)i + 3];
  const float4 hh = G.hh[G.node_mat[i]];
  const v3 h3 = mk(fmaxf(hh.x, 1e-4f), fmaxf(hh.y, 1e-4f), fmaxf(hh.z, 1e-4f));
  const v3 L = mk((B.x + Ef.x * st.w * h3.x) * INV_PI_F / h3.x, (B.y + Ef.y * st.w * h3.y) * INV_PI_F / h3.y,
                  (B.z + Ef.z * st.w * h3.z) * INV_PI_F / h3.z);
  G.node_state[4 * (size_t)i] = make_float4(Ef.x, Ef.y, Ef.z, st.w);
  G.node_state[4 * (size_t)i + G.slot_out] = make_float4(L.x, L.y, L.z, 1.0f);
}

hipError_t launch_jacobi_gather(const GatherDev &G, hipStream_t s) {
  if (G.n_active == 0) return hipSuccess;
  if (G.gw) hipLaunchKernelGGL(k_jacobi_gather<true>, dim3((G.n_active + 3) / 4), dim3(256), 0, s, G);
  else hipLaunchKernelGGL(k_jacobi_gather<false>, dim3((G.n_active + 3) / 4), dim3(256), 0, s, G);
  return hipGetLastError();
}
hipError_t launch_jacobi_weights(const GatherDev &G, hipStream_t s) {
  if (G.n_active == 0 || !G.gw) return hipSuccess;
  const uint64_t total = (uint64_t)G.n_active * G.gather_k;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(k_jacobi_weights, dim3(grid), dim3(256), 0, s, G);
  return hipGetLastError();
}
hipError_t launch_jacobi_update(const GatherDev &G, hipStream_t s) {
  if (G.n_active == 0) return hipSuccess;
  if (G.nbr_k > 32) return hipErrorInvalidValue;
  const uint32_t gr = G.nbr_k <= 8 ? 8u : G.nbr_k <= 16 ? 16u : 32u;
  const dim3 grid((unsigned)(((uint64_t)G.n_active * gr + 255) / 256));
  if (gr == 8) hipLaunchKernelGGL(k_jacobi_update<8>, grid, dim3(256), 0, s, G);
  else if (gr == 16) hipLaunchKernelGGL(k_jacobi_update<16>, grid, dim3(256), 0, s, G);
  else hipLaunchKernelGGL(k_jacobi_update<32>, grid, dim3(256), 0, s, G);
  return hipGetLastError();
}

// ------------------------------------------------------------------ dense sweeps
// y = e + F x over a 128-row x 48-RHS block; K is split KSPLIT ways and the partial sums are added to y
// (which the host first fills with e).
constexpr int DM = 128, DK = 32, DR = 48, KSPLIT = 4;
constexpr int SWEEP_KSPLIT = 8;   // K split of the MFMA sweeps

template <typename T>
__global__ __launch_bounds__(256) void k_sweep_vec(const float *F, uint32_t n, uint32_t rp, const T *x, T *y) {
  __shared__ T sx[DK * DR];
  __shared__ float sf[DM * (DK + 1)];
  const uint32_t row0 = blockIdx.x * DM;
  const uint32_t kper = (n + KSPLIT - 1) / KSPLIT;
  const uint32_t kbeg = blockIdx.y * kper, kend = min(n, kbeg + kper);
  const int t = threadIdx.x;
  const int rg = t >> 2, cg = t & 3;        // rows 2*rg, 2*rg+1; RHS 12*cg .. 12*cg+11
  T acc[2][12];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 12; b++) acc[a][b] = (T)0;
  for (uint32_t k0 = kbeg; k0 < kend; k0 += DK) {
    __syncthreads();
    for (int idx = t; idx < DK * DR; idx += 256) {
      const uint32_t kk = k0 + idx / DR;
      const int c = idx % DR;
      sx[idx] = (kk < kend && c < (int)rp) ? x[(size_t)kk * rp + c] : (T)0;
    }
    for (int idx = t; idx < DM * DK; idx += 256) {
      const int r = idx / DK, kk = idx % DK;
      const uint32_t gr = row0 + r, gk = k0 + kk;
      sf[r * (DK + 1) + kk] = (gr < n && gk < kend) ? F[(size_t)gr * n + gk] : 0.0f;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < DK; kk++) {
      const T f0 = (T)sf[(2 * rg) * (DK + 1) + kk], f1 = (T)sf[(2 * rg + 1) * (DK + 1) + kk];
#pragma unroll
      for (int b = 0; b < 12; b++) {
        const T xv = sx[kk * DR + 12 * cg + b];
        acc[0][b] += f0 * xv;
        acc[1][b] += f1 * xv;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; a++) {
    const uint32_t gr = row0 + 2 * rg + a;
    if (gr >= n) continue;
#pragma unroll
    for (int b = 0; b < 12; b++) {
      const int c = 12 * cg + b;
      if (c < (int)rp) atomicAdd(&y[(size_t)gr * rp + c], acc[a][b]);
    }
  }
}

template <typename T>
__global__ void k_fill_e(const float *e, T *y, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (T)e[i];
}


// ------------------------------------------------------------------ BF16 MFMA sweep
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);   // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                              // round to nearest even
  return (uint16_t)(u >> 16);
}

__global__ void k_f32_to_bf16(const float *src, uint16_t *dst, uint64_t n) {
  uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i + 3 < n) {
    float4 v = *reinterpret_cast<const float4 *>(src + i);
    ushort4 o = make_ushort4(f2bf(v.x), f2bf(v.y), f2bf(v.z), f2bf(v.w));
    *reinterpret_cast<ushort4 *>(dst + i) = o;
  } else {
    for (; i < n; i++) dst[i] = f2bf(src[i]);
  }
}
// x (n x rp, f32) -> xT (rp x n, bf16)
__global__ void k_x_to_bf16T(const float *x, uint32_t n, uint32_t rp, uint16_t *xT) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)n * rp) return;
  const uint32_t k = (uint32_t)(i / rp), c = (uint32_t)(i % rp);
  xT[(size_t)c * n + k] = f2bf(x[i]);
}

// One wave: 64 rows (four 16-row MFMA tiles) x 48 RHS (three 16-column tiles); a block of 4 waves covers 256
// rows of one of SWEEP_KSPLIT K slices. The slice is walked in stages of 64: every lane loads 32 contiguous
// bytes of each of its four rows (k = kb + 16 q + 0..15, q = lane >> 4), so a row's stage is one whole 128-B line
// of F instead of two half lines, and MFMA step j of the stage pairs the lane's k = kb + 16 q + 8 j .. + 7 on both
// operands. x's stage (48 x 64 BF16, 6 KB) is loaded once per block into LDS (rows padded to 144 B against bank
// conflicts) and read by all four waves; the next stage's F and x are loaded while this stage's MFMAs run (two
// register sets for F, two LDS buffers for x).
constexpr int BM_WAVE = 64, BWAVES = 4, BK_STAGE = 64, BK_CHUNKS = BK_STAGE / 32, XS_STRIDE = BK_STAGE + 8;
// K slices of the BF16 sweep: 4 gives N / 256 x 4 = 512 blocks at N = 32,768, one round of two blocks per CU
// (0.367 ms per sweep, against 0.386 ms with 8 slices and 0.416 ms with 16)
constexpr int BF16_KSPLIT = 4;
__global__ __launch_bounds__(256, 2) void k_sweep_bf16(const uint16_t *F16, const uint16_t *xT, uint32_t n,
                                                       uint32_t rp, float *y) {
  __shared__ __attribute__((aligned(16))) uint16_t xs[2][48 * XS_STRIDE];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t row0 = blockIdx.x * (BM_WAVE * BWAVES) + wv * BM_WAVE;
  const uint32_t kper = (n / BF16_KSPLIT + BK_STAGE - 1) / BK_STAGE * BK_STAGE;
  const uint32_t kbeg = blockIdx.y * kper;
  if (kbeg >= n) return;                          // the whole block: kbeg depends on blockIdx.y only
  const uint32_t stages = (min(n, kbeg + kper) - kbeg) / BK_STAGE;   // n is a multiple of 128
  const int li = lane & 15, q = lane >> 4;
  f32x4 acc[4][3];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) acc[a][b] = (f32x4){0.f, 0.f, 0.f, 0.f};
  auto loadA = [&](uint32_t kb, uint4 (&A)[4][BK_CHUNKS]) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const uint32_t r = row0 + 16 * a + li;
      const uint4 *src = reinterpret_cast<const uint4 *>(F16 + (size_t)r * n + kb + 8 * BK_CHUNKS * q);
#pragma unroll
      for (int j = 0; j < BK_CHUNKS; j++) A[a][j] = r < n ? src[j] : make_uint4(0, 0, 0, 0);
    }
  };
  // x's stage: 48 rows x 8 chunks of 16 B = 384 chunks; thread t takes chunk t and, below 128, chunk t + 256 (rows
  // past rp are zeros)
  constexpr int XCH = 48 * BK_STAGE / 8, XV = (XCH + 255) / 256;
  auto loadB = [&](uint32_t kb, uint4 (&Bt)[XV]) {
#pragma unroll
    for (int v = 0; v < XV; v++) {
      const int ch = threadIdx.x + 256 * v, c = ch / (BK_STAGE / 8), kk = (ch % (BK_STAGE / 8)) * 8;
      Bt[v] = ch < XCH && (uint32_t)c < rp ? *reinterpret_cast<const uint4 *>(xT + (size_t)c * n + kb + kk)
                                           : make_uint4(0, 0, 0, 0);
    }
  };
  auto storeB = [&](int buf, const uint4 (&Bt)[XV]) {
#pragma unroll
    for (int v = 0; v < XV; v++) {
      const int ch = threadIdx.x + 256 * v, c = ch / (BK_STAGE / 8), kk = (ch % (BK_STAGE / 8)) * 8;
      if (ch < XCH) *reinterpret_cast<uint4 *>(&xs[buf][c * XS_STRIDE + kk]) = Bt[v];
    }
  };
  auto compute = [&](int buf, const uint4 (&A)[4][BK_CHUNKS]) {
#pragma unroll
    for (int j = 0; j < BK_CHUNKS; j++) {
      uint4 B[3];
#pragma unroll
      for (int b = 0; b < 3; b++)
        B[b] = *reinterpret_cast<const uint4 *>(&xs[buf][(16 * b + li) * XS_STRIDE + 8 * BK_CHUNKS * q + 8 * j]);
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, A[a][j]),
                                                               __builtin_bit_cast(bf16x8, B[b]), acc[a][b], 0, 0, 0);
    }
  };
  uint4 A0[4][BK_CHUNKS], A1[4][BK_CHUNKS], Bt[XV];
  loadA(kbeg, A0);
  loadB(kbeg, Bt);
  storeB(0, Bt);
  __syncthreads();
  // every branch below depends on the stage count only, the same for the whole block
  for (uint32_t st = 0; st < stages; st += 2) {
    const bool more = st + 1 < stages;
    if (more) {
      loadA(kbeg + (st + 1) * BK_STAGE, A1);
      loadB(kbeg + (st + 1) * BK_STAGE, Bt);
    }
    compute(0, A0);
    if (more) storeB(1, Bt);
    __syncthreads();
    if (!more) break;
    const bool more2 = st + 2 < stages;
    if (more2) {
      loadA(kbeg + (st + 2) * BK_STAGE, A0);
      loadB(kbeg + (st + 2) * BK_STAGE, Bt);
    }
    compute(1, A1);
    if (more2) storeB(0, Bt);
    __syncthreads();
  }
  // D layout: lane holds column (lane & 15), rows 4 * (lane >> 4) + v
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const uint32_t c = 16 * b + li;
      if (c >= rp) continue;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const uint32_t r = row0 + 16 * a + 4 * q + v;
        if (r < n) atomicAdd(&y[(size_t)r * rp + c], acc[a][b][v]);
      }
    }
}

// ------------------------------------------------------------------ FP32 sweep on MFMA
// y = e + F x with F and x in binary32 on v_mfma_f32_16x16x4_f32 (FP32 inputs, FP32 accumulation). One wave:
// 64 rows (four 16-row tiles) x 48 RHS (three 16-column tiles). A K step of 16: lane (i = lane & 15,
// q = lane >> 4) loads 4 consecutive F entries F[row i][k0 + 4q .. 4q + 3] of each row tile and the matching
// 4 entries of xT (x transposed, RHS-major) of each column tile; MFMA m pairs the m-th of them, so the 16 k of
// the step are summed in the order (q, m) -> k0 + 4q + m, the same pairing on both operands. The next step's
// operands are loaded before this step's MFMAs (two register sets), and K is split SWEEP_KSPLIT ways.
typedef float f32x4v __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_sweep_f32_mfma(const float *F, const float *xT, uint32_t n, uint32_t rp,
                                                        float *y) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t row0 = blockIdx.x * 256 + wv * 64;
  const uint32_t kper = (n / SWEEP_KSPLIT + 15) / 16 * 16;
  const uint32_t kbeg = blockIdx.y * kper, kend = min(n, kbeg + kper);
  const int li = lane & 15, lq = (lane >> 4) * 4;
  f32x4v acc[4][3];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) acc[a][b] = (f32x4v){0.f, 0.f, 0.f, 0.f};
  auto load = [&](uint32_t k0, float4 (&A)[4], float4 (&B)[3]) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const uint32_t r = row0 + 16 * a + li;
      A[a] = (r < n && k0 + lq + 4 <= kend) ? *reinterpret_cast<const float4 *>(F + (size_t)r * n + k0 + lq)
                                            : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const uint32_t c = 16 * b + li;
      B[b] = (c < rp && k0 + lq + 4 <= kend) ? *reinterpret_cast<const float4 *>(xT + (size_t)c * n + k0 + lq)
                                             : make_float4(0.f, 0.f, 0.f, 0.f);
    }
  };
  float4 A0[4], B0[3], A1[4], B1[3];
  auto step = [&](const float4 (&A)[4], const float4 (&B)[3]) {
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[a].x, B[b].x, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[a].y, B[b].y, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[a].z, B[b].z, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[a].w, B[b].w, acc[a][b], 0, 0, 0);
      }
  };
  uint32_t k0 = kbeg;
  if (k0 < kend) load(k0, A0, B0);
  while (k0 < kend) {
    const uint32_t k1 = k0 + 16;
    if (k1 < kend) load(k1, A1, B1);
    step(A0, B0);
    if (k1 >= kend) break;
    const uint32_t k2 = k1 + 16;
    if (k2 < kend) load(k2, A0, B0);
    step(A1, B1);
    k0 = k2;
  }
  // D layout: lane holds column (lane & 15), rows 4 * (lane >> 4) + v
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const uint32_t c = 16 * b + li;
      if (c >= rp) continue;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const uint32_t r = row0 + 16 * a + 4 * (lane >> 4) + v;
        if (r < n) atomicAdd(&y[(size_t)r * rp + c], acc[a][b][v]);
      }
    }
}

// x (n x rp) -> xT (rp x n), binary32
__global__ void k_x_to_f32T(const float *x, uint32_t n, uint32_t rp, float *xT) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)n * rp) return;
  const uint32_t k = (uint32_t)(i / rp), c = (uint32_t)(i % rp);
  xT[(size_t)c * n + k] = x[i];
}

hipError_t launch_sweep_fp32_mfma(const float *F32, uint32_t n, uint32_t rp, const float *x, float *xT,
                                  const float *e, float *y, hipStream_t s) {
  if (n % 16 || rp > 48) return hipErrorInvalidValue;
  const uint64_t m = (uint64_t)n * rp;
  hipLaunchKernelGGL(k_x_to_f32T, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, x, n, rp, xT);
  hipLaunchKernelGGL(k_fill_e<float>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, e, y, m);
  hipLaunchKernelGGL(k_sweep_f32_mfma, dim3((n + 255) / 256, SWEEP_KSPLIT), dim3(256), 0, s, F32, xT, n, rp, y);
  return hipGetLastError();
}

hipError_t launch_convert_bf16(const float *src, uint16_t *dst, uint64_t n, hipStream_t s) {
  uint64_t th = (n + 3) / 4;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((th + 255) / 256)), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_sweep_bf16_T(const uint16_t *F16, const uint16_t *xT, uint32_t n, uint32_t rp, const float *e,
                               float *y, hipStream_t s) {
  const uint64_t m = (uint64_t)n * rp;
  hipLaunchKernelGGL(k_fill_e<float>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, e, y, m);
  hipLaunchKernelGGL(k_sweep_bf16, dim3((n + BM_WAVE * BWAVES - 1) / (BM_WAVE * BWAVES), BF16_KSPLIT), dim3(256), 0,
                     s, F16, xT, n, rp, y);
  return hipGetLastError();
}
hipError_t launch_x_to_bf16T(const float *x, uint32_t n, uint32_t rp, uint16_t *xT, hipStream_t s) {
  const uint64_t m = (uint64_t)n * rp;
  hipLaunchKernelGGL(k_x_to_bf16T, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, x, n, rp, xT);
  return hipGetLastError();
}

// FP64 sweep: y = e + F x with F widened from FP32 (exact) and FP64 accumulation
hipError_t launch_sweep_fp64(const float *F32, uint32_t n, uint32_t rp, const double *x, const float *e, double *y,
                             hipStream_t s) {
  const uint64_t m = (uint64_t)n * rp;
  hipLaunchKernelGGL(k_fill_e<double>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, e, y, m);
  hipLaunchKernelGGL(k_sweep_vec<double>, dim3((n + DM - 1) / DM, KSPLIT), dim3(256), 0, s, F32, n, rp, x, y);
  return hipGetLastError();
}

// ------------------------------------------------------------------ vector helpers
// partial sums per block: [0] sum (y - x)^2, [1] sum e^2, [2] sum (x - ref)^2, [3] sum ref^2
template <typename TX, typename TY>
__global__ void k_resid(const TY *y, const TX *x, const float *e, const double *ref, uint64_t m, double *part) {
  __shared__ double s[4][256];
  double a = 0, b = 0, c = 0, d = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
    const double r = (double)y[i] - (double)x[i];
    a += r * r;
    b += (double)e[i] * (double)e[i];
    if (ref) {
      const double q = (double)x[i] - ref[i];
      c += q * q;
      d += ref[i] * ref[i];
    }
  }
  s[0][threadIdx.x] = a; s[1][threadIdx.x] = b; s[2][threadIdx.x] = c; s[3][threadIdx.x] = d;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int q = 0; q < 4; q++) s[q][threadIdx.x] += s[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; q++) part[4 * blockIdx.x + q] = s[q][0];
}
constexpr int RES_BLOCKS = 256;
hipError_t launch_resid_f32(const float *y, const float *x, const float *e, const double *ref, uint64_t m,
                            double *part, hipStream_t s) {
  hipLaunchKernelGGL((k_resid<float, float>), dim3(RES_BLOCKS), dim3(256), 0, s, y, x, e, ref, m, part);
  return hipGetLastError();
}
hipError_t launch_resid_f64(const double *y, const double *x, const float *e, const double *ref, uint64_t m,
                            double *part, hipStream_t s) {
  hipLaunchKernelGGL((k_resid<double, double>), dim3(RES_BLOCKS), dim3(256), 0, s, y, x, e, ref, m, part);
  return hipGetLastError();
}
// x64 = (double) x32, for the FP64 check of an FP32 solution
__global__ void k_widen(const float *x, double *y, uint64_t m) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) y[i] = (double)x[i];
}
hipError_t launch_widen(const float *x, double *y, uint64_t m, hipStream_t s) {
  hipLaunchKernelGGL(k_widen, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, x, y, m);
  return hipGetLastError();
}
// a <- a + b ; c <- b (f32)
__global__ void k_axpy(float *a, const float *b, uint64_t m) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) a[i] += b[i];
}
hipError_t launch_add(float *a, const float *b, uint64_t m, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, a, b, m);
  return hipGetLastError();
}
// r = y - x
__global__ void k_sub(const float *y, const float *x, float *r, uint64_t m) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) r[i] = y[i] - x[i];
}
hipError_t launch_sub(const float *y, const float *x, float *r, uint64_t m, hipStream_t s) {
  hipLaunchKernelGGL(k_sub, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, y, x, r, m);
  return hipGetLastError();
}

// ------------------------------------------------------------------ form factors
// F_ij = V_ij cos_i cos_j A_j / (pi r^2); one lane per (i, j) pair, j fastest. V by an any-hit ray from
// p_i + 1e-4 n_i toward p_j, ending 1e-4 m short of it.
__global__ __launch_bounds__(DEFAULT_BLOCK) void k_formfactor(SceneDev S, uint32_t *overflow, uint2 *spill,
                                                              uint64_t spill_stride, uint32_t n, const float4 *nodes,
                                                              float *F) {
  __shared__ uint32_t s_node[STACK_LDS * DEFAULT_BLOCK];
  __shared__ float s_t[STACK_LDS * DEFAULT_BLOCK];
  Stack st{(lds_u32 *)s_node, (lds_f32 *)s_t, spill, threadIdx.x, DEFAULT_BLOCK, STACK_LDS, STACK_SPILL, spill_stride,
           (uint64_t)blockIdx.x * DEFAULT_BLOCK + threadIdx.x, overflow};
  const uint64_t total = (uint64_t)n * n;
  for (uint64_t q = (uint64_t)blockIdx.x * DEFAULT_BLOCK + threadIdx.x; q < total;
       q += (uint64_t)gridDim.x * DEFAULT_BLOCK) {
    const uint32_t i = (uint32_t)(q / n), j = (uint32_t)(q % n);
    float f = 0.0f;
    if (i != j) {
      const float4 pi = nodes[2 * i], ni = nodes[2 * i + 1], pj = nodes[2 * j], nj = nodes[2 * j + 1];
      v3 d = ld3(pj) - ld3(pi);
      const float r2 = dot(d, d);
      const float r = sqrtf(r2);
      d = d * (1.0f / r);
      const float ci = dot(ld3(ni), d), cj = -dot(ld3(nj), d);
      if (ci > 0.0f && cj > 0.0f && r > 2e-4f) {
        const v3 o = ld3(pi) + ld3(ni) * 1e-4f;
        v3 dd = ld3(pj) - o;
        const float rr = sqrtf(dot(dd, dd));
        dd = dd * (1.0f / rr);
        Hit h;
        if (!traverse<true>(S, S.root, o, dd, 0.0f, rr - 1e-4f, h, st)) f = ci * cj * pj.w / (PI_F * r2);
      }
    }
    F[q] = f;
  }
}
hipError_t launch_formfactors(const SceneDev &S, uint32_t *overflow, uint2 *spill, uint64_t spill_stride,
                              uint32_t n, const float4 *nodes, float *F, hipStream_t s) {
  const uint32_t blocks = (uint32_t)(spill_stride / DEFAULT_BLOCK);
  hipLaunchKernelGGL(k_formfactor, dim3(blocks), dim3(DEFAULT_BLOCK), 0, s, S, overflow, spill, spill_stride, n,
                     nodes, F);
  return hipGetLastError();
}
// F_i* *= row_max / max(1, sum_j F_ij): row sums at most row_max
__global__ void k_row_scale(float *F, uint32_t n, float row_max) {
  __shared__ double s[256];
  const uint32_t i = blockIdx.x;
  double a = 0;
  for (uint32_t j = threadIdx.x; j < n; j += 256) a += F[(size_t)i * n + j];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  const float sc = (float)((double)row_max / fmax(1.0, s[0]));
  for (uint32_t j = threadIdx.x; j < n; j += 256) F[(size_t)i * n + j] *= sc;
}
hipError_t launch_row_scale(float *F, uint32_t n, float row_max, hipStream_t s) {
  hipLaunchKernelGGL(k_row_scale, dim3(n), dim3(256), 0, s, F, n, row_max);
  return hipGetLastError();
}

}  // namespace vt
