// USFFT Radon transform for gfx950.
//
// Replaces class radonusfft (src/cuda/radonusfft.cu + kernels.cuh kernels
// divphi, unpaddivphi, fftshiftc, fftshift1c, wrap, wrapadj, takexy, gather,
// scatter, circ, pad, pad2) and its three cuFFT plans.
//
// Same mathematics as the reference chain (Gaussian gridding on the 2N x 2N
// oversampled grid, N = 3n/2, mu and M from eps = 1e-6), restructured:
//  * The checkerboard multiplications around each FFT are centred DFTs; they
//    are realised as index rotations on load/store (free).
//  * The 2N-point centred DFT of an input that is non-zero only on the
//    central N points splits into two N-point DFTs (even outputs: plain
//    input; odd outputs: input twisted by exp(-i pi j/N)).  All FFTs are
//    N-point LDS Stockham transforms (mixed radix 2/3/4/5/8).
//  * pad2 + circ + divphi are applied while loading the object rows; the
//    adjoint's unpaddivphi + circ + crop are applied while storing them, and
//    only the n object rows are transformed.
//  * The periodic halo of the gridding grid (wrap/wrapadj) is an index
//    modulo 2N.
//  * gather: one workgroup per (angle, slice) line; taps with separable
//    weights exp(-beta*dx^2)*exp(-beta*dy^2); the 1D inverse DFT, the
//    fftshift1c signs and the detector crop are fused in the same kernel.
//  * scatter: deterministic and free of global atomics.  Polar samples are
//    binned at set_theta time by the T x T grid tile of their floor
//    position; one wavefront owns a (bin chunk, slice) work item and adds its
//    samples' 15 x 15 footprints into an LDS tile extended by M on each side
//    (LDS adds from one wavefront only); a fold kernel sums the extended
//    tiles covering each grid point in a fixed order (this also performs
//    wrapadj).
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

namespace {
constexpr double kPiRef = 3.1415926535;  // kernels.cuh:1
constexpr int kTile = 32;                // scatter tile edge (grid points)
constexpr int kWorkMax = 1024;           // samples per scatter work item
}  // namespace

struct pt_tomo {
  int ntheta, pz, n, N, M;
  float mu;
  pt::LinePlan plan;
  float2* tw = nullptr;      // [N] exp(-2 pi i m/N)
  float2* twist = nullptr;   // [N] exp(-i pi j/N), j = idx - N/2
  float* dphi = nullptr;     // [N] exp(mu a^2)/(2N), a = idx - N/2
  int2* mrange = nullptr;    // [N] circ mask column range [lo, hi) per row
  float2* spos = nullptr;    // [ntheta*N] sample positions (x, y)
  int2* sfloor = nullptr;    // [ntheta*N] floor(2N*x), floor(2N*y)
  float2* A = nullptr;       // [pz][N][2N]
  float2* Bg = nullptr;      // [pz][2N][2N]
  float2* Q = nullptr;       // [pz][ntheta*N] polar samples (adjoint)
  // scatter bins
  int ntile1 = 0;            // tiles per axis
  int* wi_tile = nullptr;    // [nwi] tile index
  int2* wi_range = nullptr;  // [nwi] [s0, s1) into bin_sample
  int* bin_sample = nullptr; // [ntheta*N]
  int* tile_wi = nullptr;    // [ntile1^2 + 1] CSR: work items of each tile
  int nwi = 0;
  float2* scratch = nullptr; // [nwi][pz][(T+2M)^2]
  int rows_per_wg = 1, cols_per_wg = 1;
  bool ready = false;
};

namespace pt {
namespace {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int wrap2n(int v, int twoN) {
  v %= twoN;
  return v < 0 ? v + twoN : v;
}

// ------------------------------------------------------------------ forward row pass
// For rows q of slice z: x[p] = mask*u[clamp]*dphi[q]*dphi[p] (pad2, circ,
// divphi), centred 2N-point DFT along p -> A[z][q][kx], kx = k' + N.
__global__ void k_row_fwd(float2* __restrict__ A, const float2* __restrict__ u,
                          const float2* __restrict__ tw, const float2* __restrict__ twist,
                          const float* __restrict__ dphi, const int2* __restrict__ mrange,
                          LinePlan plan, int n, int rows) {
  extern __shared__ float2 lds[];
  const int N = plan.N, z = blockIdx.y, q0 = blockIdx.x * rows;
  const int o = N / 6;
  const int nl = 2 * rows;  // line 2r: plain, 2r+1: twisted
  for (int i = threadIdx.x; i < nl * N; i += blockDim.x) {
    const int line = i / N, c = i - line * N;
    const int r = line >> 1, q = q0 + r;
    float2 val = make_float2(0.f, 0.f);
    if (q < N) {
      const int2 mr = mrange[q];
      if (c >= mr.x && c < mr.y) {
        const int uy = min(max(q - o, 0), n - 1), ux = min(max(c - o, 0), n - 1);
        val = cscale(u[((long)z * n + uy) * n + ux], dphi[q] * dphi[c]);
        if (line & 1) val = cmul(val, twist[c]);
      }
    }
    // centred -> rotated index a' = (c + N/2) mod N
    int a = c + N / 2;
    if (a >= N) a -= N;
    lds[line * N + a] = val;
  }
  __syncthreads();
  lds_fft_lines(lds, N, nl, plan, tw, -1, threadIdx.x, blockDim.x);
  for (int i = threadIdx.x; i < nl * N; i += blockDim.x) {
    const int line = i / N, b = i - line * N;
    const int r = line >> 1, q = q0 + r;
    if (q >= N) continue;
    const int m = b < N / 2 ? b : b - N;  // centred frequency
    const int kx = 2 * m + N + (line & 1);
    A[((long)z * N + q) * (2 * N) + kx] = lds[line * N + b];
  }
}

// ------------------------------------------------------------------ forward column pass
// For columns kx of slice z: y[q] = A[z][q][kx] -> centred 2N-point DFT along
// q -> B[z][ky][kx].
__global__ void k_col_fwd(float2* __restrict__ Bg, const float2* __restrict__ A,
                          const float2* __restrict__ tw, const float2* __restrict__ twist,
                          LinePlan plan, int cols) {
  extern __shared__ float2 lds[];
  const int N = plan.N, z = blockIdx.y, k0 = blockIdx.x * cols;
  const int nl = 2 * cols;
  for (int i = threadIdx.x; i < N * cols; i += blockDim.x) {
    const int q = i / cols, cc = i - q * cols;
    const float2 val = A[((long)z * N + q) * (2 * N) + k0 + cc];
    int a = q + N / 2;
    if (a >= N) a -= N;
    lds[(2 * cc) * N + a] = val;
    lds[(2 * cc + 1) * N + a] = cmul(val, twist[q]);
  }
  __syncthreads();
  lds_fft_lines(lds, N, nl, plan, tw, -1, threadIdx.x, blockDim.x);
  for (int i = threadIdx.x; i < 2 * N * cols; i += blockDim.x) {
    const int ky = i / cols, cc = i - ky * cols;
    const int kp = ky - N;                    // centred k' in [-N, N)
    const int par = kp & 1;                   // odd -> twisted line
    const int m = (kp - par) / 2;             // [-N/2, N/2)
    const int b = m < 0 ? m + N : m;
    Bg[((long)z * 2 * N + ky) * (2 * N) + k0 + cc] = lds[(2 * cc + par) * N + b];
  }
}

// ------------------------------------------------------------------ gather + 1D IDFT
__global__ void k_gather_line(float2* __restrict__ g, const float2* __restrict__ Bg,
                              const float2* __restrict__ spos, const int2* __restrict__ sfloor,
                              const float2* __restrict__ tw, LinePlan plan, int M, float beta,
                              float oscale, int n, int ntheta, int pz) {
  extern __shared__ float2 lds[];
  const int N = plan.N, twoN = 2 * N;
  const int th = blockIdx.x, z = blockIdx.y;
  const float inv2N = 1.0f / (float)twoN;
  const float2* Bz = Bg + (long)z * twoN * twoN;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const long s = (long)th * N + i;
    const float2 p = spos[s];
    const int2 fl = sfloor[s];
    float ex[15], ey[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      const float dx = (float)(fl.x - M + k) * inv2N - p.x;
      const float dy = (float)(fl.y - M + k) * inv2N - p.y;
      ex[k] = (k <= 2 * M) ? __expf(-beta * dx * dx) : 0.f;
      ey[k] = (k <= 2 * M) ? __expf(-beta * dy * dy) : 0.f;
    }
    float2 acc = make_float2(0.f, 0.f);
    int cx[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) cx[k] = wrap2n(N + fl.x - M + k, twoN);
    for (int k1 = 0; k1 <= 2 * M; ++k1) {
      const float2* row = Bz + (long)wrap2n(N + fl.y - M + k1, twoN) * twoN;
      float2 r = make_float2(0.f, 0.f);
#pragma unroll
      for (int k0 = 0; k0 < 15; ++k0) {
        if (k0 <= 2 * M) {
          const float2 v = row[cx[k0]];
          r.x += ex[k0] * v.x;
          r.y += ex[k0] * v.y;
        }
      }
      acc.x += ey[k1] * r.x;
      acc.y += ey[k1] * r.y;
    }
    // kappa = i - N/2 -> rotated index kappa mod N
    int a = i - N / 2;
    if (a < 0) a += N;
    lds[a] = acc;
  }
  __syncthreads();
  lds_fft_lines(lds, N, 1, plan, tw, +1, threadIdx.x, blockDim.x);
  // crop: output j in [0, n) <-> sigma = j - n/2
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    int b = j - n / 2;
    if (b < 0) b += N;
    g[((long)th * pz + z) * n + j] = cscale(lds[b], oscale);
  }
}

// ------------------------------------------------------------------ adjoint line: pad + 1D DFT
__global__ void k_adj_line(float2* __restrict__ Q, const float2* __restrict__ g,
                           const float2* __restrict__ tw, LinePlan plan, float oscale, int n,
                           int ntheta, int pz) {
  extern __shared__ float2 lds[];
  const int N = plan.N, th = blockIdx.x, z = blockIdx.y;
  const int o = N / 6;
  const float2* gl = g + ((long)th * pz + z) * n;
  for (int s = threadIdx.x; s < N; s += blockDim.x) {
    const int j = min(max(s - o, 0), n - 1);  // pad (kernels.cuh:189-206)
    int a = s + N / 2;
    if (a >= N) a -= N;
    lds[a] = gl[j];
  }
  __syncthreads();
  lds_fft_lines(lds, N, 1, plan, tw, -1, threadIdx.x, blockDim.x);
  float2* Ql = Q + ((long)z * ntheta + th) * N;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    int b = i - N / 2;  // kappa
    if (b < 0) b += N;
    Ql[i] = cscale(lds[b], oscale);
  }
}

// ------------------------------------------------------------------ scatter
// One wavefront per (work item, slice).  Lanes: slot = lane/16 (4 samples in
// flight), c = lane%16 (footprint column, c <= 2M).
__global__ __launch_bounds__(64) void k_scatter(float2* __restrict__ scratch,
                                                const float2* __restrict__ Q,
                                                const float2* __restrict__ spos,
                                                const int2* __restrict__ sfloor,
                                                const int* __restrict__ wi_tile,
                                                const int2* __restrict__ wi_range,
                                                const int* __restrict__ bin_sample, int N, int M,
                                                float beta, int ntile1, int ns_total, int pz) {
  extern __shared__ float2 acc[];
  const int wi = blockIdx.x, z = blockIdx.y;
  const int E = kTile + 2 * M;
  const int lane = threadIdx.x, slot = lane >> 4, c = lane & 15;
  for (int i = lane; i < E * E; i += 64) acc[i] = make_float2(0.f, 0.f);
  __syncthreads();
  const int tile = wi_tile[wi];
  const int ty0 = (tile / ntile1) * kTile - N, tx0 = (tile % ntile1) * kTile - N;
  const int2 rg = wi_range[wi];
  const float inv2N = 1.0f / (float)(2 * N);
  const float2* Qz = Q + (long)z * ns_total;
  for (int b = rg.x; b < rg.y; b += 4) {
    const int bi = b + slot;
    const bool active = bi < rg.y && c <= 2 * M;
    const int s = bi < rg.y ? bin_sample[bi] : 0;
    const float2 p = spos[s];
    const int2 fl = sfloor[s];
    const float2 qv = Qz[s];
    const float dx = (float)(fl.x - M + c) * inv2N - p.x;
    const float dy = (float)(fl.y - M + c) * inv2N - p.y;
    const float ex = __expf(-beta * dx * dx);
    const float eyc = __expf(-beta * dy * dy);
    const float2 qx = make_float2(qv.x * ex, qv.y * ex);
    const int col = fl.x - tx0 + c;  // in [0, E)
    const int rowb = fl.y - ty0;     // row of k1 = 0 is rowb (+M offset cancels)
    for (int k1 = 0; k1 <= 2 * M; ++k1) {
      const float ey = __shfl(eyc, (slot << 4) + k1, 64);
      if (active) {
        float2* dst = acc + (rowb + k1) * E + col;
        atomicAdd(&dst->x, qx.x * ey);
        atomicAdd(&dst->y, qx.y * ey);
      }
    }
  }
  __syncthreads();
  float2* out = scratch + ((long)wi * pz + z) * E * E;
  for (int i = lane; i < E * E; i += 64) out[i] = acc[i];
}

// fold the extended tiles into the periodic 2N x 2N grid (fixed order)
__global__ void k_fold_grid(float2* __restrict__ Bg, const float2* __restrict__ scratch,
                            const int* __restrict__ tile_wi, int N, int M, int ntile1, int pz) {
  const int twoN = 2 * N;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long total = (long)pz * twoN * twoN;
  if (idx >= total) return;
  const int kx = (int)(idx % twoN);
  const long r = idx / twoN;
  const int ky = (int)(r % twoN);
  const int z = (int)(r / twoN);
  const int E = kTile + 2 * M;
  float2 s = make_float2(0.f, 0.f);
  for (int iy = -1; iy <= 1; ++iy) {
    const int ey = ky - N + iy * twoN;  // image of ell_y
    // tiles ty with ty*T - N - M <= ey < ty*T - N + T + M
    const int tlo = (int)floorf((float)(ey + N + M - E + 1) / kTile);
    const int thi = (int)floorf((float)(ey + N + M) / kTile);
    for (int ty = max(tlo, 0); ty <= min(thi, ntile1 - 1); ++ty) {
      const int ly = ey - (ty * kTile - N - M);
      if (ly < 0 || ly >= E) continue;
      for (int ix = -1; ix <= 1; ++ix) {
        const int ex = kx - N + ix * twoN;
        const int slo = (int)floorf((float)(ex + N + M - E + 1) / kTile);
        const int shi = (int)floorf((float)(ex + N + M) / kTile);
        for (int tx = max(slo, 0); tx <= min(shi, ntile1 - 1); ++tx) {
          const int lx = ex - (tx * kTile - N - M);
          if (lx < 0 || lx >= E) continue;
          const int tile = ty * ntile1 + tx;
          for (int w = tile_wi[tile]; w < tile_wi[tile + 1]; ++w)
            s = cadd(s, scratch[((long)w * pz + z) * E * E + ly * E + lx]);
        }
      }
    }
  }
  Bg[idx] = s;
}

// ------------------------------------------------------------------ adjoint column pass
// For columns kx: centred inverse 2N-point DFT along ky evaluated at the n
// object rows q in [o, o+n): even/odd split into two N-point inverse DFTs.
__global__ void k_col_adj(float2* __restrict__ A, const float2* __restrict__ Bg,
                          const float2* __restrict__ tw, const float2* __restrict__ twist,
                          LinePlan plan, int n, int cols) {
  extern __shared__ float2 lds[];
  const int N = plan.N, z = blockIdx.y, k0 = blockIdx.x * cols;
  const int nl = 2 * cols;
  for (int i = threadIdx.x; i < 2 * N * cols; i += blockDim.x) {
    const int ky = i / cols, cc = i - ky * cols;
    const int kp = ky - N;
    const int par = kp & 1;
    const int m = (kp - par) / 2;
    const int b = m < 0 ? m + N : m;
    lds[(2 * cc + par) * N + b] = Bg[((long)z * 2 * N + ky) * (2 * N) + k0 + cc];
  }
  __syncthreads();
  lds_fft_lines(lds, N, nl, plan, tw, +1, threadIdx.x, blockDim.x);
  const int o = N / 6;
  for (int i = threadIdx.x; i < n * cols; i += blockDim.x) {
    const int qi = i / cols, cc = i - qi * cols;
    const int q = qi + o;             // row index, centred q - N/2
    int a = q - N / 2;
    if (a < 0) a += N;
    const float2 e = lds[(2 * cc) * N + a];
    const float2 od = lds[(2 * cc + 1) * N + a];
    A[((long)z * n + qi) * (2 * N) + k0 + cc] = cadd(e, cmulc(od, twist[q]));
  }
}

// ------------------------------------------------------------------ adjoint row pass
__global__ void k_row_adj(float2* __restrict__ f, const float2* __restrict__ A,
                          const float2* __restrict__ tw, const float2* __restrict__ twist,
                          const float* __restrict__ dphi, const int2* __restrict__ mrange,
                          LinePlan plan, int n, int rows, float scale) {
  extern __shared__ float2 lds[];
  const int N = plan.N, z = blockIdx.y, r0 = blockIdx.x * rows;
  const int o = N / 6;
  const int nl = 2 * rows;
  for (int i = threadIdx.x; i < 2 * N * rows; i += blockDim.x) {
    const int r = i / (2 * N), kx = i - r * 2 * N;
    const int qi = r0 + r;
    float2 val = make_float2(0.f, 0.f);
    if (qi < n) val = A[((long)z * n + qi) * (2 * N) + kx];
    const int kp = kx - N;
    const int par = kp & 1;
    const int m = (kp - par) / 2;
    const int b = m < 0 ? m + N : m;
    lds[(2 * r + par) * N + b] = val;
  }
  __syncthreads();
  lds_fft_lines(lds, N, nl, plan, tw, +1, threadIdx.x, blockDim.x);
  for (int i = threadIdx.x; i < n * rows; i += blockDim.x) {
    const int r = i / n, pi = i - r * n;
    const int qi = r0 + r;
    if (qi >= n) continue;
    const int q = qi + o, p = pi + o;
    int a = p - N / 2;
    if (a < 0) a += N;
    const float2 e = lds[(2 * r) * N + a];
    const float2 od = lds[(2 * r + 1) * N + a];
    float2 v = cadd(e, cmulc(od, twist[p]));
    const int2 mr = mrange[q];
    const float w = (p >= mr.x && p < mr.y) ? dphi[q] * dphi[p] * scale : 0.f;
    f[((long)z * n + qi) * n + pi] = cscale(v, w);
  }
}

// ------------------------------------------------------------------ host helpers
int lines_fit(const LinePlan& p, int nthr, int maxlines, size_t lds_cap) {
  int best = 1;
  for (int l = 1; l <= maxlines; ++l) {
    if (!lds_fft_fits(p, l, nthr)) break;
    if ((size_t)l * p.N * sizeof(float2) > lds_cap) break;
    best = l;
  }
  return best;
}

int run_fwd(pt_tomo* h, float2* g, const float2* f, float scale, hipStream_t st) {
  const int N = h->N;
  const int nthr = 256;
  {
    const int rows = h->rows_per_wg;
    dim3 grid((N + rows - 1) / rows, h->pz);
    size_t lds = sizeof(float2) * 2 * rows * N;
    hipLaunchKernelGGL(k_row_fwd, grid, dim3(nthr), lds, st, h->A, f, h->tw, h->twist, h->dphi,
                       h->mrange, h->plan, h->n, rows);
    PT_LAUNCH_CHECK();
  }
  {
    const int cols = h->cols_per_wg;
    dim3 grid(2 * N / cols, h->pz);
    size_t lds = sizeof(float2) * 2 * cols * N;
    hipLaunchKernelGGL(k_col_fwd, grid, dim3(nthr), lds, st, h->Bg, h->A, h->tw, h->twist, h->plan,
                       cols);
    PT_LAUNCH_CHECK();
  }
  {
    const float beta = (float)(kPiRef * kPiRef / (double)h->mu);
    const double sgn = ((N / 2) % 2) ? -1.0 : 1.0;
    const float oscale = (float)(sgn * kPiRef / (double)h->mu / (double)N * (double)scale);
    dim3 grid(h->ntheta, h->pz);
    int thr = std::min(256, ((N + 63) / 64) * 64);
    if (!lds_fft_fits(h->plan, 1, thr)) thr = 256;
    hipLaunchKernelGGL(k_gather_line, grid, dim3(thr), sizeof(float2) * N, st, g, h->Bg, h->spos,
                       h->sfloor, h->tw, h->plan, h->M, beta, oscale, h->n, h->ntheta, h->pz);
    PT_LAUNCH_CHECK();
  }
  return PT_OK;
}

int run_adj(pt_tomo* h, float2* f, const float2* g, float scale, hipStream_t st) {
  const int N = h->N, M = h->M;
  const int nthr = 256;
  {
    const double sgn = ((N / 2) % 2) ? -1.0 : 1.0;
    const float oscale = (float)(sgn * kPiRef / (double)h->mu / (double)N);
    dim3 grid(h->ntheta, h->pz);
    int thr = std::min(256, ((N + 63) / 64) * 64);
    if (!lds_fft_fits(h->plan, 1, thr)) thr = 256;
    hipLaunchKernelGGL(k_adj_line, grid, dim3(thr), sizeof(float2) * N, st, h->Q, g, h->tw,
                       h->plan, oscale, h->n, h->ntheta, h->pz);
    PT_LAUNCH_CHECK();
  }
  {
    const float beta = (float)(kPiRef * kPiRef / (double)h->mu);
    const int E = kTile + 2 * M;
    dim3 grid(h->nwi, h->pz);
    hipLaunchKernelGGL(k_scatter, grid, dim3(64), sizeof(float2) * E * E, st, h->scratch, h->Q,
                       h->spos, h->sfloor, h->wi_tile, h->wi_range, h->bin_sample, N, M, beta,
                       h->ntile1, h->ntheta * N, h->pz);
    PT_LAUNCH_CHECK();
    long total = (long)h->pz * 4 * N * N;
    hipLaunchKernelGGL(k_fold_grid, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, h->Bg,
                       h->scratch, h->tile_wi, N, M, h->ntile1, h->pz);
    PT_LAUNCH_CHECK();
  }
  {
    const int cols = h->cols_per_wg;
    dim3 grid(2 * N / cols, h->pz);
    size_t lds = sizeof(float2) * 2 * cols * N;
    hipLaunchKernelGGL(k_col_adj, grid, dim3(nthr), lds, st, h->A, h->Bg, h->tw, h->twist, h->plan,
                       h->n, cols);
    PT_LAUNCH_CHECK();
  }
  {
    const int rows = h->rows_per_wg;
    dim3 grid((h->n + rows - 1) / rows, h->pz);
    size_t lds = sizeof(float2) * 2 * rows * N;
    hipLaunchKernelGGL(k_row_adj, grid, dim3(nthr), lds, st, f, h->A, h->tw, h->twist, h->dphi,
                       h->mrange, h->plan, h->n, rows, scale);
    PT_LAUNCH_CHECK();
  }
  return PT_OK;
}

}  // namespace

bool make_line_plan(int N, LinePlan* p) {
  p->N = N;
  p->npass = 0;
  int rem = N;
  // prefer radix 8, then 4, then 2; odd radices 3 and 5
  const int order[] = {8, 4, 2, 3, 5};
  // take odd factors first so that power-of-two passes come last
  for (int r : {3, 5}) {
    while (rem % r == 0) {
      if (p->npass >= 12) return false;
      p->radix[p->npass++] = r;
      rem /= r;
    }
  }
  (void)order;
  while (rem > 1) {
    int r = (rem % 8 == 0 && rem != 16) ? 8 : (rem % 4 == 0 ? 4 : (rem % 2 == 0 ? 2 : 0));
    if (!r || p->npass >= 12) return false;
    p->radix[p->npass++] = r;
    rem /= r;
  }
  return true;
}

}  // namespace pt

using namespace pt;

extern "C" int pt_tomo_create(int ntheta, int pnz, int n, pt_tomo** out) {
  PT_REQUIRE(out, PT_ERR_ARG, "pt_tomo_create: null output handle");
  *out = nullptr;
  PT_REQUIRE(ntheta > 0 && pnz > 0 && n > 0, PT_ERR_SIZE, "pt_tomo_create: sizes must be positive");
  PT_REQUIRE(n % 4 == 0, PT_ERR_SIZE, "pt_tomo_create: n must be a multiple of 4");
  const int N = n * 3 / 2;
  LinePlan plan;
  PT_REQUIRE(make_line_plan(N, &plan), PT_ERR_SIZE,
             "pt_tomo_create: 3n/2 must factor into radices 2, 3, 4, 5, 8");
  PT_REQUIRE(lds_fft_fits(plan, 1, 256), PT_ERR_SIZE, "pt_tomo_create: n too large");
  pt_tomo* h = new pt_tomo();
  h->ntheta = ntheta; h->pz = pnz; h->n = n; h->N = N; h->plan = plan;
  // radonusfft.cu:10-12
  const float eps = 1e-6f;
  h->mu = (float)(-std::log((double)eps) / (2.0 * N * N));
  h->M = (int)std::ceil(2 * N / kPiRef *
                        std::sqrt(-(double)h->mu * std::log((double)eps) +
                                  ((double)h->mu * N) * ((double)h->mu * N) / 4));
  if (h->M > 7) {
    set_error("pt_tomo_create: gridding half-width above 7 is not supported");
    delete h;
    return PT_ERR_SIZE;
  }
  // LDS lines per workgroup for the row/column passes (<= 48 KB)
  const size_t cap = 48 * 1024;
  int lines = lines_fit(plan, 256, 16, cap);
  h->rows_per_wg = std::max(1, lines / 2);
  int cols = 1;
  while (cols * 2 <= lines / 2 && (2 * N) % (cols * 2) == 0) cols *= 2;
  h->cols_per_wg = cols;
  // host tables
  std::vector<float2> tw(N), twist(N);
  std::vector<float> dphi(N);
  std::vector<int2> mr(N);
  for (int m = 0; m < N; ++m) {
    double a = -2.0 * kPi * m / N;
    tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    const int j = m - N / 2;
    double b = -kPi * j / N;
    twist[m] = make_float2((float)std::cos(b), (float)std::sin(b));
    dphi[m] = (float)(std::exp((double)h->mu * (double)j * (double)j) / (2.0 * N));
  }
  // circ mask (kernels.cuh:160-172), float32 arithmetic without contraction
  for (int q = 0; q < N; ++q) {
    volatile float y = (float)(q - N / 2) / (float)N;
    volatile float yy = 4.0f * y;
    yy = yy * y;
    volatile float thr = 1.0f - 1.0f / (float)N;
    int lo = N, hi = 0;
    for (int c = 0; c < N; ++c) {
      volatile float x = (float)(c - N / 2) / (float)N;
      volatile float xx = 4.0f * x;
      xx = xx * x;
      volatile float v = xx + yy;
      if (v < thr) {
        lo = std::min(lo, c);
        hi = std::max(hi, c + 1);
      }
    }
    if (lo >= hi) lo = hi = 0;
    mr[q] = make_int2(lo, hi);
  }
  auto fail = [&](hipError_t e, const char* what) {
    set_error(std::string("pt_tomo_create: ") + what + ": " + hipGetErrorString(e));
    pt_tomo_destroy(h);
    return PT_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&h->tw, sizeof(float2) * N))) return fail(e, "tables");
  if ((e = hipMalloc(&h->twist, sizeof(float2) * N))) return fail(e, "tables");
  if ((e = hipMalloc(&h->dphi, sizeof(float) * N))) return fail(e, "tables");
  if ((e = hipMalloc(&h->mrange, sizeof(int2) * N))) return fail(e, "tables");
  if ((e = hipMalloc(&h->spos, sizeof(float2) * (size_t)ntheta * N))) return fail(e, "positions");
  if ((e = hipMalloc(&h->sfloor, sizeof(int2) * (size_t)ntheta * N))) return fail(e, "positions");
  if ((e = hipMalloc(&h->bin_sample, sizeof(int) * (size_t)ntheta * N))) return fail(e, "bins");
  if ((e = hipMalloc(&h->A, sizeof(float2) * (size_t)pnz * N * 2 * N))) return fail(e, "row buffer");
  if ((e = hipMalloc(&h->Bg, sizeof(float2) * (size_t)pnz * 4 * N * N))) return fail(e, "grid");
  if ((e = hipMalloc(&h->Q, sizeof(float2) * (size_t)pnz * ntheta * N))) return fail(e, "samples");
  if ((e = hipMemcpy(h->tw, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice))) return fail(e, "copy");
  if ((e = hipMemcpy(h->twist, twist.data(), sizeof(float2) * N, hipMemcpyHostToDevice))) return fail(e, "copy");
  if ((e = hipMemcpy(h->dphi, dphi.data(), sizeof(float) * N, hipMemcpyHostToDevice))) return fail(e, "copy");
  if ((e = hipMemcpy(h->mrange, mr.data(), sizeof(int2) * N, hipMemcpyHostToDevice))) return fail(e, "copy");
  *out = h;
  return PT_OK;
}

extern "C" int pt_tomo_destroy(pt_tomo* h) {
  if (!h) return PT_OK;
  (void)hipFree(h->tw); (void)hipFree(h->twist); (void)hipFree(h->dphi); (void)hipFree(h->mrange);
  (void)hipFree(h->spos); (void)hipFree(h->sfloor); (void)hipFree(h->A); (void)hipFree(h->Bg); (void)hipFree(h->Q);
  (void)hipFree(h->wi_tile); (void)hipFree(h->wi_range); (void)hipFree(h->bin_sample); (void)hipFree(h->tile_wi);
  (void)hipFree(h->scratch);
  delete h;
  return PT_OK;
}

extern "C" int pt_tomo_set_theta(pt_tomo* h, const float* theta, pt_stream stream) {
  PT_REQUIRE(h && theta, PT_ERR_ARG, "pt_tomo_set_theta: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int N = h->N, nt = h->ntheta;
  std::vector<float> th(nt);
  PT_HIP(hipMemcpyAsync(th.data(), theta, sizeof(float) * nt, hipMemcpyDeviceToHost, st));
  PT_HIP(hipStreamSynchronize(st));
  // takexy (kernels.cuh:84-95): FP64 then float32, clamp at 0.5
  const long ns = (long)nt * N;
  std::vector<float2> pos(ns);
  std::vector<int2> fl(ns);
  const float lim = (float)(0.5 - 1e-5);
  const float twoN = (float)(2 * N);
  for (int t = 0; t < nt; ++t) {
    const double c = std::cos((double)th[t]), s = std::sin((double)th[t]);
    for (int i = 0; i < N; ++i) {
      const double k = (double)(i - N / 2);
      float x = (float)(k / N * c);
      float y = (float)(-k / N * s);
      if (x >= 0.5f) x = lim;
      if (y >= 0.5f) y = lim;
      volatile float px = twoN * x, py = twoN * y;
      pos[(long)t * N + i] = make_float2(x, y);
      fl[(long)t * N + i] = make_int2((int)std::floor(px), (int)std::floor(py));
    }
  }
  // bins by kTile x kTile tile of the floor position
  const int nt1 = (2 * N + kTile - 1) / kTile;
  const long ntiles = (long)nt1 * nt1;
  std::vector<int> cnt(ntiles + 1, 0);
  for (long s = 0; s < ns; ++s) {
    const int tyi = (fl[s].y + N) / kTile, txi = (fl[s].x + N) / kTile;
    if (fl[s].x < -N || fl[s].x >= N || fl[s].y < -N || fl[s].y >= N) {
      set_error("pt_tomo_set_theta: sample outside the gridding domain");
      return PT_ERR_ARG;
    }
    cnt[(long)tyi * nt1 + txi + 1]++;
  }
  for (long i = 0; i < ntiles; ++i) cnt[i + 1] += cnt[i];
  std::vector<int> fillp(cnt.begin(), cnt.end() - 1), list(ns);
  for (long s = 0; s < ns; ++s) {
    const int tyi = (fl[s].y + N) / kTile, txi = (fl[s].x + N) / kTile;
    list[fillp[(long)tyi * nt1 + txi]++] = (int)s;
  }
  std::vector<int> wtile;
  std::vector<int2> wrange;
  std::vector<int> twi(ntiles + 1, 0);
  for (long tI = 0; tI < ntiles; ++tI) {
    twi[tI] = (int)wtile.size();
    for (int s0 = cnt[tI]; s0 < cnt[tI + 1]; s0 += kWorkMax) {
      wtile.push_back((int)tI);
      wrange.push_back(make_int2(s0, std::min(s0 + kWorkMax, cnt[tI + 1])));
    }
  }
  twi[ntiles] = (int)wtile.size();
  const int nwi = (int)wtile.size();
  const int E = kTile + 2 * h->M;
  (void)hipFree(h->wi_tile); (void)hipFree(h->wi_range); (void)hipFree(h->tile_wi); (void)hipFree(h->scratch);
  h->wi_tile = nullptr; h->wi_range = nullptr; h->tile_wi = nullptr; h->scratch = nullptr;
  h->ready = false;
  PT_HIP(hipMalloc(&h->wi_tile, sizeof(int) * std::max(nwi, 1)));
  PT_HIP(hipMalloc(&h->wi_range, sizeof(int2) * std::max(nwi, 1)));
  PT_HIP(hipMalloc(&h->tile_wi, sizeof(int) * (ntiles + 1)));
  PT_HIP(hipMalloc(&h->scratch, sizeof(float2) * (size_t)std::max(nwi, 1) * h->pz * E * E));
  PT_HIP(hipMemcpyAsync(h->spos, pos.data(), sizeof(float2) * ns, hipMemcpyHostToDevice, st));
  PT_HIP(hipMemcpyAsync(h->sfloor, fl.data(), sizeof(int2) * ns, hipMemcpyHostToDevice, st));
  PT_HIP(hipMemcpyAsync(h->bin_sample, list.data(), sizeof(int) * ns, hipMemcpyHostToDevice, st));
  if (nwi) {
    PT_HIP(hipMemcpyAsync(h->wi_tile, wtile.data(), sizeof(int) * nwi, hipMemcpyHostToDevice, st));
    PT_HIP(hipMemcpyAsync(h->wi_range, wrange.data(), sizeof(int2) * nwi, hipMemcpyHostToDevice, st));
  }
  PT_HIP(hipMemcpyAsync(h->tile_wi, twi.data(), sizeof(int) * (ntiles + 1), hipMemcpyHostToDevice, st));
  PT_HIP(hipStreamSynchronize(st));
  h->ntile1 = nt1;
  h->nwi = nwi;
  h->ready = true;
  return PT_OK;
}

static int tomo_fwd_impl(pt_tomo* h, void* g, const void* f, float scale, pt_stream stream) {
  PT_REQUIRE(h && g && f, PT_ERR_ARG, "pt_tomo_fwd: null argument");
  PT_REQUIRE(h->ready, PT_ERR_STATE, "pt_tomo_fwd: set_theta not called");
  return run_fwd(h, (float2*)g, (const float2*)f, scale, (hipStream_t)stream);
}
static int tomo_adj_impl(pt_tomo* h, void* f, const void* g, float scale, pt_stream stream) {
  PT_REQUIRE(h && g && f, PT_ERR_ARG, "pt_tomo_adj: null argument");
  PT_REQUIRE(h->ready, PT_ERR_STATE, "pt_tomo_adj: set_theta not called");
  return run_adj(h, (float2*)f, (const float2*)g, scale, (hipStream_t)stream);
}

extern "C" int pt_tomo_fwd(pt_tomo* h, void* g, const void* f, pt_stream stream) {
  return tomo_fwd_impl(h, g, f, 1.0f, stream);
}
extern "C" int pt_tomo_adj(pt_tomo* h, void* f, const void* g, pt_stream stream) {
  return tomo_adj_impl(h, f, g, 1.0f, stream);
}
extern "C" int pt_tomo_fwd_scaled(pt_tomo* h, void* g, const void* f, float scale, pt_stream stream) {
  return tomo_fwd_impl(h, g, f, scale, stream);
}
extern "C" int pt_tomo_adj_scaled(pt_tomo* h, void* f, const void* g, float scale, pt_stream stream) {
  return tomo_adj_impl(h, f, g, scale, stream);
}
