// Mixed-precision kernels of the evaluate pass (gfx950).  Built with FP
// contraction: no rounding-order contract applies here.
//
//   near field (U, W, X fallbacks, M2T, L2T)  FP32 kernel evaluation in
//       coordinates relative to the target leaf centre, FP32 partial sums per
//       staged source block, FP64 running sums; near and far (L2T) parts summed
//       in FP64 at the end
//   S2M / S2L                                 same scheme, FP64 check values
//   V-list Hadamard                           FP32 spectra, FP32 FMA accumulation
//   translations and pseudo-inverse applies   FP64 GEMM with FMA
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "device.hpp"
#include "gemm.hpp"

namespace kifmm {

static KParams<float> to_f32(const KParams<double>& p) { return KParams<float>{float(p.a), float(p.b), float(p.c)}; }

// ---------------------------------------------------------------------------
// den: NR right-hand sides of SD values; acc[r][t].  The kernel block of a pair
// is formed once and applied to every right-hand side.
template <int KT, int NR>
__device__ inline void pair_f32(const KParams<float>& p, float dx, float dy, float dz, float nx, float ny,
                                float nz, const float* den, float (*acc)[trg_dof(KT)]) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  if constexpr (KT == 111) {
    const float r2 = dx * dx + dy * dy + dz * dz;
    const float g = r2 < float(kMinDif * kMinDif) ? 0.0f : p.a * rsqrtf(r2);
#pragma unroll
    for (int r = 0; r < NR; r++) acc[r][0] = fmaf(g, den[r], acc[r][0]);
  } else {
    float blk[SD * TD];
    kblock<KT, float>(p, dx, dy, dz, nx, ny, nz, blk);
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int t = 0; t < TD; t++)
#pragma unroll
        for (int s = 0; s < SD; s++) acc[r][t] = fmaf(blk[t * SD + s], den[r * SD + s], acc[r][t]);
  }
}

// Double-layer kernels: r and r.n from FP64 positions (r.n cancels for close
// pairs on a smooth surface), the rest of the kernel in FP32.
template <int KT, int NR>
__device__ inline void pair_f32_dl(const KParams<float>& p, double dx, double dy, double dz, float nx, float ny,
                                   float nz, const float* den, float (*acc)[trg_dof(KT)]) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  const float rn = float(dx * double(nx) + dy * double(ny) + dz * double(nz));
  float blk[SD * TD];
  kblock_rn<KT, float>(p, float(dx), float(dy), float(dz), nx, ny, nz, rn, blk);
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int t = 0; t < TD; t++)
#pragma unroll
      for (int s = 0; s < SD; s++) acc[r][t] = fmaf(blk[t * SD + s], den[r * SD + s], acc[r][t]);
}

// NR right-hand sides per launch (see NearArgs): the sources are staged once
// with all their densities.
template <int KT, int NR>
struct NearShared {
  static constexpr int KM = mm_kernel(KT);
  static constexpr int SD = src_dof(KT), TD = trg_dof(KT), sd = src_dof(KM);
  static constexpr int MAXD = SD > sd ? SD : sd;
  static constexpr bool NOR = needs_normals_ct(KT);
  float sp[kNearChunk * 3];
  float sn[NOR ? kNearChunk * 3 : 1];
  double sp64[NOR ? kNearChunk * 3 : 1];
  float sv[kNearChunk * MAXD * NR];
  double red[kNearChunk * TD * 2];
  int next;  // work queue: the chunk this workgroup takes next
};

template <int KT, int NR>
__device__ __forceinline__ void near_mx_chunk(const NearArgs a, int chunk, const KParams<float>& pu,
                                              const KParams<float>& pm, NearShared<KT, NR>& sh) {
  constexpr int KM = mm_kernel(KT);
  constexpr int SD = src_dof(KT), TD = trg_dof(KT), sd = src_dof(KM);
  constexpr bool NOR = needs_normals_ct(KT);
  float* sp = sh.sp;
  float* sn = sh.sn;
  double* sp64 = sh.sp64;
  float* sv = sh.sv;
  double* red = sh.red;
  const NearChunk ch = a.chunks[chunk];
  const double4 cb = a.geom[ch.node];
  const int lane = threadIdx.x;
  int S = kNearChunk / max(ch.tnum, 1);
  S = S > 16 ? 16 : S;
  const int ti = lane % ch.tnum, sub = lane / ch.tnum;
  const bool active = sub < S;
  float tx = 0, ty = 0, tz = 0;
  double tx64 = 0, ty64 = 0, tz64 = 0;
  if (active) {
    const double* p = a.tpos + 3 * int64_t(ch.tbeg + ti);
    tx64 = p[0]; ty64 = p[1]; tz64 = p[2];
    tx = float(p[0] - cb.x); ty = float(p[1] - cb.y); tz = float(p[2] - cb.z);
  }
  double near[NR][TD], far[NR][TD];
#pragma unroll
  for (int r = 0; r < NR; r++)
#pragma unroll
    for (int t = 0; t < TD; t++) near[r][t] = far[r][t] = 0.0;

  for (int it = ch.ibeg; it < ch.iend; it++) {
    const int code = a.items[it];
    const int node = code >> 2, kind = code & 3;
    float acc[NR][TD];
#pragma unroll
    for (int r = 0; r < NR; r++)
#pragma unroll
      for (int t = 0; t < TD; t++) acc[r][t] = 0.0f;
    if (kind == ITEM_EXACT) {
      const int b = a.src_beg[node], ns = a.src_num[node];
      for (int base = 0; base < ns; base += kNearChunk) {
        const int m = min(kNearChunk, ns - base);
        __syncthreads();
        if (lane < m) {
          const int64_t g = int64_t(b + base + lane);
          sp[3 * lane] = float(a.spos[3 * g] - cb.x);
          sp[3 * lane + 1] = float(a.spos[3 * g + 1] - cb.y);
          sp[3 * lane + 2] = float(a.spos[3 * g + 2] - cb.z);
          if constexpr (NOR)
            for (int d = 0; d < 3; d++) { sn[3 * lane + d] = float(a.snor[3 * g + d]); sp64[3 * lane + d] = a.spos[3 * g + d]; }
#pragma unroll
          for (int r = 0; r < NR; r++)
            for (int s = 0; s < SD; s++)
              sv[(lane * NR + r) * SD + s] = r < a.nr ? float(a.sden[(g * a.nrhs + a.r0 + r) * SD + s]) : 0.f;
        }
        __syncthreads();
        if (active) {
          for (int j = sub; j < m; j += S) {
            if constexpr (NOR) {
              pair_f32_dl<KT, NR>(pu, tx64 - sp64[3 * j], ty64 - sp64[3 * j + 1], tz64 - sp64[3 * j + 2], sn[3 * j],
                                  sn[3 * j + 1], sn[3 * j + 2], &sv[j * NR * SD], acc);
            } else {
              pair_f32<KT, NR>(pu, tx - sp[3 * j], ty - sp[3 * j + 1], tz - sp[3 * j + 2], 0.f, 0.f, 0.f,
                               &sv[j * NR * SD], acc);
            }
          }
        }
      }
#pragma unroll
      for (int r = 0; r < NR; r++)
#pragma unroll
        for (int t = 0; t < TD; t++) near[r][t] += double(acc[r][t]);
    } else {
      const double4 g = a.geom[node];
      const bool ue = kind == ITEM_UE;
      const double* sam = ue ? a.sam_ue : a.sam_de;
      const int ns = ue ? a.n_ue : a.n_de;
      const double* den = ue ? a.UE + int64_t(node) * a.sz_ue : a.DE + int64_t(node) * a.sz_de;
      const int64_t rstep = a.nstride * (ue ? a.sz_ue : a.sz_de);
      for (int base = 0; base < ns; base += kNearChunk) {
        const int m = min(kNearChunk, ns - base);
        __syncthreads();
        if (lane < m) {
          const int q = base + lane;
          sp[3 * lane] = float((g.x - cb.x) + g.w * sam[3 * q]);
          sp[3 * lane + 1] = float((g.y - cb.y) + g.w * sam[3 * q + 1]);
          sp[3 * lane + 2] = float((g.z - cb.z) + g.w * sam[3 * q + 2]);
#pragma unroll
          for (int r = 0; r < NR; r++)
            for (int s = 0; s < sd; s++)
              sv[(lane * NR + r) * sd + s] = r < a.nr ? float(den[(a.r0 + r) * rstep + q * sd + s]) : 0.f;
        }
        __syncthreads();
        if (active) {
          for (int j = sub; j < m; j += S) {
            const float x = sp[3 * j], y = sp[3 * j + 1], z = sp[3 * j + 2];
            pair_f32<KM, NR>(pm, tx - x, ty - y, tz - z, x, y, z, &sv[j * NR * sd], acc);
          }
        }
      }
      if (ue) {
#pragma unroll
        for (int r = 0; r < NR; r++)
#pragma unroll
          for (int t = 0; t < TD; t++) near[r][t] += double(acc[r][t]);
      } else {
#pragma unroll
        for (int r = 0; r < NR; r++)
#pragma unroll
          for (int t = 0; t < TD; t++) far[r][t] += double(acc[r][t]);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NR; r++) {
    __syncthreads();
    if (active && sub > 0)
      for (int t = 0; t < TD; t++) {
        red[((lane - ch.tnum) * TD + t) * 2] = near[r][t];
        red[((lane - ch.tnum) * TD + t) * 2 + 1] = far[r][t];
      }
    __syncthreads();
    if (lane < ch.tnum && r < a.nr) {
      double nv[TD], fv[TD];
#pragma unroll
      for (int t = 0; t < TD; t++) { nv[t] = near[r][t]; fv[t] = far[r][t]; }
      for (int u = 1; u < S; u++)
        for (int t = 0; t < TD; t++) {
          nv[t] += red[(((u - 1) * ch.tnum + lane) * TD + t) * 2];
          fv[t] += red[(((u - 1) * ch.tnum + lane) * TD + t) * 2 + 1];
        }
      double* o = a.tval + int64_t(a.r0 + r) * a.tstride + int64_t(ch.tbeg + lane) * TD;
#pragma unroll
      for (int t = 0; t < TD; t++) o[t] = nv[t] + fv[t];
    }
  }
}

// Grid or persistent work-queue launch, as k_near_lap_mx.
template <int KT, int NR>
__global__ __launch_bounds__(kNearChunk) void k_near_mx(NearArgs a, KParams<float> pu, KParams<float> pm) {
  __shared__ NearShared<KT, NR> sh;
  if (!a.queue) {
    const int c = int(blockIdx.x);
    near_mx_chunk<KT, NR>(a, a.order ? a.order[c] : c, pu, pm, sh);
    return;
  }
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) sh.next = atomicAdd(a.queue, 1);
    __syncthreads();
    const int c = sh.next;
    if (c >= a.nchunks) break;
    near_mx_chunk<KT, NR>(a, a.order ? a.order[c] : c, pu, pm, sh);
  }
}

// Laplace single layer (user and equivalent-density kernel 111).  The sources of
// a chunk's items are packed, item after item, into blocks of kNearChunk float4
// (x, y, z relative to the leaf centre, density / 4 pi): a lane finds its item
// from an LDS prefix of the item sizes (a wave-shuffle scan), walking the few
// items that start inside the block in a uniform loop, so the ~50 small items
// of a leaf (37 points on average) share a few staging rounds.  Exact and UE
// items are packed together (near part), the DE item (last in the list)
// separately (far part).  One broadcast LDS read per pair; lanes form S copies
// of the chunk's targets and each copy sweeps a contiguous share of the block
// (source count rounded up to 4, split S ways in multiples of 4), unrolled by 4
// with the same (uniform) trip count for every copy: the shares of the last
// copies run on into padding, which the block buffer carries up to kNearPad
// entries.  Padding entries have zero weight and lie 1e18 away, so they add
// exact zeros.
// Pairs closer than the reference's 1e-8 cut contribute nothing.  Only a
// source of the target leaf itself can come that close, unless the host found
// another item's sources within 1e-7 of the chunk's targets (then unchk = 0):
// blocks without the own item's sources use the sweep without the test (and
// the bare v_rsq_f32: r^2 >= 1e-16 there, a normal float).  The FP32 partial
// sums per block are added in FP64, near (exact, UE) and far (DE) parts
// separately.
__device__ __forceinline__ void lap_sweep(const float4* s4, int n4, float tx, float ty, float tz, float& acc0,
                                          float& acc1) {
  const float eps2 = float(kMinDif * kMinDif);
  for (int i = 0; i < n4; i++, s4 += 4) {
#pragma unroll
    for (int u = 0; u < 4; u += 2) {
      const float4 s0 = s4[u], s1 = s4[u + 1];
      const float dx0 = tx - s0.x, dy0 = ty - s0.y, dz0 = tz - s0.z;
      const float dx1 = tx - s1.x, dy1 = ty - s1.y, dz1 = tz - s1.z;
      const float r0 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;
      const float r1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
      acc0 = fmaf(r0 < eps2 ? 0.f : rsqrtf(r0), s0.w, acc0);
      acc1 = fmaf(r1 < eps2 ? 0.f : rsqrtf(r1), s1.w, acc1);
    }
  }
}

__device__ __forceinline__ void lap_sweep_nc(const float4* s4, int n4, float tx, float ty, float tz, float& acc0,
                                             float& acc1) {
  for (int i = 0; i < n4; i++, s4 += 4) {
#pragma unroll
    for (int u = 0; u < 4; u += 2) {
      const float4 s0 = s4[u], s1 = s4[u + 1];
      const float dx0 = tx - s0.x, dy0 = ty - s0.y, dz0 = tz - s0.z;
      const float dx1 = tx - s1.x, dy1 = ty - s1.y, dz1 = tz - s1.z;
      const float r0 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;
      const float r1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
      acc0 = fmaf(__builtin_amdgcn_rsqf(r0), s0.w, acc0);
      acc1 = fmaf(__builtin_amdgcn_rsqf(r1), s1.w, acc1);
    }
  }
}

// Several right-hand sides (NR > 1): the positions sit in s4 and the NR
// weights of a source side by side in sw; one rsqrt per pair serves all of
// them.  Each right-hand side accumulates in the order of lap_sweep (even and
// odd sources in two sums), so its values are those of a one-RHS launch.
template <int NR, bool CUT>
__device__ __forceinline__ void lap_sweep_nr(const float4* s4, const float* sw, int n4, float tx, float ty, float tz,
                                             float (&acc0)[NR], float (&acc1)[NR]) {
  const float eps2 = float(kMinDif * kMinDif);
  for (int i = 0; i < n4; i++, s4 += 4, sw += 4 * NR) {
#pragma unroll
    for (int u = 0; u < 4; u += 2) {
      const float4 s0 = s4[u], s1 = s4[u + 1];
      const float dx0 = tx - s0.x, dy0 = ty - s0.y, dz0 = tz - s0.z;
      const float dx1 = tx - s1.x, dy1 = ty - s1.y, dz1 = tz - s1.z;
      const float r0 = dx0 * dx0 + dy0 * dy0 + dz0 * dz0;
      const float r1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
      const float g0 = CUT ? (r0 < eps2 ? 0.f : rsqrtf(r0)) : __builtin_amdgcn_rsqf(r0);
      const float g1 = CUT ? (r1 < eps2 ? 0.f : rsqrtf(r1)) : __builtin_amdgcn_rsqf(r1);
#pragma unroll
      for (int r = 0; r < NR; r++) {
        acc0[r] = fmaf(g0, sw[u * NR + r], acc0[r]);
        acc1[r] = fmaf(g1, sw[(u + 1) * NR + r], acc1[r]);
      }
    }
  }
}

constexpr int kItemWin = 128;
typedef const __attribute__((address_space(1))) double* gdptr;
// An item of the staging window: slot j of the item lies at
// (cw[0..2]) + cw[3] * pos[3j .. 3j + 2] relative to the leaf centre (exact
// items: (-leaf centre, 1) and their source points, surface items: (node
// centre - leaf centre, radius) and the surface samples); den is its first
// density.
struct NearItem {
  double cw[4];
  gdptr pos;
  gdptr den;
};
// S copies of L = ceil4(ceil(m4 / S)) sources reach at most m4 + 4 S - 1 < 192
constexpr int kNearPad = kNearChunk + 64;
template <int NR>
struct LapShared {
  float4 s4[kNearPad];
  float sw[NR > 1 ? kNearPad * NR : 1];
  int pre[kItemWin + 1];
  unsigned char ikind_[kItemWin];
  int wsum[2];
  union {
    NearItem irec_[kItemWin];
    double red[kNearChunk * 2];  // the final reduction (after the last window)
  };
  int next;  // work queue: the chunk this workgroup takes next
};

template <int NR>
__device__ __forceinline__ void near_lap_chunk(const NearArgs a, int chunk, float pa_u, float pa_m,
                                               LapShared<NR>& sh) {
  float4* s4 = sh.s4;
  float* sw = sh.sw;
  int* pre = sh.pre;
  NearItem* irec_ = sh.irec_;
  unsigned char* ikind_ = sh.ikind_;
  int* wsum = sh.wsum;
  double* red = sh.red;
  const NearChunk ch = a.chunks[chunk];
  const double4 cb = a.geom[ch.node];
  const int lane = threadIdx.x;
  const int tn = max(ch.tnum, 1);
  int S = kNearChunk / tn;
  S = S > 16 ? 16 : S;
  const int ti = lane % tn, sub = lane / tn;
  const bool active = sub < S;
  float tx = 0, ty = 0, tz = 0;
  if (active) {
    const double* p = a.tpos + 3 * int64_t(ch.tbeg + ti);
    tx = float(p[0] - cb.x); ty = float(p[1] - cb.y); tz = float(p[2] - cb.z);
  }
  // padding past the block (read by the last copies' shares, never restaged)
  if (lane < kNearPad - kNearChunk) {
    s4[kNearChunk + lane] = make_float4(1e18f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int r = 0; r < NR; r++)
      if (NR > 1) sw[(kNearChunk + lane) * NR + r] = 0.f;
  }
  const int64_t rs_ue = a.nstride * a.sz_ue, rs_de = a.nstride * a.sz_de;  // right-hand-side steps
  const int idiv = ch.iend > ch.ibeg && (a.items[ch.iend - 1] & 3) == ITEM_DE ? ch.iend - 1 : ch.iend;
  // the near (exact, UE) and far (DE) sums; each pass accumulates into pp and
  // hands it over at its end (static indices: no scratch for a dynamic pass)
  double part[2][NR];
#pragma unroll
  for (int r = 0; r < NR; r++) part[0][r] = part[1][r] = 0.0;
  for (int pass = 0; pass < 2; pass++) {
    double pp[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) pp[r] = 0.0;
    const int i0 = pass ? idiv : ch.ibeg, i1 = pass ? ch.iend : idiv;
    for (int w0 = i0; w0 < i1; w0 += kItemWin) {
      const int nw = min(i1 - w0, kItemWin);
      __syncthreads();
      int sz = 0;
      if (lane < nw) {
        const int code = a.items[w0 + lane], node = code >> 2, kind = code & 3;
        const bool ex = kind == ITEM_EXACT, ue = kind == ITEM_UE;
        sz = ex ? a.src_num[node] : (ue ? a.n_ue : a.n_de);
        NearItem& it = irec_[lane];
        if (ex) {
          const int64_t b = a.src_beg[node];
          it.pos = (gdptr)(a.spos + 3 * b);
          it.den = (gdptr)(a.sden + (b * a.nrhs + a.r0));
          it.cw[0] = -cb.x; it.cw[1] = -cb.y; it.cw[2] = -cb.z; it.cw[3] = 1.0;
        } else {
          const double4 g = a.geom[node];
          it.pos = (gdptr)(ue ? a.sam_ue : a.sam_de);
          it.den = (gdptr)(ue ? a.UE + int64_t(node) * a.sz_ue + a.r0 * rs_ue
                              : a.DE + int64_t(node) * a.sz_de + a.r0 * rs_de);
          it.cw[0] = g.x - cb.x; it.cw[1] = g.y - cb.y; it.cw[2] = g.z - cb.z; it.cw[3] = g.w;
        }
        ikind_[lane] = (unsigned char)kind;
      }
      int inc = sz;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(inc, d, 64);
        if ((lane & 63) >= d) inc += y;
      }
      if ((lane & 63) == 63) wsum[lane >> 6] = inc;
      __syncthreads();
      if (lane >= 64) inc += wsum[0];
      pre[lane + 1] = inc;
      if (lane == 0) pre[0] = 0;
      __syncthreads();
      const int total = pre[nw];
      // slots of the target leaf's own item in this window (empty if absent)
      const int ow = ch.own >= 0 ? ch.ibeg + ch.own - w0 : -1;
      const int own0 = ow >= 0 && ow < nw ? pre[ow] : 0, own1 = ow >= 0 && ow < nw ? pre[ow + 1] : 0;
      // icur (uniform): the item of slot `base`, the last j with pre[j] <= base
      int icur = 0;
      for (int base = 0; base < total; base += kNearChunk) {
        const int m = min(kNearChunk, total - base);
        const bool cut = !ch.unchk || (own0 < base + m && own1 > base);
        // item of slot base + lane: the last j >= icur with pre[j] <= base + lane.
        // The items starting inside the block (2.7 on average for leaves of
        // ~40 points) are walked in a uniform loop.
        int lo = icur, j = icur + 1;
        for (; j < nw; j++) {
          const int sj = __builtin_amdgcn_readfirstlane(pre[j]);
          if (sj >= base + m) break;
          if (base + lane >= sj) lo = j;
        }
        int inext = j - 1;
        while (inext + 1 < nw && __builtin_amdgcn_readfirstlane(pre[inext + 1]) <= base + kNearChunk) inext++;
        icur = inext;
        __syncthreads();
        float4 v = make_float4(1e18f, 0.f, 0.f, 0.f);
        float wv[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) wv[r] = 0.f;
        if (lane < m) {
          const int p = base + lane;
          // one branch-free form for both kinds: exact items carry
          // (-centre, 1), surface items (node centre - leaf centre, radius)
          const unsigned off = unsigned(p - pre[lo]);
          const int kind = ikind_[lo];
          const bool ex = kind == ITEM_EXACT;
          const NearItem& it = irec_[lo];
          const gdptr ps = it.pos + ((off << 1) + off);
          const double w = it.cw[3];
          v = make_float4(float(fma(w, ps[0], it.cw[0])), float(fma(w, ps[1], it.cw[1])),
                          float(fma(w, ps[2], it.cw[2])), 0.f);
          const gdptr pd = it.den + (ex ? off * unsigned(a.nrhs) : off);
          const float sc = ex ? pa_u : pa_m;
#pragma unroll
          for (int r = 0; r < NR; r++)
            if (r < a.nr) wv[r] = float(pd[r * (ex ? int64_t(1) : (kind == ITEM_UE ? rs_ue : rs_de))]) * sc;
        }
        if constexpr (NR == 1) {
          v.w = wv[0];
        } else {
#pragma unroll
          for (int r = 0; r < NR; r++) sw[lane * NR + r] = wv[r];
        }
        s4[lane] = v;
        __syncthreads();
        if (active) {
          const int m4 = (m + 3) & ~3;
          const int L = ((m4 + S - 1) / S + 3) & ~3;
          const int j0 = sub * L, n4 = L >> 2;
          if constexpr (NR == 1) {
            float acc0 = 0.f, acc1 = 0.f;
            if (cut) lap_sweep(s4 + j0, n4, tx, ty, tz, acc0, acc1);
            else lap_sweep_nc(s4 + j0, n4, tx, ty, tz, acc0, acc1);
            pp[0] += double(acc0 + acc1);
          } else {
            float acc0[NR], acc1[NR];
#pragma unroll
            for (int r = 0; r < NR; r++) acc0[r] = acc1[r] = 0.f;
            if (cut) lap_sweep_nr<NR, true>(s4 + j0, sw + j0 * NR, n4, tx, ty, tz, acc0, acc1);
            else lap_sweep_nr<NR, false>(s4 + j0, sw + j0 * NR, n4, tx, ty, tz, acc0, acc1);
#pragma unroll
            for (int r = 0; r < NR; r++) pp[r] += double(acc0[r] + acc1[r]);
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if (pass == 0) part[0][r] = pp[r];
      else part[1][r] = pp[r];
    }
  }
#pragma unroll
  for (int r = 0; r < NR; r++) {
    __syncthreads();
    if (active && sub > 0) {
      red[(lane - tn) * 2] = part[0][r];
      red[(lane - tn) * 2 + 1] = part[1][r];
    }
    __syncthreads();
    if (lane < ch.tnum && r < a.nr) {
      double near = part[0][r], far = part[1][r];
      for (int u = 1; u < S; u++) {
        near += red[((u - 1) * tn + lane) * 2];
        far += red[((u - 1) * tn + lane) * 2 + 1];
      }
      const int64_t o = ch.tbeg + lane;
      a.tval[int64_t(a.r0 + r) * a.tstride + (a.tperm ? int64_t(a.tperm[o]) : o)] = near + far;
    }
  }
}

// One workgroup per chunk (a.queue null; chunk blockIdx.x of a.order, or of
// the chunk list), or a persistent grid whose workgroups take chunks from the
// global counter *a.queue (zeroed before the launch) in a.order until none is
// left: every workgroup leaves the loop when the counter passes a.nchunks.
template <int NR, bool QUEUE>
__global__ __launch_bounds__(kNearChunk) void k_near_lap_mx(NearArgs a, float pa_u, float pa_m) {
  __shared__ LapShared<NR> sh;
  if constexpr (!QUEUE) {
    const int c = int(blockIdx.x);
    near_lap_chunk<NR>(a, a.order ? a.order[c] : c, pa_u, pa_m, sh);
    return;
  }
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) sh.next = atomicAdd(a.queue, 1);
    __syncthreads();
    const int c = sh.next;
    if (c >= a.nchunks) break;
    near_lap_chunk<NR>(a, a.order ? a.order[c] : c, pa_u, pa_m, sh);
  }
}

// ---------------------------------------------------------------------------
constexpr int kSurfMx = 256;

template <int KT, int NR>
__global__ __launch_bounds__(kSurfMx) void k_surface_mx(SurfArgs a, KParams<float> pu) {
  constexpr int SD = src_dof(KT), TD = trg_dof(KT);
  constexpr bool NOR = needs_normals_ct(KT);
  // Laplace, one right-hand side: sources packed as float4 (x, y, z, density)
  // and swept four at a time (one broadcast LDS read a source, the same
  // summation order as the scalar loop); blocks padded to a multiple of 4 with
  // zero-weight points 1e18 away
  constexpr bool PACK = KT == 111 && NR == 1;
  __shared__ float sp[PACK ? 1 : kSurfMx * 3];
  __shared__ float4 s4[PACK ? kSurfMx : 1];
  __shared__ float sn[NOR ? kSurfMx * 3 : 1];
  __shared__ float sv[PACK ? 1 : kSurfMx * SD * NR];
  int node, ib, ie, sb1 = 0, sn1 = 0;
  if (a.leaf) {
    const int4 L = a.leaf[blockIdx.x];
    node = L.x; sb1 = L.y; sn1 = L.z; ib = 0; ie = 1;
  } else {
    const SurfTask tk = a.tasks[blockIdx.x];
    node = tk.node; ib = tk.ibeg; ie = tk.iend;
  }
  const double4 g = a.geom[node];
  const int lane = threadIdx.x;
  float tp[2][3];
  double acc[2][NR][TD];
  bool act[2];
  for (int u = 0; u < 2; u++) {
    const int q = lane + u * kSurfMx;
    act[u] = q < a.n_sam;
    const int qq = act[u] ? q : 0;
    for (int d = 0; d < 3; d++) tp[u][d] = float(g.w * a.sam[3 * qq + d]);
#pragma unroll
    for (int r = 0; r < NR; r++)
      for (int t = 0; t < TD; t++) acc[u][r][t] = 0.0;
  }
  for (int it = ib; it < ie; it++) {
    int b = sb1, ns = sn1;
    if (!a.leaf) {
      const int src = a.items[it];
      b = a.src_beg[src]; ns = a.src_num[src];
    }
    for (int base = 0; base < ns; base += kSurfMx) {
      const int m = min(kSurfMx, ns - base);
      __syncthreads();
      if constexpr (PACK) {
        if (lane < ((m + 3) & ~3)) {
          float4 v = make_float4(1e18f, 0.f, 0.f, 0.f);
          if (lane < m) {
            const int64_t gg = int64_t(b + base + lane);
            v = make_float4(float(a.spos[3 * gg] - g.x), float(a.spos[3 * gg + 1] - g.y),
                            float(a.spos[3 * gg + 2] - g.z),
                            float(a.sden[gg * a.nrhs + a.r0]) * pu.a);
          }
          s4[lane] = v;
        }
      } else if (lane < m) {
        const int64_t gg = int64_t(b + base + lane);
        sp[3 * lane] = float(a.spos[3 * gg] - g.x);
        sp[3 * lane + 1] = float(a.spos[3 * gg + 1] - g.y);
        sp[3 * lane + 2] = float(a.spos[3 * gg + 2] - g.z);
        if constexpr (NOR) for (int d = 0; d < 3; d++) sn[3 * lane + d] = float(a.snor[3 * gg + d]);
        // Laplace: the 1/(4 pi) factor goes with the density
#pragma unroll
        for (int r = 0; r < NR; r++)
          for (int s = 0; s < SD; s++)
            sv[(lane * NR + r) * SD + s] =
                r < a.nr ? float(a.sden[(gg * a.nrhs + a.r0 + r) * SD + s]) * (KT == 111 ? pu.a : 1.f) : 0.f;
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < 2; u++) {
        if (!act[u]) continue;
        float part[NR][TD];
#pragma unroll
        for (int r = 0; r < NR; r++)
          for (int t = 0; t < TD; t++) part[r][t] = 0.0f;
        if constexpr (PACK) {
          // (no pair within the 1e-8 cut, as below)
          const float tx = tp[u][0], ty = tp[u][1], tz = tp[u][2];
          float sum = 0.f;
          for (int j = 0; j < m; j += 4) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const float4 q = s4[j + k];
              const float dx = tx - q.x, dy = ty - q.y, dz = tz - q.z;
              sum = fmaf(__builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz), q.w, sum);
            }
          }
          part[0][0] = sum;
        } else if constexpr (KT == 111) {
          // check surfaces never meet their sources: the upward check surface
          // lies outside the leaf, and an X-list leaf is not adjacent to the
          // target node whose downward check surface is its boundary, so no
          // pair is within the 1e-8 cut and r^2 is a normal float
          for (int j = 0; j < m; j++) {
            const float dx = tp[u][0] - sp[3 * j], dy = tp[u][1] - sp[3 * j + 1], dz = tp[u][2] - sp[3 * j + 2];
            const float gk = __builtin_amdgcn_rsqf(dx * dx + dy * dy + dz * dz);
#pragma unroll
            for (int r = 0; r < NR; r++) part[r][0] = fmaf(gk, sv[j * NR + r], part[r][0]);
          }
        } else {
          for (int j = 0; j < m; j++) {
            float nx = 0, ny = 0, nz = 0;
            if constexpr (NOR) { nx = sn[3 * j]; ny = sn[3 * j + 1]; nz = sn[3 * j + 2]; }
            pair_f32<KT, NR>(pu, tp[u][0] - sp[3 * j], tp[u][1] - sp[3 * j + 1], tp[u][2] - sp[3 * j + 2], nx, ny, nz,
                             &sv[j * NR * SD], part);
          }
        }
#pragma unroll
        for (int r = 0; r < NR; r++)
          for (int t = 0; t < TD; t++) acc[u][r][t] += double(part[r][t]);
      }
    }
  }
  // one-leaf tasks (S2M) are the only writers of their node's check surface:
  // store; other tasks (S2L) add to what the V list already put there
  for (int u = 0; u < 2; u++) {
    if (!act[u]) continue;
    const int q = lane + u * kSurfMx;
#pragma unroll
    for (int r = 0; r < NR; r++) {
      if (r >= a.nr) continue;
      double* o = a.out + (int64_t(a.r0 + r) * a.nstride + node) * a.sz_out;
      for (int t = 0; t < TD; t++) {
        if (a.leaf) o[q * TD + t] = acc[u][r][t];
        else o[q * TD + t] += acc[u][r][t];
      }
    }
  }
}

// ---------------------------------------------------------------------------
typedef float tf2 __attribute__((ext_vector_type(2)));

// V-list Hadamard with FP32 spectra: one wave per sibling group and 64
// frequencies, lane = frequency, the 8 targets' accumulators per lane.  Spectra
// rows are read through buffer resources, the (uniform) row offset in the
// scalar offset and the lane's frequency in the vector offset, so the 16 loads
// of a pair need no vector address arithmetic (the translation table always
// fits 32-bit offsets; the source spectra when BUFX).  SYM: the table's DOF
// pairs (j, i) and (i, j), i, j < 3, hold the same spectra, so 9 of 12 (Stokes)
// or 6 of 9 (Navier) translation loads a pair serve all the products.
template <int SD, int TD, bool BUFX, bool SYM>
__global__ __launch_bounds__(64) void k_hadamard_mx(VlistArgs v, const float2* __restrict__ dspec,
                                                   const float2* __restrict__ m2l, float2* __restrict__ espec,
                                                   float nrm) {
  const int g = blockIdx.x / v.nrhs, rh = blockIdx.x % v.nrhs;  // virtual group g * nrhs + rh
  const int soff = rh * v.nsrc, toff = rh * v.ntrg;
  const int f = blockIdx.y * 64 + threadIdx.x;
  const int H = v.H;
  const bool ok = f < H;
  const int ff = ok ? f : 0;
  const unsigned H8 = unsigned(H) * 8u, ff8 = unsigned(ff) * 8u;
  const __amdgpu_buffer_rsrc_t rs_t =
      __builtin_amdgcn_make_buffer_rsrc((void*)m2l, (short)0, int(343u * SD * TD * H8), 0x00020000);
  const __amdgpu_buffer_rsrc_t rs_x = __builtin_amdgcn_make_buffer_rsrc(
      (void*)dspec, (short)0, int(unsigned(v.nsrc) * unsigned(v.nrhs) * SD * H8), 0x00020000);
  auto ldt = [&](int row) -> float2 {
    const tf2 t = __builtin_bit_cast(tf2, __builtin_amdgcn_raw_buffer_load_b64(rs_t, ff8, unsigned(row) * H8, 0));
    return make_float2(t.x, t.y);
  };
  auto ldx = [&](int row) -> float2 {
    if constexpr (BUFX) {
      const tf2 x = __builtin_bit_cast(tf2, __builtin_amdgcn_raw_buffer_load_b64(rs_x, ff8, unsigned(row) * H8, 0));
      return make_float2(x.x, x.y);
    } else {
      return dspec[int64_t(row) * H + ff];
    }
  };
  float2 acc[8][TD];
#pragma unroll
  for (int t = 0; t < 8; t++)
#pragma unroll
    for (int i = 0; i < TD; i++) acc[t][i] = make_float2(0.f, 0.f);
  const int cnt = v.pcnt[g];
  const int* ps = v.psrc + int64_t(g) * kPairStride;
  const int* pr = v.pslot + int64_t(g) * kPairStride;
  for (int k = 0; k < cnt; k++) {
    const int s = ps[k], word = pr[k];
    if (s < 0) continue;
    float2 d[SD];
#pragma unroll
    for (int j = 0; j < SD; j++) d[j] = ldx((s + soff) * SD + j);
    const int rs = (word >> 8) & 0x3ff, tm = word >> 20;
#pragma unroll
    for (int tau = 0; tau < 8; tau++) {
      if (!((tm >> tau) & 1)) continue;
      const int slot = rs - ((tau >> 2) * 49 + ((tau >> 1) & 1) * 7 + (tau & 1));
      float2 tv[SD][TD];
#pragma unroll
      for (int j = 0; j < SD; j++)
#pragma unroll
        for (int i = 0; i < TD; i++)
          if (!(SYM && i < j && j < 3)) tv[j][i] = ldt(slot * (SD * TD) + j * TD + i);
      if constexpr (SYM) {
#pragma unroll
        for (int j = 1; j < 3; j++)
#pragma unroll
          for (int i = 0; i < j; i++) tv[j][i] = tv[i][j];
      }
#pragma unroll
      for (int i = 0; i < TD; i++)
#pragma unroll
        for (int j = 0; j < SD; j++) {
          const float2 t = tv[j][i];
          acc[tau][i].x = fmaf(t.x, d[j].x, fmaf(-t.y, d[j].y, acc[tau][i].x));
          acc[tau][i].y = fmaf(t.x, d[j].y, fmaf(t.y, d[j].x, acc[tau][i].y));
        }
    }
  }
  if (!ok) return;
#pragma unroll
  for (int tau = 0; tau < 8; tau++) {
    const int t = v.tgt[g * 8 + tau];
    if (t < 0) continue;
#pragma unroll
    for (int i = 0; i < TD; i++)
      espec[(int64_t(t + toff) * TD + i) * H + f] = make_float2(acc[tau][i].x * nrm, acc[tau][i].y * nrm);
  }
}

// ---------------------------------------------------------------------------
// Scalar-kernel Hadamard (sd = td = 1) over tiles of sibling groups.  The V list
// of a target child tau of parent P is every source child c of a colleague Q of
// P (Q = P + q, q in [-1,1]^3, q != 0) that is not adjacent to tau; the
// translation slot is 2q + (c - tau), and the 27 slots with |offset| <= 1 hold
// zeros, so the whole halo can be swept with no per-pair lists.  Lane =
// frequency (64 per wave), so every index, halo lookup and branch is uniform
// across the wave; for one colleague offset q the 27 translation spectra
// 2q + [-1,1]^3 sit in registers and serve all kTileG groups of the tile.
//
// * The halo rows of the tile's groups (216 source slots each) are staged in
//   LDS once, so a step's eight source indices come from LDS (a short, fixed
//   latency) instead of scalar loads that miss the scalar cache on a table of
//   ngrp * 864 bytes.
// * The source spectra of the next (q, group) are loaded while the current
//   group's multiply-adds run.
// * Three real accumulators per complex output (Gauss's three-product form).
//   With T = a + ib and X = c + id: A1 += a c, A2 += b d, A3 += (a + b)(c + d),
//   so Re = A1 - A2 = sum (ac - bd) and Im = A3 - A1 - A2 = sum (ad + bc).
//   a + b is formed once per translation spectrum and q, c + d once per source
//   spectrum and step: 3 FMAs per complex multiply-add instead of 4, and one
//   extra add per source.
//
// Absent halo sources (-1) are skipped; targets outside the V-list target set
// are computed and dropped.  Frequency chunks are the slow grid dimension and
// each XCD (workgroup id mod 8) sweeps a contiguous run of tiles.
//
// BUF: spectra addressed through buffer resources, the row offset (uniform) in
// the scalar offset and the lane's frequency in the vector offset, so the
// loads need no vector address arithmetic (both tables must fit in 32-bit byte
// offsets).
//
// MIRROR: the table's spectra of opposite slots are conjugates (an even kernel
// on a grid without its unpaired wrap displacement), so offset -q (= 26 - q)
// runs right after q on the same registers: slot -2q + delta reads register
// 26 - d with the imaginary part negated (A2 takes -d, A3 takes a - b), and each
// pair of offsets loads 27 translation spectra instead of 54.
constexpr int kTileG = 5;

template <bool BUF, bool MIRROR>
__global__ __launch_bounds__(64) void k_m2l_tile(VlistArgs v, const float2* __restrict__ dspec,
                                                 const float2* __restrict__ m2l, float2* __restrict__ espec,
                                                 float nrm, int ntile, int per_xcd) {
  constexpr int G = kTileG;
  const int bx = blockIdx.x;
  const int tile = (bx & 7) * per_xcd + (bx >> 3);
  if (tile >= ntile) return;
  const int H = v.H;
  const int f = blockIdx.y * 64 + threadIdx.x;
  const bool fok = f < H;
  const int fs = fok ? f : H - 1;
  // virtual groups g * nrhs + r: right-hand side r of sibling group g reads
  // source spectra r * nsrc + slot and writes target spectra r * ntrg + slot,
  // and consecutive virtual groups share the tile's translation spectra
  const int g0 = tile * G;
  const int ng = min(G, v.ngrp * v.nrhs - g0);
  const unsigned H8 = unsigned(H) * 8u, fs8 = unsigned(fs) * 8u;
  // (unused when !BUF)
  const __amdgpu_buffer_rsrc_t rs_x = __builtin_amdgcn_make_buffer_rsrc(
      (void*)dspec, (short)0, int(unsigned(v.nsrc) * unsigned(v.nrhs) * H8), 0x00020000);
  const __amdgpu_buffer_rsrc_t rs_t =
      __builtin_amdgcn_make_buffer_rsrc((void*)m2l, (short)0, int(343u * H8), 0x00020000);
  auto ldx = [&](int row) -> float2 {
    if constexpr (BUF) {
      const tf2 x = __builtin_bit_cast(tf2, __builtin_amdgcn_raw_buffer_load_b64(rs_x, fs8, unsigned(row) * H8, 0));
      return make_float2(x.x, x.y);
    } else {
      return dspec[int64_t(row) * H + fs];
    }
  };
  auto ldt = [&](int slot) -> float2 {
    if constexpr (BUF) {
      const tf2 t = __builtin_bit_cast(tf2, __builtin_amdgcn_raw_buffer_load_b64(rs_t, fs8, unsigned(slot) * H8, 0));
      return make_float2(t.x, t.y);
    } else {
      return m2l[int64_t(slot) * H + fs];
    }
  };
  __shared__ int shal[G * 216];
  for (int i = threadIdx.x; i < G * 216; i += 64) {
    int h = -1;
    if (i < ng * 216) {
      const int vg = g0 + i / 216;
      h = v.halo[int64_t(vg / v.nrhs) * 216 + i % 216];
      if (h >= 0) h += (vg % v.nrhs) * v.nsrc;
    }
    shal[i] = h;
  }
  __syncthreads();
  float k1[G][8], k2[G][8], k3[G][8];
#pragma unroll
  for (int gi = 0; gi < G; gi++)
#pragma unroll
    for (int t = 0; t < 8; t++) k1[gi][t] = k2[gi][t] = k3[gi][t] = 0.f;
  auto halo_base = [](int q) {
    const int qx = q / 9 - 1, qy = (q / 3) % 3 - 1, qz = q % 3 - 1;
    return (2 * qx + 2) * 36 + (2 * qy + 2) * 6 + (2 * qz + 2);
  };
  // halo source slots of group gi for colleague offset q (wave-uniform), and
  // their spectra
  auto fetch = [&](int q, int gi, int* si, float2* x) {
    const int* hal = shal + gi * 216 + halo_base(q);
#pragma unroll
    for (int c = 0; c < 8; c++)
      si[c] = __builtin_amdgcn_readfirstlane(hal[(c >> 2) * 36 + ((c >> 1) & 1) * 6 + (c & 1)]);
#pragma unroll
    for (int c = 0; c < 8; c++) x[c] = ldx(si[c] < 0 ? 0 : si[c]);
  };
  int sc[8];
  float2 xc[8];
  float ta[27], tb[27], ts[27];
  // the 27 translation spectra 2q + [-1,1]^3 of colleague offset q
  auto load_t = [&](int q) {
    const int qx = q / 9 - 1, qy = (q / 3) % 3 - 1, qz = q % 3 - 1;
    const int sbase = (2 * qx + 3) * 49 + (2 * qy + 3) * 7 + (2 * qz + 3);
#pragma unroll
    for (int d = 0; d < 27; d++) {
      const int sl = sbase + (d / 9 - 1) * 49 + ((d / 3) % 3 - 1) * 7 + (d % 3 - 1);
      const float2 t = ldt(sl);
      ta[d] = t.x;
      tb[d] = t.y;
      ts[d] = t.x + t.y;
    }
  };
  // the G groups' products for colleague offset q; qnext: the offset whose
  // first group's sources are loaded during the last group (-1: none).  With
  // mirrored (MIRROR), q is the opposite of the offset whose translation
  // spectra are in the registers: the spectrum of 2q + delta is the conjugate
  // of the one of -(2q + delta), register 26 - d, and ts must hold ta - tb.
  auto sweep = [&](auto mirrored, int q, int qnext) {
    constexpr bool M = decltype(mirrored)::value;
#pragma unroll
    for (int gi = 0; gi < G; gi++) {
      int sn[8];
      float2 xn[8];
      const int qn = gi + 1 < G ? q : qnext;
      if (qn >= 0) fetch(qn, gi + 1 < G ? gi + 1 : 0, sn, xn);
#pragma unroll
      for (int c = 0; c < 8; c++) {
        if (sc[c] < 0) continue;
        const float xr = xc[c].x, xi = xc[c].y, xs = xc[c].x + xc[c].y;
#pragma unroll
        for (int t = 0; t < 8; t++) {
          const int d = ((c >> 2) - (t >> 2) + 1) * 9 + (((c >> 1) & 1) - ((t >> 1) & 1) + 1) * 3 +
                        ((c & 1) - (t & 1) + 1);
          const int e = M ? 26 - d : d;
          k1[gi][t] = fmaf(xr, ta[e], k1[gi][t]);
          k2[gi][t] = fmaf(M ? -xi : xi, tb[e], k2[gi][t]);
          k3[gi][t] = fmaf(xs, ts[e], k3[gi][t]);
        }
      }
      if (qn >= 0) {
#pragma unroll
        for (int c = 0; c < 8; c++) { sc[c] = sn[c]; xc[c] = xn[c]; }
      }
    }
  };
  fetch(0, 0, sc, xc);
  if constexpr (MIRROR) {
    // offsets q and 26 - q (= -q) in turn on one load of translation spectra
    for (int q = 0; q < 13; q++) {
      load_t(q);
      sweep(std::false_type{}, q, 26 - q);
#pragma unroll
      for (int d = 0; d < 27; d++) ts[d] = ta[d] - tb[d];
      sweep(std::true_type{}, 26 - q, q + 1 < 13 ? q + 1 : -1);
    }
  } else {
    for (int q = 0; q < 27; q++) {
      if (q == 13) continue;  // the parent itself: siblings are adjacent
      load_t(q);
      const int qn = q + 1 == 13 ? 14 : q + 1;
      sweep(std::false_type{}, q, qn < 27 ? qn : -1);
    }
  }
  if (!fok) return;
#pragma unroll
  for (int gi = 0; gi < G; gi++) {
    if (gi >= ng) break;
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const int vg = g0 + gi;
      int tg = v.tgt[(vg / v.nrhs) * 8 + t];
      if (tg >= 0) tg += (vg % v.nrhs) * v.ntrg;
      if (tg >= 0)
        espec[int64_t(tg) * H + f] =
            make_float2((k1[gi][t] - k2[gi][t]) * nrm, (k3[gi][t] - k1[gi][t] - k2[gi][t]) * nrm);
    }
  }
}

// ---------------------------------------------------------------------------
// Vector-kernel Hadamard (sd, td > 1) as the same halo sweep: for colleague
// offset q and one (source DOF j, target DOF i) component, the 27 translation
// spectra 2q + [-1,1]^3 sit in registers and serve the 8 x 8 (source child,
// target child) pairs of all G groups of the tile; the groups' source spectra of
// DOF j are loaded once per (q, j) and reused for every i.  Per frequency, a
// translation value feeds 8 G complex multiply-adds (the per-pair kernel
// k_hadamard_mx reads one per multiply-add).  Pairs of adjacent children fall
// on the zero spectra of the table's |offset| <= 1 slots, as in k_m2l_tile.
template <int SD, int TD, int G, bool BUFX, bool MIRROR>
__global__ __launch_bounds__(64) void k_m2l_tile_vec(VlistArgs v, const float2* __restrict__ dspec,
                                                     const float2* __restrict__ m2l, float2* __restrict__ espec,
                                                     float nrm, int ntile, int per_xcd) {
  const int bx = blockIdx.x;
  const int tile = (bx & 7) * per_xcd + (bx >> 3);
  if (tile >= ntile) return;
  const int H = v.H;
  const int f = blockIdx.y * 64 + threadIdx.x;
  const bool fok = f < H;
  const int fs = fok ? f : H - 1;
  const int g0 = tile * G;
  const int ng = min(G, v.ngrp * v.nrhs - g0);
  // spectra rows through buffer resources: the row offset (uniform) in the
  // scalar offset, the lane's frequency in the vector offset, so no load needs
  // vector address arithmetic (the translation table always fits 32-bit
  // offsets; the source spectra when BUFX)
  const unsigned H8 = unsigned(H) * 8u, fs8 = unsigned(fs) * 8u;
  const __amdgpu_buffer_rsrc_t rs_t =
      __builtin_amdgcn_make_buffer_rsrc((void*)m2l, (short)0, int(343u * SD * TD * H8), 0x00020000);
  const __amdgpu_buffer_rsrc_t rs_x = __builtin_amdgcn_make_buffer_rsrc(
      (void*)dspec, (short)0, int(unsigned(v.nsrc) * unsigned(v.nrhs) * SD * H8), 0x00020000);
  auto ldt = [&](int row) -> float2 {
    const tf2 t = __builtin_bit_cast(tf2, __builtin_amdgcn_raw_buffer_load_b64(rs_t, fs8, unsigned(row) * H8, 0));
    return make_float2(t.x, t.y);
  };
  auto ldx = [&](int row) -> float2 {
    if constexpr (BUFX) {
      const tf2 x = __builtin_bit_cast(tf2, __builtin_amdgcn_raw_buffer_load_b64(rs_x, fs8, unsigned(row) * H8, 0));
      return make_float2(x.x, x.y);
    } else {
      return dspec[int64_t(row) * H + fs];
    }
  };
  __shared__ int shal[G * 216];
  for (int i = threadIdx.x; i < G * 216; i += 64) {
    int h = -1;
    if (i < ng * 216) {
      const int vg = g0 + i / 216;
      h = v.halo[int64_t(vg / v.nrhs) * 216 + i % 216];
      if (h >= 0) h += (vg % v.nrhs) * v.nsrc;
    }
    shal[i] = h;
  }
  __syncthreads();
  float2 acc[G][8][TD];
#pragma unroll
  for (int gi = 0; gi < G; gi++)
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
      for (int i = 0; i < TD; i++) acc[gi][t][i] = make_float2(0.f, 0.f);
  // colleague offset q: a bit per present (group, source child), and the
  // sources' spectra of DOF j
  auto presence = [&](int q) -> unsigned {
    const int qx = q / 9 - 1, qy = (q / 3) % 3 - 1, qz = q % 3 - 1;
    const int hbase = (2 * qx + 2) * 36 + (2 * qy + 2) * 6 + (2 * qz + 2);
    unsigned pres = 0;
#pragma unroll
    for (int gi = 0; gi < G; gi++)
#pragma unroll
      for (int c = 0; c < 8; c++)
        if (shal[gi * 216 + hbase + (c >> 2) * 36 + ((c >> 1) & 1) * 6 + (c & 1)] >= 0) pres |= 1u << (gi * 8 + c);
    return __builtin_amdgcn_readfirstlane(pres);
  };
  auto load_x = [&](int q, int j, float2 (&x)[G][8]) {
    const int qx = q / 9 - 1, qy = (q / 3) % 3 - 1, qz = q % 3 - 1;
    const int hbase = (2 * qx + 2) * 36 + (2 * qy + 2) * 6 + (2 * qz + 2);
#pragma unroll
    for (int gi = 0; gi < G; gi++)
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const int sc = __builtin_amdgcn_readfirstlane(
            shal[gi * 216 + hbase + (c >> 2) * 36 + ((c >> 1) & 1) * 6 + (c & 1)]);
        x[gi][c] = ldx((sc < 0 ? 0 : sc) * SD + j);
      }
  };
  // the 27 translation spectra 2q + [-1,1]^3 of DOFs (j, i)
  auto load_t = [&](int q, int j, int i, float2 (&T)[27]) {
    const int qx = q / 9 - 1, qy = (q / 3) % 3 - 1, qz = q % 3 - 1;
    const int sbase = (2 * qx + 3) * 49 + (2 * qy + 3) * 7 + (2 * qz + 3);
#pragma unroll
    for (int d = 0; d < 27; d++) {
      const int sl = sbase + (d / 9 - 1) * 49 + ((d / 3) % 3 - 1) * 7 + (d % 3 - 1);
      T[d] = ldt(sl * (SD * TD) + j * TD + i);
    }
  };
  // products of target DOF i; mirrored: the offset is the opposite of the one
  // whose spectra T holds, so the spectrum of 2q + delta is the conjugate of
  // T[26 - d]
  auto products = [&](auto mirrored, unsigned pres, const float2 (&x)[G][8], const float2 (&T)[27], int i) {
    constexpr bool M = decltype(mirrored)::value;
#pragma unroll
    for (int gi = 0; gi < G; gi++)
#pragma unroll
      for (int c = 0; c < 8; c++) {
        if (!((pres >> (gi * 8 + c)) & 1)) continue;
        const float2 xv = x[gi][c];
#pragma unroll
        for (int t = 0; t < 8; t++) {
          const int d = ((c >> 2) - (t >> 2) + 1) * 9 + (((c >> 1) & 1) - ((t >> 1) & 1) + 1) * 3 +
                        ((c & 1) - (t & 1) + 1);
          const int e = M ? 26 - d : d;
          const float tx = T[e].x, ty = M ? -T[e].y : T[e].y;
          acc[gi][t][i].x = fmaf(tx, xv.x, fmaf(-ty, xv.y, acc[gi][t][i].x));
          acc[gi][t][i].y = fmaf(tx, xv.y, fmaf(ty, xv.x, acc[gi][t][i].y));
        }
      }
  };
  if constexpr (MIRROR) {
    // offsets q and 26 - q (= -q) on one load of translation spectra
    for (int q = 0; q < 13; q++) {
      const unsigned pq = presence(q), pm = presence(26 - q);
      if (!(pq | pm)) continue;  // no source in either colleague
      for (int j = 0; j < SD; j++) {
        float2 xq[G][8], xm[G][8];
        if (pq) load_x(q, j, xq);
        if (pm) load_x(26 - q, j, xm);
#pragma unroll
        for (int i = 0; i < TD; i++) {
          float2 T[27];
          load_t(q, j, i, T);
          if (pq) products(std::false_type{}, pq, xq, T, i);
          if (pm) products(std::true_type{}, pm, xm, T, i);
        }
      }
    }
  } else {
    for (int q = 0; q < 27; q++) {
      if (q == 13) continue;  // the parent itself: siblings are adjacent
      const unsigned pres = presence(q);
      if (!pres) continue;  // no source in this colleague
      for (int j = 0; j < SD; j++) {
        float2 x[G][8];
        load_x(q, j, x);
#pragma unroll
        for (int i = 0; i < TD; i++) {
          float2 T[27];
          load_t(q, j, i, T);
          products(std::false_type{}, pres, x, T, i);
        }
      }
    }
  }
  if (!fok) return;
#pragma unroll
  for (int gi = 0; gi < G; gi++) {
    if (gi >= ng) break;
    const int vg = g0 + gi;
#pragma unroll
    for (int t = 0; t < 8; t++) {
      int tg = v.tgt[(vg / v.nrhs) * 8 + t];
      if (tg < 0) continue;
      tg += (vg % v.nrhs) * v.ntrg;
#pragma unroll
      for (int i = 0; i < TD; i++)
        espec[(int64_t(tg) * TD + i) * H + f] = make_float2(acc[gi][t][i].x * nrm, acc[gi][t][i].y * nrm);
    }
  }
}

// ---------------------------------------------------------------------------
#define KIFMM_USER_KERNELS_MX(X) X(111) X(121) X(311) X(312) X(321) X(322) X(511) X(521)

__global__ void k_zero_int(int* p) { *p = 0; }

// Resident workgroups of a near kernel over the whole device (persistent grid).
template <typename K>
static unsigned persistent_blocks(K kernel) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kNearChunk, 0) != hipSuccess)
    return 0;
  return unsigned(std::max(1, cus) * std::max(1, per));
}

hipError_t launch_near_mixed(int kt, const NearArgs& a0, const KParams<double>& pu, const KParams<double>& pm,
                             hipStream_t s) {
  if (a0.nchunks <= 0) return hipSuccess;
  if (a0.nrhs < 1) return hipErrorInvalidValue;
  const KParams<float> fu = to_f32(pu), fm = to_f32(pm);
  const dim3 b(kNearChunk);
  for (int r0 = 0; r0 < a0.nrhs; r0 += kMaxNr) {
    NearArgs a = a0;
    a.r0 = r0;
    a.nr = std::min(kMaxNr, a0.nrhs - r0);
    const int nr = rhs_group(a.nr);
    dim3 g(unsigned(a0.nchunks));
    if (a.queue) {
      // a grid of resident workgroups over the work queue (at most one per chunk)
      unsigned pb = 0;
      if (kt == 111) {
        switch (nr) {
          case 1: pb = persistent_blocks(k_near_lap_mx<1, true>); break;
          case 2: pb = persistent_blocks(k_near_lap_mx<2, true>); break;
          case 4: pb = persistent_blocks(k_near_lap_mx<4, true>); break;
          default: pb = persistent_blocks(k_near_lap_mx<8, true>); break;
        }
      } else {
        switch (kt * 16 + nr) {
#define KIFMM_CASE_NR(K, R) case K * 16 + R: pb = persistent_blocks(k_near_mx<K, R>); break;
#define KIFMM_CASE(K) KIFMM_CASE_NR(K, 1) KIFMM_CASE_NR(K, 2) KIFMM_CASE_NR(K, 4) KIFMM_CASE_NR(K, 8)
          KIFMM_USER_KERNELS_MX(KIFMM_CASE)
#undef KIFMM_CASE
#undef KIFMM_CASE_NR
          default: return hipErrorInvalidValue;
        }
      }
      if (pb == 0) return hipErrorInvalidValue;
      g = dim3(std::min(pb, unsigned(a0.nchunks)));
      k_zero_int<<<1, 1, 0, s>>>(a.queue);
    }
    if (kt == 111) {
      if (a.queue) {
        switch (nr) {
          case 1: k_near_lap_mx<1, true><<<g, b, 0, s>>>(a, fu.a, fm.a); break;
          case 2: k_near_lap_mx<2, true><<<g, b, 0, s>>>(a, fu.a, fm.a); break;
          case 4: k_near_lap_mx<4, true><<<g, b, 0, s>>>(a, fu.a, fm.a); break;
          default: k_near_lap_mx<8, true><<<g, b, 0, s>>>(a, fu.a, fm.a); break;
        }
      } else {
        switch (nr) {
          case 1: k_near_lap_mx<1, false><<<g, b, 0, s>>>(a, fu.a, fm.a); break;
          case 2: k_near_lap_mx<2, false><<<g, b, 0, s>>>(a, fu.a, fm.a); break;
          case 4: k_near_lap_mx<4, false><<<g, b, 0, s>>>(a, fu.a, fm.a); break;
          default: k_near_lap_mx<8, false><<<g, b, 0, s>>>(a, fu.a, fm.a); break;
        }
      }
    } else {
      switch (kt * 16 + nr) {
#define KIFMM_CASE_NR(K, R) case K * 16 + R: k_near_mx<K, R><<<g, b, 0, s>>>(a, fu, fm); break;
#define KIFMM_CASE(K) KIFMM_CASE_NR(K, 1) KIFMM_CASE_NR(K, 2) KIFMM_CASE_NR(K, 4) KIFMM_CASE_NR(K, 8)
        KIFMM_USER_KERNELS_MX(KIFMM_CASE)
#undef KIFMM_CASE
#undef KIFMM_CASE_NR
        default: return hipErrorInvalidValue;
      }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_surface_mixed(int kt, const SurfArgs& a0, const KParams<double>& pu, hipStream_t s) {
  if (a0.ntasks <= 0) return hipSuccess;
  if (a0.n_sam > 2 * kSurfMx || a0.nrhs < 1) return hipErrorInvalidValue;
  const KParams<float> fu = to_f32(pu);
  const dim3 g(unsigned(a0.ntasks)), b(kSurfMx);
  for (int r0 = 0; r0 < a0.nrhs; r0 += kMaxNr) {
    SurfArgs a = a0;
    a.r0 = r0;
    a.nr = std::min(kMaxNr, a0.nrhs - r0);
    switch (kt * 16 + rhs_group(a.nr)) {
#define KIFMM_CASE_NR(K, R) case K * 16 + R: k_surface_mx<K, R><<<g, b, 0, s>>>(a, fu); break;
#define KIFMM_CASE(K) KIFMM_CASE_NR(K, 1) KIFMM_CASE_NR(K, 2) KIFMM_CASE_NR(K, 4) KIFMM_CASE_NR(K, 8)
      KIFMM_USER_KERNELS_MX(KIFMM_CASE)
#undef KIFMM_CASE
#undef KIFMM_CASE_NR
      default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_hadamard_mixed(const VlistArgs& v, const float2* dspec, const float2* m2l, float2* espec,
                                 hipStream_t s) {
  if (v.ngrp <= 0) return hipSuccess;
  if (v.nrhs < 1) return hipErrorInvalidValue;
  const float nrm = float(1.0 / double(v.n * v.n * v.n));
  const int64_t nvg = int64_t(v.ngrp) * v.nrhs;  // virtual groups
  dim3 grid(unsigned(nvg), unsigned((v.H + 63) / 64));
  if (v.sd == 1 && v.td == 1) {
    const int ntile = int((nvg + kTileG - 1) / kTileG);
    const int per_xcd = (ntile + 7) / 8;
    const dim3 grid2(unsigned(8 * per_xcd), unsigned((v.H + 63) / 64));
    const bool buf = uint64_t(v.nsrc) * v.nrhs * v.H * 8 <= 0xFFFFFF00ull;
    if (v.mirror) {
      if (buf) k_m2l_tile<true, true><<<grid2, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm, ntile, per_xcd);
      else k_m2l_tile<false, true><<<grid2, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm, ntile, per_xcd);
    } else {
      if (buf) k_m2l_tile<true, false><<<grid2, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm, ntile, per_xcd);
      else k_m2l_tile<false, false><<<grid2, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm, ntile, per_xcd);
    }
  }
  else if (v.had_form == 0) {  // one wave per group, translation spectra read per pair
    // source spectra through a buffer resource when they fit 32-bit offsets
    const bool bufx = uint64_t(v.nsrc) * v.nrhs * v.sd * v.H * 8 <= 0xFFFFFF00ull;
#define KIFMM_PAIR_B(SD_, TD_, SYM_)                                                                          \
  if (bufx) k_hadamard_mx<SD_, TD_, true, SYM_><<<grid, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm);         \
  else k_hadamard_mx<SD_, TD_, false, SYM_><<<grid, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm);
#define KIFMM_PAIR(SD_, TD_)                                                                                  \
  if (v.sd == SD_ && v.td == TD_) {                                                                           \
    if (v.sym3) { KIFMM_PAIR_B(SD_, TD_, true) }                                                              \
    else { KIFMM_PAIR_B(SD_, TD_, false) }                                                                    \
    return hipGetLastError();                                                                                  \
  }
    KIFMM_PAIR(4, 3)
    KIFMM_PAIR(3, 3)
    KIFMM_PAIR(3, 4)
#undef KIFMM_PAIR
#undef KIFMM_PAIR_B
    return hipErrorInvalidValue;
  } else {  // halo sweep over tiles of 1 or 2 groups
    const int G = v.had_form == 1 ? 1 : 2;
    const int ntile = int((nvg + G - 1) / G);
    const int per_xcd = (ntile + 7) / 8;
    const dim3 grid2(unsigned(8 * per_xcd), unsigned((v.H + 63) / 64));
    // source spectra through a buffer resource when they fit 32-bit offsets
    const bool bufx = uint64_t(v.nsrc) * v.nrhs * v.sd * v.H * 8 <= 0xFFFFFF00ull;
#define KIFMM_TILE_GM(SD_, TD_, G_, M_)                                                                          \
  if (bufx) k_m2l_tile_vec<SD_, TD_, G_, true, M_><<<grid2, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm, ntile, per_xcd); \
  else k_m2l_tile_vec<SD_, TD_, G_, false, M_><<<grid2, dim3(64), 0, s>>>(v, dspec, m2l, espec, nrm, ntile, per_xcd);
#define KIFMM_TILE_G(SD_, TD_, G_)                                                                          \
  if (v.mirror) { KIFMM_TILE_GM(SD_, TD_, G_, true) }                                                       \
  else { KIFMM_TILE_GM(SD_, TD_, G_, false) }
#define KIFMM_TILE(SD_, TD_)                                                                                \
  if (v.sd == SD_ && v.td == TD_) {                                                                         \
    if (G == 1) { KIFMM_TILE_G(SD_, TD_, 1) }                                                               \
    else { KIFMM_TILE_G(SD_, TD_, 2) }                                                                      \
    return hipGetLastError();                                                                               \
  }
    KIFMM_TILE(4, 3)
    KIFMM_TILE(3, 3)
    KIFMM_TILE(3, 4)
#undef KIFMM_TILE
#undef KIFMM_TILE_G
#undef KIFMM_TILE_GM
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Mixed path: FP64 matrix cores (v_mfma_f64_16x16x4f64, FMA, no order
// contract).  160 x 64 tile per workgroup of 8 waves (M = 152 or 296 needs 1 or
// 2 row tiles, 5% or 8% padding); wave w owns columns [16 (w % 4), +16) and the
// five 16-row blocks of row half w / 4.  K is staged through LDS in slabs of 16
// with the next slab prefetched into registers while the current one feeds 20
// MFMAs per wave.  108 VGPRs and 33 KB of LDS: 4 waves per SIMD (the same tile
// on 4 waves of ten row blocks each needed 168 VGPRs, 3 waves per SIMD, and was
// 13% slower on the down pass).
// With ksplit > 1, blockIdx.z = matrix * ksplit + K chunk: each chunk writes
// its partial product to g.work, and k_splitk_reduce sums the chunks in chunk
// order into C (deterministic), so a level with few columns still spreads over
// enough workgroups.  A split launch without a workspace is rejected.  With a
// two-part B (k1 > 0: the downward step [P_dn | P_dn L2L_z] [DC; DE_parent])
// the first k1 rows of a column come from B1 at the output node's row.  Operand
// layout of the 16x16x4 instruction: lane l supplies A[l%16][l/16] and
// B[l/16][l%16], and its result register r holds D[4r + l/16][l%16] (rows
// interleaved across the four lane groups).
typedef double kd4 __attribute__((ext_vector_type(4)));
constexpr int GM_M = 160, GM_N = 64, GM_K = 16, GM_T = 512, GM_MB = GM_M / 32;
constexpr int GM_RA = GM_M * GM_K / GM_T, GM_RB = GM_K * GM_N / GM_T;

__global__ __launch_bounds__(GM_T, 2) void k_gemm_mfma(GemmArgs g) {
  __shared__ double As[2][GM_K][GM_M + 16];  // two slabs: one multiplied, the next stored
  __shared__ double Bs[2][GM_K][GM_N + 16];
  const bool split = g.ksplit > 1;
  const int zk = split ? int(blockIdx.z) % g.ksplit : 0;
  const int z = split ? int(blockIdx.z) / g.ksplit : int(blockIdx.z);
  const int cbeg = g.zoff ? g.zoff[z] : 0, cend = g.zoff ? g.zoff[z + 1] : g.ncol;
  const int n0 = cbeg + int(blockIdx.y) * GM_N;
  if (n0 >= cend) return;
  const double* A = g.A + int64_t(z) * g.a_zstride;
  const int kbeg = zk * g.kchunk;
  const int kend = split ? min(g.K, kbeg + g.kchunk) : g.K;
  if (kbeg >= kend) return;
  const int m0 = blockIdx.x * GM_M;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
  const int wc = wave & 3, wr = wave >> 2;  // column block, row half
  kd4 acc[GM_MB];
#pragma unroll
  for (int mi = 0; mi < GM_MB; mi++) acc[mi] = kd4{0.0, 0.0, 0.0, 0.0};
  // the next slab is loaded into registers while this one is multiplied: the
  // loads are unconditional (an element outside the matrix reads element 0),
  // and the bounds mask and the K scale are applied when the slab is stored,
  // so no wait for them sits between the loads and the products
  double ra[GM_RA], rb[GM_RB], rs[GM_RB];
  // the B column this thread stages (8 threads a column); with a two-part B its
  // first k1 rows come from B1 at the output column's row
  const int jb = n0 + (tid >> 3);
  const bool colok = jb < cend;
  const double* bcol = colok ? g.B + int64_t(g.bidx[jb]) * g.ldb - g.k1 : g.B;
  const double* bcol1 = colok && g.k1 > 0 ? g.B1 + int64_t(g.cidx[jb]) * g.ldb1 : g.B;
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < GM_RA; u++) {
      const int e = tid + GM_T * u;
      const int m = m0 + e % GM_M, k = k0 + e / GM_M;
      ra[u] = A[m < g.M && k < kend ? int64_t(k) * g.M + m : 0];
    }
    const int kq0 = k0 + GM_RB * (tid & 7);
#pragma unroll
    for (int u = 0; u < GM_RB; u++) {
      const int kq = kq0 + u;
      const bool p1 = kq < g.k1;
      rb[u] = colok && kq < kend ? (p1 ? bcol1 : bcol)[kq] : g.B[0];
      rs[u] = p1 ? 1.0 : g.kscale[p1 ? 0 : (kq - g.k1) % g.ksd];
    }
  };
  auto store = [&](int buf, int k0) {
#pragma unroll
    for (int u = 0; u < GM_RA; u++) {
      const int e = tid + GM_T * u;
      const int m = m0 + e % GM_M, k = k0 + e / GM_M;
      As[buf][e / GM_M][e % GM_M] = m < g.M && k < kend ? ra[u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < GM_RB; u++) {
      const int kq = k0 + GM_RB * (tid & 7) + u;
      Bs[buf][GM_RB * (tid & 7) + u][tid >> 3] = colok && kq < kend ? rb[u] * rs[u] : 0.0;
    }
  };
  fetch(kbeg);
  store(0, kbeg);
  __syncthreads();
  // one barrier a slab: the store into the other buffer follows this slab's
  // products, and every wave has finished reading that buffer (two slabs back)
  // before the previous barrier
  int buf = 0;
  for (int k0 = kbeg; k0 < kend; k0 += GM_K) {
    const bool more = k0 + GM_K < kend;
    if (more) fetch(k0 + GM_K);
#pragma unroll
    for (int ks = 0; ks < GM_K / 4; ks++) {
      const double b = Bs[buf][4 * ks + lk][16 * wc + li];
#pragma unroll
      for (int mi = 0; mi < GM_MB; mi++)
        acc[mi] = __builtin_amdgcn_mfma_f64_16x16x4f64(As[buf][4 * ks + lk][80 * wr + 16 * mi + li], b, acc[mi], 0, 0,
                                                       0);
    }
    if (more) store(buf ^ 1, k0 + GM_K);
    __syncthreads();
    buf ^= 1;
  }
  const int n = n0 + 16 * wc + li;
  if (n >= cend) return;
  double* col = g.C + int64_t(g.cidx[n]) * g.ldc;
#pragma unroll
  for (int mi = 0; mi < GM_MB; mi++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int m = m0 + 80 * wr + 16 * mi + 4 * r + lk;
      if (m >= g.M) continue;
      const double v = g.rscale[m % g.rsd] * acc[mi][r];
      if (split) g.work[(int64_t(zk) * g.ncol + n) * g.M + m] = v;
      else col[m] = g.accumulate ? col[m] + v : v;
    }
}

// Sum of the split-K partial products in chunk order (one thread per output).
__global__ void k_splitk_reduce(const double* __restrict__ work, int ks, int ncol, int M, const int* __restrict__ cidx,
                                double* __restrict__ C, int64_t ldc, int accumulate) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= int64_t(ncol) * M) return;
  const int n = int(i / M), m = int(i % M);
  // chunk order; the loads of eight chunks are issued before their adds
  double v = 0.0;
  int z = 0;
  for (; z + 8 <= ks; z += 8) {
    double t[8];
#pragma unroll
    for (int u = 0; u < 8; u++) t[u] = work[(int64_t(z + u) * ncol + n) * M + m];
#pragma unroll
    for (int u = 0; u < 8; u++) v += t[u];
  }
  for (; z < ks; z++) v += work[(int64_t(z) * ncol + n) * M + m];
  double* dst = C + int64_t(cidx[n]) * ldc + m;
  *dst = accumulate ? *dst + v : v;
}

static hipError_t launch_gemm_mfma(const GemmArgs& g, hipStream_t s) {
  if (g.ncol <= 0) return hipSuccess;
  const bool split = g.ksplit > 1;
  if (split && (g.kchunk <= 0 || int64_t(g.ksplit) * g.kchunk < g.K || !g.work)) return hipErrorInvalidValue;
  const int nz = (g.zoff ? g.nz : 1) * (split ? g.ksplit : 1);
  dim3 grid(unsigned((g.M + GM_M - 1) / GM_M), unsigned((g.max_zcol + GM_N - 1) / GM_N), unsigned(nz));
  k_gemm_mfma<<<grid, dim3(GM_T), 0, s>>>(g);
  if (split) {
    const int64_t n = int64_t(g.ncol) * g.M;
    k_splitk_reduce<<<dim3(unsigned((n + 255) / 256)), dim3(256), 0, s>>>(g.work, g.ksplit, g.ncol, g.M, g.cidx, g.C,
                                                                           g.ldc, g.accumulate);
  }
  return hipGetLastError();
}

hipError_t launch_gemm_fma(const GemmArgs& g, hipStream_t s) { return launch_gemm_mfma(g, s); }

}  // namespace kifmm
