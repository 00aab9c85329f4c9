// State movement around the step:
//   * the corrector network's input pack, 101 channels in NCHW per member
//     (src/neuralpom/hybrid/state.py:181-226, constants.py:2-7), and its
//     adjoint;
//   * the volume-weighted conservative 4 x 4 restriction from the fine
//     reference grid to the coarse forecast grid (paper/paper_zh.md:625);
//   * a batched device copy for state snapshots (the capture / restore of
//     src/neuralpom/v1/checkpoint.py:51-95, kept on the device).
// All three are HBM-bound data movement: one pass over their arrays.
#include <algorithm>

#include "../../include/neuralpom_amd.h"
#include "pom_batch.h"

using namespace pom;

namespace {

constexpr int PACK3 = 6, PACK2 = 5, PACKC = PACK3 * KB + PACK2;   // 101 channels
constexpr int NPACK = PACK3 + PACK2;

template <class T> struct PackArgs {
  const T* f[NPACK];       // u ub v vb t s (3-D), el ua uab va vab (2-D)
  const T* mean[NPACK];    // per-cell statistics of one member's shape, or null
  const T* scale[NPACK];
};

// member m's channels: PACKC planes of im x jm starting at PACKC * off2[m]
template <class T>
__global__ void __launch_bounds__(TPB) k_pack(Batch B, PackArgs<T> a, T* __restrict__ out) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t n2 = (int64_t)p.im * p.jm;
  const int64_t cell = (int64_t)p.i * p.jm + p.j;   // member-local column
  T* o = out + (int64_t)PACKC * p.o2 + cell;
#pragma unroll
  for (int f = 0; f < PACK3; ++f) {
    const T* src = a.f[f] + p.o3 + cell * KB;
    T v[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) v[k] = src[k];
    if (a.mean[f]) {
#pragma unroll
      for (int k = 0; k < KB; ++k) v[k] = (v[k] - a.mean[f][cell * KB + k]) / a.scale[f][cell * KB + k];
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) o[(f * KB + k) * n2] = v[k];
  }
#pragma unroll
  for (int f = 0; f < PACK2; ++f) {
    const int g = PACK3 + f;
    T v = a.f[g][p.o2 + cell];
    if (a.mean[g]) v = (v - a.mean[g][cell]) / a.scale[g][cell];
    o[(PACK3 * KB + f) * n2] = v;
  }
}

template <class T> struct PackAdjArgs {
  const T* scale[NPACK];
  T* grad[NPACK];          // accumulated; null entries skipped
};

template <class T>
__global__ void __launch_bounds__(TPB) k_pack_adj(Batch B, PackAdjArgs<T> a, const T* __restrict__ cot) {
  Pos p;
  if (!pos2(B, p)) return;
  const int64_t n2 = (int64_t)p.im * p.jm;
  const int64_t cell = (int64_t)p.i * p.jm + p.j;
  const T* c = cot + (int64_t)PACKC * p.o2 + cell;
#pragma unroll
  for (int f = 0; f < PACK3; ++f) {
    if (!a.grad[f]) continue;
    T g[KB], x[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) g[k] = c[(f * KB + k) * n2];
    if (a.scale[f]) {
#pragma unroll
      for (int k = 0; k < KB; ++k) g[k] = g[k] / a.scale[f][cell * KB + k];
    }
    T* dst = a.grad[f] + p.o3 + cell * KB;
#pragma unroll
    for (int k = 0; k < KB; ++k) x[k] = dst[k];
#pragma unroll
    for (int k = 0; k < KB; ++k) dst[k] = x[k] + g[k];
  }
#pragma unroll
  for (int f = 0; f < PACK2; ++f) {
    const int g = PACK3 + f;
    if (!a.grad[g]) continue;
    T v = c[(PACK3 * KB + f) * n2];
    if (a.scale[g]) v = v / a.scale[g][cell];
    a.grad[g][p.o2 + cell] += v;
  }
}

// ---------------------------------------------------------------- restriction
// Coarse interior cell (I, J) (1 <= I <= im_c - 2) averages the fine interior
// cells i = 4 (I - 1) + 1 .. 4 I, j likewise, with weights
//   area:   art * mask               (2-D fields such as el)
//   volume: art * mask * (h + el) [* dz(k) for 3-D fields]
// value = sum w f / sum w (0 where the block holds no weight); the coarse
// edge ring is 0 (land ring of the closed basin).  Sum_c W_c f_c = sum w f
// with W_c = sum of the block's fine weights, per level and member.
template <class T> struct RestrictArgs {
  const T* f;
  const T *art, *mask, *h, *el;   // fine 2-D fields; el may be null (D = h)
  const T* vg;                     // fine vertical grids [M][4][KB]
  T* out;
  int three_d, volume;
};

template <class T>
__global__ void __launch_bounds__(TPB) k_restrict4(Batch fine, Batch coarse, RestrictArgs<T> a) {
  Pos q;
  if (a.three_d ? !pos3(coarse, q) : !pos2(coarse, q)) return;
  const int64_t co = a.three_d ? ix3(q, q.i, q.j, q.k) : ix2(q, q.i, q.j);
  if (q.i < 1 || q.i > q.im - 2 || q.j < 1 || q.j > q.jm - 2) {
    a.out[co] = T(0);
    return;
  }
  const Member fm = fine.mem[q.m];
  const int i0 = 4 * (q.i - 1) + 1, j0 = 4 * (q.j - 1) + 1;
  const T dz = a.three_d && a.volume ? a.vg[(q.m * 4 + 2) * KB + q.k] : T(1);
  T sw = T(0), sf = T(0);
#pragma unroll
  for (int di = 0; di < 4; ++di)
#pragma unroll
    for (int dj = 0; dj < 4; ++dj) {
      const int64_t c2 = fm.off2 + (int64_t)(i0 + di) * fm.jm + (j0 + dj);
      T w = a.art[c2] * a.mask[c2];
      if (a.volume) w = w * ((a.h[c2] + (a.el ? a.el[c2] : T(0))) * dz);
      const T v = a.three_d ? a.f[fm.off3 + ((int64_t)(i0 + di) * fm.jm + (j0 + dj)) * KB + q.k] : a.f[c2];
      sw += w;
      sf += w * v;
    }
  a.out[co] = sw > T(0) ? sf / sw : T(0);
}

// Adjoint of the restriction: one thread per fine water column.  A fine cell
// of block (I, J) receives w / W * cbar of its coarse cell (per level for 3-D
// fields); with volume weights and el given, el also receives
// sum_k cbar (f - c) / W * dw/del, dw/del = art * mask [* dz], through the
// weights of the layer means.
template <class T> struct RestrictAdjArgs {
  RestrictArgs<T> fw;
  const T* cbar;   // coarse cotangent
  T* gf;           // fine field cotangent (accumulated)
  T* gel;          // fine el cotangent (accumulated), may be null
};

template <class T>
__global__ void __launch_bounds__(TPB) k_restrict4_adj(Batch fine, Batch coarse, RestrictAdjArgs<T> r) {
  Pos p;
  if (!pos2(fine, p)) return;
  if (p.i < 1 || p.i > p.im - 2 || p.j < 1 || p.j > p.jm - 2) return;   // the fine edge ring is not sampled
  const RestrictArgs<T>& a = r.fw;
  const Member cm = coarse.mem[p.m];
  const int I = (p.i - 1) / 4 + 1, J = (p.j - 1) / 4 + 1;
  const int i0 = 4 * (I - 1) + 1, j0 = 4 * (J - 1) + 1;
  const int64_t my2 = ix2(p, p.i, p.j);
  const T base_w = a.art[my2] * a.mask[my2];
  const T dmy = a.volume ? a.h[my2] + (a.el ? a.el[my2] : T(0)) : T(1);
  const int nk = a.three_d ? KB : 1;
  T gel = T(0);
  for (int k = 0; k < nk; ++k) {
    const T dz = a.three_d && a.volume ? a.vg[(p.m * 4 + 2) * KB + k] : T(1);
    T sw = T(0), sf = T(0);
#pragma unroll
    for (int di = 0; di < 4; ++di)
#pragma unroll
      for (int dj = 0; dj < 4; ++dj) {
        const int64_t c2 = p.o2 + (int64_t)(i0 + di) * p.jm + (j0 + dj);
        T w = a.art[c2] * a.mask[c2];
        if (a.volume) w = w * ((a.h[c2] + (a.el ? a.el[c2] : T(0))) * dz);
        const T v = a.three_d ? a.f[p.o3 + ((int64_t)(i0 + di) * p.jm + (j0 + dj)) * KB + k] : a.f[c2];
        sw += w;
        sf += w * v;
      }
    if (!(sw > T(0))) continue;
    const int64_t co = a.three_d ? cm.off3 + ((int64_t)I * cm.jm + J) * KB + k : cm.off2 + (int64_t)I * cm.jm + J;
    const T cb = r.cbar[co];
    const T w = a.volume ? base_w * dmy * dz : base_w;
    const int64_t fo = a.three_d ? ix3(p, p.i, p.j, k) : my2;
    r.gf[fo] += w / sw * cb;
    if (r.gel && a.volume && a.el) {
      const T c = sf / sw;
      const T v = a.f[fo];
      gel += cb * (v - c) / sw * base_w * dz;
    }
  }
  if (r.gel && a.volume && a.el) r.gel[my2] += gel;
}

// ---------------------------------------------------------------- batched copy
constexpr int MAXCOPY = 64;
struct CopyTab {
  uint32_t* dst[MAXCOPY];
  const uint32_t* src[MAXCOPY];
  int64_t words[MAXCOPY];
};

__global__ void __launch_bounds__(TPB) k_copy_arrays(CopyTab t) {
  const int e = blockIdx.y;
  uint32_t* d = t.dst[e];
  const uint32_t* s = t.src[e];
  const int64_t n = t.words[e];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = s[i];
}

template <class T>
int run_pack(const pom_batch* b, const void* const* fields, const void* const* mean, const void* const* scale,
             void* out, hipStream_t st) {
  PackArgs<T> a{};
  for (int f = 0; f < NPACK; ++f) {
    if (!fields[f]) return POM_E_ARG;
    a.f[f] = static_cast<const T*>(fields[f]);
    a.mean[f] = mean ? static_cast<const T*>(mean[f]) : nullptr;
    a.scale[f] = scale ? static_cast<const T*>(scale[f]) : nullptr;
    if ((a.mean[f] == nullptr) != (a.scale[f] == nullptr)) return POM_E_ARG;
  }
  const Batch B{b->d_mem, b->d_blk3, b->d_blk2, b->M};
  if (b->nblk2 > 0)
    hipLaunchKernelGGL(k_pack<T>, dim3(b->nblk2), dim3(TPB), 0, st, B, a, static_cast<T*>(out));
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

template <class T>
int run_pack_adj(const pom_batch* b, const void* const* scale, const void* cot, void* const* grad, hipStream_t st) {
  PackAdjArgs<T> a{};
  for (int f = 0; f < NPACK; ++f) {
    a.scale[f] = scale ? static_cast<const T*>(scale[f]) : nullptr;
    a.grad[f] = static_cast<T*>(grad[f]);
  }
  const Batch B{b->d_mem, b->d_blk3, b->d_blk2, b->M};
  if (b->nblk2 > 0)
    hipLaunchKernelGGL(k_pack_adj<T>, dim3(b->nblk2), dim3(TPB), 0, st, B, a, static_cast<const T*>(cot));
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

// every member's statistics share one shape: they apply to uniform batches only
bool uniform(const pom_batch* b) {
  for (const Member& m : b->h_mem)
    if (m.im != b->h_mem[0].im || m.jm != b->h_mem[0].jm) return false;
  return true;
}

}  // namespace

extern "C" {

int pom_pack_input(const pom_batch* b, const void* const* fields, const void* const* mean,
                   const void* const* scale, void* out, void* stream) {
  if (!b || !fields || !out) return POM_E_ARG;
  if ((mean || scale) && !uniform(b)) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_pack<float>(b, fields, mean, scale, out, st);
  return run_pack<double>(b, fields, mean, scale, out, st);
}

int pom_pack_input_adj(const pom_batch* b, const void* const* scale, const void* cot_out, void* const* grad_fields,
                       void* stream) {
  if (!b || !cot_out || !grad_fields) return POM_E_ARG;
  if (scale && !uniform(b)) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (b->dtype == POM_F32) return run_pack_adj<float>(b, scale, cot_out, grad_fields, st);
  return run_pack_adj<double>(b, scale, cot_out, grad_fields, st);
}

int pom_restrict4(const pom_batch* fine, const pom_batch* coarse, int32_t three_d, int32_t volume_weighted,
                  const void* field, const void* art, const void* mask, const void* h, const void* el,
                  const void* fine_vgrid, void* out, void* stream) {
  if (!fine || !coarse || !field || !art || !mask || !out) return POM_E_ARG;
  if (volume_weighted && !h) return POM_E_ARG;
  if (three_d && volume_weighted && !fine_vgrid) return POM_E_ARG;
  if (fine->M != coarse->M || fine->dtype != coarse->dtype) return POM_E_ARG;
  for (int m = 0; m < fine->M; ++m) {
    const Member& f = fine->h_mem[m];
    const Member& c = coarse->h_mem[m];
    if ((f.im - 2) % 4 || (f.jm - 2) % 4 || c.im != (f.im - 2) / 4 + 2 || c.jm != (f.jm - 2) / 4 + 2)
      return POM_E_ARG;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Batch F{fine->d_mem, fine->d_blk3, fine->d_blk2, fine->M};
  const Batch C{coarse->d_mem, coarse->d_blk3, coarse->d_blk2, coarse->M};
  const int nb = three_d ? coarse->nblk3 : coarse->nblk2;
  if (nb == 0) return POM_OK;
  auto go = [&](auto tag) {
    using T = decltype(tag);
    RestrictArgs<T> a{static_cast<const T*>(field), static_cast<const T*>(art), static_cast<const T*>(mask),
                      static_cast<const T*>(h), static_cast<const T*>(el), static_cast<const T*>(fine_vgrid),
                      static_cast<T*>(out), three_d != 0, volume_weighted != 0};
    hipLaunchKernelGGL(k_restrict4<T>, dim3(nb), dim3(TPB), 0, st, F, C, a);
  };
  if (fine->dtype == POM_F32) go(float{});
  else go(double{});
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

int pom_restrict4_adj(const pom_batch* fine, const pom_batch* coarse, int32_t three_d, int32_t volume_weighted,
                      const void* field, const void* art, const void* mask, const void* h, const void* el,
                      const void* fine_vgrid, const void* cot_out, void* grad_field, void* grad_el, void* stream) {
  if (!fine || !coarse || !field || !art || !mask || !cot_out || !grad_field) return POM_E_ARG;
  if (volume_weighted && !h) return POM_E_ARG;
  if (three_d && volume_weighted && !fine_vgrid) return POM_E_ARG;
  if (fine->M != coarse->M || fine->dtype != coarse->dtype) return POM_E_ARG;
  for (int m = 0; m < fine->M; ++m) {
    const Member& f = fine->h_mem[m];
    const Member& c = coarse->h_mem[m];
    if ((f.im - 2) % 4 || (f.jm - 2) % 4 || c.im != (f.im - 2) / 4 + 2 || c.jm != (f.jm - 2) / 4 + 2)
      return POM_E_ARG;
  }
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Batch F{fine->d_mem, fine->d_blk3, fine->d_blk2, fine->M};
  const Batch C{coarse->d_mem, coarse->d_blk3, coarse->d_blk2, coarse->M};
  if (fine->nblk2 == 0) return POM_OK;
  auto go = [&](auto tag) {
    using T = decltype(tag);
    RestrictAdjArgs<T> r{};
    r.fw = RestrictArgs<T>{static_cast<const T*>(field), static_cast<const T*>(art), static_cast<const T*>(mask),
                           static_cast<const T*>(h), static_cast<const T*>(el), static_cast<const T*>(fine_vgrid),
                           nullptr, three_d != 0, volume_weighted != 0};
    r.cbar = static_cast<const T*>(cot_out);
    r.gf = static_cast<T*>(grad_field);
    r.gel = static_cast<T*>(grad_el);
    hipLaunchKernelGGL(k_restrict4_adj<T>, dim3(fine->nblk2), dim3(TPB), 0, st, F, C, r);
  };
  if (fine->dtype == POM_F32) go(float{});
  else go(double{});
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

int pom_copy_arrays(int32_t n, void* const* dst, const void* const* src, const int64_t* bytes, void* stream) {
  if (n < 0 || (n > 0 && (!dst || !src || !bytes))) return POM_E_ARG;
  hipStream_t st = static_cast<hipStream_t>(stream);
  for (int32_t s0 = 0; s0 < n; s0 += MAXCOPY) {
    CopyTab t{};
    const int cnt = std::min<int32_t>(MAXCOPY, n - s0);
    int64_t most = 0;
    for (int e = 0; e < cnt; ++e) {
      const int64_t nbytes = bytes[s0 + e];
      if (nbytes < 0 || nbytes % 4 || !dst[s0 + e] || !src[s0 + e] ||
          reinterpret_cast<uintptr_t>(dst[s0 + e]) % 4 || reinterpret_cast<uintptr_t>(src[s0 + e]) % 4)
        return POM_E_ARG;
      t.dst[e] = static_cast<uint32_t*>(dst[s0 + e]);
      t.src[e] = static_cast<const uint32_t*>(src[s0 + e]);
      t.words[e] = nbytes / 4;
      most = std::max(most, t.words[e]);
    }
    if (most == 0) continue;
    const int64_t bl = (most + 4 * TPB - 1) / (4 * TPB);
    const int blocks = (int)std::min<int64_t>(std::max<int64_t>(bl, 1), 1024);
    hipLaunchKernelGGL(k_copy_arrays, dim3(blocks, cnt), dim3(TPB), 0, st, t);
  }
  return hipGetLastError() == hipSuccess ? POM_OK : POM_E_HIP;
}

}  // extern "C"
