// ibm_kernels.hip -- marker pack/scatter/repulsion, immersed-boundary
// interpolation and spreading lists, and the concentration step (gfx950).
//
// Built with -ffp-contract=off.  Every sum runs in the reference's loop
// order (stencil x outer, z inner; markers in increasing index), so with the
// same inputs these kernels reproduce the FP32 restatement bit for bit.
//
// Design (vs. the reference's one-thread-per-marker loops with atomicAdd,
// fluid_kernels.cu:244-459):
//   * markers only move between list rebuilds, so each marker's 64 stencil
//     cells (4x4x4 Peskin support, wall-clipped or wrapped), its raw weights
//     phi_x phi_y phi_z, their sum and the spreading weights are tabulated
//     once per rebuild;
//   * interpolation: one 64-lane wavefront per marker, one lane per stencil
//     cell; the 64 products are summed in stencil order with lane reads, which
//     keeps the reference's summation order without LDS round trips;
//   * spreading: gathered, not scattered.  Per-cell lists of (marker,
//     weight), sorted by marker, are rebuilt with the tables (count, block
//     scan, fill, sort).  The fused LBM kernel and the scalar kernel sum a
//     cell's contributions in marker order -- no atomics, the same bits every
//     run.
#include <algorithm>
#include <atomic>
#include <cstddef>

#include "glsm_internal.hpp"

namespace glsm {
namespace {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / 64;
constexpr int kFlagCells = 256;  // cells per spreading block flag

inline int blocks_for(long n, int per) { return (int)((n + per - 1) / per); }

__device__ inline float lane_read(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

// Peskin 4-point kernel, fluid_kernels.cu:60-71
__device__ inline float delta4(float r) {
  r = fabsf(r);
  if (r <= 1.f) return 0.125f * (3.f - 2.f * r + sqrtf(1.f + 4.f * r - 4.f * r * r));
  if (r <= 2.f) return 0.125f * (5.f - 2.f * r - sqrtf(-7.f + 12.f * r - 4.f * r * r));
  return 0.f;
}

// One lane's stencil cell of marker position p: cell index (or -1 when the
// cell is clipped by a wall) and the weight phi_x*phi_y*phi_z.
struct StencilLane {
  int cell;
  float w;
};
__device__ inline int axis_cell(int i, int n, int per, bool& ok) {
  if (per) {
    ok = true;
    return ((i % n) + n) % n;
  }
  ok = (i >= 0 && i < n);
  return i;
}
__device__ inline StencilLane stencil_lane(const LatticeGeom& g, float3 p, int lane) {
  const float gx = p.x / g.h, gy = p.y / g.h, gz = p.z / g.h;
  const int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
  const int ii = ix - 1 + (lane >> 4), jj = iy - 1 + ((lane >> 2) & 3), kk = iz - 1 + (lane & 3);
  bool okx, oky, okz;
  const int cx = axis_cell(ii, g.Nx, g.perX, okx);
  const int cy = axis_cell(jj, g.Ny, g.perY, oky);
  const int cz = axis_cell(kk, g.Nz, g.perZ, okz);
  StencilLane s;
  if (okx && oky && okz) {
    const float phix = delta4(gx - (float)ii), phiy = delta4(gy - (float)jj), phiz = delta4(gz - (float)kk);
    s.w = phix * phiy * phiz;
    s.cell = cx + cy * g.Nx + cz * g.Nx * g.Ny;
  } else {
    s.w = 0.f;
    s.cell = -1;
  }
  return s;
}

// exclusive scan of one int per thread over the 256-thread block:
// wavefront shuffles, then one LDS exchange of the four wave totals
__device__ inline int block_excl_scan(int v, int* s_wave, int& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_wave[w] = x;
  __syncthreads();
  int pre = 0;
  total = 0;
#pragma unroll
  for (int k = 0; k < kWavesPerBlock; ++k) {
    const int t = s_wave[k];
    if (k < w) pre += t;
    total += t;
  }
  __syncthreads();
  return pre + x - v;
}

// ---------------------------------------------------------------- coupler
// k_gather_boundary, coupling_kernels.cu:42-59, all gels in one launch
__device__ __forceinline__ void pack_at(const MarkerArrays& m, float g2l, int l) {
  const GelMarkerDesc d = m.desc[m.owner[l]];
  const int id = m.bIndex[l];
  const double3 r = d.rn[id], v = d.Vels[id];
  m.lag[l] = make_float3((float)r.x, (float)r.y, (float)r.z);
  const double s = (double)g2l;
  m.Vl[l] = make_float3((float)(v.x * s), (float)(v.y * s), (float)(v.z * s));
  m.Cl[l] = (float)d.un_norm[id];
}

// k_add_reaction_to_gel + k_add_drag_to_gel, coupling_kernels.cu:8-34;
// every boundary node owns exactly one marker, so no two threads collide
__device__ __forceinline__ void scatter_at(const MarkerArrays& m, int l) {
  const GelMarkerDesc d = m.desc[m.owner[l]];
  const int id = m.bIndex[l];
  const float3 fr = m.Frep[l], fd = m.Fdrag[l];
  const float s = m.Sl[l];
  double3 a = d.Fn_robin[id];
  a.x -= (double)fr.x; a.y -= (double)fr.y; a.z -= (double)fr.z;
  d.Fn_robin[id] = a;
  d.un_robin[id] -= (double)s;
  double3 b = d.Fdrag_robin[id];
  b.x -= (double)fd.x; b.y -= (double)fd.y; b.z -= (double)fd.z;
  d.Fdrag_robin[id] = b;
}

// Lennard-Jones repulsion between markers of different gels
// (k_gel_repulsion / bodyBodyInteraction, coupling_kernels.cu:61-146).  One
// wavefront per marker a; its 64 lanes evaluate 64 partners at a time, and
// the partners inside the cut-off (a ballot mask) are added in increasing
// index -- the reference's accumulation order -- by every lane alike.  The
// partners' owners and positions are loaded kRepAhead chunks of 64 at a
// time, so a wave waits for one round of loads per kRepAhead chunks (8
// measured best of 1, 4, 8 and 16: 126.6, 72.0, 69.4 and 81.0 us at
// 10,098 markers).
constexpr int kRepAhead = 8;
// partners [b0, b1) of marker a (position pa, owner oa), added to f in
// increasing index
__device__ __forceinline__ void repulsion_range(const MarkerArrays& m, float delta, float gamma, float3 pa, int oa,
                                                int b0, int b1, int lane, float& fx, float& fy, float& fz) {
  const float rcut = 2.5f * gamma, rc2 = rcut * rcut, sig2 = gamma * gamma;
  for (int base = b0; base < b1; base += 64 * kRepAhead) {
    int ob[kRepAhead];
    float3 pb[kRepAhead];
#pragma unroll
    for (int j = 0; j < kRepAhead; ++j) {
      const int b = min(base + 64 * j + lane, b1 - 1);  // (a partner past the end is not counted)
      ob[j] = m.owner[b];
      pb[j] = m.lag[b];
    }
#pragma unroll
    for (int j = 0; j < kRepAhead; ++j) {
      const int b = base + 64 * j + lane;
      float cx = 0.f, cy = 0.f, cz = 0.f;
      bool hit = false;
      if (b < b1 && ob[j] != oa) {
        const float dx = pa.x - pb[j].x, dy = pa.y - pb[j].y, dz = pa.z - pb[j].z;
        const float d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > 0.0f && d2 < rc2) {
          const float inv = 1.0f / d2;
          const float sr2 = sig2 * inv;
          const float sr6 = sr2 * sr2 * sr2;
          const float sr12 = sr6 * sr6;
          const float fac = 24.0f * delta * (2.0f * sr12 - sr6) * inv;
          cx = fac * dx;
          cy = fac * dy;
          cz = fac * dz;
          hit = true;
        }
      }
      unsigned long long mask = __ballot(hit);
      while (mask) {
        const int k = __ffsll(mask) - 1;
        fx += lane_read(cx, k);
        fy += lane_read(cy, k);
        fz += lane_read(cz, k);
        mask &= mask - 1;
      }
    }
  }
}
__device__ __forceinline__ void repulsion_wave(const MarkerArrays& m, float delta, float gamma, int a, int lane) {
  float fx = 0.f, fy = 0.f, fz = 0.f;
  repulsion_range(m, delta, gamma, m.lag[a], m.owner[a], 0, m.M, lane, fx, fy, fz);
  if (lane == 0) m.Frep[a] = make_float3(0.f + fx, 0.f + fy, 0.f + fz);
}
// The same sum with the markers of each gel in turn (a gel's markers are
// consecutive, so gel by gel is still increasing index), skipping a gel
// whose markers' bounding box lies beyond the cut-off from marker a: none
// of them would count (the 0.1% margin covers the rounding of both
// distances).  gel_off[g] = first marker of gel g, gel_off[ngel] = M; box =
// per gel (min x, y, z, max x, y, z).
__device__ __forceinline__ void repulsion_wave_culled(const MarkerArrays& m, int ngel, const int* __restrict__ gel_off,
                                                      const float* __restrict__ box, float delta, float gamma, int a,
                                                      int lane) {
  const float rcut = 2.5f * gamma, rc2 = rcut * rcut;
  const float3 pa = m.lag[a];
  const int oa = m.owner[a];
  float fx = 0.f, fy = 0.f, fz = 0.f;
  for (int g = 0; g < ngel; ++g) {
    if (g == oa) continue;
    const float* bx = box + 6 * g;
    const float ex = fmaxf(fmaxf(bx[0] - pa.x, pa.x - bx[3]), 0.f);
    const float ey = fmaxf(fmaxf(bx[1] - pa.y, pa.y - bx[4]), 0.f);
    const float ez = fmaxf(fmaxf(bx[2] - pa.z, pa.z - bx[5]), 0.f);
    if (ex * ex + ey * ey + ez * ez > rc2 * 1.001f) continue;
    repulsion_range(m, delta, gamma, pa, oa, gel_off[g], gel_off[g + 1], lane, fx, fy, fz);
  }
  if (lane == 0) m.Frep[a] = make_float3(0.f + fx, 0.f + fy, 0.f + fz);
}

// per gel, the bounding box of its markers (one work-group per gel)
__device__ __forceinline__ void gel_marker_box(const MarkerArrays& m, const int* __restrict__ gel_off,
                                               float* __restrict__ box, int g) {
  __shared__ float s[kWavesPerBlock][6];
  float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = gel_off[g] + threadIdx.x; b < gel_off[g + 1]; b += kThreads) {
    const float3 p = m.lag[b];
    v[0] = fminf(v[0], p.x); v[1] = fminf(v[1], p.y); v[2] = fminf(v[2], p.z);
    v[3] = fmaxf(v[3], p.x); v[4] = fmaxf(v[4], p.y); v[5] = fmaxf(v[5], p.z);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const float w = __shfl_xor(v[k], o, 64);
      v[k] = k < 3 ? fminf(v[k], w) : fmaxf(v[k], w);
    }
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 6; ++k) s[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int k = threadIdx.x;
    float r = s[0][k];
    for (int w = 1; w < kWavesPerBlock; ++w) r = k < 3 ? fminf(r, s[w][k]) : fmaxf(r, s[w][k]);
    box[6 * g + k] = r;
  }
}

// ---------------------------------------------------------------- IBM
// Ul = sum(u*phi)/sum(phi) over the stencil, Fdrag = beta_eff*(Vl - Ul)
// (k_ibm_interpolate_velocity + k_scale_negbeta, fluid_kernels.cu:244-291)
// Stencil tables are stored lane-major: entry (marker l, stencil lane k) at
// k * M + l, so one thread per marker reads them coalesced.
__device__ __forceinline__ size_t tab(const MarkerArrays& m, int l, int k) { return (size_t)k * m.M + l; }

// Ul / Fdrag from the 64 stencil sums, the weight sum ws and the marker
// velocity vl (shared by both interpolation forms)
__device__ __forceinline__ void interp_finish(const MarkerArrays& m, float beta_eff, int l, float ws, float3 vl,
                                              float ux, float uy, float uz) {
  float3 ul = make_float3(0.f, 0.f, 0.f);
  if (ws > 0.0f) ul = make_float3(ux / ws, uy / ws, uz / ws);
  m.Ul[l] = ul;
  m.Fdrag[l] = make_float3(beta_eff * (vl.x - ul.x), beta_eff * (vl.y - ul.y), beta_eff * (vl.z - ul.z));
}

// One wavefront per marker, one lane per stencil cell, the 64 products
// summed in lane order (for small marker counts: more parallelism).
__device__ __forceinline__ void interp_wave(const MarkerArrays& m, const SpreadLists& sl,
                                            const float3* __restrict__ u, float beta_eff, int l, int lane) {
  // the marker's weight sum and velocity are loaded ahead of the stencil
  const float ws = gmem(sl.wsum)[l];
  const GLSM_GLOBAL float* vp = gmem(reinterpret_cast<const float*>(m.Vl)) + 3 * (size_t)l;
  const float3 vl = make_float3(vp[0], vp[1], vp[2]);
  const size_t k = tab(m, l, lane);
  const int cell = gmem(sl.ctab)[k];
  float px = 0.f, py = 0.f, pz = 0.f;
  if (cell >= 0) {
    const float w = gmem(sl.rtab)[k];
    const GLSM_GLOBAL float* up = gmem(reinterpret_cast<const float*>(u)) + 3 * (size_t)cell;
    px = up[0] * w;
    py = up[1] * w;
    pz = up[2] * w;
  }
  float ux = 0.f, uy = 0.f, uz = 0.f;
#pragma unroll
  for (int j = 0; j < 64; ++j) {
    ux += lane_read(px, j);
    uy += lane_read(py, j);
    uz += lane_read(pz, j);
  }
  if (lane == 0) interp_finish(m, beta_eff, l, ws, vl, ux, uy, uz);
}

// One thread per (marker, velocity component) summing the 64 stencil cells
// in order (for many markers: the serial sums run in all lanes at once).
// The loads are unconditional -- a cell outside the lattice reads cell 0
// with weight 0 -- so 32 iterations' loads are in flight together.  The
// same products in the same order as the wavefront form, except that cells
// outside the lattice add (+-0) instead of nothing: the running sum starts
// at +0 and is never -0, so adding a zero leaves it unchanged and the two
// forms agree bit for bit.
__device__ __forceinline__ float stencil_sum(const MarkerArrays& m, const SpreadLists& sl,
                                             const float* __restrict__ field, int stride, int comp, int l) {
  constexpr int kChunk = 16;  // table entries, then field values, loaded ahead per chunk
  const GLSM_GLOBAL int* ctab = gmem(sl.ctab);
  const GLSM_GLOBAL float* rtab = gmem(sl.rtab);
  const GLSM_GLOBAL float* gf = gmem(field);
  float acc = 0.f;
  for (int k0 = 0; k0 < 64; k0 += kChunk) {
    int cell[kChunk];
    float w[kChunk], f[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const size_t t = tab(m, l, k0 + j);
      cell[j] = max(ctab[t], 0);
      w[j] = rtab[t];
    }
#pragma unroll
    for (int j = 0; j < kChunk; ++j) f[j] = gf[(size_t)stride * cell[j] + comp];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) acc += f[j] * w[j];
  }
  return acc;
}

__device__ __forceinline__ void interp_component(const MarkerArrays& m, const SpreadLists& sl,
                                                 const float3* __restrict__ u, float beta_eff, int l, int comp) {
  const float ws = gmem(sl.wsum)[l];  // loaded ahead of the stencil sums
  const float vl = gmem(reinterpret_cast<const float*>(m.Vl))[3 * (size_t)l + comp];
  const float acc = stencil_sum(m, sl, reinterpret_cast<const float*>(u), 3, comp, l);
  const float ul = ws > 0.0f ? acc / ws : 0.f;
  gmem(reinterpret_cast<float*>(m.Ul))[3 * (size_t)l + comp] = ul;
  gmem(reinterpret_cast<float*>(m.Fdrag))[3 * (size_t)l + comp] = beta_eff * (vl - ul);
}

// Staged form: a work-group takes kStageMarkers markers.  Phase 1, one
// thread per (marker, kStageEntries-entry part of the stencil): the part's
// table entries, then its velocities, are loaded together and the products
// stored in LDS ([component][entry][marker], so both phases access
// consecutive words across a wavefront).  Phase 2, one thread per (marker,
// component): the 64 products added in stencil order from LDS.  Each marker
// then waits for two dependent rounds of loads instead of eight, and every
// table load is coalesced.  Same products, same order, same bits as the
// other forms (cells outside the lattice add +-0 as in the thread form).
// 8 parts of 8 entries (32 markers, 24 KB of LDS per work-group) measured
// fastest of 4, 8 and 16 parts from 436 to 29k markers.
#ifndef GLSM_STAGE_PARTS
#define GLSM_STAGE_PARTS 8
#endif
constexpr int kStageParts = GLSM_STAGE_PARTS;             // threads per marker in phase 1
constexpr int kStageMarkers = kThreads / kStageParts;     // markers per work-group
constexpr int kStageEntries = 64 / kStageParts;           // stencil entries per phase-1 thread
static_assert(kStageParts >= 3 && 64 % kStageParts == 0, "phase 2 needs a thread per component");
struct InterpStage {
  float p[3][64][kStageMarkers];
};
__device__ __forceinline__ void interp_stage_load(const SpreadLists& sl, int M, const float3* __restrict__ u, int l,
                                                  int part, int mi, InterpStage& st) {
  const GLSM_GLOBAL int* ctab = gmem(sl.ctab);
  const GLSM_GLOBAL float* rtab = gmem(sl.rtab);
  const GLSM_GLOBAL float* gu = gmem(reinterpret_cast<const float*>(u));
  const int k0 = kStageEntries * part;
  int cell[kStageEntries];
  float w[kStageEntries], fx[kStageEntries], fy[kStageEntries], fz[kStageEntries];
#pragma unroll
  for (int j = 0; j < kStageEntries; ++j) {
    const size_t t = (size_t)(k0 + j) * M + l;
    cell[j] = max(ctab[t], 0);
    w[j] = rtab[t];
  }
#pragma unroll
  for (int j = 0; j < kStageEntries; ++j) {
    fx[j] = gu[3 * (size_t)cell[j]];
    fy[j] = gu[3 * (size_t)cell[j] + 1];
    fz[j] = gu[3 * (size_t)cell[j] + 2];
  }
#pragma unroll
  for (int j = 0; j < kStageEntries; ++j) {
    st.p[0][k0 + j][mi] = fx[j] * w[j];
    st.p[1][k0 + j][mi] = fy[j] * w[j];
    st.p[2][k0 + j][mi] = fz[j] * w[j];
  }
}
__device__ __forceinline__ float interp_stage_sum(const InterpStage& st, int comp, int mi) {
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < 64; ++k) acc += st.p[comp][k][mi];
  return acc;
}
__device__ __forceinline__ void interp_stage_finish(const MarkerArrays& m, float beta_eff, int l, int comp, float ws,
                                                    float vl, float acc) {
  const float ul = ws > 0.0f ? acc / ws : 0.f;
  gmem(reinterpret_cast<float*>(m.Ul))[3 * (size_t)l + comp] = ul;
  gmem(reinterpret_cast<float*>(m.Fdrag))[3 * (size_t)l + comp] = beta_eff * (vl - ul);
}

// Dl = sum(c*phi)/sum(phi), Sl = Cl - Dl (k_ibm_sample_concentration +
// k_robin_boundary, fluid_kernels.cu:350-387, 293-298)
__device__ __forceinline__ void sample_finish(const MarkerArrays& m, const SpreadLists& sl, int l, float cl) {
  const float ws = sl.wsum[l];
  const float dl = ws > 0.0f ? cl / ws : 0.0f;
  m.Dl[l] = dl;
  m.Sl[l] = 1 * (m.Cl[l] - dl);
}

__device__ __forceinline__ void sample_wave(const MarkerArrays& m, const SpreadLists& sl, const float* __restrict__ c,
                                            int l, int lane) {
  const size_t k = tab(m, l, lane);
  const int cell = sl.ctab[k];
  const float p = cell >= 0 ? c[cell] * sl.rtab[k] : 0.f;
  float cl = 0.f;
#pragma unroll
  for (int j = 0; j < 64; ++j) cl += lane_read(p, j);
  if (lane == 0) sample_finish(m, sl, l, cl);
}

__device__ __forceinline__ void sample_thread(const MarkerArrays& m, const SpreadLists& sl,
                                              const float* __restrict__ c, int l) {
  sample_finish(m, sl, l, stencil_sum(m, sl, c, 1, 0, l));
}

// ---- spreading lists ----
// Per marker: stencil cells and weights; the weight sum in stencil order;
// the spreading weight phi*(1/h^3)/sum(phi) (k_ibm_spread_forces,
// fluid_kernels.cu:300-348); a marker whose weights sum to zero neither
// interpolates nor spreads.  Counts contributions per cell and flags the
// 256-cell blocks that receive any.
__device__ __forceinline__ void spread_count_wave(const LatticeGeom& g, const MarkerArrays& m, const SpreadLists& sl,
                                                  int l, int lane) {
  const StencilLane st = stencil_lane(g, m.lag[l], lane);
  float ws = 0.f;
#pragma unroll
  for (int j = 0; j < 64; ++j) ws += lane_read(st.w, j);
  const size_t k = tab(m, l, lane);
  if (lane == 0) sl.wsum[l] = ws;
  if (st.cell < 0 || ws == 0.0f) {
    sl.ctab[k] = -1;
    sl.rtab[k] = 0.f;
    return;
  }
  const float inv_h3 = (float)(1.0 / (double)(g.h * g.h * g.h));
  const float scale = 1 * inv_h3 / ws;
  sl.ctab[k] = st.cell;
  sl.rtab[k] = st.w;
  sl.wtab[k] = st.w * scale;
  atomicAdd(&sl.cnt[st.cell], 1);
  // the flag byte of the cell's wavefront (every writer stores the same 1)
  reinterpret_cast<unsigned char*>(sl.blkflag + st.cell / kFlagCells)[(st.cell % kFlagCells) / 64] = 1;
}

__device__ __forceinline__ void clear_flag_at(const SpreadLists& sl, int b) { sl.blkflag[b] = 0; }

// per flagged block: total contributions
__device__ __forceinline__ void block_sum_block(const LatticeGeom& g, const SpreadLists& sl, int b) {
  __shared__ int s_wave[kWavesPerBlock];
  if (!sl.blkflag[b]) {
    if (threadIdx.x == 0) sl.bsum[b] = 0;
    return;
  }
  const int i = b * kFlagCells + threadIdx.x;
  const int v = i < g.N ? sl.cnt[i] : 0;
  int total;
  block_excl_scan(v, s_wave, total);
  if (threadIdx.x == 0) sl.bsum[b] = total;
}

// exclusive scan of the block totals (one work-group, tiles of 256 x 8
// entries); bsum[nblk] = total
__device__ __forceinline__ void scan_blocks_block(const SpreadLists& sl) {
  __shared__ int s_wave[kWavesPerBlock];
  constexpr int kPer = 8;
  int carry = 0;
  for (int base = 0; base < sl.nblk; base += kThreads * kPer) {
    const int b0 = base + threadIdx.x * kPer;
    int v[kPer];
    int sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      v[j] = b0 + j < sl.nblk ? sl.bsum[b0 + j] : 0;
      sum += v[j];
    }
    int total;
    int run = carry + block_excl_scan(sum, s_wave, total);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if (b0 + j < sl.nblk) sl.bsum[b0 + j] = run;
      run += v[j];
    }
    carry += total;
  }
  if (threadIdx.x == 0) sl.bsum[sl.nblk] = carry;
}

// per flagged block: each cell's first entry and count; counts reset to 0
__device__ __forceinline__ void local_block(const LatticeGeom& g, const SpreadLists& sl, int b) {
  __shared__ int s_wave[kWavesPerBlock];
  if (!sl.blkflag[b]) return;
  const int i = b * kFlagCells + threadIdx.x;
  const int v = i < g.N ? sl.cnt[i] : 0;
  int total;
  const int ex = block_excl_scan(v, s_wave, total);
  if (i < g.N) {
    sl.rng[i] = make_int2(sl.bsum[b] + ex, v);
    sl.cnt[i] = 0;
  }
}

// table entry k (of marker l)
__device__ __forceinline__ void fill_at(const SpreadLists& sl, size_t k, int l) {
  const int cell = sl.ctab[k];
  if (cell < 0) return;
  const int pos = sl.rng[cell].x + atomicAdd(&sl.fillpos[cell], 1);
  SpreadEntry e;
  e.marker = l;
  e.w = sl.wtab[k];
  sl.ent[pos] = e;
}

// order each cell's entries by marker (insertion sort; a few dozen at most)
__device__ __forceinline__ void sort_at(const LatticeGeom& g, const SpreadLists& sl, int b) {
  if (!sl.blkflag[b]) return;
  const int i = b * kFlagCells + threadIdx.x;
  if (i >= g.N) return;
  const int2 r = sl.rng[i];
  sl.fillpos[i] = 0;
  SpreadEntry* e = sl.ent + r.x;
  for (int a = 1; a < r.y; ++a) {
    const SpreadEntry key = e[a];
    int j = a - 1;
    while (j >= 0 && e[j].marker > key.marker) {
      e[j + 1] = e[j];
      --j;
    }
    e[j + 1] = key;
  }
}

// ---------------------------------------------------------------- scalar
// Element `idx` of a lattice field (u as 3 floats per cell, or c): a 32-bit
// byte offset from a scalar base (every field of a lattice whose
// distributions fit the library's 2^32-element limit is below 4 GiB).
__device__ __forceinline__ float field_at(const GLSM_GLOBAL float* f, unsigned idx) {
  return *reinterpret_cast<const GLSM_GLOBAL float*>(reinterpret_cast<const GLSM_GLOBAL char*>(f) + idx * 4u);
}

// first-order upwind advection + 7-point diffusion + spread IBM source,
// forward Euler (k_zero_scalar + k_ibm_spread_concentration +
// k_convection_diffusion, fluid_kernels.cu:119-123, 414-459, 177-242)
static_assert(sizeof(SpreadEntry) == 8 && offsetof(SpreadEntry, w) == 4, "SpreadEntry is read as (int, float bits)");
__device__ __forceinline__ void scalar_at(const LatticeGeom& g, const ScalarParams& p, const float3* __restrict__ u,
                                          const float* __restrict__ c1, float* __restrict__ c2, const SpreadLists& sl,
                                          const float* __restrict__ Sl, int b) {
  const int i = b * kFlagCells + threadIdx.x;
  const int flagged = wave_flag(gmem(sl.blkflag)[b], threadIdx.x);
  if (i >= g.N) return;
  int x, y, z;
  decode_cell(g, i, x, y, z);
  const int Nx = g.Nx, Ny = g.Ny, Nz = g.Nz;
  const GLSM_GLOBAL float* gu = gmem(reinterpret_cast<const float*>(u));
  const GLSM_GLOBAL float* gc = gmem(c1);
  const int xp = (x + 1 < Nx) ? x + 1 : (g.perX ? 0 : x);
  const int xm = (x - 1 >= 0) ? x - 1 : (g.perX ? Nx - 1 : x);
  const int yp = (y + 1 < Ny) ? y + 1 : (g.perY ? 0 : y);
  const int ym = (y - 1 >= 0) ? y - 1 : (g.perY ? Ny - 1 : y);
  const int zp = (z + 1 < Nz) ? z + 1 : (g.perZ ? 0 : z);
  const int zm = (z - 1 >= 0) ? z - 1 : (g.perZ ? Nz - 1 : z);
  const int sxy = Nx * Ny;
  const int ixp = xp + y * Nx + z * sxy, ixm = xm + y * Nx + z * sxy;
  const int iyp = x + yp * Nx + z * sxy, iym = x + ym * Nx + z * sxy;
  const int izp = x + y * Nx + zp * sxy, izm = x + y * Nx + zm * sxy;
  float src = 0.f;
  if (flagged) {
    const GLSM_GLOBAL int* rp = gmem(reinterpret_cast<const int*>(sl.rng)) + 2 * i;
    const GLSM_GLOBAL int* ent = gmem(reinterpret_cast<const int*>(sl.ent));  // (marker, w bits) pairs
    const GLSM_GLOBAL float* gS = gmem(Sl);
    const int e0 = rp[0], e1 = e0 + rp[1];
    for (int k = e0; k < e1; k += 4) {  // four entries' loads ahead of the ordered adds
      int mk[4];
      float w[4], sv[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int kk = min(k + j, e1 - 1);
        mk[j] = ent[2 * kk];
        w[j] = __int_as_float(ent[2 * kk + 1]);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) sv[j] = gS[mk[j]];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (k + j < e1) src += sv[j] * w[j];
    }
  }
  const float ihx = p.ihx, ihy = p.ihy, ihz = p.ihz;
  const float ihx2 = ihx * ihx, ihy2 = ihy * ihy, ihz2 = ihz * ihz;
  const float c = field_at(gc, i);
  const float cxp = field_at(gc, ixp), cxm = field_at(gc, ixm), cyp = field_at(gc, iyp), cym = field_at(gc, iym),
              czp = field_at(gc, izp), czm = field_at(gc, izm);
  // the cell's own u as one 12-byte access (three 4-byte loads would touch
  // its lines three times)
  struct U3 {
    float x, y, z;
  };
  const U3 u0 = *reinterpret_cast<const GLSM_GLOBAL U3*>(reinterpret_cast<const GLSM_GLOBAL char*>(gu) + 12u * (unsigned)i);
  const float u0x = u0.x, u0y = u0.y, u0z = u0.z;
  const float uL = 0.5f * (field_at(gu, 3u * ixm) + u0x), uR = 0.5f * (u0x + field_at(gu, 3u * ixp));
  const float vL = 0.5f * (field_at(gu, 3u * iym + 1) + u0y), vR = 0.5f * (u0y + field_at(gu, 3u * iyp + 1));
  const float wL = 0.5f * (field_at(gu, 3u * izm + 2) + u0z), wR = 0.5f * (u0z + field_at(gu, 3u * izp + 2));
  const float FLx = uL * ((uL >= 0.0f) ? cxm : c), FLy = vL * ((vL >= 0.0f) ? cym : c), FLz = wL * ((wL >= 0.0f) ? czm : c);
  const float FRx = uR * ((uR >= 0.0f) ? c : cxp), FRy = vR * ((vR >= 0.0f) ? c : cyp), FRz = wR * ((wR >= 0.0f) ? c : czp);
  const float conv = -((FRx - FLx) * ihx + (FRy - FLy) * ihy + (FRz - FLz) * ihz);
  const float diff = p.D * ((cxm - 2.0f * c + cxp) * ihx2 + (cym - 2.0f * c + cyp) * ihy2 + (czm - 2.0f * c + czp) * ihz2);
  *reinterpret_cast<GLSM_GLOBAL float*>(reinterpret_cast<GLSM_GLOBAL char*>(gmem(c2)) + (unsigned)i * 4u) =
      c + p.dt * (conv + diff + src);
}

// ---------------------------------------------------------------- kernels
// Single-lattice kernels: one marker per thread / wavefront, one 256-cell
// block per work-group.
__global__ void __launch_bounds__(kThreads) k_pack(MarkerArrays m, float g2l) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l < m.M) pack_at(m, g2l, l);
}
__global__ void __launch_bounds__(kThreads) k_scatter(MarkerArrays m) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l < m.M) scatter_at(m, l);
}
__global__ void __launch_bounds__(kThreads) k_gel_marker_box(MarkerArrays m, const int* gel_off, float* box) {
  gel_marker_box(m, gel_off, box, blockIdx.x);
}
__global__ void __launch_bounds__(kThreads) k_repulsion_culled(MarkerArrays m, int ngel, const int* gel_off,
                                                               const float* box, float delta, float gamma) {
  // one marker per wavefront: its own fields as scalar loads
  const int a = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  if (a < m.M) repulsion_wave_culled(m, ngel, gel_off, box, delta, gamma, a, threadIdx.x & 63);
}
__global__ void __launch_bounds__(kThreads) k_repulsion(MarkerArrays m, float delta, float gamma) {
  const int a = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  if (a < m.M) repulsion_wave(m, delta, gamma, a, threadIdx.x & 63);  // whole wavefront
}
__global__ void __launch_bounds__(kThreads) k_ibm_interp(MarkerArrays m, SpreadLists sl, const float3* __restrict__ u,
                                                         float beta_eff) {
  const int l = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (l < m.M) interp_wave(m, sl, u, beta_eff, l, threadIdx.x & 63);
}
__global__ void __launch_bounds__(kThreads) k_ibm_sample(MarkerArrays m, SpreadLists sl, const float* __restrict__ c) {
  const int l = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (l < m.M) sample_wave(m, sl, c, l, threadIdx.x & 63);
}
__global__ void __launch_bounds__(kThreads) k_ibm_interp_t(MarkerArrays m, SpreadLists sl, const float3* __restrict__ u,
                                                           float beta_eff) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 3 * m.M) interp_component(m, sl, u, beta_eff, t / 3, t % 3);
}
__global__ void __launch_bounds__(kThreads) k_ibm_interp_s(MarkerArrays m, SpreadLists sl, const float3* __restrict__ u,
                                                           float beta_eff) {
  __shared__ InterpStage st;
  const int base = blockIdx.x * kStageMarkers, mi = threadIdx.x % kStageMarkers, part = threadIdx.x / kStageMarkers;
  const int l = base + mi;
  // phase 2's inputs (part < 3 = component): the marker's weight sum and
  // velocity, loaded with phase 1's
  const bool fin = part < 3 && l < m.M;
  const float ws = fin ? gmem(sl.wsum)[l] : 0.f;
  const float vl = fin ? gmem(reinterpret_cast<const float*>(m.Vl))[3 * (size_t)l + part] : 0.f;
  if (l < m.M) interp_stage_load(sl, m.M, u, l, part, mi, st);
  __syncthreads();
  if (fin) interp_stage_finish(m, beta_eff, l, part, ws, vl, interp_stage_sum(st, part, mi));
}
__global__ void __launch_bounds__(kThreads) k_ibm_sample_t(MarkerArrays m, SpreadLists sl, const float* __restrict__ c) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l < m.M) sample_thread(m, sl, c, l);
}
__global__ void __launch_bounds__(kThreads) k_spread_count(LatticeGeom g, MarkerArrays m, SpreadLists sl) {
  const int l = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (l < m.M) spread_count_wave(g, m, sl, l, threadIdx.x & 63);
}
__global__ void k_spread_clear_flags(SpreadLists sl) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < sl.nblk) clear_flag_at(sl, b);
}
__global__ void __launch_bounds__(kThreads) k_spread_block_sum(LatticeGeom g, SpreadLists sl) {
  block_sum_block(g, sl, blockIdx.x);
}
__global__ void __launch_bounds__(kThreads) k_spread_scan_blocks(SpreadLists sl) { scan_blocks_block(sl); }
__global__ void __launch_bounds__(kThreads) k_spread_local(LatticeGeom g, SpreadLists sl) {
  local_block(g, sl, blockIdx.x);
}
__global__ void __launch_bounds__(kThreads) k_spread_fill(int M, SpreadLists sl) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;  // lane-major table entry
  if (k < 64 * M) fill_at(sl, k, k % M);
}
__global__ void __launch_bounds__(kThreads) k_spread_sort(LatticeGeom g, SpreadLists sl) { sort_at(g, sl, blockIdx.x); }
__global__ void __launch_bounds__(kThreads) k_scalar(LatticeGeom g, ScalarParams p, const float3* __restrict__ u,
                                                     const float* __restrict__ c1, float* __restrict__ c2,
                                                     SpreadLists sl, const float* __restrict__ Sl) {
  scalar_at(g, p, u, c1, c2, sl, Sl, xcd_block<0>(blockIdx.x, gridDim.x));
}

// Batched kernels: markers and 256-cell blocks of all members are numbered
// consecutively (member i owns markers [mk0, mk0 + M) and blocks
// [blk0, blk0 + nblk)); mk_inst / blk_inst map a global number to its member.
__device__ __forceinline__ const BatchInst& marker_member(const BatchTables& t, int gm, int& l) {
  const BatchInst& I = t.inst[t.mk_inst[gm]];
  l = gm - I.mk0;
  return I;
}
__device__ __forceinline__ const BatchInst& block_member(const BatchTables& t, int gb, int& b) {
  const BatchInst& I = t.inst[t.blk_inst[gb]];
  b = gb - I.blk0;
  return I;
}

__global__ void __launch_bounds__(kThreads) kb_pack(BatchTables t) {
  const int gm = blockIdx.x * blockDim.x + threadIdx.x;
  if (gm >= t.M) return;
  int l;
  const BatchInst& I = marker_member(t, gm, l);
  pack_at(I.m, I.g2l, l);
}
__global__ void __launch_bounds__(kThreads) kb_scatter(BatchTables t) {
  const int gm = blockIdx.x * blockDim.x + threadIdx.x;
  if (gm >= t.M) return;
  int l;
  const BatchInst& I = marker_member(t, gm, l);
  scatter_at(I.m, l);
}
__global__ void __launch_bounds__(kThreads) kb_repulsion(BatchTables t) {
  // one marker per wavefront: its member lookup and own fields as scalar loads
  const int gm = __builtin_amdgcn_readfirstlane(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  if (gm >= t.M) return;
  int l;
  const BatchInst& I = marker_member(t, gm, l);
  repulsion_wave(I.m, I.delta, I.gamma, l, threadIdx.x & 63);
}
// beta_eff = beta * min(1, (iter + 1) / ramp) in FP32 (fluid.cu:265-270)
__device__ __forceinline__ float beta_eff_of(const BatchInst& I, int iter) {
  const float ramp = (I.ramp <= 0.0f) ? 1.0f : fminf(1.0f, (float)(iter + 1) / I.ramp);
  return I.beta * ramp;
}
__global__ void __launch_bounds__(kThreads) kb_ibm_interp(BatchTables t) {
  const int gm = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (gm >= t.M) return;
  int l;
  const BatchInst& I = marker_member(t, gm, l);
  interp_wave(I.m, I.sl, I.u, beta_eff_of(I, t.iters[t.mk_inst[gm]]), l, threadIdx.x & 63);
}
__global__ void __launch_bounds__(kThreads) kb_ibm_interp_t(BatchTables t) {
  const int g3 = blockIdx.x * blockDim.x + threadIdx.x;
  if (g3 >= 3 * t.M) return;
  const int gm = g3 / 3;
  int l;
  const BatchInst& I = marker_member(t, gm, l);
  const MarkerArrays m = I.m;  // values in registers: the table loop unrolls
  const SpreadLists sl = I.sl;
  interp_component(m, sl, I.u, beta_eff_of(I, t.iters[t.mk_inst[gm]]), l, g3 % 3);
}
// one work-group's staged interpolation of markers [base, base + 32) with
// member I holding marker gm (member-local index l)
__device__ __forceinline__ void interp_s_member(const BatchInst& I, int iter, int l, bool live, int part, int mi,
                                                InterpStage& st) {
  const bool fin = part < 3 && live;
  const float beta_eff = fin ? beta_eff_of(I, iter) : 0.f;
  const float ws = fin ? gmem(I.sl.wsum)[l] : 0.f;
  const float vl = fin ? gmem(reinterpret_cast<const float*>(I.m.Vl))[3 * (size_t)l + part] : 0.f;
  if (live) interp_stage_load(I.sl, I.m.M, I.u, l, part, mi, st);
  __syncthreads();
  if (fin) interp_stage_finish(I.m, beta_eff, l, part, ws, vl, interp_stage_sum(st, part, mi));
}
__global__ void __launch_bounds__(kThreads) kb_ibm_interp_s(BatchTables t) {
  __shared__ InterpStage st;
  const int base = blockIdx.x * kStageMarkers, mi = threadIdx.x % kStageMarkers, part = threadIdx.x / kStageMarkers;
  const int gm = base + mi;
  const bool live = gm < t.M;
  // Most work-groups' markers all belong to one member (members are
  // contiguous in the marker numbering): its descriptor and iteration are
  // then read once, as scalar loads, instead of per lane after a per-lane
  // member lookup.
  const int first = t.mk_inst[base], last = t.mk_inst[min(base + kStageMarkers, t.M) - 1];
  if (first == last) {
    const BatchInst& I = t.inst[first];
    interp_s_member(I, t.iters[first], (live ? gm : base) - I.mk0, live, part, mi, st);
  } else {
    int l = 0;
    const BatchInst& I = marker_member(t, live ? gm : 0, l);
    interp_s_member(I, live ? t.iters[t.mk_inst[gm]] : 0, l, live, part, mi, st);
  }
}
__global__ void __launch_bounds__(kThreads) kb_ibm_sample_t(BatchTables t, int cpar) {
  const int base = blockIdx.x * blockDim.x, gm = base + threadIdx.x;
  // one member for the whole work-group (the usual case): its descriptor as
  // scalar loads, as in the staged interpolation
  const int first = t.mk_inst[base], last = t.mk_inst[min(base + (int)blockDim.x, t.M) - 1];
  if (gm >= t.M) return;
  if (first == last) {
    const BatchInst& I = t.inst[first];
    sample_thread(I.m, I.sl, I.c[cpar], gm - I.mk0);
    return;
  }
  int l;
  const BatchInst& I = marker_member(t, gm, l);
  const MarkerArrays m = I.m;
  const SpreadLists sl = I.sl;
  sample_thread(m, sl, I.c[cpar], l);
}
__global__ void __launch_bounds__(kThreads) kb_ibm_sample(BatchTables t, int cpar) {
  const int gm = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (gm >= t.M) return;
  int l;
  const BatchInst& I = marker_member(t, gm, l);
  sample_wave(I.m, I.sl, I.c[cpar], l, threadIdx.x & 63);
}
__global__ void __launch_bounds__(kThreads) kb_spread_count(BatchTables t) {
  const int gm = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  if (gm >= t.M) return;
  int l;
  const BatchInst& I = marker_member(t, gm, l);
  spread_count_wave(I.g, I.m, I.sl, l, threadIdx.x & 63);
}
__global__ void kb_spread_clear_flags(BatchTables t) {
  const int gb = blockIdx.x * blockDim.x + threadIdx.x;
  if (gb >= t.nblk) return;
  int b;
  const BatchInst& I = block_member(t, gb, b);
  clear_flag_at(I.sl, b);
}
__global__ void __launch_bounds__(kThreads) kb_spread_block_sum(BatchTables t) {
  int b;
  const BatchInst& I = block_member(t, blockIdx.x, b);
  block_sum_block(I.g, I.sl, b);
}
__global__ void __launch_bounds__(kThreads) kb_spread_scan_blocks(BatchTables t) {
  scan_blocks_block(t.inst[blockIdx.x].sl);  // one work-group per member
}
__global__ void __launch_bounds__(kThreads) kb_spread_local(BatchTables t) {
  int b;
  const BatchInst& I = block_member(t, blockIdx.x, b);
  local_block(I.g, I.sl, b);
}
__global__ void __launch_bounds__(kThreads) kb_spread_fill(BatchTables t) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= 64L * t.M) return;
  int l;
  const BatchInst& I = marker_member(t, (int)(k >> 6), l);
  fill_at(I.sl, tab(I.m, l, (int)(k & 63)), l);
}
__global__ void __launch_bounds__(kThreads) kb_spread_sort(BatchTables t) {
  int b;
  const BatchInst& I = block_member(t, blockIdx.x, b);
  sort_at(I.g, I.sl, b);
}
// contiguous runs of blocks per XCD: the y/z neighbour rows a block reads
// were mostly fetched by the same XCD's L2
__global__ void __launch_bounds__(kThreads) kb_scalar(BatchTables t, int cpar) {
  int b;
  const BatchInst& I = block_member(t, xcd_block<0>(blockIdx.x, gridDim.x), b);
  scalar_at(I.g, I.sp, I.u, I.c[cpar], I.c[cpar ^ 1], I.sl, I.m.Sl, b);
}

}  // namespace

void launch_pack(const MarkerArrays& m, float g2l, hipStream_t s) {
  if (m.M == 0) return;
  k_pack<<<blocks_for(m.M, kThreads), kThreads, 0, s>>>(m, g2l);
  GLSM_LAUNCH_CHECK();
}
void launch_scatter(const MarkerArrays& m, hipStream_t s) {
  if (m.M == 0) return;
  k_scatter<<<blocks_for(m.M, kThreads), kThreads, 0, s>>>(m);
  GLSM_LAUNCH_CHECK();
}
// Process-wide switches, read when a launch is issued: a setter called from
// another host thread takes effect from the next launch on.
namespace {
std::atomic<int> g_rep_cull_markers{4096};  // from this many markers on (and two gels), skip far gels
}  // namespace
void set_repulsion_cull(int min_markers) {
  if (min_markers < 0) throw std::invalid_argument("repulsion cull threshold must be >= 0 (0 = never)");
  g_rep_cull_markers = min_markers;
}
void launch_repulsion(const MarkerArrays& m, float delta, float gamma, hipStream_t s, int ngel, const int* gel_off,
                      float* box) {
  if (m.M == 0) return;
  const int cull = g_rep_cull_markers.load(std::memory_order_relaxed);
  if (gel_off && box && ngel > 1 && cull > 0 && m.M >= cull) {
    k_gel_marker_box<<<ngel, kThreads, 0, s>>>(m, gel_off, box);
    k_repulsion_culled<<<blocks_for(m.M, kWavesPerBlock), kThreads, 0, s>>>(m, ngel, gel_off, box, delta, gamma);
  } else {
    k_repulsion<<<blocks_for(m.M, kWavesPerBlock), kThreads, 0, s>>>(m, delta, gamma);
  }
  GLSM_LAUNCH_CHECK();
}
// Interpolation form: 0 = default, 1 = wavefront per marker, 2 = thread
// per marker, 3 = staged through LDS.  All forms give the same bits.  The
// default interpolates in the staged form (measured fastest from a few
// hundred to ~30k markers); concentration sampling, which has the wavefront
// and thread forms only, takes the thread form from kThreadFormMarkers
// markers on (and in modes 2 and 3).
namespace {
std::atomic<int> g_interp_mode{0};
constexpr int kThreadFormMarkers = 2048;  // from this many markers on, one thread per marker
bool thread_form(int M) {
  const int mode = g_interp_mode.load(std::memory_order_relaxed);
  return mode >= 2 || (mode == 0 && M >= kThreadFormMarkers);
}
bool staged_form() {
  const int mode = g_interp_mode.load(std::memory_order_relaxed);
  return mode == 0 || mode == 3;
}
}  // namespace
void set_interp_mode(int mode) {
  if (mode < 0 || mode > 3) throw std::invalid_argument("interpolation mode must be 0, 1, 2 or 3");
  g_interp_mode = mode;
}

void launch_ibm_interp(const MarkerArrays& m, const SpreadLists& sl, const float3* u, float beta_eff,
                       hipStream_t s) {
  if (m.M == 0) return;
  if (staged_form())
    k_ibm_interp_s<<<blocks_for(m.M, kStageMarkers), kThreads, 0, s>>>(m, sl, u, beta_eff);
  else if (thread_form(m.M))
    k_ibm_interp_t<<<blocks_for(3L * m.M, kThreads), kThreads, 0, s>>>(m, sl, u, beta_eff);
  else
    k_ibm_interp<<<blocks_for(m.M, kWavesPerBlock), kThreads, 0, s>>>(m, sl, u, beta_eff);
  GLSM_LAUNCH_CHECK();
}
void launch_ibm_sample(const MarkerArrays& m, const SpreadLists& sl, const float* c, hipStream_t s) {
  if (m.M == 0) return;
  if (thread_form(m.M))
    k_ibm_sample_t<<<blocks_for(m.M, kThreads), kThreads, 0, s>>>(m, sl, c);
  else
    k_ibm_sample<<<blocks_for(m.M, kWavesPerBlock), kThreads, 0, s>>>(m, sl, c);
  GLSM_LAUNCH_CHECK();
}

void launch_spread_rebuild(const LatticeGeom& g, const MarkerArrays& m, const SpreadLists& sl, hipStream_t s) {
  k_spread_clear_flags<<<blocks_for(sl.nblk, kThreads), kThreads, 0, s>>>(sl);
  if (m.M > 0) k_spread_count<<<blocks_for(m.M, kWavesPerBlock), kThreads, 0, s>>>(g, m, sl);
  k_spread_block_sum<<<sl.nblk, kThreads, 0, s>>>(g, sl);
  k_spread_scan_blocks<<<1, kThreads, 0, s>>>(sl);
  k_spread_local<<<sl.nblk, kThreads, 0, s>>>(g, sl);
  if (m.M > 0) k_spread_fill<<<blocks_for((long)m.M * 64, kThreads), kThreads, 0, s>>>(m.M, sl);
  k_spread_sort<<<sl.nblk, kThreads, 0, s>>>(g, sl);
  GLSM_LAUNCH_CHECK();
}

void launch_scalar(const LatticeGeom& g, const ScalarParams& p, const float3* u, const float* c1, float* c2,
                   const SpreadLists& sl, const float* Sl, hipStream_t s) {
  k_scalar<<<blocks_for(g.N, kThreads), kThreads, 0, s>>>(g, p, u, c1, c2, sl, Sl);
  GLSM_LAUNCH_CHECK();
}

// ---- batched launches (all members in one launch per phase) ----
void launch_pack_batch(const BatchTables& t, hipStream_t s) {
  if (t.M == 0) return;
  kb_pack<<<blocks_for(t.M, kThreads), kThreads, 0, s>>>(t);
  GLSM_LAUNCH_CHECK();
}
void launch_scatter_batch(const BatchTables& t, hipStream_t s) {
  if (t.M == 0) return;
  kb_scatter<<<blocks_for(t.M, kThreads), kThreads, 0, s>>>(t);
  GLSM_LAUNCH_CHECK();
}
void launch_repulsion_batch(const BatchTables& t, hipStream_t s) {
  if (t.M == 0) return;
  kb_repulsion<<<blocks_for(t.M, kWavesPerBlock), kThreads, 0, s>>>(t);
  GLSM_LAUNCH_CHECK();
}
void launch_ibm_interp_batch(const BatchTables& t, hipStream_t s) {
  if (t.M == 0) return;
  if (staged_form())
    kb_ibm_interp_s<<<blocks_for(t.M, kStageMarkers), kThreads, 0, s>>>(t);
  else if (thread_form(t.M))
    kb_ibm_interp_t<<<blocks_for(3L * t.M, kThreads), kThreads, 0, s>>>(t);
  else
    kb_ibm_interp<<<blocks_for(t.M, kWavesPerBlock), kThreads, 0, s>>>(t);
  GLSM_LAUNCH_CHECK();
}
void launch_ibm_sample_batch(const BatchTables& t, int cpar, hipStream_t s) {
  if (t.M == 0) return;
  if (thread_form(t.M))
    kb_ibm_sample_t<<<blocks_for(t.M, kThreads), kThreads, 0, s>>>(t, cpar);
  else
    kb_ibm_sample<<<blocks_for(t.M, kWavesPerBlock), kThreads, 0, s>>>(t, cpar);
  GLSM_LAUNCH_CHECK();
}
void launch_spread_rebuild_batch(const BatchTables& t, hipStream_t s) {
  if (t.nblk == 0) return;
  kb_spread_clear_flags<<<blocks_for(t.nblk, kThreads), kThreads, 0, s>>>(t);
  if (t.M > 0) kb_spread_count<<<blocks_for(t.M, kWavesPerBlock), kThreads, 0, s>>>(t);
  kb_spread_block_sum<<<t.nblk, kThreads, 0, s>>>(t);
  kb_spread_scan_blocks<<<t.ninst, kThreads, 0, s>>>(t);
  kb_spread_local<<<t.nblk, kThreads, 0, s>>>(t);
  if (t.M > 0) kb_spread_fill<<<blocks_for(64L * t.M, kThreads), kThreads, 0, s>>>(t);
  kb_spread_sort<<<t.nblk, kThreads, 0, s>>>(t);
  GLSM_LAUNCH_CHECK();
}
void launch_scalar_batch(const BatchTables& t, int cpar, hipStream_t s) {
  if (t.nblk == 0) return;
  kb_scalar<<<t.nblk, kThreads, 0, s>>>(t, cpar);
  GLSM_LAUNCH_CHECK();
}

}  // namespace glsm
