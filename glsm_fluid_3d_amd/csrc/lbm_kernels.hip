// lbm_kernels.hip -- fused D3Q19 pull-scheme stream + BGK/Guo collide for
// gfx950.
//
// One launch per LBM sub-step replaces the reference's seven per-cell passes
// (k_set_force, k_vec_add, k_macros, k_collide, k_stream_bounce,
// k_zero_vector + the swap, fluid.cu:273-282, ~496 B/cell).  Each thread owns
// one cell and
//   1. pulls f_k(x,q) = fpost(x - c_q, q), or fpost(x, opp q) when x - c_q
//      lies outside a wall -- the pull form of the reference's push with
//      halfway bounce-back (fluid_kernels.cu:144-165); a periodic axis wraps
//      instead.  The buffers are in the 256-cell block layout (dist_index):
//      a work-group's cells store one contiguous region;
//   2. forms rho, u = (sum f c + F/2)/rho with F = F_const + F_ibm, where
//      F_ibm is gathered from the marker forces through the cell's sorted
//      (marker, weight) list -- only in the 64-cell wavefronts that some
//      marker stencil touches (one flag byte per wavefront, loaded with the
//      block's flag word);
//   3. relaxes to the second-order equilibrium with the Guo forcing term
//      (fluid_kernels.cu:98-142) and stores the 19 post-collision values.
// Algorithmic traffic per cell update: 19 loads + 19 stores of the storage
// type (152 B in FP32, 76 B in FP16/BF16); u (12 B) is stored only in the
// cells some marker stencil covers (those the interpolation reads) and, for
// every cell, on the last sub-step.  Adjacent lanes own adjacent cells, so
// every population access of a 64-lane wavefront is one contiguous run
// (shifted by the stream offset, and split where it crosses a block).
//
// Storage policies (collision always in FP32):
//   F32_FULL : f itself, the reference's representation;
//   F32_DEV  : d = f - w_q.  At the small velocities of this problem
//              (|u| ~ 1e-7 .. 1e-4) FP32 full f keeps only ~1e-8 absolute
//              precision, which is the whole velocity signal; deviations
//              keep ~1e-7 relative precision on it;
//   F16_DEV  : d * 2^s in binary16 (s = scale exponent, default 20);
//   BF16_DEV : d in bfloat16 (round to nearest even).
#include <atomic>
#include <cstddef>
#include <map>
#include <mutex>
#include <type_traits>

#include <hip/hip_ext.h>

#include "glsm_internal.hpp"

namespace glsm {
namespace {

constexpr int kThreads = 256;  // == cells per spreading block flag
constexpr int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0};
constexpr int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, 1, -1, -1};
constexpr int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, -1, -1, 1};
constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16};  // fluid.cu:215

// ---- storage policies ----
// Loads are plain (cached): non-temporal loads measured ~11% slower on
// gfx950.  Stores are non-temporal only when the two distribution buffers
// exceed the 256 MB Infinity Cache (NT): below that, plain stores leave the
// next sub-step's input in the cache.
template <class T, bool NT, class P>
__device__ __forceinline__ void put(P p, T v) {
  if (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}
// plain 3-float record for u and the precomputed cell forces: trivially
// copyable, so it is read and written through GLSM_GLOBAL pointers as one
// 12-byte access
struct Vec3 {
  float x, y, z;
};

template <int S>
struct Store;
// Values move in stored units: f (full storage), d = f - w (deviations),
// d * 2^scale_exp (FP16 deviations, `scaled`).  st() stores v and returns
// the value kept, read back as ld() would.
template <>
struct Store<F32_FULL> {
  using T = float;
  static constexpr bool dev = false, scaled = false;
  template <class P>
  __device__ static float ld(P p) {
    return *p;
  }
  template <bool NT, class P>
  __device__ static float st(P p, float v) {
    put<T, NT>(p, v);
    return v;
  }
};
template <>
struct Store<F32_DEV> {
  using T = float;
  static constexpr bool dev = true, scaled = false;
  template <class P>
  __device__ static float ld(P p) {
    return *p;
  }
  template <bool NT, class P>
  __device__ static float st(P p, float v) {
    put<T, NT>(p, v);
    return v;
  }
};
template <>
struct Store<F16_DEV> {
  using T = _Float16;
  static constexpr bool dev = true, scaled = true;
  template <class P>
  __device__ static float ld(P p) {
    return (float)*p;
  }
  template <bool NT, class P>
  __device__ static float st(P p, float v) {
    // v is rounded to FP32 first, then to FP16: the empty asm keeps the
    // compiler from fusing the multiply-add that produced v with the
    // conversion (a mixed-precision FMA rounds once, straight to FP16), a
    // choice it would otherwise make differently in different kernels
    asm("" : "+v"(v));
    const _Float16 h = (_Float16)v;
    put<T, NT>(p, h);
    return (float)h;
  }
  // stores v as st() does and returns acc + the value kept, as one
  // mixed-precision multiply-add (h * 1.0 + acc, rounded once: the same
  // value as converting h to FP32 and adding)
  template <bool NT, class P>
  __device__ static float st_add(P p, float v, float acc) {
    asm("" : "+v"(v));
    const _Float16 h = (_Float16)v;
    put<T, NT>(p, h);
    float r;
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(acc));
    return r;
  }
};
template <>
struct Store<BF16_DEV> {
  using T = uint16_t;
  static constexpr bool dev = true, scaled = false;
  template <class P>
  __device__ static float ld(P p) {
    return __uint_as_float((uint32_t)*p << 16);
  }
  template <bool NT, class P>
  __device__ static float st(P p, float v) {
    const uint32_t b = __float_as_uint(v);
    const uint32_t r = (b + 0x7FFFu + ((b >> 16) & 1u)) >> 16;  // round to nearest even (finite v)
    put<T, NT>(p, (uint16_t)r);
    return __uint_as_float(r << 16);
  }
  template <bool NT, class P>
  __device__ static float st_add(P p, float v, float acc) {
    return acc + st<NT>(p, v);
  }
};
// stored units <-> the population (or deviation) itself
template <int S>
__device__ __forceinline__ float to_value(float raw, const LbmParams& p) {
  return Store<S>::scaled ? raw * p.inv_scale : raw;
}
template <int S>
__device__ __forceinline__ float to_raw(float val, const LbmParams& p) {
  return Store<S>::scaled ? val * p.scale : val;
}

// Pull-source addressing of one cell: for each axis, the flat offset to the
// upstream neighbour of a +1 / -1 velocity and whether it exists.  Offsets
// are in units of E bytes: E = 1 gives cell offsets, E = sizeof(T) byte
// offsets (src_rel).
template <unsigned E = 1>
struct PullAddrT {
  int oxp, oxm, oyp, oym, ozp, ozm;
  bool kxp, kxm, kyp, kym, kzp, kzm;
  __device__ PullAddrT(const LatticeGeom& g, int x, int y, int z) {
    const int sx = (int)E, sy = g.Nx * (int)E, sz = g.Nx * g.Ny * (int)E;
    const bool xl = x > 0, xh = x < g.Nx - 1, yl = y > 0, yh = y < g.Ny - 1, zl = z > g.zlo, zh = z < g.zhi;
    kxp = xl || g.perX; oxp = xl ? -sx : (g.Nx - 1) * sx;
    kxm = xh || g.perX; oxm = xh ? sx : -(g.Nx - 1) * sx;
    kyp = yl || g.perY; oyp = yl ? -sy : (g.Ny - 1) * sy;
    kym = yh || g.perY; oym = yh ? sy : -(g.Ny - 1) * sy;
    kzp = zl || g.perZ; ozp = zl ? -sz : (g.Nz - 1) * sz;
    kzm = zh || g.perZ; ozm = zh ? sz : -(g.Nz - 1) * sz;
  }
  // Byte form (E = sizeof(T)): the source of population q arriving at the
  // cell whose offset in bytes is ib = i * E and whose population 0 sits at
  // byte eib = dist_index(i, 0) * E, as a byte offset from pull_base(f, q)
  // (kPullBias runs before population q's run; every offset is then
  // non-negative, also for a bounce-back to a population whose run lies
  // before q's in the buffer's first block).  Per source: one add, one
  // shift and one multiply-add; the run offset is uniform and goes to the
  // scalar base.
  template <int q>
  __device__ __forceinline__ unsigned src_rel(unsigned ib, unsigned eib) const {
    static_assert(E == 2 || E == 4, "byte form for 16- and 32-bit storage");
    constexpr unsigned SH = E == 2 ? 9 : 10;  // bytes of one run: kDistBlock * E = 2^SH
    static_assert(kDistBlock * E == (1u << SH), "256-value runs");
    bool ok = true;
    int off = 0;
    if (CX[q] > 0) { ok = ok && kxp; off += oxp; }
    if (CX[q] < 0) { ok = ok && kxm; off += oxm; }
    if (CY[q] > 0) { ok = ok && kyp; off += oyp; }
    if (CY[q] < 0) { ok = ok && kym; off += oym; }
    if (CZ[q] > 0) { ok = ok && kzp; off += ozp; }
    if (CZ[q] < 0) { ok = ok && kzm; off += ozm; }
    constexpr unsigned bias = kPullBias << SH;
    if (q == 0) return eib + bias;
    // jbb = (i + off) * E + 2 runs: jbb >> SH is then the source's block
    // + 2, and the multiply-add lands 2 * 18 + 2 = kPullBias runs past
    // dist_index(i + off, 0) * E
    static_assert(kPullBias == 2 * 18 + 2, "bias of the multiply-add form");
    const unsigned jbb = (unsigned)((int)ib + off) + (2u << SH);
    const unsigned ejb = jbb + (jbb >> SH) * (18u << SH);
    return ok ? ejb : eib + (unsigned)((OPP[q] - q + (int)kPullBias) * (int)(1u << SH));
  }
  // distribution element (dist_index) of the source of population q
  // arriving at cell i (E = 1)
  template <int q>
  __device__ __forceinline__ unsigned src(unsigned i) const {
    bool ok = true;
    int off = 0;
    if (CX[q] > 0) { ok = ok && kxp; off += oxp; }
    if (CX[q] < 0) { ok = ok && kxm; off += oxm; }
    if (CY[q] > 0) { ok = ok && kyp; off += oyp; }
    if (CY[q] < 0) { ok = ok && kym; off += oym; }
    if (CZ[q] > 0) { ok = ok && kzp; off += ozp; }
    if (CZ[q] < 0) { ok = ok && kzm; off += ozm; }
    // (cell, population) = (i + off, q), or (i, opp q) at a wall: the cell
    // part is selected and q's run offset added after, a uniform term
    const unsigned j = (unsigned)((int)i + off);
    const unsigned cell = ok ? dist_index(j, 0) : dist_index(i, 0) + (unsigned)(OPP[q] - q) * kDistBlock;
    return cell + (unsigned)q * kDistBlock;
  }
  // index into a plane-major array (population q of cell i at q * plane + i)
  // of the destination of population q leaving cell i (push): the cell it
  // streams into, or (i, opp q) when it bounces off a wall
  template <int q>
  __device__ __forceinline__ unsigned dst(unsigned i, unsigned plane) const {
    bool ok = true;
    int off = 0;
    if (CX[q] < 0) { ok = ok && kxp; off += oxp; }
    if (CX[q] > 0) { ok = ok && kxm; off += oxm; }
    if (CY[q] < 0) { ok = ok && kyp; off += oyp; }
    if (CY[q] > 0) { ok = ok && kym; off += oym; }
    if (CZ[q] < 0) { ok = ok && kzp; off += ozp; }
    if (CZ[q] > 0) { ok = ok && kzm; off += ozm; }
    return ok ? (unsigned)q * plane + (unsigned)((int)i + off) : (unsigned)OPP[q] * plane + i;
  }
};

using PullAddr = PullAddrT<1>;

template <int q>
struct Unroll {
  template <class F>
  __device__ __forceinline__ static void run(F&& f) {
    Unroll<q - 1>::run(f);
    f(std::integral_constant<int, q - 1>{});
  }
};
template <>
struct Unroll<0> {
  template <class F>
  __device__ __forceinline__ static void run(F&&) {}
};

__device__ __forceinline__ float wq(const LbmParams& p, int q) { return q == 0 ? p.w0 : (q < 7 ? p.w1 : p.w2); }

// Element `idx` of a population buffer.  B32 (buffer below 4 GiB): the base
// plus a 32-bit byte offset, which the load or store takes as an SGPR base +
// VGPR offset instead of a 64-bit lane address; otherwise 64-bit indexing.
template <bool B32, class T>
__device__ __forceinline__ T* elem_at(T* f, unsigned idx) {
  using C = typename std::conditional<std::is_const<T>::value, const GLSM_GLOBAL char, GLSM_GLOBAL char>::type;
  if (B32) return reinterpret_cast<T*>(reinterpret_cast<C*>(f) + idx * (unsigned)sizeof(T));
  return f + idx;
}
// Loads of the stream/collide kernel: population q's run start minus
// kPullBias runs, as a uniform address (formed as an integer: it may lie
// before the buffer; src_rel's offsets from it are at least kPullBias runs,
// so every address loaded is inside the buffer).
// A pointer every lane of the wavefront holds, as a scalar value (the
// batched kernels of several blocks per work-group pick the member per
// group of four wavefronts, which the compiler cannot see to be uniform).
template <class T>
__device__ __forceinline__ T* wave_uniform(T* p) {
  const uint64_t v = reinterpret_cast<uintptr_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v), hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

template <class T>
__device__ __forceinline__ const GLSM_GLOBAL T* pull_at(const GLSM_GLOBAL T* f, int q, unsigned rel) {
  uintptr_t base =
      reinterpret_cast<uintptr_t>(f) + (uintptr_t)(intptr_t)((q - (int)kPullBias) * (int)(kDistBlock * sizeof(T)));
  // kept a scalar-register value the compiler cannot fold into the per-lane
  // offset: the load is then scalar base + 32-bit lane offset, not a 64-bit
  // lane address built with four vector instructions
  asm("" : "+s"(base));
  return reinterpret_cast<const GLSM_GLOBAL T*>(base + rel);
}

// Population q of a cell in the block layout, from the start of q's run of
// the buffer's first block (f + q * kDistBlock, a uniform address) plus the
// cell's offset `rel`: bytes when B32, else values.
template <bool B32, class T>
__device__ __forceinline__ T* run_at(T* f, unsigned q, unsigned rel) {
  using C = typename std::conditional<std::is_const<T>::value, const GLSM_GLOBAL char, GLSM_GLOBAL char>::type;
  if (B32) return reinterpret_cast<T*>(reinterpret_cast<C*>(f + q * kDistBlock) + rel);
  return f + q * kDistBlock + rel;
}

// c_q . u: the sum of the nonzero components (c has components 0 and +-1,
// at most two of them nonzero), the value cx*ux + cy*uy + cz*uz rounds to
// (up to the sign of a zero result)
template <int q>
__device__ __forceinline__ float c_dot(float ux, float uy, float uz) {
  constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
  const float tx = cx > 0 ? ux : -ux, ty = cy > 0 ? uy : -uy, tz = cz > 0 ? uz : -uz;
  if (cx != 0 && cy != 0) return tx + ty;
  if (cx != 0 && cz != 0) return tx + tz;
  if (cy != 0 && cz != 0) return ty + tz;
  if (cx != 0) return tx;
  if (cy != 0) return ty;
  if (cz != 0) return tz;
  return 0.f;
}

// BGK + Guo on one cell.  v[] holds the stored units (f, d = f - w, or d
// scaled by S = 2^scale_exp); the same units are written back.  The scale
// enters per cell, not per population: rho = 1 + s/S, u = m / (rho S), the
// equilibrium deviation w (s + rho S poly), the Guo term times S.
// 16-bit storages: the moments are summed in FP64.  Every loaded value is a
// 16-bit number times a power of two, and the 19 of a cell span far fewer
// than 53 bits, so these sums are exact: rho and rho*u + F/2 are the exact
// moments of the stored populations, each rounded once to FP32 for the
// collision.  The collision conserves mass exactly (sum of the equilibria =
// rho, the Guo term sums to 0), so the rest population is stored last as the
// remainder of the cell's mass after the 18 moving ones were rounded: a cell
// keeps its stored mass up to the one rounding of that remainder instead of
// accumulating 19 roundings per sub-step.
template <int S, bool FORCED, bool NT, bool B32>
__device__ __forceinline__ void collide_store(const float (&v)[19], float Fx, float Fy, float Fz, const LbmParams& p,
                                              GLSM_GLOBAL typename Store<S>::T* __restrict__ fout, unsigned own,
                                              unsigned i, GLSM_GLOBAL Vec3* __restrict__ u, bool store_u) {
  constexpr bool DEV = Store<S>::dev;
  constexpr bool MOM64 = S == F16_DEV || S == BF16_DEV;
  constexpr bool SCALED = Store<S>::scaled;
  const float sc = SCALED ? p.scale : 1.0f, isc = SCALED ? p.inv_scale : 1.0f;
  float s, rho, mx, my, mz;  // (stored units) mass deviation or mass, momentum incl. F/2; rho itself
  double sd = 0.0;           // MOM64: exact mass deviation (stored units)
  if (MOM64) {
    // exact sums, so any grouping gives the same values: the six one-sided
    // momentum sums are shared with the mass sum
    double d[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) d[q] = (double)v[q];
    const double px = (d[1] + d[7]) + (d[10] + d[11]) + d[14], nx = (d[2] + d[8]) + (d[9] + d[12]) + d[13];
    const double py = (d[3] + d[7]) + (d[8] + d[15]) + d[16], ny = (d[4] + d[9]) + (d[10] + d[17]) + d[18];
    const double pz = (d[5] + d[11]) + (d[12] + d[15]) + d[18], nz = (d[6] + d[13]) + (d[14] + d[16]) + d[17];
    sd = ((d[0] + px) + nx) + (((d[3] + d[4]) + (d[5] + d[6])) + ((d[15] + d[16]) + (d[17] + d[18])));
    double mxd = px - nx, myd = py - ny, mzd = pz - nz;
    if (FORCED) {
      const double h = 0.5 * (double)sc;  // F/2 in stored units (exact)
      mxd += h * (double)Fx;
      myd += h * (double)Fy;
      mzd += h * (double)Fz;
    }
    s = (float)sd;
    rho = (float)(1.0 + sd * (double)isc);
    mx = (float)mxd;
    my = (float)myd;
    mz = (float)mzd;
  } else {
    s = 0.f;
#pragma unroll
    for (int q = 0; q < 19; ++q) s += v[q];
    rho = DEV ? 1.0f + s : s;  // DEV: s = rho - 1
    mx = (v[1] - v[2]) + (v[7] - v[8]) + (v[10] - v[9]) + (v[11] - v[12]) + (v[14] - v[13]);
    my = (v[3] - v[4]) + (v[7] - v[10]) + (v[8] - v[9]) + (v[15] - v[17]) + (v[16] - v[18]);
    mz = (v[5] - v[6]) + (v[11] - v[14]) + (v[12] - v[13]) + (v[15] - v[16]) + (v[18] - v[17]);
    if (FORCED) {
      mx += 0.5f * Fx;
      my += 0.5f * Fy;
      mz += 0.5f * Fz;
    }
  }
  const float inv_rho = isc / rho;
  const float ux = mx * inv_rho, uy = my * inv_rho, uz = mz * inv_rho;
  if (store_u) u[i] = Vec3{ux, uy, uz};
  // The multiply-adds below are written as fmaf: each rounds once, the same
  // way in every kernel this body is inlined into.
  const float omega = p.omega;
  const float uu15 = 1.5f * fmaf(ux, ux, fmaf(uy, uy, uz * uz));
  const float uF = FORCED ? fmaf(ux, Fx, fmaf(uy, Fy, uz * Fz)) : 0.f;
  const float kF = 1.0f - 0.5f * omega;
  const float rhos = SCALED ? rho * sc : rho;
  // The weights enter each population's equilibrium separately, and the
  // relaxation is v + omega (feq - v): precomputed per-weight coefficients
  // (omega w rho) or the form (1 - omega) v + omega feq round the same way
  // for every population of a weight class, and the lattice momentum then
  // drifted the same way every sub-step (201^3 four-gel u error against
  // FP64 8.9e-6 -> 1.4e-5 after 50 steps).
  float kept = 0.f;  // MOM64: sum of the stored moving populations
  Unroll<19>::run([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    if (MOM64 && q == 0) return;  // stored last
    const float w = wq(p, q);
    const float cu = c_dot<q>(ux, uy, uz);
    const float poly = fmaf(cu, fmaf(4.5f, cu, 3.0f), -uu15);
    // feq = w rho (1 + poly); deviation form: feq - w = w (s + rho poly);
    // out = v + omega (feq - v)
    const float neq = DEV ? fmaf(w, fmaf(rhos, poly, s), -v[q]) : fmaf(w * rho, 1.0f + poly, -v[q]);
    float out = fmaf(omega, neq, v[q]);
    if (FORCED) {
      // Guo: w (1 - omega/2) (3 (c.F - u.F) + 9 (c.u)(c.F))
      const float ciF = c_dot<q>(Fx, Fy, Fz);
      out = fmaf(SCALED ? w * kF * sc : w * kF, fmaf(9.0f * cu, ciF, 3.0f * (ciF - uF)), out);
    }
    if constexpr (MOM64)
      kept = Store<S>::template st_add<NT>(run_at<B32>(fout, q, own), out, kept);
    else
      Store<S>::template st<NT>(run_at<B32>(fout, q, own), out);
  });
  if (MOM64) Store<S>::template st<NT>(run_at<B32>(fout, 0, own), (float)(sd - (double)kept));
}

// Every access goes through GLSM_GLOBAL pointers: the batched kernels take
// their pointers from a device-side table, which would otherwise compile to
// flat memory instructions.
// PRIO: the wavefront runs at raised priority while it forms addresses and
// issues its 19 loads, and at normal priority for the collision, so that
// wavefronts about to load get issue slots ahead of those computing
template <int S, bool NT, bool B32, bool PRIO = false>
__device__ __forceinline__ void lbm_block(const LatticeGeom& g, const LbmParams& p,
                                          const typename Store<S>::T* __restrict__ fin,
                                          typename Store<S>::T* __restrict__ fout, float3* __restrict__ u,
                                          int write_u_all, const SpreadLists& sl,
                                          const float3* __restrict__ Fl, int const_force, unsigned b,
                                          unsigned lane) {
  const unsigned i = b * kThreads + lane;
  const int flag_word = gmem(sl.blkflag)[b];  // loaded first, tested after the population loads
  if (i >= (unsigned)g.N) return;
  if constexpr (PRIO) __builtin_amdgcn_s_setprio(3);
  int x, y, z;
  decode_cell(g, (int)i, x, y, z);
  using T = typename Store<S>::T;
  constexpr unsigned E = sizeof(T);
  const GLSM_GLOBAL T* gin = gmem(wave_uniform(fin));
  float v[19];
  // own: the cell's population-0 offset, bytes (B32) or values
  unsigned own;
  if constexpr (B32) {
    constexpr unsigned SH = E == 2 ? 9 : 10;
    const unsigned ib = i * E;
    own = ib + (ib >> SH) * (18u << SH);  // dist_index(i, 0) * E
    const PullAddrT<E> a(g, x, y, z);
    Unroll<19>::run([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      v[q] = Store<S>::ld(pull_at(gin, q, a.template src_rel<q>(ib, own)));
    });
  } else {
    own = dist_index(i, 0);
    const PullAddr a(g, x, y, z);
    Unroll<19>::run([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      v[q] = Store<S>::ld(gin + a.src<q>(i));
    });
  }
  if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
  const int flagged = wave_flag(flag_word, lane);
  bool store_u = write_u_all;
  GLSM_GLOBAL Vec3* gu = gmem(reinterpret_cast<Vec3*>(u));
  if (flagged || const_force) {
    float Fx = 0.f, Fy = 0.f, Fz = 0.f;
    if (flagged) {
      // sum of Fl[marker] * w over the cell's list, in list (marker) order;
      // four entries, then their four marker forces, are loaded ahead of
      // the ordered adds (an entry past the end re-reads the last one and
      // is not added).  Products and sums are rounded one by one.
#pragma clang fp contract(off)
      const GLSM_GLOBAL int* rp = gmem(reinterpret_cast<const int*>(sl.rng)) + 2 * i;
      const int e0 = rp[0], e1 = e0 + rp[1];
      store_u = store_u || e1 > e0;  // a stencil cell: the interpolation reads its u
      static_assert(sizeof(SpreadEntry) == 8 && offsetof(SpreadEntry, w) == 4, "entries are (int, float bits)");
      const GLSM_GLOBAL int* ent = gmem(reinterpret_cast<const int*>(sl.ent));
      const GLSM_GLOBAL float* gF = gmem(reinterpret_cast<const float*>(Fl));
      for (int k = e0; k < e1; k += 4) {
        int mk[4];
        float w[4], fx[4], fy[4], fz[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int kk = min(k + j, e1 - 1);
          mk[j] = ent[2 * kk];
          w[j] = __int_as_float(ent[2 * kk + 1]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          fx[j] = gF[3 * mk[j]];
          fy[j] = gF[3 * mk[j] + 1];
          fz[j] = gF[3 * mk[j] + 2];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (k + j < e1) {
            Fx += fx[j] * w[j];
            Fy += fy[j] * w[j];
            Fz += fz[j] * w[j];
          }
      }
    }
    collide_store<S, true, NT, B32>(v, p.Fc.x + Fx, p.Fc.y + Fy, p.Fc.z + Fz, p, gmem(fout), own, i, gu, store_u);
  } else {
    collide_store<S, false, NT, B32>(v, 0.f, 0.f, 0.f, p, gmem(fout), own, i, gu, store_u);
  }
}

// 8 waves per SIMD (<= 64 VGPRs): keeps enough loads in flight per CU.
// RUN: the XCD run length of the block order (xcd_block).  16-bit storages
// raise the wave priority around the loads (measured: +0.7% at 201^3; the
// FP32 and batched kernels were slower with it).
template <int S, bool NT, bool B32, int RUN>
__global__ void __launch_bounds__(kThreads, 8) k_lbm_step(LatticeGeom g, LbmParams p,
                                                       const typename Store<S>::T* __restrict__ fin,
                                                       typename Store<S>::T* __restrict__ fout,
                                                       float3* __restrict__ u, int write_u_all, SpreadLists sl,
                                                       const float3* __restrict__ Fl, int const_force) {
  const unsigned b = xcd_block<RUN>(blockIdx.x, gridDim.x);
  lbm_block<S, NT, B32, sizeof(typename Store<S>::T) == 2>(g, p, fin, fout, u, write_u_all, sl, Fl, const_force, b,
                                                            threadIdx.x);
}

// All members of a batch: the blocks of every lattice form one flattened
// index (member i's at [blk0, blk0 + nblk)); each work-group walks it with
// stride gridDim.x (persistent when the grid is smaller than the block
// count), the XCD remap applied to the flattened index.
template <int S, bool NT, int RUN>
__global__ void __launch_bounds__(kThreads, 8) kb_lbm_step(BatchTables t, int sel, int write_u_all) {
  using T = typename Store<S>::T;
  for (unsigned w = blockIdx.x; w < (unsigned)t.nblk; w += gridDim.x) {
    const unsigned gb = xcd_block<RUN>(w, (unsigned)t.nblk);
    const BatchInst& I = t.inst[t.blk_inst[gb]];
    lbm_block<S, NT, true>(I.g, I.p, static_cast<const T*>(I.f[sel]), static_cast<T*>(I.f[sel ^ 1]),
                           I.u, write_u_all, I.sl, I.m.Fdrag, I.const_force, gb - (unsigned)I.blk0, threadIdx.x);
  }
}

// The same in work units of WGB * chunk consecutive 256-cell blocks: a
// work-group of WGB * 256 threads takes one unit, its WGB groups of four
// waves each one block per round, for `chunk` rounds over consecutive
// blocks (the shifted sources' edge lines of round r are read again in round
// r + 1 by the same CU).  One work-group per unit, the XCD remap applied to
// the unit index.
template <int S, bool NT, int RUN, int WGB>
__global__ void __launch_bounds__(kThreads * WGB, 8) kb_lbm_step_units(BatchTables t, int sel, int write_u_all,
                                                                       int chunk) {
  using T = typename Store<S>::T;
  const unsigned per_unit = (unsigned)(WGB * chunk);
  const unsigned nunits = ((unsigned)t.nblk + per_unit - 1) / per_unit;
  const unsigned base = xcd_block<RUN>(blockIdx.x, nunits) * per_unit + threadIdx.x / kThreads;
  for (int r = 0; r < chunk; ++r) {
    const unsigned gb = base + (unsigned)(r * WGB);
    if (gb >= (unsigned)t.nblk) break;
    const BatchInst& I = t.inst[t.blk_inst[gb]];
    lbm_block<S, NT, true>(I.g, I.p, static_cast<const T*>(I.f[sel]), static_cast<T*>(I.f[sel ^ 1]),
                           I.u, write_u_all, I.sl, I.m.Fdrag, I.const_force, gb - (unsigned)I.blk0,
                           threadIdx.x % kThreads);
  }
}

__global__ void k_advance_iters(int* __restrict__ iters, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) iters[i] += 1;
}

// k_init, fluid_kernels.cu:84-96: f = w_q (a rest state, which is its own
// post-collision state), u = 0, c1 = c2 = 0
template <int S>
__global__ void __launch_bounds__(kThreads) k_lbm_init(LatticeGeom g, LbmParams p,
                                                       typename Store<S>::T* __restrict__ f,
                                                       float3* __restrict__ u, float* __restrict__ c1,
                                                       float* __restrict__ c2) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= (unsigned)g.N) return;
#pragma unroll
  for (int q = 0; q < 19; ++q)
    Store<S>::template st<false>(f + dist_index(i, q), Store<S>::dev ? 0.f : wq(p, q));
  u[i] = make_float3(0.f, 0.f, 0.f);
  c1[i] = 0.f;
  c2[i] = 0.f;
}

// stored post-collision state -> full FP32 populations; PULL = the
// pre-collision populations of the next sub-step, else the stored ones
template <int S, bool PULL>
__global__ void __launch_bounds__(kThreads) k_lbm_export(LatticeGeom g, LbmParams p,
                                                         const typename Store<S>::T* __restrict__ fpost,
                                                         float* __restrict__ out, unsigned out_plane) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= (unsigned)g.N) return;
  int x, y, z;
  decode_cell(g, (int)i, x, y, z);
  const PullAddr a(g, x, y, z);
  Unroll<19>::run([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const unsigned s = PULL ? a.src<q>(i) : dist_index(i, q);
    const float v = to_value<S>(Store<S>::ld(fpost + s), p);
    out[(unsigned)q * out_plane + i] = Store<S>::dev ? wq(p, q) + v : v;
  });
}

// inverse of the PULL export: fpost(y,q) = f(y + c_q, q), or f(y, opp q) at
// a wall, converted to the storage representation
template <int S>
__global__ void __launch_bounds__(kThreads) k_lbm_import(LatticeGeom g, LbmParams p, const float* __restrict__ in,
                                                         typename Store<S>::T* __restrict__ fpost, unsigned in_plane) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= (unsigned)g.N) return;
  int x, y, z;
  decode_cell(g, (int)i, x, y, z);
  const PullAddr a(g, x, y, z);
  Unroll<19>::run([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    const float v = in[a.dst<q>(i, in_plane)];  // from plane q or opp q; w_opp(q) == w_q
    Store<S>::template st<false>(fpost + dist_index(i, q), to_raw<S>(Store<S>::dev ? v - wq(p, q) : v, p));
  });
}

// ---- FP64 conservation reductions ----
// Mass excess sum(rho - 1) and momentum sum(rho u) = sum_q f_q c_q of the
// pre-collision populations of the next sub-step (what k_macros sums,
// fluid_kernels.cu:98-111, without the F/2 term).  Per cell the 19 values
// are summed in FP64; each work-group reduces its cells with wave shuffles
// and a fixed-order combine of its 4 waves; k_reduce_partials adds the
// per-group partials in a fixed order.  The result is deterministic.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int K>
__device__ __forceinline__ void block_sum_store(double (&acc)[K], double* __restrict__ out) {
  __shared__ double part[kThreads / 64][K];
  const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = wave_sum(acc[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) part[wave][k] = acc[k];
  __syncthreads();
  if (threadIdx.x < K) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) t += part[w][threadIdx.x];
    out[threadIdx.x] = t;
  }
}

template <int S>
__global__ void __launch_bounds__(kThreads) k_lbm_moments(LatticeGeom g, LbmParams p,
                                                          const typename Store<S>::T* __restrict__ fpost,
                                                          double* __restrict__ partial) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < (unsigned)g.N) {
    int x, y, z;
    decode_cell(g, (int)i, x, y, z);
    const PullAddr a(g, x, y, z);
    double m = Store<S>::dev ? 0.0 : -1.0;
    Unroll<19>::run([&](auto qc) {
      constexpr int q = decltype(qc)::value;
      const double f = (double)to_value<S>(Store<S>::ld(fpost + a.src<q>(i)), p);
      // deviation storage: sum d = rho - 1 (the weights sum to 1), and
      // sum d c = sum f c (the weights of opposite velocities are equal)
      m += f;
      if (CX[q]) acc[1] += CX[q] > 0 ? f : -f;
      if (CY[q]) acc[2] += CY[q] > 0 ? f : -f;
      if (CZ[q]) acc[3] += CZ[q] > 0 ? f : -f;
    });
    acc[0] = m;
  }
  block_sum_store<4>(acc, partial + 4ull * blockIdx.x);
}

__global__ void __launch_bounds__(kThreads) k_reduce_partials(const double* __restrict__ partial, int nblk,
                                                              double* __restrict__ out) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblk; b += kThreads)
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[k] += partial[4ll * b + k];
  block_sum_store<4>(acc, out);
}

// ---- health scan and density field ----
// k_lbm_health: per cell, whether any of its 19 stored values, its velocity
// or its concentration is not finite (a binary16 deviation whose magnitude
// passed 65504 / 2^s was stored as +-Inf: saturation shows up here), and the
// largest finite |f_q - w_q| in value units.  Each wavefront combines its
// lanes with shuffles and adds its count / raises the maximum with one
// integer atomic: a count and a maximum do not depend on the order of the
// atomics.  The maximum is kept as the bit pattern of a non-negative double,
// whose integer order is its numeric order.
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <int S>
__global__ void __launch_bounds__(kThreads) k_lbm_health(LatticeGeom g, LbmParams p,
                                                         const typename Store<S>::T* __restrict__ fpost,
                                                         const float3* __restrict__ u,
                                                         const float* __restrict__ c,
                                                         unsigned long long* __restrict__ out2) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  unsigned long long bad = 0;
  double mx = 0.0;
  if (i < (unsigned)g.N) {
    bool ok = true;
#pragma unroll
    for (int q = 0; q < 19; ++q) {
      const float v = to_value<S>(Store<S>::ld(fpost + dist_index(i, q)), p);
      const float d = Store<S>::dev ? v : v - wq(p, q);
      if (__builtin_isfinite(d))
        mx = fmax(mx, fabs((double)d));
      else
        ok = false;
    }
    const float3 ui = u[i];
    ok = ok && __builtin_isfinite(ui.x) && __builtin_isfinite(ui.y) && __builtin_isfinite(ui.z) &&
         __builtin_isfinite(c[i]);
    bad = ok ? 0ull : 1ull;
  }
  bad = wave_sum_u64(bad);
  mx = wave_max(mx);
  if ((threadIdx.x & 63u) == 0) {
    if (bad) atomicAdd(out2, bad);
    atomicMax(out2 + 1, (unsigned long long)__double_as_longlong(mx));
  }
}

// n values of a device array that are not finite, added to *out
__global__ void __launch_bounds__(kThreads) k_count_nonfinite(const double* __restrict__ a, size_t n,
                                                              unsigned long long* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  unsigned long long bad = (i < n && !__builtin_isfinite(a[i])) ? 1ull : 0ull;
  bad = wave_sum_u64(bad);
  if ((threadIdx.x & 63u) == 0 && bad) atomicAdd(out, bad);
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// rho - 1 of the pre-collision populations of the next sub-step, per cell,
// summed in FP64 from the stored values (deviation storage: sum of the
// deviations, since the weights sum to 1)
template <int S>
__global__ void __launch_bounds__(kThreads) k_lbm_density(LatticeGeom g, LbmParams p,
                                                          const typename Store<S>::T* __restrict__ fpost,
                                                          double* __restrict__ out) {
  const unsigned i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= (unsigned)g.N) return;
  int x, y, z;
  decode_cell(g, (int)i, x, y, z);
  const PullAddr a(g, x, y, z);
  double m = Store<S>::dev ? 0.0 : -1.0;
  Unroll<19>::run([&](auto qc) {
    constexpr int q = decltype(qc)::value;
    m += (double)to_value<S>(Store<S>::ld(fpost + a.src<q>(i)), p);
  });
  out[i] = m;
}

// plain device copy (16 B per lane): the reference point for achievable HBM
// bandwidth and for calibrating the FETCH_SIZE / WRITE_SIZE counters
__global__ void __launch_bounds__(kThreads) k_copy16(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
  const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n4) b[i] = a[i];
}

inline int blocks_for(long n) { return (int)((n + kThreads - 1) / kThreads); }

// A timed launch passes start/stop events that the dispatch itself
// timestamps at kernel start and end (hipExtLaunchKernelGGL), so the sample
// is the kernel's own duration, without launch latency.
// LDS reserved per stream/collide work-group, which caps the work-groups
// resident per CU (160 KB of LDS per CU): -1 = automatic, else bytes.
// Fewer resident work-groups keep fewer DRAM streams open: with FP32 storage
// streamed from HBM the 201^3 step ran 3.5% faster at 3 work-groups per CU
// than at 8 (the register limit), and a 64-instance FP32 batch 6% faster at
// 5 (its autotuner picks); 16-bit storage and cache-resident lattices need
// the full occupancy to cover their arithmetic.
std::atomic<int> g_lds_reserve{-1};
size_t lds_reserve_lattice(size_t elem, bool nt) {
  const int v = g_lds_reserve.load(std::memory_order_relaxed);
  if (v >= 0) return (size_t)v;
  return elem == 4 && nt ? (size_t)kLatticeLdsReserve : 0;
}
size_t lds_reserve_batch(const LbmLaunch& sh) {
  const int v = g_lds_reserve.load(std::memory_order_relaxed);
  return v >= 0 ? (size_t)v : (size_t)sh.lds;
}

template <class K, class... A>
void launch_maybe_timed(K kernel, int grid, size_t lds, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, A... args) {
  if (ev0)
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads), lds, s, ev0, ev1, 0, args...);
  else
    kernel<<<grid, kThreads, lds, s>>>(args...);
}

// a distribution buffer whose every byte offset fits in 32 bits
bool fits_b32(size_t plane, size_t elem) { return dist_fits_b32(plane, elem); }

std::atomic<int> g_lattice_run{-1};  // forced XCD run length of the single-lattice kernel, -1 = automatic

template <int S>
void step_impl(const LatticeGeom& g, const LbmParams& p, const void* fin, void* fout, size_t plane, float3* u,
               bool write_u_all, const SpreadLists& sl, const float3* Fl, hipStream_t s, hipEvent_t ev0,
               hipEvent_t ev1) {
  using T = typename Store<S>::T;
  const int const_force = (p.Fc.x != 0.f || p.Fc.y != 0.f || p.Fc.z != 0.f) ? 1 : 0;
  // both ping-pong buffers larger than ~3/4 of the 256 MB Infinity Cache:
  // stream the stores past it
  const bool nt = 2.0 * 19.0 * (double)plane * sizeof(T) > 192.0 * (1 << 20);
  const T* in = static_cast<const T*>(fin);
  T* out = static_cast<T*>(fout);
  const int wu = write_u_all ? 1 : 0;
  const bool b32 = fits_b32(plane, sizeof(T));
  auto go = [&](auto kernel) {
    launch_maybe_timed(kernel, blocks_for(g.N), lds_reserve_lattice(sizeof(T), nt), s, ev0, ev1, g, p, in, out, u,
                       wu, sl, Fl, const_force);
  };
  // XCD run length: 16-block runs once the state streams from HBM, one
  // contiguous run per XCD while it stays in the Infinity Cache, unless
  // set_lattice_xcd_run() forced another
  const int forced = g_lattice_run.load(std::memory_order_relaxed);
  const int run = forced >= 0 ? forced : (nt ? 16 : 0);
  if (!b32) {
    nt ? go(k_lbm_step<S, true, false, 16>) : go(k_lbm_step<S, false, false, 0>);
    return;
  }
  switch (run * 2 + (nt ? 1 : 0)) {
    case 0: go(k_lbm_step<S, false, true, 0>); break;
    case 1: go(k_lbm_step<S, true, true, 0>); break;
    case 8: go(k_lbm_step<S, false, true, 4>); break;
    case 9: go(k_lbm_step<S, true, true, 4>); break;
    case 32: go(k_lbm_step<S, false, true, 16>); break;
    case 33: go(k_lbm_step<S, true, true, 16>); break;
    case 128: go(k_lbm_step<S, false, true, 64>); break;
    case 129: go(k_lbm_step<S, true, true, 64>); break;
    default: throw std::invalid_argument("XCD run length must be 0, 4, 16 or 64");
  }
}

}  // namespace

size_t storage_bytes(int storage) { return storage == F16_DEV || storage == BF16_DEV ? 2 : 4; }

void set_lbm_lds_reserve(int bytes) {
  if (bytes < -1 || bytes > 64 * 1024) throw std::invalid_argument("LDS reserve must be -1 (automatic) or 0..65536 bytes");
  g_lds_reserve = bytes;
}

void set_lattice_xcd_run(int run) {
  if (run != -1 && run != 0 && run != 4 && run != 16 && run != 64)
    throw std::invalid_argument("XCD run length must be -1 (automatic), 0, 4, 16 or 64");
  g_lattice_run = run;
}

void launch_copy16(const void* a, void* b, size_t bytes, hipStream_t s) {
  const size_t n4 = bytes / 16;
  k_copy16<<<(unsigned)((n4 + kThreads - 1) / kThreads), kThreads, 0, s>>>(static_cast<const float4*>(a),
                                                                            static_cast<float4*>(b), n4);
  GLSM_LAUNCH_CHECK();
}

void launch_lbm_step(int storage, const LatticeGeom& g, const LbmParams& p, const void* fin, void* fout,
                     size_t plane, float3* u, bool write_u_all, const SpreadLists& sl, const float3* Fl,
                     hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  switch (storage) {
    case F32_FULL: step_impl<F32_FULL>(g, p, fin, fout, plane, u, write_u_all, sl, Fl, s, ev0, ev1); break;
    case F32_DEV: step_impl<F32_DEV>(g, p, fin, fout, plane, u, write_u_all, sl, Fl, s, ev0, ev1); break;
    case F16_DEV: step_impl<F16_DEV>(g, p, fin, fout, plane, u, write_u_all, sl, Fl, s, ev0, ev1); break;
    case BF16_DEV: step_impl<BF16_DEV>(g, p, fin, fout, plane, u, write_u_all, sl, Fl, s, ev0, ev1); break;
    default: throw std::invalid_argument("unknown storage mode");
  }
  GLSM_LAUNCH_CHECK();
}

#define GLSM_STORAGE_DISPATCH(storage, KERNEL_CALL)                 \
  switch (storage) {                                                \
    case F32_FULL: { constexpr int S = F32_FULL; KERNEL_CALL; } break; \
    case F32_DEV: { constexpr int S = F32_DEV; KERNEL_CALL; } break;   \
    case F16_DEV: { constexpr int S = F16_DEV; KERNEL_CALL; } break;   \
    case BF16_DEV: { constexpr int S = BF16_DEV; KERNEL_CALL; } break; \
    default: throw std::invalid_argument("unknown storage mode");   \
  }

template <int S, int WGB>
void step_batch_units(const BatchTables& t, int sel, int wu, const LbmLaunch& sh, hipStream_t s, hipEvent_t ev0,
                      hipEvent_t ev1) {
  const int per_unit = WGB * sh.chunk;
  const int grid = (t.nblk + per_unit - 1) / per_unit;
  auto go = [&](auto kernel) {
    if (ev0)
      hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(kThreads * WGB), 0, s, ev0, ev1, 0, t, sel, wu, sh.chunk);
    else
      kernel<<<grid, kThreads * WGB, 0, s>>>(t, sel, wu, sh.chunk);
  };
  switch (sh.run * 2 + (sh.nt ? 1 : 0)) {
    case 0: go(kb_lbm_step_units<S, false, 0, WGB>); break;
    case 1: go(kb_lbm_step_units<S, true, 0, WGB>); break;
    case 8: go(kb_lbm_step_units<S, false, 4, WGB>); break;
    case 9: go(kb_lbm_step_units<S, true, 4, WGB>); break;
    case 32: go(kb_lbm_step_units<S, false, 16, WGB>); break;
    case 33: go(kb_lbm_step_units<S, true, 16, WGB>); break;
    default: throw std::invalid_argument("work units of several blocks take an XCD run length of 0, 4 or 16");
  }
}

template <int S>
void step_batch_impl(const BatchTables& t, int sel, bool write_u_all, const LbmLaunch& sh, hipStream_t s,
                     hipEvent_t ev0, hipEvent_t ev1) {
  const int wu = write_u_all ? 1 : 0;
  if (sh.wgb != 1 || sh.chunk != 1) {
    if (sh.grid != 0) throw std::invalid_argument("work units of several blocks take one work-group per unit");
    switch (sh.wgb) {
      case 1: step_batch_units<S, 1>(t, sel, wu, sh, s, ev0, ev1); break;
      case 2: step_batch_units<S, 2>(t, sel, wu, sh, s, ev0, ev1); break;
      case 4: step_batch_units<S, 4>(t, sel, wu, sh, s, ev0, ev1); break;
      default: throw std::invalid_argument("work-group size must be 256, 512 or 1024 threads");
    }
    return;
  }
  const int grid = sh.grid > 0 && sh.grid < t.nblk ? sh.grid : t.nblk;
  const size_t lds = lds_reserve_batch(sh);
  auto go = [&](auto kernel) { launch_maybe_timed(kernel, grid, lds, s, ev0, ev1, t, sel, wu); };
  switch (sh.run * 2 + (sh.nt ? 1 : 0)) {
    case 0: go(kb_lbm_step<S, false, 0>); break;
    case 1: go(kb_lbm_step<S, true, 0>); break;
    case 8: go(kb_lbm_step<S, false, 4>); break;
    case 9: go(kb_lbm_step<S, true, 4>); break;
    case 32: go(kb_lbm_step<S, false, 16>); break;
    case 33: go(kb_lbm_step<S, true, 16>); break;
    case 128: go(kb_lbm_step<S, false, 64>); break;
    case 129: go(kb_lbm_step<S, true, 64>); break;
    default: throw std::invalid_argument("XCD run length must be 0, 4, 16 or 64");
  }
}

void launch_lbm_step_batch(int storage, const BatchTables& t, int sel, bool write_u_all, const LbmLaunch& shape,
                           hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  if (t.nblk == 0) return;
  // (members' distribution buffers are below 4 GiB: checked when they join)
  if (shape.run != 0 && shape.run != 4 && shape.run != 16 && shape.run != 64)
    throw std::invalid_argument("XCD run length must be 0, 4, 16 or 64");
  if (shape.chunk < 1 || shape.chunk > 64) throw std::invalid_argument("blocks per round chain must be in [1, 64]");
  GLSM_STORAGE_DISPATCH(storage, (step_batch_impl<S>(t, sel, write_u_all, shape, s, ev0, ev1)));
  GLSM_LAUNCH_CHECK();
}

void launch_advance_iters(int* iters, int n, hipStream_t s) {
  if (n == 0) return;
  k_advance_iters<<<blocks_for(n), kThreads, 0, s>>>(iters, n);
  GLSM_LAUNCH_CHECK();
}

void launch_lbm_init(int storage, const LatticeGeom& g, const LbmParams& p, void* f, float3* u, float* c1,
                     float* c2, hipStream_t s) {
  GLSM_STORAGE_DISPATCH(storage, (k_lbm_init<S><<<blocks_for(g.N), kThreads, 0, s>>>(
                                     g, p, static_cast<typename Store<S>::T*>(f), u, c1, c2)));
  GLSM_LAUNCH_CHECK();
}

void launch_lbm_export(int storage, bool pull, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                       float* out, size_t out_plane, hipStream_t s) {
  if (pull) {
    GLSM_STORAGE_DISPATCH(storage, (k_lbm_export<S, true><<<blocks_for(g.N), kThreads, 0, s>>>(
                                       g, p, static_cast<const typename Store<S>::T*>(fpost), out,
                                       (unsigned)out_plane)));
  } else {
    GLSM_STORAGE_DISPATCH(storage, (k_lbm_export<S, false><<<blocks_for(g.N), kThreads, 0, s>>>(
                                       g, p, static_cast<const typename Store<S>::T*>(fpost), out,
                                       (unsigned)out_plane)));
  }
  GLSM_LAUNCH_CHECK();
}

void launch_lbm_moments(int storage, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                        double* partial, double* out4, hipStream_t s) {
  const int nb = blocks_for(g.N);
  GLSM_STORAGE_DISPATCH(storage, (k_lbm_moments<S><<<nb, kThreads, 0, s>>>(
                                     g, p, static_cast<const typename Store<S>::T*>(fpost), partial)));
  GLSM_LAUNCH_CHECK();
  k_reduce_partials<<<1, kThreads, 0, s>>>(partial, nb, out4);
  GLSM_LAUNCH_CHECK();
}

void launch_lbm_health(int storage, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                       const float3* u, const float* c, unsigned long long* out2, hipStream_t s) {
  k_fill_u64<<<1, 64, 0, s>>>(out2, 2);
  GLSM_LAUNCH_CHECK();
  GLSM_STORAGE_DISPATCH(storage, (k_lbm_health<S><<<blocks_for(g.N), kThreads, 0, s>>>(
                                     g, p, static_cast<const typename Store<S>::T*>(fpost), u, c, out2)));
  GLSM_LAUNCH_CHECK();
}

void launch_zero_u64(unsigned long long* p, int n, hipStream_t s) {
  if (n <= 0) return;
  k_fill_u64<<<(n + 63) / 64, 64, 0, s>>>(p, n);
  GLSM_LAUNCH_CHECK();
}

void launch_count_nonfinite(const double* a, size_t n, unsigned long long* out, hipStream_t s) {
  if (n == 0) return;
  k_count_nonfinite<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, s>>>(a, n, out);
  GLSM_LAUNCH_CHECK();
}

void launch_lbm_density(int storage, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                        double* out, hipStream_t s) {
  GLSM_STORAGE_DISPATCH(storage, (k_lbm_density<S><<<blocks_for(g.N), kThreads, 0, s>>>(
                                     g, p, static_cast<const typename Store<S>::T*>(fpost), out)));
  GLSM_LAUNCH_CHECK();
}

// ---- allocation-free FP64 conservation sums ----
// Scratch of conservation_static(): the four FP64 partial sums of up to
// kStaticBlocks 256-cell blocks, then the four results -- 4 MiB in a
// file-scope device array that every device the library runs on holds until
// the process ends.  In host code `scratch` names a host-side stand-in, not
// device memory: its device address is looked up once per device
// (hipGetSymbolAddress, with that device current) and only that address is
// passed to kernels and copies.  Every entry a call reads was written by the
// same call first (each block stores all four of its partials), so nothing
// needs clearing; calls that overlap on one device share the array and race.
namespace {
constexpr int kStaticBlocks = 131072;  // lattices of up to 33,554,432 cells
__device__ __align__(256) double scratch[4 * kStaticBlocks + 4];
std::mutex g_scratch_mu;
std::map<int, double*> g_scratch_ptr;  // device -> device address of `scratch`
}  // namespace

double* static_scratch() {
  int dev = 0;
  GLSM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it != g_scratch_ptr.end()) return it->second;
  void* ptr = nullptr;
  GLSM_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) throw std::runtime_error("hipGetSymbolAddress returned no address for the conservation scratch");
  g_scratch_ptr[dev] = static_cast<double*>(ptr);
  return static_cast<double*>(ptr);
}

long static_scratch_max_cells() { return (long)kStaticBlocks * kThreads; }

void conservation_static(int storage, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                         double* host_out4, hipStream_t s) {
  const int nb = blocks_for(g.N);
  if (nb > kStaticBlocks)
    throw std::invalid_argument("conservation without allocation: the lattice has " + std::to_string(g.N) +
                                " cells, the static scratch holds at most " +
                                std::to_string(static_scratch_max_cells()));
  double* partial = static_scratch();
  double* out4 = partial + 4ull * kStaticBlocks;
  GLSM_STORAGE_DISPATCH(storage, (k_lbm_moments<S><<<nb, kThreads, 0, s>>>(
                                     g, p, static_cast<const typename Store<S>::T*>(fpost), partial)));
  GLSM_LAUNCH_CHECK();
  k_reduce_partials<<<1, kThreads, 0, s>>>(partial, nb, out4);
  GLSM_LAUNCH_CHECK();
  GLSM_HIP(hipMemcpyAsync(host_out4, out4, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
  GLSM_HIP(hipStreamSynchronize(s));
}

void launch_lbm_import(int storage, const LatticeGeom& g, const LbmParams& p, const float* in, void* fpost,
                       size_t in_plane, hipStream_t s) {
  GLSM_STORAGE_DISPATCH(storage, (k_lbm_import<S><<<blocks_for(g.N), kThreads, 0, s>>>(
                                     g, p, in, static_cast<typename Store<S>::T*>(fpost), (unsigned)in_plane)));
  GLSM_LAUNCH_CHECK();
}

}  // namespace glsm
