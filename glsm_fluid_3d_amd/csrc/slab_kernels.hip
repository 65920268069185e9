// slab_kernels.hip -- data movement of the z-slab decomposition (gfx950):
// copying planes between a whole lattice and a slab with halos, and packing
// / unpacking the five populations that cross a slab face.  Pure copies of
// the stored representation (2- or 4-byte elements), no arithmetic.
//
// Layout on both sides: the distribution block layout (dist_index) over
// cells x + Nx*(y + Ny*z).  A slab's z-planes hold z' = 0 (halo below),
// 1..nz (own planes z0..z0+nz-1), nz+1 (halo above).
#include <cstdint>

#include "glsm_internal.hpp"

namespace glsm {
namespace {

constexpr int kThreads = 256;
// populations with c_z = +1 (they cross a face upwards) and c_z = -1
// (fluid.cu:202-207 velocity table)
__constant__ int kUp[5] = {5, 11, 12, 15, 18};
__constant__ int kDown[5] = {6, 13, 14, 16, 17};

inline unsigned blocks_for(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// slab plane zs' (0..nz+1) <- lattice plane z0 - 1 + zs' (wrapped when
// periodic; a plane outside a walled lattice is left as it is)
template <class E>
__global__ void __launch_bounds__(kThreads) k_slab_gather(const E* __restrict__ src, E* __restrict__ dst, size_t nxy,
                                                          int z0, int nzp2, int Nz, int perZ) {
  const size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const size_t per_q = nxy * nzp2;
  if (t >= 19 * per_q) return;
  const int q = (int)(t / per_q);
  const size_t r = t - (size_t)q * per_q;
  const int zs = (int)(r / nxy);
  const size_t j = r - (size_t)zs * nxy;
  int z = z0 - 1 + zs;
  if (z < 0 || z >= Nz) {
    if (!perZ) return;
    z = (z + Nz) % Nz;
  }
  dst[dist_index((unsigned)((size_t)zs * nxy + j), q)] = src[dist_index((unsigned)((size_t)z * nxy + j), q)];
}

// lattice planes z0..z0+nz-1 <- slab planes 1..nz
template <class E>
__global__ void __launch_bounds__(kThreads) k_slab_scatter(const E* __restrict__ src, E* __restrict__ dst, size_t nxy,
                                                           int z0, int nz) {
  const size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  const size_t per_q = nxy * nz;
  if (t >= 19 * per_q) return;
  const int q = (int)(t / per_q);
  const size_t r = t - (size_t)q * per_q;
  dst[dist_index((unsigned)((size_t)z0 * nxy + r), q)] = src[dist_index((unsigned)(nxy + r), q)];
}

// face buffer (5 x nxy) <- populations `which` (0: c_z = +1, 1: c_z = -1)
// of slab plane z
template <class E>
__global__ void __launch_bounds__(kThreads) k_face_pack(const E* __restrict__ f, size_t nxy, int z, int which,
                                                        E* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= 5 * nxy) return;
  const int k = (int)(t / nxy);
  const size_t j = t - (size_t)k * nxy;
  const int q = which ? kDown[k] : kUp[k];
  out[t] = f[dist_index((unsigned)((size_t)z * nxy + j), q)];
}

// slab plane z, populations `which` <- face buffer
template <class E>
__global__ void __launch_bounds__(kThreads) k_face_unpack(const E* __restrict__ in, E* __restrict__ f, size_t nxy,
                                                          int z, int which) {
  const size_t t = (size_t)blockIdx.x * kThreads + threadIdx.x;
  if (t >= 5 * nxy) return;
  const int k = (int)(t / nxy);
  const size_t j = t - (size_t)k * nxy;
  const int q = which ? kDown[k] : kUp[k];
  f[dist_index((unsigned)((size_t)z * nxy + j), q)] = in[t];
}

}  // namespace

#define GLSM_ELEM_DISPATCH(bytes, CALL)                     \
  do {                                                      \
    if ((bytes) == 2) {                                     \
      using E = uint16_t;                                   \
      CALL;                                                 \
    } else if ((bytes) == 4) {                              \
      using E = uint32_t;                                   \
      CALL;                                                 \
    } else {                                                \
      throw std::invalid_argument("element size must be 2 or 4"); \
    }                                                       \
  } while (0)

void launch_slab_gather(size_t elem, const void* src, void* dst, size_t nxy, int z0, int nz, int Nz, int perZ,
                        hipStream_t s) {
  GLSM_ELEM_DISPATCH(elem, (k_slab_gather<E><<<blocks_for(19 * nxy * (nz + 2)), kThreads, 0, s>>>(
                               static_cast<const E*>(src), static_cast<E*>(dst), nxy, z0, nz + 2, Nz, perZ)));
  GLSM_LAUNCH_CHECK();
}

void launch_slab_scatter(size_t elem, const void* src, void* dst, size_t nxy, int z0, int nz, hipStream_t s) {
  GLSM_ELEM_DISPATCH(elem, (k_slab_scatter<E><<<blocks_for(19 * nxy * nz), kThreads, 0, s>>>(
                               static_cast<const E*>(src), static_cast<E*>(dst), nxy, z0, nz)));
  GLSM_LAUNCH_CHECK();
}

void launch_face_pack(size_t elem, const void* f, size_t nxy, int z, int which, void* out, hipStream_t s) {
  GLSM_ELEM_DISPATCH(elem, (k_face_pack<E><<<blocks_for(5 * nxy), kThreads, 0, s>>>(
                               static_cast<const E*>(f), nxy, z, which, static_cast<E*>(out))));
  GLSM_LAUNCH_CHECK();
}

void launch_face_unpack(size_t elem, const void* in, void* f, size_t nxy, int z, int which, hipStream_t s) {
  GLSM_ELEM_DISPATCH(elem, (k_face_unpack<E><<<blocks_for(5 * nxy), kThreads, 0, s>>>(
                               static_cast<const E*>(in), static_cast<E*>(f), nxy, z, which)));
  GLSM_LAUNCH_CHECK();
}

}  // namespace glsm
