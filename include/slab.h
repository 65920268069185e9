// slab.h -- one GPU's share of a z-decomposed lattice.
//
// Extension with no counterpart in the reference (which runs one lattice per
// process on one device: fluid.cu:263-284).  A lattice too large for one
// device, or to be spread over several, splits into z-slabs, one per GPU.
// A Slab holds planes [z0, z0 + nz) of a Fluid's lattice plus one halo plane
// below and one above, in the Fluid's storage and layout, and advances them
// with the same stream/collide kernel.  Between sub-steps the five
// populations that cross each face move to the neighbour:
//   pack     the own boundary plane's five populations into a contiguous
//            face buffer (5 x Nx*Ny values);
//   transfer the face buffer to the neighbour's receive buffer (across GPUs
//            an RCCL send/recv or peer copy over xGMI; here a device copy);
//   unpack   the received buffer into the halo plane.
// Across a global z wall there is no neighbour: the slab's own first or last
// plane bounces back, as in the whole lattice.  The marker (IBM) force is not
// part of a slab: with replicated markers each slab would gather the force
// of its own cells; a slab applies F_const only.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "fluid.h"

namespace glsm {
struct SlabDevice;
}

class Slab {
 public:
  // planes [z0, z0 + nz) of `like`'s lattice (its geometry, parameters,
  // periodicity and storage); nz >= 1
  Slab(const Fluid& like, int z0, int nz);
  ~Slab();
  Slab(const Slab&) = delete;
  Slab& operator=(const Slab&) = delete;

  // own planes and halos from `src`'s stored (post-collision) state; `src`
  // must have the lattice and storage of `like`
  void load(const Fluid& src);
  // own planes into `dst`'s stored state
  void store(Fluid& dst) const;
  // one stream/collide sub-step over all nz + 2 planes (the halos' results
  // are discarded by the next unpack); with write_u the velocities of every
  // own cell are stored
  void step(bool write_u = false);
  // face buffers: up = the last own plane's c_z = +1 populations, down = the
  // first own plane's c_z = -1 populations (5 * Nx * Ny elements each)
  void pack();
  // neighbour buffers into this slab's receive buffers (the transfer step)
  void receiveFromBelow(const Slab& below);
  void receiveFromAbove(const Slab& above);
  // receive buffers into the halo planes
  void unpack();

  int z0() const { return z0_; }
  int nz() const { return nz_; }
  int latticeNz() const { return Nz_; }
  size_t faceBytes() const;  // one face buffer
  bool wallBelow() const { return wall_lo_; }
  bool wallAbove() const { return wall_hi_; }

  // event timing of the stream/collide launch and of pack + unpack (ms,
  // averaged over the calls since enableTiming)
  void enableTiming(bool on);
  double lbmKernelMs() const;
  double exchangeKernelMs() const;

 private:
  int Nx_, Ny_, Nz_, z0_, nz_, storage_, scale_exp_;
  bool perZ_, wall_lo_, wall_hi_;
  size_t plane_;  // cells of the slab buffers (halos included) rounded up to whole 256-cell layout blocks
  std::unique_ptr<glsm::SlabDevice> dev_;  // geometry and collision parameters
  void* f_[2] = {nullptr, nullptr};
  float3* u_ = nullptr;
  int* flags_ = nullptr;  // no marker force: every block flag 0
  void* send_up_ = nullptr;
  void* send_down_ = nullptr;
  void* recv_below_ = nullptr;
  void* recv_above_ = nullptr;
  int cur_ = 0;
  hipStream_t s_;
  bool timing_ = false;
  mutable double lbm_ms_ = 0, xch_ms_ = 0;
  mutable int lbm_n_ = 0, xch_n_ = 0;
  hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
};

// One sub-step of a chain of slabs that covers the lattice in z order:
// every slab steps, packs its faces, receives its neighbours' faces
// (wrapping in a periodic lattice) and unpacks them into its halos.
void step_slab_chain(const std::vector<Slab*>& chain, bool write_u = false);
