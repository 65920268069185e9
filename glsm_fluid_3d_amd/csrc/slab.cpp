// slab.cpp -- z-slab decomposition of one lattice (include/slab.h).
//
// A slab is a lattice of Nx x Ny x (nz + 2) planes advanced by the same
// stream/collide kernel as a whole lattice (k_lbm_step).  Its geometry says
// which planes have a lower / upper neighbour (LatticeGeom::zlo, zhi): the
// halos do, so the own planes pull across the face from them; across a
// global wall the first / last own plane bounces back.  The halo planes are
// computed too (two planes of wasted work) and overwritten by the unpack.
#include "slab.h"

#include <stdexcept>
#include <vector>

#include "glsm_internal.hpp"

namespace glsm {
struct SlabDevice {
  LatticeGeom g;
  LbmParams p;
  SpreadLists sl;
};
}  // namespace glsm

Slab::Slab(const Fluid& like, int z0, int nz) : dev_(new glsm::SlabDevice()) {
  const glsm::LatticeGeom lg = glsm::fluid_geom(like);
  if (nz < 1 || z0 < 0 || z0 + nz > lg.Nz) throw std::invalid_argument("Slab: planes outside the lattice");
  Nx_ = lg.Nx;
  Ny_ = lg.Ny;
  Nz_ = lg.Nz;
  z0_ = z0;
  nz_ = nz;
  storage_ = like.storage();
  scale_exp_ = like.storageScaleExp();
  perZ_ = lg.perZ != 0;
  wall_lo_ = !perZ_ && z0 == 0;
  wall_hi_ = !perZ_ && z0 + nz == Nz_;
  glsm::LatticeGeom& g = dev_->g;
  g = lg;
  g.Nz = nz + 2;
  g.N = Nx_ * Ny_ * (nz + 2);
  g.perZ = 0;
  g.zlo = wall_lo_ ? 1 : 0;
  g.zhi = wall_hi_ ? nz : nz + 1;
  dev_->p = glsm::fluid_lbm_params(like);
  plane_ = glsm::dist_cells((size_t)g.N);
  if (19ull * plane_ >= (1ull << 32)) throw std::invalid_argument("Slab: too large for 32-bit element indexing");
  const size_t elem = glsm::storage_bytes(storage_);
  for (auto& f : f_) f = glsm::dalloc<char>(19 * plane_ * elem);
  u_ = glsm::dalloc<float3>(g.N);
  const int nblk = (g.N + 255) / 256;
  flags_ = glsm::dalloc<int>(nblk);
  const size_t face = faceBytes();
  send_up_ = glsm::dalloc<char>(face);
  send_down_ = glsm::dalloc<char>(face);
  recv_below_ = glsm::dalloc<char>(face);
  recv_above_ = glsm::dalloc<char>(face);
  glsm::SpreadLists& sl = dev_->sl;
  sl = glsm::SpreadLists{};
  sl.nblk = nblk;
  sl.blkflag = flags_;
  s_ = glsm::stream();
  for (auto& e : ev_) e = glsm::timing_event();
}

Slab::~Slab() {
  (void)hipStreamSynchronize(s_);
  for (auto& e : ev_) (void)hipEventDestroy(e);
  for (void* p : {f_[0], f_[1], (void*)u_, (void*)flags_, send_up_, send_down_, recv_below_, recv_above_})
    glsm::dfree(p);
}

size_t Slab::faceBytes() const { return 5ull * Nx_ * Ny_ * glsm::storage_bytes(storage_); }

void Slab::load(const Fluid& src) {
  const glsm::LatticeGeom sg = glsm::fluid_geom(src);
  if (sg.Nx != Nx_ || sg.Ny != Ny_ || sg.Nz != Nz_ || src.storage() != storage_)
    throw std::invalid_argument("Slab::load: fluid has another lattice or storage");
  glsm::launch_slab_gather(glsm::storage_bytes(storage_), src.d_f, f_[cur_], (size_t)Nx_ * Ny_, z0_, nz_, Nz_,
                           perZ_ ? 1 : 0, s_);
  GLSM_HIP(hipStreamSynchronize(s_));
}

void Slab::store(Fluid& dst) const {
  const glsm::LatticeGeom dg = glsm::fluid_geom(dst);
  if (dg.Nx != Nx_ || dg.Ny != Ny_ || dg.Nz != Nz_ || dst.storage() != storage_)
    throw std::invalid_argument("Slab::store: fluid has another lattice or storage");
  glsm::launch_slab_scatter(glsm::storage_bytes(storage_), f_[cur_], dst.d_f, (size_t)Nx_ * Ny_, z0_, nz_, s_);
  GLSM_HIP(hipStreamSynchronize(s_));
}

void Slab::step(bool write_u) {
  hipEvent_t e0 = timing_ ? ev_[0] : nullptr, e1 = timing_ ? ev_[1] : nullptr;
  glsm::launch_lbm_step(storage_, dev_->g, dev_->p, f_[cur_], f_[cur_ ^ 1], plane_, u_, write_u, dev_->sl, nullptr,
                        s_, e0, e1);
  cur_ ^= 1;
  if (timing_) {
    GLSM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GLSM_HIP(hipEventElapsedTime(&ms, e0, e1));
    lbm_ms_ += ms;
    ++lbm_n_;
  }
}

void Slab::pack() {
  const size_t elem = glsm::storage_bytes(storage_), nxy = (size_t)Nx_ * Ny_;
  if (timing_) GLSM_HIP(hipEventRecord(ev_[2], s_));
  if (!wall_hi_) glsm::launch_face_pack(elem, f_[cur_], nxy, nz_, 0, send_up_, s_);
  if (!wall_lo_) glsm::launch_face_pack(elem, f_[cur_], nxy, 1, 1, send_down_, s_);
}

void Slab::receiveFromBelow(const Slab& below) {
  if (wall_lo_) throw std::invalid_argument("Slab: nothing below a wall");
  if (below.Nx_ != Nx_ || below.Ny_ != Ny_ || below.storage_ != storage_)
    throw std::invalid_argument("Slab: neighbour of another lattice");
  GLSM_HIP(hipMemcpyAsync(recv_below_, below.send_up_, faceBytes(), hipMemcpyDeviceToDevice, s_));
}

void Slab::receiveFromAbove(const Slab& above) {
  if (wall_hi_) throw std::invalid_argument("Slab: nothing above a wall");
  if (above.Nx_ != Nx_ || above.Ny_ != Ny_ || above.storage_ != storage_)
    throw std::invalid_argument("Slab: neighbour of another lattice");
  GLSM_HIP(hipMemcpyAsync(recv_above_, above.send_down_, faceBytes(), hipMemcpyDeviceToDevice, s_));
}

void Slab::unpack() {
  const size_t elem = glsm::storage_bytes(storage_), nxy = (size_t)Nx_ * Ny_;
  if (!wall_lo_) glsm::launch_face_unpack(elem, recv_below_, f_[cur_], nxy, 0, 0, s_);
  if (!wall_hi_) glsm::launch_face_unpack(elem, recv_above_, f_[cur_], nxy, nz_ + 1, 1, s_);
  if (timing_) {
    GLSM_HIP(hipEventRecord(ev_[3], s_));
    GLSM_HIP(hipEventSynchronize(ev_[3]));
    float ms = 0;
    GLSM_HIP(hipEventElapsedTime(&ms, ev_[2], ev_[3]));
    xch_ms_ += ms;
    ++xch_n_;
  }
}

void Slab::enableTiming(bool on) {
  GLSM_HIP(hipStreamSynchronize(s_));
  timing_ = on;
  lbm_ms_ = xch_ms_ = 0;
  lbm_n_ = xch_n_ = 0;
}

double Slab::lbmKernelMs() const { return lbm_n_ ? lbm_ms_ / lbm_n_ : 0.0; }
double Slab::exchangeKernelMs() const { return xch_n_ ? xch_ms_ / xch_n_ : 0.0; }

void step_slab_chain(const std::vector<Slab*>& chain, bool write_u) {
  const int n = (int)chain.size();
  if (n == 0) throw std::invalid_argument("slab chain: empty");
  for (int i = 0; i + 1 < n; ++i)
    if (chain[i]->z0() + chain[i]->nz() != chain[i + 1]->z0())
      throw std::invalid_argument("slab chain: slabs must be contiguous and in z order");
  if (chain[0]->z0() != 0 || chain[n - 1]->z0() + chain[n - 1]->nz() != chain[0]->latticeNz())
    throw std::invalid_argument("slab chain: slabs must cover the lattice");
  for (Slab* s : chain) s->step(write_u);
  for (Slab* s : chain) s->pack();
  for (int i = 0; i < n; ++i) {
    if (!chain[i]->wallBelow()) chain[i]->receiveFromBelow(*chain[(i + n - 1) % n]);
    if (!chain[i]->wallAbove()) chain[i]->receiveFromAbove(*chain[(i + 1) % n]);
  }
  for (Slab* s : chain) s->unpack();
}
