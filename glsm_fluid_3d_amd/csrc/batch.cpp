// batch.cpp -- many gel-fluid instances advanced by one set of launches.
//
// Per time step, in sim.cpp's order (sim.cpp:66-81) but phase-major across
// the members:
//   gel phases of every gel (elasticity where the member's iteration is a
//   multiple of 5, setZero, chemistry, centre record on multiples of 1000);
//   pack of every marker; re-binning of the IBM spreading lists of every
//   lattice; 25 x (interpolation of every marker, one stream/collide launch
//   over the blocks of every lattice); concentration sample and step;
//   repulsion; scatter; iteration counters + 1.
// The members' iteration counters live in device memory, so the cadences
// and the IBM ramp are evaluated on the device and the same launch sequence
// serves every step: it is captured once per buffer parity in a HIP graph.
// Marker positions change only on elasticity steps (iteration multiple of
// 5) or from outside the batch, so the spreading lists are re-binned only on
// steps where some member needs it (a second graph variant).
#include "batch.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "glsm_internal.hpp"

namespace glsm {

template <class T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  void assign(const std::vector<T>& v, hipStream_t s) {
    if (v.size() > n) {
      dfree(p);
      p = dalloc<T>(v.size());
      n = v.size();
    }
    if (!v.empty()) GLSM_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  ~DevArray() { dfree(p); }
};

struct BatchDevice {
  DevArray<BatchInst> inst;
  DevArray<BatchInst> tune_inst;  // copy with u redirected to scratch (autotuning)
  DevArray<int> blk_inst, mk_inst, iters;
  DevArray<GelView> views;
  DevArray<int> gel_inst;
  DevArray<float3> scratch_u;
  BatchTables tables{};
  int ngel = 0;
  int max_L = 0;
};

}  // namespace glsm

namespace {
glsm::LbmLaunch to_internal(const Batch::Launch& l) {
  glsm::LbmLaunch s;
  s.nt = l.nt;
  s.run = l.run;
  s.grid = l.grid;
  s.wgb = l.wgb;
  s.chunk = l.chunk;
  s.lds = l.lds;
  return s;
}
}  // namespace

Batch::Batch(int storage, int scale_exp) : storage_(storage), scale_exp_(scale_exp), dev_(new glsm::BatchDevice) {
  if (storage < Fluid::F32_FULL || storage > Fluid::BF16_DEV) throw std::invalid_argument("Batch: unknown storage");
}

Batch::~Batch() {
  (void)hipStreamSynchronize(glsm::stream());
  dropGraphs();
  for (void* e : timing_events_) (void)hipEventDestroy((hipEvent_t)e);
}

void Batch::dropGraphs() {
  for (auto& kv : graphs_) (void)hipGraphExecDestroy((hipGraphExec_t)kv.second);
  graphs_.clear();
}

int Batch::add(std::vector<Gel*>& gels, Coupler* coupler, Fluid* fluid, int iter, bool write_files,
               const std::string& out_dir) {
  if (!coupler || !fluid) throw std::invalid_argument("Batch::add: null coupler or fluid");
  if (&coupler->gels != &gels) throw std::invalid_argument("Batch::add: gels is not the coupler's vector");
  if (fluid->coupler != coupler) throw std::invalid_argument("Batch::add: fluid belongs to another coupler");
  if (iter < 0) throw std::invalid_argument("Batch::add: negative iteration");
  for (const Member& m : members_)
    if (m.fluid == fluid || m.coupler == coupler) throw std::invalid_argument("Batch::add: instance already a member");
  if (nsub_ == 0) nsub_ = fluid->Nsub;
  if (fluid->Nsub != nsub_) throw std::invalid_argument("Batch::add: members must share the sub-step count");
  // the batched stream/collide kernel addresses a member's distributions by
  // 32-bit byte offsets
  if (!glsm::dist_fits_b32(fluid->plane, glsm::storage_bytes(storage_)))
    throw std::invalid_argument("Batch::add: lattice too large for a batch member (distributions over 4 GiB)");
  if (fluid->storage() != storage_ || (storage_ == Fluid::F16_DEV && fluid->storageScaleExp() != scale_exp_))
    fluid->setStorage(storage_, scale_exp_);
  GLSM_HIP(hipStreamSynchronize(glsm::stream()));
  members_.push_back(Member{next_id_, &gels, coupler, fluid, iter, write_files, out_dir, {nullptr, nullptr},
                            {nullptr, nullptr}});
  dirty_ = true;
  return next_id_++;
}

void Batch::remove(int id) {
  auto it = std::find_if(members_.begin(), members_.end(), [&](const Member& m) { return m.id == id; });
  if (it == members_.end()) throw std::invalid_argument("Batch::remove: no member " + std::to_string(id));
  GLSM_HIP(hipStreamSynchronize(glsm::stream()));
  members_.erase(it);
  dirty_ = true;
}

std::vector<int> Batch::ids() const {
  std::vector<int> v;
  for (const Member& m : members_) v.push_back(m.id);
  return v;
}

int Batch::iteration(int id) const {
  for (const Member& m : members_)
    if (m.id == id) return m.iter;
  throw std::invalid_argument("Batch::iteration: no member " + std::to_string(id));
}

long Batch::cells() const {
  long n = 0;
  for (const Member& m : members_) n += m.fluid->N;
  return n;
}

void Batch::useGraph(bool on) { use_graph_ = on; }

Batch::Launch Batch::launch() const { return launch_; }

void Batch::setLaunch(const Launch& l) {
  if (l.run != 0 && l.run != 4 && l.run != 16 && l.run != 64)
    throw std::invalid_argument("Batch::setLaunch: run must be 0, 4, 16 or 64");
  if (l.grid < 0) throw std::invalid_argument("Batch::setLaunch: negative grid");
  if (l.wgb != 1 && l.wgb != 2 && l.wgb != 4) throw std::invalid_argument("Batch::setLaunch: wgb must be 1, 2 or 4");
  if (l.chunk < 1 || l.chunk > 64) throw std::invalid_argument("Batch::setLaunch: chunk must be in [1, 64]");
  if ((l.wgb != 1 || l.chunk != 1) && (l.grid != 0 || l.run == 64))
    throw std::invalid_argument("Batch::setLaunch: work units of several blocks take grid 0 and run 0, 4 or 16");
  if (l.lds < 0 || l.lds > 65536) throw std::invalid_argument("Batch::setLaunch: lds must be in [0, 65536]");
  if (l.lds != 0 && (l.wgb != 1 || l.chunk != 1))
    throw std::invalid_argument("Batch::setLaunch: an LDS reserve takes one block per work-group");
  launch_ = l;
  launch_set_ = true;
  dropGraphs();
}

// The parameter values a member's device descriptors copy from its objects.
std::vector<char> Batch::memberParams(const Member& m) const {
  std::vector<char> out;
  auto put = [&](const void* p, size_t n) {
    const char* c = static_cast<const char*>(p);
    out.insert(out.end(), c, c + n);
  };
  const Fluid& f = *m.fluid;
  const glsm::LbmParams p = glsm::fluid_lbm_params(f);
  const glsm::ScalarParams sp = glsm::scalar_params(f.h_fp->dt, f.h_fp->D, f.h_fp->dx, f.h_fp->dy, f.h_fp->dz);
  const float vals[6] = {m.coupler->h_cp->gel_to_lbm_vel, f.h_fp->beta, f.ramp_steps, m.coupler->h_cp->delta,
                         m.coupler->h_cp->gamma, 0.f};
  put(&p, sizeof p);
  put(&sp, sizeof sp);
  put(vals, sizeof vals);
  for (Gel* g : *m.gels) {
    const glsm::GelView v = glsm::gel_view(*g);
    put(&v, sizeof v);
  }
  return out;
}

// Device tables from the members' objects; the members' current buffers
// are bound to parity (pf_, pc_).
void Batch::rebuild() {
  hipStream_t s = glsm::stream();
  GLSM_HIP(hipStreamSynchronize(s));
  dropGraphs();
  std::vector<glsm::BatchInst> inst;
  std::vector<int> blk_inst, mk_inst, iters, gel_inst;
  std::vector<glsm::GelView> views;
  int blk = 0, mk = 0, maxL = 0;
  size_t fbytes = 0, maxN = 0;
  for (size_t i = 0; i < members_.size(); ++i) {
    Member& m = members_[i];
    Fluid& f = *m.fluid;
    if (f.storage() != storage_ || (storage_ == Fluid::F16_DEV && f.storageScaleExp() != scale_exp_))
      throw std::invalid_argument("Batch: a member's storage changed; remove it and add it again");
    m.ftab[pf_] = f.d_f;
    m.ftab[pf_ ^ 1] = f.d_fnext;
    m.ctab[pc_] = f.d_c1;
    m.ctab[pc_ ^ 1] = f.d_c2;
    m.params = memberParams(m);
    glsm::BatchInst b{};
    b.g = glsm::fluid_geom(f);
    b.p = glsm::fluid_lbm_params(f);
    b.sp = glsm::scalar_params(f.h_fp->dt, f.h_fp->D, f.h_fp->dx, f.h_fp->dy, f.h_fp->dz);
    b.f[pf_] = f.d_f;
    b.f[pf_ ^ 1] = f.d_fnext;
    b.u = f.d_u;
    b.c[pc_] = f.d_c1;
    b.c[pc_ ^ 1] = f.d_c2;
    b.sl = f.spreadLists();
    b.m = glsm::markers_with_gels(*m.coupler);
    b.g2l = m.coupler->h_cp->gel_to_lbm_vel;
    b.beta = f.h_fp->beta;
    b.ramp = f.ramp_steps;
    b.delta = m.coupler->h_cp->delta;
    b.gamma = m.coupler->h_cp->gamma;
    b.const_force = (b.p.Fc.x != 0.f || b.p.Fc.y != 0.f || b.p.Fc.z != 0.f) ? 1 : 0;
    b.blk0 = blk;
    b.mk0 = mk;
    inst.push_back(b);
    blk_inst.insert(blk_inst.end(), b.sl.nblk, (int)i);
    mk_inst.insert(mk_inst.end(), b.m.M, (int)i);
    blk += b.sl.nblk;
    mk += b.m.M;
    iters.push_back(m.iter);
    for (Gel* g : *m.gels) {
      views.push_back(glsm::gel_view(*g));
      gel_inst.push_back((int)i);
      maxL = std::max(maxL, std::max(g->m_gelNodeGrid.x, std::max(g->m_gelNodeGrid.y, g->m_gelNodeGrid.z)));
    }
    fbytes += 2 * 19 * f.plane * glsm::storage_bytes(storage_);
    maxN = std::max(maxN, (size_t)f.N);
  }
  glsm::BatchDevice& d = *dev_;
  d.inst.assign(inst, s);
  d.blk_inst.assign(blk_inst, s);
  d.mk_inst.assign(mk_inst, s);
  d.iters.assign(iters, s);
  d.views.assign(views, s);
  d.gel_inst.assign(gel_inst, s);
  d.ngel = (int)views.size();
  d.max_L = maxL;
  d.tables.inst = d.inst.p;
  d.tables.blk_inst = d.blk_inst.p;
  d.tables.mk_inst = d.mk_inst.p;
  d.tables.iters = d.iters.p;
  d.tables.nblk = blk;
  d.tables.M = mk;
  d.tables.ninst = (int)inst.size();
  // the stream/collide tuning copy writes u to scratch instead of the state
  if (d.scratch_u.n < maxN) {
    glsm::dfree(d.scratch_u.p);
    d.scratch_u.p = glsm::dalloc<float3>(maxN);
    d.scratch_u.n = maxN;
  }
  for (auto& b : inst) b.u = d.scratch_u.p;
  d.tune_inst.assign(inst, s);
  if (!launch_set_) {
    // same rule as the single-lattice path: stream the stores past the
    // 256 MB Infinity Cache once the state no longer fits in it
    launch_.nt = fbytes > (size_t)192 << 20 ? 1 : 0;
    launch_.run = launch_.nt ? 16 : 0;
    launch_.grid = 0;
    launch_.wgb = 1;
    launch_.chunk = 1;
    // 4-byte storage streamed from HBM: 5 work-groups per CU (measured on
    // the 64-instance FP32 batch; the autotuner also times 4 and 3)
    launch_.lds = launch_.nt && glsm::storage_bytes(storage_) == 4 ? 32 * 1024 : 0;
  }
  GLSM_HIP(hipStreamSynchronize(s));
  dirty_ = false;
  lists_fresh_ = false;
}

// Re-bin this step when the tables are new, some member's gels move in
// this step's elasticity (iteration multiple of 5), or some gel moved
// outside the batch since its last step.
bool Batch::listsNeeded() const {
  if (!lists_fresh_) return true;
  size_t k = 0;
  for (const Member& m : members_) {
    if (m.iter % 5 == 0) return true;
    for (Gel* g : *m.gels)
      if (k >= seen_pos_.size() || g->positionVersion() != seen_pos_[k++]) return true;
  }
  return false;
}

void Batch::enqueueStep(bool timed, bool lists) {
  hipStream_t s = glsm::stream();
  glsm::BatchDevice& d = *dev_;
  const glsm::BatchTables& t = d.tables;
  glsm::launch_gel_step_batch(d.views.p, d.gel_inst.p, d.iters.p, d.ngel, d.max_L, s);
  glsm::launch_pack_batch(t, s);
  if (lists) glsm::launch_spread_rebuild_batch(t, s);
  const glsm::LbmLaunch shape = to_internal(launch_);
  for (int k = 0; k < nsub_; ++k) {
    glsm::launch_ibm_interp_batch(t, s);
    const int sel = (pf_ + k) & 1;
    if (timed && k == nsub_ / 2) {
      hipEvent_t e[2];
      for (auto& x : e) x = glsm::timing_event();
      glsm::launch_lbm_step_batch(storage_, t, sel, k == nsub_ - 1, shape, s, e[0], e[1]);
      timing_events_.push_back(e[0]);
      timing_events_.push_back(e[1]);
    } else {
      glsm::launch_lbm_step_batch(storage_, t, sel, k == nsub_ - 1, shape, s);
    }
  }
  glsm::launch_ibm_sample_batch(t, pc_, s);
  glsm::launch_scalar_batch(t, pc_, s);
  glsm::launch_repulsion_batch(t, s);
  glsm::launch_scatter_batch(t, s);
  glsm::launch_advance_iters(d.iters.p, t.ninst, s);
}

void Batch::step(int nsteps) {
  if (nsteps < 0) throw std::invalid_argument("Batch::step: negative step count");
  if (members_.empty()) return;
  for (const Member& m : members_)
    if (m.fluid->d_f != m.ftab[pf_] || m.fluid->d_c1 != m.ctab[pc_] || memberParams(m) != m.params) dirty_ = true;
  if (dirty_) rebuild();
  hipStream_t s = glsm::stream();
  for (int n = 0; n < nsteps; ++n) {
    const bool timed = timing_every_ > 0 && (timing_calls_++ % timing_every_) == 0 && timing_events_.size() < 8192;
    const bool lists = listsNeeded();
    if (use_graph_ && !timed) {
      const int key = pf_ * 4 + pc_ * 2 + (lists ? 1 : 0);
      auto it = graphs_.find(key);
      if (it == graphs_.end()) {
        hipGraph_t g;
        GLSM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          enqueueStep(false, lists);
        } catch (...) {
          (void)hipStreamEndCapture(s, &g);
          throw;
        }
        GLSM_HIP(hipStreamEndCapture(s, &g));
        hipGraphExec_t ge;
        GLSM_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        it = graphs_.emplace(key, (void*)ge).first;
      }
      GLSM_HIP(hipGraphLaunch((hipGraphExec_t)it->second, s));
    } else {
      enqueueStep(timed, lists);
    }
    lists_fresh_ = true;
    // the objects follow the device state: buffers swapped as their own
    // step methods would have, positions marked as changed
    for (Member& m : members_) {
      Fluid& f = *m.fluid;
      if (nsub_ & 1) std::swap(f.d_f, f.d_fnext);
      std::swap(f.d_c1, f.d_c2);
      for (Gel* g : *m.gels) g->markPositionsChanged();
    }
    seen_pos_.clear();
    for (Member& m : members_)
      for (Gel* g : *m.gels) seen_pos_.push_back(g->positionVersion());
    if (nsub_ & 1) pf_ ^= 1;
    pc_ ^= 1;
    for (Member& m : members_) {
      const int it = m.iter++;
      if (!m.write_files) continue;
      if (!output_) {
        for (Gel* g : *m.gels) g->writeFiles(it);
        m.fluid->writeFiles(it);
        continue;
      }
      // the objects' cadence and output times (gel.cu:543-553, fluid.cu:140-150)
      for (Gel* g : *m.gels)
        if (it % 1000 == 0) output_->gel(g, int(it * g->m_dt), m.out_dir);
      if (it % 2000 == 0) output_->fluid(m.fluid, int((double)(it * m.fluid->dt)), m.out_dir);
    }
  }
}

Batch::Launch Batch::autotune(int reps, double* best_ms) {
  if (reps < 1) throw std::invalid_argument("Batch::autotune: reps must be >= 1");
  if (members_.empty()) throw std::invalid_argument("Batch::autotune: empty batch");
  if (dirty_) rebuild();
  hipStream_t s = glsm::stream();
  glsm::BatchTables t = dev_->tables;
  t.inst = dev_->tune_inst.p;
  int cus = 256;
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
      cus = prop.multiProcessorCount;
  }
  std::vector<Launch> cand;
  for (int nt = 0; nt < 2; ++nt)
    for (int run : {0, 4, 16, 64}) {
      for (int g : {0, 2, 4, 8, 16}) cand.push_back(Launch{nt, run, g ? g * cus : 0, 1, 1, 0});
      for (int kb : {32, 40, 52}) cand.push_back(Launch{nt, run, 0, 1, 1, kb * 1024});
    }
  for (int nt = 0; nt < 2; ++nt)
    for (int run : {0, 4, 16}) {
      for (int wgb : {2, 4}) cand.push_back(Launch{nt, run, 0, wgb, 1});
      for (int chunk : {2, 4, 8}) cand.push_back(Launch{nt, run, 0, 1, chunk});
    }
  hipEvent_t e0 = glsm::timing_event(), e1 = glsm::timing_event();
  // one trial: a warm-up launch, then `reps` timed launches
  auto trial = [&](const Launch& c) {
    const glsm::LbmLaunch sh = to_internal(c);
    glsm::launch_lbm_step_batch(storage_, t, pf_, false, sh, s);
    GLSM_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) glsm::launch_lbm_step_batch(storage_, t, pf_, false, sh, s);
    GLSM_HIP(hipEventRecord(e1, s));
    GLSM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GLSM_HIP(hipEventElapsedTime(&ms, e0, e1));
    return (double)ms / reps;
  };
  // Each candidate's time is the median of three trials, so that one
  // mistimed trial, too fast or too slow, cannot decide the pick.
  double best = 1e30;
  Launch pick = launch_;
  for (const Launch& c : cand) {
    double tr[3];
    for (double& x : tr) x = trial(c);
    std::sort(tr, tr + 3);
    if (tr[1] < best) {
      best = tr[1];
      pick = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  setLaunch(pick);
  if (best_ms) *best_ms = best;
  return pick;
}

void Batch::enableLbmTiming(int every) {
  if (every < 0) throw std::invalid_argument("Batch::enableLbmTiming: negative period");
  timing_every_ = every;
  timing_calls_ = 0;
  for (void* e : timing_events_) (void)hipEventDestroy((hipEvent_t)e);
  timing_events_.clear();
}

double Batch::lbmKernelTiming(int* samples) {
  GLSM_HIP(hipStreamSynchronize(glsm::stream()));
  double sum = 0;
  int n = 0;
  for (size_t i = 0; i + 1 < timing_events_.size(); i += 2) {
    float ms = 0;
    GLSM_HIP(hipEventElapsedTime(&ms, (hipEvent_t)timing_events_[i], (hipEvent_t)timing_events_[i + 1]));
    sum += ms;
    ++n;
  }
  if (samples) *samples = n;
  return n ? sum / n : 0.0;
}
