// output.cpp -- asynchronous .dat output (see include/output.h).
//
// Per object a slot keeps a device staging buffer, a pinned host buffer and
// two events.  A request:
//   1. library stream: device-to-device copies of the output fields into
//      the staging buffer, then event `snap` (the instance may step on
//      right after: the snapshot is taken in stream order);
//   2. copy stream: waits for `snap`, copies staging -> pinned host, event
//      `landed`;
//   3. writer thread: waits for `landed`, formats the reference's files.
// A second request for the same object waits until the first was written.
#include "output.h"

#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "fluid.h"
#include "gel.h"
#include "glsm_internal.hpp"

namespace glsm {

struct OutSlot {
  bool is_gel = false;
  size_t bytes = 0;
  char* dev = nullptr;
  char* host = nullptr;
  hipEvent_t snap = nullptr, landed = nullptr;
  bool busy = false;
  ~OutSlot() {
    dfree(dev);
    if (host) (void)hipHostFree(host);
    if (snap) (void)hipEventDestroy(snap);
    if (landed) (void)hipEventDestroy(landed);
  }
};

struct OutRequest {
  OutSlot* slot;
  int time;
  std::string dir;
  int dims[3];
  int gel_id;
  bool body;  // append the centre ring to gel<id>bodycenter.dat
};

struct OutputState {
  hipStream_t copy = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  std::map<const void*, std::unique_ptr<OutSlot>> slots;
  std::deque<OutRequest> queue;
  long pending = 0, done = 0;
  std::string err;
  bool stop = false;
  std::vector<std::thread> writers;
};

namespace {

// gel staging layout (doubles): rn, Fn, Veln (3 nN each), rm (3 nE),
// um, vm, wm (nE each), then the centre ring: rm, Fn, Veln (3000 each),
// vm, wm (1000 each)
size_t gel_doubles(const Gel& g) {
  const size_t nN = g.m_numGelNodes, nE = g.m_numGelElements;
  return 9 * nN + 6 * nE + 11000;
}

void body_center(const std::string& dir, int gel_id, int time, const double* c) {
  const double3* rm = reinterpret_cast<const double3*>(c);
  const double3* Fn = rm + 1000;
  const double3* Vn = Fn + 1000;
  const double* vm = reinterpret_cast<const double*>(Vn + 1000);
  const double* wm = vm + 1000;
  const std::string name = "gel" + std::to_string(gel_id) + "bodycenter.dat";
  const std::string path = dir.empty() || dir == "." ? name : dir + "/" + name;
  FILE* fp = std::fopen(path.c_str(), "a");
  if (!fp) throw std::runtime_error("cannot write " + path);
  // gel.cu:528-537: setw(9) << to_string(...) columns joined by six spaces
  auto col = [&](double v, bool last) {
    const std::string s = std::to_string(v);
    std::fprintf(fp, "%9s%s", s.c_str(), last ? "\n" : "      ");
  };
  for (int i = 0; i < 1000; ++i) {
    const std::string t = std::to_string(time - 1000 + i);
    std::fprintf(fp, "%9s      ", t.c_str());
    col(rm[i].x, false); col(rm[i].y, false); col(rm[i].z, false);
    col(Fn[i].x, false); col(Fn[i].y, false); col(Fn[i].z, false);
    col(Vn[i].x, false); col(Vn[i].y, false); col(Vn[i].z, false);
    col(vm[i], false); col(wm[i], true);
  }
  std::fclose(fp);
}

void write_request(const OutRequest& r) {
  GLSM_HIP(hipEventSynchronize(r.slot->landed));
  if (r.slot->is_gel) {
    const int LX = r.dims[0], LY = r.dims[1], LZ = r.dims[2];
    const size_t nN = (size_t)(LX + 2) * (LY + 2) * (LZ + 2), nE = (size_t)(LX + 1) * (LY + 1) * (LZ + 1);
    const double* d = reinterpret_cast<const double*>(r.slot->host);
    const double3* rn = reinterpret_cast<const double3*>(d);
    const double3* Fn = rn + nN;
    const double3* Vn = Fn + nN;
    const double3* rm = Vn + nN;
    const double* um = reinterpret_cast<const double*>(rm + nE);
    const double* vm = um + nE;
    const double* wm = vm + nE;
    write_gel_dat(r.dir, r.gel_id, r.time, LX, LY, LZ, rn, Fn, Vn, rm, um, vm, wm);
    if (r.body) body_center(r.dir, r.gel_id, r.time, wm + nE);
  } else {
    const size_t N = (size_t)r.dims[0] * r.dims[1] * r.dims[2];
    const float3* u = reinterpret_cast<const float3*>(r.slot->host);
    const float* c = reinterpret_cast<const float*>(u + N);
    write_fluid_dat(r.dir, r.time, r.dims[0], r.dims[1], r.dims[2], u, c);
  }
}

void writer_loop(OutputState* st) {
  for (;;) {
    OutRequest r;
    {
      std::unique_lock<std::mutex> lk(st->mu);
      st->cv.wait(lk, [&] { return st->stop || !st->queue.empty(); });
      if (st->queue.empty()) return;  // stop requested and drained
      r = st->queue.front();
      st->queue.pop_front();
    }
    std::string err;
    try {
      write_request(r);
    } catch (const std::exception& e) {
      err = e.what();
    }
    std::lock_guard<std::mutex> lk(st->mu);
    if (!err.empty()) st->err = err;
    r.slot->busy = false;
    --st->pending;
    ++st->done;
    st->cv.notify_all();
  }
}

OutSlot* acquire(OutputState& st, const void* obj, bool is_gel, size_t bytes) {
  std::unique_lock<std::mutex> lk(st.mu);
  auto& sp = st.slots[obj];
  if (sp) st.cv.wait(lk, [&] { return !sp->busy; });
  if (!sp || sp->bytes != bytes) {
    sp.reset(new OutSlot);
    sp->is_gel = is_gel;
    sp->bytes = bytes;
    sp->dev = dalloc<char>(bytes);
    GLSM_HIP(hipHostMalloc((void**)&sp->host, bytes, hipHostMallocDefault));
    GLSM_HIP(hipEventCreateWithFlags(&sp->snap, hipEventDisableTiming));
    GLSM_HIP(hipEventCreateWithFlags(&sp->landed, hipEventDisableTiming));
  }
  sp->busy = true;
  ++st.pending;
  return sp.get();
}

void submit(OutputState& st, const OutRequest& r) {
  hipStream_t s = stream();
  GLSM_HIP(hipEventRecord(r.slot->snap, s));
  GLSM_HIP(hipStreamWaitEvent(st.copy, r.slot->snap, 0));
  GLSM_HIP(hipMemcpyAsync(r.slot->host, r.slot->dev, r.slot->bytes, hipMemcpyDeviceToHost, st.copy));
  GLSM_HIP(hipEventRecord(r.slot->landed, st.copy));
  std::lock_guard<std::mutex> lk(st.mu);
  st.queue.push_back(r);
  st.cv.notify_all();
}

void d2d(char*& dst, const void* src, size_t bytes) {
  GLSM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
  dst += bytes;
}

}  // namespace
}  // namespace glsm

OutputStream::OutputStream(int writers) : st_(new glsm::OutputState) {
  if (writers < 1 || writers > 64) throw std::invalid_argument("OutputStream: writers must be in [1, 64]");
  GLSM_HIP(hipStreamCreate(&st_->copy));  // default priority
  for (int i = 0; i < writers; ++i) st_->writers.emplace_back(glsm::writer_loop, st_.get());
}

OutputStream::~OutputStream() {
  {
    std::lock_guard<std::mutex> lk(st_->mu);
    st_->stop = true;
    st_->cv.notify_all();
  }
  for (auto& w : st_->writers)
    if (w.joinable()) w.join();
  (void)hipStreamSynchronize(st_->copy);
  st_->slots.clear();
  (void)hipStreamDestroy(st_->copy);
}

// gel.cu:452-553: the files of one output time, plus the centre ring on
// every 1000th output time after the first call (the reference's `flag`)
void OutputStream::gel(Gel* g, int time, const std::string& dir) {
  if (!g) throw std::invalid_argument("OutputStream::gel: null gel");
  const size_t nN = g->m_numGelNodes, nE = g->m_numGelElements;
  glsm::OutSlot* slot = glsm::acquire(*st_, g, true, glsm::gel_doubles(*g) * sizeof(double));
  char* p = slot->dev;
  glsm::d2d(p, g->m_drn, nN * sizeof(double3));
  glsm::d2d(p, g->m_dFn, nN * sizeof(double3));
  glsm::d2d(p, g->m_dVeln, nN * sizeof(double3));
  glsm::d2d(p, g->m_drm, nE * sizeof(double3));
  glsm::d2d(p, g->m_dum, nE * sizeof(double));
  glsm::d2d(p, g->m_dvm, nE * sizeof(double));
  glsm::d2d(p, g->m_dwm, nE * sizeof(double));
  glsm::d2d(p, g->m_drm_center, 1000 * sizeof(double3));
  glsm::d2d(p, g->m_dFn_center, 1000 * sizeof(double3));
  glsm::d2d(p, g->m_dVeln_center, 1000 * sizeof(double3));
  glsm::d2d(p, g->m_dvm_center, 1000 * sizeof(double));
  glsm::d2d(p, g->m_dwm_center, 1000 * sizeof(double));
  glsm::OutRequest r{slot, time, dir, {g->m_gelNodeGrid.x, g->m_gelNodeGrid.y, g->m_gelNodeGrid.z}, g->m_gelId,
                     g->flag && time % 1000 == 0};
  if (!r.body) g->flag = true;
  glsm::submit(*st_, r);
}

// fluid.cu:115-150
void OutputStream::fluid(Fluid* f, int time, const std::string& dir) {
  if (!f) throw std::invalid_argument("OutputStream::fluid: null fluid");
  const size_t N = f->N;
  glsm::OutSlot* slot = glsm::acquire(*st_, f, false, N * (sizeof(float3) + sizeof(float)));
  char* p = slot->dev;
  glsm::d2d(p, f->d_u, N * sizeof(float3));
  glsm::d2d(p, f->d_c1, N * sizeof(float));
  glsm::OutRequest r{slot, time, dir, {f->fluidNodeGrid.x, f->fluidNodeGrid.y, f->fluidNodeGrid.z}, 0, false};
  glsm::submit(*st_, r);
}

void OutputStream::release(const void* object) {
  std::unique_lock<std::mutex> lk(st_->mu);
  auto it = st_->slots.find(object);
  if (it == st_->slots.end()) return;
  glsm::OutSlot* s = it->second.get();
  st_->cv.wait(lk, [&] { return !s->busy; });
  st_->slots.erase(it);
}

void OutputStream::flush() {
  std::unique_lock<std::mutex> lk(st_->mu);
  st_->cv.wait(lk, [&] { return st_->pending == 0; });
}

long OutputStream::pending() const {
  std::lock_guard<std::mutex> lk(st_->mu);
  return st_->pending;
}

long OutputStream::completed() const {
  std::lock_guard<std::mutex> lk(st_->mu);
  return st_->done;
}

std::string OutputStream::lastError() {
  std::lock_guard<std::mutex> lk(st_->mu);
  return st_->err;
}
