// fluid.cpp -- host side of the Fluid class (reference: src/fluid.cu).
#include "fluid.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "gel.h"
#include "glsm_internal.hpp"

using namespace std;

namespace {

glsm::LatticeGeom geom(const Fluid& f) {
  glsm::LatticeGeom g;
  g.Nx = f.fluidNodeGrid.x;
  g.Ny = f.fluidNodeGrid.y;
  g.Nz = f.fluidNodeGrid.z;
  g.N = f.N;
  g.perX = f.h_fp->perX;
  g.perY = f.h_fp->perY;
  g.perZ = f.h_fp->perZ;
  g.zlo = 0;
  g.zhi = g.Nz - 1;
  g.h = f.h_fp->h;
  g.rNx = 1.0 / g.Nx;
  g.rNxNy = 1.0 / ((double)g.Nx * g.Ny);
  return g;
}

glsm::LbmParams lbm_params(const FluidParams& fp, int storage, int scale_exp) {
  glsm::LbmParams p;
  p.omega = 1.f / fp.tau;
  p.w0 = fp.w[0];
  p.w1 = fp.w[1];
  p.w2 = fp.w[7];
  p.Fc = fp.F_const;
  p.scale = storage == Fluid::F16_DEV ? ldexpf(1.0f, scale_exp) : 1.0f;
  p.inv_scale = 1.0f / p.scale;
  return p;
}

int storage_from_env(int dflt) {
  const char* e = getenv("GLSM_FLUID_STORAGE");
  if (!e) return dflt;
  const string s = e;
  if (s == "f32") return Fluid::F32_DEV;
  if (s == "f32full") return Fluid::F32_FULL;
  if (s == "f16") return Fluid::F16_DEV;
  if (s == "bf16") return Fluid::BF16_DEV;
  throw std::invalid_argument("GLSM_FLUID_STORAGE must be f32, f32full, f16 or bf16");
}

}  // namespace

namespace glsm {
LatticeGeom fluid_geom(const Fluid& f) { return geom(f); }
LbmParams fluid_lbm_params(const Fluid& f) { return lbm_params(*f.h_fp, f.storage(), f.storageScaleExp()); }
}  // namespace glsm

int Fluid::idx3(int x, int y, int z, int Nx, int Ny) { return x + y * Nx + z * Nx * Ny; }

// fluid.cu:152-232
Fluid::Fluid(int3 size, int time, Coupler* cp)
    : h_u(nullptr), h_c1(nullptr), h_fp(nullptr), d_f(nullptr), d_fnext(nullptr), d_u(nullptr), d_c1(nullptr),
      d_c2(nullptr), coupler(cp), fluidSize(size) {
  if (!coupler) throw std::invalid_argument("Fluid: coupler is null");
  dt = 1e-3f;
  h = 0.5f;
  fluidNodeGrid.x = int(fluidSize.x / h) + 1;
  fluidNodeGrid.y = int(fluidSize.y / h) + 1;
  fluidNodeGrid.z = int(fluidSize.z / h) + 1;
  if (fluidNodeGrid.x < 2 || fluidNodeGrid.y < 2 || fluidNodeGrid.z < 2)
    throw std::invalid_argument("Fluid: lattice needs at least 2 nodes per axis");
  const long long Nll = (long long)fluidNodeGrid.x * fluidNodeGrid.y * fluidNodeGrid.z;
  plane = glsm::dist_cells((size_t)Nll);
  if (19ull * plane >= (1ull << 32))
    throw std::invalid_argument("Fluid: lattice too large for 32-bit distribution indexing");
  N = (int)Nll;
  Nd = 19 * plane;
  threads = coupler->threads;
  blocksN = (N + threads - 1) / threads;
  blocksM = coupler->blocksM;
  h_fp = new FluidParams;
  memset(h_fp, 0, sizeof(FluidParams));
  h_fp->dt = dt;
  h_fp->perX = 0;
  h_fp->perY = 0;
  h_fp->perZ = 0;
  h_fp->cs2 = 1.f / 3.f;
  h_fp->tau = 0.8f;
  h_fp->nu = h_fp->cs2 * (h_fp->tau - 0.5f);
  h_fp->N = N;
  h_fp->M = coupler->sumGelBoundaryCount;
  h_fp->L = fluidNodeGrid;
  h_fp->h = h;
  h_fp->dx = 1;
  h_fp->dy = 1;
  h_fp->dz = 1;
  h_fp->D = 1;
  niu = 1e-6f;
  dx_fluid = 40e-6f * h_fp->h;
  dt_fluid = (h_fp->tau - 0.5f) * dx_fluid * dx_fluid / (3 * niu);
  Nsub = std::max(1, (int)llround(dt / dt_fluid));
  h_fp->beta = 1.0f;
  if (const char* beta_env = getenv("IBM_BETA")) h_fp->beta = (float)atof(beta_env);
  ramp_steps = 20000.0f;
  if (const char* ramp_env = getenv("IBM_RAMP_STEPS")) ramp_steps = fmaxf(0.0f, (float)atof(ramp_env));
  static const int c[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},
                               {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {-1, 1, 0},  {-1, -1, 0},
                               {1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {-1, 0, -1}, {1, 0, -1},
                               {0, 1, 1},  {0, 1, -1}, {0, -1, -1}, {0, -1, 1}};  // fluid.cu:202-208
  static const int opp[19] = {0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16};
  for (int q = 0; q < 19; ++q) {
    h_fp->c[q] = make_int3(c[q][0], c[q][1], c[q][2]);
    h_fp->w[q] = q == 0 ? 1.f / 3.f : (q < 7 ? 1.f / 18.f : 1.f / 36.f);
    h_fp->opp[q] = opp[q];
  }
  h_fp->F_const = make_float3(0, 0, 0);

  // unit conversion pushed into the coupler and every gel (fluid.cu:221-229)
  const float gel_to_lbm_vel = h_fp->h * dt_fluid / dt;
  const float lbm_to_gel_vel = 1.0f / gel_to_lbm_vel;
  coupler->h_cp->gel_to_lbm_vel = gel_to_lbm_vel;
  coupler->h_cp->lbm_to_gel_vel = lbm_to_gel_vel;
  fluid_stream = glsm::stream();
  GLSM_HIP(hipMemcpyAsync(coupler->d_cp, coupler->h_cp, sizeof(CouplerParams), hipMemcpyHostToDevice, fluid_stream));
  for (auto g : coupler->gels) {
    g->m_hgp->lbm_to_gel_vel = lbm_to_gel_vel;
    GLSM_HIP(hipMemcpyAsync(g->m_dgp, g->m_hgp, sizeof(GelParams), hipMemcpyHostToDevice, fluid_stream));
  }
  GLSM_HIP(hipStreamSynchronize(fluid_stream));

  // storage
  GLSM_HIP(hipHostMalloc((void**)&h_u, N * sizeof(float3), hipHostMallocPortable));
  memset(h_u, 0, N * sizeof(float3));
  GLSM_HIP(hipHostMalloc((void**)&h_c1, N * sizeof(float), hipHostMallocPortable));
  memset(h_c1, 0, N * sizeof(float));
  storage_ = storage_from_env(F32_DEV);
  d_f = glsm::dalloc<char>(Nd * glsm::storage_bytes(storage_));
  d_fnext = glsm::dalloc<char>(Nd * glsm::storage_bytes(storage_));
  d_u = glsm::dalloc<float3>(N);
  d_c1 = glsm::dalloc<float>(N);
  d_c2 = glsm::dalloc<float>(N);
  nblk_ = (N + 255) / 256;
  d_blkflag_ = glsm::dalloc<int>(nblk_);
  d_cnt_ = glsm::dalloc<int>(N);
  d_fillpos_ = glsm::dalloc<int>(N);
  d_rng_ = glsm::dalloc<int2>(N);
  d_bsum_ = glsm::dalloc<int>(nblk_ + 1);
  d_red_ = glsm::dalloc<double>(4ull * nblk_ + 4);
  ent_cap_ = 64 * std::max(1, coupler->sumGelBoundaryCount);
  d_ent_ = glsm::dalloc<SpreadEntry>(ent_cap_);
  d_ctab_ = glsm::dalloc<int>(ent_cap_);
  d_wtab_ = glsm::dalloc<float>(ent_cap_);
  d_rtab_ = glsm::dalloc<float>(ent_cap_);
  d_wsum_ = glsm::dalloc<float>(std::max(1, coupler->sumGelBoundaryCount));

  glsm::launch_lbm_init(storage_, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), d_f, d_u, d_c1, d_c2,
                        fluid_stream);
  if (time) setGoonValue(time);
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
}

Fluid::~Fluid() {
  if (file_writer_thread.joinable()) file_writer_thread.join();
  (void)hipStreamSynchronize(fluid_stream);
  for (hipEvent_t e : timing_events_) (void)hipEventDestroy(e);
  (void)hipHostFree(h_u);
  (void)hipHostFree(h_c1);
  delete h_fp;
  for (void* p : {(void*)d_f, (void*)d_fnext, (void*)d_u, (void*)d_c1, (void*)d_c2, (void*)d_blkflag_, (void*)d_cnt_,
                  (void*)d_fillpos_, (void*)d_rng_, (void*)d_bsum_, (void*)d_red_, (void*)d_ent_, (void*)d_ctab_, (void*)d_wtab_,
                  (void*)d_rtab_, (void*)d_wsum_})
    glsm::dfree(p);
}

// fluid.cu:57-83: f back at rest, u and c from the text snapshots
void Fluid::setGoonValue(int time) {
  ifstream fVelb("Velb" + to_string(time) + ".dat");
  ifstream fC1("Conc" + to_string(time) + ".dat");
  if (!fVelb || !fC1) throw std::runtime_error("Fluid::setGoonValue: missing Velb/Conc" + to_string(time) + ".dat");
  for (int zi = 0; zi < fluidNodeGrid.z; zi++)
    for (int yi = 0; yi < fluidNodeGrid.y; yi++)
      for (int xi = 0; xi < fluidNodeGrid.x; xi++) {
        const int gi = idx3(xi, yi, zi, fluidNodeGrid.x, fluidNodeGrid.y);
        fVelb >> h_u[gi].x >> h_u[gi].y >> h_u[gi].z;
        fC1 >> h_c1[gi];
      }
  GLSM_HIP(hipMemcpyAsync(d_u, h_u, sizeof(float3) * N, hipMemcpyHostToDevice, fluid_stream));
  GLSM_HIP(hipMemcpyAsync(d_c1, h_c1, sizeof(float) * N, hipMemcpyHostToDevice, fluid_stream));
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
}

void Fluid::setPeriodic(int px, int py, int pz) {
  // a 4-wide delta stencil wrapped on an axis of fewer than 4 nodes would
  // put two lanes of one marker on the same cell
  const int3 n = fluidNodeGrid;
  if ((px && n.x < 4) || (py && n.y < 4) || (pz && n.z < 4))
    throw std::invalid_argument("setPeriodic: a periodic axis needs at least 4 nodes");
  h_fp->perX = px ? 1 : 0;
  h_fp->perY = py ? 1 : 0;
  h_fp->perZ = pz ? 1 : 0;
  built_lag_version_ = ~0ull;  // stencils wrap differently now
}

void Fluid::setIBM(float beta, float ramp) {
  h_fp->beta = beta;
  ramp_steps = fmaxf(0.0f, ramp);
}

// fluid.cu:265-270, FP32 on the host as there
float Fluid::betaEff(int iter) const {
  const float ramp = (ramp_steps <= 0.0f) ? 1.0f : fminf(1.0f, (iter + 1) / ramp_steps);
  return h_fp->beta * ramp;
}

glsm::SpreadLists Fluid::spreadLists() const {
  glsm::SpreadLists sl;
  sl.nblk = nblk_;
  sl.blkflag = d_blkflag_;
  sl.cnt = d_cnt_;
  sl.fillpos = d_fillpos_;
  sl.rng = d_rng_;
  sl.bsum = d_bsum_;
  sl.ent = d_ent_;
  sl.ctab = d_ctab_;
  sl.wtab = d_wtab_;
  sl.rtab = d_rtab_;
  sl.wsum = d_wsum_;
  return sl;
}

void Fluid::ensureSpreadLists() {
  if (built_lag_version_ == coupler->lagVersion()) return;
  glsm::launch_spread_rebuild(geom(*this), glsm::markers_of(*coupler), spreadLists(), fluid_stream);
  built_lag_version_ = coupler->lagVersion();
}

// fluid.cu:263-284: Nsub fused sub-steps, each = marker interpolation +
// one stream/collide pass with the spread force gathered in place
void Fluid::stepVelocity(int iter) {
  ensureSpreadLists();
  const float beta_eff = betaEff(iter);
  const glsm::LatticeGeom g = geom(*this);
  const glsm::LbmParams p = lbm_params(*h_fp, storage_, scale_exp_);
  const glsm::MarkerArrays m = glsm::markers_of(*coupler);
  const glsm::SpreadLists sl = spreadLists();
  const bool timed = timing_every_ > 0 && (timing_calls_++ % timing_every_) == 0 &&
                     (int)timing_events_.size() < 2 * timing_max_;
  for (int kk = 0; kk < Nsub; kk++) {
    glsm::launch_ibm_interp(m, sl, d_u, beta_eff, fluid_stream);
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (timed && kk == Nsub / 2) {
      for (auto& e : ev) e = glsm::timing_event();
      timing_events_.push_back(ev[0]);
      timing_events_.push_back(ev[1]);
    }
    glsm::launch_lbm_step(storage_, g, p, d_f, d_fnext, plane, d_u, kk == Nsub - 1, sl, m.Fdrag, fluid_stream, ev[0],
                          ev[1]);
    std::swap(d_f, d_fnext);
  }
}

void Fluid::enableLbmTiming(int every, int max_samples) {
  timing_every_ = std::max(0, every);
  timing_max_ = std::max(0, max_samples);
  timing_calls_ = 0;
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
  for (hipEvent_t e : timing_events_) (void)hipEventDestroy(e);
  timing_events_.clear();
}

double Fluid::lbmKernelTiming(int* samples) {
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
  double total = 0;
  const int n = (int)timing_events_.size() / 2;
  for (int k = 0; k < n; ++k) {
    float ms = 0;
    GLSM_HIP(hipEventElapsedTime(&ms, timing_events_[2 * k], timing_events_[2 * k + 1]));
    total += ms;
  }
  if (samples) *samples = n;
  return n ? total / n : 0.0;
}

void Fluid::subStep(float beta_eff, bool write_u_all) {
  ensureSpreadLists();
  const glsm::MarkerArrays m = glsm::markers_of(*coupler);
  glsm::launch_ibm_interp(m, spreadLists(), d_u, beta_eff, fluid_stream);
  glsm::launch_lbm_step(storage_, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), d_f, d_fnext, plane, d_u,
                        write_u_all, spreadLists(), m.Fdrag, fluid_stream);
  std::swap(d_f, d_fnext);
}

// The state is converted on the device: exported as FP32 pre-collision
// populations into a scratch buffer, then imported in the new storage (the
// same values a host round trip would carry, without the PCIe copies).
void Fluid::setStorage(int storage, int scale_exp) {
  if (storage < F32_FULL || storage > BF16_DEV) throw std::invalid_argument("setStorage: unknown storage mode");
  if (scale_exp < 0 || scale_exp > 30) throw std::invalid_argument("setStorage: scale_exp must be in [0, 30]");
  if (storage == storage_ && (storage != F16_DEV || scale_exp == scale_exp_)) {
    scale_exp_ = scale_exp;
    return;
  }
  float* tmp = glsm::dalloc<float>(Nd);
  glsm::launch_lbm_export(storage_, true, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), d_f, tmp, plane,
                          fluid_stream);
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
  glsm::dfree(d_f);
  glsm::dfree(d_fnext);
  d_f = d_fnext = nullptr;
  storage_ = storage;
  scale_exp_ = scale_exp;
  d_f = glsm::dalloc<char>(Nd * glsm::storage_bytes(storage_));
  d_fnext = glsm::dalloc<char>(Nd * glsm::storage_bytes(storage_));
  glsm::launch_lbm_import(storage_, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), tmp, d_f, plane,
                          fluid_stream);
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
  glsm::dfree(tmp);
}

// fluid.cu:251-261
void Fluid::stepConcentration(int /*iter*/) {
  ensureSpreadLists();
  const glsm::LatticeGeom g = geom(*this);
  const glsm::MarkerArrays m = glsm::markers_of(*coupler);
  glsm::launch_ibm_sample(m, spreadLists(), d_c1, fluid_stream);
  const glsm::ScalarParams sp = glsm::scalar_params(h_fp->dt, h_fp->D, h_fp->dx, h_fp->dy, h_fp->dz);
  glsm::launch_scalar(g, sp, d_u, d_c1, d_c2, spreadLists(), m.Sl, fluid_stream);
  std::swap(d_c1, d_c2);
}

// fluid.cu:85-90
void Fluid::copyDataToHost() {
  GLSM_HIP(hipMemcpyAsync(h_u, d_u, sizeof(float3) * N, hipMemcpyDeviceToHost, fluid_stream));
  GLSM_HIP(hipMemcpyAsync(h_c1, d_c1, sizeof(float) * N, hipMemcpyDeviceToHost, fluid_stream));
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
}

// fluid.cu:115-138
void Fluid::recordData(int time) {
  glsm::write_fluid_dat("", time, fluidNodeGrid.x, fluidNodeGrid.y, fluidNodeGrid.z, h_u, h_c1);
}

// fluid.cu:140-150 (iter * dt is a float product there)
void Fluid::writeFiles(int iter) {
  double time = iter * dt;
  if (iter % 2000 == 0) {
    if (file_writer_thread.joinable()) file_writer_thread.join();
    copyDataToHost();
    // a failed write is reported, not thrown: the reference's ofstreams fail silently
    file_writer_thread = std::thread([this, t = int(time)] {
      try {
        recordData(t);
      } catch (const std::exception& e) {
        std::fprintf(stderr, "fluid: %s\n", e.what());
      }
    });
  }
}

void Fluid::conservation(double out[4]) {
  double* dout = d_red_ + 4ull * nblk_;
  glsm::launch_lbm_moments(storage_, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), d_f, d_red_, dout,
                           fluid_stream);
  GLSM_HIP(hipMemcpyAsync(out, dout, 4 * sizeof(double), hipMemcpyDeviceToHost, fluid_stream));
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
}

long long Fluid::checkFinite(double* max_abs_dev, double* limit) {
  // d_red_ holds 4 * nblk_ + 4 >= 8 doubles: the first two serve as counters
  auto* out2 = reinterpret_cast<unsigned long long*>(d_red_);
  glsm::launch_lbm_health(storage_, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), d_f, d_u, d_c1, out2,
                          fluid_stream);
  unsigned long long h[2] = {0, 0};
  GLSM_HIP(hipMemcpyAsync(h, out2, sizeof h, hipMemcpyDeviceToHost, fluid_stream));
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
  if (max_abs_dev) {
    double m;
    std::memcpy(&m, &h[1], sizeof m);
    *max_abs_dev = m;
  }
  if (limit) {
    switch (storage_) {
      case F16_DEV: *limit = 65504.0 * std::ldexp(1.0, -scale_exp_); break;
      case BF16_DEV: *limit = 3.3895313892515355e38; break;  // largest finite bfloat16
      default: *limit = 3.4028234663852886e38;               // largest finite float
    }
  }
  return (long long)h[0];
}

void Fluid::readDensity(double* host) {
  double* tmp = glsm::dalloc<double>(N);
  glsm::launch_lbm_density(storage_, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), d_f, tmp,
                           fluid_stream);
  GLSM_HIP(hipMemcpyAsync(host, tmp, N * sizeof(double), hipMemcpyDeviceToHost, fluid_stream));
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
  glsm::dfree(tmp);
}

void Fluid::readPopulations(float* host) {
  float* tmp = glsm::dalloc<float>(Nd);
  glsm::launch_lbm_export(storage_, true, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), d_f, tmp, plane,
                          fluid_stream);
  GLSM_HIP(hipMemcpy2DAsync(host, N * sizeof(float), tmp, plane * sizeof(float), N * sizeof(float), 19,
                            hipMemcpyDeviceToHost, fluid_stream));
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
  glsm::dfree(tmp);
}

void Fluid::writePopulations(const float* host) {
  float* tmp = glsm::dalloc<float>(Nd);
  GLSM_HIP(hipMemcpy2DAsync(tmp, plane * sizeof(float), host, N * sizeof(float), N * sizeof(float), 19,
                            hipMemcpyHostToDevice, fluid_stream));
  glsm::launch_lbm_import(storage_, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), tmp, d_f, plane,
                          fluid_stream);
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
  glsm::dfree(tmp);
}

void Fluid::readPostCollision(float* host) {
  float* tmp = glsm::dalloc<float>(Nd);
  glsm::launch_lbm_export(storage_, false, geom(*this), lbm_params(*h_fp, storage_, scale_exp_), d_f, tmp, plane,
                          fluid_stream);
  GLSM_HIP(hipMemcpy2DAsync(host, N * sizeof(float), tmp, plane * sizeof(float), N * sizeof(float), 19,
                            hipMemcpyDeviceToHost, fluid_stream));
  GLSM_HIP(hipStreamSynchronize(fluid_stream));
  glsm::dfree(tmp);
}
