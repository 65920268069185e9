// fluid.h -- D3Q19 BGK lattice Boltzmann fluid (Guo forcing, halfway
// bounce-back walls) with a passive concentration scalar and immersed-
// boundary coupling, on one MI355X.  Drop-in for the reference's
// src/fluid.h:1-84: same constructor, step methods, output, and the members
// its drivers read (Nsub, h_fp->h).
//
// State layout in HBM (differs from the reference on purpose):
//   * distributions: two ping-pong buffers in blocks of 256 cells, each
//     block holding the 19 populations as 19 runs of 256 values (see d_f
//     below for the index formula); they hold POST-collision
//     populations (pull scheme: one fused stream+collide pass per sub-step),
//     by default as FP32 deviations f - w_q (see Storage below);
//   * d_u (float3, x fastest), d_c1/d_c2 (float) as in the reference;
//   * no F / F_ibm / F_tot / rho fields: the immersed-boundary force is
//     gathered per cell inside the fused kernel from marker forces, using
//     per-cell contribution lists rebuilt when the markers move.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <thread>
#include <vector>

#include "coupling.h"

// fluid.h:7-22
struct FluidParams {
  int3 c[19];
  float w[19];
  int opp[19];
  float tau, cs2, nu;
  float3 F_const;
  float beta;
  int N;
  int M;
  int3 L;
  float h;
  float dx, dy, dz;
  float dt;
  float D;
  int perX, perY, perZ;
};

struct SpreadEntry;  // (marker, weight) pair, defined in the library
namespace glsm {
struct SpreadLists;
}

class Fluid {
 public:
  Fluid(int3 fluidSize, int time, Coupler* coupler);  // fluid.cu:152-232
  ~Fluid();
  Fluid(const Fluid&) = delete;
  Fluid& operator=(const Fluid&) = delete;

  void stepVelocity(int iter);       // fluid.cu:263-284
  void subStep(float beta_eff, bool write_u_all);  // one pass of fluid.cu:273-282
  void stepConcentration(int iter);  // fluid.cu:251-261
  void writeFiles(int iter);         // fluid.cu:140-150
  void copyDataToHost();             // fluid.cu:85-90
  void recordData(int time);         // fluid.cu:115-138
  void setGoonValue(int time);       // fluid.cu:57-83
  int idx3(int x, int y, int z, int Nx, int Ny);

  // Extensions: FluidParams::perX/Y/Z honoured by stream and IBM too;
  // IBM stiffness / ramp without environment variables.
  void setPeriodic(int perX, int perY, int perZ);
  void setIBM(float beta, float ramp_steps);
  float betaEff(int iter) const;  // fluid.cu:265-270

  // Distribution storage (collision is FP32 in every mode):
  //   F32_DEV  (default) FP32 deviations f - w_q
  //   F32_FULL FP32 populations, the reference's representation
  //   F16_DEV  binary16 deviations scaled by 2^scale_exp
  //   BF16_DEV bfloat16 deviations
  // $GLSM_FLUID_STORAGE (f32 | f32full | f16 | bf16) sets it at construction;
  // setStorage converts the current state.
  enum Storage { F32_FULL = 0, F32_DEV = 1, F16_DEV = 2, BF16_DEV = 3 };
  void setStorage(int storage, int scale_exp = 20);
  int storage() const { return storage_; }
  int storageScaleExp() const { return scale_exp_; }

  // Measurement hook: when enabled, every `every`-th stepVelocity launches
  // its middle sub-step's stream/collide kernel with start/stop events that
  // the dispatch timestamps at kernel start and end (at most `max_samples`
  // pairs).  lbmKernelTiming() waits for the stream and returns the mean
  // kernel duration in ms.
  void enableLbmTiming(int every, int max_samples = 4096);
  double lbmKernelTiming(int* samples);

  // FP64 conservation reductions over the pre-collision populations of the
  // next sub-step: out = {sum(rho - 1), sum(rho ux), sum(rho uy), sum(rho uz)}
  // with rho u = sum_q f_q c_q (fluid_kernels.cu:98-111 without F/2).
  // Deterministic; waits for the fluid stream.
  void conservation(double out[4]);

  // Extension (failure detection, the check the reference's drivers make
  // on their own: sim_gelonly.cpp:112-115): number of cells whose stored
  // populations, velocity or concentration hold a non-finite value (a
  // binary16 deviation that overflowed is stored as Inf and counts), the
  // largest finite |f_q - w_q| (max_abs_dev) and the largest magnitude the
  // storage represents (limit: 65504 / 2^scale_exp for F16_DEV).  Waits for
  // the fluid stream.
  long long checkFinite(double* max_abs_dev = nullptr, double* limit = nullptr);
  // rho - 1 per cell (N doubles) of the pre-collision populations of the
  // next sub-step, summed in FP64 from the stored values.
  void readDensity(double* hostN);

  // Pre-collision populations f_k(x,q) (what the reference's d_f holds) and
  // their inverse; host arrays of 19*N floats, plane-major.
  void readPopulations(float* host19N);
  void writePopulations(const float* host19N);
  void readPostCollision(float* host19N);
  // device view of the IBM spreading lists (valid after the first step)
  glsm::SpreadLists spreadLists() const;

  // host data
  float3* h_u;
  float* h_c1;
  FluidParams* h_fp;
  // device data
  // post-collision populations (storage representation) in the 256-cell
  // block layout: population q of cell i at (i / 256) * 19 * 256 + q * 256 +
  // i % 256 (the reference's d_f holds one plane of N values per
  // population; readPopulations / writePopulations convert)
  void* d_f;      // current
  void* d_fnext;  // next
  float3* d_u;
  float* d_c1;
  float* d_c2;

  Coupler* coupler;
  int3 fluidSize;
  int3 fluidNodeGrid;
  int df;
  float dt;
  float h;
  int N;
  size_t Nd;        // values per distribution buffer: 19 * plane
  size_t plane;     // cells rounded up to whole 256-cell layout blocks
  float dt_fluid;
  int Nsub;
  float niu;
  float dx_fluid;
  float ramp_steps;
  std::thread file_writer_thread;
  hipStream_t fluid_stream;
  int threads;
  int blocksN;
  int blocksM;

 private:
  void ensureSpreadLists();
  int storage_ = F32_DEV;
  int scale_exp_ = 20;
  int timing_every_ = 0;
  int timing_max_ = 0;
  long timing_calls_ = 0;
  std::vector<hipEvent_t> timing_events_;  // start/stop pairs
  int nblk_ = 0;              // 256-cell blocks
  int* d_blkflag_ = nullptr;  // block touched by some marker stencil
  int* d_cnt_ = nullptr;      // per-cell contribution count (scratch, kept 0)
  int* d_fillpos_ = nullptr;  // per-cell fill cursor (scratch, kept 0)
  int2* d_rng_ = nullptr;     // per-cell (first entry, count), flagged blocks
  int* d_bsum_ = nullptr;     // per-block entry totals -> offsets
  double* d_red_ = nullptr;   // 4 FP64 partials per block + the 4 totals
  SpreadEntry* d_ent_ = nullptr;
  int* d_ctab_ = nullptr;     // per (marker, stencil lane): cell or -1
  float* d_wtab_ = nullptr;   // per (marker, stencil lane): spreading weight
  float* d_rtab_ = nullptr;   // per (marker, stencil lane): interpolation weight
  float* d_wsum_ = nullptr;   // per marker: sum of interpolation weights
  int ent_cap_ = 0;
  uint64_t built_lag_version_ = ~0ull;
};
