// coupling.h -- immersed-boundary glue between gels and the lattice.
// Drop-in for the reference's src/coupling.h:1-58: it packs the boundary
// nodes of all gels into flat marker arrays (gel-major, surface nodes in
// z,y,x order) and scatters the lattice's concentration exchange back.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

class Gel;

// coupling.h:6-12
struct CouplerParams {
  int M;
  float delta;
  float gamma;
  float gel_to_lbm_vel;
  float lbm_to_gel_vel;
};

struct GelMarkerDesc;  // device table entry, defined in the library

class Coupler {
 public:
  explicit Coupler(std::vector<Gel*>& gels);  // coupling.cu:50-80
  ~Coupler();
  Coupler(const Coupler&) = delete;
  Coupler& operator=(const Coupler&) = delete;

  void packFromGels();       // coupling.cu:82-98
  void scatterToGels();      // coupling.cu:100-118
  void applyGelRepulsion();  // coupling.cu:120-123

  // Bumped by packFromGels when the packed marker positions may differ from
  // the previous pack (some gel moved); the fluid re-bins its spreading
  // lists only then.
  uint64_t lagVersion() const { return m_lag_version; }
  // device table of the gels' boundary-node arrays (one entry per gel)
  const GelMarkerDesc* deviceGelTable() const { return d_desc_; }

  CouplerParams* h_cp;
  CouplerParams* d_cp;
  std::vector<int> h_offsets;
  std::vector<int> h_gelBoundaryCount;
  std::vector<int> h_owner;
  int* d_bIndex_all_;
  float3* d_lag_all_;
  float3* d_Ul_all_;
  float3* d_Vl_all_;
  float3* d_Fdrag_all_;
  float3* d_Frep_all_;
  float* d_Cl_all_;
  float* d_Dl_all_;
  float* d_Sl_all_;
  int* d_owner;

  double dt = 1e-3;
  int numGels;
  int sumGelBoundaryCount;
  hipStream_t coupler_stream;
  int threads;
  int blocksM;
  std::vector<Gel*>& gels;

 private:
  GelMarkerDesc* d_desc_ = nullptr;
  int* d_gel_off_ = nullptr;    // first marker of each gel, then the total
  float* d_gel_box_ = nullptr;  // per gel: its markers' bounding box (repulsion)
  std::vector<uint64_t> m_seen_versions;
  uint64_t m_lag_version = 0;
};
