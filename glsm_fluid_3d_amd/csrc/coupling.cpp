// coupling.cpp -- host side of the Coupler class (reference: src/coupling.cu).
#include "coupling.h"

#include <algorithm>
#include <cstring>
#include <vector>

#include "gel.h"
#include "glsm_internal.hpp"

namespace {
glsm::MarkerArrays marker_arrays(Coupler& c, const GelMarkerDesc* desc) {
  glsm::MarkerArrays m;
  m.M = c.sumGelBoundaryCount;
  m.owner = c.d_owner;
  m.bIndex = c.d_bIndex_all_;
  m.desc = desc;
  m.lag = c.d_lag_all_;
  m.Ul = c.d_Ul_all_;
  m.Vl = c.d_Vl_all_;
  m.Fdrag = c.d_Fdrag_all_;
  m.Frep = c.d_Frep_all_;
  m.Cl = c.d_Cl_all_;
  m.Dl = c.d_Dl_all_;
  m.Sl = c.d_Sl_all_;
  return m;
}
}  // namespace

// coupling.cu:50-80 + setInitValue/copyDataToDevice (coupling.cu:28-49)
Coupler::Coupler(std::vector<Gel*>& g)
    : h_cp(nullptr), d_cp(nullptr), d_bIndex_all_(nullptr), d_lag_all_(nullptr), d_Ul_all_(nullptr),
      d_Vl_all_(nullptr), d_Fdrag_all_(nullptr), d_Frep_all_(nullptr), d_Cl_all_(nullptr), d_Dl_all_(nullptr),
      d_Sl_all_(nullptr), d_owner(nullptr), gels(g) {
  numGels = (int)gels.size();
  sumGelBoundaryCount = 0;
  for (int i = 0; i < numGels; ++i) sumGelBoundaryCount += gels[i]->m_boundaryCount;
  threads = 256;
  blocksM = (sumGelBoundaryCount + threads - 1) / threads;
  h_cp = new CouplerParams;
  memset(h_cp, 0, sizeof(CouplerParams));
  h_cp->M = sumGelBoundaryCount;
  h_cp->delta = 1e-4f;
  h_cp->gamma = 1.5f;
  h_cp->gel_to_lbm_vel = 1.0f;
  h_cp->lbm_to_gel_vel = 1.0f;
  coupler_stream = glsm::stream();

  h_offsets.assign(numGels, 0);
  h_gelBoundaryCount.assign(numGels, 0);
  h_owner.assign(sumGelBoundaryCount, 0);
  std::vector<int> bidx(sumGelBoundaryCount);
  std::vector<GelMarkerDesc> desc(numGels);
  int sumM = 0;
  for (int i = 0; i < numGels; ++i) {
    const int Mi = gels[i]->m_boundaryCount;
    h_gelBoundaryCount[i] = Mi;
    h_offsets[i] = sumM;
    for (int l = 0; l < Mi; ++l) {
      h_owner[sumM + l] = i;
      bidx[sumM + l] = gels[i]->m_hbIndex[l];
    }
    sumM += Mi;
    desc[i] = GelMarkerDesc{gels[i]->m_drn, gels[i]->m_dVels, gels[i]->m_dun_norm, gels[i]->m_dun_robin,
                            gels[i]->m_dFn_robin, gels[i]->m_dFdrag_robin};
  }
  const size_t M = sumGelBoundaryCount;
  d_cp = glsm::dalloc<CouplerParams>(1);
  d_lag_all_ = glsm::dalloc<float3>(M);
  d_Ul_all_ = glsm::dalloc<float3>(M);
  d_Vl_all_ = glsm::dalloc<float3>(M);
  d_Fdrag_all_ = glsm::dalloc<float3>(M);
  d_Frep_all_ = glsm::dalloc<float3>(M);
  d_Cl_all_ = glsm::dalloc<float>(M);
  d_Sl_all_ = glsm::dalloc<float>(M);
  d_Dl_all_ = glsm::dalloc<float>(M);
  d_bIndex_all_ = glsm::dalloc<int>(M);
  d_owner = glsm::dalloc<int>(M);
  d_desc_ = glsm::dalloc<GelMarkerDesc>(numGels);
  d_gel_off_ = glsm::dalloc<int>(numGels + 1);
  d_gel_box_ = glsm::dalloc<float>(6 * (size_t)std::max(numGels, 1));
  hipStream_t s = coupler_stream;
  GLSM_HIP(hipMemcpyAsync(d_cp, h_cp, sizeof(CouplerParams), hipMemcpyHostToDevice, s));
  if (M) {
    GLSM_HIP(hipMemcpyAsync(d_bIndex_all_, bidx.data(), M * sizeof(int), hipMemcpyHostToDevice, s));
    GLSM_HIP(hipMemcpyAsync(d_owner, h_owner.data(), M * sizeof(int), hipMemcpyHostToDevice, s));
  }
  if (numGels)
    GLSM_HIP(hipMemcpyAsync(d_desc_, desc.data(), numGels * sizeof(GelMarkerDesc), hipMemcpyHostToDevice, s));
  std::vector<int> gel_off(h_offsets.begin(), h_offsets.end());
  gel_off.push_back(sumGelBoundaryCount);
  GLSM_HIP(hipMemcpyAsync(d_gel_off_, gel_off.data(), gel_off.size() * sizeof(int), hipMemcpyHostToDevice, s));
  GLSM_HIP(hipStreamSynchronize(s));
  m_seen_versions.assign(numGels, 0);
}

Coupler::~Coupler() {
  (void)hipStreamSynchronize(coupler_stream);
  for (void* p : {(void*)d_cp, (void*)d_lag_all_, (void*)d_Ul_all_, (void*)d_Vl_all_, (void*)d_Fdrag_all_,
                  (void*)d_Frep_all_, (void*)d_Cl_all_, (void*)d_Sl_all_, (void*)d_Dl_all_, (void*)d_bIndex_all_,
                  (void*)d_owner, (void*)d_desc_, (void*)d_gel_off_, (void*)d_gel_box_})
    glsm::dfree(p);
  delete h_cp;
}

// coupling.cu:82-98: all gels' boundary nodes in one launch
void Coupler::packFromGels() {
  bool moved = false;
  for (int i = 0; i < numGels; ++i) {
    const uint64_t v = gels[i]->positionVersion();
    if (v != m_seen_versions[i]) {
      m_seen_versions[i] = v;
      moved = true;
    }
  }
  if (moved) ++m_lag_version;
  glsm::launch_pack(marker_arrays(*this, d_desc_), h_cp->gel_to_lbm_vel, coupler_stream);
}

// coupling.cu:100-118
void Coupler::scatterToGels() { glsm::launch_scatter(marker_arrays(*this, d_desc_), coupler_stream); }

// coupling.cu:120-123
void Coupler::applyGelRepulsion() {
  glsm::launch_repulsion(marker_arrays(*this, d_desc_), h_cp->delta, h_cp->gamma, coupler_stream, numGels,
                         d_gel_off_, d_gel_box_);
}

namespace glsm {
// marker arrays for the lattice-side kernels (which never touch the gels)
MarkerArrays markers_of(Coupler& c) { return marker_arrays(c, nullptr); }
MarkerArrays markers_with_gels(Coupler& c) { return marker_arrays(c, c.deviceGelTable()); }
}  // namespace glsm
