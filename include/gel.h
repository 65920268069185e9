// gel.h -- gLSM gel (hexahedral lattice-spring gel with BZ chemistry) on one
// MI355X.  Drop-in for the reference's src/gel.h:1-165: same class name,
// constructor, step methods, phase controls and the public members its
// drivers read (m_df, m_dt, m_gelNodeGrid, m_numGelNodes, m_numGelElements,
// m_h*/m_d* arrays in the reference's AoS layouts and index order).
//
// All device work is enqueued on the library stream (glsm::stream()), so
// calls execute in program order.  Device arrays are zero-filled at
// construction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <thread>

// gel.h:5-31 -- kept field for field (drivers and tests inspect it)
struct GelParams {
  int LX, LY, LZ;
  double f, q, ep, P1, P2;
  double dt, dtx;
  double dx, dy, dz;
  double CH0, CH1, CHS, C0, AZ0, FA0;
  double c0_bis, b;
  double uss, vss, wss, I;
  int3 rn_offset[27];
  int3 um_offset_noflux[27];
  int3 um_offset_periodic[27];
  int gelType;
  int maxFilamentlen;
  float lbm_to_gel_vel;
  int tube_mode;  // 0 solid, 1 square tube, 2 cylinder tube (along X)
  float tube_cy;
  float tube_cz;
  float tube_inner_hy;
  float tube_inner_hz;
};

class Gel {
 public:
  // gel.cu:555-663.  time != 0 restarts from gel<id>{rn,um,vm,wm}<time>.dat
  Gel(int3 gelSize, double3 gelPosition, int gelType, int gel_id, int time);
  ~Gel();
  Gel(const Gel&) = delete;
  Gel& operator=(const Gel&) = delete;

  void setGoonValue(int time);     // gel.cu:221-282
  void stepChemistry(int iter);    // gel.cu:700-704
  void stepElasticity(int iter);   // gel.cu:682-698
  void recordCenter(int iter);     // gel.cu:706-713
  void writeFiles(int iter);       // gel.cu:543-553

  void resetToQuiescent();         // gel.cu:731-747
  void fireExcitationPulse();      // gel.cu:751-783
  void setAnchorZ(double anchor_z);// gel.cu:788-795
  void buildTubeMask(int wall_thickness, bool circular = false);  // gel.cu:811-867

  // Extension: seeded additive chemistry perturbation (splitmix64).
  void perturbChemistry(uint64_t seed, double amplitude);
  // Extension: rate constants f and ep (GelParams, gel.cu:633-652) of this
  // gel; the steady state is solved again (gel.cu:424-443) and the interior
  // chemistry set to it, as at construction (gel.cu:179-219).
  void setChemistryRates(double f, double ep);

  // Extension (failure detection, as the reference's drivers check vm and
  // the tip position for NaN: sim_gelonly.cpp:102-115): number of
  // non-finite values in rn, Veln (node arrays) and um, vm, wm.  Waits for
  // the gel stream.
  long long checkFinite();

  int get_index(int xi, int yi, int zi, int size);  // gel.cu:116-119
  int idx3(int x, int y, int z, int Nx, int Ny);    // gel.cu:156-159
  double fu_h(double u, double v, double w, double phi);            // gel.cu:445-450
  void steadyStateValue(double& u, double& v, double& w, double phi);  // gel.cu:424-443
  void setType(int* a, int size);  // gel.cu:284-344
  void buildBoundaryIndex();       // gel.cu:161-177
  void copyDataToDevice();         // gel.cu:346-356
  void copyDataToHost();           // gel.cu:358-374
  void recordData(int time);       // gel.cu:452-541

  // Monotone counter bumped whenever node positions may have changed; the
  // coupler uses it to skip re-binning unchanged IBM markers.
  uint64_t positionVersion() const { return m_rn_version; }
  void markPositionsChanged() { ++m_rn_version; }

  // ---- host data (reference names) ----
  GelParams* m_hgp;
  double* m_hum;
  double* m_hvm;   // sized to the node array, see gel.cpp
  double* m_hwm;
  double* m_hvm_center;
  double* m_hwm_center;
  double3* m_hfilament;
  double3* m_hrn;
  double3* m_hrm;
  double3* m_hrm_center;
  double3* m_hFn_center;
  double3* m_hVeln_center;
  double3* m_hVeln;
  double3* m_hFn;
  int* m_hmap_node;
  int* m_hmap_element;
  int* m_hbIndex;
  double m_htime = 0;

  // ---- device data (reference names and layouts) ----
  GelParams* m_dgp;
  double* m_dum;
  double* m_dum_norm;
  double* m_dun_norm;
  double* m_dun_robin;
  double* m_dvm;
  double* m_dvm_norm;
  double* m_dvn_norm;
  double* m_dwm;
  double* m_dvm_center;
  double* m_dwm_center;
  double* m_dwmp;
  double* m_dT0m;
  double* m_dT1m;
  double* m_dT2m;
  double3* m_dfilament;
  double3* m_drn;
  double3* m_drm;
  double3* m_drm_center;
  double3* m_dFn_center;
  double3* m_dVeln_center;
  double3* m_dVeln;
  double3* m_dVels;
  double3* m_dFn;
  double3* m_dFn_robin;
  double3* m_dFdrag_robin;
  double3* m_drm_loc;
  double3* m_dnmSm;
  double* m_dVolm;
  double* m_dPrem;
  int* m_dmap_node;
  int* m_dmap_element;
  int* m_dbIndex;

  // ---- parameters ----
  int m_gelId;
  int m_gelType;
  double3 m_gelPosition;
  int3 m_gelSize;
  double m_dx;
  double m_dy;
  double m_dz;
  int3 m_gelNodeGrid;
  double m_dt;
  int m_df;
  int m_numGelElements;
  int m_numGelNodes;
  int m_boundaryCount;
  hipStream_t m_gel_stream;
  std::thread m_file_writer_thread;
  bool flag = false;
  double m_anchor_z = 0.0;
  bool m_use_anchor = false;

 private:
  void allocateHostStorage();
  void allocateDeviceStorage();
  void freeHostMemory();
  void freeDeviceMemory();
  void setInitValue();
  void uploadChemistry(bool with_w);
  uint64_t m_rn_version = 1;
  unsigned long long* m_dcheck = nullptr;  // checkFinite's counter
};
