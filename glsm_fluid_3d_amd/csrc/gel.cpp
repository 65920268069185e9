// gel.cpp -- host side of the Gel class (reference: src/gel.cu).
#include "gel.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "glsm_internal.hpp"

using namespace std;

namespace {

// The 27 neighbour offsets indexed by the type codes of setType
// (gel.cu:655-659): ghost-node mirror direction and no-flux copy direction.
const int kOffset27[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
    {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},
    {-1, 0, -1}, {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0}, {1, 1, 1},   {1, 1, -1},
    {1, -1, 1},  {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

uint64_t splitmix64_next(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double splitmix64_unit(uint64_t& s) { return (double)(splitmix64_next(s) >> 11) * (1.0 / 9007199254740992.0); }

glsm::GelView make_view(Gel& g) {
  glsm::GelView v;
  std::memset(&v, 0, sizeof v);  // padding too: views are compared bytewise
  const GelParams& p = *g.m_hgp;
  v.LX = p.LX; v.LY = p.LY; v.LZ = p.LZ;
  v.gelType = p.gelType;
  v.f = p.f; v.ep = p.ep; v.dt = p.dt; v.dtx = p.dtx; v.dx = p.dx; v.dy = p.dy; v.dz = p.dz;
  v.CH0 = p.CH0; v.CH1 = p.CH1; v.CHS = p.CHS; v.C0 = p.C0; v.AZ0 = p.AZ0; v.FA0 = p.FA0;
  v.c0_bis = p.c0_bis; v.b = p.b; v.uss = p.uss; v.vss = p.vss; v.wss = p.wss;
  v.tube_mode = p.tube_mode; v.tube_cy = p.tube_cy; v.tube_cz = p.tube_cz;
  v.tube_inner_hy = p.tube_inner_hy; v.tube_inner_hz = p.tube_inner_hz;
  v.use_anchor = g.m_use_anchor ? 1 : 0;
  v.anchor_z = g.m_anchor_z;
  v.maxFilamentlen = p.maxFilamentlen;
  for (int t = 0; t < 27; ++t) {
    v.off[t][0] = p.rn_offset[t].x; v.off[t][1] = p.rn_offset[t].y; v.off[t][2] = p.rn_offset[t].z;
  }
  v.um = g.m_dum; v.vm = g.m_dvm; v.wm = g.m_dwm; v.um_norm = g.m_dum_norm; v.vm_norm = g.m_dvm_norm;
  v.wmp = g.m_dwmp; v.T0 = g.m_dT0m; v.T1 = g.m_dT1m; v.T2 = g.m_dT2m; v.vol = g.m_dVolm; v.pm = g.m_dPrem;
  v.rm = g.m_drm; v.rm_loc = g.m_drm_loc; v.nmSm = g.m_dnmSm;
  v.rn = g.m_drn; v.Veln = g.m_dVeln; v.Vels = g.m_dVels; v.Fn = g.m_dFn; v.Fn_robin = g.m_dFn_robin;
  v.Fdrag_robin = g.m_dFdrag_robin;
  v.un_norm = g.m_dun_norm; v.un_robin = g.m_dun_robin; v.vn_norm = g.m_dvn_norm;
  v.map_node = g.m_dmap_node; v.map_element = g.m_dmap_element;
  v.vm_center = g.m_dvm_center; v.wm_center = g.m_dwm_center; v.rm_center = g.m_drm_center;
  v.Fn_center = g.m_dFn_center; v.Veln_center = g.m_dVeln_center; v.filament = g.m_dfilament;
  return v;
}

template <class T>
void h2d(T* d, const T* h, size_t n, hipStream_t s) {
  GLSM_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <class T>
void d2h(T* h, const T* d, size_t n, hipStream_t s) {
  GLSM_HIP(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

}  // namespace

namespace glsm {
GelView gel_view(Gel& g) { return make_view(g); }
}  // namespace glsm

// gel.cu:555-663
Gel::Gel(int3 gelSize, double3 gelPosition, int gelType, int gelId, int time)
    : m_hgp(nullptr), m_hum(nullptr), m_hvm(nullptr), m_hwm(nullptr), m_hvm_center(nullptr),
      m_hwm_center(nullptr), m_hfilament(nullptr), m_hrn(nullptr), m_hrm(nullptr), m_hrm_center(nullptr),
      m_hFn_center(nullptr), m_hVeln_center(nullptr), m_hVeln(nullptr), m_hFn(nullptr), m_hmap_node(nullptr),
      m_hmap_element(nullptr), m_hbIndex(nullptr), m_dgp(nullptr), m_gelId(gelId), m_gelType(gelType),
      m_gelPosition(gelPosition), m_gelSize(gelSize) {
  m_dt = 1e-3;
  m_df = int(1 / m_dt);
  m_dx = 1;
  m_dy = 1;
  m_dz = 1;
  m_gelNodeGrid.x = int(gelSize.x / m_dx) + 1;
  m_gelNodeGrid.y = int(gelSize.y / m_dy) + 1;
  m_gelNodeGrid.z = int(gelSize.z / m_dz) + 1;
  if (m_gelNodeGrid.x < 2 || m_gelNodeGrid.y < 2 || m_gelNodeGrid.z < 2)
    throw std::invalid_argument("Gel: every gel size must be >= 1 element");
  m_numGelElements = (m_gelNodeGrid.x + 1) * (m_gelNodeGrid.y + 1) * (m_gelNodeGrid.z + 1);
  m_numGelNodes = (m_gelNodeGrid.x + 2) * (m_gelNodeGrid.y + 2) * (m_gelNodeGrid.z + 2);
  m_boundaryCount = m_gelNodeGrid.x * m_gelNodeGrid.y * m_gelNodeGrid.z -
                    (m_gelNodeGrid.x - 2) * (m_gelNodeGrid.y - 2) * (m_gelNodeGrid.z - 2);
  m_hgp = new GelParams;
  memset(m_hgp, 0, sizeof(GelParams));
  m_hgp->LX = m_gelNodeGrid.x;
  m_hgp->LY = m_gelNodeGrid.y;
  m_hgp->LZ = m_gelNodeGrid.z;
  m_hgp->I = 0.0;
  m_hgp->f = 0.9;
  m_hgp->ep = 0.3;
  m_hgp->q = 1e-4;
  m_hgp->P1 = 0.0124;
  m_hgp->P2 = 0.77;
  m_hgp->dt = m_dt;
  m_hgp->dtx = 5 * m_dt;
  m_hgp->dx = m_dx;
  m_hgp->dy = m_dy;
  m_hgp->dz = m_dz;
  m_hgp->CH0 = 0.338;
  m_hgp->CH1 = 0.518;
  m_hgp->CHS = 0.1;
  m_hgp->C0 = 1.3e-3;
  m_hgp->c0_bis = 5e-4;
  m_hgp->b = 0.01;
  m_hgp->AZ0 = 100.0;
  m_hgp->FA0 = 0.139;
  m_hgp->maxFilamentlen = 300;
  m_hgp->gelType = m_gelType;
  steadyStateValue(m_hgp->uss, m_hgp->vss, m_hgp->wss, 0);
  for (int t = 0; t < 27; ++t) {
    m_hgp->rn_offset[t] = make_int3(kOffset27[t][0], kOffset27[t][1], kOffset27[t][2]);
    m_hgp->um_offset_noflux[t] = m_hgp->rn_offset[t];
  }
  m_gel_stream = glsm::stream();
  allocateHostStorage();
  allocateDeviceStorage();
  if (time)
    setGoonValue(time);
  else
    setInitValue();
  copyDataToDevice();
}

Gel::~Gel() {
  if (m_file_writer_thread.joinable()) m_file_writer_thread.join();
  (void)hipStreamSynchronize(m_gel_stream);
  freeHostMemory();
  freeDeviceMemory();
}

void Gel::allocateHostStorage() {
  const size_t nE = m_numGelElements, nN = m_numGelNodes;
  m_hum = new double[nE]();
  m_hvm = new double[nN]();  // node-array size: sim_gelonly.cpp:99-101 copies (L+2)^3
  m_hwm = new double[nE]();
  m_hfilament = new double3[1000 * (size_t)m_hgp->maxFilamentlen]();
  m_hrn = new double3[nN]();
  m_hrm = new double3[nE]();
  m_hFn = new double3[nN]();
  m_hVeln = new double3[nN]();
  m_hmap_element = new int[nE]();
  m_hmap_node = new int[nN]();
  m_hvm_center = new double[1000]();
  m_hwm_center = new double[1000]();
  m_hrm_center = new double3[1000]();
  m_hFn_center = new double3[1000]();
  m_hVeln_center = new double3[1000]();
  m_hbIndex = new int[m_boundaryCount]();
}

void Gel::allocateDeviceStorage() {
  using glsm::dalloc;
  const size_t nE = m_numGelElements, nN = m_numGelNodes;
  m_dgp = dalloc<GelParams>(1);
  m_dum = dalloc<double>(nE);
  m_dum_norm = dalloc<double>(nE);
  m_dun_norm = dalloc<double>(nN);
  m_dun_robin = dalloc<double>(nN);
  m_dvm = dalloc<double>(nN);
  m_dvm_norm = dalloc<double>(nE);
  m_dvn_norm = dalloc<double>(nN);
  m_dwm = dalloc<double>(nE);
  m_dwmp = dalloc<double>(nE);
  m_dT0m = dalloc<double>(nE);
  m_dT1m = dalloc<double>(nE);
  m_dT2m = dalloc<double>(nE);
  m_dfilament = dalloc<double3>(1000 * (size_t)m_hgp->maxFilamentlen);
  m_drn = dalloc<double3>(nN);
  m_drm = dalloc<double3>(nE);
  m_dFn = dalloc<double3>(nN);
  m_dFn_robin = dalloc<double3>(nN);
  m_dFdrag_robin = dalloc<double3>(nN);
  m_dVeln = dalloc<double3>(nN);
  m_dVels = dalloc<double3>(nN);
  m_drm_loc = dalloc<double3>(8 * nE);
  m_dnmSm = dalloc<double3>(6 * nE);
  m_dVolm = dalloc<double>(nE);
  m_dPrem = dalloc<double>(nE);
  m_dmap_element = dalloc<int>(nE);
  m_dmap_node = dalloc<int>(nN);
  m_dvm_center = dalloc<double>(1000);
  m_dwm_center = dalloc<double>(1000);
  m_drm_center = dalloc<double3>(1000);
  m_dFn_center = dalloc<double3>(1000);
  m_dVeln_center = dalloc<double3>(1000);
  m_dbIndex = dalloc<int>(m_boundaryCount);
}

void Gel::freeHostMemory() {
  delete[] m_hum; delete[] m_hvm; delete[] m_hwm; delete[] m_hfilament; delete[] m_hrn; delete[] m_hrm;
  delete[] m_hFn; delete[] m_hVeln; delete[] m_hmap_element; delete[] m_hmap_node; delete[] m_hvm_center;
  delete[] m_hwm_center; delete[] m_hrm_center; delete[] m_hFn_center; delete[] m_hVeln_center;
  delete[] m_hbIndex;
  delete m_hgp;
}

void Gel::freeDeviceMemory() {
  for (void* p : {(void*)m_dgp, (void*)m_dum, (void*)m_dum_norm, (void*)m_dun_norm, (void*)m_dun_robin,
                  (void*)m_dvm, (void*)m_dvm_norm, (void*)m_dvn_norm, (void*)m_dwm, (void*)m_dwmp, (void*)m_dT0m,
                  (void*)m_dT1m, (void*)m_dT2m, (void*)m_dfilament, (void*)m_drn, (void*)m_drm, (void*)m_dFn,
                  (void*)m_dFn_robin, (void*)m_dFdrag_robin, (void*)m_dVeln, (void*)m_dVels, (void*)m_drm_loc,
                  (void*)m_dnmSm, (void*)m_dVolm, (void*)m_dPrem, (void*)m_dmap_element, (void*)m_dmap_node,
                  (void*)m_dvm_center, (void*)m_dwm_center, (void*)m_drm_center, (void*)m_dFn_center,
                  (void*)m_dVeln_center, (void*)m_dbIndex, (void*)m_dcheck})
    glsm::dfree(p);
}

long long Gel::checkFinite() {
  if (!m_dcheck) m_dcheck = glsm::dalloc<unsigned long long>(1);
  const size_t nE = m_numGelElements, nN = m_numGelNodes;
  glsm::launch_zero_u64(m_dcheck, 1, m_gel_stream);
  glsm::launch_count_nonfinite(reinterpret_cast<const double*>(m_drn), 3 * nN, m_dcheck, m_gel_stream);
  glsm::launch_count_nonfinite(reinterpret_cast<const double*>(m_dVeln), 3 * nN, m_dcheck, m_gel_stream);
  glsm::launch_count_nonfinite(m_dum, nE, m_dcheck, m_gel_stream);
  glsm::launch_count_nonfinite(m_dvm, nN, m_dcheck, m_gel_stream);
  glsm::launch_count_nonfinite(m_dwm, nE, m_dcheck, m_gel_stream);
  unsigned long long h = 0;
  GLSM_HIP(hipMemcpyAsync(&h, m_dcheck, sizeof h, hipMemcpyDeviceToHost, m_gel_stream));
  GLSM_HIP(hipStreamSynchronize(m_gel_stream));
  return (long long)h;
}

int Gel::get_index(int xi, int yi, int zi, int size) {
  return xi + yi * (m_gelNodeGrid.x + size) + zi * (m_gelNodeGrid.y + size) * (m_gelNodeGrid.x + size);
}

int Gel::idx3(int x, int y, int z, int Nx, int Ny) { return x + y * Nx + z * Nx * Ny; }

// gel.cu:445-450
double Gel::fu_h(double u, double v, double w, double phi) {
  double ww = (1.0 - w) * (1.0 - w);
  return ww * u - u * u - (1.0 - w) * (m_hgp->f * v + phi) * (u - m_hgp->q * ww) / (u + m_hgp->q * ww);
}

// gel.cu:424-443: bisection for the chemistry steady state at w = FA0
void Gel::steadyStateValue(double& um, double& vm, double& wm, double phi) {
  wm = m_hgp->FA0;
  double vsm, usm, usa;
  double vsa = 1.0, vsb = 0.0;
  while (std::abs(vsa - vsb) > 1e-10) {
    vsm = 0.5 * (vsa + vsb);
    usm = vsm / (1 - wm);
    usa = vsa / (1 - wm);
    if (fu_h(usm, vsm, wm, phi) * fu_h(usa, vsa, wm, phi) < 0)
      vsb = vsm;
    else
      vsa = vsm;
  }
  um = vsb / (1.0 - wm);
  vm = vsb;
}

// gel.cu:284-344: type codes (0 interior, 1-6 faces, 7-18 edges, 19-26
// corners) for nodes (size 2) or elements (size 1)
void Gel::setType(int* a, int size) {
  const int LX = m_gelNodeGrid.x + size - 1, LY = m_gelNodeGrid.y + size - 1, LZ = m_gelNodeGrid.z + size - 1;
  auto at = [&](int x, int y, int z) -> int& { return a[get_index(x, y, z, size)]; };
  for (int x = 1; x < LX; x++)
    for (int y = 1; y < LY; y++)
      for (int z = 1; z < LZ; z++) at(x, y, z) = 0;
  for (int y = 1; y < LY; y++)
    for (int z = 1; z < LZ; z++) { at(0, y, z) = 1; at(LX, y, z) = 2; }
  for (int x = 1; x < LX; x++)
    for (int z = 1; z < LZ; z++) { at(x, 0, z) = 3; at(x, LY, z) = 4; }
  for (int x = 1; x < LX; x++)
    for (int y = 1; y < LY; y++) { at(x, y, 0) = 5; at(x, y, LZ) = 6; }
  for (int x = 1; x < LX; x++) { at(x, 0, 0) = 7; at(x, 0, LZ) = 8; at(x, LY, 0) = 9; at(x, LY, LZ) = 10; }
  for (int y = 1; y < LY; y++) { at(0, y, 0) = 11; at(0, y, LZ) = 12; at(LX, y, 0) = 13; at(LX, y, LZ) = 14; }
  for (int z = 1; z < LZ; z++) { at(0, 0, z) = 15; at(0, LY, z) = 16; at(LX, 0, z) = 17; at(LX, LY, z) = 18; }
  at(0, 0, 0) = 19; at(0, 0, LZ) = 20; at(0, LY, 0) = 21; at(0, LY, LZ) = 22;
  at(LX, 0, 0) = 23; at(LX, 0, LZ) = 24; at(LX, LY, 0) = 25; at(LX, LY, LZ) = 26;
}

// gel.cu:161-177: real nodes on the block surface, z,y,x order
void Gel::buildBoundaryIndex() {
  const int LX = m_gelNodeGrid.x, LY = m_gelNodeGrid.y, LZ = m_gelNodeGrid.z;
  int count = 0;
  for (int k = 1; k < LZ + 1; ++k)
    for (int j = 1; j < LY + 1; ++j)
      for (int i = 1; i < LX + 1; ++i)
        if (i == 1 || i == LX || j == 1 || j == LY || k == 1 || k == LZ)
          m_hbIndex[count++] = idx3(i, j, k, LX + 2, LY + 2);
}

// gel.cu:179-219
void Gel::setInitValue() {
  for (int zi = 1; zi < m_gelNodeGrid.z; zi++)
    for (int yi = 1; yi < m_gelNodeGrid.y; yi++)
      for (int xi = 1; xi < m_gelNodeGrid.x; xi++) {
        int gi = get_index(xi, yi, zi, 1);
        m_hum[gi] = m_hgp->uss;
        m_hvm[gi] = m_hgp->vss;
        m_hwm[gi] = m_hgp->wss;
      }
  double lam = pow(m_hgp->FA0 / m_hgp->wss, 1 / 3.0);
  for (int zi = 0; zi < m_gelNodeGrid.z + 2; zi++)
    for (int yi = 0; yi < m_gelNodeGrid.y + 2; yi++)
      for (int xi = 0; xi < m_gelNodeGrid.x + 2; xi++) {
        int hi = get_index(xi, yi, zi, 2);
        m_hrn[hi].x = m_gelPosition.x + (2 * double(xi) - double(m_gelNodeGrid.x) - 1) * m_hgp->dx * lam / 2;
        m_hrn[hi].y = m_gelPosition.y + (2 * double(yi) - double(m_gelNodeGrid.y) - 1) * m_hgp->dy * lam / 2;
        m_hrn[hi].z = m_gelPosition.z + (2 * double(zi) - double(m_gelNodeGrid.z) - 1) * m_hgp->dz * lam / 2;
      }
  setType(m_hmap_node, 2);
  setType(m_hmap_element, 1);
  buildBoundaryIndex();
}

// gel.cu:221-282: restart from the text snapshots gel<id>{rn,um,vm,wm}<time>.dat
// (interior elements, real nodes); ghosts are rebuilt by the first
// elasticity step, as in the reference.
void Gel::setGoonValue(int time) {
  const int LX = m_gelNodeGrid.x, LY = m_gelNodeGrid.y, LZ = m_gelNodeGrid.z;
  const string id = to_string(m_gelId), t = to_string(time);
  ifstream frn("gel" + id + "rn" + t + ".dat"), fum("gel" + id + "um" + t + ".dat"),
      fvm("gel" + id + "vm" + t + ".dat"), fwm("gel" + id + "wm" + t + ".dat");
  if (!frn || !fum || !fvm || !fwm)
    throw std::runtime_error("Gel::setGoonValue: missing restart file gel" + id + "*" + t + ".dat");
  string line;
  for (int zi = 1; zi < LZ; zi++)
    for (int yi = 1; yi < LY; yi++)
      for (int xi = 1; xi < LX; xi++) {
        int gi = get_index(xi, yi, zi, 1);
        double v;
        getline(fum, line); { istringstream s(line); s >> v; m_hum[gi] = v; }
        getline(fvm, line); { istringstream s(line); s >> v; m_hvm[gi] = v; }
        getline(fwm, line); { istringstream s(line); s >> v; m_hwm[gi] = v; }
      }
  for (int zi = 1; zi < LZ + 1; zi++)
    for (int yi = 1; yi < LY + 1; yi++)
      for (int xi = 1; xi < LX + 1; xi++) {
        int gi = get_index(xi, yi, zi, 2);
        double x, y, z;
        getline(frn, line);
        istringstream s(line);
        s >> x >> y >> z;
        m_hrn[gi] = make_double3(x, y, z);
      }
  setType(m_hmap_node, 2);
  setType(m_hmap_element, 1);
  buildBoundaryIndex();
}

// gel.cu:346-356
void Gel::copyDataToDevice() {
  hipStream_t s = m_gel_stream;
  h2d(m_dum, m_hum, m_numGelElements, s);
  h2d(m_dvm, m_hvm, m_numGelElements, s);
  h2d(m_dwm, m_hwm, m_numGelElements, s);
  h2d(m_drn, m_hrn, m_numGelNodes, s);
  h2d(m_dmap_element, m_hmap_element, m_numGelElements, s);
  h2d(m_dmap_node, m_hmap_node, m_numGelNodes, s);
  h2d(m_dgp, m_hgp, 1, s);
  h2d(m_dbIndex, m_hbIndex, m_boundaryCount, s);
  GLSM_HIP(hipStreamSynchronize(s));
  ++m_rn_version;
}

// gel.cu:358-374
void Gel::copyDataToHost() {
  hipStream_t s = m_gel_stream;
  d2h(m_hum, m_dum, m_numGelElements, s);
  d2h(m_hvm, m_dvm, m_numGelElements, s);
  d2h(m_hwm, m_dwm, m_numGelElements, s);
  d2h(m_hfilament, m_dfilament, 1000 * (size_t)m_hgp->maxFilamentlen, s);
  d2h(m_hrm, m_drm, m_numGelElements, s);
  d2h(m_hrn, m_drn, m_numGelNodes, s);
  d2h(m_hFn, m_dFn, m_numGelNodes, s);
  d2h(m_hVeln, m_dVeln, m_numGelNodes, s);
  d2h(m_hvm_center, m_dvm_center, 1000, s);
  d2h(m_hwm_center, m_dwm_center, 1000, s);
  d2h(m_hrm_center, m_drm_center, 1000, s);
  d2h(m_hFn_center, m_dFn_center, 1000, s);
  d2h(m_hVeln_center, m_dVeln_center, 1000, s);
  GLSM_HIP(hipStreamSynchronize(s));
}

// Uploads the host chemistry mirrors whole, as the reference's phase
// controls do (gel.cu:744-746, 781-782, 857-860).
void Gel::uploadChemistry(bool with_w) {
  hipStream_t s = m_gel_stream;
  h2d(m_dum, m_hum, m_numGelElements, s);
  h2d(m_dvm, m_hvm, m_numGelElements, s);
  if (with_w) h2d(m_dwm, m_hwm, m_numGelElements, s);
  GLSM_HIP(hipStreamSynchronize(s));
}

// gel.cu:682-698
void Gel::stepElasticity(int iter) {
  const glsm::GelView v = make_view(*this);
  if (iter % 5 == 0) {
    glsm::launch_gel_elasticity(v, m_gel_stream);
    ++m_rn_version;
  }
  glsm::launch_gel_set_zero(v, m_gel_stream);
}

// gel.cu:700-704
void Gel::stepChemistry(int /*iter*/) { glsm::launch_gel_chemistry(make_view(*this), m_gel_stream); }

// gel.cu:706-713
void Gel::recordCenter(int iter) {
  if (iter % 1000 == 0) glsm::launch_gel_record_center(make_view(*this), (iter / 1000) % 1000, m_gel_stream);
}

// gel.cu:452-541: text snapshot of one output time
void Gel::recordData(int time) {
  const string id = to_string(m_gelId);
  glsm::write_gel_dat("", m_gelId, time, m_gelNodeGrid.x, m_gelNodeGrid.y, m_gelNodeGrid.z, m_hrn, m_hFn, m_hVeln,
                      m_hrm, m_hum, m_hvm, m_hwm);
  if (flag && time % 1000 == 0) {
    ofstream fb("gel" + id + "bodycenter" + ".dat", ios::app);
    for (int i = 0; i < 1000; i++) {
      fb << setw(9) << to_string(time - 1000 + i * 1) << "      " << setw(9) << to_string(m_hrm_center[i].x)
         << "      " << setw(9) << to_string(m_hrm_center[i].y) << "      " << setw(9)
         << to_string(m_hrm_center[i].z) << "      " << setw(9) << to_string(m_hFn_center[i].x) << "      "
         << setw(9) << to_string(m_hFn_center[i].y) << "      " << setw(9) << to_string(m_hFn_center[i].z)
         << "      " << setw(9) << to_string(m_hVeln_center[i].x) << "      " << setw(9)
         << to_string(m_hVeln_center[i].y) << "      " << setw(9) << to_string(m_hVeln_center[i].z) << "      "
         << setw(9) << to_string(m_hvm_center[i]) << "      " << setw(9) << to_string(m_hwm_center[i]) << "\n";
    }
  } else {
    flag = true;
  }
}

// gel.cu:543-553
void Gel::writeFiles(int iter) {
  double time = iter * m_dt;
  if (iter % 1000 == 0) {
    if (m_file_writer_thread.joinable()) m_file_writer_thread.join();
    copyDataToHost();
    // a failed write is reported, not thrown: the reference's ofstreams fail silently
    m_file_writer_thread = std::thread([this, t = int(time)] {
      try {
        recordData(t);
      } catch (const std::exception& e) {
        std::fprintf(stderr, "gel %d: %s\n", m_gelId, e.what());
      }
    });
  }
}

// Extension: rate constants f and ep of GelParams (gel.cu:633-652), with the
// steady state solved again (gel.cu:424-443) and written to the interior
// elements as setInitValue does (gel.cu:179-219)
void Gel::setChemistryRates(double f, double ep) {
  if (!(f > 0.0) || !(ep > 0.0)) throw std::invalid_argument("setChemistryRates: f and ep must be positive");
  m_hgp->f = f;
  m_hgp->ep = ep;
  steadyStateValue(m_hgp->uss, m_hgp->vss, m_hgp->wss, 0);
  for (int zi = 1; zi < m_gelNodeGrid.z; zi++)
    for (int yi = 1; yi < m_gelNodeGrid.y; yi++)
      for (int xi = 1; xi < m_gelNodeGrid.x; xi++) {
        const int gi = get_index(xi, yi, zi, 1);
        m_hum[gi] = m_hgp->uss;
        m_hvm[gi] = m_hgp->vss;
        m_hwm[gi] = m_hgp->wss;
      }
  GLSM_HIP(hipMemcpyAsync(m_dgp, m_hgp, sizeof(GelParams), hipMemcpyHostToDevice, m_gel_stream));
  uploadChemistry(true);
}

// gel.cu:731-747
void Gel::resetToQuiescent() {
  const int LX = m_gelNodeGrid.x, LY = m_gelNodeGrid.y, LZ = m_gelNodeGrid.z;
  for (int zi = 1; zi <= LZ; zi++)
    for (int yi = 1; yi <= LY; yi++)
      for (int xi = 1; xi <= LX; xi++) {
        int gi = get_index(xi, yi, zi, 1);
        m_hum[gi] = m_hgp->uss * 0.01;
        m_hvm[gi] = m_hgp->vss * 0.01;
        m_hwm[gi] = m_hgp->wss;
      }
  uploadChemistry(true);
}

// gel.cu:751-783
void Gel::fireExcitationPulse() {
  const int LX = m_gelNodeGrid.x, LY = m_gelNodeGrid.y, LZ = m_gelNodeGrid.z;
  const int pulse_width = std::max(1, LX / 10);
  for (int zi = 1; zi <= LZ; zi++)
    for (int yi = 1; yi <= LY; yi++) {
      bool is_void = false;
      const float cy = m_hgp->tube_cy, cz = m_hgp->tube_cz;
      const float dy = (float)yi - cy, dz = (float)zi - cz;
      if (m_hgp->tube_mode == 1) {
        is_void = (fabsf(dy) < m_hgp->tube_inner_hy && fabsf(dz) < m_hgp->tube_inner_hz);
      } else if (m_hgp->tube_mode == 2) {
        const float rhy = m_hgp->tube_inner_hy, rhz = m_hgp->tube_inner_hz;
        if (rhy >= 0.5f && rhz >= 0.5f) is_void = ((dy * dy) / (rhy * rhy) + (dz * dz) / (rhz * rhz) < 1.0f);
      }
      if (is_void) continue;
      for (int xi = 1; xi <= pulse_width; xi++) {
        int gi = get_index(xi, yi, zi, 1);
        m_hum[gi] = m_hgp->uss * 3.0;
        m_hvm[gi] = m_hgp->vss * 0.1;
      }
    }
  uploadChemistry(false);
}

// gel.cu:788-795
void Gel::setAnchorZ(double anchor_z) {
  m_anchor_z = anchor_z;
  m_use_anchor = true;
}

// gel.cu:811-867
void Gel::buildTubeMask(int wall_thickness, bool circular) {
  const int LX = m_gelNodeGrid.x, LY = m_gelNodeGrid.y, LZ = m_gelNodeGrid.z;
  const float cy = (LY + 1) * 0.5f, cz = (LZ + 1) * 0.5f;
  float inner_hy = (LY * 0.5f) - wall_thickness;
  float inner_hz = (LZ * 0.5f) - wall_thickness;
  if (inner_hy <= 0) inner_hy = 0;
  if (inner_hz <= 0) inner_hz = 0;
  m_hgp->tube_mode = circular ? 2 : 1;
  m_hgp->tube_cy = cy;
  m_hgp->tube_cz = cz;
  m_hgp->tube_inner_hy = inner_hy;
  m_hgp->tube_inner_hz = inner_hz;
  for (int zi = 1; zi <= LZ; zi++)
    for (int yi = 1; yi <= LY; yi++) {
      const float dy = (float)yi - cy, dz = (float)zi - cz;
      bool is_void;
      if (circular) {
        const float r2 = dy * dy / (inner_hy * inner_hy + 1e-6f) + dz * dz / (inner_hz * inner_hz + 1e-6f);
        is_void = (inner_hy > 0 && inner_hz > 0 && r2 < 1.0f);
      } else {
        is_void = (fabsf(dy) < inner_hy && fabsf(dz) < inner_hz);
      }
      if (!is_void) continue;
      for (int xi = 1; xi <= LX; xi++) {
        int gi = get_index(xi, yi, zi, 1);
        m_hum[gi] = m_hgp->uss;
        m_hvm[gi] = m_hgp->vss;
        m_hwm[gi] = m_hgp->wss;
      }
    }
  h2d(m_dgp, m_hgp, 1, m_gel_stream);
  uploadChemistry(true);
  printf("[buildTubeMask] mode=%s  wall_thickness=%d\n", circular ? "cylinder" : "square", wall_thickness);
  printf("  LY=%d LZ=%d  cy=%.1f cz=%.1f  inner_hy=%.1f inner_hz=%.1f\n", LY, LZ, cy, cz, inner_hy, inner_hz);
  fflush(stdout);
}

// Extension: u += amp*U(-1,1), v += amp*U(-1,1) on interior elements in z,y,x
// order, from splitmix64(seed).  Applied to the device state; the host
// mirrors are refreshed with it.
void Gel::perturbChemistry(uint64_t seed, double amplitude) {
  hipStream_t s = m_gel_stream;
  d2h(m_hum, m_dum, m_numGelElements, s);
  d2h(m_hvm, m_dvm, m_numGelElements, s);
  d2h(m_hwm, m_dwm, m_numGelElements, s);
  GLSM_HIP(hipStreamSynchronize(s));
  uint64_t st = seed;
  for (int zi = 1; zi < m_gelNodeGrid.z; zi++)
    for (int yi = 1; yi < m_gelNodeGrid.y; yi++)
      for (int xi = 1; xi < m_gelNodeGrid.x; xi++) {
        const int gi = get_index(xi, yi, zi, 1);
        const double ru = splitmix64_unit(st);
        const double rv = splitmix64_unit(st);
        m_hum[gi] += amplitude * (2.0 * ru - 1.0);
        m_hvm[gi] += amplitude * (2.0 * rv - 1.0);
      }
  uploadChemistry(false);
}
