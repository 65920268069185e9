// glsm_oracle.hpp -- CPU restatement of the gLSM_fluid_3D time step.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, calls or
// includes this file.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it (through oracle/liboracle.so), and only as the
// oracle the library's results are compared with or as the timed CPU
// baseline.
//
// What it restates (reference file:line for every routine is given at the
// routine):
//   * Fluid  -- src/fluid.cu, src/fluid_kernels.cu   (D3Q19 BGK + Guo forcing,
//               push stream with halfway bounce-back, IBM interpolate/spread,
//               concentration advection-diffusion)
//   * Coupler-- src/coupling.cu, src/coupling_kernels.cu
//   * Gel    -- src/gel.cu, src/gel_kernels.cu        (always FP64)
//
// The fluid and coupler are templated on the real type R:
//   R = double : the FP64 correctness oracle;
//   R = float  : the reference's own FP32 arithmetic, used as the OpenMP CPU
//                baseline and as the one-sub-step FP32 comparison.
// Arithmetic is written in the reference's evaluation order (left-assoc
// sums, the same products), and the build uses -ffp-contract=off, so the
// float instantiation reproduces the reference's FP32 rounding operation by
// operation.  Per-cell loops are OpenMP-parallel where cells are independent;
// spreading runs serially in marker order, which makes every result
// deterministic (the reference's atomicAdd order is not).
//
// Deliberate resolutions of reference behaviour (all also in the library):
//   * every buffer starts at zero (the reference reads uninitialised F_ibm on
//     the first sub-step and uninitialised un_robin on the first elasticity
//     step, fluid.cu:41/276, gel.cu:73/689);
//   * one program order (the reference's per-object streams are unordered);
//   * perX/perY/perZ, which the reference only honours in the scalar kernel
//     (fluid_kernels.cu:190-195), are honoured in the stream and in the IBM
//     stencils too; the default 0 keeps the reference's walls;
//   * calFilamentD's hit order (atomic counter, gel_kernels.cu:861) is
//     replaced by element order and bounded by maxFilamentlen;
//   * a seeded additive chemistry perturbation (splitmix64) is an extension.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

namespace glsm_oracle {

template <class T>
struct V3 {
  T x, y, z;
};
using D3 = V3<double>;
struct I3 {
  int x, y, z;
};

// double3 helpers in the reference's operator forms (gel_kernels.cu:7-93).
static inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline D3 neg(D3 a) { return {-a.x, -a.y, -a.z}; }
static inline D3 smul(double a, D3 b) { return {a * b.x, a * b.y, a * b.z}; }
static inline D3 muls(D3 a, double b) { return {a.x * b, a.y * b, a.z * b}; }
static inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline D3 hmul(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline D3 hdiv(D3 a, D3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
static inline D3 divs(D3 a, double b) { return {a.x / b, a.y / b, a.z / b}; }
static inline void addto(D3& a, D3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
static inline D3 cross(D3 a, D3 b) {
  // gel_kernels.cu:80-83
  return {a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y};
}
// gel_kernels.cu:85-88 (pow(d,2) is the correctly rounded square, d*d)
static inline double distance(D3 a, D3 b) {
  double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// ---------------------------------------------------------------------------
// splitmix64 -- the seeded perturbation's generator (extension, no reference
// counterpart; the reference's disabled branch used unseeded rand(),
// gel.cu:188-191).
static inline uint64_t splitmix64_next(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline double splitmix64_unit(uint64_t& s) {
  return (double)(splitmix64_next(s) >> 11) * (1.0 / 9007199254740992.0);
}

// ===========================================================================
// Gel (gLSM), FP64.  src/gel.h, src/gel.cu, src/gel_kernels.cu
// ===========================================================================
struct GelParams {  // gel.h:5-31
  int LX, LY, LZ;
  double f, q, ep, P1, P2;
  double dt, dtx;
  double dx, dy, dz;
  double CH0, CH1, CHS, C0, AZ0, FA0;
  double c0_bis, b;
  double uss, vss, wss, I;
  int rn_offset[27][3];
  int gelType;
  int maxFilamentlen;
  float lbm_to_gel_vel;
  int tube_mode;
  float tube_cy, tube_cz, tube_inner_hy, tube_inner_hz;
};

// The 27 neighbour offsets of gel.cu:655-659, indexed by the type codes of
// setType (gel.cu:284-344).  Used both as the ghost-node mirror direction and
// as the no-flux ghost-element copy direction.
static const int kOffset27[27][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},   {0, 0, -1},
    {0, 1, 1},   {0, 1, -1},  {0, -1, 1},  {0, -1, -1}, {1, 0, 1},   {1, 0, -1},  {-1, 0, 1},
    {-1, 0, -1}, {1, 1, 0},   {1, -1, 0},  {-1, 1, 0},  {-1, -1, 0}, {1, 1, 1},   {1, 1, -1},
    {1, -1, 1},  {1, -1, -1}, {-1, 1, 1},  {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

class Gel {
 public:
  // Gel::Gel (gel.cu:555-663) + _initialize/setInitValue (gel.cu:670-680, 179-219)
  Gel(int sx, int sy, int sz, double px, double py, double pz, int gelType, int gelId)
      : id(gelId), type(gelType), pos{px, py, pz} {
    dt = 1e-3;
    df = int(1 / dt);
    dx = dy = dz = 1;
    L.x = int(sx / dx) + 1;
    L.y = int(sy / dy) + 1;
    L.z = int(sz / dz) + 1;
    nE = (L.x + 1) * (L.y + 1) * (L.z + 1);
    nN = (L.x + 2) * (L.y + 2) * (L.z + 2);
    M = L.x * L.y * L.z - (L.x - 2) * (L.y - 2) * (L.z - 2);
    std::memset(&gp, 0, sizeof(gp));
    gp.LX = L.x; gp.LY = L.y; gp.LZ = L.z;
    gp.I = 0.0; gp.f = 0.9; gp.ep = 0.3; gp.q = 1e-4; gp.P1 = 0.0124; gp.P2 = 0.77;
    gp.dt = dt; gp.dtx = 5 * dt; gp.dx = dx; gp.dy = dy; gp.dz = dz;
    gp.CH0 = 0.338; gp.CH1 = 0.518; gp.CHS = 0.1; gp.C0 = 1.3e-3; gp.c0_bis = 5e-4;
    gp.b = 0.01; gp.AZ0 = 100.0; gp.FA0 = 0.139; gp.maxFilamentlen = 300;
    gp.gelType = gelType;
    steadyStateValue(gp.uss, gp.vss, gp.wss, 0);
    std::memcpy(gp.rn_offset, kOffset27, sizeof(kOffset27));
    allocate();
    setInitValue();
  }

  // ---- geometry helpers --------------------------------------------------
  // gel.cu:116-119 / gel_kernels.cu:95-98
  int idx(int xi, int yi, int zi, int size) const {
    return xi + yi * (L.x + size) + zi * (L.y + size) * (L.x + size);
  }
  int elem(int xi, int yi, int zi) const { return idx(xi, yi, zi, 1); }
  int node(int xi, int yi, int zi) const { return idx(xi, yi, zi, 2); }

  // gel.cu:445-450
  double fu_h(double u, double v, double w, double phi) const {
    double ww = (1.0 - w) * (1.0 - w);
    return ww * u - u * u - (1.0 - w) * (gp.f * v + phi) * (u - gp.q * ww) / (u + gp.q * ww);
  }
  // gel.cu:424-443 (bisection on v in [0,1])
  void steadyStateValue(double& um_, double& vm_, double& wm_, double phi) const {
    wm_ = gp.FA0;
    double vsm, usm, usa;
    double vsa = 1.0, vsb = 0.0;
    while (std::abs(vsa - vsb) > 1e-10) {
      vsm = 0.5 * (vsa + vsb);
      usm = vsm / (1 - wm_);
      usa = vsa / (1 - wm_);
      if (fu_h(usm, vsm, wm_, phi) * fu_h(usa, vsa, wm_, phi) < 0)
        vsb = vsm;
      else
        vsa = vsm;
    }
    um_ = vsb / (1.0 - wm_);
    vm_ = vsb;
  }

  // gel.cu:284-344: face/edge/corner type codes, size 2 for nodes, 1 for elements
  void setType(std::vector<int>& a, int size) const {
    int LX = L.x + size - 1, LY = L.y + size - 1, LZ = L.z + size - 1;
    for (int xi = 1; xi < LX; xi++)
      for (int yi = 1; yi < LY; yi++)
        for (int zi = 1; zi < LZ; zi++) a[idx(xi, yi, zi, size)] = 0;
    for (int yi = 1; yi < LY; yi++)
      for (int zi = 1; zi < LZ; zi++) {
        a[idx(0, yi, zi, size)] = 1;
        a[idx(LX, yi, zi, size)] = 2;
      }
    for (int xi = 1; xi < LX; xi++)
      for (int zi = 1; zi < LZ; zi++) {
        a[idx(xi, 0, zi, size)] = 3;
        a[idx(xi, LY, zi, size)] = 4;
      }
    for (int xi = 1; xi < LX; xi++)
      for (int yi = 1; yi < LY; yi++) {
        a[idx(xi, yi, 0, size)] = 5;
        a[idx(xi, yi, LZ, size)] = 6;
      }
    for (int xi = 1; xi < LX; xi++) {
      a[idx(xi, 0, 0, size)] = 7;
      a[idx(xi, 0, LZ, size)] = 8;
      a[idx(xi, LY, 0, size)] = 9;
      a[idx(xi, LY, LZ, size)] = 10;
    }
    for (int yi = 1; yi < LY; yi++) {
      a[idx(0, yi, 0, size)] = 11;
      a[idx(0, yi, LZ, size)] = 12;
      a[idx(LX, yi, 0, size)] = 13;
      a[idx(LX, yi, LZ, size)] = 14;
    }
    for (int zi = 1; zi < LZ; zi++) {
      a[idx(0, 0, zi, size)] = 15;
      a[idx(0, LY, zi, size)] = 16;
      a[idx(LX, 0, zi, size)] = 17;
      a[idx(LX, LY, zi, size)] = 18;
    }
    a[idx(0, 0, 0, size)] = 19;
    a[idx(0, 0, LZ, size)] = 20;
    a[idx(0, LY, 0, size)] = 21;
    a[idx(0, LY, LZ, size)] = 22;
    a[idx(LX, 0, 0, size)] = 23;
    a[idx(LX, 0, LZ, size)] = 24;
    a[idx(LX, LY, 0, size)] = 25;
    a[idx(LX, LY, LZ, size)] = 26;
  }

  // gel.cu:161-177: surface real nodes in z,y,x order
  void buildBoundaryIndex() {
    int count = 0;
    for (int k = 1; k < L.z + 1; ++k)
      for (int j = 1; j < L.y + 1; ++j)
        for (int i = 1; i < L.x + 1; ++i)
          if (i == 1 || i == L.x || j == 1 || j == L.y || k == 1 || k == L.z)
            bIndex[count++] = i + j * (L.x + 2) + k * (L.x + 2) * (L.y + 2);
  }

  void allocate() {
    // m_hvm/m_dvm are sized to the node array so that sim_gelonly.cpp's
    // (L+2)^3 copy (sim_gelonly.cpp:99-101) stays in bounds.
    um.assign(nE, 0); vm.assign(nN, 0); wm.assign(nE, 0);
    um_norm.assign(nE, 0); vm_norm.assign(nE, 0); wmp.assign(nE, 0);
    T0.assign(nE, 0); T1.assign(nE, 0); T2.assign(nE, 0);
    vol.assign(nE, 0); pm.assign(nE, 0);
    rm.assign(nE, D3{0, 0, 0}); rm_loc.assign(8 * (size_t)nE, D3{0, 0, 0});
    nmSm.assign(6 * (size_t)nE, D3{0, 0, 0});
    rn.assign(nN, D3{0, 0, 0}); Veln.assign(nN, D3{0, 0, 0}); Vels.assign(nN, D3{0, 0, 0});
    Fn.assign(nN, D3{0, 0, 0}); Fn_robin.assign(nN, D3{0, 0, 0}); Fdrag_robin.assign(nN, D3{0, 0, 0});
    un_norm.assign(nN, 0); un_robin.assign(nN, 0); vn_norm.assign(nN, 0);
    map_node.assign(nN, 0); map_element.assign(nE, 0); bIndex.assign(M, 0);
    vm_center.assign(1000, 0); wm_center.assign(1000, 0);
    rm_center.assign(1000, D3{0, 0, 0}); Fn_center.assign(1000, D3{0, 0, 0});
    Veln_center.assign(1000, D3{0, 0, 0});
    filament.assign(1000 * (size_t)gp.maxFilamentlen, D3{0, 0, 0});
  }

  // gel.cu:179-219
  void setInitValue() {
    for (int zi = 1; zi < L.z; zi++)
      for (int yi = 1; yi < L.y; yi++)
        for (int xi = 1; xi < L.x; xi++) {
          int gi = elem(xi, yi, zi);
          um[gi] = gp.uss;
          vm[gi] = gp.vss;
          wm[gi] = gp.wss;
        }
    double lam = std::pow(gp.FA0 / gp.wss, 1 / 3.0);
    for (int zi = 0; zi < L.z + 2; zi++)
      for (int yi = 0; yi < L.y + 2; yi++)
        for (int xi = 0; xi < L.x + 2; xi++) {
          int hi = node(xi, yi, zi);
          rn[hi].x = pos.x + (2 * double(xi) - double(L.x) - 1) * gp.dx * lam / 2;
          rn[hi].y = pos.y + (2 * double(yi) - double(L.y) - 1) * gp.dy * lam / 2;
          rn[hi].z = pos.z + (2 * double(zi) - double(L.z) - 1) * gp.dz * lam / 2;
        }
    setType(map_node, 2);
    setType(map_element, 1);
    buildBoundaryIndex();
    snapshotToHost();
  }

  // The reference keeps host mirrors (m_hum, m_hvm, m_hwm) that are refreshed
  // only by copyDataToHost (gel.cu:358-374); the phase controls edit those
  // mirrors and upload them whole (gel.cu:744-746, 781-782, 857-860).
  void snapshotToHost() { h_um = um; h_vm = vm; h_wm = wm; }

  // gel.cu:731-747
  void resetToQuiescent() {
    for (int zi = 1; zi <= L.z; zi++)
      for (int yi = 1; yi <= L.y; yi++)
        for (int xi = 1; xi <= L.x; xi++) {
          int gi = elem(xi, yi, zi);
          h_um[gi] = gp.uss * 0.01;
          h_vm[gi] = gp.vss * 0.01;
          h_wm[gi] = gp.wss;
        }
    um = h_um; vm = h_vm; wm = h_wm;
  }

  // host-side void test shared by fireExcitationPulse (gel.cu:759-772)
  bool pulseVoid(int yi, int zi) const {
    float cy = gp.tube_cy, cz = gp.tube_cz;
    float ddy = (float)yi - cy, ddz = (float)zi - cz;
    if (gp.tube_mode == 1) return std::fabs(ddy) < gp.tube_inner_hy && std::fabs(ddz) < gp.tube_inner_hz;
    if (gp.tube_mode == 2) {
      float rhy = gp.tube_inner_hy, rhz = gp.tube_inner_hz;
      if (rhy >= 0.5f && rhz >= 0.5f) return (ddy * ddy) / (rhy * rhy) + (ddz * ddz) / (rhz * rhz) < 1.0f;
    }
    return false;
  }
  // gel.cu:751-783
  void fireExcitationPulse() {
    int pulse_width = std::max(1, L.x / 10);
    for (int zi = 1; zi <= L.z; zi++)
      for (int yi = 1; yi <= L.y; yi++) {
        if (pulseVoid(yi, zi)) continue;
        for (int xi = 1; xi <= pulse_width; xi++) {
          int gi = elem(xi, yi, zi);
          h_um[gi] = gp.uss * 3.0;
          h_vm[gi] = gp.vss * 0.1;
        }
      }
    um = h_um; vm = h_vm;
  }
  // gel.cu:788-795
  void setAnchorZ(double z) { anchor_z = z; use_anchor = true; }
  // gel.cu:811-867
  void buildTubeMask(int wall_thickness, bool circular) {
    float cy = (L.y + 1) * 0.5f, cz = (L.z + 1) * 0.5f;
    float inner_hy = (L.y * 0.5f) - wall_thickness;
    float inner_hz = (L.z * 0.5f) - wall_thickness;
    if (inner_hy <= 0) inner_hy = 0;
    if (inner_hz <= 0) inner_hz = 0;
    gp.tube_mode = circular ? 2 : 1;
    gp.tube_cy = cy; gp.tube_cz = cz;
    gp.tube_inner_hy = inner_hy; gp.tube_inner_hz = inner_hz;
    for (int zi = 1; zi <= L.z; zi++)
      for (int yi = 1; yi <= L.y; yi++) {
        float ddy = (float)yi - cy, ddz = (float)zi - cz;
        bool is_void;
        if (circular) {
          float r2 = ddy * ddy / (inner_hy * inner_hy + 1e-6f) + ddz * ddz / (inner_hz * inner_hz + 1e-6f);
          is_void = (inner_hy > 0 && inner_hz > 0 && r2 < 1.0f);
        } else {
          is_void = (std::fabs(ddy) < inner_hy && std::fabs(ddz) < inner_hz);
        }
        if (!is_void) continue;
        for (int xi = 1; xi <= L.x; xi++) {
          int gi = elem(xi, yi, zi);
          h_um[gi] = gp.uss;
          h_vm[gi] = gp.vss;
          h_wm[gi] = gp.wss;
        }
      }
    um = h_um; vm = h_vm; wm = h_wm;
  }

  // Extension: rate constants f, ep; steady state solved again
  // (gel.cu:424-443) and the interior chemistry set to it (gel.cu:179-219).
  void setChemistryRates(double f, double ep) {
    gp.f = f;
    gp.ep = ep;
    steadyStateValue(gp.uss, gp.vss, gp.wss, 0);
    for (int zi = 1; zi < L.z; zi++)
      for (int yi = 1; yi < L.y; yi++)
        for (int xi = 1; xi < L.x; xi++) {
          int gi = elem(xi, yi, zi);
          um[gi] = gp.uss;
          vm[gi] = gp.vss;
          wm[gi] = gp.wss;
        }
    snapshotToHost();
  }

  // Extension: interior elements (z,y,x order) get u += amp*U(-1,1) then
  // v += amp*U(-1,1) from splitmix64(seed); w untouched.
  void perturbChemistry(uint64_t seed, double amp) {
    uint64_t s = seed;
    for (int zi = 1; zi < L.z; zi++)
      for (int yi = 1; yi < L.y; yi++)
        for (int xi = 1; xi < L.x; xi++) {
          int gi = elem(xi, yi, zi);
          double ru = splitmix64_unit(s);
          double rv = splitmix64_unit(s);
          um[gi] += amp * (2.0 * ru - 1.0);
          vm[gi] += amp * (2.0 * rv - 1.0);
        }
    snapshotToHost();
  }

  // ---- kernels (device code restated) ------------------------------------
  // calServiceNodesPositionD, gel_kernels.cu:615-631: ghost node mirrored
  // through the first real node along its type's offset.
  void serviceNodes() {
    for (int zi = 0; zi <= L.z + 1; zi++)
      for (int yi = 0; yi <= L.y + 1; yi++)
        for (int xi = 0; xi <= L.x + 1; xi++) {
          int gi = node(xi, yi, zi);
          int t = map_node[gi];
          if (t == 0) continue;
          const int* o = gp.rn_offset[t];
          int g1 = node(xi + o[0], yi + o[1], zi + o[2]);
          int g2 = node(xi + 2 * o[0], yi + 2 * o[1], zi + 2 * o[2]);
          rn[gi] = sub(smul(2, rn[g1]), rn[g2]);
        }
  }

  // Shape-function derivative dN_i/d(axis) of the trilinear hexahedron at
  // local point p, node i = 4*ix + 2*iy + iz (the node[] order of
  // gel_kernels.cu:236-246).  All values are exact dyadic rationals, equal to
  // the literal expressions of gel_kernels.cu:280-283 ... 448-455.
  static double dshape(int i, int axis, const double p[3]) {
    double sgn[3] = {double(2 * ((i >> 2) & 1) - 1), double(2 * ((i >> 1) & 1) - 1), double(2 * (i & 1) - 1)};
    double v = sgn[axis];
    for (int a = 0; a < 3; a++)
      if (a != axis) v *= (1 + p[a] * sgn[a]);
    return v / 8;
  }

  // calElementPropertiesD, gel_kernels.cu:216-483 (the live branch).
  void elementProperties() {
    // face table: node order of coordinate_st, the two in-face axes (outer,
    // inner quadrature loop), fixed axis value, and the stored sign.
    static const int fnode[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1},
                                    {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
    static const int faxes[6][3] = {{1, 2, 0}, {1, 2, 0}, {0, 2, 1}, {0, 2, 1}, {0, 1, 2}, {0, 1, 2}};
    static const double fval[6] = {-1, 1, -1, 1, -1, 1};
    static const double fsign[6] = {-1, 1, 1, -1, -1, 1};
    const double qp[2] = {-0.5, 0.5};
    for (int zi = 0; zi <= L.z; zi++)
      for (int yi = 0; yi <= L.y; yi++)
        for (int xi = 0; xi <= L.x; xi++) {
          int gi = elem(xi, yi, zi);
          D3 nd[8];
          D3 bc{0, 0, 0};
          int c = 0;
          for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
              for (int k = 0; k < 2; k++) {
                nd[c] = rn[node(xi + i, yi + j, zi + k)];
                addto(bc, nd[c]);
                c++;
              }
          bc = divs(bc, 8);
          rm[gi] = bc;
          for (int i = 0; i < 8; i++) {
            nd[i] = sub(nd[i], bc);
            rm_loc[8 * (size_t)gi + i] = nd[i];
          }
          for (int f = 0; f < 6; f++) {
            D3 detJ{0, 0, 0};
            for (int a = 0; a < 2; a++)
              for (int b = 0; b < 2; b++) {
                double p[3];
                p[faxes[f][2]] = fval[f];
                p[faxes[f][0]] = qp[a];
                p[faxes[f][1]] = qp[b];
                D3 J0{0, 0, 0}, J1{0, 0, 0};
                for (int i = 0; i < 4; i++) {
                  int n = fnode[f][i];
                  addto(J0, muls(nd[n], dshape(n, faxes[f][0], p)));
                  addto(J1, muls(nd[n], dshape(n, faxes[f][1], p)));
                }
                addto(detJ, cross(J0, J1));
              }
            detJ = muls(detJ, 1.0);  // k_s = 4/2^2
            nmSm[6 * (size_t)gi + f] = fsign[f] < 0 ? neg(detJ) : detJ;
          }
          double detv = 0;
          for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
              for (int cc = 0; cc < 2; cc++) {
                double p[3] = {qp[a], qp[b], qp[cc]};
                double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                for (int i = 0; i < 8; i++) {
                  double d0 = dshape(i, 0, p), d1 = dshape(i, 1, p), d2 = dshape(i, 2, p);
                  J[0][0] += d0 * nd[i].x; J[0][1] += d0 * nd[i].y; J[0][2] += d0 * nd[i].z;
                  J[1][0] += d1 * nd[i].x; J[1][1] += d1 * nd[i].y; J[1][2] += d1 * nd[i].z;
                  J[2][0] += d2 * nd[i].x; J[2][1] += d2 * nd[i].y; J[2][2] += d2 * nd[i].z;
                }
                detv += (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                         J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2]) +
                         J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]));
              }
          double v = detv * 1.0;  // k_v = 8/2^3
          vol[gi] = v;
          wmp[gi] = wm[gi];
          wm[gi] = gp.dx * gp.dy * gp.dz * gp.FA0 / v;
        }
  }

  // calChemBoundaryD, gel_kernels.cu:757-778 (no-flux ghost copy + norms)
  void chemBoundary() {
    for (int zi = 0; zi <= L.z; zi++)
      for (int yi = 0; yi <= L.y; yi++)
        for (int xi = 0; xi <= L.x; xi++) {
          int gi = elem(xi, yi, zi);
          const int* o = kOffset27[map_element[gi]];
          int g1 = elem(xi + o[0], yi + o[1], zi + o[2]);
          um[gi] = um[g1];
          vm[gi] = vm[g1];
          um_norm[gi] = um[gi] / (1 - wm[gi]);
          vm_norm[gi] = vm[gi] / (1 - wm[gi]);
        }
  }

  // calUnnormD, gel_kernels.cu:795-821
  void unnorm() {
    for (int zi = 1; zi <= L.z; zi++)
      for (int yi = 1; yi <= L.y; yi++)
        for (int xi = 1; xi <= L.x; xi++) {
          double su = 0, sv = 0;
          for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
              for (int k = 0; k < 2; k++) {
                int e = elem(xi - i, yi - j, zi - k);
                su += um_norm[e];
                sv += vm_norm[e];
              }
          int gi = node(xi, yi, zi);
          un_norm[gi] = su / 8 + un_robin[gi];
          vn_norm[gi] = sv / 8;
        }
  }

  // calPressureD, gel_kernels.cu:496-516
  void pressure() {
    for (int zi = 1; zi < L.z; zi++)
      for (int yi = 1; yi < L.y; yi++)
        for (int xi = 1; xi < L.x; xi++) {
          int gi = elem(xi, yi, zi);
          double w = wm[gi];
          if (gp.gelType == 1) {
            pm[gi] = -(w + std::log(1.0 - w) + (gp.CH0 + gp.CH1 * w) * w * w) + gp.C0 * w / (2.0 * gp.FA0) +
                     gp.CHS * w * vm[gi];
          } else if (gp.gelType == 2) {
            double c0 = gp.c0_bis + gp.b * vm[gi];
            pm[gi] = -(w + std::log(1.0 - w) + (gp.CH0 + gp.CH1 * w) * w * w) + c0 * w / (2.0 * gp.FA0);
          }
        }
  }

  // calNodesVelocityD, gel_kernels.cu:518-598.  The eight adjacent-element
  // blocks are visited in the reference's order: sx (+ then -) outer, sy,
  // sz inner.
  void nodesVelocity() {
    for (int zi = 1; zi <= L.z; zi++)
      for (int yi = 1; yi <= L.y; yi++)
        for (int xi = 1; xi <= L.x; xi++) {
          int gi = node(xi, yi, zi);
          double wn = 0;
          for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
              for (int k = 0; k < 2; k++) wn += wm[elem(xi - i, yi - j, zi - k)];
          wn /= 8;
          double Mn = 8. * gp.AZ0 * std::sqrt(gp.FA0 / wn) * (1 - wn) / (gp.dx * gp.dy * gp.dz);
          D3 f1{0, 0, 0}, f2{0, 0, 0};
          D3 r0 = rn[gi];
          for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++)
              for (int c = 0; c < 2; c++) {
                int sx = a ? -1 : 1, sy = b ? -1 : 1, sz = c ? -1 : 1;
                bool okx = sx > 0 ? xi < L.x : xi > 1;
                bool oky = sy > 0 ? yi < L.y : yi > 1;
                bool okz = sz > 0 ? zi < L.z : zi > 1;
                if (!(okx && oky && okz)) continue;
                int ex = xi - a, ey = yi - b, ez = zi - c;
                size_t fb = 6 * (size_t)elem(ex, ey, ez);
                D3 t = add(add(add(rn[node(xi + sx, yi + sy, zi + sz)], rn[node(xi, yi + sy, zi + sz)]),
                               rn[node(xi + sx, yi, zi + sz)]),
                           rn[node(xi + sx, yi + sy, zi)]);
                addto(f1, sub(t, smul(4, r0)));
                D3 s3 = add(add(nmSm[fb + (sx > 0 ? 0 : 1)], nmSm[fb + (sy > 0 ? 2 : 3)]), nmSm[fb + (sz > 0 ? 4 : 5)]);
                addto(f2, smul(pm[elem(ex, ey, ez)], s3));
              }
          // pow(dx*dy*dz, 1/3) with integer 1/3 == 0 is pow(., 0) == 1
          if (gp.gelType == 1) {
            f1 = muls(f1, std::pow(gp.dx * gp.dy * gp.dz, 1 / 3) * gp.C0 / 12.);
          } else if (gp.gelType == 2) {
            double c0 = gp.c0_bis + gp.b * vn_norm[gi] * (1 - wn);
            f1 = muls(f1, std::pow(gp.dx * gp.dy * gp.dz, 1 / 3) * c0 / 12.);
          }
          f2 = divs(f2, 4.);
          Fn[gi] = add(f1, f2);
          Veln[gi] = smul(Mn, Fn[gi]);
          Vels[gi] = divs(smul(-wn, Veln[gi]), (1 - wn));
        }
  }

  // calInternalNodesPositionD, gel_kernels.cu:600-613
  void nodesPosition() {
    for (int zi = 1; zi <= L.z; zi++)
      for (int yi = 1; yi <= L.y; yi++)
        for (int xi = 1; xi <= L.x; xi++) {
          int gi = node(xi, yi, zi);
          addto(rn[gi], smul(gp.dtx, Veln[gi]));
        }
  }

  // anchorBottomNodesD, gel_kernels.cu:870-880, launched on gridDim0 with no
  // +1 shift (gel.cu:694): x and y are covered only below 8*ceil(L/8).
  void anchorBottom() {
    int xcov = 8 * ((L.x + 7) / 8), ycov = 8 * ((L.y + 7) / 8);
    for (int yi = 1; yi <= L.y && yi < ycov; yi++)
      for (int xi = 1; xi <= L.x && xi < xcov; xi++) {
        int gi = node(xi, yi, 1);
        Veln[gi] = D3{0.0, 0.0, 0.0};
        rn[gi].z = anchor_z;
      }
  }

  // calTermsD, gel_kernels.cu:633-703
  void terms() {
    for (int zi = 1; zi < L.z; zi++)
      for (int yi = 1; yi < L.y; yi++)
        for (int xi = 1; xi < L.x; xi++) {
          int gi = elem(xi, yi, zi);
          double w0 = wm[gi], u0 = um_norm[gi];
          D3 r0 = rm[gi];
          T0[gi] = (1 - w0 / wmp[gi]) / gp.dtx;
          D3 vmum{0, 0, 0};
          int c = 0;
          for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
              for (int k = 0; k < 2; k++) {
                D3 rl = rm_loc[8 * (size_t)gi + c];
                c++;
                int n = node(xi + i, yi + j, zi + k);
                double N = 0.125 * (1 + (2 * double(i) - 1) * rl.x) * (1 + (2 * double(j) - 1) * rl.y) *
                           (1 + (2 * double(k) - 1) * rl.z);
                addto(vmum, muls(smul(N, Veln[n]), un_norm[n]));
              }
          double dT1 = 0;
          for (int f = 0; f < 6; f++) dT1 += dot(nmSm[6 * (size_t)gi + f], vmum);
          T1[gi] = dT1 / vol[gi];
          int g1 = elem(xi + 1, yi, zi), g2 = elem(xi - 1, yi, zi);
          int g3 = elem(xi, yi + 1, zi), g4 = elem(xi, yi - 1, zi);
          int g5 = elem(xi, yi, zi + 1), g6 = elem(xi, yi, zi - 1);
          D3 aa{distance(rm[g1], r0), distance(rm[g3], r0), distance(rm[g5], r0)};
          D3 as{distance(rm[g2], r0), distance(rm[g4], r0), distance(rm[g6], r0)};
          D3 aa2 = hmul(aa, aa), as2 = hmul(as, as);
          D3 aas = hmul(hmul(aa, as), add(aa, as));
          D3 wa{wm[g1], wm[g3], wm[g5]}, ws{wm[g2], wm[g4], wm[g6]};
          D3 ua{um_norm[g1], um_norm[g3], um_norm[g5]}, us{um_norm[g2], um_norm[g4], um_norm[g6]};
          D3 dw = hdiv(sub(add(smul(w0, sub(aa2, as2)), hmul(wa, as2)), hmul(ws, aa2)), aas);
          D3 du = hdiv(sub(add(smul(u0, sub(aa2, as2)), hmul(ua, as2)), hmul(us, aa2)), aas);
          D3 d2u = hdiv(smul(2, sub(add(hmul(ua, as), hmul(us, aa)), smul(u0, add(aa, as)))), aas);
          T2[gi] = -dot(dw, du) + (1 - w0) * (d2u.x + d2u.y + d2u.z);
        }
  }

  // isVoidElement, gel_kernels.cu:104-117 (device-side void test)
  bool isVoid(int yi, int zi) const {
    if (gp.tube_mode == 0) return false;
    float ddy = (float)yi - gp.tube_cy, ddz = (float)zi - gp.tube_cz;
    if (gp.tube_mode == 1) return std::fabs(ddy) < gp.tube_inner_hy && std::fabs(ddz) < gp.tube_inner_hz;
    float rhy = gp.tube_inner_hy, rhz = gp.tube_inner_hz;
    if (rhy < 0.5f || rhz < 0.5f) return false;
    float r2 = (ddy * ddy) / (rhy * rhy) + (ddz * ddz) / (rhz * rhz);
    return r2 < 1.0f;
  }
  // fv / fu, gel_kernels.cu:129-145 (q, P1, P2 hard-coded as in the kernels)
  static double fv(double u, double v, double w, double I) {
    double P1 = 0.0124, P2 = 0.77;
    return (1.0 - w) * (1.0 - w) * u - (1.0 - w) * v + I * (0.5 * P1 + P2);
  }
  static double fu(double u, double v, double w, double f, double I) {
    double P1 = 0.0124, P2 = 0.77, q = 1e-4;
    double ww = (1.0 - w) * (1.0 - w);
    return ww * u - u * u - (1.0 - w) * (f * v + I * P1) * (u - q * ww) / (u + q * ww) + I * P2;
  }
  // calChemD, gel_kernels.cu:705-755 (forward Euler from pre-step values)
  void chem() {
    for (int zi = 1; zi < L.z; zi++)
      for (int yi = 1; yi < L.y; yi++)
        for (int xi = 1; xi < L.x; xi++) {
          int gi = elem(xi, yi, zi);
          if (isVoid(yi, zi)) {
            vm[gi] = gp.vss;
            um[gi] = gp.uss;
            wm[gi] = gp.wss;
            continue;
          }
          double I = 0;
          double v0 = vm[gi], u0 = um[gi], w0 = wm[gi];
          vm[gi] += gp.dt * (-v0 * T0[gi] + gp.ep * fv(u0, v0, w0, I));
          um[gi] += gp.dt * (-u0 * T0[gi] + T1[gi] + T2[gi] + fu(u0, v0, w0, gp.f, I));
        }
  }
  // setZero, gel_kernels.cu:780-793 (Fdrag_robin is not cleared)
  void setZero() {
    for (int zi = 1; zi <= L.z; zi++)
      for (int yi = 1; yi <= L.y; yi++)
        for (int xi = 1; xi <= L.x; xi++) {
          int gi = node(xi, yi, zi);
          un_robin[gi] = 0;
          Fn_robin[gi] = D3{0., 0., 0.};
        }
  }

  // Gel::stepElasticity, gel.cu:682-698
  void stepElasticity(int iter) {
    if (iter % 5 == 0) {
      serviceNodes();
      elementProperties();
      chemBoundary();
      unnorm();
      pressure();
      nodesVelocity();
      nodesPosition();
      if (use_anchor) anchorBottom();
      terms();
    }
    setZero();
  }
  // Gel::stepChemistry, gel.cu:700-704
  void stepChemistry(int /*iter*/) { chem(); }

  // recordCenterElementD, gel_kernels.cu:823-835
  void recordCenterElement(int slot) {
    int gi = elem(L.x / 2, L.y / 2, L.z / 2);
    int hi = node((L.x + 1) / 2, (L.y + 1) / 2, (L.z + 1) / 2);
    vm_center[slot] = vm[gi];
    wm_center[slot] = wm[gi];
    rm_center[slot] = rn[hi];
    Fn_center[slot] = Fn[hi];
    Veln_center[slot] = Veln[hi];
  }
  // solveBilinear, gel_kernels.cu:165-214
  static bool solveBilinear(double A, double B, double C, double D, double E, double F, double G, double H,
                            double vIso, double& xb, double& yb, double eps = 1e-12) {
    const double a00 = A - vIso, a10 = B - A, a01 = C - A, a11 = A - B + D - C;
    const double b00 = E - vIso, b10 = F - E, b01 = G - E, b11 = E - F + H - G;
    const double c2 = b10 * a11 - b11 * a10;
    const double c1 = b00 * a11 + b10 * a01 - b01 * a10 - b11 * a00;
    const double c0 = b00 * a01 - b01 * a00;
    double roots[2];
    int n = 0;
    if (std::fabs(c2) < eps) {
      if (std::fabs(c1) < eps) return false;
      roots[n++] = -c0 / c1;
    } else {
      double disc = c1 * c1 - 4.0 * c2 * c0;
      if (disc < 0.0) return false;
      double s = std::sqrt(disc);
      roots[n++] = (-c1 + s) / (2.0 * c2);
      roots[n++] = (-c1 - s) / (2.0 * c2);
    }
    for (int i = 0; i < n; ++i) {
      double x = roots[i];
      if (x < -eps || x > 1.0 + eps) continue;
      double denom = a01 + a11 * x;
      if (std::fabs(denom) < eps) continue;
      double y = -(a00 + a10 * x) / denom;
      if (y < -eps || y > 1.0 + eps) continue;
      xb = x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
      yb = y < 0.0 ? 0.0 : (y > 1.0 ? 1.0 : y);
      return true;
    }
    return false;
  }
  // calFilamentD, gel_kernels.cu:837-866, with hits in element order
  void filamentPoints(int slot) {
    int cnt = 0;
    D3* out = &filament[(size_t)slot * gp.maxFilamentlen];
    for (int zi = 1; zi < L.z; zi++)
      for (int yi = 1; yi < L.y; yi++)
        for (int xi = 1; xi < L.x; xi++) {
          double xb, yb;
          bool ok = solveBilinear(un_norm[node(xi, yi, zi)], un_norm[node(xi + 1, yi, zi)],
                                  un_norm[node(xi, yi + 1, zi)], un_norm[node(xi + 1, yi + 1, zi)],
                                  vn_norm[node(xi, yi, zi)], vn_norm[node(xi + 1, yi, zi)],
                                  vn_norm[node(xi, yi + 1, zi)], vn_norm[node(xi + 1, yi + 1, zi)], 0.15, xb, yb);
          if (ok && cnt < gp.maxFilamentlen) {
            out[cnt].x = xi + xb;
            out[cnt].y = yi + yb;
            out[cnt].z = zi;
            cnt++;
          }
        }
    filament_count = cnt;
  }
  // Gel::recordCenter, gel.cu:706-713
  void recordCenter(int iter) {
    if (iter % 1000 == 0) {
      int slot = (iter / 1000) % 1000;
      recordCenterElement(slot);
      filamentPoints(slot);
    }
  }

  // ---- state ---------------------------------------------------------------
  int id, type;
  D3 pos;
  double dt, dx, dy, dz;
  int df;
  I3 L;
  int nE, nN, M;
  GelParams gp;
  bool use_anchor = false;
  double anchor_z = 0.0;
  std::vector<double> um, vm, wm, um_norm, vm_norm, wmp, T0, T1, T2, vol, pm;
  std::vector<double> h_um, h_vm, h_wm;
  std::vector<D3> rm, rm_loc, nmSm;
  std::vector<D3> rn, Veln, Vels, Fn, Fn_robin, Fdrag_robin;
  std::vector<double> un_norm, un_robin, vn_norm;
  std::vector<int> map_node, map_element, bIndex;
  std::vector<double> vm_center, wm_center;
  std::vector<D3> rm_center, Fn_center, Veln_center, filament;
  int filament_count = 0;
};

// ===========================================================================
// Coupler.  src/coupling.h, src/coupling.cu, src/coupling_kernels.cu
// ===========================================================================
template <class R>
class Coupler {
 public:
  using R3 = V3<R>;
  // Coupler::Coupler + setInitValue, coupling.cu:50-80, 39-49
  explicit Coupler(std::vector<Gel*> g) : gels(std::move(g)) {
    numGels = (int)gels.size();
    M = 0;
    for (auto* gl : gels) {
      offsets.push_back(M);
      for (int l = 0; l < gl->M; ++l) owner.push_back((int)offsets.size() - 1);
      M += gl->M;
    }
    delta = (R)1e-4f;
    gamma = (R)1.5f;
    gel_to_lbm_vel = 1;
    lbm_to_gel_vel = 1;
    lag.assign(M, R3{0, 0, 0}); Ul.assign(M, R3{0, 0, 0}); Vl.assign(M, R3{0, 0, 0});
    Fdrag.assign(M, R3{0, 0, 0}); Frep.assign(M, R3{0, 0, 0});
    Cl.assign(M, 0); Dl.assign(M, 0); Sl.assign(M, 0);
  }
  // packFromGels + k_gather_boundary, coupling.cu:82-98, coupling_kernels.cu:42-59
  void packFromGels() {
    for (int g = 0; g < numGels; ++g) {
      Gel* gl = gels[g];
      for (int l = 0; l < gl->M; ++l) {
        int id = gl->bIndex[l];
        D3 r = gl->rn[id], v = gl->Vels[id];
        int o = offsets[g] + l;
        lag[o] = R3{(R)r.x, (R)r.y, (R)r.z};
        Vl[o] = R3{(R)(v.x * gel_to_lbm_vel), (R)(v.y * gel_to_lbm_vel), (R)(v.z * gel_to_lbm_vel)};
        Cl[o] = (R)gl->un_norm[id];
      }
    }
  }
  // scatterToGels + k_add_reaction_to_gel / k_add_drag_to_gel,
  // coupling.cu:100-118, coupling_kernels.cu:8-34
  void scatterToGels() {
    for (int g = 0; g < numGels; ++g) {
      Gel* gl = gels[g];
      for (int l = 0; l < gl->M; ++l) {
        int id = gl->bIndex[l];
        int o = offsets[g] + l;
        gl->Fn_robin[id].x -= (double)Frep[o].x;
        gl->Fn_robin[id].y -= (double)Frep[o].y;
        gl->Fn_robin[id].z -= (double)Frep[o].z;
        gl->un_robin[id] -= (double)Sl[o];
      }
      for (int l = 0; l < gl->M; ++l) {
        int id = gl->bIndex[l];
        int o = offsets[g] + l;
        gl->Fdrag_robin[id].x -= (double)Fdrag[o].x;
        gl->Fdrag_robin[id].y -= (double)Fdrag[o].y;
        gl->Fdrag_robin[id].z -= (double)Fdrag[o].z;
      }
    }
  }
  // applyGelRepulsion + k_gel_repulsion/bodyBodyInteraction,
  // coupling.cu:120-123, coupling_kernels.cu:61-146 (pair order b = 0..M-1)
  void applyGelRepulsion() {
    const R rcut = (R)2.5 * gamma, rc2 = rcut * rcut, sig2 = gamma * gamma;
    for (int a = 0; a < M; ++a) {
      R3 acc{0, 0, 0};
      for (int b = 0; b < M; ++b) {
        if (owner[a] == owner[b]) continue;
        R dx = lag[a].x - lag[b].x, dy = lag[a].y - lag[b].y, dz = lag[a].z - lag[b].z;
        R d2 = dx * dx + dy * dy + dz * dz;
        if (d2 <= (R)0 || d2 >= rc2) continue;
        R inv = (R)1 / d2;
        R sr2 = sig2 * inv;
        R sr6 = sr2 * sr2 * sr2;
        R sr12 = sr6 * sr6;
        R fac = (R)24 * delta * ((R)2 * sr12 - sr6) * inv;
        acc.x += fac * dx;
        acc.y += fac * dy;
        acc.z += fac * dz;
      }
      Frep[a] = R3{(R)0 + acc.x, (R)0 + acc.y, (R)0 + acc.z};
    }
  }

  std::vector<Gel*> gels;
  int numGels, M;
  std::vector<int> offsets, owner;
  R delta, gamma, gel_to_lbm_vel, lbm_to_gel_vel;
  std::vector<R3> lag, Ul, Vl, Fdrag, Frep;
  std::vector<R> Cl, Dl, Sl;
};

// ===========================================================================
// Fluid.  src/fluid.h, src/fluid.cu, src/fluid_kernels.cu
// ===========================================================================
static const int kC[19][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},  {0, -1, 0},
                              {0, 0, 1},  {0, 0, -1},  {1, 1, 0},   {-1, 1, 0}, {-1, -1, 0},
                              {1, -1, 0}, {1, 0, 1},   {-1, 0, 1},  {-1, 0, -1}, {1, 0, -1},
                              {0, 1, 1},  {0, 1, -1},  {0, -1, -1}, {0, -1, 1}};  // fluid.cu:202-208
static const int kOpp[19] = {0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16};  // fluid.cu:215

// delta4, fluid_kernels.cu:60-71 (Peskin 4-point)
template <class R>
static inline R delta4(R r) {
  r = std::fabs(r);
  if (r <= (R)1) return (R)0.125 * ((R)3 - (R)2 * r + std::sqrt((R)1 + (R)4 * r - (R)4 * r * r));
  if (r <= (R)2) return (R)0.125 * ((R)5 - (R)2 * r - std::sqrt((R)-7 + (R)12 * r - (R)4 * r * r));
  return (R)0;
}

template <class R>
class Fluid {
 public:
  using R3 = V3<R>;
  // Fluid::Fluid, fluid.cu:152-232 (constants evaluated in R)
  Fluid(int sx, int sy, int sz, Coupler<R>* cp) : coupler(cp) {
    dt = (R)1e-3;
    h = (R)0.5;
    Nx = int(sx / h) + 1;
    Ny = int(sy / h) + 1;
    Nz = int(sz / h) + 1;
    N = Nx * Ny * Nz;
    perX = perY = perZ = 0;
    cs2 = (R)1 / (R)3;
    tau = (R)0.8;
    nu = cs2 * (tau - (R)0.5);
    M = coupler->M;
    pdx = pdy = pdz = 1;
    D = 1;
    R niu = (R)1e-6;
    dx_fluid = (R)40e-6 * h;
    dt_fluid = (tau - (R)0.5) * dx_fluid * dx_fluid / (3 * niu);
    Nsub = std::max(1, (int)std::llround(dt / dt_fluid));
    beta = (R)1.0;
    if (const char* e = std::getenv("IBM_BETA")) beta = (R)std::atof(e);
    ramp_steps = (R)20000.0;
    if (const char* e = std::getenv("IBM_RAMP_STEPS")) ramp_steps = std::max((R)0, (R)std::atof(e));
    w[0] = (R)1 / (R)3;
    for (int q = 1; q < 7; ++q) w[q] = (R)1 / (R)18;
    for (int q = 7; q < 19; ++q) w[q] = (R)1 / (R)36;
    Fc = R3{0, 0, 0};
    const R g2l = h * dt_fluid / dt;
    coupler->gel_to_lbm_vel = g2l;
    coupler->lbm_to_gel_vel = (R)1 / g2l;
    for (auto* g : coupler->gels) g->gp.lbm_to_gel_vel = (float)((R)1 / g2l);
    size_t Nd = (size_t)N * 19;
    f.assign(Nd, 0); fpost.assign(Nd, 0); fnext.assign(Nd, 0);
    rho.assign(N, 0); c1.assign(N, 0); c2.assign(N, 0); source.assign(N, 0);
    u.assign(N, R3{0, 0, 0}); F.assign(N, R3{0, 0, 0}); Fibm.assign(N, R3{0, 0, 0});
    Ftot.assign(N, R3{0, 0, 0});
    init();
  }

  // k_init, fluid_kernels.cu:84-96
  void init() {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      rho[i] = 1;
      u[i] = R3{0, 0, 0};
      c1[i] = 0;
      c2[i] = 0;
      for (int q = 0; q < 19; ++q) f[i + q * (size_t)N] = w[q];
    }
  }

  int id3(int x, int y, int z) const { return x + y * Nx + z * Nx * Ny; }
  static int wrap(int i, int n) { return ((i % n) + n) % n; }

  // Stencil range of one axis: clipped to the box (fluid_kernels.cu:262-266)
  // or the full 4 points when the axis is periodic.
  void axisRange(int i0, int n, int per, int& lo, int& hi) const {
    if (per) { lo = i0 - 1; hi = i0 + 2; }
    else { lo = std::max(0, i0 - 1); hi = std::min(n - 1, i0 + 2); }
  }

  // k_ibm_interpolate_velocity / k_ibm_sample_concentration,
  // fluid_kernels.cu:244-281, 350-387.  Generic over the sampled field.
  template <class Get, class Acc>
  void stencilAccumulate(const R3& p, Get get, Acc& acc, R& wsum) const {
    const R gx = p.x / h, gy = p.y / h, gz = p.z / h;
    const int ix = (int)std::floor(gx), iy = (int)std::floor(gy), iz = (int)std::floor(gz);
    int xl, xh, yl, yh, zl, zh;
    axisRange(ix, Nx, perX, xl, xh);
    axisRange(iy, Ny, perY, yl, yh);
    axisRange(iz, Nz, perZ, zl, zh);
    for (int ii = xl; ii <= xh; ++ii) {
      const R phix = delta4<R>(gx - (R)ii);
      for (int jj = yl; jj <= yh; ++jj) {
        const R phiy = delta4<R>(gy - (R)jj);
        for (int kk = zl; kk <= zh; ++kk) {
          const R phiz = delta4<R>(gz - (R)kk);
          const R ww = phix * phiy * phiz;
          get(id3(wrap(ii, Nx), wrap(jj, Ny), wrap(kk, Nz)), ww, acc);
          wsum += ww;
        }
      }
    }
  }

  // k_ibm_spread_forces / k_ibm_spread_concentration, fluid_kernels.cu:300-348,
  // 414-459: first pass sums the weights, second adds val*phi*(1/h^3)/wsum.
  template <class Put>
  void stencilSpread(const R3& p, Put put) const {
    const R gx = p.x / h, gy = p.y / h, gz = p.z / h;
    const int ix = (int)std::floor(gx), iy = (int)std::floor(gy), iz = (int)std::floor(gz);
    int xl, xh, yl, yh, zl, zh;
    axisRange(ix, Nx, perX, xl, xh);
    axisRange(iy, Ny, perY, yl, yh);
    axisRange(iz, Nz, perZ, zl, zh);
    R wsum = 0;
    for (int ii = xl; ii <= xh; ++ii) {
      const R phix = delta4<R>(gx - (R)ii);
      for (int jj = yl; jj <= yh; ++jj) {
        const R phiy = delta4<R>(gy - (R)jj);
        for (int kk = zl; kk <= zh; ++kk) wsum += phix * phiy * delta4<R>(gz - (R)kk);
      }
    }
    if (wsum == (R)0) return;
    const R inv_h3 = (R)(1.0 / (h * h * h));
    const R scale = 1 * inv_h3 / wsum;
    for (int ii = xl; ii <= xh; ++ii) {
      const R phix = delta4<R>(gx - (R)ii);
      for (int jj = yl; jj <= yh; ++jj) {
        const R phiy = delta4<R>(gy - (R)jj);
        for (int kk = zl; kk <= zh; ++kk) {
          const R phiz = delta4<R>(gz - (R)kk);
          const R ww = phix * phiy * phiz * scale;
          put(id3(wrap(ii, Nx), wrap(jj, Ny), wrap(kk, Nz)), ww);
        }
      }
    }
  }

  // fluid.cu:265-270: host FP32 (here R) ramp of the IBM stiffness
  R betaEff(int iter) const {
    R ramp = (ramp_steps <= (R)0) ? (R)1 : std::min((R)1, (R)(iter + 1) / ramp_steps);
    return beta * ramp;
  }

  // One LBM sub-step in the reference's launch order, fluid.cu:273-282.
  void subStep(R beta_eff) {
    auto& cp = *coupler;
    // k_set_force
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) F[i] = Fc;
    // k_ibm_interpolate_velocity + k_scale_negbeta
    for (int l = 0; l < M; ++l) {
      R3 ul{0, 0, 0};
      R wsum = 0;
      stencilAccumulate(cp.lag[l], [&](int id, R ww, R3& a) {
        a.x += u[id].x * ww; a.y += u[id].y * ww; a.z += u[id].z * ww;
      }, ul, wsum);
      cp.Ul[l] = wsum > (R)0 ? R3{ul.x / wsum, ul.y / wsum, ul.z / wsum} : R3{0, 0, 0};
      const R3 d{cp.Vl[l].x - cp.Ul[l].x, cp.Vl[l].y - cp.Ul[l].y, cp.Vl[l].z - cp.Ul[l].z};
      cp.Fdrag[l] = R3{beta_eff * d.x, beta_eff * d.y, beta_eff * d.z};
    }
    // k_ibm_spread_forces (serial, marker order)
    for (int l = 0; l < M; ++l) {
      const R3 fL = cp.Fdrag[l];
      stencilSpread(cp.lag[l], [&](int id, R ww) {
        Fibm[id].x += fL.x * ww; Fibm[id].y += fL.y * ww; Fibm[id].z += fL.z * ww;
      });
    }
    const R inv_tau = (R)1 / tau;
    // k_vec_add + k_macros + k_collide (per cell, fluid_kernels.cu:98-142)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
      Ftot[i] = R3{F[i].x + Fibm[i].x, F[i].y + Fibm[i].y, F[i].z + Fibm[i].z};
      R rh = 0;
      R3 s{0, 0, 0};
      for (int q = 0; q < 19; ++q) {
        R fi = f[i + q * (size_t)N];
        rh += fi;
        s.x += fi * (R)kC[q][0]; s.y += fi * (R)kC[q][1]; s.z += fi * (R)kC[q][2];
      }
      rho[i] = rh;
      const R3 Ft = Ftot[i];
      u[i] = R3{(s.x + (R)0.5 * Ft.x) / rh, (s.y + (R)0.5 * Ft.y) / rh, (s.z + (R)0.5 * Ft.z) / rh};
      const R3 ui = u[i];
      const R uu = ui.x * ui.x + ui.y * ui.y + ui.z * ui.z;
      for (int q = 0; q < 19; ++q) {
        const R cu = (R)kC[q][0] * ui.x + (R)kC[q][1] * ui.y + (R)kC[q][2] * ui.z;
        const R feq = w[q] * rh * ((R)1 + (cu / cs2) + (R)0.5 * (cu * cu) / (cs2 * cs2) - (R)0.5 * (uu / cs2));
        const R ciF = (R)kC[q][0] * Ft.x + (R)kC[q][1] * Ft.y + (R)kC[q][2] * Ft.z;
        const R uF = ui.x * Ft.x + ui.y * Ft.y + ui.z * Ft.z;
        const R Fterm = w[q] * ((R)1 - (R)0.5 * inv_tau) * ((ciF / cs2) + (cu * ciF) / (cs2 * cs2) - (uF / cs2));
        const R fi = f[i + q * (size_t)N];
        fpost[i + q * (size_t)N] = fi - inv_tau * (fi - feq) + Fterm;
      }
    }
    // k_stream_bounce, fluid_kernels.cu:144-165 (push; halfway bounce-back,
    // or wrap on a periodic axis)
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Nz; ++z)
      for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x) {
          int i = id3(x, y, z);
          for (int q = 0; q < 19; ++q) {
            int xn = x + kC[q][0], yn = y + kC[q][1], zn = z + kC[q][2];
            if (perX) xn = wrap(xn, Nx);
            if (perY) yn = wrap(yn, Ny);
            if (perZ) zn = wrap(zn, Nz);
            bool inside = 0 <= xn && xn < Nx && 0 <= yn && yn < Ny && 0 <= zn && zn < Nz;
            if (inside)
              fnext[id3(xn, yn, zn) + q * (size_t)N] = fpost[i + q * (size_t)N];
            else
              fnext[i + kOpp[q] * (size_t)N] = fpost[i + q * (size_t)N];
          }
        }
    std::swap(f, fnext);
    // k_zero_vector(F_ibm)
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) Fibm[i] = R3{0, 0, 0};
  }

  // Fluid::stepVelocity, fluid.cu:263-284
  void stepVelocity(int iter) {
    const R be = betaEff(iter);
    for (int kk = 0; kk < Nsub; ++kk) subStep(be);
  }

  // Fluid::stepConcentration, fluid.cu:251-261
  void stepConcentration(int /*iter*/) {
    auto& cp = *coupler;
    // k_ibm_sample_concentration (writes Dl) + k_robin_boundary
    for (int l = 0; l < M; ++l) {
      R cl = 0, wsum = 0;
      stencilAccumulate(cp.lag[l], [&](int id, R ww, R& a) { a += c1[id] * ww; }, cl, wsum);
      cp.Dl[l] = wsum > (R)0 ? cl / wsum : (R)0;
    }
    for (int l = 0; l < M; ++l) cp.Sl[l] = 1 * (cp.Cl[l] - cp.Dl[l]);
    // k_zero_scalar + k_ibm_spread_concentration
    std::fill(source.begin(), source.end(), (R)0);
    for (int l = 0; l < M; ++l) {
      const R dL = cp.Sl[l];
      stencilSpread(cp.lag[l], [&](int id, R ww) { source[id] += dL * ww; });
    }
    convectionDiffusion();
    std::swap(c1, c2);
  }

  // k_convection_diffusion, fluid_kernels.cu:177-242
  void convectionDiffusion() {
    const R3 hh{pdx, pdy, pdz};
    const R3 ih{(R)1 / hh.x, (R)1 / hh.y, (R)1 / hh.z};
    const R3 ih2{ih.x * ih.x, ih.y * ih.y, ih.z * ih.z};
#pragma omp parallel for schedule(static)
    for (int z = 0; z < Nz; ++z)
      for (int y = 0; y < Ny; ++y)
        for (int x = 0; x < Nx; ++x) {
          const int xp = (x + 1 < Nx) ? x + 1 : (perX ? 0 : x);
          const int xm = (x - 1 >= 0) ? x - 1 : (perX ? Nx - 1 : x);
          const int yp = (y + 1 < Ny) ? y + 1 : (perY ? 0 : y);
          const int ym = (y - 1 >= 0) ? y - 1 : (perY ? Ny - 1 : y);
          const int zp = (z + 1 < Nz) ? z + 1 : (perZ ? 0 : z);
          const int zm = (z - 1 >= 0) ? z - 1 : (perZ ? Nz - 1 : z);
          const int idx = id3(x, y, z);
          const int ixp = id3(xp, y, z), ixm = id3(xm, y, z);
          const int iyp = id3(x, yp, z), iym = id3(x, ym, z);
          const int izp = id3(x, y, zp), izm = id3(x, y, zm);
          const R c = c1[idx];
          const R cxp = c1[ixp], cxm = c1[ixm], cyp = c1[iyp], cym = c1[iym], czp = c1[izp], czm = c1[izm];
          const R uL = (R)0.5 * (u[ixm].x + u[idx].x);
          const R uR = (R)0.5 * (u[idx].x + u[ixp].x);
          const R vL = (R)0.5 * (u[iym].y + u[idx].y);
          const R vR = (R)0.5 * (u[idx].y + u[iyp].y);
          const R wL = (R)0.5 * (u[izm].z + u[idx].z);
          const R wR = (R)0.5 * (u[idx].z + u[izp].z);
          const R cLx = (uL >= (R)0) ? cxm : c, cLy = (vL >= (R)0) ? cym : c, cLz = (wL >= (R)0) ? czm : c;
          const R cRx = (uR >= (R)0) ? c : cxp, cRy = (vR >= (R)0) ? c : cyp, cRz = (wR >= (R)0) ? c : czp;
          const R FLx = uL * cLx, FLy = vL * cLy, FLz = wL * cLz;
          const R FRx = uR * cRx, FRy = vR * cRy, FRz = wR * cRz;
          const R dFx = FRx - FLx, dFy = FRy - FLy, dFz = FRz - FLz;
          const R conv = -(dFx * ih.x + dFy * ih.y + dFz * ih.z);
          const R sx = cxm - (R)2 * c + cxp, sy = cym - (R)2 * c + cyp, sz = czm - (R)2 * c + czp;
          const R diff = D * (sx * ih2.x + sy * ih2.y + sz * ih2.z);
          c2[idx] = c + dt * (conv + diff + source[idx]);
        }
  }

  Coupler<R>* coupler;
  int Nx, Ny, Nz, N, M, Nsub;
  int perX, perY, perZ;
  R dt, h, cs2, tau, nu, pdx, pdy, pdz, D, dx_fluid, dt_fluid, beta, ramp_steps;
  R w[19];
  R3 Fc;
  std::vector<R> f, fpost, fnext, rho, c1, c2, source;
  std::vector<R3> u, F, Fibm, Ftot;
};

// ===========================================================================
// One time step in sim.cpp's program order (sim.cpp:60-111): every gel's
// elasticity + chemistry + centre record, pack, fluid velocity (Nsub LBM
// sub-steps), concentration, repulsion, scatter.
// ===========================================================================
template <class R>
struct System {
  std::vector<Gel*> gels;
  Coupler<R>* coupler = nullptr;
  Fluid<R>* fluid = nullptr;
  void step(int iter) {
    for (auto* g : gels) {
      g->stepElasticity(iter);
      g->stepChemistry(iter);
      g->recordCenter(iter);
    }
    coupler->packFromGels();
    fluid->stepVelocity(iter);
    fluid->stepConcentration(iter);
    coupler->applyGelRepulsion();
    coupler->scatterToGels();
  }
};

}  // namespace glsm_oracle
