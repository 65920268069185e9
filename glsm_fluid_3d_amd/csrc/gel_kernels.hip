// gel_kernels.hip -- FP64 gLSM gel kernels for gfx950.
//
// Built with -ffp-contract=off: every product and sum is rounded separately
// and in the reference's order, so these kernels agree with an FP64 CPU
// restatement to the last bit except where device libm (log, sqrt of a
// quotient) differs from the host's.
//
// The reference launches 9-10 kernels per elasticity step on 8x8x8 blocks
// (gel.cu:686-695).  Here the same phases run as 6 flat-indexed launches:
// phases with no cross-item dependency are fused (no-flux ghosts + pressure;
// node averages + node forces; node advance + bottom anchor).
#include "glsm_internal.hpp"

namespace glsm {
namespace {

constexpr int kThreads = 256;

__device__ inline double3 d3(double x, double y, double z) { return make_double3(x, y, z); }
__device__ inline double3 add3(double3 a, double3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline double3 sub3(double3 a, double3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline double3 scal(double a, double3 b) { return d3(a * b.x, a * b.y, a * b.z); }
__device__ inline double3 mulsc(double3 a, double b) { return d3(a.x * b, a.y * b, a.z * b); }
__device__ inline double3 divsc(double3 a, double b) { return d3(a.x / b, a.y / b, a.z / b); }
__device__ inline double3 had(double3 a, double3 b) { return d3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ inline double3 hdv(double3 a, double3 b) { return d3(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ inline double dot3(double3 a, double3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline void acc3(double3& a, double3 b) { a.x += b.x; a.y += b.y; a.z += b.z; }
__device__ inline double3 cross3(double3 a, double3 b) {
  return d3(a.y * b.z - b.y * a.z, b.x * a.z - a.x * b.z, a.x * b.y - b.x * a.y);
}
__device__ inline double dist3(double3 a, double3 b) {
  double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

struct Range3 {  // inclusive box [lo, hi]^3 (per axis), flattened x fastest
  int x0, y0, z0, ex, ey, ez;
  __device__ bool decode(int t, int& x, int& y, int& z) const {
    if (t >= ex * ey * ez) return false;
    x = x0 + t % ex;
    int r = t / ex;
    y = y0 + r % ey;
    z = z0 + r / ey;
    return true;
  }
  __host__ __device__ int count() const { return ex * ey * ez; }
};
__host__ __device__ inline Range3 box(int lo, int hx, int hy, int hz) {
  Range3 r{lo, lo, lo, hx - lo + 1, hy - lo + 1, hz - lo + 1};
  if (r.ex < 0) r.ex = 0;
  if (r.ey < 0) r.ey = 0;
  if (r.ez < 0) r.ez = 0;
  return r;
}
inline int blocks_for(int n) { return (n + kThreads - 1) / kThreads; }

__device__ inline int elemIdx(const GelView& g, int x, int y, int z) {
  return x + y * (g.LX + 1) + z * (g.LY + 1) * (g.LX + 1);
}
__device__ inline int nodeIdx(const GelView& g, int x, int y, int z) {
  return x + y * (g.LX + 2) + z * (g.LY + 2) * (g.LX + 2);
}

// ghost node = 2*(first real node) - (second real node) along the type's
// offset (calServiceNodesPositionD, gel_kernels.cu:615-631)
__device__ __forceinline__ void service_nodes_at(const GelView& g, int x, int y, int z) {
  const int gi = nodeIdx(g, x, y, z);
  const int t = g.map_node[gi];
  if (t == 0) return;
  const int ox = g.off[t][0], oy = g.off[t][1], oz = g.off[t][2];
  const double3 r1 = g.rn[nodeIdx(g, x + ox, y + oy, z + oz)];
  const double3 r2 = g.rn[nodeIdx(g, x + 2 * ox, y + 2 * oy, z + 2 * oz)];
  g.rn[gi] = sub3(scal(2.0, r1), r2);
}

// trilinear shape-function derivative dN_i/d(axis) at local point (p0,p1,p2);
// node i = 4*ix + 2*iy + iz.  Exact dyadic values (folded at compile time).
__device__ inline double dshape(int i, int axis, double p0, double p1, double p2) {
  const double s0 = double(2 * ((i >> 2) & 1) - 1), s1 = double(2 * ((i >> 1) & 1) - 1),
               s2 = double(2 * (i & 1) - 1);
  double v = axis == 0 ? s0 : (axis == 1 ? s1 : s2);
  if (axis != 0) v *= (1 + p0 * s0);
  if (axis != 1) v *= (1 + p1 * s1);
  if (axis != 2) v *= (1 + p2 * s2);
  return v / 8;
}

// centroid, local node offsets, six face area vectors (2x2 midpoint rule of
// the Jacobian cross products), volume (2x2x2 rule of det J), wm update
// (calElementPropertiesD, gel_kernels.cu:216-483)
__device__ __forceinline__ void element_props_at(const GelView& g, int x, int y, int z) {
  const int gi = elemIdx(g, x, y, z);
  double3 nd[8];
  double3 bc = d3(0, 0, 0);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    nd[c] = g.rn[nodeIdx(g, x + (c >> 2), y + ((c >> 1) & 1), z + (c & 1))];
    acc3(bc, nd[c]);
  }
  bc = divsc(bc, 8);
  g.rm[gi] = bc;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    nd[c] = sub3(nd[c], bc);
    g.rm_loc[8 * (size_t)gi + c] = nd[c];
  }
  // face f: nodes (in-face order), outer/inner quadrature axes, fixed axis
  // and its value, stored sign (outward normal)
  constexpr int fn[6][4] = {{0, 1, 3, 2}, {4, 6, 7, 5}, {0, 4, 5, 1}, {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 5, 7, 3}};
  constexpr int fa[6][3] = {{1, 2, 0}, {1, 2, 0}, {0, 2, 1}, {0, 2, 1}, {0, 1, 2}, {0, 1, 2}};
  constexpr double fv[6] = {-1, 1, -1, 1, -1, 1};
  constexpr bool fneg[6] = {true, false, false, true, true, false};
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    double3 dJ = d3(0, 0, 0);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        double p[3];
        p[fa[f][2]] = fv[f];
        p[fa[f][0]] = a ? 0.5 : -0.5;
        p[fa[f][1]] = b ? 0.5 : -0.5;
        double3 J0 = d3(0, 0, 0), J1 = d3(0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int n = fn[f][i];
          acc3(J0, mulsc(nd[n], dshape(n, fa[f][0], p[0], p[1], p[2])));
          acc3(J1, mulsc(nd[n], dshape(n, fa[f][1], p[0], p[1], p[2])));
        }
        acc3(dJ, cross3(J0, J1));
      }
    g.nmSm[6 * (size_t)gi + f] = fneg[f] ? d3(-dJ.x, -dJ.y, -dJ.z) : dJ;
  }
  double detv = 0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const double p0 = a ? 0.5 : -0.5, p1 = b ? 0.5 : -0.5, p2 = c ? 0.5 : -0.5;
        double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int i = 0; i < 8; ++i) {
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            const double d = dshape(i, ax, p0, p1, p2);
            J[ax][0] += d * nd[i].x;
            J[ax][1] += d * nd[i].y;
            J[ax][2] += d * nd[i].z;
          }
        }
        detv += (J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
                 J[0][1] * (J[1][0] * J[2][2] - J[2][0] * J[1][2]) +
                 J[0][2] * (J[1][0] * J[2][1] - J[2][0] * J[1][1]));
      }
  g.vol[gi] = detv;
  g.wmp[gi] = g.wm[gi];
  g.wm[gi] = g.dx * g.dy * g.dz * g.FA0 / detv;
}

// no-flux ghost elements copy their interior neighbour; normalised u,v for
// every element (calChemBoundaryD, gel_kernels.cu:757-778); interior
// elements also get their osmotic pressure (calPressureD,
// gel_kernels.cu:496-516), which only reads the element's own w and v.
__device__ __forceinline__ void chem_boundary_pressure_at(const GelView& g, int x, int y, int z) {
  const int gi = elemIdx(g, x, y, z);
  const int t = g.map_element[gi];
  const int g1 = elemIdx(g, x + g.off[t][0], y + g.off[t][1], z + g.off[t][2]);
  const double u = g.um[g1], v = g.vm[g1];
  g.um[gi] = u;
  g.vm[gi] = v;
  const double w = g.wm[gi];
  g.um_norm[gi] = u / (1 - w);
  g.vm_norm[gi] = v / (1 - w);
  if (x >= 1 && y >= 1 && z >= 1 && x < g.LX && y < g.LY && z < g.LZ) {
    if (g.gelType == 1) {
      g.pm[gi] = -(w + log(1.0 - w) + (g.CH0 + g.CH1 * w) * w * w) + g.C0 * w / (2.0 * g.FA0) + g.CHS * w * v;
    } else if (g.gelType == 2) {
      const double c0 = g.c0_bis + g.b * v;
      g.pm[gi] = -(w + log(1.0 - w) + (g.CH0 + g.CH1 * w) * w * w) + c0 * w / (2.0 * g.FA0);
    }
  }
}

// node averages of the normalised chemistry (calUnnormD,
// gel_kernels.cu:795-821) then the node force, mobility and velocities
// (calNodesVelocityD, gel_kernels.cu:518-598).  Both are per real node.
__device__ __forceinline__ void unnorm_velocity_at(const GelView& g, int x, int y, int z) {
  const int gi = nodeIdx(g, x, y, z);
  double su = 0, sv = 0, wn = 0;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int e = elemIdx(g, x - i, y - j, z - k);
        su += g.um_norm[e];
        sv += g.vm_norm[e];
        wn += g.wm[e];
      }
  g.un_norm[gi] = su / 8 + g.un_robin[gi];
  const double vn = sv / 8;
  g.vn_norm[gi] = vn;

  wn /= 8;
  const double Mn = 8. * g.AZ0 * sqrt(g.FA0 / wn) * (1 - wn) / (g.dx * g.dy * g.dz);
  double3 f1 = d3(0., 0., 0.), f2 = d3(0., 0., 0.);
  const double3 r0 = g.rn[gi];
  // the eight adjacent elements, +side before -side, x outer, z inner
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int sx = a ? -1 : 1, sy = b ? -1 : 1, sz = c ? -1 : 1;
        const bool ok = (a ? x > 1 : x < g.LX) && (b ? y > 1 : y < g.LY) && (c ? z > 1 : z < g.LZ);
        if (!ok) continue;
        const int e = elemIdx(g, x - a, y - b, z - c);
        const double3 t = add3(add3(add3(g.rn[nodeIdx(g, x + sx, y + sy, z + sz)], g.rn[nodeIdx(g, x, y + sy, z + sz)]),
                                    g.rn[nodeIdx(g, x + sx, y, z + sz)]),
                               g.rn[nodeIdx(g, x + sx, y + sy, z)]);
        acc3(f1, sub3(t, scal(4.0, r0)));
        const size_t fb = 6 * (size_t)e;
        const double3 s3 = add3(add3(g.nmSm[fb + a], g.nmSm[fb + 2 + b]), g.nmSm[fb + 4 + c]);
        acc3(f2, scal(g.pm[e], s3));
      }
  // pow(dx*dy*dz, 1/3) with integer 1/3 (== pow(., 0) == 1), gel_kernels.cu:587
  if (g.gelType == 1) {
    f1 = mulsc(f1, 1.0 * g.C0 / 12.);
  } else if (g.gelType == 2) {
    const double c0 = g.c0_bis + g.b * vn * (1 - wn);
    f1 = mulsc(f1, 1.0 * c0 / 12.);
  }
  f2 = divsc(f2, 4.);
  const double3 Fn = add3(f1, f2);
  const double3 ven = scal(Mn, Fn);
  g.Fn[gi] = Fn;
  g.Veln[gi] = ven;
  g.Vels[gi] = divsc(scal(-wn, ven), (1 - wn));
}

// rn += dtx * ven (calInternalNodesPositionD, gel_kernels.cu:600-613), then
// the optional bottom anchor (anchorBottomNodesD, gel_kernels.cu:870-880).
// The reference launches the anchor on gridDim0 with no +1 shift
// (gel.cu:694), so it reaches x, y only below 8*ceil(L/8); kept here.
__device__ __forceinline__ void position_anchor_at(const GelView& g, int x, int y, int z) {
  const int gi = nodeIdx(g, x, y, z);
  double3 p = g.rn[gi];
  acc3(p, scal(g.dtx, g.Veln[gi]));
  if (g.use_anchor && z == 1 && x < 8 * ((g.LX + 7) / 8) && y < 8 * ((g.LY + 7) / 8)) {
    g.Veln[gi] = d3(0.0, 0.0, 0.0);
    p.z = g.anchor_z;
  }
  g.rn[gi] = p;
}

// chemistry source terms T0 (dilation), T1 (advective flux through the six
// faces), T2 (non-uniform 3-point gradient/Laplacian) (calTermsD,
// gel_kernels.cu:633-703)
__device__ __forceinline__ void terms_at(const GelView& g, int x, int y, int z) {
  const int gi = elemIdx(g, x, y, z);
  const double w0 = g.wm[gi], u0 = g.um_norm[gi];
  const double3 r0 = g.rm[gi];
  g.T0[gi] = (1 - w0 / g.wmp[gi]) / g.dtx;
  double3 vmum = d3(0, 0, 0);
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int i = c >> 2, j = (c >> 1) & 1, k = c & 1;
    const double3 rl = g.rm_loc[8 * (size_t)gi + c];
    const int n = nodeIdx(g, x + i, y + j, z + k);
    const double N = 0.125 * (1 + (2 * double(i) - 1) * rl.x) * (1 + (2 * double(j) - 1) * rl.y) *
                     (1 + (2 * double(k) - 1) * rl.z);
    acc3(vmum, mulsc(scal(N, g.Veln[n]), g.un_norm[n]));
  }
  double dT1 = 0;
#pragma unroll
  for (int f = 0; f < 6; ++f) dT1 += dot3(g.nmSm[6 * (size_t)gi + f], vmum);
  g.T1[gi] = dT1 / g.vol[gi];

  const int g1 = elemIdx(g, x + 1, y, z), g2 = elemIdx(g, x - 1, y, z);
  const int g3 = elemIdx(g, x, y + 1, z), g4 = elemIdx(g, x, y - 1, z);
  const int g5 = elemIdx(g, x, y, z + 1), g6 = elemIdx(g, x, y, z - 1);
  const double3 aa = d3(dist3(g.rm[g1], r0), dist3(g.rm[g3], r0), dist3(g.rm[g5], r0));
  const double3 as = d3(dist3(g.rm[g2], r0), dist3(g.rm[g4], r0), dist3(g.rm[g6], r0));
  const double3 aa2 = had(aa, aa), as2 = had(as, as);
  const double3 aas = had(had(aa, as), add3(aa, as));
  const double3 wa = d3(g.wm[g1], g.wm[g3], g.wm[g5]), ws = d3(g.wm[g2], g.wm[g4], g.wm[g6]);
  const double3 ua = d3(g.um_norm[g1], g.um_norm[g3], g.um_norm[g5]);
  const double3 us = d3(g.um_norm[g2], g.um_norm[g4], g.um_norm[g6]);
  const double3 dw = hdv(sub3(add3(scal(w0, sub3(aa2, as2)), had(wa, as2)), had(ws, aa2)), aas);
  const double3 du = hdv(sub3(add3(scal(u0, sub3(aa2, as2)), had(ua, as2)), had(us, aa2)), aas);
  const double3 d2u = hdv(scal(2.0, sub3(add3(had(ua, as), had(us, aa)), scal(u0, add3(aa, as)))), aas);
  g.T2[gi] = -dot3(dw, du) + (1 - w0) * (d2u.x + d2u.y + d2u.z);
}

__device__ __forceinline__ void set_zero_at(const GelView& g, int x, int y, int z) {
  const int gi = nodeIdx(g, x, y, z);
  g.un_robin[gi] = 0;
  g.Fn_robin[gi] = d3(0., 0., 0.);
}

// isVoidElement, gel_kernels.cu:104-117 (float arithmetic, as there)
__device__ inline bool void_element(const GelView& g, int y, int z) {
  if (g.tube_mode == 0) return false;
  const float dy = (float)y - g.tube_cy, dz = (float)z - g.tube_cz;
  if (g.tube_mode == 1) return fabsf(dy) < g.tube_inner_hy && fabsf(dz) < g.tube_inner_hz;
  const float rhy = g.tube_inner_hy, rhz = g.tube_inner_hz;
  if (rhy < 0.5f || rhz < 0.5f) return false;
  return (dy * dy) / (rhy * rhy) + (dz * dz) / (rhz * rhz) < 1.0f;
}

// forward-Euler modified Oregonator (calChemD + fu/fv, gel_kernels.cu:705-755,
// 129-145; q, P1, P2 are the kernels' literals)
__device__ __forceinline__ void chem_at(const GelView& g, int x, int y, int z) {
  const int gi = elemIdx(g, x, y, z);
  if (void_element(g, y, z)) {
    g.vm[gi] = g.vss;
    g.um[gi] = g.uss;
    g.wm[gi] = g.wss;
    return;
  }
  const double I = 0;
  const double P1 = 0.0124, P2 = 0.77, q = 1e-4;
  const double v = g.vm[gi], u = g.um[gi], w = g.wm[gi];
  const double fvv = (1.0 - w) * (1.0 - w) * u - (1.0 - w) * v + I * (0.5 * P1 + P2);
  const double ww = (1.0 - w) * (1.0 - w);
  const double fuu = ww * u - u * u - (1.0 - w) * (g.f * v + I * P1) * (u - q * ww) / (u + q * ww) + I * P2;
  const double t0 = g.T0[gi];
  g.vm[gi] = v + g.dt * (-v * t0 + g.ep * fvv);
  g.um[gi] = u + g.dt * (-u * t0 + g.T1[gi] + g.T2[gi] + fuu);
}

// solveBilinear, gel_kernels.cu:165-214
__device__ bool solve_bilinear(double A, double B, double C, double D, double E, double F, double G, double H,
                               double vIso, double& xb, double& yb) {
  const double eps = 1e-12;
  const double a00 = A - vIso, a10 = B - A, a01 = C - A, a11 = A - B + D - C;
  const double b00 = E - vIso, b10 = F - E, b01 = G - E, b11 = E - F + H - G;
  const double c2 = b10 * a11 - b11 * a10;
  const double c1 = b00 * a11 + b10 * a01 - b01 * a10 - b11 * a00;
  const double c0 = b00 * a01 - b01 * a00;
  double roots[2];
  int n = 0;
  if (fabs(c2) < eps) {
    if (fabs(c1) < eps) return false;
    roots[n++] = -c0 / c1;
  } else {
    const double disc = c1 * c1 - 4.0 * c2 * c0;
    if (disc < 0.0) return false;
    const double s = sqrt(disc);
    roots[n++] = (-c1 + s) / (2.0 * c2);
    roots[n++] = (-c1 - s) / (2.0 * c2);
  }
  for (int i = 0; i < n; ++i) {
    const double xx = roots[i];
    if (xx < -eps || xx > 1.0 + eps) continue;
    const double denom = a01 + a11 * xx;
    if (fabs(denom) < eps) continue;
    const double yy = -(a00 + a10 * xx) / denom;
    if (yy < -eps || yy > 1.0 + eps) continue;
    xb = xx < 0.0 ? 0.0 : (xx > 1.0 ? 1.0 : xx);
    yb = yy < 0.0 ? 0.0 : (yy > 1.0 ? 1.0 : yy);
    return true;
  }
  return false;
}

// centre sample into the ring slot (recordCenterElementD,
// gel_kernels.cu:823-835) and the u/v iso-line points (calFilamentD,
// gel_kernels.cu:837-866).  One work-group; hits are compacted in element
// order with a block prefix sum (the reference's atomic counter gives an
// arbitrary order) and bounded by maxFilamentlen.
__device__ __forceinline__ void record_center_block(const GelView& g, int slot) {
  __shared__ int s_scan[kThreads];
  __shared__ int s_base;
  if (threadIdx.x == 0) {
    const int gi = elemIdx(g, g.LX / 2, g.LY / 2, g.LZ / 2);
    const int hi = nodeIdx(g, (g.LX + 1) / 2, (g.LY + 1) / 2, (g.LZ + 1) / 2);
    g.vm_center[slot] = g.vm[gi];
    g.wm_center[slot] = g.wm[gi];
    g.rm_center[slot] = g.rn[hi];
    g.Fn_center[slot] = g.Fn[hi];
    g.Veln_center[slot] = g.Veln[hi];
    s_base = 0;
  }
  __syncthreads();
  const int ex = g.LX - 1, ey = g.LY - 1, ez = g.LZ - 1;
  const int total = ex * ey * ez;
  double3* out = g.filament + (size_t)slot * g.maxFilamentlen;
  for (int base = 0; base < total; base += kThreads) {
    const int t = base + threadIdx.x;
    bool ok = false;
    double xb = 0, yb = 0;
    int x = 0, y = 0, z = 0;
    if (t < total) {
      x = 1 + t % ex;
      y = 1 + (t / ex) % ey;
      z = 1 + t / (ex * ey);
      ok = solve_bilinear(g.un_norm[nodeIdx(g, x, y, z)], g.un_norm[nodeIdx(g, x + 1, y, z)],
                          g.un_norm[nodeIdx(g, x, y + 1, z)], g.un_norm[nodeIdx(g, x + 1, y + 1, z)],
                          g.vn_norm[nodeIdx(g, x, y, z)], g.vn_norm[nodeIdx(g, x + 1, y, z)],
                          g.vn_norm[nodeIdx(g, x, y + 1, z)], g.vn_norm[nodeIdx(g, x + 1, y + 1, z)], 0.15, xb, yb);
    }
    s_scan[threadIdx.x] = ok ? 1 : 0;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {  // inclusive Hillis-Steele scan
      const int v = threadIdx.x >= o ? s_scan[threadIdx.x - o] : 0;
      __syncthreads();
      s_scan[threadIdx.x] += v;
      __syncthreads();
    }
    const int pos = s_base + s_scan[threadIdx.x] - 1;
    if (ok && pos < g.maxFilamentlen) out[pos] = d3(x + xb, y + yb, (double)z);
    __syncthreads();
    if (threadIdx.x == kThreads - 1) s_base += s_scan[kThreads - 1];
    __syncthreads();
  }
}

// ---- launch shapes ----
// Each phase runs over one box of a gel's node or element array; the
// single-gel kernel takes the box, the batched kernel derives it per gel.
enum GelOp { OP_SERVICE, OP_PROPS, OP_CHEMB, OP_UNNORM, OP_POS, OP_TERMS, OP_SETZERO, OP_CHEM };

template <int OP>
__device__ __forceinline__ void gel_op(const GelView& g, int x, int y, int z) {
  if constexpr (OP == OP_SERVICE) service_nodes_at(g, x, y, z);
  if constexpr (OP == OP_PROPS) element_props_at(g, x, y, z);
  if constexpr (OP == OP_CHEMB) chem_boundary_pressure_at(g, x, y, z);
  if constexpr (OP == OP_UNNORM) unnorm_velocity_at(g, x, y, z);
  if constexpr (OP == OP_POS) position_anchor_at(g, x, y, z);
  if constexpr (OP == OP_TERMS) terms_at(g, x, y, z);
  if constexpr (OP == OP_SETZERO) set_zero_at(g, x, y, z);
  if constexpr (OP == OP_CHEM) chem_at(g, x, y, z);
}

// nodes_all, elements_all, real nodes, interior elements
template <int OP>
__host__ __device__ inline Range3 op_range(int LX, int LY, int LZ) {
  if (OP == OP_SERVICE) return box(0, LX + 1, LY + 1, LZ + 1);
  if (OP == OP_PROPS || OP == OP_CHEMB) return box(0, LX, LY, LZ);
  if (OP == OP_UNNORM || OP == OP_POS || OP == OP_SETZERO) return box(1, LX, LY, LZ);
  return box(1, LX - 1, LY - 1, LZ - 1);
}

template <int OP>
__global__ void __launch_bounds__(kThreads) k_gel(GelView g, Range3 r) {
  int x, y, z;
  if (!r.decode(blockIdx.x * blockDim.x + threadIdx.x, x, y, z)) return;
  gel_op<OP>(g, x, y, z);
}

// All gels of a batch in one launch: blockIdx.y selects the gel; a gel
// takes part when its instance's iteration is a multiple of `cadence`.
template <int OP>
__global__ void __launch_bounds__(kThreads) k_gel_batch(const GelView* __restrict__ views,
                                                        const int* __restrict__ gel_inst,
                                                        const int* __restrict__ iters, int cadence) {
  const GelView& g = views[blockIdx.y];
  if (cadence > 1 && iters[gel_inst[blockIdx.y]] % cadence != 0) return;
  const Range3 r = op_range<OP>(g.LX, g.LY, g.LZ);
  int x, y, z;
  if (!r.decode(blockIdx.x * blockDim.x + threadIdx.x, x, y, z)) return;
  gel_op<OP>(g, x, y, z);
}

__global__ void __launch_bounds__(kThreads) k_record_center(GelView g, int slot) { record_center_block(g, slot); }

// every 1000th iteration of the gel's instance: slot (iter / 1000) % 1000
__global__ void __launch_bounds__(kThreads) k_record_center_batch(const GelView* __restrict__ views,
                                                                  const int* __restrict__ gel_inst,
                                                                  const int* __restrict__ iters) {
  const int it = iters[gel_inst[blockIdx.x]];
  if (it % 1000 != 0) return;
  record_center_block(views[blockIdx.x], (it / 1000) % 1000);
}

template <int OP>
void launch_op(const GelView& g, hipStream_t s) {
  const Range3 r = op_range<OP>(g.LX, g.LY, g.LZ);
  if (r.count() > 0) k_gel<OP><<<blocks_for(r.count()), kThreads, 0, s>>>(g, r);
}

template <int OP>
void launch_op_batch(const GelView* views, const int* gel_inst, const int* iters, int ngel, int max_count,
                     int cadence, hipStream_t s) {
  if (ngel > 0 && max_count > 0)
    k_gel_batch<OP><<<dim3(blocks_for(max_count), ngel), kThreads, 0, s>>>(views, gel_inst, iters, cadence);
}

}  // namespace

void launch_gel_elasticity(const GelView& g, hipStream_t s) {
  launch_op<OP_SERVICE>(g, s);
  launch_op<OP_PROPS>(g, s);
  launch_op<OP_CHEMB>(g, s);
  launch_op<OP_UNNORM>(g, s);
  launch_op<OP_POS>(g, s);
  launch_op<OP_TERMS>(g, s);
  GLSM_LAUNCH_CHECK();
}

void launch_gel_set_zero(const GelView& g, hipStream_t s) {
  launch_op<OP_SETZERO>(g, s);
  GLSM_LAUNCH_CHECK();
}

void launch_gel_chemistry(const GelView& g, hipStream_t s) {
  launch_op<OP_CHEM>(g, s);
  GLSM_LAUNCH_CHECK();
}

void launch_gel_record_center(const GelView& g, int slot, hipStream_t s) {
  k_record_center<<<1, kThreads, 0, s>>>(g, slot);
  GLSM_LAUNCH_CHECK();
}

// gel.cu:682-713 for every gel of a batch: elasticity on iterations that are
// multiples of 5, setZero and chemistry always, the centre record on
// multiples of 1000.  `max_node_grid` is the largest L over the gels.
void launch_gel_step_batch(const GelView* views, const int* gel_inst, const int* iters, int ngel,
                           int max_node_grid, hipStream_t s) {
  const int L = max_node_grid;
  const int nodes_all = (L + 2) * (L + 2) * (L + 2), elems_all = (L + 1) * (L + 1) * (L + 1);
  const int nodes_real = L * L * L, elems_int = (L - 1) * (L - 1) * (L - 1);
  launch_op_batch<OP_SERVICE>(views, gel_inst, iters, ngel, nodes_all, 5, s);
  launch_op_batch<OP_PROPS>(views, gel_inst, iters, ngel, elems_all, 5, s);
  launch_op_batch<OP_CHEMB>(views, gel_inst, iters, ngel, elems_all, 5, s);
  launch_op_batch<OP_UNNORM>(views, gel_inst, iters, ngel, nodes_real, 5, s);
  launch_op_batch<OP_POS>(views, gel_inst, iters, ngel, nodes_real, 5, s);
  launch_op_batch<OP_TERMS>(views, gel_inst, iters, ngel, elems_int, 5, s);
  launch_op_batch<OP_SETZERO>(views, gel_inst, iters, ngel, nodes_real, 1, s);
  launch_op_batch<OP_CHEM>(views, gel_inst, iters, ngel, elems_int, 1, s);
  if (ngel > 0) k_record_center_batch<<<ngel, kThreads, 0, s>>>(views, gel_inst, iters);
  GLSM_LAUNCH_CHECK();
}

// the cadence decided on the host: the kernels read no iteration counters
void launch_gel_step_multi(const GelView* views, int ngel, int max_node_grid, int iter, hipStream_t s) {
  const int L = max_node_grid;
  const int nodes_all = (L + 2) * (L + 2) * (L + 2), elems_all = (L + 1) * (L + 1) * (L + 1);
  const int nodes_real = L * L * L, elems_int = (L - 1) * (L - 1) * (L - 1);
  if (iter % 5 == 0) {
    launch_op_batch<OP_SERVICE>(views, nullptr, nullptr, ngel, nodes_all, 1, s);
    launch_op_batch<OP_PROPS>(views, nullptr, nullptr, ngel, elems_all, 1, s);
    launch_op_batch<OP_CHEMB>(views, nullptr, nullptr, ngel, elems_all, 1, s);
    launch_op_batch<OP_UNNORM>(views, nullptr, nullptr, ngel, nodes_real, 1, s);
    launch_op_batch<OP_POS>(views, nullptr, nullptr, ngel, nodes_real, 1, s);
    launch_op_batch<OP_TERMS>(views, nullptr, nullptr, ngel, elems_int, 1, s);
  }
  launch_op_batch<OP_SETZERO>(views, nullptr, nullptr, ngel, nodes_real, 1, s);
  launch_op_batch<OP_CHEM>(views, nullptr, nullptr, ngel, elems_int, 1, s);
  GLSM_LAUNCH_CHECK();
}

}  // namespace glsm
