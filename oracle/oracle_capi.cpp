// oracle_capi.cpp -- flat C binding of the CPU restatement (glsm_oracle.hpp).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg through oracle/__init__.py.  The entry points
// mirror include/glsm_fluid.h (same field numbering) so a test can drive the
// oracle and the HIP library with the same call sequence.  Fluid/coupler
// handles carry a precision (32 = the reference's FP32 arithmetic, 64 = the
// FP64 oracle); gels are always FP64.
#include <omp.h>

#include <cstring>
#include <stdexcept>
#include <vector>

#include "../include/glsm_fluid.h"
#include "glsm_oracle.hpp"

using namespace glsm_oracle;

struct OCoupler {
  int prec;
  Coupler<float>* f32 = nullptr;
  Coupler<double>* f64 = nullptr;
};
struct OFluid {
  int prec;
  Fluid<float>* f32 = nullptr;
  Fluid<double>* f64 = nullptr;
};

namespace {

template <class T>
int copy_out(const std::vector<T>& v, size_t n, void* out, size_t bytes) {
  if (bytes != n * sizeof(T) || n > v.size()) return GLSM_EINVAL;
  std::memcpy(out, v.data(), bytes);
  return GLSM_OK;
}
template <class T>
int copy_in(std::vector<T>& v, size_t n, const void* in, size_t bytes) {
  if (bytes != n * sizeof(T) || n > v.size()) return GLSM_EINVAL;
  std::memcpy(v.data(), in, bytes);
  return GLSM_OK;
}

// (vector, logical length) of each gel field
int gel_field(Gel* g, int field, void* out, const void* in, size_t bytes) {
  const size_t nE = g->nE, nN = g->nN;
  auto io = [&](auto& v, size_t n) { return out ? copy_out(v, n, out, bytes) : copy_in(v, n, in, bytes); };
  switch (field) {
    case GLSM_GEL_UM: return io(g->um, nE);
    case GLSM_GEL_VM: return io(g->vm, nE);
    case GLSM_GEL_WM: return io(g->wm, nE);
    case GLSM_GEL_UM_NORM: return io(g->um_norm, nE);
    case GLSM_GEL_VM_NORM: return io(g->vm_norm, nE);
    case GLSM_GEL_WMP: return io(g->wmp, nE);
    case GLSM_GEL_T0: return io(g->T0, nE);
    case GLSM_GEL_T1: return io(g->T1, nE);
    case GLSM_GEL_T2: return io(g->T2, nE);
    case GLSM_GEL_VOL: return io(g->vol, nE);
    case GLSM_GEL_PM: return io(g->pm, nE);
    case GLSM_GEL_RM: return io(g->rm, nE);
    case GLSM_GEL_RM_LOC: return io(g->rm_loc, 8 * nE);
    case GLSM_GEL_NMSM: return io(g->nmSm, 6 * nE);
    case GLSM_GEL_RN: return io(g->rn, nN);
    case GLSM_GEL_VELN: return io(g->Veln, nN);
    case GLSM_GEL_VELS: return io(g->Vels, nN);
    case GLSM_GEL_FN: return io(g->Fn, nN);
    case GLSM_GEL_FN_ROBIN: return io(g->Fn_robin, nN);
    case GLSM_GEL_FDRAG_ROBIN: return io(g->Fdrag_robin, nN);
    case GLSM_GEL_UN_NORM: return io(g->un_norm, nN);
    case GLSM_GEL_UN_ROBIN: return io(g->un_robin, nN);
    case GLSM_GEL_VN_NORM: return io(g->vn_norm, nN);
    case GLSM_GEL_MAP_NODE: return io(g->map_node, nN);
    case GLSM_GEL_MAP_ELEMENT: return io(g->map_element, nE);
    case GLSM_GEL_BINDEX: return io(g->bIndex, (size_t)g->M);
    case GLSM_GEL_CENTER_VM: return io(g->vm_center, 1000);
    case GLSM_GEL_CENTER_WM: return io(g->wm_center, 1000);
    case GLSM_GEL_CENTER_RM: return io(g->rm_center, 1000);
    case GLSM_GEL_CENTER_FN: return io(g->Fn_center, 1000);
    case GLSM_GEL_CENTER_VELN: return io(g->Veln_center, 1000);
    case GLSM_GEL_FILAMENT: return io(g->filament, 1000 * (size_t)g->gp.maxFilamentlen);
    default: return GLSM_EINVAL;
  }
}

// Coupler/fluid fields are converted to double for the caller.
template <class T>
int out_double(const std::vector<T>& v, size_t n, double* out, size_t count) {
  if (count != n || n > v.size()) return GLSM_EINVAL;
  for (size_t i = 0; i < n; ++i) out[i] = (double)v[i];
  return GLSM_OK;
}
template <class T>
int out_double3(const std::vector<V3<T>>& v, size_t n, double* out, size_t count) {
  if (count != 3 * n || n > v.size()) return GLSM_EINVAL;
  for (size_t i = 0; i < n; ++i) {
    out[3 * i] = v[i].x; out[3 * i + 1] = v[i].y; out[3 * i + 2] = v[i].z;
  }
  return GLSM_OK;
}

template <class R>
int coupler_read(Coupler<R>* c, int field, double* out, size_t count) {
  const size_t M = c->M;
  switch (field) {
    case GLSM_CP_LAG: return out_double3(c->lag, M, out, count);
    case GLSM_CP_UL: return out_double3(c->Ul, M, out, count);
    case GLSM_CP_VL: return out_double3(c->Vl, M, out, count);
    case GLSM_CP_FDRAG: return out_double3(c->Fdrag, M, out, count);
    case GLSM_CP_FREP: return out_double3(c->Frep, M, out, count);
    case GLSM_CP_CL: return out_double(c->Cl, M, out, count);
    case GLSM_CP_DL: return out_double(c->Dl, M, out, count);
    case GLSM_CP_SL: return out_double(c->Sl, M, out, count);
    case GLSM_CP_OWNER: return out_double(c->owner, M, out, count);
    default: return GLSM_EINVAL;
  }
}

template <class R>
int fluid_read(Fluid<R>* f, int field, double* out, size_t count) {
  const size_t N = f->N;
  switch (field) {
    case GLSM_FL_U: return out_double3(f->u, N, out, count);
    case GLSM_FL_C: return out_double(f->c1, N, out, count);
    case GLSM_FL_F: return out_double(f->f, 19 * N, out, count);
    case GLSM_FL_FPOST: return out_double(f->fpost, 19 * N, out, count);
    default: return GLSM_EINVAL;
  }
}
template <class R>
int fluid_write(Fluid<R>* f, int field, const double* in, size_t count) {
  const size_t N = f->N;
  if (field == GLSM_FL_C) {
    if (count != N) return GLSM_EINVAL;
    for (size_t i = 0; i < N; ++i) f->c1[i] = (R)in[i];
    return GLSM_OK;
  }
  if (field == GLSM_FL_F) {
    if (count != 19 * N) return GLSM_EINVAL;
    for (size_t i = 0; i < 19 * N; ++i) f->f[i] = (R)in[i];
    return GLSM_OK;
  }
  if (field == GLSM_FL_U) {
    if (count != 3 * N) return GLSM_EINVAL;
    for (size_t i = 0; i < N; ++i) f->u[i] = V3<R>{(R)in[3 * i], (R)in[3 * i + 1], (R)in[3 * i + 2]};
    return GLSM_OK;
  }
  return GLSM_EINVAL;
}

}  // namespace

extern "C" {

int oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

// ---- gel ----
void* oracle_gel_create(int sx, int sy, int sz, double px, double py, double pz, int type, int id) {
  return new Gel(sx, sy, sz, px, py, pz, type, id);
}
void oracle_gel_destroy(void* g) { delete static_cast<Gel*>(g); }
void oracle_gel_step_elasticity(void* g, int iter) { static_cast<Gel*>(g)->stepElasticity(iter); }
void oracle_gel_step_chemistry(void* g, int iter) { static_cast<Gel*>(g)->stepChemistry(iter); }
void oracle_gel_record_center(void* g, int iter) { static_cast<Gel*>(g)->recordCenter(iter); }
void oracle_gel_reset_to_quiescent(void* g) { static_cast<Gel*>(g)->resetToQuiescent(); }
void oracle_gel_fire_excitation_pulse(void* g) { static_cast<Gel*>(g)->fireExcitationPulse(); }
void oracle_gel_set_anchor_z(void* g, double z) { static_cast<Gel*>(g)->setAnchorZ(z); }
void oracle_gel_build_tube_mask(void* g, int wall, int circular) {
  static_cast<Gel*>(g)->buildTubeMask(wall, circular != 0);
}
void oracle_gel_perturb_chemistry(void* g, uint64_t seed, double amp) {
  static_cast<Gel*>(g)->perturbChemistry(seed, amp);
}
void oracle_gel_set_chemistry_rates(void* g, double f, double ep) {
  static_cast<Gel*>(g)->setChemistryRates(f, ep);
}
void oracle_gel_dims(void* gv, int* grid3, int* nN, int* nE, int* M) {
  Gel* g = static_cast<Gel*>(gv);
  grid3[0] = g->L.x; grid3[1] = g->L.y; grid3[2] = g->L.z;
  *nN = g->nN; *nE = g->nE; *M = g->M;
}
void oracle_gel_params(void* gv, double* uss, double* vss, double* wss) {
  Gel* g = static_cast<Gel*>(gv);
  *uss = g->gp.uss; *vss = g->gp.vss; *wss = g->gp.wss;
}
int oracle_gel_read(void* g, int field, void* out, size_t bytes) {
  return gel_field(static_cast<Gel*>(g), field, out, nullptr, bytes);
}
int oracle_gel_write(void* g, int field, const void* in, size_t bytes) {
  int rc = gel_field(static_cast<Gel*>(g), field, nullptr, in, bytes);
  if (rc == GLSM_OK) static_cast<Gel*>(g)->snapshotToHost();
  return rc;
}

// ---- coupler ----
void* oracle_coupler_create(void** gels, int n, int prec) {
  std::vector<Gel*> v;
  for (int i = 0; i < n; ++i) v.push_back(static_cast<Gel*>(gels[i]));
  OCoupler* c = new OCoupler{prec};
  if (prec == 32) c->f32 = new Coupler<float>(v);
  else c->f64 = new Coupler<double>(v);
  return c;
}
void oracle_coupler_destroy(void* cv) {
  OCoupler* c = static_cast<OCoupler*>(cv);
  delete c->f32; delete c->f64; delete c;
}
#define DISPATCH_C(c, ...) \
  do { OCoupler* _c = static_cast<OCoupler*>(c); if (_c->f32) { auto* p = _c->f32; __VA_ARGS__; } else { auto* p = _c->f64; __VA_ARGS__; } } while (0)
void oracle_coupler_pack_from_gels(void* c) { DISPATCH_C(c, p->packFromGels()); }
void oracle_coupler_scatter_to_gels(void* c) { DISPATCH_C(c, p->scatterToGels()); }
void oracle_coupler_apply_gel_repulsion(void* c) { DISPATCH_C(c, p->applyGelRepulsion()); }
int oracle_coupler_marker_count(void* c) { int m = 0; DISPATCH_C(c, m = p->M); return m; }
int oracle_coupler_read(void* c, int field, double* out, size_t count) {
  int rc = 0; DISPATCH_C(c, rc = coupler_read(p, field, out, count)); return rc;
}
// direct write of the marker positions / velocities (unit tests of IBM)
int oracle_coupler_write(void* c, int field, const double* in, size_t count) {
  int rc = GLSM_EINVAL;
  DISPATCH_C(c, {
    using R = std::remove_reference_t<decltype(p->Cl[0])>;
    size_t M = p->M;
    if ((field == GLSM_CP_LAG || field == GLSM_CP_VL) && count == 3 * M) {
      auto& v = field == GLSM_CP_LAG ? p->lag : p->Vl;
      for (size_t i = 0; i < M; ++i) v[i] = {(R)in[3 * i], (R)in[3 * i + 1], (R)in[3 * i + 2]};
      rc = GLSM_OK;
    } else if (field == GLSM_CP_CL && count == M) {
      for (size_t i = 0; i < M; ++i) p->Cl[i] = (R)in[i];
      rc = GLSM_OK;
    }
  });
  return rc;
}

// ---- fluid ----
void* oracle_fluid_create(int sx, int sy, int sz, void* cv) {
  OCoupler* c = static_cast<OCoupler*>(cv);
  OFluid* f = new OFluid{c->prec};
  if (c->f32) f->f32 = new Fluid<float>(sx, sy, sz, c->f32);
  else f->f64 = new Fluid<double>(sx, sy, sz, c->f64);
  return f;
}
void oracle_fluid_destroy(void* fv) {
  OFluid* f = static_cast<OFluid*>(fv);
  delete f->f32; delete f->f64; delete f;
}
#define DISPATCH_F(f, ...) \
  do { OFluid* _f = static_cast<OFluid*>(f); if (_f->f32) { auto* p = _f->f32; __VA_ARGS__; } else { auto* p = _f->f64; __VA_ARGS__; } } while (0)
void oracle_fluid_step_velocity(void* f, int iter) { DISPATCH_F(f, p->stepVelocity(iter)); }
void oracle_fluid_sub_step(void* f, double beta_eff) {
  DISPATCH_F(f, p->subStep((std::remove_reference_t<decltype(p->dt)>)beta_eff));
}
void oracle_fluid_step_concentration(void* f, int iter) { DISPATCH_F(f, p->stepConcentration(iter)); }
void oracle_fluid_set_periodic(void* f, int px, int py, int pz) {
  DISPATCH_F(f, { p->perX = px; p->perY = py; p->perZ = pz; });
}
void oracle_fluid_set_ibm(void* f, double beta, double ramp) {
  DISPATCH_F(f, { p->beta = (decltype(p->beta))beta; p->ramp_steps = (decltype(p->beta))ramp; });
}
void oracle_fluid_set_force(void* f, double fx, double fy, double fz) {
  DISPATCH_F(f, { using R = decltype(p->dt); p->Fc = V3<R>{(R)fx, (R)fy, (R)fz}; });
}
void oracle_fluid_dims(void* f, int* grid3, int* nsub) {
  DISPATCH_F(f, { grid3[0] = p->Nx; grid3[1] = p->Ny; grid3[2] = p->Nz; *nsub = p->Nsub; });
}
double oracle_fluid_beta_eff(void* f, int iter) { double b = 0; DISPATCH_F(f, b = (double)p->betaEff(iter)); return b; }
int oracle_fluid_read(void* f, int field, double* out, size_t count) {
  int rc = 0; DISPATCH_F(f, rc = fluid_read(p, field, out, count)); return rc;
}
int oracle_fluid_write(void* f, int field, const double* in, size_t count) {
  int rc = 0; DISPATCH_F(f, rc = fluid_write(p, field, in, count)); return rc;
}

// ---- whole step, sim.cpp:66-81 ----
void oracle_system_step(void** gels, int n, void* cv, void* fv, int iter) {
  OCoupler* c = static_cast<OCoupler*>(cv);
  OFluid* f = static_cast<OFluid*>(fv);
  std::vector<Gel*> gs;
  for (int i = 0; i < n; ++i) gs.push_back(static_cast<Gel*>(gels[i]));
  if (c->f32) {
    System<float> s{gs, c->f32, f->f32};
    s.step(iter);
  } else {
    System<double> s{gs, c->f64, f->f64};
    s.step(iter);
  }
}

// ---- scalar helpers for known-answer tests ----
double oracle_delta4_f64(double r) { return delta4<double>(r); }
float oracle_delta4_f32(float r) { return delta4<float>(r); }
void oracle_lattice_tables(int* c57, double* w19, int* opp19) {
  for (int q = 0; q < 19; ++q) {
    for (int a = 0; a < 3; ++a) c57[3 * q + a] = kC[q][a];
    opp19[q] = kOpp[q];
    w19[q] = q == 0 ? 1.0 / 3.0 : (q < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
  }
}
// one phase of the gel step, for pinning the phases one at a time:
// 0 serviceNodes, 1 elementProperties, 2 chemBoundary, 3 unnorm,
// 4 pressure, 5 nodesVelocity, 6 nodesPosition, 7 terms, 8 chem, 9 setZero
int oracle_gel_phase(void* gv, int phase) {
  Gel* g = static_cast<Gel*>(gv);
  switch (phase) {
    case 0: g->serviceNodes(); break;
    case 1: g->elementProperties(); break;
    case 2: g->chemBoundary(); break;
    case 3: g->unnorm(); break;
    case 4: g->pressure(); break;
    case 5: g->nodesVelocity(); break;
    case 6: g->nodesPosition(); break;
    case 7: g->terms(); break;
    case 8: g->chem(); break;
    case 9: g->setZero(); break;
    default: return GLSM_EINVAL;
  }
  g->snapshotToHost();
  return GLSM_OK;
}
int oracle_solve_bilinear(const double* AH, double v_iso, double* xb, double* yb) {
  return Gel::solveBilinear(AH[0], AH[1], AH[2], AH[3], AH[4], AH[5], AH[6], AH[7], v_iso, *xb, *yb) ? 1 : 0;
}
double oracle_fu(double u, double v, double w, double f) { return Gel::fu(u, v, w, f, 0); }
double oracle_fv(double u, double v, double w) { return Gel::fv(u, v, w, 0); }

}  // extern "C"
