// capi.cpp -- the C ABI of include/glsm_fluid.h over the Gel / Coupler /
// Fluid classes.  Every entry point catches C++ exceptions and HIP errors and
// returns a status; the message is kept per thread for glsm_last_error().
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "batch.h"
#include "slab.h"
#include "output.h"
#include "coupling.h"
#include "fluid.h"
#include "gel.h"
#include "glsm_fluid.h"
#include "glsm_internal.hpp"

struct glsm_coupler {
  std::vector<Gel*> gels;  // Coupler keeps a reference to this vector
  Coupler* c = nullptr;
  // the gels' device views, for glsm_system_step's one launch per gel phase;
  // uploaded again whenever a gel's view changes (its parameters are set
  // through its own methods between steps)
  glsm::GelView* d_views = nullptr;
  std::vector<glsm::GelView> h_views;
  int max_L = 0;
  ~glsm_coupler() { glsm::dfree(d_views); }
};

namespace {

thread_local std::string t_last_error;

struct IoError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

template <class F>
int guarded(F&& f) {
  try {
    t_last_error.clear();
    return f();
  } catch (const std::invalid_argument& e) {
    t_last_error = e.what();
    return GLSM_EINVAL;
  } catch (const IoError& e) {
    t_last_error = e.what();
    return GLSM_EIO;
  } catch (const std::bad_alloc& e) {
    t_last_error = e.what();
    return GLSM_ENOMEM;
  } catch (const std::runtime_error& e) {
    t_last_error = e.what();
    const std::string m = e.what();
    if (m.find("failed: hip") != std::string::npos) return GLSM_EDEVICE;
    if (m.find("missing") != std::string::npos) return GLSM_EIO;
    return GLSM_ESTATE;
  } catch (...) {
    t_last_error = "unknown C++ exception";
    return GLSM_ESTATE;
  }
}

#define REQUIRE(cond, msg)                   \
  do {                                       \
    if (!(cond)) throw std::invalid_argument(msg); \
  } while (0)

Gel* G(glsm_gel* g) {
  REQUIRE(g, "null gel handle");
  return reinterpret_cast<Gel*>(g);
}
Fluid* F(glsm_fluid* f) {
  REQUIRE(f, "null fluid handle");
  return reinterpret_cast<Fluid*>(f);
}
glsm_coupler* CP(glsm_coupler* c) {
  REQUIRE(c && c->c, "null coupler handle");
  return c;
}

struct FieldRef {
  void* dev;
  size_t bytes;
};

FieldRef gel_field(Gel* g, int field) {
  const size_t nE = g->m_numGelElements, nN = g->m_numGelNodes, M = g->m_boundaryCount;
  const size_t d = sizeof(double), d3 = sizeof(double3), I = sizeof(int);
  const size_t fil = 1000 * (size_t)g->m_hgp->maxFilamentlen;
  switch (field) {
    case GLSM_GEL_UM: return {g->m_dum, nE * d};
    case GLSM_GEL_VM: return {g->m_dvm, nE * d};
    case GLSM_GEL_WM: return {g->m_dwm, nE * d};
    case GLSM_GEL_UM_NORM: return {g->m_dum_norm, nE * d};
    case GLSM_GEL_VM_NORM: return {g->m_dvm_norm, nE * d};
    case GLSM_GEL_WMP: return {g->m_dwmp, nE * d};
    case GLSM_GEL_T0: return {g->m_dT0m, nE * d};
    case GLSM_GEL_T1: return {g->m_dT1m, nE * d};
    case GLSM_GEL_T2: return {g->m_dT2m, nE * d};
    case GLSM_GEL_VOL: return {g->m_dVolm, nE * d};
    case GLSM_GEL_PM: return {g->m_dPrem, nE * d};
    case GLSM_GEL_RM: return {g->m_drm, nE * d3};
    case GLSM_GEL_RM_LOC: return {g->m_drm_loc, 8 * nE * d3};
    case GLSM_GEL_NMSM: return {g->m_dnmSm, 6 * nE * d3};
    case GLSM_GEL_RN: return {g->m_drn, nN * d3};
    case GLSM_GEL_VELN: return {g->m_dVeln, nN * d3};
    case GLSM_GEL_VELS: return {g->m_dVels, nN * d3};
    case GLSM_GEL_FN: return {g->m_dFn, nN * d3};
    case GLSM_GEL_FN_ROBIN: return {g->m_dFn_robin, nN * d3};
    case GLSM_GEL_FDRAG_ROBIN: return {g->m_dFdrag_robin, nN * d3};
    case GLSM_GEL_UN_NORM: return {g->m_dun_norm, nN * d};
    case GLSM_GEL_UN_ROBIN: return {g->m_dun_robin, nN * d};
    case GLSM_GEL_VN_NORM: return {g->m_dvn_norm, nN * d};
    case GLSM_GEL_MAP_NODE: return {g->m_dmap_node, nN * I};
    case GLSM_GEL_MAP_ELEMENT: return {g->m_dmap_element, nE * I};
    case GLSM_GEL_BINDEX: return {g->m_dbIndex, M * I};
    case GLSM_GEL_CENTER_VM: return {g->m_dvm_center, 1000 * d};
    case GLSM_GEL_CENTER_WM: return {g->m_dwm_center, 1000 * d};
    case GLSM_GEL_CENTER_RM: return {g->m_drm_center, 1000 * d3};
    case GLSM_GEL_CENTER_FN: return {g->m_dFn_center, 1000 * d3};
    case GLSM_GEL_CENTER_VELN: return {g->m_dVeln_center, 1000 * d3};
    case GLSM_GEL_FILAMENT: return {g->m_dfilament, fil * d3};
    default: throw std::invalid_argument("unknown gel field " + std::to_string(field));
  }
}

FieldRef coupler_field(Coupler* c, int field) {
  const size_t M = c->sumGelBoundaryCount;
  switch (field) {
    case GLSM_CP_LAG: return {c->d_lag_all_, M * sizeof(float3)};
    case GLSM_CP_UL: return {c->d_Ul_all_, M * sizeof(float3)};
    case GLSM_CP_VL: return {c->d_Vl_all_, M * sizeof(float3)};
    case GLSM_CP_FDRAG: return {c->d_Fdrag_all_, M * sizeof(float3)};
    case GLSM_CP_FREP: return {c->d_Frep_all_, M * sizeof(float3)};
    case GLSM_CP_CL: return {c->d_Cl_all_, M * sizeof(float)};
    case GLSM_CP_DL: return {c->d_Dl_all_, M * sizeof(float)};
    case GLSM_CP_SL: return {c->d_Sl_all_, M * sizeof(float)};
    case GLSM_CP_OWNER: return {c->d_owner, M * sizeof(int)};
    case GLSM_CP_BINDEX: return {c->d_bIndex_all_, M * sizeof(int)};
    default: throw std::invalid_argument("unknown coupler field " + std::to_string(field));
  }
}

size_t fluid_field_bytes(Fluid* f, int field) {
  switch (field) {
    case GLSM_FL_U: return f->N * sizeof(float3);
    case GLSM_FL_C: return f->N * sizeof(float);
    case GLSM_FL_F:
    case GLSM_FL_FPOST: return 19ull * f->N * sizeof(float);
    case GLSM_FL_RHO: return (size_t)f->N * sizeof(double);
    default: throw std::invalid_argument("unknown fluid field " + std::to_string(field));
  }
}

void d2h(void* host, const void* dev, size_t bytes) {
  hipStream_t s = glsm::stream();
  GLSM_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s));
  GLSM_HIP(hipStreamSynchronize(s));
}
void h2d(void* dev, const void* host, size_t bytes) {
  hipStream_t s = glsm::stream();
  GLSM_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s));
  GLSM_HIP(hipStreamSynchronize(s));
}

// ---- binary state snapshots ----------------------------------------------
// File = header, then records (id, byte count, bytes).  Gel records are the
// device fields by their glsm_gel_field number plus the parameter block;
// fluid records are the stored populations in their storage representation,
// u, c and the parameter block.  A snapshot restores the state bit for bit.
constexpr uint32_t kSnapVersion = 2;  // 2: distributions in the 256-cell block layout
constexpr uint32_t kSnapGel = 1, kSnapFluid = 2;
constexpr uint32_t kRecGelParams = 1000, kRecGelAnchor = 1001;
constexpr uint32_t kRecFluidParams = 2000, kRecFluidRamp = 2001, kRecFluidF = 2002, kRecFluidU = 2003,
                   kRecFluidC = 2004;

struct SnapHeader {
  char magic[8];
  uint32_t version, kind;
  int32_t dims[3];
  int32_t extra[5];
  uint64_t records;
};

struct SnapWriter {
  FILE* fp;
  explicit SnapWriter(const char* path) : fp(std::fopen(path, "wb")) {
    if (!fp) throw IoError(std::string("snapshot: cannot open ") + path + " for writing");
  }
  ~SnapWriter() {
    if (fp) std::fclose(fp);
  }
  void raw(const void* p, size_t n) {
    if (std::fwrite(p, 1, n, fp) != n) throw IoError("snapshot: write failed");
  }
  void host(uint32_t id, const void* p, size_t n) {
    const uint64_t b = n;
    raw(&id, sizeof id);
    raw(&b, sizeof b);
    raw(p, n);
  }
  void device(uint32_t id, const void* dev, size_t n) {
    std::vector<char> tmp(n);
    d2h(tmp.data(), dev, n);
    host(id, tmp.data(), n);
  }
};

struct SnapReader {
  FILE* fp;
  explicit SnapReader(const char* path) : fp(std::fopen(path, "rb")) {
    if (!fp) throw IoError(std::string("snapshot: cannot open ") + path);
  }
  ~SnapReader() {
    if (fp) std::fclose(fp);
  }
  void raw(void* p, size_t n) {
    if (std::fread(p, 1, n, fp) != n) throw IoError("snapshot: truncated file");
  }
  // next record, which must be `id` of exactly `n` bytes
  void expect(uint32_t id, size_t n, void* dst) {
    uint32_t rid;
    uint64_t b;
    raw(&rid, sizeof rid);
    raw(&b, sizeof b);
    if (rid != id || b != n)
      throw std::invalid_argument("snapshot: record " + std::to_string(rid) + " (" + std::to_string(b) +
                                  " B) where " + std::to_string(id) + " (" + std::to_string(n) + " B) was expected");
    raw(dst, n);
  }
  void expect_device(uint32_t id, size_t n, void* dev) {
    std::vector<char> tmp(n);
    expect(id, n, tmp.data());
    h2d(dev, tmp.data(), n);
  }
};

SnapHeader make_header(uint32_t kind, int3 dims) {
  SnapHeader h{};
  std::memcpy(h.magic, "GLSMSNAP", 8);
  h.version = kSnapVersion;
  h.kind = kind;
  h.dims[0] = dims.x;
  h.dims[1] = dims.y;
  h.dims[2] = dims.z;
  return h;
}

void check_header(const SnapHeader& h, const SnapHeader& want, const char* what) {
  if (std::memcmp(h.magic, "GLSMSNAP", 8) != 0 || h.version != kSnapVersion)
    throw std::invalid_argument(std::string(what) + " snapshot: not a snapshot file of this version");
  if (h.kind != want.kind) throw std::invalid_argument(std::string(what) + " snapshot: file holds another object kind");
  for (int i = 0; i < 3; ++i)
    if (h.dims[i] != want.dims[i]) throw std::invalid_argument(std::string(what) + " snapshot: grid differs");
}

void gel_save(Gel* g, const char* path) {
  SnapHeader h = make_header(kSnapGel, g->m_gelNodeGrid);
  h.extra[0] = g->m_gelType;
  h.extra[1] = g->m_boundaryCount;
  h.extra[2] = g->m_hgp->maxFilamentlen;
  h.extra[3] = g->flag ? 1 : 0;
  h.extra[4] = g->m_use_anchor ? 1 : 0;
  h.records = 2 + GLSM_GEL_NFIELDS;
  GLSM_HIP(hipStreamSynchronize(glsm::stream()));
  SnapWriter w(path);
  w.raw(&h, sizeof h);
  w.host(kRecGelParams, g->m_hgp, sizeof(GelParams));
  w.host(kRecGelAnchor, &g->m_anchor_z, sizeof(double));
  for (int f = 0; f < GLSM_GEL_NFIELDS; ++f) {
    const FieldRef r = gel_field(g, f);
    w.device((uint32_t)f, r.dev, r.bytes);
  }
}

void gel_load(Gel* g, const char* path) {
  SnapReader rd(path);
  SnapHeader h;
  rd.raw(&h, sizeof h);
  check_header(h, make_header(kSnapGel, g->m_gelNodeGrid), "gel");
  if (h.extra[0] != g->m_gelType || h.extra[1] != g->m_boundaryCount || h.extra[2] != g->m_hgp->maxFilamentlen)
    throw std::invalid_argument("gel snapshot: gel type or sizes differ");
  GelParams gp;
  rd.expect(kRecGelParams, sizeof gp, &gp);
  double anchor;
  rd.expect(kRecGelAnchor, sizeof anchor, &anchor);
  for (int f = 0; f < GLSM_GEL_NFIELDS; ++f) {
    const FieldRef r = gel_field(g, f);
    rd.expect_device((uint32_t)f, r.bytes, r.dev);
  }
  *g->m_hgp = gp;
  g->m_anchor_z = anchor;
  g->flag = h.extra[3] != 0;
  g->m_use_anchor = h.extra[4] != 0;
  g->copyDataToHost();    // host mirrors follow the restored device state
  g->copyDataToDevice();  // uploads params; marks positions as changed
}

void fluid_save(Fluid* f, const char* path) {
  SnapHeader h = make_header(kSnapFluid, f->fluidNodeGrid);
  h.extra[0] = f->storage();
  h.extra[1] = f->storageScaleExp();
  h.extra[2] = f->Nsub;
  h.extra[3] = f->h_fp->M;
  h.records = 5;
  GLSM_HIP(hipStreamSynchronize(glsm::stream()));
  SnapWriter w(path);
  w.raw(&h, sizeof h);
  w.host(kRecFluidParams, f->h_fp, sizeof(FluidParams));
  w.host(kRecFluidRamp, &f->ramp_steps, sizeof(float));
  w.device(kRecFluidF, f->d_f, f->Nd * glsm::storage_bytes(f->storage()));
  w.device(kRecFluidU, f->d_u, f->N * sizeof(float3));
  w.device(kRecFluidC, f->d_c1, f->N * sizeof(float));
}

void fluid_load(Fluid* f, const char* path) {
  SnapReader rd(path);
  SnapHeader h;
  rd.raw(&h, sizeof h);
  check_header(h, make_header(kSnapFluid, f->fluidNodeGrid), "fluid");
  if (h.extra[2] != f->Nsub || h.extra[3] != f->h_fp->M)
    throw std::invalid_argument("fluid snapshot: sub-step or marker count differs");
  FluidParams fp;
  rd.expect(kRecFluidParams, sizeof fp, &fp);
  float ramp;
  rd.expect(kRecFluidRamp, sizeof ramp, &ramp);
  if (h.extra[0] != f->storage() || h.extra[1] != f->storageScaleExp()) f->setStorage(h.extra[0], h.extra[1]);
  rd.expect_device(kRecFluidF, f->Nd * glsm::storage_bytes(f->storage()), f->d_f);
  rd.expect_device(kRecFluidU, f->N * sizeof(float3), f->d_u);
  rd.expect_device(kRecFluidC, f->N * sizeof(float), f->d_c1);
  *f->h_fp = fp;
  f->ramp_steps = ramp;
  f->setPeriodic(fp.perX, fp.perY, fp.perZ);  // also forces the IBM lists to be rebuilt
}

}  // namespace

extern "C" {

int glsm_init(int device) {
  return guarded([&] {
    REQUIRE(device >= 0, "device must be >= 0");
    glsm::set_device(device);
    (void)glsm::stream();
    (void)glsm::static_scratch();  // the allocation-free entry point's scratch on this device
    return GLSM_OK;
  });
}

int glsm_synchronize(void) {
  return guarded([&] {
    GLSM_HIP(hipDeviceSynchronize());
    return GLSM_OK;
  });
}

const char* glsm_last_error(void) { return t_last_error.c_str(); }
const char* glsm_version(void) { return "glsm_fluid_3d_amd 0.1 (gfx950)"; }

// ---------------------------------------------------------------- gel
int glsm_gel_create(int sx, int sy, int sz, double px, double py, double pz, int gel_type, int gel_id, int time,
                    glsm_gel** out) {
  return guarded([&] {
    REQUIRE(out, "out is null");
    REQUIRE(sx >= 1 && sy >= 1 && sz >= 1, "gel size must be >= 1 element per axis");
    *out = reinterpret_cast<glsm_gel*>(
        new Gel(make_int3(sx, sy, sz), make_double3(px, py, pz), gel_type, gel_id, time));
    return GLSM_OK;
  });
}
int glsm_gel_destroy(glsm_gel* g) {
  return guarded([&] {
    delete G(g);
    return GLSM_OK;
  });
}
int glsm_gel_step_elasticity(glsm_gel* g, int iter) {
  return guarded([&] { G(g)->stepElasticity(iter); return GLSM_OK; });
}
int glsm_gel_step_chemistry(glsm_gel* g, int iter) {
  return guarded([&] { G(g)->stepChemistry(iter); return GLSM_OK; });
}
int glsm_gel_record_center(glsm_gel* g, int iter) {
  return guarded([&] { G(g)->recordCenter(iter); return GLSM_OK; });
}
int glsm_gel_write_files(glsm_gel* g, int iter) {
  return guarded([&] { G(g)->writeFiles(iter); return GLSM_OK; });
}
int glsm_gel_save_state(glsm_gel* g, const char* path) {
  return guarded([&] {
    REQUIRE(path, "null path");
    gel_save(G(g), path);
    return GLSM_OK;
  });
}
int glsm_gel_load_state(glsm_gel* g, const char* path) {
  return guarded([&] {
    REQUIRE(path, "null path");
    gel_load(G(g), path);
    return GLSM_OK;
  });
}
int glsm_gel_reset_to_quiescent(glsm_gel* g) {
  return guarded([&] { G(g)->resetToQuiescent(); return GLSM_OK; });
}
int glsm_gel_fire_excitation_pulse(glsm_gel* g) {
  return guarded([&] { G(g)->fireExcitationPulse(); return GLSM_OK; });
}
int glsm_gel_set_anchor_z(glsm_gel* g, double z) {
  return guarded([&] { G(g)->setAnchorZ(z); return GLSM_OK; });
}
int glsm_gel_build_tube_mask(glsm_gel* g, int wall, int circular) {
  return guarded([&] { G(g)->buildTubeMask(wall, circular != 0); return GLSM_OK; });
}
int glsm_gel_get_index(glsm_gel* g, int xi, int yi, int zi, int size) {
  int r = 0;
  int rc = guarded([&] { r = G(g)->get_index(xi, yi, zi, size); return GLSM_OK; });
  return rc == GLSM_OK ? r : rc;
}
int glsm_gel_dims(glsm_gel* g, int grid[3], int* nN, int* nE, int* M) {
  return guarded([&] {
    Gel* p = G(g);
    REQUIRE(grid && nN && nE && M, "null output pointer");
    grid[0] = p->m_gelNodeGrid.x;
    grid[1] = p->m_gelNodeGrid.y;
    grid[2] = p->m_gelNodeGrid.z;
    *nN = p->m_numGelNodes;
    *nE = p->m_numGelElements;
    *M = p->m_boundaryCount;
    return GLSM_OK;
  });
}
int glsm_gel_set_chemistry_rates(glsm_gel* g, double f, double ep) {
  return guarded([&] { G(g)->setChemistryRates(f, ep); return GLSM_OK; });
}
int glsm_gel_perturb_chemistry(glsm_gel* g, uint64_t seed, double amplitude) {
  return guarded([&] { G(g)->perturbChemistry(seed, amplitude); return GLSM_OK; });
}
size_t glsm_gel_field_bytes(glsm_gel* g, int field) {
  size_t b = 0;
  guarded([&] { b = gel_field(G(g), field).bytes; return GLSM_OK; });
  return b;
}
int glsm_gel_read(glsm_gel* g, int field, void* host, size_t bytes) {
  return guarded([&] {
    const FieldRef r = gel_field(G(g), field);
    REQUIRE(host && bytes == r.bytes, "gel read: size mismatch");
    d2h(host, r.dev, bytes);
    return GLSM_OK;
  });
}
int glsm_gel_check_finite(glsm_gel* g, long long* bad_values) {
  return guarded([&] {
    REQUIRE(bad_values, "null output pointer");
    *bad_values = G(g)->checkFinite();
    return GLSM_OK;
  });
}
int glsm_gel_write(glsm_gel* g, int field, const void* host, size_t bytes) {
  return guarded([&] {
    Gel* p = G(g);
    const FieldRef r = gel_field(p, field);
    REQUIRE(host && bytes == r.bytes, "gel write: size mismatch");
    h2d(r.dev, host, bytes);
    // keep the host chemistry mirrors in step (the phase controls upload them)
    if (field == GLSM_GEL_UM) std::memcpy(p->m_hum, host, bytes);
    if (field == GLSM_GEL_VM) std::memcpy(p->m_hvm, host, bytes);
    if (field == GLSM_GEL_WM) std::memcpy(p->m_hwm, host, bytes);
    if (field == GLSM_GEL_RN) {
      p->copyDataToHost();
      // new marker positions: the IBM stencil tables and spreading lists of
      // every coupler, fluid and batch holding this gel are rebuilt
      p->markPositionsChanged();
    }
    return GLSM_OK;
  });
}

// ---------------------------------------------------------------- coupler
int glsm_coupler_create(glsm_gel* const* gels, int n, glsm_coupler** out) {
  return guarded([&] {
    REQUIRE(out && (gels || n == 0) && n >= 0, "bad coupler arguments");
    auto* h = new glsm_coupler;
    try {
      for (int i = 0; i < n; ++i) h->gels.push_back(G(gels[i]));
      h->c = new Coupler(h->gels);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return GLSM_OK;
  });
}
int glsm_coupler_destroy(glsm_coupler* c) {
  return guarded([&] {
    delete CP(c)->c;
    delete c;
    return GLSM_OK;
  });
}
int glsm_coupler_pack_from_gels(glsm_coupler* c) {
  return guarded([&] { CP(c)->c->packFromGels(); return GLSM_OK; });
}
int glsm_coupler_scatter_to_gels(glsm_coupler* c) {
  return guarded([&] { CP(c)->c->scatterToGels(); return GLSM_OK; });
}
int glsm_coupler_apply_gel_repulsion(glsm_coupler* c) {
  return guarded([&] { CP(c)->c->applyGelRepulsion(); return GLSM_OK; });
}
int glsm_coupler_marker_count(glsm_coupler* c) {
  int m = 0;
  int rc = guarded([&] { m = CP(c)->c->sumGelBoundaryCount; return GLSM_OK; });
  return rc == GLSM_OK ? m : rc;
}
size_t glsm_coupler_field_bytes(glsm_coupler* c, int field) {
  size_t b = 0;
  guarded([&] { b = coupler_field(CP(c)->c, field).bytes; return GLSM_OK; });
  return b;
}
int glsm_coupler_read(glsm_coupler* c, int field, void* host, size_t bytes) {
  return guarded([&] {
    const FieldRef r = coupler_field(CP(c)->c, field);
    REQUIRE(host && bytes == r.bytes, "coupler read: size mismatch");
    d2h(host, r.dev, bytes);
    return GLSM_OK;
  });
}

// ---------------------------------------------------------------- fluid
int glsm_fluid_create(int sx, int sy, int sz, int time, glsm_coupler* c, glsm_fluid** out) {
  return guarded([&] {
    REQUIRE(out, "out is null");
    REQUIRE(sx >= 1 && sy >= 1 && sz >= 1, "fluid size must be >= 1 per axis");
    *out = reinterpret_cast<glsm_fluid*>(new Fluid(make_int3(sx, sy, sz), time, CP(c)->c));
    return GLSM_OK;
  });
}
int glsm_fluid_destroy(glsm_fluid* f) {
  return guarded([&] {
    delete F(f);
    return GLSM_OK;
  });
}
int glsm_fluid_step_velocity(glsm_fluid* f, int iter) {
  return guarded([&] { F(f)->stepVelocity(iter); return GLSM_OK; });
}
int glsm_fluid_sub_step(glsm_fluid* f, float beta_eff) {
  return guarded([&] { F(f)->subStep(beta_eff, true); return GLSM_OK; });
}
int glsm_fluid_step_concentration(glsm_fluid* f, int iter) {
  return guarded([&] { F(f)->stepConcentration(iter); return GLSM_OK; });
}
int glsm_fluid_write_files(glsm_fluid* f, int iter) {
  return guarded([&] { F(f)->writeFiles(iter); return GLSM_OK; });
}
int glsm_fluid_save_state(glsm_fluid* f, const char* path) {
  return guarded([&] {
    REQUIRE(path, "null path");
    fluid_save(F(f), path);
    return GLSM_OK;
  });
}
int glsm_fluid_load_state(glsm_fluid* f, const char* path) {
  return guarded([&] {
    REQUIRE(path, "null path");
    fluid_load(F(f), path);
    return GLSM_OK;
  });
}
int glsm_fluid_conservation(glsm_fluid* f, double out[4]) {
  return guarded([&] {
    REQUIRE(out, "null output pointer");
    F(f)->conservation(out);
    return GLSM_OK;
  });
}
int glsm_fluid_conservation_static(glsm_fluid* f, double out[4]) {
  return guarded([&] {
    REQUIRE(out, "null output pointer");
    Fluid* p = F(f);
    glsm::conservation_static(p->storage(), glsm::fluid_geom(*p), glsm::fluid_lbm_params(*p), p->d_f, out,
                              p->fluid_stream);
    return GLSM_OK;
  });
}
int glsm_fluid_check_finite(glsm_fluid* f, long long* bad_cells, double* max_abs_dev, double* limit) {
  return guarded([&] {
    REQUIRE(bad_cells, "null output pointer");
    *bad_cells = F(f)->checkFinite(max_abs_dev, limit);
    return GLSM_OK;
  });
}
int glsm_fluid_set_periodic(glsm_fluid* f, int px, int py, int pz) {
  return guarded([&] { F(f)->setPeriodic(px, py, pz); return GLSM_OK; });
}
int glsm_fluid_set_ibm(glsm_fluid* f, float beta, float ramp) {
  return guarded([&] { F(f)->setIBM(beta, ramp); return GLSM_OK; });
}
int glsm_fluid_set_storage(glsm_fluid* f, int storage, int scale_exp) {
  return guarded([&] { F(f)->setStorage(storage, scale_exp); return GLSM_OK; });
}
int glsm_fluid_enable_lbm_timing(glsm_fluid* f, int every, int max_samples) {
  return guarded([&] { F(f)->enableLbmTiming(every, max_samples); return GLSM_OK; });
}
int glsm_fluid_lbm_timing(glsm_fluid* f, double* mean_ms, int* samples) {
  return guarded([&] {
    REQUIRE(mean_ms, "null output pointer");
    *mean_ms = F(f)->lbmKernelTiming(samples);
    return GLSM_OK;
  });
}
int glsm_fluid_dims(glsm_fluid* f, int grid[3], int* nsub) {
  return guarded([&] {
    Fluid* p = F(f);
    REQUIRE(grid && nsub, "null output pointer");
    grid[0] = p->fluidNodeGrid.x;
    grid[1] = p->fluidNodeGrid.y;
    grid[2] = p->fluidNodeGrid.z;
    *nsub = p->Nsub;
    return GLSM_OK;
  });
}
size_t glsm_fluid_field_bytes(glsm_fluid* f, int field) {
  size_t b = 0;
  guarded([&] { b = fluid_field_bytes(F(f), field); return GLSM_OK; });
  return b;
}
int glsm_fluid_read(glsm_fluid* f, int field, void* host, size_t bytes) {
  return guarded([&] {
    Fluid* p = F(f);
    REQUIRE(host && bytes == fluid_field_bytes(p, field), "fluid read: size mismatch");
    switch (field) {
      case GLSM_FL_U: d2h(host, p->d_u, bytes); break;
      case GLSM_FL_C: d2h(host, p->d_c1, bytes); break;
      case GLSM_FL_F: p->readPopulations(static_cast<float*>(host)); break;
      case GLSM_FL_FPOST: p->readPostCollision(static_cast<float*>(host)); break;
      case GLSM_FL_RHO: p->readDensity(static_cast<double*>(host)); break;
    }
    return GLSM_OK;
  });
}
int glsm_fluid_write(glsm_fluid* f, int field, const void* host, size_t bytes) {
  return guarded([&] {
    Fluid* p = F(f);
    REQUIRE(host && bytes == fluid_field_bytes(p, field), "fluid write: size mismatch");
    switch (field) {
      case GLSM_FL_U: h2d(p->d_u, host, bytes); break;
      case GLSM_FL_C: h2d(p->d_c1, host, bytes); break;
      case GLSM_FL_F: p->writePopulations(static_cast<const float*>(host)); break;
      default: throw std::invalid_argument("fluid field is read-only");
    }
    return GLSM_OK;
  });
}

// ---------------------------------------------------------------- measurement
int glsm_copy_bandwidth(size_t bytes, int reps, double* gbps) {
  return guarded([&] {
    REQUIRE(gbps && reps > 0 && bytes >= 16, "bad copy benchmark arguments");
    bytes = bytes / 16 * 16;
    hipStream_t s = glsm::stream();
    char* a = glsm::dalloc<char>(bytes);
    char* b = glsm::dalloc<char>(bytes);
    hipEvent_t e0 = glsm::timing_event(), e1 = glsm::timing_event();
    glsm::launch_copy16(a, b, bytes, s);
    GLSM_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) glsm::launch_copy16(r % 2 ? b : a, r % 2 ? a : b, bytes, s);
    GLSM_HIP(hipEventRecord(e1, s));
    GLSM_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GLSM_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    glsm::dfree(a);
    glsm::dfree(b);
    *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return GLSM_OK;
  });
}

// ---------------------------------------------------------------- batch
struct glsm_batch {
  Batch* b = nullptr;
};
namespace {
Batch* B(glsm_batch* b) {
  REQUIRE(b && b->b, "null batch handle");
  return b->b;
}
}  // namespace
int glsm_batch_create(int storage, int scale_exp, glsm_batch** out) {
  return guarded([&] {
    REQUIRE(out, "null output pointer");
    auto* h = new glsm_batch;
    try {
      h->b = new Batch(storage, scale_exp);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return GLSM_OK;
  });
}
int glsm_batch_destroy(glsm_batch* b) {
  return guarded([&] {
    delete B(b);
    delete b;
    return GLSM_OK;
  });
}
int glsm_batch_add(glsm_batch* b, glsm_coupler* c, glsm_fluid* f, int iter, int write_files,
                   const char* out_dir) {
  return guarded([&] {
    glsm_coupler* cp = CP(c);
    return B(b)->add(cp->gels, cp->c, F(f), iter, write_files != 0, out_dir ? out_dir : "");
  });
}
int glsm_batch_remove(glsm_batch* b, int id) {
  return guarded([&] {
    B(b)->remove(id);
    return GLSM_OK;
  });
}
int glsm_batch_size(glsm_batch* b) {
  return guarded([&] { return B(b)->size(); });
}
int glsm_batch_iteration(glsm_batch* b, int id) {
  return guarded([&] { return B(b)->iteration(id); });
}
int glsm_batch_step(glsm_batch* b, int nsteps) {
  return guarded([&] {
    B(b)->step(nsteps);
    return GLSM_OK;
  });
}
int glsm_batch_use_graph(glsm_batch* b, int on) {
  return guarded([&] {
    B(b)->useGraph(on != 0);
    return GLSM_OK;
  });
}
int glsm_batch_set_launch(glsm_batch* b, int nt, int run, int grid, int wgb, int chunk, int lds) {
  return guarded([&] {
    B(b)->setLaunch(Batch::Launch{nt, run, grid, wgb, chunk, lds});
    return GLSM_OK;
  });
}
int glsm_batch_get_launch(glsm_batch* b, int* nt, int* run, int* grid, int* wgb, int* chunk, int* lds) {
  return guarded([&] {
    REQUIRE(nt && run && grid && wgb && chunk && lds, "null output pointer");
    const Batch::Launch l = B(b)->launch();
    *nt = l.nt;
    *run = l.run;
    *grid = l.grid;
    *wgb = l.wgb;
    *chunk = l.chunk;
    *lds = l.lds;
    return GLSM_OK;
  });
}
int glsm_batch_autotune(glsm_batch* b, int reps, double* best_ms) {
  return guarded([&] {
    B(b)->autotune(reps, best_ms);
    return GLSM_OK;
  });
}
int glsm_batch_enable_lbm_timing(glsm_batch* b, int every) {
  return guarded([&] {
    B(b)->enableLbmTiming(every);
    return GLSM_OK;
  });
}
int glsm_batch_lbm_timing(glsm_batch* b, double* mean_ms, int* samples) {
  return guarded([&] {
    REQUIRE(mean_ms && samples, "null output pointer");
    *mean_ms = B(b)->lbmKernelTiming(samples);
    return GLSM_OK;
  });
}

struct glsm_output {
  OutputStream* o = nullptr;
};
namespace {
OutputStream* O(glsm_output* o) {
  REQUIRE(o && o->o, "null output handle");
  return o->o;
}
}  // namespace
int glsm_output_create(int writers, glsm_output** out) {
  return guarded([&] {
    REQUIRE(out, "null output pointer");
    auto* h = new glsm_output;
    try {
      h->o = new OutputStream(writers);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return GLSM_OK;
  });
}
int glsm_output_destroy(glsm_output* o) {
  return guarded([&] {
    delete O(o);
    delete o;
    return GLSM_OK;
  });
}
int glsm_output_gel(glsm_output* o, glsm_gel* g, int time, const char* dir) {
  return guarded([&] {
    O(o)->gel(G(g), time, dir ? dir : "");
    return GLSM_OK;
  });
}
int glsm_output_fluid(glsm_output* o, glsm_fluid* f, int time, const char* dir) {
  return guarded([&] {
    O(o)->fluid(F(f), time, dir ? dir : "");
    return GLSM_OK;
  });
}
int glsm_output_release(glsm_output* o, const void* object) {
  return guarded([&] {
    O(o)->release(object);
    return GLSM_OK;
  });
}
int glsm_output_flush(glsm_output* o) {
  return guarded([&] {
    O(o)->flush();
    return GLSM_OK;
  });
}
int glsm_output_status(glsm_output* o, long* pending, long* completed, char* err, size_t err_len) {
  return guarded([&] {
    OutputStream* s = O(o);
    if (pending) *pending = s->pending();
    if (completed) *completed = s->completed();
    if (err && err_len) {
      const std::string e = s->lastError();
      std::snprintf(err, err_len, "%s", e.c_str());
    }
    return GLSM_OK;
  });
}
int glsm_batch_set_output(glsm_batch* b, glsm_output* o) {
  return guarded([&] {
    B(b)->setOutput(o ? O(o) : nullptr);
    return GLSM_OK;
  });
}
int glsm_device_memory(size_t* free_bytes, size_t* total_bytes) {
  return guarded([&] {
    REQUIRE(free_bytes && total_bytes, "null output pointer");
    GLSM_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return GLSM_OK;
  });
}

int glsm_set_lbm_lds_reserve(int bytes) {
  return guarded([&] {
    glsm::set_lbm_lds_reserve(bytes);
    return GLSM_OK;
  });
}
int glsm_set_lattice_xcd_run(int run) {
  return guarded([&] {
    glsm::set_lattice_xcd_run(run);
    return GLSM_OK;
  });
}
int glsm_set_interp_mode(int mode) {
  return guarded([&] {
    glsm::set_interp_mode(mode);
    return GLSM_OK;
  });
}

int glsm_set_repulsion_cull(int min_markers) {
  return guarded([&] {
    glsm::set_repulsion_cull(min_markers);
    return GLSM_OK;
  });
}

// ---------------------------------------------------------------- slabs
struct glsm_slab {
  Slab* s = nullptr;
};
namespace {
Slab* SL(glsm_slab* s) {
  REQUIRE(s && s->s, "null slab handle");
  return s->s;
}
}  // namespace

int glsm_slab_create(glsm_fluid* like, int z0, int nz, glsm_slab** out) {
  return guarded([&] {
    REQUIRE(out, "null output pointer");
    Fluid* f = F(like);
    auto* h = new glsm_slab;
    try {
      h->s = new Slab(*f, z0, nz);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return GLSM_OK;
  });
}
int glsm_slab_destroy(glsm_slab* s) {
  return guarded([&] {
    delete SL(s);
    delete s;
    return GLSM_OK;
  });
}
int glsm_slab_load(glsm_slab* s, glsm_fluid* src) {
  return guarded([&] {
    SL(s)->load(*F(src));
    return GLSM_OK;
  });
}
int glsm_slab_store(glsm_slab* s, glsm_fluid* dst) {
  return guarded([&] {
    SL(s)->store(*F(dst));
    return GLSM_OK;
  });
}
int glsm_slab_chain_step(glsm_slab* const* chain, int n, int steps, int write_u) {
  return guarded([&] {
    REQUIRE(chain && n > 0, "empty slab chain");
    REQUIRE(steps >= 0, "negative step count");
    std::vector<Slab*> v;
    for (int i = 0; i < n; ++i) v.push_back(SL(chain[i]));
    for (int k = 0; k < steps; ++k) step_slab_chain(v, write_u != 0);
    return GLSM_OK;
  });
}
int glsm_slab_face_bytes(glsm_slab* s, size_t* bytes) {
  return guarded([&] {
    REQUIRE(bytes, "null output pointer");
    *bytes = SL(s)->faceBytes();
    return GLSM_OK;
  });
}
int glsm_slab_enable_timing(glsm_slab* s, int on) {
  return guarded([&] {
    SL(s)->enableTiming(on != 0);
    return GLSM_OK;
  });
}
int glsm_slab_timing(glsm_slab* s, double* lbm_ms, double* exchange_ms) {
  return guarded([&] {
    REQUIRE(lbm_ms && exchange_ms, "null output pointer");
    *lbm_ms = SL(s)->lbmKernelMs();
    *exchange_ms = SL(s)->exchangeKernelMs();
    return GLSM_OK;
  });
}

// ---------------------------------------------------------------- system
int glsm_system_step(glsm_gel* const* gels, int n, glsm_coupler* c, glsm_fluid* f, int iter) {
  return guarded([&] {
    glsm_coupler* cp = CP(c);
    Fluid* fl = F(f);
    REQUIRE((int)cp->gels.size() == n, "gel count differs from the coupler's");
    for (int i = 0; i < n; ++i) REQUIRE(G(gels[i]) == cp->gels[i], "gel order differs from the coupler's");
    // the gels' stepElasticity + stepChemistry (gel.cu:682-704), one launch
    // per phase for all of them: the gels are independent, so this is the
    // per-gel sequence reordered
    if (n > 0) {
      std::vector<glsm::GelView> views;
      for (Gel* g : cp->gels) views.push_back(glsm::gel_view(*g));
      if (!cp->d_views) {
        cp->d_views = glsm::dalloc<glsm::GelView>(views.size());
        for (Gel* g : cp->gels)
          cp->max_L = std::max(cp->max_L, std::max(g->m_gelNodeGrid.x, std::max(g->m_gelNodeGrid.y, g->m_gelNodeGrid.z)));
      }
      if (views.size() != cp->h_views.size() ||
          std::memcmp(views.data(), cp->h_views.data(), views.size() * sizeof(glsm::GelView)) != 0) {
        // blocking copy: the stream (blocking with respect to the null
        // stream) finishes the steps that read the old views first
        GLSM_HIP(hipMemcpy(cp->d_views, views.data(), views.size() * sizeof(glsm::GelView), hipMemcpyHostToDevice));
        cp->h_views = views;
      }
      glsm::launch_gel_step_multi(cp->d_views, n, cp->max_L, iter, glsm::stream());
      if (iter % 5 == 0)
        for (Gel* g : cp->gels) g->markPositionsChanged();  // as stepElasticity does
    }
    for (int i = 0; i < n; ++i) cp->gels[i]->recordCenter(iter);
    cp->c->packFromGels();
    fl->stepVelocity(iter);
    fl->stepConcentration(iter);
    cp->c->applyGelRepulsion();
    cp->c->scatterToGels();
    return GLSM_OK;
  });
}

}  // extern "C"
