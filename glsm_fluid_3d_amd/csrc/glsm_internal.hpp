// glsm_internal.hpp -- device-side views, launch entry points and runtime
// helpers shared by the library's translation units.  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace glsm {

// ---- runtime ----------------------------------------------------------------
hipStream_t stream();  // the one library stream (created on first use)
void set_device(int device);

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);
#define GLSM_HIP(call)                                              \
  do {                                                              \
    hipError_t _e = (call);                                         \
    if (_e != hipSuccess) ::glsm::throw_hip(_e, #call, __FILE__, __LINE__); \
  } while (0)
#define GLSM_LAUNCH_CHECK() GLSM_HIP(hipGetLastError())

// Device global memory (address space 1).  A pointer read from a device-side
// table (the batch descriptors) is a generic pointer to the compiler, which
// then emits flat memory instructions: each is counted against both the
// vector-memory and the scalar/LDS wait counters, so a wait for a scalar
// descriptor load also waits for every population load in flight.  Every
// buffer of this library is global memory; accessing it through GLSM_GLOBAL
// pointers gives global_* instructions (SGPR base + 32-bit lane offset).
// (The host pass parses device code too; it has no such address space.)
#ifdef __HIP_DEVICE_COMPILE__
#define GLSM_GLOBAL __attribute__((address_space(1)))
#else
#define GLSM_GLOBAL
#endif
template <class T>
__device__ __forceinline__ GLSM_GLOBAL T* gmem(T* p) {
  return (GLSM_GLOBAL T*)p;
}

// Work-groups are dealt to the 8 XCDs round-robin (id mod 8), and each XCD
// has its own L2.  Consecutive 256-cell blocks share the edge cache line of
// every shifted plane read, so the logical block of a work-group is chosen
// to put neighbours on the same XCD, where they meet in one L2 instead of
// fetching the shared line twice (measured: 1.10x -> 1.01x of the
// algorithmic bytes at 201^3).
//   RUN = 0: one contiguous run per XCD -- best while the state lives in the
//            Infinity Cache;
//   RUN > 0: runs of RUN blocks, dealt round-robin -- keeps the DRAM access
//            front compact for HBM-resident lattices (RUN = 16 measured best
//            with the 256-cell block layout: 201^3 FP32 205.1 vs 206.6 us for
//            RUN = 4, FP16 104.3 vs 106.6 us).
template <int RUN>
__device__ __forceinline__ unsigned xcd_block(unsigned id, unsigned nb) {
  constexpr unsigned kXcd = 8;
  if (RUN == 0) {
    const unsigned q = nb / kXcd, r = nb % kXcd, x = id % kXcd, k = id / kXcd;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
  }
  const unsigned grp = kXcd * RUN, full = nb / grp * grp;
  if (id >= full) return id;
  const unsigned base = id / grp * grp, w = id - base;
  return base + (w % kXcd) * RUN + w / kXcd;
}

template <class T>
T* dalloc(size_t n) {  // zero-filled device allocation
  void* p = nullptr;
  if (n == 0) n = 1;
  GLSM_HIP(hipMalloc(&p, n * sizeof(T)));
  GLSM_HIP(hipMemset(p, 0, n * sizeof(T)));
  return static_cast<T*>(p);
}
inline void dfree(void* p) {
  if (p) (void)hipFree(p);
}
// An event used only to time device work, read after the stream has been
// synchronised: recorded without a system-scope release, which would write
// back and invalidate the L2 at the timed launch's end and count that in its
// duration (and slow the launches after it).
inline hipEvent_t timing_event() {
  hipEvent_t e = nullptr;
  GLSM_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  return e;
}

// ---- gel --------------------------------------------------------------------
// Everything a gel kernel reads: dimensions, constants, the 27-entry offset
// table (ghost mirror / no-flux copy direction per type code) and the
// device arrays (reference layouts).
struct GelView {
  int LX, LY, LZ;
  int gelType;
  double f, ep, dt, dtx, dx, dy, dz;
  double CH0, CH1, CHS, C0, AZ0, FA0, c0_bis, b;
  double uss, vss, wss;
  int tube_mode;
  float tube_cy, tube_cz, tube_inner_hy, tube_inner_hz;
  int use_anchor;
  double anchor_z;
  int maxFilamentlen;
  int off[27][3];
  double *um, *vm, *wm, *um_norm, *vm_norm, *wmp, *T0, *T1, *T2, *vol, *pm;
  double3 *rm, *rm_loc, *nmSm;
  double3 *rn, *Veln, *Vels, *Fn, *Fn_robin, *Fdrag_robin;
  double *un_norm, *un_robin, *vn_norm;
  const int *map_node, *map_element;
  double *vm_center, *wm_center;
  double3 *rm_center, *Fn_center, *Veln_center, *filament;
};

void launch_gel_elasticity(const GelView& g, hipStream_t s);  // iter % 5 == 0 block
void launch_gel_set_zero(const GelView& g, hipStream_t s);
void launch_gel_chemistry(const GelView& g, hipStream_t s);
void launch_gel_record_center(const GelView& g, int slot, hipStream_t s);
// one gel step (elasticity on multiples of 5, setZero, chemistry, centre
// record on multiples of 1000) for `ngel` gels; views / gel_inst in device
// memory, iters[gel_inst[g]] = the iteration of gel g's instance
void launch_gel_step_batch(const GelView* views, const int* gel_inst, const int* iters, int ngel,
                           int max_node_grid, hipStream_t s);
// the same for `ngel` gels at one host-side iteration, without the centre
// record (Gel::recordCenter, one gel at a time): views in device memory
void launch_gel_step_multi(const GelView* views, int ngel, int max_node_grid, int iter, hipStream_t s);

// ---- coupler ----------------------------------------------------------------
}  // namespace glsm

// per-gel pointers used by the marker pack / scatter kernels
struct GelMarkerDesc {
  const double3* rn;
  const double3* Vels;
  const double* un_norm;
  double* un_robin;
  double3* Fn_robin;
  double3* Fdrag_robin;
};

// one (marker, weight) contribution to a lattice cell; entries of a cell are
// stored in increasing marker order
struct SpreadEntry {
  int marker;
  float w;
};

namespace glsm {

struct MarkerArrays {
  int M;
  const int* owner;
  const int* bIndex;  // node index inside the owner gel
  const GelMarkerDesc* desc;
  float3 *lag, *Ul, *Vl, *Fdrag, *Frep;
  float *Cl, *Dl, *Sl;
};

}  // namespace glsm
class Coupler;
namespace glsm {
MarkerArrays markers_of(Coupler& c);  // lattice-side view (no gel table)
MarkerArrays markers_with_gels(Coupler& c);  // with the gel table (pack / scatter)

void launch_pack(const MarkerArrays& m, float gel_to_lbm_vel, hipStream_t s);
void launch_scatter(const MarkerArrays& m, hipStream_t s);
// With the gels' marker ranges (gel_off, ngel + 1 entries) and a box buffer
// (6 floats per gel), large marker sets skip the gels out of range (same
// bits as the full scan); set_repulsion_cull(n): from n markers on (0 =
// never; a test and measurement switch)
void launch_repulsion(const MarkerArrays& m, float delta, float gamma, hipStream_t s, int ngel = 0,
                      const int* gel_off = nullptr, float* box = nullptr);
void set_repulsion_cull(int min_markers);

// ---- lattice ------------------------------------------------------------------
struct LatticeGeom {
  int Nx, Ny, Nz, N;
  int perX, perY, perZ;
  // planes z in (zlo, zhi] have a lower neighbour plane and [zlo, zhi) an
  // upper one (a whole lattice: 0 and Nz - 1; a z-slab with halo planes:
  // its halos, or its first / last own plane where a wall bounds it)
  int zlo, zhi;
  float h;
  double rNx, rNxNy;  // 1/Nx, 1/(Nx*Ny) for exact index decoding
};

// i = x + Nx (y + Ny z).  z = floor((i + 1/2) / (Nx Ny)), then y from the
// remainder the same way: the quotient's fractional part lies at least
// 1/(2n) from an integer, far more than the FP64 product's rounding error
// (relative 2^-52 of a quotient below 2^51 / n), so truncating gives the
// exact quotient without a correction step.
__device__ __forceinline__ void decode_cell(const LatticeGeom& g, int i, int& x, int& y, int& z) {
  const int sxy = g.Nx * g.Ny;
  z = (int)fma((double)i, g.rNxNy, 0.5 * g.rNxNy);
  const int r = i - z * sxy;
  y = (int)fma((double)r, g.rNx, 0.5 * g.rNx);
  x = r - (int)__umul24((unsigned)y, (unsigned)g.Nx);  // y * Nx < 2^32, both factors below 2^24
}

struct SpreadLists {
  int nblk;
  // per 256-cell block: byte w (of 4) is nonzero when some cell of the
  // block's w-th run of 64 cells (one wavefront) has a contribution, so the
  // int is nonzero when any cell of the block has one
  int* blkflag;
  int* cnt;
  int* fillpos;
  int2* rng;
  int* bsum;  // nblk + 1
  SpreadEntry* ent;
  // per marker, rebuilt with the lists (markers are fixed between rebuilds):
  // 64 entries per marker, lane-major (entry k of marker l at k * M + l):
  int* ctab;     // stencil cell (x outer, z inner) or -1
  float* wtab;   // spreading weight phi * (1/h^3) / sum(phi)
  float* rtab;   // interpolation weight phi_x phi_y phi_z
  float* wsum;   // per marker: sum of rtab in stencil order
};

// the flag byte, in a block's flag word, of the wavefront holding `lane`
// (0..255), as a wavefront-uniform value
__device__ __forceinline__ int wave_flag(int flag_word, unsigned lane) {
  const unsigned shift = __builtin_amdgcn_readfirstlane(8u * (lane >> 6));
  return (int)(((unsigned)flag_word >> shift) & 0xffu);
}

// rebuild the per-cell contribution lists for the current marker positions
void launch_spread_rebuild(const LatticeGeom& g, const MarkerArrays& m, const SpreadLists& sl,
                           hipStream_t s);
// Ul = interp(u), Fdrag = beta_eff (Vl - Ul), from the rebuilt stencil tables
void launch_ibm_interp(const MarkerArrays& m, const SpreadLists& sl, const float3* u, float beta_eff,
                       hipStream_t s);
// 0 = choose by marker count, 1 = one wavefront per marker, 2 = one thread
// per marker (identical results; a measurement and test switch)
void set_interp_mode(int mode);
// Dl = interp(c), Sl = Cl - Dl
void launch_ibm_sample(const MarkerArrays& m, const SpreadLists& sl, const float* c, hipStream_t s);

// distribution storage representations (collision is FP32 in all of them)
enum StorageMode { F32_FULL = 0, F32_DEV = 1, F16_DEV = 2, BF16_DEV = 3 };

// Layout of a distribution buffer: blocks of kDistBlock consecutive cells,
// each holding its cells' 19 populations as 19 consecutive runs of
// kDistBlock values.  Element (cell i, population q) is at
//   (i / 256) * 19 * 256 + q * 256 + i % 256.
// A 256-cell work-group of the stream/collide kernel then stores one
// contiguous 19 * 256-value region and reads its pulled values from a few
// such regions, instead of 19 runs in 19 planes a lattice apart (measured on
// the pull pattern alone: 6% faster in FP32, 7.5% in FP16 at 201^3).  A
// buffer holds `plane` = the cell count rounded up to whole blocks, times 19
// values; every element index fits in 32 bits (checked at allocation).
constexpr unsigned kDistBlock = 256;
__host__ __device__ __forceinline__ unsigned dist_index(unsigned i, unsigned q) {
  return i + (i / kDistBlock) * (18u * kDistBlock) + q * kDistBlock;
}
__host__ __device__ __forceinline__ size_t dist_cells(size_t n) {  // n rounded up to whole blocks
  return (n + kDistBlock - 1) / kDistBlock * kDistBlock;
}
// The stream/collide kernel's loads address a population's source by a
// 32-bit byte offset from a base kPullBias runs before that population's
// run (so that a bounce-back source, whose run can lie before q's, still has
// a non-negative offset).  A buffer of `plane` cells takes these offsets
// when dist_fits_b32(plane, bytes per value).
constexpr unsigned kPullBias = 38;
inline bool dist_fits_b32(size_t plane, size_t elem) {
  return (19ull * plane + (unsigned long long)kPullBias * kDistBlock) * elem < (1ull << 32);
}
size_t storage_bytes(int storage);
// device copy of `bytes` (multiple of 16), 16 B per lane
void launch_copy16(const void* a, void* b, size_t bytes, hipStream_t s);

struct LbmParams {
  float omega;     // 1/tau
  float w0, w1, w2;
  float3 Fc;       // FluidParams::F_const
  float scale;     // FP16 deviation scale 2^s (1 otherwise)
  float inv_scale;
};
// one fused pull stream + collide sub-step: fin -> fout (post-collision)
// ev0/ev1 (optional): start/stop events timestamped by the dispatch itself
void launch_lbm_step(int storage, const LatticeGeom& g, const LbmParams& p, const void* fin, void* fout,
                     size_t plane, float3* u, bool write_u_all, const SpreadLists& sl, const float3* Fl,
                     hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// XCD run length of the single-lattice stream/collide block order
// (-1 = automatic: 16 when the state streams from HBM, else 0); a
// measurement switch, process-wide
void set_lattice_xcd_run(int run);
// LDS reserved per stream/collide work-group (single-lattice and batched,
// one work-group per block), which caps the work-groups resident per CU:
// -1 = automatic (single lattice: kLatticeLdsReserve for 4-byte storage
// streamed from HBM, else 0; batch: its launch shape's), else 0..65536
// bytes for every launch; a measurement switch, process-wide
void set_lbm_lds_reserve(int bytes);
constexpr int kLatticeLdsReserve = 52 * 1024;  // 3 work-groups per CU
void launch_lbm_init(int storage, const LatticeGeom& g, const LbmParams& p, void* f, float3* u, float* c1,
                     float* c2, hipStream_t s);
// stored state -> FP32 full populations (pull = pre-collision of the next
// sub-step, else the stored post-collision ones), and the inverse of pull.
// The FP32 side is plane-major: population q of cell i at q * out_plane + i.
void launch_lbm_export(int storage, bool pull, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                       float* out, size_t out_plane, hipStream_t s);
// FP64 sum(rho - 1) and sum(rho u) of the pre-collision populations into
// out4[0..3]; `partial` holds 4 doubles per 256-cell block
void launch_lbm_moments(int storage, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                        double* partial, double* out4, hipStream_t s);
void launch_lbm_import(int storage, const LatticeGeom& g, const LbmParams& p, const float* in, void* fpost,
                       size_t in_plane, hipStream_t s);
// health scan of a lattice: out2[0] = cells with a non-finite stored value,
// velocity or concentration; out2[1] = bits of the largest finite
// |f_q - w_q| (a non-negative double).  out2 is cleared first.
void launch_lbm_health(int storage, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                       const float3* u, const float* c, unsigned long long* out2, hipStream_t s);
// p[0..n) = 0, and *out += number of non-finite values of a[0..n)
void launch_zero_u64(unsigned long long* p, int n, hipStream_t s);
void launch_count_nonfinite(const double* a, size_t n, unsigned long long* out, hipStream_t s);
// The FP64 conservation sums of launch_lbm_moments with their scratch in a
// static device array instead of an allocation (lattices of at most
// static_scratch_max_cells() cells); results copied to host_out4, stream
// synchronised.  static_scratch() looks up the array's address on the
// current device (once per device).
double* static_scratch();
long static_scratch_max_cells();
void conservation_static(int storage, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                         double* host_out4, hipStream_t s);
// rho - 1 per cell (FP64) of the pre-collision populations of the next sub-step
void launch_lbm_density(int storage, const LatticeGeom& g, const LbmParams& p, const void* fpost,
                        double* out, hipStream_t s);

struct ScalarParams {
  float dt, D, dx, dy, dz;
  float ihx, ihy, ihz;  // 1/dx, 1/dy, 1/dz, correctly rounded (as the device's division would give)
};
inline ScalarParams scalar_params(float dt, float D, float dx, float dy, float dz) {
  return ScalarParams{dt, D, dx, dy, dz, 1.0f / dx, 1.0f / dy, 1.0f / dz};
}
void launch_scalar(const LatticeGeom& g, const ScalarParams& p, const float3* u, const float* c1,
                   float* c2, const SpreadLists& sl, const float* Sl, hipStream_t s);

// ---- .dat output (reference text layout, from host arrays) -----------------
void write_fluid_dat(const std::string& dir, int time, int Nx, int Ny, int Nz, const float3* u, const float* c);
void write_gel_dat(const std::string& dir, int gel_id, int time, int LX, int LY, int LZ, const double3* rn,
                   const double3* Fn, const double3* Veln, const double3* rm, const double* um, const double* vm,
                   const double* wm);

// ---- batches --------------------------------------------------------------
}  // namespace glsm
class Gel;
class Fluid;
namespace glsm {
GelView gel_view(Gel& g);
LatticeGeom fluid_geom(const Fluid& f);
LbmParams fluid_lbm_params(const Fluid& f);

// One member (gel-fluid instance) of a batch as the batched kernels see it.
// Its 256-cell blocks are [blk0, blk0 + sl.nblk) and its markers
// [mk0, mk0 + m.M) in the batch-wide numbering.
struct BatchInst {
  LatticeGeom g;
  LbmParams p;
  ScalarParams sp;
  void* f[2];     // the two post-collision buffers (storage representation, block layout)
  float3* u;
  float* c[2];    // concentration ping-pong
  SpreadLists sl;
  MarkerArrays m;
  float g2l;      // gel -> lattice velocity factor of the pack
  float beta, ramp;
  float delta, gamma;
  int const_force;
  int blk0, mk0;
};

// Device tables of a batch: members, block -> member, marker -> member and
// the members' iteration counters.
struct BatchTables {
  const BatchInst* inst;
  const int* blk_inst;
  const int* mk_inst;
  const int* iters;
  int nblk, M, ninst;
};

void launch_pack_batch(const BatchTables& t, hipStream_t s);
void launch_scatter_batch(const BatchTables& t, hipStream_t s);
void launch_repulsion_batch(const BatchTables& t, hipStream_t s);
void launch_ibm_interp_batch(const BatchTables& t, hipStream_t s);  // beta_eff from iters
void launch_ibm_sample_batch(const BatchTables& t, int cpar, hipStream_t s);
void launch_spread_rebuild_batch(const BatchTables& t, hipStream_t s);
void launch_scalar_batch(const BatchTables& t, int cpar, hipStream_t s);


// Launch shape of the batched stream/collide kernel (the autotuner's knobs):
// non-temporal stores, XCD run length (0 = one contiguous run per XCD), the
// number of work-groups (0 = one per work unit; fewer = persistent
// work-groups looping over the units), 256-cell blocks stepped together by
// one work-group (wgb: 256, 512 or 1024 threads) and consecutive rounds of
// blocks per work-group (chunk: cells per thread).  lds: bytes of LDS
// reserved per work-group of a one-block-per-work-group shape, which caps
// the work-groups resident per CU (160 KB / lds).
struct LbmLaunch {
  int nt = 0;
  int run = 0;
  int grid = 0;
  int wgb = 1;
  int chunk = 1;
  int lds = 0;
};
// one sub-step of every member: f[sel] -> f[sel ^ 1]
void launch_lbm_step_batch(int storage, const BatchTables& t, int sel, bool write_u_all, const LbmLaunch& shape,
                           hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// iters[i] += 1 for the n members
void launch_advance_iters(int* iters, int n, hipStream_t s);

// ---- z-slab decomposition (slab_kernels.hip): copies of the stored
// representation, `elem` = 2 or 4 bytes per value ----
// slab planes 0..nz+1 <- lattice planes z0-1..z0+nz (wrapped if perZ)
void launch_slab_gather(size_t elem, const void* src, void* dst, size_t nxy, int z0, int nz, int Nz, int perZ,
                        hipStream_t s);
// lattice planes z0..z0+nz-1 <- slab planes 1..nz
void launch_slab_scatter(size_t elem, const void* src, void* dst, size_t nxy, int z0, int nz, hipStream_t s);
// face buffer (5 x nxy) <-> the c_z = +1 (which = 0) or c_z = -1 (which = 1)
// populations of slab plane z
void launch_face_pack(size_t elem, const void* f, size_t nxy, int z, int which, void* out, hipStream_t s);
void launch_face_unpack(size_t elem, const void* in, void* f, size_t nxy, int z, int which, hipStream_t s);

}  // namespace glsm
