/* glsm_fluid.h -- C ABI of the MI355X gLSM_fluid_3D time step.
 *
 * The reference exposes this path as three C++ classes (src/gel.h,
 * src/fluid.h, src/coupling.h).  This header is the flat C binding of those
 * classes: every entry point below replaces one constructor or method (cited
 * next to it) with plain pointers, sizes and an int status.  The same classes
 * are also available as C++ (gel.h, fluid.h, coupling.h in this directory),
 * which is what the sim*.cpp drivers compile against.
 *
 * Conventions
 *   - return 0 on success, a negative GLSM_E* code on failure;
 *     glsm_last_error() returns the message of the calling thread's last
 *     failure ("" if none).
 *   - all work is enqueued on one library stream in call order (the
 *     reference's per-object streams are unordered; program order is what its
 *     drivers assume).  Reads synchronise that stream before returning.
 *   - array reads/writes copy whole fields between device and host memory;
 *     `bytes` must equal glsm_*_field_bytes() for that field.
 */
#ifndef GLSM_FLUID_H_
#define GLSM_FLUID_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum glsm_status {
  GLSM_OK = 0,
  GLSM_EINVAL = -1,   /* bad argument (null handle, size mismatch, range) */
  GLSM_EDEVICE = -2,  /* HIP runtime error */
  GLSM_ESTATE = -3,   /* call not valid in the object's current state */
  GLSM_EIO = -4,      /* file could not be read or written */
  GLSM_ENOMEM = -5
};

/* ---- gel fields (layouts are the reference's: element arrays (L+1)^3,
 *      node arrays (L+2)^3, x fastest; double3 stored as 3 doubles) ---- */
enum glsm_gel_field {
  GLSM_GEL_UM = 0, GLSM_GEL_VM, GLSM_GEL_WM, GLSM_GEL_UM_NORM, GLSM_GEL_VM_NORM,
  GLSM_GEL_WMP, GLSM_GEL_T0, GLSM_GEL_T1, GLSM_GEL_T2, GLSM_GEL_VOL, GLSM_GEL_PM,
  GLSM_GEL_RM,          /* element double3 */
  GLSM_GEL_RM_LOC,      /* element double3 x 8 */
  GLSM_GEL_NMSM,        /* element double3 x 6 */
  GLSM_GEL_RN, GLSM_GEL_VELN, GLSM_GEL_VELS, GLSM_GEL_FN, GLSM_GEL_FN_ROBIN,
  GLSM_GEL_FDRAG_ROBIN, /* node double3 */
  GLSM_GEL_UN_NORM, GLSM_GEL_UN_ROBIN, GLSM_GEL_VN_NORM, /* node double */
  GLSM_GEL_MAP_NODE, GLSM_GEL_MAP_ELEMENT, GLSM_GEL_BINDEX, /* int32 */
  GLSM_GEL_CENTER_VM, GLSM_GEL_CENTER_WM,               /* double[1000] */
  GLSM_GEL_CENTER_RM, GLSM_GEL_CENTER_FN, GLSM_GEL_CENTER_VELN, /* double3[1000] */
  GLSM_GEL_FILAMENT,    /* double3[1000 * maxFilamentlen] */
  GLSM_GEL_NFIELDS
};

/* ---- coupler fields (marker arrays of length M, gel-major) ---- */
enum glsm_coupler_field {
  GLSM_CP_LAG = 0, GLSM_CP_UL, GLSM_CP_VL, GLSM_CP_FDRAG, GLSM_CP_FREP, /* float3 */
  GLSM_CP_CL, GLSM_CP_DL, GLSM_CP_SL,                                   /* float */
  GLSM_CP_OWNER, GLSM_CP_BINDEX,                                        /* int32 */
  GLSM_CP_NFIELDS
};

/* ---- fluid fields (lattice arrays of length N = Nx*Ny*Nz, x fastest) ---- */
enum glsm_fluid_field {
  GLSM_FL_U = 0,   /* float3: velocity of the last sub-step (reference d_u) */
  GLSM_FL_C,       /* float : concentration (reference d_c1) */
  GLSM_FL_F,       /* float[19][N]: pre-collision populations f_k(x,q) of the
                      next sub-step (reference d_f after the swap) */
  GLSM_FL_FPOST,   /* float[19][N]: post-collision populations (the stored
                      state of the pull scheme) */
  GLSM_FL_RHO,     /* double: rho - 1 of the GLSM_FL_F populations, summed in
                      FP64 from the stored values (read-only; extension) */
  GLSM_FL_NFIELDS
};

typedef struct glsm_gel glsm_gel;
typedef struct glsm_coupler glsm_coupler;
typedef struct glsm_fluid glsm_fluid;

/* library */
int glsm_init(int device);               /* cudaSetDevice (sim.cpp:19) + stream */
int glsm_synchronize(void);              /* cudaDeviceSynchronize (sim.cpp:85) */
const char* glsm_last_error(void);
const char* glsm_version(void);

/* ---- Gel: src/gel.h:33-165 ---- */
/* Gel::Gel(int3 gelSize, double3 gelPosition, int gelType, int gel_id, int time)
   gel.cu:555-663 */
int glsm_gel_create(int size_x, int size_y, int size_z, double pos_x, double pos_y, double pos_z,
                    int gel_type, int gel_id, int time, glsm_gel** out);
int glsm_gel_destroy(glsm_gel* g);                          /* Gel::~Gel gel.cu:665 */
int glsm_gel_step_elasticity(glsm_gel* g, int iter);        /* gel.cu:682-698 */
int glsm_gel_step_chemistry(glsm_gel* g, int iter);         /* gel.cu:700-704 */
int glsm_gel_record_center(glsm_gel* g, int iter);          /* gel.cu:706-713 */
int glsm_gel_write_files(glsm_gel* g, int iter);            /* gel.cu:543-553 */
/* Extension: lossless binary snapshot of every device field, the params and
   the output flags.  It complements the reference's text restart
   Gel::setGoonValue (gel.cu:221-282), which keeps 6 decimals of rn/um and 6
   significant digits of vm/wm and rebuilds the rest.  Loading needs a gel of
   the same size and type; on a failed load the gel's state is unspecified. */
int glsm_gel_save_state(glsm_gel* g, const char* path);
int glsm_gel_load_state(glsm_gel* g, const char* path);
int glsm_gel_reset_to_quiescent(glsm_gel* g);               /* gel.cu:731-747 */
int glsm_gel_fire_excitation_pulse(glsm_gel* g);            /* gel.cu:751-783 */
int glsm_gel_set_anchor_z(glsm_gel* g, double anchor_z);    /* gel.cu:788-795 */
int glsm_gel_build_tube_mask(glsm_gel* g, int wall_thickness, int circular); /* gel.cu:811-867 */
int glsm_gel_get_index(glsm_gel* g, int xi, int yi, int zi, int size);       /* gel.cu:116-119 */
/* node grid (L), array sizes and marker count (gel.cu:605-610) */
int glsm_gel_dims(glsm_gel* g, int node_grid[3], int* num_nodes, int* num_elements,
                  int* boundary_count);
/* Extension (no reference counterpart): interior elements in z,y,x order get
   u += amp*U(-1,1), then v += amp*U(-1,1), drawn from splitmix64(seed). */
int glsm_gel_perturb_chemistry(glsm_gel* g, uint64_t seed, double amplitude);
/* Extension: rate constants f and ep (GelParams, gel.cu:633-652); the steady
   state is solved again (gel.cu:424-443) and the interior chemistry set to
   it (gel.cu:179-219).  Call before stepping. */
int glsm_gel_set_chemistry_rates(glsm_gel* g, double f, double ep);
/* Extension (failure detection; the reference's drivers test vm and the
   tip position for NaN themselves, sim_gelonly.cpp:102-115): number of
   non-finite values in the gel's rn, Veln, um, vm and wm. */
int glsm_gel_check_finite(glsm_gel* g, long long* bad_values);
size_t glsm_gel_field_bytes(glsm_gel* g, int field);
int glsm_gel_read(glsm_gel* g, int field, void* host, size_t bytes);
int glsm_gel_write(glsm_gel* g, int field, const void* host, size_t bytes);

/* ---- Coupler: src/coupling.h:14-58 ---- */
/* Coupler::Coupler(std::vector<Gel*>&) coupling.cu:50-80 */
int glsm_coupler_create(glsm_gel* const* gels, int num_gels, glsm_coupler** out);
int glsm_coupler_destroy(glsm_coupler* c);                  /* coupling.cu:161-164 */
int glsm_coupler_pack_from_gels(glsm_coupler* c);           /* coupling.cu:82-98 */
int glsm_coupler_scatter_to_gels(glsm_coupler* c);          /* coupling.cu:100-118 */
int glsm_coupler_apply_gel_repulsion(glsm_coupler* c);      /* coupling.cu:120-123 */
int glsm_coupler_marker_count(glsm_coupler* c);             /* sumGelBoundaryCount */
size_t glsm_coupler_field_bytes(glsm_coupler* c, int field);
int glsm_coupler_read(glsm_coupler* c, int field, void* host, size_t bytes);

/* ---- Fluid: src/fluid.h:24-84 ---- */
/* Fluid::Fluid(int3 fluidSize, int time, Coupler*) fluid.cu:152-232 */
int glsm_fluid_create(int size_x, int size_y, int size_z, int time, glsm_coupler* c,
                      glsm_fluid** out);
int glsm_fluid_destroy(glsm_fluid* f);                      /* fluid.cu:234-237 */
int glsm_fluid_step_velocity(glsm_fluid* f, int iter);      /* fluid.cu:263-284 */
/* one LBM sub-step (one pass of the fluid.cu:272-283 loop body) with the
   given IBM stiffness; u is stored for every cell */
int glsm_fluid_sub_step(glsm_fluid* f, float beta_eff);
int glsm_fluid_step_concentration(glsm_fluid* f, int iter); /* fluid.cu:251-261 */
int glsm_fluid_write_files(glsm_fluid* f, int iter);        /* fluid.cu:140-150 */
/* Extension (precision form): FP64 conservation reductions over the
   pre-collision populations of the next sub-step (the reference's d_f):
   out[0] = sum over cells of (rho - 1), out[1..3] = sum of rho*u with
   rho*u = sum_q f_q c_q as in k_macros (fluid_kernels.cu:98-111) without the
   F/2 term.  Per-cell sums in FP64, fixed-order block reduction. */
int glsm_fluid_conservation(glsm_fluid* f, double out[4]);
/* No counterpart in the reference.  glsm_fluid_conservation for callers that
   cannot allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc): the same kernels and the same
   bits, with the per-block FP64 partials in a static 4 MiB device array
   instead of library-allocated memory.  No hipMalloc/hipFree or allocating
   library call.  Lattices of at most 33,554,432 cells (GLSM_EINVAL beyond).
   Every call on a device shares that device's array: calls that overlap on
   one device (several host threads, or streams) race. */
int glsm_fluid_conservation_static(glsm_fluid* f, double out[4]);
/* Extension (failure detection): bad_cells = cells whose stored populations,
   velocity or concentration hold a non-finite value -- a binary16 deviation
   beyond the storage's range is stored as Inf and counts; max_abs_dev (may
   be NULL) = the largest finite |f_q - w_q| of the stored state; limit (may
   be NULL) = the largest magnitude the storage represents (65504 / 2^s for
   FP16 deviations scaled by 2^s).  One pass over the lattice. */
int glsm_fluid_check_finite(glsm_fluid* f, long long* bad_cells, double* max_abs_dev, double* limit);
/* Extension: lossless binary snapshot of the stored populations (in their
   storage representation), u, c and the params; loading switches the storage
   mode to the saved one.  It complements Fluid::setGoonValue
   (fluid.cu:57-83), which restarts f at rest and reads u and c as text.  The
   fluid must have the same lattice and marker count. */
int glsm_fluid_save_state(glsm_fluid* f, const char* path);
int glsm_fluid_load_state(glsm_fluid* f, const char* path);
/* FluidParams::perX/perY/perZ (fluid.h:21); honoured by stream, IBM and scalar */
int glsm_fluid_set_periodic(glsm_fluid* f, int per_x, int per_y, int per_z);
/* FluidParams::beta and the IBM ramp length, normally from $IBM_BETA /
   $IBM_RAMP_STEPS (fluid.cu:198-201, 265-268) */
int glsm_fluid_set_ibm(glsm_fluid* f, float beta, float ramp_steps);
/* Extension: distribution storage of the post-collision state (collision is
   FP32 in every mode).  The current state is converted.  scale_exp scales
   FP16 deviations by 2^scale_exp (ignored otherwise). */
enum glsm_storage {
  GLSM_STORAGE_F32_FULL = 0, /* f, the reference's representation */
  GLSM_STORAGE_F32_DEV = 1,  /* f - w_q in FP32 (default) */
  GLSM_STORAGE_F16_DEV = 2,  /* (f - w_q) * 2^scale_exp in binary16 */
  GLSM_STORAGE_BF16_DEV = 3  /* f - w_q in bfloat16 */
};
int glsm_fluid_set_storage(glsm_fluid* f, int storage, int scale_exp);
/* Extension (measurement): sample the stream/collide kernel duration every
   `every`-th stepVelocity (0 = off) with events the dispatch timestamps at
   kernel start and end; the result call waits for the stream and returns
   the mean in ms. */
int glsm_fluid_enable_lbm_timing(glsm_fluid* f, int every, int max_samples);
int glsm_fluid_lbm_timing(glsm_fluid* f, double* mean_ms, int* samples);
/* lattice (Nx,Ny,Nz) and Nsub (fluid.cu:167-197) */
int glsm_fluid_dims(glsm_fluid* f, int grid[3], int* nsub);
size_t glsm_fluid_field_bytes(glsm_fluid* f, int field);
int glsm_fluid_read(glsm_fluid* f, int field, void* host, size_t bytes);
int glsm_fluid_write(glsm_fluid* f, int field, const void* host, size_t bytes);

/* Extension (measurement, no reference counterpart): bandwidth of a plain
   16-B-per-lane device copy of `bytes` bytes, `reps` times; reported as
   (bytes read + bytes written) / s in GB/s.  The reference point for the
   roofline and for calibrating HBM traffic counters. */
int glsm_copy_bandwidth(size_t bytes, int reps, double* gbps);

/* Extension (measurement and tests): form of the IBM interpolation and
   concentration sampling kernels, 0 = default (interpolation staged as in
   3; sampling by marker count: one wavefront per marker below 2048 markers,
   else one thread per marker), 1 = one wavefront per marker, 2 = one
   thread per marker, 3 = products staged in LDS by eight threads per marker
   and summed by one per component (sampling in the thread form).  Every
   form sums the 64 stencil products in the same order, so results are
   identical.  Process-wide; launches issued after the call use it. */
int glsm_set_interp_mode(int mode);

/* Extension (measurement): XCD run length of the single-lattice
   stream/collide kernel's block order -- consecutive 256-cell blocks dealt
   to one XCD (whose L2 then serves the cache lines neighbouring blocks
   share): -1 = automatic (4 when the state streams from HBM, 0 = one
   contiguous run per XCD while it fits the Infinity Cache), else 0, 4, 16
   or 64.  Results are identical.  Process-wide. */
int glsm_set_lattice_xcd_run(int run);

/* Extension (measurement): LDS reserved per work-group of the
   stream/collide kernels (single-lattice and batched), which caps the
   work-groups resident per CU at 160 KB / reserve: -1 = automatic, else
   0..65536 bytes.  Results are identical.  Process-wide. */
int glsm_set_lbm_lds_reserve(int bytes);

/* Extension (measurement and tests): from `min_markers` markers on (default
   4096; 0 = never), Coupler::applyGelRepulsion skips, for each marker, the
   gels whose markers' bounding box lies beyond the cut-off.  The sum is
   still taken in increasing partner index, so results are identical to the
   full scan.  Process-wide; launches issued after the call use it. */
int glsm_set_repulsion_cull(int min_markers);

/* ---- Batch (include/batch.h; no reference counterpart): many independent
   instances, each a coupler's gels + the coupler + its fluid, advanced one
   time step at a time by one launch per phase over all of them (HIP-graph
   captured).  Every instance gives the same bits as glsm_system_step on its
   own objects.  The objects stay owned by the caller and must outlive their
   membership. ---- */
typedef struct glsm_batch glsm_batch;
int glsm_batch_create(int storage, int scale_exp, glsm_batch** out);
int glsm_batch_destroy(glsm_batch* b);
/* iter = the instance's next iteration; write_files = produce the .dat
   output (Gel::writeFiles / Fluid::writeFiles cadence) after each step,
   into out_dir when an output stream is attached (NULL or "" = the working
   directory).  Returns the member id (> 0) or a negative status. */
int glsm_batch_add(glsm_batch* b, glsm_coupler* c, glsm_fluid* f, int iter, int write_files,
                   const char* out_dir);
int glsm_batch_remove(glsm_batch* b, int id);
int glsm_batch_size(glsm_batch* b);
int glsm_batch_iteration(glsm_batch* b, int id);
int glsm_batch_step(glsm_batch* b, int nsteps);
int glsm_batch_use_graph(glsm_batch* b, int on);
/* stream/collide launch shape: nt = non-temporal stores, run = XCD run length
   (0, 4, 16 or 64), grid = work-groups (0 = one per work unit, else
   persistent), wgb = 256-cell blocks stepped at once by a work-group (1, 2
   or 4: 256, 512 or 1024 threads), chunk = rounds of consecutive blocks per
   work-group (cells per thread, 1..64); a work unit is wgb * chunk blocks;
   lds = LDS bytes reserved per work-group (0..65536, one block per
   work-group only), which caps the work-groups resident per CU */
int glsm_batch_set_launch(glsm_batch* b, int nt, int run, int grid, int wgb, int chunk, int lds);
int glsm_batch_get_launch(glsm_batch* b, int* nt, int* run, int* grid, int* wgb, int* chunk, int* lds);
/* time every launch shape on the current state (unchanged) and keep the
   fastest; best_ms = its mean stream/collide launch time */
int glsm_batch_autotune(glsm_batch* b, int reps, double* best_ms);
int glsm_batch_enable_lbm_timing(glsm_batch* b, int every);
int glsm_batch_lbm_timing(glsm_batch* b, double* mean_ms, int* samples);

/* ---- asynchronous .dat output (include/output.h; no reference
   counterpart: the reference's writeFiles copies synchronously).  A request
   snapshots the fields device-to-device in stream order, copies them to
   pinned host memory on a copy stream and formats the reference's files on
   writer threads while stepping goes on. ---- */
typedef struct glsm_output glsm_output;
int glsm_output_create(int writers, glsm_output** out);
int glsm_output_destroy(glsm_output* o);       /* waits for pending writes */
int glsm_output_gel(glsm_output* o, glsm_gel* g, int time, const char* dir);
int glsm_output_fluid(glsm_output* o, glsm_fluid* f, int time, const char* dir);
int glsm_output_release(glsm_output* o, const void* object);  /* a gel or fluid handle */
int glsm_output_flush(glsm_output* o);
/* requests not yet written / written; the last write error ("" if none) */
int glsm_output_status(glsm_output* o, long* pending, long* completed, char* err, size_t err_len);
/* attach (or detach with NULL) the stream to a batch */
int glsm_batch_set_output(glsm_batch* b, glsm_output* o);

/* ---- z-slab decomposition (include/slab.h; no reference counterpart: the
   reference runs one lattice on one device).  A slab holds planes
   [z0, z0 + nz) of a fluid's lattice plus one halo plane on each side and
   advances them with the same stream/collide kernel (F_const only, no
   marker force).  A chain of slabs covering the lattice in z order steps
   together: each slab steps, packs the five populations crossing each face
   into a face buffer, receives its neighbours' buffers (a device copy here;
   an RCCL send/recv or peer copy across GPUs) and unpacks them into its
   halos.  A chain reproduces the whole lattice's stream/collide sub-steps
   bit for bit. ---- */
typedef struct glsm_slab glsm_slab;
int glsm_slab_create(glsm_fluid* like, int z0, int nz, glsm_slab** out);
int glsm_slab_destroy(glsm_slab* s);
/* own planes and halos <- src's stored state; own planes -> dst's */
int glsm_slab_load(glsm_slab* s, glsm_fluid* src);
int glsm_slab_store(glsm_slab* s, glsm_fluid* dst);
/* `steps` sub-steps of the chain (slabs in z order covering the lattice) */
int glsm_slab_chain_step(glsm_slab* const* chain, int n, int steps, int write_u);
/* bytes of one face buffer (5 x Nx x Ny stored values) */
int glsm_slab_face_bytes(glsm_slab* s, size_t* bytes);
/* event timing (ms, mean per sub-step since enabled): the slab's
   stream/collide launch, and its pack -> transfer -> unpack phase */
int glsm_slab_enable_timing(glsm_slab* s, int on);
int glsm_slab_timing(glsm_slab* s, double* lbm_ms, double* exchange_ms);

/* device memory of the current device, bytes (hipMemGetInfo) */
int glsm_device_memory(size_t* free_bytes, size_t* total_bytes);

/* ---- one time step in sim.cpp's order (sim.cpp:66-81): for each gel
   stepElasticity, stepChemistry, recordCenter; packFromGels; stepVelocity;
   stepConcentration; applyGelRepulsion; scatterToGels.  The gels must be the
   ones the coupler was built from, in the same order. ---- */
int glsm_system_step(glsm_gel* const* gels, int num_gels, glsm_coupler* c, glsm_fluid* f,
                     int iter);

#ifdef __cplusplus
}
#endif
#endif /* GLSM_FLUID_H_ */
