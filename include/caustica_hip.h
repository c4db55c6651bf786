/* caustica_hip.h -- C ABI of the MI355X (gfx950) CW k-space + Pennes/CEM43 library.
 *
 * Plain C: pointers, DLPack tensors (cst_dlpack.h), sizes and status codes. Every call
 * returns CST_OK (0) or a nonzero status; cst_last_error() returns the message of the
 * last failure on the calling thread. All volumes are C-order float32 [nx][ny][nz] with z
 * contiguous (the upstream numpy layout of the padded grid).
 *
 * Ownership. Arrays passed in -- plain pointers (host memory) or DLTensors (host memory:
 * kDLCPU, kDLCUDAHost, kDLROCMHost; or device memory: kDLROCM on the session's GPU,
 * compact C order) -- are borrowed for the duration of the call; the library keeps no
 * reference. Tensors the library exports (cst_*_export, cst_dl_alloc) are
 * DLManagedTensors owned by the caller, released by calling their deleter. The
 * plain-pointer calls wrap their arguments in host DLTensors and call the DLPack form.
 *
 * Upstream interfaces these entry points replace (paths in the upstream caustica tree):
 *   the backend seam       src/caustica/core/backend.py:141-190, :236-270 (get_backend -> xp, fft)
 *   the CW solve           src/caustica/solvers/kspace/engine.py:167-586
 *                          (setup :225-278, step() :458-482, settle loop :484-510,
 *                           record window :522-556, result :558-586)
 *   checkpoint state       src/caustica/solvers/kspace/engine.py:344-371
 *   the Pennes solve       src/caustica/thermal/pennes.py:451-689 (hot loop :608-654)
 *   CEM43 rate             src/caustica/thermal/dose.py:74-89
 *   heat source            src/caustica/sensors.py:193-295
 */
#ifndef CAUSTICA_HIP_H
#define CAUSTICA_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "cst_dlpack.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CST_OK 0
#define CST_ERR_ARG 1         /* invalid argument (shape, slot, size, pointer) */
#define CST_ERR_HIP 2         /* HIP runtime failure (allocation, launch, copy) */
#define CST_ERR_UNSUPPORTED 3 /* axis length without an instantiated kernel */
#define CST_ERR_STATE 4       /* call out of order (e.g. member slot empty) */

#define CST_MAX_BATCH 16      /* members advanced by one launch sequence */
#define CST_MAX_HARMONICS 3

/* Library version (major * 10000 + minor * 100 + patch). */
int cst_version(void);
/* Copy the calling thread's last error message into buf (always NUL-terminated). */
int cst_last_error(char* buf, size_t n);
/* Number of visible HIP devices. */
int cst_device_count(int* n);
/* 1 if an axis of length n is supported: axis 0/1 (x, y: complex passes) or 2 (z: real,
 * n even). Mirrors the 2/3/5-smooth padding rule of upstream operators.py:23-39. */
int cst_supported_length(int n, int axis);

/* ------------------------------------------------------------------ CW k-space solve
 * A session holds up to max_members solves ("members") of ONE padded shape and one HIP
 * stream; cst_advance() steps any subset of them with one launch sequence (batched
 * launch, blockIdx.z = member). A single solve is a session with one member.
 * Replaces: engine.py:225-278 (backend resolve, upload of maps/operators/source). */
typedef struct cst_session cst_session;

int cst_session_create(int nx, int ny, int nz, int max_members, int device, cst_session** out);
int cst_session_destroy(cst_session* s);
/* The session's HIP stream (hipStream_t), for callers that order their own work. */
int cst_session_stream(cst_session* s, void** stream);

typedef struct {
  double dx;             /* grid spacing [m] */
  int pml_vox;           /* sponge width [voxels] (upstream grid.pml_vox) */
  double pml_edge;       /* sponge edge factor (upstream PMLSpec.edge) */
  double dt;             /* time step [s] (exact-period, engine.py:86-113) */
  double c_ref;          /* kappa reference speed (c_max, engine.py:252) */
  int spp;               /* steps per period */
  double f0;             /* drive frequency [Hz] */
  double amplitude;      /* drive amplitude [Pa] */
  double ramp_periods;   /* cosine ramp length [periods] (sources.py:98-102) */
  /* padded-shape coefficient maps (engine.py:239-249); beta2_dt NULL = linear */
  const float* dt_over_rho;
  const float* rhoc2_dt;
  const float* absorb;
  const float* beta2_dt;
  /* float32 wavenumbers (operators.py:60-73, 90-142): kx[nx] kx_d[nx] ky[ny] ky_d[ny]
   * kz[nz/2+1] kz_d[nz/2+1], _d with the even-axis Nyquist bin zeroed */
  const float* ktab;
  /* float32 sponge profiles on the padded axes (pml.py:53-73): sz[nz] sx[nx] sy[ny] */
  const float* sponge;
  /* source voxels (engine.py:268-278): n_src x 3 int32 indices, phase [rad],
   * scale = float32(2 c dt / dx) at the voxel; duplicates are refused */
  int n_src;
  const int32_t* src_idx;
  const float* src_phase;
  const float* src_scale;
  int conv[6];           /* convergence box x0,x1,y0,y1,z0,z1 (engine.py:292-296) */
  int rec[6];            /* record box (engine.py:298-301) */
  int n_harmonics;       /* 1..3 */
  int harmonics[CST_MAX_HARMONICS];
  int max_periods;       /* capacity of the per-period peak array */
} cst_member_desc;

/* Upload a member; returns its slot. Allocates its device memory (state, spectra,
 * maps, record box) or reuses a retired block of the same size. The upload runs on the
 * session's upload stream and is complete when the call returns; members already in
 * the session may be running meanwhile (admission at any step boundary). */
int cst_member_add(cst_session* s, const cst_member_desc* d, int* slot);
/* The same with the coefficient maps as DLPack tensors (float32, nx * ny * nz elements,
 * host or device memory; beta2_dt NULL = linear); the desc's map pointers are ignored.
 * Device maps are copied device to device: nothing crosses the bus. */
int cst_member_add_dl(cst_session* s, const cst_member_desc* d, const DLTensor* dt_over_rho,
                      const DLTensor* rhoc2_dt, const DLTensor* absorb, const DLTensor* beta2_dt,
                      int* slot);
/* ------------------------------------------------ opt-in staggered scheme
 * No counterpart in the reference: its step (engine.py:458-482) is collocated with
 * exp(-alpha c dt) absorption and a multiplicative sponge, and power-law absorption is
 * its unstarted milestone M16 (MILESTONES.md:1201). A session created with
 * CST_SCHEME_STAGGERED steps the first-order equations on a staggered grid: half-cell
 * shifted, kappa-corrected spectral derivatives, split density rho_a with a split-field
 * PML, fractional-Laplacian power-law absorption and dispersion (alpha0 omega^y, when
 * `absorbing`) and the Westervelt term m (beta - 1) rho^2 / rho0. The schedule, source
 * drive, period peaks, record window, results and state calls are the reference
 * scheme's; the source adds q(n) / (nd c^2) to rho_a for the grid's axes (src_scale =
 * float32(2 c dt / dx / (nd c^2)) at each voxel). Members of a reference-scheme session
 * are added with cst_member_add*, of a staggered session with cst_member_add_staggered*
 * (the other form is refused). cst_session_create = scheme CST_SCHEME_REFERENCE. */
#define CST_SCHEME_REFERENCE 0
#define CST_SCHEME_STAGGERED 1

int cst_session_create_scheme(int nx, int ny, int nz, int max_members, int device, int scheme,
                              cst_session** out);
int cst_session_scheme(cst_session* s, int* scheme);

typedef struct {
  /* complex64 [nx + ny + nz/2 + 1] (interleaved re, im): f_a(k) = k_a exp(+i k_a dx / 2)
   * per axis, k_a = 2 pi fftfreq (rfftfreq along z); i kappa f_a is the derivative to
   * x_a + dx/2, i kappa conj(f_a) the one back */
  const float* kfac;
  /* float32 [4][nz + nx + ny], each table sz[nz] sx[nx] sy[ny] of exp(-sigma dt / 2):
   * table 0 at grid points on every axis (density), table 1 + a at staggered points
   * x_a + dx/2 along axis a and ones along the others (velocity of axis a) */
  const float* pml;
  double alpha_power;   /* y of alpha0 omega^y (0 < y < 3, y != 1) */
  int absorbing;        /* 1: maps 5, 6 are given and the fractional Laplacians run */
  int src_axes;         /* bit a: the source adds to rho_a (the grid's axes) */
} cst_staggered_desc;

/* Upload a member of a staggered session. The desc's ktab (full wavenumbers; the _d
 * entries are unused), source, boxes, timing and harmonics mean what they mean for
 * cst_member_add; its map and sponge pointers are ignored. maps: 9 float32 tensors of
 * nx * ny * nz elements (host or device): [0..2] dt / rho0 at x_a + dx/2, [3] dt rho0,
 * [4] c^2, [5] c^2 tau / dt and [6] c^2 eta (absorbing members, else NULL; tau =
 * -2 alpha0 c^(y-1), eta = 2 alpha0 c^y tan(pi y / 2)), [7] 2 m dt and [8] c^2 m
 * (beta - 1) / rho0 (nonlinear members, else NULL; m = 1 where beta > 0). Members
 * advanced together must agree on `absorbing` (CST_ERR_ARG otherwise). */
int cst_member_add_staggered_dl(cst_session* s, const cst_member_desc* d, const cst_staggered_desc* sd,
                                const DLTensor* const* maps, int* slot);
int cst_member_add_staggered(cst_session* s, const cst_member_desc* d, const cst_staggered_desc* sd,
                             const float* const* maps, int* slot);
/* The split density rho_0, rho_1, rho_2 (real volumes of the padded shape, host or device
 * tensors; any may be NULL in get) of a staggered member. set_density after
 * cst_member_set_state_dl (which restarts the record window and sets n) completes the
 * state of step n. Synchronizes. */
int cst_member_get_density_into(cst_session* s, int slot, DLTensor* rho0, DLTensor* rho1, DLTensor* rho2);
int cst_member_set_density_dl(cst_session* s, int slot, const DLTensor* rho0, const DLTensor* rho1,
                              const DLTensor* rho2);

/* Retire a member once its own queued steps are done (steps of other members queued
 * after them may still run). Its device block stays with the session and is reused by
 * the next member of the same footprint (no hipFree on the serving path). */
int cst_member_remove(cst_session* s, int slot);
/* Free the session's retired blocks; *released (may be NULL) = bytes returned. */
int cst_session_trim(cst_session* s, size_t* released);
int cst_session_pooled_bytes(cst_session* s, size_t* bytes);
/* Free and total device memory (hipMemGetInfo) of a device, for admission budgets. */
int cst_device_memory(int device, size_t* free_bytes, size_t* total_bytes);
/* Calibration kernel for hardware-counter passes (no counterpart in the reference):
 * allocates two buffers of `bytes` (a multiple of 16), runs one plain copy kernel that
 * reads `bytes` and writes `bytes`, synchronizes and frees them. A counter's value for
 * that kernel converts its units into bytes. */
int cst_diag_copy(int device, size_t bytes);
/* Zero the member's state, spectra, record sums and peaks and set its step counter to
 * 0, keeping its uploaded maps and source (a fresh solve of the same setup). */
int cst_member_reset(cst_session* s, int slot);
/* Steps n >= record_from accumulate the phasors and p_max (engine.py:535-550). */
int cst_member_schedule(cst_session* s, int slot, int64_t record_from);
/* Current step counter of a member. */
int cst_member_step(cst_session* s, int slot, int64_t* n);

/* Advance every listed member by n_steps steps of engine.py:458-482, each from its own
 * step counter, as one batched launch sequence per step (asynchronous on the session
 * stream). Per-period max |p| over the convergence box goes to the member's peak slot
 * floor(n / spp) during steps before record_from. The step scalars (drive amplitude and
 * phase, record flag, period slot) are derived on the device from each member's step
 * clock, so every step takes the same launch arguments: the sequence is captured once
 * per member list as HIP graphs (16 steps, and 1 step for a remainder) and replayed
 * (cst_session_set_graphs). A member whose phasors were finalized by cst_member_result
 * is refused (CST_ERR_STATE) until cst_member_reset, set_state or set_raw. */
int cst_advance(cst_session* s, int n_slots, const int* slots, int n_steps);
/* Replay captured graphs (on = 1, the default) or launch every pass (0; timing with
 * cst_session_set_profiling always launches). Clears the session's graph cache. */
int cst_session_set_graphs(cst_session* s, int on);
/* Graph use: *enabled (0 once a capture failed: the session then launches directly),
 * graph launches and captures so far. */
int cst_session_graph_stats(cst_session* s, int* enabled, int64_t* launches, int64_t* captures);
/* Block until the session stream is idle. */
int cst_sync(cst_session* s);
/* *idle = 1 when every call queued on the session's stream has completed (non-blocking
 * hipStreamQuery), 0 while work is in flight: lets one host thread drive several
 * sessions without waiting on any of them. */
int cst_session_query(cst_session* s, int* idle);

/* Period peaks (engine.py:490-495): out[i] = max |p| over the convergence box during
 * period first + i (0-based). Synchronizes. NaN/inf propagate (engine.py:137-164). */
int cst_member_peaks(cst_session* s, int slot, int first, int count, float* out);
/* Record-window result (engine.py:551-567): phasor of harmonic index hi as complex64
 * (interleaved re, im) over the record box, = (Kahan sum - compensation) * scale
 * evaluated in float64 on the device; pmax (may be NULL) = float32 max |p|. The
 * reduction runs once per record window, in place of the compensation: a later call for
 * the same harmonic returns the same phasor and refuses another scale (CST_ERR_ARG), and
 * the member cannot advance until it is reset or given a new state. Runs on the
 * session's io stream after the member's last queued step (other members' steps may
 * run meanwhile). Synchronizes the io stream. */
int cst_member_result(cst_session* s, int slot, int hi, double scale, float* phasor, float* pmax);
/* The same into caller tensors (complex64 / float32, one element per record-box voxel,
 * host or device memory; either may be NULL). */
int cst_member_result_into(cst_session* s, int slot, int hi, double scale, DLTensor* phasor,
                           DLTensor* pmax);
/* The same into new tensors of the record-box shape [rx][ry][rz] on the host
 * (device_type kDLCPU) or the session's GPU (kDLROCM); the caller owns them. pmax may be
 * NULL (not exported). */
int cst_member_result_export(cst_session* s, int slot, int hi, double scale, int device_type,
                             DLManagedTensor** phasor, DLManagedTensor** pmax);
/* Solver state as real volumes (the state engine.py:344-371 checkpoints): p, u0, u1, u2
 * (padded shape, each may be NULL in get_state) and step n. The velocity is held on the
 * device as its z-spectrum: get_state forms u by a C2R along z, set_state by an R2C, so a
 * get/set round trip changes u by float32 transform rounding. set_state and set_raw
 * restart the record window (sums and p_max cleared: checkpoints are taken before the
 * first record step). Synchronizes. */
int cst_member_get_state(cst_session* s, int slot, float* p, float* u0, float* u1, float* u2,
                         int64_t* n);
int cst_member_set_state(cst_session* s, int slot, int64_t n, const float* p, const float* u0,
                         const float* u1, const float* u2);
/* DLPack forms of get_state / set_state: float32 tensors of nx * ny * nz elements in host
 * or device memory (a device-resident checkpoint never crosses the bus). */
int cst_member_get_state_into(cst_session* s, int slot, DLTensor* p, DLTensor* u0, DLTensor* u1,
                              DLTensor* u2, int64_t* n);
int cst_member_set_state_dl(cst_session* s, int slot, int64_t n, const DLTensor* p, const DLTensor* u0,
                            const DLTensor* u1, const DLTensor* u2);
/* The exact device state as bytes (p, then the three velocity spectra, then on a
 * staggered session the three density components), for checkpoints
 * that resume bit for bit (io/checkpoint.py:127-226 with the resume contract of
 * tests/test_checkpoint.py:61-83). state_size gives the byte count for the session's
 * shape; get_raw / set_raw refuse any other size. Synchronizes. */
int cst_member_state_size(cst_session* s, size_t* bytes);
int cst_member_get_raw(cst_session* s, int slot, void* buf, size_t bytes, int64_t* n);
int cst_member_set_raw(cst_session* s, int slot, const void* buf, size_t bytes, int64_t n);

/* Select the variant of a tunable pass for every later step of the session; the
 * autotuner times them per shape and keeps the fastest. Pass 0 (the y passes) and 1 (the
 * x passes): column-tile width 0 (default), 1 (half), 2 (twice; lengths >= 64). Pass 2
 * (zu) and 3 (zp): thread layout (zu 0..2, zp 0..1) + 4 x radix plan (0: stage-1 count
 * the smallest divisor of nz/2 at least sqrt(nz/2), 1: its cofactor). */
int cst_session_set_variant(cst_session* s, int pass, int variant);

/* Kernel timing: when enabled, every launch is bracketed by HIP events on the session
 * stream and its duration accumulated per kernel kind. Kinds: 0 xgrad, 1 y (3 spectra),
 * 2 zu, 3 xdiv, 4 y (1 spectrum), 5 zp, 6 record-finalize. */
int cst_session_set_profiling(cst_session* s, int on);
int cst_session_kernel_time(cst_session* s, int kind, double* total_ms, int64_t* launches);

/* DLPack helpers for callers without an array library of their own: a new tensor
 * (dtype code/bits as DLDataType, one lane) on the host (kDLCPU) or GPU `device`
 * (kDLROCM), owned by the caller; and a copy between two tensors of the same dtype and
 * element count (host or device, either direction). */
int cst_dl_alloc(int device_type, int device, int dtype_code, int dtype_bits, const int64_t* shape,
                 int ndim, DLManagedTensor** out);
int cst_dl_copy(const DLTensor* src, DLTensor* dst);

/* ------------------------------------------------------------------ Pennes + CEM43
 * One thermal solve on a grid nx x ny x nz; state theta = T - T_ref (float32), CEM43 dose
 * summed in float64. Replaces the hot loop of thermal/pennes.py:608-654 (coefficients
 * :404-447 are assembled by the caller, as upstream does before the loop). */
typedef struct cst_thermal cst_thermal;

typedef struct {
  double dx;            /* grid spacing [m] */
  const float* k;       /* conductivity [W/m/K] (faces use the harmonic mean, pennes.py:404-412) */
  const float* perf;    /* w_b rho_b C_b [W/m^3/K] (pennes.py:440-443) */
  const float* coeff;   /* dt_sub / (rho C) [m^3 K / J] (pennes.py:561) */
  const float* q;       /* heat source [W/m^3] or NULL */
  const float* theta0;  /* initial T - T_ref */
  const float* dose0;   /* CEM43 already accumulated [min] or NULL */
  int boundary;         /* Dirichlet wall mask: bit 0 x, bit 1 y, bit 2 z walls held at
                         * theta_wall; 0 = insulated everywhere */
  float theta_wall;     /* wall temperature - T_ref */
  float t_ref;          /* reference temperature [C] (dose uses T = theta + t_ref) */
  double half_dt_min;   /* 0.5 * dt_sub / 60 (trapezoid weight, minutes) */
  int dose;             /* accumulate CEM43 */
} cst_thermal_desc;

int cst_thermal_create(int nx, int ny, int nz, int device, cst_thermal** out);
int cst_thermal_destroy(cst_thermal* t);
/* Upload coefficients and the initial state; theta_max starts at theta0. */
int cst_thermal_setup(cst_thermal* t, const cst_thermal_desc* d);
/* Run n_substeps explicit sub-steps (asynchronous on the thermal stream). */
int cst_thermal_run(cst_thermal* t, int n_substeps);
/* x planes per block of the two-step stencil kernel: 0 (default) chooses per grid (the
 * longest of 64, 32, 16, 8 that gives every CU as many blocks as it holds at once); 8, 16,
 * 32 or 64 force it. The results do not depend on it (each plane's update is the same
 * arithmetic). */
int cst_thermal_set_chunk(cst_thermal* t, int planes);
/* Set-up finds whether the k, perfusion and dt/(rho C) maps are all uniform (every
 * element with element 0's bits); runs on such maps take stencil kernels that hold the
 * three values in registers and read no map (the same arithmetic on the same values:
 * the same bits). allow = 0 makes runs read the maps regardless; *uniform (optional)
 * receives what set-up found (0 before set-up). No reference counterpart: the upstream
 * medium always holds per-voxel volumes (thermal/properties.py:83-130). */
int cst_thermal_set_uniform_maps(cst_thermal* t, int allow, int* uniform);
/* 1 if any theta or dose value has stopped being finite (pennes.py:756-781). Synchronizes. */
int cst_thermal_status(cst_thermal* t, int* nonfinite);
/* Copy theta, theta_max (float32) and the dose (float64 sum rounded to float32); any may be NULL. */
int cst_thermal_read(cst_thermal* t, float* theta, float* theta_max, float* dose);
/* Acoustic -> thermal join on the device (replaces sensors.py:193-295 and the host round
 * trip of the phasor and Q). For every voxel of member `slot`'s record region:
 *   q = float32(scale * w * sum_j factor[j] |P_j|^2) - perf * q_offset      (float32 ops)
 * with P_j the member's finalized phasor of harmonic index harmonic_index[j]
 * (cst_member_result has run for it), |P_j|^2 and the sum in float64, w = alpha_f0 / (rho c)
 * a host float32 map over the region, factor[j] the harmonic's absorption scale (n^y for a
 * power law, 1 at f0) and q_offset = T_ref - T_a, the perfusion offset that
 * thermal/pennes.py folds into Q when the reference temperature is not arterial (0
 * otherwise). The thermal session must be set up on the record region's shape; it runs
 * with this heat source from then on. Synchronizes both sessions' streams. */
int cst_thermal_heat_from_member(cst_thermal* t, cst_session* s, int slot, int n_terms,
                                 const int* harmonic_index, const double* factor, const float* w,
                                 double scale, float q_offset);
/* Copy the session's heat source Q [W/m^3] (float32, the thermal grid) to the host, e.g.
 * the one cst_thermal_heat_from_member formed on the device. Synchronizes. */
int cst_thermal_read_source(cst_thermal* t, float* q);
int cst_thermal_set_profiling(cst_thermal* t, int on);
int cst_thermal_kernel_time(cst_thermal* t, double* total_ms, int64_t* launches);

/* ------------------------------------------------ Pennes + CEM43 without allocation
 * No counterpart in the reference. For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * sub-steps of cst_thermal_run on caller-owned device buffers, with the library's
 * scratch (the second theta buffer and the non-finite flag) in a fixed device array of
 * CST_NOALLOC_MAX_VOXELS * 4 + 256 bytes that every GPU the library runs on holds until
 * the process ends. Nothing is allocated or freed; the call runs on the current HIP
 * device and the given stream (NULL: the default stream). The array is shared by every
 * call on a GPU and keeps the last call's contents: calls that overlap on one GPU race.
 * Arrays are device pointers of C-order [nx][ny][nz] float32 (dose float64), the same
 * quantities cst_thermal_desc describes; theta is updated in place, theta_max and dose
 * accumulate in place (initialize theta_max to theta and the dose to the dose already
 * taken). nonfinite (may be NULL) receives the finite-guard flag of this call and then
 * synchronizes the stream; with NULL nothing synchronizes (capturable). Grids above
 * CST_NOALLOC_MAX_VOXELS are refused with CST_ERR_ARG. */
#define CST_NOALLOC_MAX_VOXELS (1L << 24)

typedef struct {
  int nx, ny, nz;
  double dx;            /* grid spacing [m] */
  float* theta;         /* T - T_ref, in/out */
  float* theta_max;     /* running max of theta, in/out */
  double* dose;         /* CEM43 [min] (float64), in/out; NULL when dose_on = 0 */
  const float* k;       /* conductivity [W/m/K] */
  const float* perf;    /* w_b rho_b C_b [W/m^3/K] */
  const float* coeff;   /* dt_sub / (rho C) */
  const float* q;       /* heat source [W/m^3] or NULL */
  int boundary;         /* Dirichlet wall mask, as cst_thermal_desc */
  float theta_wall;
  float t_ref;
  double half_dt_min;   /* 0.5 * dt_sub / 60 */
  int dose_on;          /* accumulate CEM43 */
} cst_thermal_dev_desc;

int cst_thermal_run_noalloc(const cst_thermal_dev_desc* d, int n_substeps, void* stream, int* nonfinite);

#ifdef __cplusplus
}
#endif
#endif /* CAUSTICA_HIP_H */
