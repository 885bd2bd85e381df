/*
 * vse.h -- C-ABI of libvse, the MI355X (gfx950) staggered-grid
 * velocity-stress time loop (elastic and Zener viscoelastic, CPML, 2D/3D).
 *
 * The reference has no FFI: its drop-in point is the Python plugin hook
 * Simulator.implementation() (sim_support/simulator.py:608-611, called at
 * :337; Simulator3D.py:619-622, called at :356) overridden by each backend
 * (sim_webgpu_viscoelastic.py:43-555, sim_webgpu_3D_viscoelastic.py:44-713,
 * sim_taichi_elastic.py:28-351).  These entry points are what such a backend
 * needs from a native library: each one replaces the backend step named in
 * its comment.  Python binds them with ctypes (taichi_sim_amd/_lib.py).
 *
 * Conventions: plain pointers and sizes; host buffers are caller-owned and
 * copied; device memory is owned by the handle; every call is synchronous
 * with respect to host buffers; return 0 on success, a negative VSE_ERR_*
 * code on failure (vse_last_error() gives the message).  One handle per
 * thread.  A handle belongs to the device current at vse_create; every call
 * on it makes that device current for its duration and restores the
 * caller's.  Grids are C order with the last axis fastest (2D (nx, ny) with
 * y = the ROI's z axis; 3D (nx, ny, nz)), exactly as the reference lays out
 * its buffers.
 */
#ifndef VSE_H
#define VSE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VSE_ABI_VERSION 3

#define VSE_OK 0
#define VSE_ERR_ARG (-1)        /* invalid argument / shape */
#define VSE_ERR_HIP (-2)        /* HIP runtime error */
#define VSE_ERR_UNSTABLE (-3)   /* |v|^2 reached +inf or NaN (StabilityError) */
#define VSE_ERR_STATE (-4)      /* call order (e.g. run before set_material) */

#define VSE_VARIANT_WGSL 0      /* shader_webgpu_(3D_)viscoelastic.wgsl semantics */
#define VSE_VARIANT_TAICHI 1    /* sim_taichi_elastic.py semantics (2D elastic) */

/* field ids for vse_read_field / vse_field_device_ptr */
#define VSE_F_VX 0
#define VSE_F_VY 1
#define VSE_F_VZ 2
#define VSE_F_SXX 3
#define VSE_F_SYY 4
#define VSE_F_SZZ 5
#define VSE_F_SXY 6
#define VSE_F_SXZ 7
#define VSE_F_SYZ 8

typedef struct vse_sim vse_sim;

/* Simulation description.  Replaces the parameter buffers of the WGSL
 * backends (sim_webgpu_viscoelastic.py:57-71: params_i32/params_f32,
 * tau_attenuation, alpha_sum; sim_webgpu_3D_viscoelastic.py:62-76). */
typedef struct vse_desc {
    int ndim;               /* 2 or 3 */
    int nx, ny, nz;         /* grid incl. CPML and pad (nz ignored in 2D) */
    float dx, dy, dz, dt;   /* mm, mm, mm, us */
    int ord;                /* number of FD coefficients, 2..6 */
    float coefs[8];         /* staggered coefficients (coefs_Lui row) */
    int variant;            /* VSE_VARIANT_*; 3D supports WGSL only */
    int visco;              /* viscoelastic_attn flag */
    int n_sls;              /* Zener mechanisms L, 1..8 */
    float tau[32];          /* (4, L) row-major: tau_eps_p, tau_sig_p, tau_eps_s, tau_sig_s */
    float sum_alpha[2];     /* sum_l tau_eps/tau_sig for P and S */
    int n_rows;             /* rows of the source and trace buffers (= n_steps) */
} vse_desc;

int vse_abi_version(void);

/* Allocate device state for one simulation.  Replaces the ~60 buffer
 * creations per implementation() call (sim_webgpu_viscoelastic.py:83-191). */
int vse_create(const vse_desc *desc, vse_sim **out);

/* rho, cp, cs on the full grid (nx*ny[*nz] floats each); rho is the kernel
 * density the host prepared (the reference's aliased, averaged
 * _rho_grid_vx; simulator.py:88-131).  Replaces b_rho_map/b_cp_map/b_cs_map
 * (sim_webgpu_viscoelastic.py:131-133). */
int vse_set_material(vse_sim *s, const float *rho, const float *cp, const float *cs);

/* CPML profiles of one axis (0=x, 1=y, 2=z), each of length
 * n_axis - 2*(ord-1): full grid a, b, k and half grid a, b, k.  Replaces
 * b_a_x ... b_k_y_h (sim_webgpu_viscoelastic.py:95-106). */
int vse_set_cpml(vse_sim *s, int axis, const float *a, const float *b, const float *k,
                 const float *a_h, const float *b_h, const float *k_h);

/* Source points: n_pts distinct linear cell indices with their source column
 * (a cell listed twice is VSE_ERR_ARG: the reference's per-cell map holds one
 * column per cell), and the (n_rows, n_cols) row-major source matrix (value
 * of column c at step s).
 * WGSL variants add src[s][c]*dt/rho after the velocity update; the Taichi
 * variant adds it inside the vy update.  Replaces b_force and b_idx_src
 * (sim_webgpu_viscoelastic.py:89-92; sim_taichi_elastic.py:120-130). */
int vse_set_sources(vse_sim *s, int n_pts, const int *cells, const int *cols,
                    int n_cols, const float *src);

/* Receivers: n_rec receivers; receiver r owns the point cells
 * cells[beg[r] .. beg[r+1]) (beg has n_rec+1 entries) and records from the
 * first step s with s + gate >= delay[r] (gate 0: WGSL, 1: Taichi).
 * channel: the field (VSE_F_*) the receivers sum at the end of each step;
 * -1 = the reference's own, sigma_yy in 2D and sigma_zz in 3D
 * (shader_webgpu_viscoelastic.wgsl:1130, shader_webgpu_3D_viscoelastic.wgsl:1839).
 * A velocity channel (2D VSE_F_VY is the ROI's vertical "vz") is recorded
 * after the velocity update and its source term.  The gradient handle takes
 * sigma_yy only.
 * Replaces b_delay_rec, b_info_rec_pt, b_offset_sensors and the trace
 * buffers (sim_webgpu_viscoelastic.py:174-191). */
int vse_set_receivers(vse_sim *s, int n_rec, const int *beg, const int *cells,
                      const int *delay, int gate, int channel);

/* Zero every state array, the traces and the step counter. */
int vse_reset(vse_sim *s);

/* Run n_steps time steps (stress, velocity, sources+receivers per step),
 * captured in hipGraphs of up to check_every steps; after each chunk the
 * running max |v|^2 is checked and VSE_ERR_UNSTABLE returned if it is +inf
 * or NaN (StabilityError, sim_webgpu_viscoelastic.py:533-535).  The step
 * loop of sim_webgpu_viscoelastic.py:470-537 / sim_taichi_elastic.py:312-335.
 * If elapsed_ms is not NULL it receives the HIP-event time of the loop. */
int vse_run(vse_sim *s, int n_steps, int check_every, float *elapsed_ms);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: vse_run's time steps, launch for launch the same kernels
 * (results identical to vse_run's), enqueued on `stream` (a hipStream_t; NULL =
 * the handle's own stream) with no allocation, no graph or event creation
 * and no host synchronisation.  The step counter, the running max of |v|^2
 * and the 2D per-workgroup |v|^2 slots live in the library's static device
 * scratch (2 + VSE_NOALLOC_SLOTS words per GPU, held until the process
 * ends) instead of the handle's: the call copies the handle's counter and
 * max in, clears the slots, and copies both back after the last step, so
 * vse_get_step / vse_read_v2max report them.  The stability check needs the
 * host: after synchronising the stream, read vse_read_v2max (+inf or NaN =
 * unstable).  Every call on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (other streams, other threads) race.  The handle must
 * be ready (vse_reset or vse_run called once, which allocates its state);
 * a 2D grid needing more than VSE_NOALLOC_SLOTS velocity workgroups (about
 * 4096^2 cells) returns VSE_ERR_ARG. */
#define VSE_NOALLOC_SLOTS 65536
int vse_run_noalloc(vse_sim *s, int n_steps, void *stream);

/* Run n_steps steps launching each kernel between its own HIP events on the
 * library's stream; ms[0..2] receive the average duration of the stress,
 * velocity and sparse (source/receiver/counter) kernels.  In 2D the sources
 * and receivers run inside the velocity kernel, and ms[2] times the one-block
 * step-counter advance.  For roofline
 * measurement; no reference counterpart (the reference only times the whole
 * loop, sim_webgpu_viscoelastic.py:470,537). */
int vse_profile(vse_sim *s, int n_steps, float *ms);

/* Launch shape of a kernel (kernel 0 = stress, 1 = velocity).  2D: xchunk is
 * the number of 4-row groups each 64x4 workgroup walks in x (0 or 1: one,
 * at most 16) and remap the tile order (1, the default: each XCD takes a
 * contiguous band of tiles; 0: round-robin order); nt = 1 (velocity pass
 * only) makes each wavefront walk xchunk consecutive rows with its
 * x-stencil operands in register windows instead; ty does not apply.
 * 3D: workgroup
 * tile of 64 (z) x ty (y) cells, ty in {1,4,8}; nt bit 0 streams the fields
 * touched only at the own cell with non-temporal loads/stores; nt bit 1
 * (velocity pass only, ty >= 2 * FD order, xchunk >= 1) stages the y-stencil
 * operands of sigma_xy, sigma_yy and sigma_yz in LDS per x-plane; each
 * workgroup marches through xchunk x-planes (0 = the plain kernel, one cell
 * per thread on a (z, y, x) grid, where ty and nt do not apply; the velocity
 * pass of a grid of 2^30 cells or more always runs the plain kernel);
 * remap = 1 gives each XCD a contiguous block of tiles instead of the
 * round-robin dispatch order; remap = 2 gives each XCD a contiguous band of
 * y-tiles that it walks x-chunk after x-chunk (2D: 2 acts as 1).  Results are identical for every shape; only speed changes.  Picked by the
 * autotuner (taichi_sim_amd/autotune.py); no reference counterpart (the
 * reference fixes its workgroup to gcd(n, 16) per axis,
 * sim_webgpu_3D_viscoelastic.py:40-42). */
int vse_set_tuning(vse_sim *s, int kernel, int ty, int nt, int xchunk, int remap);

/* Schedule of the step.  VSE_SCHED_TWO_PASS (default): a stress pass and a
 * velocity pass per step, as the reference dispatches them
 * (sim_webgpu_3D_viscoelastic.py:646-651, sim_webgpu_viscoelastic.py:488-509).
 * VSE_SCHED_FUSED: one kernel per step does the velocity update of step n and
 * the stress update of step n+1 (velocities recomputed on a halo of each tile;
 * sigma, v and the velocity-pass CPML memory double-buffered), so the fields
 * cross HBM once per step; results are bit-identical to the two-pass
 * schedule.
 * 3D: viscoelastic simulations with 3 mechanisms and FD order 2 (acc = 2);
 * anything else returns VSE_ERR_ARG.  Its launch shape is vse_set_tuning
 * kernel 2 (ty in {4, 8}, nt 0, xchunk >= 1, remap).
 * 2D: every variant and order; a wavefront owns a strip of 64 - 2*ord columns
 * and marches xchunk rows (vse_set_tuning kernel 2: nt 0, xchunk at least
 * about 4 so that the launch keeps one |v|^2 slot per workgroup; default 16).
 * vse_profile then reports (fused kernel, 0, sparse kernel or step-counter
 * kernel).  After a run the handle's fields may sit in the second buffer set:
 * vse_field_device_ptr changes, vse_read_* follow.  vse_run_noalloc always
 * runs the two-pass schedule.  Both fused kernels measured slower than the
 * two-pass step on MI355X (register pressure: fewer waves per SIMD); the
 * two-pass schedule stays the default. */
#define VSE_SCHED_TWO_PASS 0
#define VSE_SCHED_FUSED 1
int vse_set_schedule(vse_sim *s, int sched);

/* Arithmetic mode of the forward kernels.  VSE_MATH_STRICT (default) keeps the
 * reference's FP32 operation order term for term (coef*(a-b)/dx as in
 * shader_webgpu_3D_viscoelastic.wgsl:1367-1378 and
 * shader_webgpu_viscoelastic.wgsl:850 in 2D).  VSE_MATH_FAST multiplies
 * by f32(coef/dx), 1/k and (3D) 1/3 instead of dividing: one extra rounding
 * per term, within the FP32 noise of the reference.  2D: the WGSL variant
 * only (the Taichi variant already multiplies by f32(1/dx)); a batch takes
 * its members' mode, which must be the same for all of them. */
#define VSE_MATH_STRICT 0
#define VSE_MATH_FAST 1
int vse_set_math(vse_sim *s, int mode);

/* Copy a field (VSE_F_*) to host, n cells in the reference layout.
 * Replaces the sigma_yy / sigma_zz read-back (sim_webgpu_viscoelastic.py:540-541). */
int vse_read_field(vse_sim *s, int field, float *out);

/* Copy the (n_rows, n_rec) traces to host (sim_webgpu_viscoelastic.py:542-543). */
int vse_read_traces(vse_sim *s, float *out);

/* Zener memory variables of component c (0=xx,1=yy,2=zz,3=xy,4=xz,5=yz;
 * 2D uses 0,1,3) in the reference's (cells, L) interleave. */
int vse_read_memory(vse_sim *s, int comp, float *out);

/* Running max of |v|^2 since the last reset (v_2 of wgsl2d :374-375). */
int vse_read_v2max(vse_sim *s, float *out);

/* Steps executed since the last reset. */
int vse_get_step(vse_sim *s, int *out);

/* Device pointer of a field, for zero-copy use by a tensor library. */
void *vse_field_device_ptr(vse_sim *s, int field);

/* Bytes of device memory held by the handle. */
size_t vse_device_bytes(vse_sim *s);

/* Name of the device the handle runs on (the reference's "gpu_str"). */
const char *vse_device_name(vse_sim *s);

const char *vse_last_error(vse_sim *s);

/* Device-to-device copy throughput (read + write bytes per second, in GB/s)
 * of a plain float4 copy kernel over two `bytes`-sized buffers, `iters`
 * launches on the current device.  No reference counterpart: the achievable-
 * HBM point the roofline is read against, and the known-size kernel that
 * calibrates rocprof's memory counters. */
int vse_copy_bandwidth(size_t bytes, int iters, float *gbs);

/* Throughput (GB/s of bytes read plus written) of the read/write mix of a
 * stencil pass without its arithmetic: n_read read-only arrays and n_rw
 * arrays read and rewritten in place, n floats each (one element of every
 * array per thread), iters launches on the current device.  No reference
 * counterpart: the in-place ceiling the stencil kernels are read against
 * (e.g. the 3D stress pass: 6 read-only, 24 read-and-written).
 * n_read + n_rw <= VSE_STREAMS_MAX. */
#define VSE_STREAMS_MAX 32
int vse_stream_bandwidth(size_t n, int n_read, int n_rw, int iters, float *gbs);

/* Release the handle.  Batches and gradient handles built on s may be
 * destroyed before or after it: destroying s first leaves them detached
 * (a gradient handle frees its device memory at once), and every later call
 * on them except their destroy returns VSE_ERR_STATE. */
void vse_destroy(vse_sim *s);

/* ---- batches: many 2D simulations per launch ---------------------------
 * The reference runs one simulation per emission law, one after another
 * (sim_support/simulator.py:328-337 loops over laws around implementation()).
 * A batch advances up to VSE_BATCH_MAX 2D handles (shots of one specimen, or
 * specimens of different sizes).  Members are packed, in the given order,
 * into groups whose device bytes fit a cache budget (vse_batch_set_group_bytes;
 * default VSE_BATCH_GROUP_BYTES_DEFAULT, sized for the 256 MB Infinity Cache
 * of MI355X).  Each group runs all its steps before the next group starts,
 * as hipGraphs of up to check_every steps in which every step is one stress
 * and one velocity launch over the tiles of all the group's members (the
 * velocity launch also injects their sources and records their receivers;
 * one launch per graph advances their step counters; members recording a
 * velocity channel are gathered by one more launch after each velocity
 * pass); a group of one member runs vse_run on it.  Members must be 2D
 * simulations sharing variant, viscoelastic flag, FD order, number of
 * mechanisms and device; they keep their own state,
 * so vse_read_* on a member reads its results, and every member's result is
 * identical to vse_run on it alone.  A batch does not own its members; a
 * batch one of whose members was destroyed returns VSE_ERR_STATE.  Do not
 * call a member while a batch call is in flight on another thread.  No
 * reference counterpart. */
#define VSE_BATCH_MAX 256
#define VSE_BATCH_GROUP_BYTES_DEFAULT ((size_t)224 << 20)
typedef struct vse_batch vse_batch;
int vse_batch_create(vse_sim *const *sims, int n, vse_batch **out);
/* Advance every member n_steps; VSE_ERR_UNSTABLE names the first unstable
 * member in vse_batch_last_error() (and in that member's vse_last_error()).
 * elapsed_ms (optional) sums the device time of every group. */
int vse_batch_run(vse_batch *b, int n_steps, int check_every, float *elapsed_ms);
/* Average (stress, velocity, step-counter advance) durations in ms of one
 * step of all members in one launch per pass (ignores the grouping), over
 * n_steps steps, each launch bracketed by HIP events.  Advances the members. */
int vse_batch_profile(vse_batch *b, int n_steps, float *ms);
/* Each workgroup of a batched launch covers 64 (y) x 4*rows (x) cells of one
 * member, walking its `rows` groups of 4 x-rows in turn (default 1, 1..64).
 * Results are identical for every value; only speed changes. */
int vse_batch_set_rows(vse_batch *b, int rows);
/* Cache budget of a group in bytes of member device memory; 0 puts every
 * member in one group (all members advance in lockstep).  Speed only. */
int vse_batch_set_group_bytes(vse_batch *b, size_t bytes);
/* Schedule of vse_batch_run.  VSE_BATCH_SCHED_GRAPH (the default) is the one
 * described above.  VSE_BATCH_SCHED_PERSISTENT runs every group, a group of
 * one member included, as cooperative launches of up to check_every steps
 * (at most 1024): the workgroups, all resident at once, stride over the
 * group's tiles, and grid-wide barriers separate the passes, so a group's
 * steps are one launch instead of two or three per step.  Results are
 * identical to the graph schedule; only speed changes.  VSE_ERR_HIP when
 * the device cannot launch cooperative kernels.  No reference counterpart:
 * the reference runs one simulation per emission law in a host loop
 * (sim_support/simulator.py:328-337). */
#define VSE_BATCH_SCHED_GRAPH 0
#define VSE_BATCH_SCHED_PERSISTENT 1
int vse_batch_set_schedule(vse_batch *b, int schedule);
/* Number of groups; writes the first member index of each (up to max_groups). */
int vse_batch_groups(vse_batch *b, int *first, int max_groups);
int vse_batch_size(vse_batch *b);
const char *vse_batch_last_error(vse_batch *b);
void vse_batch_destroy(vse_batch *b);

/* ---- gradient of the receiver misfit (2D viscoelastic, WGSL variant) ----
 * No reference counterpart: the reference only runs the forward time loop.
 * A vse_grad handle differentiates the traces of one 2D viscoelastic
 * simulation (its material, CPML, sources and receivers as set on `s`) with
 * respect to the coefficients the kernels use, per cell:
 *   VSE_GRAD_LAMBDA, VSE_GRAD_MU  lambda and mu at the normal-stress point
 *                                 (formed from rho, cp, cs at x and x+1,
 *                                 shader_webgpu_viscoelastic.wgsl:863-868)
 *   VSE_GRAD_MU_XY                mu at the shear point (y and y+1, :948-951)
 *   VSE_GRAD_RHO_VX               rho at the vx point (the cell, :1019)
 *   VSE_GRAD_RHO_VY               the 4-point rho at the vy point (:1035-1036),
 *                                 including its use by the source term
 *   VSE_GRAD_TAU                  a per-cell scale of every relaxation time
 *                                 (1 = the simulation's own tau)
 * The chain rule to rho, cp, cs (or rho, lambda, mu) is left to the caller
 * (taichi_sim_amd/adjoint.py does it in FP64 torch).
 * vse_grad_forward runs the forward with the tau scale (vse_grad_set_tau_scale)
 * keeping a checkpoint every seg_len steps (0: ceil(sqrt(n_rows))); its traces
 * equal vse_run's when the scale is 1.  vse_grad_misfit takes the observed
 * traces, returns 0.5*sum((traces - observed)^2) accumulated in FP64 and
 * makes the residual the adjoint source; vse_grad_backward (grad_traces NULL)
 * then runs the reverse sweep -- segment recompute from the checkpoints plus
 * the hand-written adjoint kernels -- accumulating the gradient in FP64.
 * grad_traces non-NULL supplies the adjoint source (d misfit / d traces)
 * directly.  Wavefields are FP32; only the misfit and the gradient
 * accumulators are FP64.  Each step's gradient terms are formed in FP32 and
 * added into FP32 per-cell partials, which go into the FP64 maps after every
 * step n with n % flush_every == 0 (vse_grad_set_flush: default 32, 1 = an
 * FP64 time accumulation of every step's terms, at one more kernel per step). */
#define VSE_GRAD_LAMBDA 0
#define VSE_GRAD_MU 1
#define VSE_GRAD_MU_XY 2
#define VSE_GRAD_RHO_VX 3
#define VSE_GRAD_RHO_VY 4
#define VSE_GRAD_TAU 5
typedef struct vse_grad vse_grad;
int vse_grad_create(vse_sim *s, int seg_len, vse_grad **out);
int vse_grad_set_tau_scale(vse_grad *g, const float *scale);   /* (n cells), NULL = all 1 */
int vse_grad_forward(vse_grad *g, int n_steps, int check_every, float *traces, float *elapsed_ms);
int vse_grad_misfit(vse_grad *g, const float *observed, double *misfit);
int vse_grad_backward(vse_grad *g, const float *grad_traces, float *elapsed_ms);
/* Steps per FP64 flush of the FP32 gradient partials, 1 .. 4096 (default 32). */
int vse_grad_set_flush(vse_grad *g, int flush_every);
/* With sqrt(N) checkpoints the reverse sweep recomputes segment s-1 on a
 * second stream while the adjoint of segment s runs (overlap 1, the default;
 * it holds one more segment of levels in device memory).  overlap 0
 * recomputes each segment before its adjoint on the simulation's stream.  The
 * gradient is the same either way. */
int vse_grad_set_overlap(vse_grad *g, int overlap);
/* FP64 gradient map VSE_GRAD_* (n cells, grid layout) of the last backward. */
int vse_grad_read(vse_grad *g, int which, double *out);
/* Field 0..4 = vx, vy, sxx, syy, sxy of the last forward's final state. */
int vse_grad_read_state(vse_grad *g, int field, float *out);
int vse_grad_segment(vse_grad *g);
size_t vse_grad_device_bytes(vse_grad *g);
const char *vse_grad_last_error(vse_grad *g);
void vse_grad_destroy(vse_grad *g);

#ifdef __cplusplus
}
#endif
#endif
