// vse_internal.h -- device-side argument blocks shared by the 2D/3D kernels
// and the C-ABI layer of libvse.
//
// Grid layout is the reference's: C order with the last axis fastest
// (2D index y + x*ny, wgsl2d :18-22; 3D index z + y*nz + x*ny*nz, wgsl3d
// :28-32).  Zener memory variables are stored as L planes per component
// (plane l of component c at r[c] + l*rs, rs = ncell plus a small skew)
// instead of the reference's (cells, L) interleave, so every load is a
// coalesced stream.
// CPML memory variables (psi) are stored only over each axis's active band.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Device code sees the argument blocks' pointers as global-memory pointers
// (address space 1): a block read from device memory (batched launches) then
// still compiles to global_load/global_store instead of flat accesses.  Host
// code sees plain pointers; the layout is the same.  Files without device
// code (the C-ABI layer) define VSE_HOST_SIDE to see plain pointers in both
// compilation passes.
#if defined(__HIP_DEVICE_COMPILE__) && !defined(VSE_HOST_SIDE)
#define VSE_G __attribute__((address_space(1)))
#else
#define VSE_G
#endif

#define VSE_MAX_ORD 6
#define VSE_MAX_SLS 8
#define VSE_VAR_WGSL 0
#define VSE_VAR_TAICHI 1

// psi slots, reference naming (wgsl2d :444-610, wgsl3d :622-997)
enum {
  P_DVX_DX = 0, P_DVY_DY, P_DVZ_DZ, P_DVY_DX, P_DVX_DY, P_DVZ_DX, P_DVX_DZ, P_DVZ_DY, P_DVY_DZ,
  P_DSXX_DX, P_DSXY_DY, P_DSXZ_DZ, P_DSXY_DX, P_DSYY_DY, P_DSYZ_DZ, P_DSXZ_DX, P_DSYZ_DY, P_DSZZ_DZ,
  P_COUNT
};

// psi slots each kernel family touches (the library allocates exactly these
// and checks them before any launch)
static const int kPsiSlots2D[] = {P_DVX_DX, P_DVY_DY, P_DVY_DX, P_DVX_DY,
                                  P_DSXX_DX, P_DSXY_DY, P_DSXY_DX, P_DSYY_DY};

// CPML profiles of one axis.  Profile index j = cell - (ord-1) runs over
// [0, np).  The profile is "active" (a != 0 or k != 1 on the full or half
// grid) only for j < lo or j >= hi; there psi lives in a compact band of
// width nb = lo + (np - hi).  Outside the band a = 0 and k = 1 exactly, so
// psi stays 0 and the stretched derivative equals the plain one bit for bit.
struct AxisProf {
  const VSE_G float *a, *b, *k, *ah, *bh, *kh;
  const VSE_G float *ki, *khi;       // 1/k, 1/k_half (fast arithmetic)
  int lo, hi, np, nb;
};

struct ZenerK {
  float phi_p[VSE_MAX_SLS], phi_s[VSE_MAX_SLS];
  float h_p[VSE_MAX_SLS], h_s[VSE_MAX_SLS];
  float m_p[VSE_MAX_SLS], m_s[VSE_MAX_SLS];
};

struct StepArgs {
  int nx, ny, nz;
  long ncell;
  float dx, dy, dz, dt;
  float one_dx, one_dy, one_dz;      // f32(1/d) (Taichi variant)
  float C[VSE_MAX_ORD];
  float Cd[3][VSE_MAX_ORD];          // f32(C[c] / d_axis) (fast arithmetic)
  AxisProf ax[3];
  VSE_G float *vx, *vy, *vz;
  VSE_G float *sxx, *syy, *szz, *sxy, *sxz, *syz;
  VSE_G float *psi[P_COUNT];
  VSE_G float *r[6];                 // r_xx, r_yy, r_zz, r_xy, r_xz, r_yz (3D); 2D: xx, yy, -, xy
  long rs;                           // mechanism l of a component at r[c] + l * rs (rs >= ncell)
  const VSE_G float *rho, *cp, *cs;
  ZenerK zk;
  int L;
  unsigned *v2max;                   // running max of |v|^2 as float bits
  // 2D: the velocity kernels keep the running max per workgroup in
  // v2slots[0..n_slots) instead; vse_launch_fold_v2 merges them into v2max
  unsigned *v2slots;
  int n_slots;
  const VSE_G int *step;             // device step counter (0-based)
  // Taichi-variant source (inline in the velocity update)
  const VSE_G int *src_map;          // full grid, -1 or column
  const VSE_G float *src;            // (n_rows, src_cols)
  int src_cols, src_rows;
  // 2D single-simulation launches: the velocity kernel also does the step's
  // sparse work -- the WGSL-variant source through src_map (rows flagged in
  // src_rowflag) and the receivers in workgroup 0
  const VSE_G int *src_rowflag;      // per x-row (3D: x-plane): 1 if it holds a source cell
  int n_rec, rec_gate, trace_rows;
  const VSE_G int *rec_beg, *rec_cell, *rec_delay;
  VSE_G float *traces;               // (trace_rows, n_rec)
  // the field the receivers sample (sigma_yy by default); a velocity channel
  // is written by the velocity pass itself, so rec_after = 1 moves its
  // gather into a one-workgroup launch after that pass
  const VSE_G float *rec_field;
  int rec_after;
  // 3D sources by (x, y) row, for the fused step: row r = x*ny + y holds the
  // source points src_rbeg[r] .. src_rbeg[r+1] (their z and column)
  const VSE_G int *src_rbeg, *src_rz, *src_rcol;
};

// Fused 3D step (vse_fused3d.hip): the velocity update of step n followed by
// the stress update of step n+1 in one pass.  `A` holds the inputs -- sigma at
// step n+1 (A.sxx .. A.syz), v at step n (A.vx .. A.vz) and the velocity-pass
// psi at step n -- and the state updated in place (memory variables, the
// stress-pass psi, the step counter, the running max); the outputs go to a
// second set of buffers, because neighbouring workgroups recompute velocities
// in their halos from the inputs.
struct FusedArgs {
  StepArgs A;
  VSE_G float *vo[3];                // v at step n+1: x, y, z
  VSE_G float *so[6];                // sigma at step n+2: xx, yy, zz, xy, xz, yz
  VSE_G float *psio[9];              // velocity-pass psi at step n+1, slots P_DSXX_DX + i
};

// Sparse per-step work: source injection (WGSL variants), receiver gather,
// and the step counter.  One block.
struct SparseArgs {
  int ndim, nx, ny, nz;
  float dt;
  float *vx, *vy, *vz;
  const float *rec_field;            // the receivers' channel (sigma_yy / sigma_zz by default)
  const float *rho;
  unsigned *v2max;
  int *step;
  // sources: n_src_pts cells; value src[s*src_cols + col]
  int n_src_pts;
  const int *src_cell, *src_col;
  const float *src;
  int src_cols, src_rows;
  // receivers: CSR over sensors
  int n_rec;
  const int *rec_beg, *rec_cell;     // rec_beg has n_rec+1 entries
  const int *rec_delay;
  int rec_gate;                      // record when s + rec_gate >= delay
  float *traces;                     // (rows, n_rec)
  int trace_rows;
};

// Launch shape of one 3D kernel (chosen by the autotuner): tile height in y
// (the tile is 64 cells wide in z), non-temporal streaming on/off, and the
// number of x planes each workgroup marches through, and whether tiles are
// remapped so each XCD owns a contiguous block of them.
struct KTune {
  int ty, nt, xchunk, remap;
};

// Arithmetic mode.  STRICT keeps the reference's FP32 operation order term
// for term (coef*(a-b)/d, D/k, div/3): bit-compatible with the CPU oracle.
// FAST multiplies by precomputed f32(coef/d), 1/k and 1/3 instead of
// dividing: one extra rounding per term, results within the FP32 noise of the
// reference (checked against the oracle at 1e-5 rel-L2), far fewer VALU ops.
enum { VSE_MATH_STRICT = 0, VSE_MATH_FAST = 1 };

// Batched 2D step: per-instance argument blocks in device memory and the
// prefix of per-instance workgroup (tile) counts.
#define VSE_MAX_BATCH 256
struct BatchArgs {
  const StepArgs *__restrict__ args;        // device array [n]
  const SparseArgs *__restrict__ sparse;    // device array [n]
  int n;
  int L;                       // Zener mechanisms of every member
  int rows;                    // BX-row groups each workgroup walks in x
  int tile_beg[VSE_MAX_BATCH + 1];
};

// ---- differentiable 2D viscoelastic path (vse_adj2d.hip) -----------------
// One time level of the 2D viscoelastic state.  psi[i] is 2D slot
// kPsiSlots2D[i] in the compact band layout of the forward kernels.
struct DLevel {
  VSE_G float *vx, *vy, *sxx, *syy, *sxy;
  VSE_G float *r[3];                 // xx, yy, xy: L planes each, plane l at r[c] + l * DiffArgs::rs
  VSE_G float *psi[8];
};

// Gradient accumulators (FP64), one per coefficient the kernels use.
enum { VSE_GR_LAMBDA = 0, VSE_GR_MU = 1, VSE_GR_MU_XY = 2, VSE_GR_RHO_VX = 3, VSE_GR_RHO_VY = 4,
       VSE_GR_TAU = 5, VSE_GR_COUNT = 6 };
// The reverse sweep adds the FP32 partials into the FP64 gradients after every
// step n with n % flush == 0, so the grouping depends on n only (not on the
// checkpoint spacing).  VSE_GRAD_FLUSH is the default flush interval.
constexpr int VSE_GRAD_FLUSH = 32;
constexpr int VSE_GRAD_FLUSH_MAX = 4096;

struct DiffArgs {
  int nx, ny, ord;
  long ncell;
  long rs;                           // mechanism stride of r in every level and in bar (rs >= ncell)
  float dx, dy, dt;
  float C[VSE_MAX_ORD];
  float Cd[2][VSE_MAX_ORD];          // f32(C[c] / d_axis): the adjoint kernels' differences
  AxisProf ax[2];
  const VSE_G float *rho, *cp, *cs;
  const VSE_G float *ts;             // per-cell relaxation-time scale
  int ts_unit;                       // 1: every scale is exactly 1 (ts is not read)
  int L;
  DLevel in, out;                    // forward: n -> n+1; adjoint: levels n and n+1
  DLevel bar;                        // adjoint state (in place)
  VSE_G float *abar[8];              // adjoints of the raw derivatives a1..a8
  VSE_G double *g[VSE_GR_COUNT];     // gradients (FP64 time accumulation)
  VSE_G float *gf[VSE_GR_COUNT];     // FP32 per-step partials, added into g every VSE_GRAD_FLUSH steps
  unsigned *v2max;
  ZenerK zk;                         // the constants at scale 1 (vse_create's values)
  // sources and receivers, handled inside the velocity kernels
  int n;                             // the step being advanced / reversed
  const VSE_G int *src_map;          // per cell: source column or -1
  const VSE_G int *src_rowflag;      // per x-row: 1 if it holds a source cell
  const VSE_G float *src;            // (src_rows, src_cols)
  int src_cols, src_rows;
  int n_rec, rec_gate;
  const VSE_G int *rec_beg, *rec_cell, *rec_delay;
  VSE_G float *traces;               // (rows, n_rec), forward
  const VSE_G float *gbar;           // (rows, n_rec), adjoint source
  int record;                        // forward: record traces of step n
};

hipError_t vse_launch_dstress(const DiffArgs &A, hipStream_t st);
hipError_t vse_launch_dvelocity(const DiffArgs &A, hipStream_t st);
hipError_t vse_launch_adj_velocity(const DiffArgs &A, hipStream_t st);
hipError_t vse_launch_src_map(int *map, long ncell, const int *cells, const int *cols, int n, hipStream_t st);
hipError_t vse_launch_adj_stress(const DiffArgs &A, hipStream_t st);
hipError_t vse_launch_grad_flush(const DiffArgs &A, hipStream_t st);
hipError_t vse_launch_misfit(const float *traces, const float *obs, float *resid, long n, double *acc,
                             hipStream_t st);

// launchers (vse_k2d.hip / vse_k3d.hip)
// 2D: t.remap = 1 hands each XCD a contiguous band of tiles (other fields of
// t are not used in 2D)
// math = VSE_MATH_FAST takes the WGSL variant's differences as (a-b)*f32(C/d)
// and its CPML 1/k as a product
hipError_t vse_launch_stress2d(int variant, int visco, int ord, int math, const StepArgs &A, const KTune &t,
                               hipStream_t st);
// 2D velocity (single simulation): advances step *A.step + i and does that
// step's sources and receivers; vse_launch_advance(step, k) moves the counter
// after k such launches
hipError_t vse_launch_velocity2d(int variant, int visco, int ord, int math, const StepArgs &A, const KTune &t, int i,
                                 hipStream_t st);
hipError_t vse_launch_advance(int *step, int n, hipStream_t st);
// receivers of a velocity channel (A.rec_after): the gather of step *A.step + i
// after the velocity pass; the batch form gathers every member with rec_after
hipError_t vse_launch_record2d(const StepArgs &A, int i, hipStream_t st);
hipError_t vse_launch_record2d_batch(const BatchArgs &B, int i, hipStream_t st);
hipError_t vse_launch_stress3d(int visco, int ord, int math, const StepArgs &A, const KTune &t, hipStream_t st);
// true when the fused kernel covers this problem: 3D, viscoelastic with 3
// mechanisms, FD order 2, and every array addressable with 32-bit byte offsets
bool vse_fused3d_supported(long ncell, long rs, int L, int visco, int ord);
// t.ty in {4, 8}: the tile height; t.xchunk: x-planes per workgroup; t.remap: tile order
hipError_t vse_launch_fused3d(int math, const FusedArgs &F, const KTune &t, hipStream_t st);
hipError_t vse_launch_velocity3d(int visco, int ord, int math, const StepArgs &A, const KTune &t, hipStream_t st);
// Fused 2D step (velocity of step *step + i, then stress of the next step)
// reading F.A and writing F.vo[0..1] (vx, vy), F.so[0], F.so[1], F.so[3] (sxx,
// syy, sxy) and F.psio[0, 1, 3, 4] (the velocity-pass psi slots); xchunk rows per
// workgroup.  vse_fused2d_blocks is its workgroup count (each keeps one |v|^2 slot).
long vse_fused2d_blocks(int nx, int ny, int ord, int xchunk);
hipError_t vse_launch_fused2d(int variant, int visco, int ord, int math, const FusedArgs &F, int xchunk, int i,
                              hipStream_t st);

hipError_t vse_launch_sparse(const SparseArgs &S, hipStream_t st);
// the velocity pass (stress = false) advances every member's step *step + i
// and does its sources and receivers; vse_launch_advance_batch moves the
// members' counters after k such steps
hipError_t vse_launch_batch2d(bool stress, int variant, int visco, int ord, int math, const BatchArgs &B,
                              int total_tiles, int i, hipStream_t st);
hipError_t vse_launch_advance_batch(const BatchArgs &B, int n, hipStream_t st);
// persistent schedule: nsteps steps of every member in one cooperative launch
// of `grid` workgroups striding over the total_tiles tiles (grid-wide
// barriers between the passes), then the counters advance by nsteps;
// vse_batch2d_persistent_grid gives the largest grid that can be resident
// (at most total_tiles, a multiple of 8 from 8 on) on the current device
// `bar` holds the launch's grid-barrier words (VSE_GRID_BAR_WORDS; the
// launcher clears the counters before the kernel).  Word VSE_GRID_BAR_FAULT
// is left to the caller: a launch sets it when a barrier wait gave up (its
// results are then invalid), so the caller clears it and reads it back.
#define VSE_GRID_BAR_SUB 64
#define VSE_GRID_BAR_STRIDE 64
#define VSE_GRID_BAR_FAULT ((VSE_GRID_BAR_SUB + 2) * VSE_GRID_BAR_STRIDE)
#define VSE_GRID_BAR_WORDS ((VSE_GRID_BAR_SUB + 3) * VSE_GRID_BAR_STRIDE)
#define VSE_GRID_BAR_SPIN (1u << 18)
// barriers per launch stay below 2^32 / (workgroups per first-level counter)
#define VSE_GRID_BAR_MAX_STEPS 4096
hipError_t vse_batch2d_persistent_grid(int variant, int visco, int ord, int math, int L, int total_tiles, int *grid);
hipError_t vse_launch_batch2d_persistent(int variant, int visco, int ord, int math, const BatchArgs &B,
                                         int total_tiles, int grid, int nsteps, int rec_after, unsigned *bar,
                                         hipStream_t st);

hipError_t vse_launch_batch_ctrl(const BatchArgs &B, unsigned *out, hipStream_t st);
hipError_t vse_launch_fold_v2(const unsigned *slots, int n, unsigned *v2max, hipStream_t st);
long vse_tiles2d(int nx, int ny, int rows);
