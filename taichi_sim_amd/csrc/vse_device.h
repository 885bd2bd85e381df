// vse_device.h -- device helpers shared by the 2D and 3D kernels.
#pragma once
#include "vse_internal.h"

// One staggered-difference term.  WGSL divides each term by the spacing
// (wgsl2d :850: coef*(a-b)/dx); Taichi multiplies by f32(1/dx)
// (sim_taichi_elastic.py:154).  Keeping the reference's form keeps FP32
// rounding term for term (the library is built with -ffp-contract=off and
// IEEE division).
template <int VAR>
__device__ __forceinline__ float fd_term(float C, float a, float b, float d, float one_d) {
  if constexpr (VAR == VSE_VAR_TAICHI) return C * (a - b) * one_d;
  else return C * (a - b) / d;
}

// A whole field as a raw buffer: a scalar (SGPR) resource descriptor over its
// bytes, addressed with a 32-bit per-lane offset (no 64-bit address per
// load).  The hardware range-checks every access: an index outside
// [0, ncell) -- a negative one wraps above the range -- reads 0, so a
// cell's stencil operands can be loaded unconditionally; operands read
// outside an update's interior are never used.  Word 3 0x00020000 is the
// gfx9-family raw 32-bit buffer format.
struct FieldBuf {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ FieldBuf(const float *base, long ncell)
      : r(__builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, (int)(ncell * 4), 0x00020000)) {}
  __device__ __forceinline__ float operator[](int i) const {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)((unsigned)i * 4u), 0, 0));
  }
};

// 3D tile assignment of the x-marching kernels.  Tiles are ordered z-tile
// fastest, then y-tile, then x-chunk.  remap = 0 dispatches them in that
// order (consecutive workgroups, which land on XCDs round-robin, work on
// neighbouring tiles of one plane); remap = 1 gives each XCD (dispatch bid
// lands on XCD bid % 8) a contiguous eighth of the tiles.  remap = 2 gives XCD
// k the y-tiles [k, k+1) * nyb/8 over every x-chunk, z fastest, then y, then
// x-chunk: the XCD's workgroups march its whole y-band chunk after chunk, so
// the planes a chunk re-reads at its start were read by the previous chunk of
// the same XCD shortly before (nyb not a multiple of 8: as remap = 1).
struct Tile { int bz, by, bxc; };
__device__ __forceinline__ Tile tile_of_block(int nzb, int nyb, int nxc, int remap) {
  const int bid = blockIdx.x;
  const int total = nzb * nyb * nxc;
  Tile r;
  if (remap == 2 && (nyb & 7) == 0) {
    const int k = bid & 7, j = bid >> 3, band = nyb >> 3;
    r.bz = j % nzb;
    r.by = k * band + (j / nzb) % band;
    r.bxc = j / (nzb * band);
    return r;
  }
  int t = bid;
  if (remap && (total & 7) == 0) t = (bid & 7) * (total >> 3) + (bid >> 3);
  r.bz = t % nzb;
  r.by = (t / nzb) % nyb;
  r.bxc = t / (nzb * nyb);
  return r;
}

// Compact band index of profile index j, or -1 when j is outside the band.
__device__ __forceinline__ int band_index(const AxisProf &P, int j) {
  return j < P.lo ? j : (j >= P.hi ? P.lo + (j - P.hi) : -1);
}

// CPML recursion psi <- b psi + a D; returns D/k + psi (wgsl2d :854-861).
// Outside the band the recursion is an exact no-op and D is returned as is.
__device__ __forceinline__ float cpml_stretch(const float *a, const float *b, const float *k,
                                              float *psi, long pidx, int j, bool in_band, float D) {
  if (!in_band) return D;
  float pn = b[j] * psi[pidx] + a[j] * D;
  psi[pidx] = pn;
  return D / k[j] + pn;
}

// The same with the FAST choice: D times the precomputed f32(1/k) instead of
// D / k (a rounding-level difference).
template <bool FAST>
__device__ __forceinline__ float cpml_stretch_k(const float *a, const float *b, const float *k, const float *ki,
                                                float *psi, long pidx, int j, bool in_band, float D) {
  if (!in_band) return D;
  float pn = b[j] * psi[pidx] + a[j] * D;
  psi[pidx] = pn;
  if constexpr (FAST) return D * ki[j] + pn;
  else return D / k[j] + pn;
}

// Max over a wavefront, then one atomicMax per wavefront into its
// workgroup's own slot.  The result is not used, so no wavefront waits on
// it, and no two workgroups share an address (a single shared word costs each
// workgroup a device-scope round trip before it can retire).
__device__ __forceinline__ void wave_max_to_slot(unsigned v, unsigned *slot) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned o = __shfl_xor(v, off, 64);
    v = v > o ? v : o;
  }
  if (((threadIdx.x + threadIdx.y * blockDim.x) & 63) == 0 && v) atomicMax(slot, v);
}

// Max over a wavefront, then lane 0 reads the running max at dst and issues
// an atomicMax only when its value would raise it: no LDS and no barrier, and
// once the running max is up to date almost no wavefront touches the shared
// word.  Same result as block_max_u32 (the max does not depend on order).
__device__ __forceinline__ void wave_max_running(unsigned v, unsigned *dst) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned o = __shfl_xor(v, off, 64);
    v = v > o ? v : o;
  }
  if (((threadIdx.x + threadIdx.y * blockDim.x) & 63) == 0 && v &&
      v > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(dst, v);
}

// Max over a 64-wide wavefront, then at most one atomicMax per block.  For |v|^2 >= 0
// the IEEE bit pattern orders like the value; NaN sorts above +inf, so the
// running max also flags NaN (the reference's `vsn2 > 1e38` test does not).
__device__ __forceinline__ void block_max_u32(unsigned v, unsigned *dst) {
  for (int off = 32; off > 0; off >>= 1) {
    unsigned o = __shfl_xor(v, off, 64);
    v = v > o ? v : o;
  }
  __shared__ unsigned wmax[16];
  const int tid = threadIdx.x + threadIdx.y * blockDim.x + threadIdx.z * blockDim.x * blockDim.y;
  const int nthr = blockDim.x * blockDim.y * blockDim.z;
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) wmax[wave] = v;
  __syncthreads();
  if (tid == 0) {
    unsigned m = 0;
    for (int w = 0; w < (nthr + 63) / 64; w++) m = m > wmax[w] ? m : wmax[w];
    // Read the running max first: once it is up to date most workgroups skip
    // the atomic, instead of every workgroup of the launch serialising on one
    // address.  A stale read only costs an atomic that changes nothing.
    if (m && m > __hip_atomic_load(dst, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(dst, m);
  }
}
