// vse_sparse.hip -- per-step sparse work in one block: source injection,
// receiver gather and the device step counter.
//
// * Sources (WGSL variants): vy (2D) / vz (3D) += src[s, col] * dt / rho at
//   each source cell, after the velocity update and anywhere on the grid,
//   with the vy-grid density (wgsl2d :1059-1065, wgsl3d :1760-1766).  The
//   reference reads a full-grid int map per cell; here the cells are a list.
// * Receivers: traces[s, r] accumulates the sigma_yy / sigma_zz values of
//   receiver r's points in table order when s + gate >= delay[r]
//   (wgsl2d :1117-1134 with gate 0; sim_taichi_elastic.py:187-190 with gate 1).
// * Step counter: s = *step at entry; s + 1 is stored at exit (wgsl2d
//   :1136-1141).  One block, so no other block can observe the increment early.
#include "vse_device.h"

namespace {

__device__ __forceinline__ void sparse_body(const SparseArgs &S) {
  __shared__ int s_shared;
  if (threadIdx.x == 0) s_shared = *S.step;
  __syncthreads();
  const int s = s_shared;
  unsigned v2bits = 0;

  if (S.n_src_pts > 0 && s < S.src_rows) {
    for (int i = threadIdx.x; i < S.n_src_pts; i += blockDim.x) {
      const int cell = S.src_cell[i], col = S.src_col[i];
      float rho;
      if (S.ndim == 2) {
        // 0.25*(rho(x,y)+rho(x+1,y)+rho(x+1,y+1)+rho(x,y+1)), out of grid -> 0
        const int x = cell / S.ny, y = cell % S.ny;
        const bool xe = x + 1 < S.nx, ye = y + 1 < S.ny;
        rho = 0.25f * (S.rho[cell] + (xe ? S.rho[cell + S.ny] : 0.f) + ((xe && ye) ? S.rho[cell + S.ny + 1] : 0.f)
                       + (ye ? S.rho[cell + 1] : 0.f));
      } else {
        // 0.5*(rho(x,y+1,z)+rho(x,y,z))
        const int yz = cell % (S.ny * S.nz);
        const int y = yz / S.nz;
        rho = 0.5f * ((y + 1 < S.ny ? S.rho[cell + S.nz] : 0.f) + S.rho[cell]);
      }
      if (rho > 0.f) {
        const float val = S.src[(long)s * S.src_cols + col];
        float *v = S.ndim == 2 ? S.vy : S.vz;
        v[cell] = v[cell] + val * S.dt / rho;
        float m = S.vx[cell] * S.vx[cell] + S.vy[cell] * S.vy[cell];
        if (S.ndim == 3) m = m + S.vz[cell] * S.vz[cell];
        const unsigned u = __float_as_uint(m);
        v2bits = v2bits > u ? v2bits : u;
      }
    }
  }

  if (s < S.trace_rows) {
    for (int r = threadIdx.x; r < S.n_rec; r += blockDim.x) {
      if (s + S.rec_gate >= S.rec_delay[r]) {
        float *t = S.traces + (long)s * S.n_rec + r;
        float acc = *t;
        for (int p = S.rec_beg[r]; p < S.rec_beg[r + 1]; p++) acc = acc + S.rec_field[S.rec_cell[p]];
        *t = acc;
      }
    }
  }

  block_max_u32(v2bits, S.v2max);
  __syncthreads();
  if (threadIdx.x == 0) *S.step = s + 1;
}

__global__ __launch_bounds__(1024) void k_sparse(SparseArgs S) { sparse_body(S); }

// Largest of a workgroup's values (every thread calls it; thread 0 gets it).
__device__ __forceinline__ unsigned block_reduce_max(unsigned v) {
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned o = __shfl_xor(v, off, 64);
    v = v > o ? v : o;
  }
  __shared__ unsigned w[16];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned m = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; k++) m = m > w[k] ? m : w[k];
  return m;
}

// (running max of |v|^2, step) of member m of a batch into out[2m], out[2m+1];
// one workgroup per member, merging the member's workgroup slots
__global__ __launch_bounds__(256) void k_batch_ctrl(BatchArgs B, unsigned *out) {
  const int m = blockIdx.x;
  const StepArgs &A = B.args[m];
  unsigned v = 0;
  for (int k = threadIdx.x; k < A.n_slots; k += blockDim.x) v = v > A.v2slots[k] ? v : A.v2slots[k];
  v = block_reduce_max(v);
  if (threadIdx.x == 0) {
    const unsigned g = *B.sparse[m].v2max;
    out[2 * m] = g > v ? g : v;
    out[2 * m + 1] = (unsigned)*B.sparse[m].step;
  }
}

// v2max = max(v2max, slots[0..n)): each workgroup folds a strided share of
// the slots and adds its maximum with one atomicMax
__global__ __launch_bounds__(256) void k_fold_v2(const unsigned *slots, int n, unsigned *v2max) {
  unsigned v = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    v = v > slots[k] ? v : slots[k];
  v = block_reduce_max(v);
  if (threadIdx.x == 0 && v) atomicMax(v2max, v);
}

}  // namespace

hipError_t vse_launch_fold_v2(const unsigned *slots, int n, unsigned *v2max, hipStream_t st) {
  const int blocks = n / 2048 + 1;      // about 8 slots per thread, at most a few hundred workgroups
  hipLaunchKernelGGL(k_fold_v2, dim3(blocks < 512 ? blocks : 512), dim3(256), 0, st, slots, n, v2max);
  return hipGetLastError();
}

hipError_t vse_launch_batch_ctrl(const BatchArgs &B, unsigned *out, hipStream_t st) {
  hipLaunchKernelGGL(k_batch_ctrl, dim3(B.n), dim3(256), 0, st, B, out);
  return hipGetLastError();
}

hipError_t vse_launch_sparse(const SparseArgs &S, hipStream_t st) {
  hipLaunchKernelGGL(k_sparse, dim3(1), dim3(1024), 0, st, S);
  return hipGetLastError();
}
