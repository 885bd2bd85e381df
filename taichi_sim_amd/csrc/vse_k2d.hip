// vse_k2d.hip -- 2D stress and velocity kernels (elastic / Zener, CPML).
//
// Update semantics follow shader_webgpu_viscoelastic.wgsl:829-1072 (WGSL
// variant) and sim_taichi_elastic.py:138-309 (Taichi variant); the CPU oracle
// restates the same order.  One thread per cell, blocks of 64 (fast y axis)
// x 4 (x axis): every wavefront reads 256 contiguous bytes per field, and the
// +-ORD stencil neighbours come from L1/L2, so each field streams from HBM
// about once per pass.
#include "vse_device.h"

namespace {

constexpr int BY = 64, BX = 4;

// LC > 0: the mechanism count is the compile-time LC and every memory-variable
// load of a point is issued before its arithmetic (same operations, same
// order: bit-identical to the runtime-L loop, which waits on each
// mechanism's loads in turn).
template <int VAR, bool VISCO, int ORD, int LC = 0>
__device__ __forceinline__ void stress2d_cell(const StepArgs &A, int x, int y) {
  const int nx = A.nx, ny = A.ny;
  if (x >= nx || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long id = (long)x * ny + y;
  const float dt = A.dt;

  // ---- normal stresses: x in the D+ range, y in the D- range (wgsl2d :842-846)
  if (x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last) {
    float dvx = 0.f, dvy = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      dvx += fd_term<VAR>(A.C[c], A.vx[id + (long)(c + 1) * ny], A.vx[id - (long)c * ny], A.dx, A.one_dx);
      dvy += fd_term<VAR>(A.C[c], A.vy[id + c], A.vy[id - c - 1], A.dy, A.one_dy);
    }
    const int jx = x - off, jy = y - off;
    const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy);
    dvx = cpml_stretch(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.psi[P_DVX_DX], (long)bx * ny + y, jx, bx >= 0, dvx);
    dvy = cpml_stretch(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.psi[P_DVY_DY], (long)x * A.ax[1].nb + by, jy, by >= 0, dvy);

    const float rho_h = 0.5f * (A.rho[id + ny] + A.rho[id]);
    const float cp_h = 0.5f * (A.cp[id + ny] + A.cp[id]);
    const float csa = A.cs[id + ny], csb = A.cs[id];
    const float cs_l = 0.5f * (csa + csb);
    const float cs_m = (fminf(csa, csb) == 0.f) ? 0.f : cs_l;
    const float lambda = rho_h * (cp_h * cp_h - 2.0f * cs_l * cs_l);
    const float mu = rho_h * (cs_m * cs_m);
    if constexpr (VAR == VSE_VAR_TAICHI) {
      const float dl = lambda * (dvx + dvy);
      A.sxx[id] += (dl + 2.0f * mu * dvx) * dt;
      A.syy[id] += (dl + 2.0f * mu * dvy) * dt;
    } else {
      const float lp2m = lambda + 2.0f * mu;
      if constexpr (VISCO && LC > 0) {
        const float lpm = lambda + mu;
        float oxx[LC], oyy[LC];
#pragma unroll
        for (int l = 0; l < LC; l++) {
          oxx[l] = A.r[0][l * A.ncell + id];
          oyy[l] = A.r[1][l * A.ncell + id];
        }
        const float sxx0 = A.sxx[id], syy0 = A.syy[id];
        float sum_xx = 0.f, sum_yy = 0.f;
#pragma unroll
        for (int l = 0; l < LC; l++) {
          const float nxx = ((oxx[l] + (dvx + dvy) * A.zk.phi_p[l] - oxx[l] * A.zk.h_p[l]) * A.zk.m_p[l]);
          const float nyy = ((oyy[l] + 0.5f * (dvx - dvy) * A.zk.phi_s[l] - oyy[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
          sum_xx += oxx[l] + nxx;
          sum_yy += oyy[l] + nyy;
          A.r[0][l * A.ncell + id] = nxx;
          A.r[1][l * A.ncell + id] = nyy;
        }
        A.sxx[id] = sxx0 + (lp2m * dvx + lambda * dvy + 0.5f * lpm * sum_xx + mu * sum_yy) * dt;
        A.syy[id] = syy0 + (lambda * dvx + lp2m * dvy + 0.5f * lpm * sum_xx - mu * sum_yy) * dt;
      } else if constexpr (VISCO) {
        const float lpm = lambda + mu;
        float sum_xx = 0.f, sum_yy = 0.f;
        for (int l = 0; l < A.L; l++) {
          float *pxx = A.r[0] + l * A.ncell + id, *pyy = A.r[1] + l * A.ncell + id;
          const float oxx = *pxx, oyy = *pyy;
          const float nxx = ((oxx + (dvx + dvy) * A.zk.phi_p[l] - oxx * A.zk.h_p[l]) * A.zk.m_p[l]);
          const float nyy = ((oyy + 0.5f * (dvx - dvy) * A.zk.phi_s[l] - oyy * A.zk.h_s[l]) * A.zk.m_s[l]);
          sum_xx += oxx + nxx;
          sum_yy += oyy + nyy;
          *pxx = nxx;
          *pyy = nyy;
        }
        A.sxx[id] = A.sxx[id] + (lp2m * dvx + lambda * dvy + 0.5f * lpm * sum_xx + mu * sum_yy) * dt;
        A.syy[id] = A.syy[id] + (lambda * dvx + lp2m * dvy + 0.5f * lpm * sum_xx - mu * sum_yy) * dt;
      } else {
        A.sxx[id] = A.sxx[id] + (lp2m * dvx + lambda * dvy) * dt;
        A.syy[id] = A.syy[id] + (lambda * dvx + lp2m * dvy) * dt;
      }
    }
  }

  // ---- shear stress: x in the D- range, y in the D+ range (wgsl2d :927-931)
  if (x >= last + 1 && x < nx - last && y >= last && y < ny - last - 1) {
    float dvydx = 0.f, dvxdy = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      dvydx += fd_term<VAR>(A.C[c], A.vy[id + (long)c * ny], A.vy[id - (long)(c + 1) * ny], A.dx, A.one_dx);
      dvxdy += fd_term<VAR>(A.C[c], A.vx[id + c + 1], A.vx[id - c], A.dy, A.one_dy);
    }
    const int jx = x - off, jy = y - off;
    const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy);
    dvydx = cpml_stretch(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.psi[P_DVY_DX], (long)bx * ny + y, jx, bx >= 0, dvydx);
    dvxdy = cpml_stretch(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.psi[P_DVX_DY], (long)x * A.ax[1].nb + by, jy, by >= 0, dvxdy);

    const float rho_h = 0.5f * (A.rho[id + 1] + A.rho[id]);
    const float csa = A.cs[id + 1], csb = A.cs[id];
    const float cs_h = (fminf(csa, csb) == 0.f) ? 0.f : 0.5f * (csa + csb);
    const float mu = rho_h * (cs_h * cs_h);
    if constexpr (VISCO && VAR == VSE_VAR_WGSL && LC > 0) {
      float o[LC];
#pragma unroll
      for (int l = 0; l < LC; l++) o[l] = A.r[3][l * A.ncell + id];
      const float sxy0 = A.sxy[id];
      float sum_xy = 0.f;
#pragma unroll
      for (int l = 0; l < LC; l++) {
        const float n = ((o[l] + (dvydx + dvxdy) * A.zk.phi_s[l] - o[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
        sum_xy += o[l] + n;
        A.r[3][l * A.ncell + id] = n;
      }
      A.sxy[id] = sxy0 + (mu * (dvydx + dvxdy) + 0.5f * mu * sum_xy) * dt;
    } else if constexpr (VISCO && VAR == VSE_VAR_WGSL) {
      float sum_xy = 0.f;
      for (int l = 0; l < A.L; l++) {
        float *pxy = A.r[3] + l * A.ncell + id;
        const float o = *pxy;
        const float n = ((o + (dvydx + dvxdy) * A.zk.phi_s[l] - o * A.zk.h_s[l]) * A.zk.m_s[l]);
        sum_xy += o + n;
        *pxy = n;
      }
      A.sxy[id] = A.sxy[id] + (mu * (dvydx + dvxdy) + 0.5f * mu * sum_xy) * dt;
    } else {
      A.sxy[id] = A.sxy[id] + (dvxdy + dvydx) * mu * dt;
    }
  }
}

template <int VAR, bool VISCO, int ORD>
__device__ __forceinline__ unsigned velocity2d_cell(const StepArgs &A, int x, int y) {
  const int nx = A.nx, ny = A.ny;
  const bool valid = x < nx && y < ny;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long id = (long)x * ny + y;
  const float dt = A.dt;
  unsigned v2bits = 0;
  if (valid) {
    float vxn = A.vx[id], vyn = A.vy[id];
    const bool in_vx = x >= last + 1 && x < nx - last && y >= last + 1 && y < ny - last;
    const bool in_vy = x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1;
    // vx (wgsl2d :998-1025)
    if (in_vx) {
      float d1 = 0.f, d2 = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += fd_term<VAR>(A.C[c], A.sxx[id + (long)c * ny], A.sxx[id - (long)(c + 1) * ny], A.dx, A.one_dx);
        d2 += fd_term<VAR>(A.C[c], A.sxy[id + c], A.sxy[id - c - 1], A.dy, A.one_dy);
      }
      const int jx = x - off, jy = y - off;
      const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy);
      d1 = cpml_stretch(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.psi[P_DSXX_DX], (long)bx * ny + y, jx, bx >= 0, d1);
      d2 = cpml_stretch(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.psi[P_DSXY_DY], (long)x * A.ax[1].nb + by, jy, by >= 0, d2);
      const float rho = A.rho[id];
      if (rho > 0.f) {
        if constexpr (VAR == VSE_VAR_TAICHI) vxn += (d1 + d2) * dt / rho;
        else vxn = (d1 + d2) * dt / rho + vxn;
      }
    } else if constexpr (VAR == VSE_VAR_TAICHI) {
      vxn = 0.f;     // per-component Dirichlet (sim_taichi_elastic.py:266-268)
    }
    // vy (wgsl2d :1027-1057; sim_taichi_elastic.py:270-304)
    if (in_vy) {
      float d1 = 0.f, d2 = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += fd_term<VAR>(A.C[c], A.sxy[id + (long)(c + 1) * ny], A.sxy[id - (long)c * ny], A.dx, A.one_dx);
        d2 += fd_term<VAR>(A.C[c], A.syy[id + c + 1], A.syy[id - c], A.dy, A.one_dy);
      }
      const int jx = x - off, jy = y - off;
      const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy);
      d1 = cpml_stretch(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.psi[P_DSXY_DX], (long)bx * ny + y, jx, bx >= 0, d1);
      d2 = cpml_stretch(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.psi[P_DSYY_DY], (long)x * A.ax[1].nb + by, jy, by >= 0, d2);
      // in the vy range x+1 and y+1 are inside the grid
      if constexpr (VAR == VSE_VAR_TAICHI) {
        const float rho = 0.25f * (A.rho[id + 1] + A.rho[id] + A.rho[id + ny + 1] + A.rho[id + ny]);
        if (rho > 0.f) {
          const int is = A.src_map[id];
          float vs = 0.f;
          const int s = *A.step;
          if (is != -1 && s < A.src_rows) vs = A.src[(long)s * A.src_cols + is] * dt / rho;
          vyn += vs + (d1 + d2) * dt / rho;
        }
      } else {
        const float rho = 0.25f * (A.rho[id] + A.rho[id + ny] + A.rho[id + ny + 1] + A.rho[id + 1]);
        if (rho > 0.f) vyn = (d1 + d2) * dt / rho + vyn;
      }
    } else {
      if constexpr (VAR == VSE_VAR_WGSL) vxn = 0.f;   // the vy else zeroes both (wgsl2d :1054-1057)
      vyn = 0.f;
    }
    A.vx[id] = vxn;
    A.vy[id] = vyn;
    v2bits = __float_as_uint(vxn * vxn + vyn * vyn);
  }
  return v2bits;
}

template <int VAR, bool VISCO, int ORD, int LC = 0>
__global__ __launch_bounds__(BX * BY) void k_stress2d(StepArgs A) {
  stress2d_cell<VAR, VISCO, ORD, LC>(A, blockIdx.y * BX + threadIdx.y, blockIdx.x * BY + threadIdx.x);
}

template <int VAR, bool VISCO, int ORD>
__global__ __launch_bounds__(BX * BY) void k_velocity2d(StepArgs A) {
  const unsigned v2 = velocity2d_cell<VAR, VISCO, ORD>(A, blockIdx.y * BX + threadIdx.y, blockIdx.x * BY + threadIdx.x);
  block_max_u32(v2, A.v2max);
}

// ---- batched: one launch over the tiles of many simulations --------------
// A batch tile is BY (y) x BX*B.rows (x) cells: the workgroup walks its
// B.rows row-groups of BX rows one after another, so the instance lookup and
// the argument loads at its start are paid once per B.rows row-groups.
// Workgroup b belongs to instance i with B.tile_beg[i] <= b < B.tile_beg[i+1];
// its tile inside that instance is b - tile_beg[i], y-tiles fastest.
__device__ __forceinline__ int batch_instance(const BatchArgs &B, int b) {
  int lo = 0, hi = B.n;          // tile_beg is increasing; find the last i with tile_beg[i] <= b
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (B.tile_beg[mid] <= b) lo = mid; else hi = mid;
  }
  return lo;
}

template <int VAR, bool VISCO, int ORD, int LC = 0>
__global__ __launch_bounds__(BX * BY) void k_stress2d_batch(BatchArgs B) {
  const int i = batch_instance(B, blockIdx.x);
  const StepArgs &A = B.args[i];
  const int t = blockIdx.x - B.tile_beg[i];
  const int nyb = (A.ny + BY - 1) / BY;
  const int y = (t % nyb) * BY + threadIdx.x;
  const int x0 = (t / nyb) * BX * B.rows + threadIdx.y;
  for (int k = 0; k < B.rows; k++) stress2d_cell<VAR, VISCO, ORD, LC>(A, x0 + k * BX, y);
}

template <int VAR, bool VISCO, int ORD>
__global__ __launch_bounds__(BX * BY) void k_velocity2d_batch(BatchArgs B) {
  const int i = batch_instance(B, blockIdx.x);
  const StepArgs &A = B.args[i];
  const int t = blockIdx.x - B.tile_beg[i];
  const int nyb = (A.ny + BY - 1) / BY;
  const int y = (t % nyb) * BY + threadIdx.x;
  const int x0 = (t / nyb) * BX * B.rows + threadIdx.y;
  unsigned v2 = 0;
  for (int k = 0; k < B.rows; k++) {
    const unsigned u = velocity2d_cell<VAR, VISCO, ORD>(A, x0 + k * BX, y);
    v2 = v2 > u ? v2 : u;
  }
  block_max_u32(v2, A.v2max);
}

template <int VAR, bool VISCO>
hipError_t launch_ord(bool stress, int ord, const StepArgs &A, hipStream_t st) {
  dim3 block(BY, BX);
  dim3 grid((A.ny + BY - 1) / BY, (A.nx + BX - 1) / BX);
  const bool l3 = VISCO && VAR == VSE_VAR_WGSL && A.L == 3;
#define VSE_CASE(O)                                                                              \
  case O:                                                                                        \
    if (stress && l3) hipLaunchKernelGGL((k_stress2d<VAR, VISCO, O, 3>), grid, block, 0, st, A); \
    else if (stress) hipLaunchKernelGGL((k_stress2d<VAR, VISCO, O>), grid, block, 0, st, A);     \
    else hipLaunchKernelGGL((k_velocity2d<VAR, VISCO, O>), grid, block, 0, st, A);               \
    break;
  switch (ord) {
    VSE_CASE(2) VSE_CASE(3) VSE_CASE(4) VSE_CASE(5) VSE_CASE(6)
    default: return hipErrorInvalidValue;
  }
#undef VSE_CASE
  return hipGetLastError();
}

template <int VAR, bool VISCO>
hipError_t launch_batch_ord(bool stress, int ord, const BatchArgs &B, int total_tiles, hipStream_t st) {
  dim3 block(BY, BX), grid(total_tiles);
  const bool l3 = VISCO && VAR == VSE_VAR_WGSL && B.L == 3;
#define VSE_CASE(O)                                                                              \
  case O:                                                                                        \
    if (stress && l3) hipLaunchKernelGGL((k_stress2d_batch<VAR, VISCO, O, 3>), grid, block, 0, st, B); \
    else if (stress) hipLaunchKernelGGL((k_stress2d_batch<VAR, VISCO, O>), grid, block, 0, st, B); \
    else hipLaunchKernelGGL((k_velocity2d_batch<VAR, VISCO, O>), grid, block, 0, st, B);         \
    break;
  switch (ord) {
    VSE_CASE(2) VSE_CASE(3) VSE_CASE(4) VSE_CASE(5) VSE_CASE(6)
    default: return hipErrorInvalidValue;
  }
#undef VSE_CASE
  return hipGetLastError();
}

hipError_t dispatch(bool stress, int variant, int visco, int ord, const StepArgs &A, hipStream_t st) {
  if (variant == VSE_VAR_TAICHI) return launch_ord<VSE_VAR_TAICHI, false>(stress, ord, A, st);
  if (visco) return launch_ord<VSE_VAR_WGSL, true>(stress, ord, A, st);
  return launch_ord<VSE_VAR_WGSL, false>(stress, ord, A, st);
}

}  // namespace

hipError_t vse_launch_batch2d(bool stress, int variant, int visco, int ord, const BatchArgs &B, int total_tiles,
                              hipStream_t st) {
  if (variant == VSE_VAR_TAICHI) return launch_batch_ord<VSE_VAR_TAICHI, false>(stress, ord, B, total_tiles, st);
  if (visco) return launch_batch_ord<VSE_VAR_WGSL, true>(stress, ord, B, total_tiles, st);
  return launch_batch_ord<VSE_VAR_WGSL, false>(stress, ord, B, total_tiles, st);
}

long vse_tiles2d(int nx, int ny, int rows) {
  return (long)((ny + BY - 1) / BY) * (((nx + BX - 1) / BX + rows - 1) / rows);
}

hipError_t vse_launch_stress2d(int variant, int visco, int ord, const StepArgs &A, hipStream_t st) {
  return dispatch(true, variant, visco, ord, A, st);
}
hipError_t vse_launch_velocity2d(int variant, int visco, int ord, const StepArgs &A, hipStream_t st) {
  return dispatch(false, variant, visco, ord, A, st);
}
