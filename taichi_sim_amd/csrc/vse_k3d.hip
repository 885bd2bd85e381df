// vse_k3d.hip -- 3D stress and velocity kernels (elastic / Zener, CPML).
//
// Update semantics follow shader_webgpu_3D_viscoelastic.wgsl:1340-1773; the
// CPU oracle restates the same order.  One thread per cell; blocks of 64
// cells along the contiguous z axis x 4 along y, grid (z-blocks, y-blocks, x).
#include "vse_device.h"

namespace {

constexpr int BZ = 64, BYY = 4;

template <bool VISCO, int ORD>
__global__ __launch_bounds__(BZ * BYY) void k_stress3d(StepArgs A) {
  const int z = blockIdx.x * BZ + threadIdx.x;
  const int y = blockIdx.y * BYY + threadIdx.y;
  const int x = blockIdx.z;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  if (z >= nz || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long sy = nz, sx = (long)ny * nz;
  const long id = x * sx + y * sy + z;
  const float dt = A.dt;
  const int jx = x - off, jy = y - off, jz = z - off;
  const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
  const int nbx = A.ax[0].nb, nby = A.ax[1].nb, nbz = A.ax[2].nb;
  // compact psi indices for x-, y- and z-derivatives at this cell
  const long px = (long)bx * sx + y * sy + z;
  const long py = (long)x * nby * nz + (long)by * nz + z;
  const long pz = ((long)x * ny + y) * nbz + bz;
  (void)nbx;

  // ---- normal stresses (wgsl3d :1356-1453)
  if (x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last && z >= last + 1 && z < nz - last) {
    float dvx = 0.f, dvy = 0.f, dvz = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      dvx += A.C[c] * (A.vx[id + (c + 1) * sx] - A.vx[id - c * sx]) / A.dx;
      dvy += A.C[c] * (A.vy[id + c * sy] - A.vy[id - (c + 1) * sy]) / A.dy;
      dvz += A.C[c] * (A.vz[id + c] - A.vz[id - c - 1]) / A.dz;
    }
    dvx = cpml_stretch(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.psi[P_DVX_DX], px, jx, bx >= 0, dvx);
    dvy = cpml_stretch(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.psi[P_DVY_DY], py, jy, by >= 0, dvy);
    dvz = cpml_stretch(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.psi[P_DVZ_DZ], pz, jz, bz >= 0, dvz);
    const float rho = A.rho[id], cp = A.cp[id], cs = A.cs[id];
    const float lambda = rho * (cp * cp - 2.0f * cs * cs);
    const float mu = rho * (cs * cs);
    const float lp2m = lambda + 2.0f * mu;
    if constexpr (VISCO) {
      float Sxx = 0.f, Syy = 0.f, Szz = 0.f;
      const float div = dvx + dvy + dvz;
      for (int l = 0; l < A.L; l++) {
        float *pxx = A.r[0] + l * A.ncell + id, *pyy = A.r[1] + l * A.ncell + id, *pzz = A.r[2] + l * A.ncell + id;
        const float oxx = *pxx, oyy = *pyy, ozz = *pzz;
        const float nxx = ((oxx + div * A.zk.phi_p[l] - oxx * A.zk.h_p[l]) * A.zk.m_p[l]);
        const float nyy = ((oyy + (dvx - div / 3.0f) * A.zk.phi_s[l] - oyy * A.zk.h_s[l]) * A.zk.m_s[l]);
        const float nzz = ((ozz + (dvy - div / 3.0f) * A.zk.phi_s[l] - ozz * A.zk.h_s[l]) * A.zk.m_s[l]);
        Sxx += oxx + nxx; Syy += oyy + nyy; Szz += ozz + nzz;
        *pxx = nxx; *pyy = nyy; *pzz = nzz;
      }
      const float l23 = lambda + 2.0f * mu / 3.0f;
      A.sxx[id] = A.sxx[id] + (lp2m * dvx + lambda * (dvy + dvz) + l23 * 0.5f * Sxx + mu * Syy) * dt;
      A.syy[id] = A.syy[id] + (lambda * (dvx + dvz) + lp2m * dvy + l23 * 0.5f * Sxx + mu * Szz) * dt;
      A.szz[id] = A.szz[id] + (lambda * (dvx + dvy) + lp2m * dvz + lp2m * 0.5f * Sxx - (mu / 3.0f) * (Syy + Szz)) * dt;
    } else {
      A.sxx[id] = A.sxx[id] + (lp2m * dvx + lambda * (dvy + dvz)) * dt;
      A.syy[id] = A.syy[id] + (lambda * (dvx + dvz) + lp2m * dvy) * dt;
      A.szz[id] = A.szz[id] + (lambda * (dvx + dvy) + lp2m * dvz) * dt;
    }
  }

  const bool xm = x >= last + 1 && x < nx - last;      // D- range in x
  const bool xp = x >= last && x < nx - last - 1;      // D+ range in x
  const bool yp = y >= last && y < ny - last - 1;
  const bool zp = z >= last && z < nz - last - 1;
  // ---- shear stresses: sigma_xy, sigma_xz, sigma_yz (wgsl3d :1457-1629)
#pragma unroll
  for (int sh = 0; sh < 3; sh++) {
    const bool ok = (sh < 2 ? xm : xp) && yp && zp;
    if (!ok) continue;
    float d1 = 0.f, d2 = 0.f, rho4, cs4;
    float *sig, *rr;
    if (sh == 0) {            // dvy/dx (D-), dvx/dy (D+)
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += A.C[c] * (A.vy[id + c * sx] - A.vy[id - (c + 1) * sx]) / A.dx;
        d2 += A.C[c] * (A.vx[id + (c + 1) * sy] - A.vx[id - c * sy]) / A.dy;
      }
      d1 = cpml_stretch(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.psi[P_DVY_DX], px, jx, bx >= 0, d1);
      d2 = cpml_stretch(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.psi[P_DVX_DY], py, jy, by >= 0, d2);
      rho4 = 0.25f * (A.rho[id + sx] + A.rho[id] + A.rho[id + sy] + A.rho[id + sx + sy]);
      cs4 = 0.25f * (A.cs[id + sx] + A.cs[id] + A.cs[id + sy] + A.cs[id + sx + sy]);
      sig = A.sxy; rr = A.r[3];
    } else if (sh == 1) {     // dvz/dx (D-), dvx/dz (D+)
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += A.C[c] * (A.vz[id + c * sx] - A.vz[id - (c + 1) * sx]) / A.dx;
        d2 += A.C[c] * (A.vx[id + c + 1] - A.vx[id - c]) / A.dz;
      }
      d1 = cpml_stretch(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.psi[P_DVZ_DX], px, jx, bx >= 0, d1);
      d2 = cpml_stretch(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.psi[P_DVX_DZ], pz, jz, bz >= 0, d2);
      rho4 = 0.25f * (A.rho[id + sx] + A.rho[id] + A.rho[id + 1] + A.rho[id + sx + 1]);
      cs4 = 0.25f * (A.cs[id + sx] + A.cs[id] + A.cs[id + 1] + A.cs[id + sx + 1]);
      sig = A.sxz; rr = A.r[4];
    } else {                  // dvz/dy (D+), dvy/dz (D+)
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        d1 += A.C[c] * (A.vz[id + (c + 1) * sy] - A.vz[id - c * sy]) / A.dy;
        d2 += A.C[c] * (A.vy[id + c + 1] - A.vy[id - c]) / A.dz;
      }
      d1 = cpml_stretch(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.psi[P_DVZ_DY], py, jy, by >= 0, d1);
      d2 = cpml_stretch(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.psi[P_DVY_DZ], pz, jz, bz >= 0, d2);
      rho4 = 0.25f * (A.rho[id + sy] + A.rho[id] + A.rho[id + 1] + A.rho[id + sy + 1]);
      cs4 = 0.25f * (A.cs[id + sy] + A.cs[id] + A.cs[id + 1] + A.cs[id + sy + 1]);
      sig = A.syz; rr = A.r[5];
    }
    const float mu = rho4 * (cs4 * cs4);
    if constexpr (VISCO) {
      float S = 0.f;
      const float X = d1 + d2;
      for (int l = 0; l < A.L; l++) {
        float *pr = rr + l * A.ncell + id;
        const float o = *pr;
        const float n = (o + X * A.zk.phi_s[l] - o * A.zk.h_s[l]) * A.zk.m_s[l];
        S += o + n;
        *pr = n;
      }
      sig[id] = sig[id] + (mu * (d1 + d2) + mu * 0.5f * S) * dt;
    } else {
      sig[id] = sig[id] + (d2 + d1) * mu * dt;
    }
  }
}

template <int ORD>
__global__ __launch_bounds__(BZ * BYY) void k_velocity3d(StepArgs A) {
  const int z = blockIdx.x * BZ + threadIdx.x;
  const int y = blockIdx.y * BYY + threadIdx.y;
  const int x = blockIdx.z;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const bool valid = z < nz && y < ny;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long sy = nz, sx = (long)ny * nz;
  const long id = x * sx + y * sy + z;
  const float dt = A.dt;
  unsigned v2bits = 0;
  if (valid) {
    const int jx = x - off, jy = y - off, jz = z - off;
    const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
    const long px = (long)bx * sx + y * sy + z;
    const long py = (long)x * A.ax[1].nb * nz + (long)by * nz + z;
    const long pz = ((long)x * ny + y) * A.ax[2].nb + bz;
    float vxn = A.vx[id], vyn = A.vy[id], vzn = A.vz[id];
    // vx (wgsl3d :1647-1681)
    if (x >= last + 1 && x < nx - last && y >= last + 1 && y < ny - last && z >= last + 1 && z < nz - last) {
      float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        a += A.C[c] * (A.sxx[id + c * sx] - A.sxx[id - (c + 1) * sx]) / A.dx;
        b += A.C[c] * (A.sxy[id + c * sy] - A.sxy[id - (c + 1) * sy]) / A.dy;
        c_ += A.C[c] * (A.sxz[id + c] - A.sxz[id - c - 1]) / A.dz;
      }
      a = cpml_stretch(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.psi[P_DSXX_DX], px, jx, bx >= 0, a);
      b = cpml_stretch(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.psi[P_DSXY_DY], py, jy, by >= 0, b);
      c_ = cpml_stretch(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.psi[P_DSXZ_DZ], pz, jz, bz >= 0, c_);
      const float rho = 0.5f * (A.rho[id + sx] + A.rho[id]);
      if (rho > 0.f) vxn = (a + b + c_) * dt / rho + vxn;
    }
    // vy (wgsl3d :1683-1717)
    if (x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1 && z >= last + 1 && z < nz - last) {
      float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        a += A.C[c] * (A.sxy[id + (c + 1) * sx] - A.sxy[id - c * sx]) / A.dx;
        b += A.C[c] * (A.syy[id + (c + 1) * sy] - A.syy[id - c * sy]) / A.dy;
        c_ += A.C[c] * (A.syz[id + c] - A.syz[id - c - 1]) / A.dz;
      }
      a = cpml_stretch(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.psi[P_DSXY_DX], px, jx, bx >= 0, a);
      b = cpml_stretch(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.psi[P_DSYY_DY], py, jy, by >= 0, b);
      c_ = cpml_stretch(A.ax[2].a, A.ax[2].b, A.ax[2].k, A.psi[P_DSYZ_DZ], pz, jz, bz >= 0, c_);
      const float rho = 0.5f * (A.rho[id + sy] + A.rho[id]);
      if (rho > 0.f) vyn = (a + b + c_) * dt / rho + vyn;
    }
    // vz, with the vy density (wgsl3d :1719-1758)
    if (x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last && z >= last && z < nz - last - 1) {
      float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        a += A.C[c] * (A.sxz[id + (c + 1) * sx] - A.sxz[id - c * sx]) / A.dx;
        b += A.C[c] * (A.syz[id + c * sy] - A.syz[id - (c + 1) * sy]) / A.dy;
        c_ += A.C[c] * (A.szz[id + c + 1] - A.szz[id - c]) / A.dz;
      }
      a = cpml_stretch(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.psi[P_DSXZ_DX], px, jx, bx >= 0, a);
      b = cpml_stretch(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.psi[P_DSYZ_DY], py, jy, by >= 0, b);
      c_ = cpml_stretch(A.ax[2].ah, A.ax[2].bh, A.ax[2].kh, A.psi[P_DSZZ_DZ], pz, jz, bz >= 0, c_);
      const float rho = 0.5f * (A.rho[id + sy] + A.rho[id]);
      if (rho > 0.f) vzn = (a + b + c_) * dt / rho + vzn;
    } else {
      vxn = 0.f; vyn = 0.f; vzn = 0.f;
    }
    A.vx[id] = vxn;
    A.vy[id] = vyn;
    A.vz[id] = vzn;
    v2bits = __float_as_uint(vxn * vxn + vyn * vyn + vzn * vzn);
  }
  block_max_u32(v2bits, A.v2max);
}

}  // namespace

hipError_t vse_launch_stress3d(int visco, int ord, const StepArgs &A, hipStream_t st) {
  dim3 block(BZ, BYY);
  dim3 grid((A.nz + BZ - 1) / BZ, (A.ny + BYY - 1) / BYY, A.nx);
#define VSE_CASE(O)                                                                              \
  case O:                                                                                        \
    if (visco) hipLaunchKernelGGL((k_stress3d<true, O>), grid, block, 0, st, A);                 \
    else hipLaunchKernelGGL((k_stress3d<false, O>), grid, block, 0, st, A);                      \
    break;
  switch (ord) {
    VSE_CASE(2) VSE_CASE(3) VSE_CASE(4) VSE_CASE(5) VSE_CASE(6)
    default: return hipErrorInvalidValue;
  }
#undef VSE_CASE
  return hipGetLastError();
}

hipError_t vse_launch_velocity3d(int visco, int ord, const StepArgs &A, hipStream_t st) {
  (void)visco;
  dim3 block(BZ, BYY);
  dim3 grid((A.nz + BZ - 1) / BZ, (A.ny + BYY - 1) / BYY, A.nx);
#define VSE_CASE(O) \
  case O: hipLaunchKernelGGL((k_velocity3d<O>), grid, block, 0, st, A); break;
  switch (ord) {
    VSE_CASE(2) VSE_CASE(3) VSE_CASE(4) VSE_CASE(5) VSE_CASE(6)
    default: return hipErrorInvalidValue;
  }
#undef VSE_CASE
  return hipGetLastError();
}
