// vse_k3d.hip -- 3D stress and velocity kernels (elastic / Zener, CPML).
//
// Update semantics follow shader_webgpu_3D_viscoelastic.wgsl:1340-1773; the
// CPU oracle restates the same order and every expression keeps it.
//
// Memory plan (both passes are HBM-bound stencils, ~1.4 flop/byte):
// * a workgroup owns a (TY x 64) column tile in (y, z) -- z is the contiguous
//   axis, so each wavefront reads 256 B per field -- and marches along x (the
//   slowest axis, plane stride ny*nz) over a chunk of planes;
// * the x-direction stencil window of each thread's own column lives in
//   registers and advances one plane per iteration, so every plane of a
//   stencil field is fetched once instead of 2*ORD times;
// * y/z neighbours at the current plane are plain loads that hit L1/L2: the
//   tile itself fetched them this iteration, and tiles adjacent in y run on
//   the same XCD (tile index remapped so each XCD owns contiguous tiles);
// * fields touched only at the thread's own cell (memory variables r, the
//   updated field, psi, cp) are streamed with non-temporal loads/stores so
//   they do not evict the stencil planes from L2.
#include <algorithm>

#include "vse_device.h"

namespace {

constexpr int BZ = 64;

// Streamed accesses: non-temporal when NT (bypass-friendly, keeps L2 for the
// stencil planes), ordinary otherwise; the autotuner picks.
template <bool NT, typename T>
__device__ __forceinline__ T sload(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void sstore(T v, T *p) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Tile assignment: linear dispatch order bid lands on XCD bid % 8; give each
// XCD a contiguous run of (z-tile fastest, y-tile, x-chunk) tiles.
struct Tile { int bz, by, bxc; };
__device__ __forceinline__ Tile tile_of_block(int nzb, int nyb, int nxc) {
  const int bid = blockIdx.x;
  const int total = nzb * nyb * nxc;
  int t = bid;
  if ((total & 7) == 0) t = (bid & 7) * (total >> 3) + (bid >> 3);
  Tile r;
  r.bz = t % nzb;
  r.by = (t / nzb) % nyb;
  r.bxc = t / (nzb * nyb);
  return r;
}

__device__ __forceinline__ float ld_plane(const float *f, long sx, int x, int nx, long col) {
  return (x >= 0 && x < nx) ? f[x * sx + col] : 0.f;
}

template <bool VISCO, int ORD, int TY, bool NT>
__global__ __launch_bounds__(BZ * TY) void k_stress3d(StepArgs A, int xchunk, int nzb, int nyb, int nxc) {
  const Tile T = tile_of_block(nzb, nyb, nxc);
  const int z = T.bz * BZ + threadIdx.x;
  const int y = T.by * TY + threadIdx.y;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  if (z >= nz || y >= ny) return;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long sy = nz, sx = (long)ny * nz;
  const long col = y * sy + z;
  const float dt = A.dt;
  const int x0 = T.bxc * xchunk, x1 = min(nx, x0 + xchunk);
  const int jy = y - off, jz = z - off;
  const int by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
  const int nby = A.ax[1].nb, nbz = A.ax[2].nb;
  const bool yp = y >= last && y < ny - last - 1;
  const bool zp = z >= last && z < nz - last - 1;
  const bool ym = y >= last + 1 && y < ny - last;
  const bool zm = z >= last + 1 && z < nz - last;

  // x windows of the own column: vx at x-(ORD-1)..x+ORD, vy/vz at x-ORD..x+ORD-1
  float wvx[2 * ORD], wvy[2 * ORD], wvz[2 * ORD];
#pragma unroll
  for (int i = 0; i < 2 * ORD - 1; i++) {
    wvx[i + 1] = ld_plane(A.vx, sx, x0 - (ORD - 1) + i, nx, col);
    wvy[i + 1] = ld_plane(A.vy, sx, x0 - ORD + i, nx, col);
    wvz[i + 1] = ld_plane(A.vz, sx, x0 - ORD + i, nx, col);
  }
  float rho_n = ld_plane(A.rho, sx, x0, nx, col), cs_n = ld_plane(A.cs, sx, x0, nx, col);

  for (int x = x0; x < x1; x++) {
#pragma unroll
    for (int i = 0; i < 2 * ORD - 1; i++) { wvx[i] = wvx[i + 1]; wvy[i] = wvy[i + 1]; wvz[i] = wvz[i + 1]; }
    wvx[2 * ORD - 1] = ld_plane(A.vx, sx, x + ORD, nx, col);
    wvy[2 * ORD - 1] = ld_plane(A.vy, sx, x + ORD - 1, nx, col);
    wvz[2 * ORD - 1] = ld_plane(A.vz, sx, x + ORD - 1, nx, col);
    const float rho0 = rho_n, cs0 = cs_n;
    rho_n = ld_plane(A.rho, sx, x + 1, nx, col);
    cs_n = ld_plane(A.cs, sx, x + 1, nx, col);
    const long id = x * sx + col;
    const int jx = x - off;
    const int bx = band_index(A.ax[0], jx);
    const long px = (long)bx * sx + col;
    const long py = (long)x * nby * nz + (long)by * nz + z;
    const long pz = ((long)x * ny + y) * nbz + bz;
    const bool xp = x >= last && x < nx - last - 1;
    const bool xm = x >= last + 1 && x < nx - last;

    // ---- normal stresses (wgsl3d :1356-1453): x D+, y D-, z D-
    if (xp && ym && zm) {
      float dvx = 0.f, dvy = 0.f, dvz = 0.f;
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        dvx += A.C[c] * (wvx[ORD + c] - wvx[ORD - 1 - c]) / A.dx;            // vx[x+c+1] - vx[x-c]
        dvy += A.C[c] * (A.vy[id + c * sy] - A.vy[id - (c + 1) * sy]) / A.dy;
        dvz += A.C[c] * (A.vz[id + c] - A.vz[id - c - 1]) / A.dz;
      }
      if (bx >= 0) {
        float *p = A.psi[P_DVX_DX] + px;
        const float pn = A.ax[0].bh[jx] * sload<NT>(p) + A.ax[0].ah[jx] * dvx;
        sstore<NT>(pn, p);
        dvx = dvx / A.ax[0].kh[jx] + pn;
      }
      if (by >= 0) {
        float *p = A.psi[P_DVY_DY] + py;
        const float pn = A.ax[1].b[jy] * sload<NT>(p) + A.ax[1].a[jy] * dvy;
        sstore<NT>(pn, p);
        dvy = dvy / A.ax[1].k[jy] + pn;
      }
      if (bz >= 0) {
        float *p = A.psi[P_DVZ_DZ] + pz;
        const float pn = A.ax[2].b[jz] * sload<NT>(p) + A.ax[2].a[jz] * dvz;
        sstore<NT>(pn, p);
        dvz = dvz / A.ax[2].k[jz] + pn;
      }
      const float cp = sload<NT>(A.cp + id);
      const float lambda = rho0 * (cp * cp - 2.0f * cs0 * cs0);
      const float mu = rho0 * (cs0 * cs0);
      const float lp2m = lambda + 2.0f * mu;
      const float sxx = sload<NT>(A.sxx + id), syy = sload<NT>(A.syy + id), szz = sload<NT>(A.szz + id);
      if constexpr (VISCO) {
        float Sxx = 0.f, Syy = 0.f, Szz = 0.f;
        const float div = dvx + dvy + dvz;
        for (int l = 0; l < A.L; l++) {
          float *pxx = A.r[0] + l * A.ncell + id, *pyy = A.r[1] + l * A.ncell + id, *pzz = A.r[2] + l * A.ncell + id;
          const float oxx = sload<NT>(pxx), oyy = sload<NT>(pyy), ozz = sload<NT>(pzz);
          const float nxx = ((oxx + div * A.zk.phi_p[l] - oxx * A.zk.h_p[l]) * A.zk.m_p[l]);
          const float nyy = ((oyy + (dvx - div / 3.0f) * A.zk.phi_s[l] - oyy * A.zk.h_s[l]) * A.zk.m_s[l]);
          const float nzz = ((ozz + (dvy - div / 3.0f) * A.zk.phi_s[l] - ozz * A.zk.h_s[l]) * A.zk.m_s[l]);
          Sxx += oxx + nxx; Syy += oyy + nyy; Szz += ozz + nzz;
          sstore<NT>(nxx, pxx); sstore<NT>(nyy, pyy); sstore<NT>(nzz, pzz);
        }
        const float l23 = lambda + 2.0f * mu / 3.0f;
        sstore<NT>(sxx + (lp2m * dvx + lambda * (dvy + dvz) + l23 * 0.5f * Sxx + mu * Syy) * dt, A.sxx + id);
        sstore<NT>(syy + (lambda * (dvx + dvz) + lp2m * dvy + l23 * 0.5f * Sxx + mu * Szz) * dt, A.syy + id);
        sstore<NT>(szz + (lambda * (dvx + dvy) + lp2m * dvz + lp2m * 0.5f * Sxx - (mu / 3.0f) * (Syy + Szz)) * dt, A.szz + id);
      } else {
        sstore<NT>(sxx + (lp2m * dvx + lambda * (dvy + dvz)) * dt, A.sxx + id);
        sstore<NT>(syy + (lambda * (dvx + dvz) + lp2m * dvy) * dt, A.syy + id);
        sstore<NT>(szz + (lambda * (dvx + dvy) + lp2m * dvz) * dt, A.szz + id);
      }
    }

    // ---- shear stresses (wgsl3d :1457-1629)
#pragma unroll
    for (int sh = 0; sh < 3; sh++) {
      const bool ok = (sh < 2 ? xm : xp) && yp && zp;
      if (!ok) continue;
      float d1 = 0.f, d2 = 0.f, rho4, cs4;
      float *sig, *rr;
      if (sh == 0) {            // sigma_xy: dvy/dx (D-), dvx/dy (D+)
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          d1 += A.C[c] * (wvy[ORD + c] - wvy[ORD - 1 - c]) / A.dx;         // vy[x+c] - vy[x-c-1]
          d2 += A.C[c] * (A.vx[id + (c + 1) * sy] - A.vx[id - c * sy]) / A.dy;
        }
        if (bx >= 0) {
          float *p = A.psi[P_DVY_DX] + px;
          const float pn = A.ax[0].b[jx] * sload<NT>(p) + A.ax[0].a[jx] * d1;
          sstore<NT>(pn, p);
          d1 = d1 / A.ax[0].k[jx] + pn;
        }
        if (by >= 0) {
          float *p = A.psi[P_DVX_DY] + py;
          const float pn = A.ax[1].bh[jy] * sload<NT>(p) + A.ax[1].ah[jy] * d2;
          sstore<NT>(pn, p);
          d2 = d2 / A.ax[1].kh[jy] + pn;
        }
        rho4 = 0.25f * (rho_n + rho0 + A.rho[id + sy] + A.rho[id + sx + sy]);
        cs4 = 0.25f * (cs_n + cs0 + A.cs[id + sy] + A.cs[id + sx + sy]);
        sig = A.sxy; rr = A.r[3];
      } else if (sh == 1) {     // sigma_xz: dvz/dx (D-), dvx/dz (D+)
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          d1 += A.C[c] * (wvz[ORD + c] - wvz[ORD - 1 - c]) / A.dx;         // vz[x+c] - vz[x-c-1]
          d2 += A.C[c] * (A.vx[id + c + 1] - A.vx[id - c]) / A.dz;
        }
        if (bx >= 0) {
          float *p = A.psi[P_DVZ_DX] + px;
          const float pn = A.ax[0].b[jx] * sload<NT>(p) + A.ax[0].a[jx] * d1;
          sstore<NT>(pn, p);
          d1 = d1 / A.ax[0].k[jx] + pn;
        }
        if (bz >= 0) {
          float *p = A.psi[P_DVX_DZ] + pz;
          const float pn = A.ax[2].bh[jz] * sload<NT>(p) + A.ax[2].ah[jz] * d2;
          sstore<NT>(pn, p);
          d2 = d2 / A.ax[2].kh[jz] + pn;
        }
        rho4 = 0.25f * (rho_n + rho0 + A.rho[id + 1] + A.rho[id + sx + 1]);
        cs4 = 0.25f * (cs_n + cs0 + A.cs[id + 1] + A.cs[id + sx + 1]);
        sig = A.sxz; rr = A.r[4];
      } else {                  // sigma_yz: dvz/dy (D+), dvy/dz (D+)
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          d1 += A.C[c] * (A.vz[id + (c + 1) * sy] - A.vz[id - c * sy]) / A.dy;
          d2 += A.C[c] * (A.vy[id + c + 1] - A.vy[id - c]) / A.dz;
        }
        if (by >= 0) {
          float *p = A.psi[P_DVZ_DY] + py;
          const float pn = A.ax[1].bh[jy] * sload<NT>(p) + A.ax[1].ah[jy] * d1;
          sstore<NT>(pn, p);
          d1 = d1 / A.ax[1].kh[jy] + pn;
        }
        if (bz >= 0) {
          float *p = A.psi[P_DVY_DZ] + pz;
          const float pn = A.ax[2].bh[jz] * sload<NT>(p) + A.ax[2].ah[jz] * d2;
          sstore<NT>(pn, p);
          d2 = d2 / A.ax[2].kh[jz] + pn;
        }
        rho4 = 0.25f * (A.rho[id + sy] + rho0 + A.rho[id + 1] + A.rho[id + sy + 1]);
        cs4 = 0.25f * (A.cs[id + sy] + cs0 + A.cs[id + 1] + A.cs[id + sy + 1]);
        sig = A.syz; rr = A.r[5];
      }
      const float mu = rho4 * (cs4 * cs4);
      const float so = sload<NT>(sig + id);
      if constexpr (VISCO) {
        float S = 0.f;
        const float X = d1 + d2;
        for (int l = 0; l < A.L; l++) {
          float *pr = rr + l * A.ncell + id;
          const float o = sload<NT>(pr);
          const float n = (o + X * A.zk.phi_s[l] - o * A.zk.h_s[l]) * A.zk.m_s[l];
          S += o + n;
          sstore<NT>(n, pr);
        }
        sstore<NT>(so + (mu * (d1 + d2) + mu * 0.5f * S) * dt, sig + id);
      } else {
        sstore<NT>(so + (d2 + d1) * mu * dt, sig + id);
      }
    }
  }
}

template <int ORD, int TY, bool NT>
__global__ __launch_bounds__(BZ * TY) void k_velocity3d(StepArgs A, int xchunk, int nzb, int nyb, int nxc) {
  const Tile T = tile_of_block(nzb, nyb, nxc);
  const int z = T.bz * BZ + threadIdx.x;
  const int y = T.by * TY + threadIdx.y;
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  const bool valid = z < nz && y < ny;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long sy = nz, sx = (long)ny * nz;
  const long col = (long)y * sy + z;
  const float dt = A.dt;
  const int x0 = T.bxc * xchunk, x1 = min(nx, x0 + xchunk);
  unsigned v2bits = 0;
  if (valid) {
    const int jy = y - off, jz = z - off;
    const int by = band_index(A.ax[1], jy), bz = band_index(A.ax[2], jz);
    const int nby = A.ax[1].nb, nbz = A.ax[2].nb;
    const bool yi_vx = y >= last + 1 && y < ny - last, zi_vx = z >= last + 1 && z < nz - last;
    const bool yi_vy = y >= last && y < ny - last - 1, zi_vy = z >= last + 1 && z < nz - last;
    const bool yi_vz = y >= last + 1 && y < ny - last, zi_vz = z >= last && z < nz - last - 1;
    // x windows: sxx at x-ORD..x+ORD-1 (D-), sxy/sxz at x-(ORD-1)..x+ORD (D+)
    float wxx[2 * ORD], wxy[2 * ORD], wxz[2 * ORD];
#pragma unroll
    for (int i = 0; i < 2 * ORD - 1; i++) {
      wxx[i + 1] = ld_plane(A.sxx, sx, x0 - ORD + i, nx, col);
      wxy[i + 1] = ld_plane(A.sxy, sx, x0 - (ORD - 1) + i, nx, col);
      wxz[i + 1] = ld_plane(A.sxz, sx, x0 - (ORD - 1) + i, nx, col);
    }
    float rho_n = ld_plane(A.rho, sx, x0, nx, col);
    for (int x = x0; x < x1; x++) {
#pragma unroll
      for (int i = 0; i < 2 * ORD - 1; i++) { wxx[i] = wxx[i + 1]; wxy[i] = wxy[i + 1]; wxz[i] = wxz[i + 1]; }
      wxx[2 * ORD - 1] = ld_plane(A.sxx, sx, x + ORD - 1, nx, col);
      wxy[2 * ORD - 1] = ld_plane(A.sxy, sx, x + ORD, nx, col);
      wxz[2 * ORD - 1] = ld_plane(A.sxz, sx, x + ORD, nx, col);
      const float rho0 = rho_n;
      rho_n = ld_plane(A.rho, sx, x + 1, nx, col);
      const long id = x * sx + col;
      const int jx = x - off;
      const int bx = band_index(A.ax[0], jx);
      const long px = (long)bx * sx + col;
      const long py = (long)x * nby * nz + (long)by * nz + z;
      const long pz = ((long)x * ny + y) * nbz + bz;
      float vxn = sload<NT>(A.vx + id), vyn = sload<NT>(A.vy + id), vzn = sload<NT>(A.vz + id);
      // vx (wgsl3d :1647-1681)
      if (x >= last + 1 && x < nx - last && yi_vx && zi_vx) {
        float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          a += A.C[c] * (wxx[ORD + c] - wxx[ORD - 1 - c]) / A.dx;          // sxx[x+c] - sxx[x-c-1]
          b += A.C[c] * (A.sxy[id + c * sy] - A.sxy[id - (c + 1) * sy]) / A.dy;
          c_ += A.C[c] * (A.sxz[id + c] - A.sxz[id - c - 1]) / A.dz;
        }
        if (bx >= 0) {
          float *p = A.psi[P_DSXX_DX] + px;
          const float pn = A.ax[0].b[jx] * sload<NT>(p) + A.ax[0].a[jx] * a;
          sstore<NT>(pn, p);
          a = a / A.ax[0].k[jx] + pn;
        }
        if (by >= 0) {
          float *p = A.psi[P_DSXY_DY] + py;
          const float pn = A.ax[1].b[jy] * sload<NT>(p) + A.ax[1].a[jy] * b;
          sstore<NT>(pn, p);
          b = b / A.ax[1].k[jy] + pn;
        }
        if (bz >= 0) {
          float *p = A.psi[P_DSXZ_DZ] + pz;
          const float pn = A.ax[2].b[jz] * sload<NT>(p) + A.ax[2].a[jz] * c_;
          sstore<NT>(pn, p);
          c_ = c_ / A.ax[2].k[jz] + pn;
        }
        const float rho = 0.5f * (rho_n + rho0);
        if (rho > 0.f) vxn = (a + b + c_) * dt / rho + vxn;
      }
      // vy (wgsl3d :1683-1717)
      const bool xi_vyz = x >= last && x < nx - last - 1;
      if (xi_vyz && yi_vy && zi_vy) {
        float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          a += A.C[c] * (wxy[ORD + c] - wxy[ORD - 1 - c]) / A.dx;          // sxy[x+c+1] - sxy[x-c]
          b += A.C[c] * (A.syy[id + (c + 1) * sy] - A.syy[id - c * sy]) / A.dy;
          c_ += A.C[c] * (A.syz[id + c] - A.syz[id - c - 1]) / A.dz;
        }
        if (bx >= 0) {
          float *p = A.psi[P_DSXY_DX] + px;
          const float pn = A.ax[0].bh[jx] * sload<NT>(p) + A.ax[0].ah[jx] * a;
          sstore<NT>(pn, p);
          a = a / A.ax[0].kh[jx] + pn;
        }
        if (by >= 0) {
          float *p = A.psi[P_DSYY_DY] + py;
          const float pn = A.ax[1].bh[jy] * sload<NT>(p) + A.ax[1].ah[jy] * b;
          sstore<NT>(pn, p);
          b = b / A.ax[1].kh[jy] + pn;
        }
        if (bz >= 0) {
          float *p = A.psi[P_DSYZ_DZ] + pz;
          const float pn = A.ax[2].b[jz] * sload<NT>(p) + A.ax[2].a[jz] * c_;
          sstore<NT>(pn, p);
          c_ = c_ / A.ax[2].k[jz] + pn;
        }
        const float rho = 0.5f * (A.rho[id + sy] + rho0);
        if (rho > 0.f) vyn = (a + b + c_) * dt / rho + vyn;
      }
      // vz, with the vy density (wgsl3d :1719-1758)
      if (xi_vyz && yi_vz && zi_vz) {
        float a = 0.f, b = 0.f, c_ = 0.f;
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          a += A.C[c] * (wxz[ORD + c] - wxz[ORD - 1 - c]) / A.dx;          // sxz[x+c+1] - sxz[x-c]
          b += A.C[c] * (A.syz[id + c * sy] - A.syz[id - (c + 1) * sy]) / A.dy;
          c_ += A.C[c] * (A.szz[id + c + 1] - A.szz[id - c]) / A.dz;
        }
        if (bx >= 0) {
          float *p = A.psi[P_DSXZ_DX] + px;
          const float pn = A.ax[0].bh[jx] * sload<NT>(p) + A.ax[0].ah[jx] * a;
          sstore<NT>(pn, p);
          a = a / A.ax[0].kh[jx] + pn;
        }
        if (by >= 0) {
          float *p = A.psi[P_DSYZ_DY] + py;
          const float pn = A.ax[1].b[jy] * sload<NT>(p) + A.ax[1].a[jy] * b;
          sstore<NT>(pn, p);
          b = b / A.ax[1].k[jy] + pn;
        }
        if (bz >= 0) {
          float *p = A.psi[P_DSZZ_DZ] + pz;
          const float pn = A.ax[2].bh[jz] * sload<NT>(p) + A.ax[2].ah[jz] * c_;
          sstore<NT>(pn, p);
          c_ = c_ / A.ax[2].kh[jz] + pn;
        }
        const float rho = 0.5f * (A.rho[id + sy] + rho0);
        if (rho > 0.f) vzn = (a + b + c_) * dt / rho + vzn;
      } else {
        vxn = 0.f; vyn = 0.f; vzn = 0.f;
      }
      sstore<NT>(vxn, A.vx + id);
      sstore<NT>(vyn, A.vy + id);
      sstore<NT>(vzn, A.vz + id);
      const unsigned u = __float_as_uint(vxn * vxn + vyn * vyn + vzn * vzn);
      v2bits = v2bits > u ? v2bits : u;
    }
  }
  block_max_u32(v2bits, A.v2max);
}

// Launch shape: tiles of 64 (z) x TY (y) columns; x split into chunks of
// `xchunk` planes (1 = one plane per workgroup, i.e. one cell per thread).
struct Launch {
  int xchunk, nzb, nyb, nxc;
};

Launch launch_shape(const StepArgs &A, int ty, int xchunk) {
  Launch L;
  L.nzb = (A.nz + BZ - 1) / BZ;
  L.nyb = (A.ny + ty - 1) / ty;
  L.xchunk = std::max(1, std::min(xchunk, A.nx));
  L.nxc = (A.nx + L.xchunk - 1) / L.xchunk;
  return L;
}

template <int ORD, int TY, bool NT>
hipError_t launch_one(bool stress, int visco, const StepArgs &A, int xchunk, hipStream_t st) {
  const Launch L = launch_shape(A, TY, xchunk);
  const dim3 grid((unsigned)(L.nzb * L.nyb * L.nxc)), block(BZ, TY);
  if (stress) {
    if (visco) hipLaunchKernelGGL((k_stress3d<true, ORD, TY, NT>), grid, block, 0, st, A, L.xchunk, L.nzb, L.nyb, L.nxc);
    else hipLaunchKernelGGL((k_stress3d<false, ORD, TY, NT>), grid, block, 0, st, A, L.xchunk, L.nzb, L.nyb, L.nxc);
  } else {
    hipLaunchKernelGGL((k_velocity3d<ORD, TY, NT>), grid, block, 0, st, A, L.xchunk, L.nzb, L.nyb, L.nxc);
  }
  return hipGetLastError();
}

template <int ORD>
hipError_t launch_tuned(bool stress, int visco, const StepArgs &A, const KTune &t, hipStream_t st) {
  // shapes compiled for every order: (TY, NT) in {1,2,4,8} x {0,1}
#define VSE_TY(TYV)                                                                          \
  case TYV:                                                                                  \
    return t.nt ? launch_one<ORD, TYV, true>(stress, visco, A, t.xchunk, st)                 \
                : launch_one<ORD, TYV, false>(stress, visco, A, t.xchunk, st);
  switch (t.ty) {
    VSE_TY(1) VSE_TY(2) VSE_TY(4) VSE_TY(8)
    default: return hipErrorInvalidValue;
  }
#undef VSE_TY
}

hipError_t dispatch(bool stress, int visco, int ord, const StepArgs &A, const KTune &t, hipStream_t st) {
  switch (ord) {
    case 2: return launch_tuned<2>(stress, visco, A, t, st);
    case 3: return launch_tuned<3>(stress, visco, A, t, st);
    case 4: return launch_tuned<4>(stress, visco, A, t, st);
    case 5: return launch_tuned<5>(stress, visco, A, t, st);
    case 6: return launch_tuned<6>(stress, visco, A, t, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t vse_launch_stress3d(int visco, int ord, const StepArgs &A, const KTune &t, hipStream_t st) {
  return dispatch(true, visco, ord, A, t, st);
}

hipError_t vse_launch_velocity3d(int visco, int ord, const StepArgs &A, const KTune &t, hipStream_t st) {
  return dispatch(false, visco, ord, A, t, st);
}
