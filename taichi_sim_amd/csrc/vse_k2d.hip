// vse_k2d.hip -- 2D stress and velocity kernels (elastic / Zener, CPML).
//
// Update semantics follow shader_webgpu_viscoelastic.wgsl:829-1072 (WGSL
// variant) and sim_taichi_elastic.py:138-309 (Taichi variant); the CPU oracle
// restates the same order.  One thread per cell, blocks of 64 (fast y axis)
// x 4 (x axis): every wavefront reads 256 contiguous bytes per field, and the
// +-ORD stencil neighbours come from L1/L2, so each field streams from HBM
// about once per pass.
#include <algorithm>

#include "vse_device.h"

namespace {

constexpr int BY = 64, BX = 4;

// One staggered-difference term along axis AX: the variant's own form, or
// with FAST (WGSL variant only) (a - b) * f32(C/d), a rounding-level
// difference from the reference order C (a - b) / d.
template <int VAR, bool FAST, int AX>
__device__ __forceinline__ float dterm2(const StepArgs &A, int c, float a, float b) {
  if constexpr (FAST) return (a - b) * A.Cd[AX][c];
  else return fd_term<VAR>(A.C[c], a, b, AX == 0 ? A.dx : A.dy, AX == 0 ? A.one_dx : A.one_dy);
}


// Velocity operands of a stress cell read from the fields: vx(dx, dy) is vx
// at (x + dx, y + dy).
struct VGlobal {
  const VSE_G float *vx_, *vy_;
  long id;
  int ny;
  __device__ __forceinline__ float vx(int dx, int dy) const { return vx_[id + (long)dx * ny + dy]; }
  __device__ __forceinline__ float vy(int dx, int dy) const { return vy_[id + (long)dx * ny + dy]; }
};

// The stress update of cell (x, y) < (nx, ny): V supplies the velocity
// operands, sigma is read from A and written to osxx / osyy / osxy (A's own
// fields in place, or another buffer set), memory variables and the
// stress-pass psi are updated in place.  LC > 0: the mechanism count is the
// compile-time LC and every memory-variable load of a point is issued before
// its arithmetic (same operations, same order: bit-identical to the
// runtime-L loop, which waits on each mechanism's loads in turn).
template <int VAR, bool VISCO, int ORD, int LC, bool FAST, class VS>
__device__ __forceinline__ void stress2d_core(const StepArgs &A, int x, int y, const VS &V, VSE_G float *osxx,
                                              VSE_G float *osyy, VSE_G float *osxy) {
  const int nx = A.nx, ny = A.ny;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long id = (long)x * ny + y;
  const float dt = A.dt;

  // ---- normal stresses: x in the D+ range, y in the D- range (wgsl2d :842-846)
  if (x >= last && x < nx - last - 1 && y >= last + 1 && y < ny - last) {
    float dvx = 0.f, dvy = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      dvx += dterm2<VAR, FAST, 0>(A, c, V.vx(c + 1, 0), V.vx(-c, 0));
      dvy += dterm2<VAR, FAST, 1>(A, c, V.vy(0, c), V.vy(0, -c - 1));
    }
    const int jx = x - off, jy = y - off;
    const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy);
    dvx = cpml_stretch_k<FAST>(A.ax[0].ah, A.ax[0].bh, A.ax[0].kh, A.ax[0].khi, A.psi[P_DVX_DX], (long)bx * ny + y, jx,
                               bx >= 0, dvx);
    dvy = cpml_stretch_k<FAST>(A.ax[1].a, A.ax[1].b, A.ax[1].k, A.ax[1].ki, A.psi[P_DVY_DY], (long)x * A.ax[1].nb + by,
                               jy, by >= 0, dvy);

    const float rho_h = 0.5f * (A.rho[id + ny] + A.rho[id]);
    const float cp_h = 0.5f * (A.cp[id + ny] + A.cp[id]);
    const float csa = A.cs[id + ny], csb = A.cs[id];
    const float cs_l = 0.5f * (csa + csb);
    const float cs_m = (fminf(csa, csb) == 0.f) ? 0.f : cs_l;
    const float lambda = rho_h * (cp_h * cp_h - 2.0f * cs_l * cs_l);
    const float mu = rho_h * (cs_m * cs_m);
    if constexpr (VAR == VSE_VAR_TAICHI) {
      const float dl = lambda * (dvx + dvy);
      osxx[id] = A.sxx[id] + (dl + 2.0f * mu * dvx) * dt;
      osyy[id] = A.syy[id] + (dl + 2.0f * mu * dvy) * dt;
    } else {
      const float lp2m = lambda + 2.0f * mu;
      if constexpr (VISCO && LC > 0) {
        const float lpm = lambda + mu;
        float oxx[LC], oyy[LC];
#pragma unroll
        for (int l = 0; l < LC; l++) {
          oxx[l] = A.r[0][l * A.rs + id];
          oyy[l] = A.r[1][l * A.rs + id];
        }
        const float sxx0 = A.sxx[id], syy0 = A.syy[id];
        float sum_xx = 0.f, sum_yy = 0.f;
#pragma unroll
        for (int l = 0; l < LC; l++) {
          const float nxx = ((oxx[l] + (dvx + dvy) * A.zk.phi_p[l] - oxx[l] * A.zk.h_p[l]) * A.zk.m_p[l]);
          const float nyy = ((oyy[l] + 0.5f * (dvx - dvy) * A.zk.phi_s[l] - oyy[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
          sum_xx += oxx[l] + nxx;
          sum_yy += oyy[l] + nyy;
          A.r[0][l * A.rs + id] = nxx;
          A.r[1][l * A.rs + id] = nyy;
        }
        osxx[id] = sxx0 + (lp2m * dvx + lambda * dvy + 0.5f * lpm * sum_xx + mu * sum_yy) * dt;
        osyy[id] = syy0 + (lambda * dvx + lp2m * dvy + 0.5f * lpm * sum_xx - mu * sum_yy) * dt;
      } else if constexpr (VISCO) {
        const float lpm = lambda + mu;
        float sum_xx = 0.f, sum_yy = 0.f;
        for (int l = 0; l < A.L; l++) {
          float *pxx = A.r[0] + l * A.rs + id, *pyy = A.r[1] + l * A.rs + id;
          const float oxx = *pxx, oyy = *pyy;
          const float nxx = ((oxx + (dvx + dvy) * A.zk.phi_p[l] - oxx * A.zk.h_p[l]) * A.zk.m_p[l]);
          const float nyy = ((oyy + 0.5f * (dvx - dvy) * A.zk.phi_s[l] - oyy * A.zk.h_s[l]) * A.zk.m_s[l]);
          sum_xx += oxx + nxx;
          sum_yy += oyy + nyy;
          *pxx = nxx;
          *pyy = nyy;
        }
        osxx[id] = A.sxx[id] + (lp2m * dvx + lambda * dvy + 0.5f * lpm * sum_xx + mu * sum_yy) * dt;
        osyy[id] = A.syy[id] + (lambda * dvx + lp2m * dvy + 0.5f * lpm * sum_xx - mu * sum_yy) * dt;
      } else {
        osxx[id] = A.sxx[id] + (lp2m * dvx + lambda * dvy) * dt;
        osyy[id] = A.syy[id] + (lambda * dvx + lp2m * dvy) * dt;
      }
    }
  }

  // ---- shear stress: x in the D- range, y in the D+ range (wgsl2d :927-931)
  if (x >= last + 1 && x < nx - last && y >= last && y < ny - last - 1) {
    float dvydx = 0.f, dvxdy = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      dvydx += dterm2<VAR, FAST, 0>(A, c, V.vy(c, 0), V.vy(-c - 1, 0));
      dvxdy += dterm2<VAR, FAST, 1>(A, c, V.vx(0, c + 1), V.vx(0, -c));
    }
    const int jx = x - off, jy = y - off;
    const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy);
    dvydx = cpml_stretch_k<FAST>(A.ax[0].a, A.ax[0].b, A.ax[0].k, A.ax[0].ki, A.psi[P_DVY_DX], (long)bx * ny + y, jx,
                                 bx >= 0, dvydx);
    dvxdy = cpml_stretch_k<FAST>(A.ax[1].ah, A.ax[1].bh, A.ax[1].kh, A.ax[1].khi, A.psi[P_DVX_DY],
                                 (long)x * A.ax[1].nb + by, jy, by >= 0, dvxdy);

    const float rho_h = 0.5f * (A.rho[id + 1] + A.rho[id]);
    const float csa = A.cs[id + 1], csb = A.cs[id];
    const float cs_h = (fminf(csa, csb) == 0.f) ? 0.f : 0.5f * (csa + csb);
    const float mu = rho_h * (cs_h * cs_h);
    if constexpr (VISCO && VAR == VSE_VAR_WGSL && LC > 0) {
      float o[LC];
#pragma unroll
      for (int l = 0; l < LC; l++) o[l] = A.r[3][l * A.rs + id];
      const float sxy0 = A.sxy[id];
      float sum_xy = 0.f;
#pragma unroll
      for (int l = 0; l < LC; l++) {
        const float n = ((o[l] + (dvydx + dvxdy) * A.zk.phi_s[l] - o[l] * A.zk.h_s[l]) * A.zk.m_s[l]);
        sum_xy += o[l] + n;
        A.r[3][l * A.rs + id] = n;
      }
      osxy[id] = sxy0 + (mu * (dvydx + dvxdy) + 0.5f * mu * sum_xy) * dt;
    } else if constexpr (VISCO && VAR == VSE_VAR_WGSL) {
      float sum_xy = 0.f;
      for (int l = 0; l < A.L; l++) {
        float *pxy = A.r[3] + l * A.rs + id;
        const float o = *pxy;
        const float n = ((o + (dvydx + dvxdy) * A.zk.phi_s[l] - o * A.zk.h_s[l]) * A.zk.m_s[l]);
        sum_xy += o + n;
        *pxy = n;
      }
      osxy[id] = A.sxy[id] + (mu * (dvydx + dvxdy) + 0.5f * mu * sum_xy) * dt;
    } else {
      osxy[id] = A.sxy[id] + (dvxdy + dvydx) * mu * dt;
    }
  }
}

template <int VAR, bool VISCO, int ORD, int LC = 0, bool FAST = false>
__device__ __forceinline__ void stress2d_cell(const StepArgs &A, int x, int y) {
  if (x >= A.nx || y >= A.ny) return;
  const VGlobal V{A.vx, A.vy, (long)x * A.ny + y, A.ny};
  stress2d_core<VAR, VISCO, ORD, LC, FAST>(A, x, y, V, A.sxx, A.syy, A.sxy);
}

// The stencil operands of one velocity cell: vx, vy at the cell, sigma_xx
// along x and sigma_xy along y for vx (D-), sigma_xy along x and sigma_yy
// along y for vy (D+), and the four densities of the vy point.
template <int ORD>
struct VelOps2 {
  float vxn, vyn;
  float xx_p[ORD], xx_m[ORD], xy_p[ORD], xy_m[ORD];
  float yx_p[ORD], yx_m[ORD], yy_p[ORD], yy_m[ORD];
  float r00, r10, r11, r01;
};

// The WGSL-variant source flag of row x (1 on rows holding a source cell),
// loaded unconditionally before a cell's stencil operands so that the
// source-map load waits only for it; 0 for the Taichi variant (its source is
// part of the update).
template <int VAR>
__device__ __forceinline__ int source_row(const StepArgs &A, int x) {
  if constexpr (VAR == VSE_VAR_WGSL) return A.src_rowflag[x];
  else return 0;
}

// The velocity update of one cell (x, y) < (nx, ny) from its operands: the
// CPML band loads, the two updates, the Dirichlet rules, the WGSL-variant
// source and the stores.  Returns |v|^2 as float bits, the larger of its
// values before and after the source.  s: the step being advanced; srow:
// source_row of the cell's row.  The WGSL source comes after the update,
// anywhere on the grid (wgsl2d :1059-1065): vy += src[s, col] * dt / rho at
// each source cell, rho the vy-grid density with out-of-grid neighbours 0.
// The update itself: psi is read from A and, when `store`, the new psi is
// written to po (the slots P_DSXX_DX, P_DSXY_DY, P_DSXY_DX, P_DSYY_DY of A in
// place, or of another buffer set); the new velocity is returned in vxo / vyo.
template <int VAR, bool VISCO, int ORD, bool FAST = false>
__device__ __forceinline__ unsigned velocity2d_update(const StepArgs &A, int x, int y, int s, int srow,
                                                      VelOps2<ORD> &o, float &vxo, float &vyo,
                                                      VSE_G float *const *po, bool store) {
  const int nx = A.nx, ny = A.ny;
  constexpr int last = ORD - 1, off = ORD - 1;
  const long id = (long)x * ny + y;
  const float dt = A.dt;
  const bool in_vx = x >= last + 1 && x < nx - last && y >= last + 1 && y < ny - last;
  const bool in_vy = x >= last && x < nx - last - 1 && y >= last && y < ny - last - 1;
  const int jx = x - off, jy = y - off;
  const int bx = band_index(A.ax[0], jx), by = band_index(A.ax[1], jy);
  const long pxi = (long)bx * ny + y, pyi = (long)x * A.ax[1].nb + by;
  float vxn = o.vxn, vyn = o.vyn;
  float q1 = 0.f, q2 = 0.f, q3 = 0.f, q4 = 0.f;
  int is = -1;
  if (in_vx) {
    if (bx >= 0) q1 = A.psi[P_DSXX_DX][pxi];
    if (by >= 0) q2 = A.psi[P_DSXY_DY][pyi];
  }
  if (in_vy) {
    if (bx >= 0) q3 = A.psi[P_DSXY_DX][pxi];
    if (by >= 0) q4 = A.psi[P_DSYY_DY][pyi];
    if constexpr (VAR == VSE_VAR_TAICHI) is = A.src_map[id];
  }
  if constexpr (VAR == VSE_VAR_WGSL)
    if (s < A.src_rows && srow) is = A.src_map[id];
  // the band's profile values, with the other loads (a wavefront holding band
  // cells would otherwise wait on each of them inside its arithmetic)
  float xa = 0.f, xb = 0.f, xk = 1.f, xah = 0.f, xbh = 0.f, xkh = 1.f;
  float ya = 0.f, yb = 0.f, yk = 1.f, yah = 0.f, ybh = 0.f, ykh = 1.f;
  if (bx >= 0) {
    const AxisProf &P = A.ax[0];
    xa = P.a[jx]; xb = P.b[jx]; xk = FAST ? P.ki[jx] : P.k[jx];
    xah = P.ah[jx]; xbh = P.bh[jx]; xkh = FAST ? P.khi[jx] : P.kh[jx];
  }
  if (by >= 0) {
    const AxisProf &P = A.ax[1];
    ya = P.a[jy]; yb = P.b[jy]; yk = FAST ? P.ki[jy] : P.k[jy];
    yah = P.ah[jy]; ybh = P.bh[jy]; ykh = FAST ? P.khi[jy] : P.kh[jy];
  }
  float sval = 0.f;
  if constexpr (VAR == VSE_VAR_WGSL)
    if (is >= 0) sval = A.src[(long)s * A.src_cols + is];
  // ---- vx (wgsl2d :998-1025)
  if (in_vx) {
    float d1 = 0.f, d2 = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      d1 += dterm2<VAR, FAST, 0>(A, c, o.xx_p[c], o.xx_m[c]);
      d2 += dterm2<VAR, FAST, 1>(A, c, o.xy_p[c], o.xy_m[c]);
    }
    if (bx >= 0) {
      const float pn = xb * q1 + xa * d1;
      if (store) po[0][pxi] = pn;
      d1 = (FAST ? d1 * xk : d1 / xk) + pn;
    }
    if (by >= 0) {
      const float pn = yb * q2 + ya * d2;
      if (store) po[1][pyi] = pn;
      d2 = (FAST ? d2 * yk : d2 / yk) + pn;
    }
    if (o.r00 > 0.f) {
      if constexpr (VAR == VSE_VAR_TAICHI) vxn += (d1 + d2) * dt / o.r00;
      else vxn = (d1 + d2) * dt / o.r00 + vxn;
    }
  } else if constexpr (VAR == VSE_VAR_TAICHI) {
    vxn = 0.f;     // per-component Dirichlet (sim_taichi_elastic.py:266-268)
  }
  // ---- vy (wgsl2d :1027-1057; sim_taichi_elastic.py:270-304)
  if (in_vy) {
    float d1 = 0.f, d2 = 0.f;
#pragma unroll
    for (int c = 0; c < ORD; c++) {
      d1 += dterm2<VAR, FAST, 0>(A, c, o.yx_p[c], o.yx_m[c]);
      d2 += dterm2<VAR, FAST, 1>(A, c, o.yy_p[c], o.yy_m[c]);
    }
    if (bx >= 0) {
      const float pn = xbh * q3 + xah * d1;
      if (store) po[2][pxi] = pn;
      d1 = (FAST ? d1 * xkh : d1 / xkh) + pn;
    }
    if (by >= 0) {
      const float pn = ybh * q4 + yah * d2;
      if (store) po[3][pyi] = pn;
      d2 = (FAST ? d2 * ykh : d2 / ykh) + pn;
    }
    if constexpr (VAR == VSE_VAR_TAICHI) {
      const float rho = 0.25f * (o.r01 + o.r00 + o.r11 + o.r10);
      if (rho > 0.f) {
        float vs = 0.f;
        if (is != -1 && s < A.src_rows) vs = A.src[(long)s * A.src_cols + is] * dt / rho;
        vyn += vs + (d1 + d2) * dt / rho;
      }
    } else {
      const float rho = 0.25f * (o.r00 + o.r10 + o.r11 + o.r01);
      if (rho > 0.f) vyn = (d1 + d2) * dt / rho + vyn;
    }
  } else {
    if constexpr (VAR == VSE_VAR_WGSL) vxn = 0.f;   // the vy else zeroes both (wgsl2d :1054-1057)
    vyn = 0.f;
  }
  unsigned u = __float_as_uint(vxn * vxn + vyn * vyn);
  if constexpr (VAR == VSE_VAR_WGSL) {
    if (is >= 0) {
      const bool xe = x + 1 < nx, ye = y + 1 < ny;
      const float rho = 0.25f * (o.r00 + (xe ? o.r10 : 0.f) + ((xe && ye) ? o.r11 : 0.f) + (ye ? o.r01 : 0.f));
      if (rho > 0.f) {
        vyn = vyn + sval * A.dt / rho;
        const unsigned w = __float_as_uint(vxn * vxn + vyn * vyn);
        u = u > w ? u : w;
      }
    }
  }
  vxo = vxn;
  vyo = vyn;
  return u;
}

// The update of cell (x, y) in place: psi and v written back to A.
template <int VAR, bool VISCO, int ORD, bool FAST = false>
__device__ __forceinline__ unsigned velocity2d_apply(const StepArgs &A, int x, int y, int s, int srow,
                                                     VelOps2<ORD> &o) {
  VSE_G float *const po[4] = {A.psi[P_DSXX_DX], A.psi[P_DSXY_DY], A.psi[P_DSXY_DX], A.psi[P_DSYY_DY]};
  float vxn, vyn;
  const unsigned u = velocity2d_update<VAR, VISCO, ORD, FAST>(A, x, y, s, srow, o, vxn, vyn, po, true);
  const long id = (long)x * A.ny + y;
  A.vx[id] = vxn;
  A.vy[id] = vyn;
  return u;
}

// All loads of a cell are issued before its arithmetic: raw stencil operands
// go into registers and the sums run afterwards, in the same order, so the
// results are the same; a division right after a load would otherwise make
// the wavefront wait on it before issuing the loads that follow.  Every
// stencil operand is loaded unconditionally through the field buffers (reads
// outside the grid give 0; operands outside an update's interior are loaded
// but not used); the CPML band and source map only where they apply.  The
// source row flag is loaded first, so that the source-map load waits for it
// alone and not for the stencil operands issued after it.
template <int VAR, bool VISCO, int ORD, bool FAST = false>
__device__ __forceinline__ unsigned velocity2d_cell(const StepArgs &A, int x, int y, int s) {
  const int nx = A.nx, ny = A.ny;
  if (x >= nx || y >= ny) return 0u;
  const FieldBuf Bvx(A.vx, A.ncell), Bvy(A.vy, A.ncell), Bsxx(A.sxx, A.ncell), Bsxy(A.sxy, A.ncell),
      Bsyy(A.syy, A.ncell), Brho(A.rho, A.ncell);
  const int i = (int)((long)x * ny + y);
  const int srow = source_row<VAR>(A, x);
  VelOps2<ORD> o;
  o.vxn = Bvx[i];
  o.vyn = Bvy[i];
#pragma unroll
  for (int c = 0; c < ORD; c++) {
    o.xx_p[c] = Bsxx[i + c * ny];
    o.xx_m[c] = Bsxx[i - (c + 1) * ny];
    o.xy_p[c] = Bsxy[i + c];
    o.xy_m[c] = Bsxy[i - c - 1];
    o.yx_p[c] = Bsxy[i + (c + 1) * ny];
    o.yx_m[c] = Bsxy[i - c * ny];
    o.yy_p[c] = Bsyy[i + c + 1];
    o.yy_m[c] = Bsyy[i - c];
  }
  // in the vy range x+1 and y+1 are inside the grid
  o.r00 = Brho[i]; o.r10 = Brho[i + ny]; o.r11 = Brho[i + ny + 1]; o.r01 = Brho[i + 1];
  return velocity2d_apply<VAR, VISCO, ORD, FAST>(A, x, y, s, srow, o);
}

// Tile of a 1D launch over nty (y) x ntx (x) tiles of BY x BX cells, y-tile
// fastest.  Blocks land on the 8 XCDs round-robin, each with its own L2; with
// remap = 1 XCD k gets the contiguous tile range [k/8, (k+1)/8) of the grid, so
// the +-ORD halo rows and columns a tile reads mostly belong to tiles of the
// same XCD (remap needs a tile count divisible by 8; otherwise it is ignored).
__device__ __forceinline__ int xcd_tile(int b, int total, int remap) {
  return (remap && (total & 7) == 0) ? (b & 7) * (total >> 3) + (b >> 3) : b;
}

struct Tile2 {
  int ty, tx;
};
__device__ __forceinline__ Tile2 tile2d(int nty, int remap) {
  const int t = xcd_tile(blockIdx.x, gridDim.x, remap);
  return Tile2{t % nty, t / nty};
}

// A tile is BY (y) x BX*rows (x) cells: the workgroup walks its `rows`
// row-groups of BX rows one after another (the autotuner picks rows).
template <int VAR, bool VISCO, int ORD, int LC = 0, bool FAST = false>
__global__ __launch_bounds__(BX * BY) void k_stress2d(StepArgs A, int nty, int remap, int rows) {
  const Tile2 t = tile2d(nty, remap);
  const int x0 = t.tx * BX * rows + threadIdx.y, y = t.ty * BY + threadIdx.x;
  for (int k = 0; k < rows; k++) stress2d_cell<VAR, VISCO, ORD, LC, FAST>(A, x0 + k * BX, y);
}

// The step's sparse work, done inside the velocity pass (vse_sparse.hip does
// it for 3D, in the same way and order):
// * the WGSL-variant source, part of each cell's update (velocity2d_apply);
// * receivers of step s, by one workgroup: traces[s, r] += the channel
//   (sigma_yy by default) over the receiver's points in table order when
//   s + gate >= delay (wgsl2d :1117-1134; sim_taichi_elastic.py:187-190).  A
//   stress channel is final after the stress pass and this pass does not
//   write it; a velocity channel (rec_after) is gathered by k_record2d after
//   this pass instead.
// The step is s = *A.step + i: a graph of k steps passes i = 0..k-1 and ends
// with an advance of the counter by k, so it moves once per graph launch and
// no workgroup waits for another (wgsl2d :1136-1141 advances it per step).
__device__ __forceinline__ void record_receivers(const StepArgs &A, int s) {
  if (s >= A.trace_rows) return;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  for (int r = tid; r < A.n_rec; r += blockDim.x * blockDim.y) {
    if (s + A.rec_gate >= A.rec_delay[r]) {
      VSE_G float *tr = A.traces + (long)s * A.n_rec + r;
      float acc = *tr;
      for (int p = A.rec_beg[r]; p < A.rec_beg[r + 1]; p++) acc = acc + A.rec_field[A.rec_cell[p]];
      *tr = acc;
    }
  }
}

// Velocity pass of one simulation, with its step's sparse work.
template <int VAR, bool VISCO, int ORD, bool FAST = false>
__global__ __launch_bounds__(BX * BY) void k_velocity2d(StepArgs A, int nty, int remap, int rows, int i) {
  const Tile2 t = tile2d(nty, remap);
  const int x0 = t.tx * BX * rows + threadIdx.y, y = t.ty * BY + threadIdx.x;
  const int s = *A.step + i;
  unsigned v2 = 0;
  for (int k = 0; k < rows; k++) {
    const unsigned u = velocity2d_cell<VAR, VISCO, ORD, FAST>(A, x0 + k * BX, y, s);
    v2 = v2 > u ? v2 : u;
  }
  if (blockIdx.x == 0 && !A.rec_after) record_receivers(A, s);
  wave_max_to_slot(v2, A.v2slots + blockIdx.x);
}

// The velocity pass with each wavefront walking `rows` consecutive x-rows of
// its 64-cell y column (a workgroup: BX wavefronts, BX*rows rows).  The
// x-stencil operands (sigma_xx for vx, sigma_xy for vy) and the x+1 densities
// move through register windows, so a row step loads one new row of each
// instead of 2*ORD; the y-stencil operands are loaded as in velocity2d_cell.
// Same operands (a window row outside the grid reads 0 through the field
// buffer, as the per-cell loads do), same arithmetic (velocity2d_apply):
// bit-identical to k_velocity2d.
template <int VAR, bool VISCO, int ORD, bool FAST = false>
__global__ __launch_bounds__(BX * BY) void k_velocity2d_march(StepArgs A, int nty, int remap, int rows, int i) {
  const Tile2 t = tile2d(nty, remap);
  const int nx = A.nx, ny = A.ny;
  const int y = t.ty * BY + threadIdx.x;
  const int x0 = (t.tx * BX + threadIdx.y) * rows;
  const int s = *A.step + i;
  unsigned v2 = 0;
  if (y < ny && x0 < nx) {
    const int x1 = min(nx, x0 + rows);
    const FieldBuf Bvx(A.vx, A.ncell), Bvy(A.vy, A.ncell), Bsxx(A.sxx, A.ncell), Bsxy(A.sxy, A.ncell),
        Bsyy(A.syy, A.ncell), Brho(A.rho, A.ncell);
    // windows over x: wxx[j] = sigma_xx(x - ORD + j), wxy[j] = sigma_xy(x - ORD + 1 + j)
    float wxx[2 * ORD], wxy[2 * ORD];
    const int i0 = (int)((long)x0 * ny + y);
#pragma unroll
    for (int j = 0; j < 2 * ORD - 1; j++) {
      wxx[j + 1] = Bsxx[i0 + (j - ORD) * ny];
      wxy[j + 1] = Bsxy[i0 + (j - ORD + 1) * ny];
    }
    float rn0 = Brho[i0], rn1 = Brho[i0 + 1];          // rho(x, y), rho(x, y+1) of the next row
#pragma unroll 1
    for (int x = x0; x < x1; x++) {
      const int ic = (int)((long)x * ny + y);
      const int srow = source_row<VAR>(A, x);
#pragma unroll
      for (int j = 0; j < 2 * ORD - 1; j++) { wxx[j] = wxx[j + 1]; wxy[j] = wxy[j + 1]; }
      wxx[2 * ORD - 1] = Bsxx[ic + (ORD - 1) * ny];
      wxy[2 * ORD - 1] = Bsxy[ic + ORD * ny];
      VelOps2<ORD> o;
      o.vxn = Bvx[ic];
      o.vyn = Bvy[ic];
#pragma unroll
      for (int c = 0; c < ORD; c++) {
        o.xx_p[c] = wxx[ORD + c];
        o.xx_m[c] = wxx[ORD - 1 - c];
        o.xy_p[c] = c == 0 ? wxy[ORD - 1] : Bsxy[ic + c];
        o.xy_m[c] = Bsxy[ic - c - 1];
        o.yx_p[c] = wxy[ORD + c];
        o.yx_m[c] = wxy[ORD - 1 - c];
        o.yy_p[c] = Bsyy[ic + c + 1];
        o.yy_m[c] = Bsyy[ic - c];
      }
      o.r00 = rn0;
      o.r01 = rn1;
      rn0 = Brho[ic + ny];
      rn1 = Brho[ic + ny + 1];
      o.r10 = rn0;
      o.r11 = rn1;
      const unsigned u = velocity2d_apply<VAR, VISCO, ORD, FAST>(A, x, y, s, srow, o);
      v2 = v2 > u ? v2 : u;
    }
  }
  if (blockIdx.x == 0 && !A.rec_after) record_receivers(A, s);
  wave_max_to_slot(v2, A.v2slots + blockIdx.x);
}

// ---- fused 2D step: velocity of step n, then stress of step n+1 -----------
// A workgroup of two wavefronts owns a strip of W = 64 - 2 ORD columns and
// marches x over a chunk of rows.  Lane j of each wavefront holds column
// y0 - ORD + j; the ORD lanes on each side are a halo.
// * Wavefront 0 (producer) computes v(n+1) row by row over all 64 columns,
//   starting 2 ORD rows before the chunk (register windows over x as in
//   k_velocity2d_march), stores the rows it owns and puts every row into an
//   LDS ring of 2 ORD + 2 rows.
// * Wavefront 1 (consumer) updates the stress of the chunk's rows in the owned
//   lanes, one row behind: at iteration t the producer writes ring row t while
//   the consumer reads rows t-1-2 ORD .. t-1 (one barrier per iteration; the
//   ring is one row longer than those, so the row being written is never one
//   being read).  Each role keeps only its own registers.
// * Halo lanes and lead-in rows compute velocities other workgroups own, from
//   the same inputs with the same arithmetic, so every copy has the same bits;
//   only the owner stores.  Those inputs must not change during the launch, so
//   sigma, v and the velocity-pass psi are read from one buffer set (F.A) and
//   written to the other (F.so, F.vo, F.psio).  Memory variables and the
//   stress-pass psi are touched only by their own cell and stay in place.
// * sigma outside the stress interiors is never written by either schedule,
//   so both sets keep the value both started with there.
// * Every expression is the two-pass kernels' (velocity2d_update, then
//   stress2d_core), in the same order: the fused step is bit-identical.
// * The receivers of step n read the stress channel from the input set, where
//   the previous stress pass left it final (a velocity channel is gathered
//   from the output set after the launch).
template <int ORD>
struct VRingL {
  const float (*vx_)[64];      // ring rows: v of this column's row r at vx_[slot(r)][lane]
  const float (*vy_)[64];
  int base, lane;              // ring slot of row x - ORD; this lane
  static constexpr int R = 2 * ORD + 2;
  __device__ __forceinline__ int slot(int dx) const { return (base + ORD + dx) % R; }
  __device__ __forceinline__ float vx(int dx, int dy) const { return vx_[slot(dx)][lane + dy]; }
  __device__ __forceinline__ float vy(int dx, int dy) const { return vy_[slot(dx)][lane + dy]; }
};

template <int VAR, bool VISCO, int ORD, int LC, bool FAST>
__global__ __launch_bounds__(128) void k_fused2d(FusedArgs F, int xchunk, int nstrip, int i) {
  constexpr int W = 64 - 2 * ORD, R = 2 * ORD + 2;
  __shared__ float ring_x[R][64], ring_y[R][64];
  const StepArgs &A = F.A;
  const int nx = A.nx, ny = A.ny;
  const int b = xcd_tile(blockIdx.x, gridDim.x, 1);
  const int strip = b % nstrip, chunk = b / nstrip;
  const int lane = threadIdx.x & 63, role = threadIdx.x >> 6;
  const int y = strip * W - ORD + lane;
  const bool col = y >= 0 && y < ny;
  const bool own = lane >= ORD && lane < 64 - ORD && y < ny;
  const int x0 = chunk * xchunk, x1 = min(nx, x0 + xchunk);
  const int T = (x1 - x0) + 2 * ORD;                 // velocity rows x0 - ORD .. x1 + ORD - 1
  const int s = *A.step + i;
  if (blockIdx.x == 0 && !A.rec_after) record_receivers(A, s);
  unsigned v2 = 0;
  if (role == 0) {
    const FieldBuf Bvx(A.vx, A.ncell), Bvy(A.vy, A.ncell), Bsxx(A.sxx, A.ncell), Bsxy(A.sxy, A.ncell),
        Bsyy(A.syy, A.ncell), Brho(A.rho, A.ncell);
    VSE_G float *const po[4] = {F.psio[0], F.psio[1], F.psio[3], F.psio[4]};
    // windows over the velocity rows xv: wxx[j] = sxx(xv - ORD + j), wxy[j] = sxy(xv - ORD + 1 + j)
    float wxx[2 * ORD], wxy[2 * ORD];
    const int xv0 = x0 - ORD;
    const int i0 = xv0 * ny + y;
#pragma unroll
    for (int j = 0; j < 2 * ORD - 1; j++) {
      wxx[j + 1] = Bsxx[i0 + (j - ORD) * ny];
      wxy[j + 1] = Bsxy[i0 + (j - ORD + 1) * ny];
    }
    float rn0 = Brho[i0], rn1 = Brho[i0 + 1];
#pragma unroll 1
    for (int t = 0; t <= T; t++) {
      if (t < T) {
        const int xv = xv0 + t;
        const int ic = xv * ny + y;
#pragma unroll
        for (int j = 0; j < 2 * ORD - 1; j++) { wxx[j] = wxx[j + 1]; wxy[j] = wxy[j + 1]; }
        wxx[2 * ORD - 1] = Bsxx[ic + (ORD - 1) * ny];
        wxy[2 * ORD - 1] = Bsxy[ic + ORD * ny];
        VelOps2<ORD> o;
        o.vxn = Bvx[ic];
        o.vyn = Bvy[ic];
#pragma unroll
        for (int c = 0; c < ORD; c++) {
          o.xx_p[c] = wxx[ORD + c];
          o.xx_m[c] = wxx[ORD - 1 - c];
          o.xy_p[c] = c == 0 ? wxy[ORD - 1] : Bsxy[ic + c];
          o.xy_m[c] = Bsxy[ic - c - 1];
          o.yx_p[c] = wxy[ORD + c];
          o.yx_m[c] = wxy[ORD - 1 - c];
          o.yy_p[c] = Bsyy[ic + c + 1];
          o.yy_m[c] = Bsyy[ic - c];
        }
        o.r00 = rn0;
        o.r01 = rn1;
        rn0 = Brho[ic + ny];
        rn1 = Brho[ic + ny + 1];
        o.r10 = rn0;
        o.r11 = rn1;
        // v(n+1) of row xv in this lane's column; 0 outside the grid, where no
        // stress operand uses it
        float vxo = 0.f, vyo = 0.f;
        if (col && xv >= 0 && xv < nx) {
          const bool mine = own && xv >= x0 && xv < x1;
          const unsigned u = velocity2d_update<VAR, VISCO, ORD, FAST>(A, xv, y, s, source_row<VAR>(A, xv), o, vxo,
                                                                      vyo, po, mine);
          if (mine) {
            const long id = (long)xv * ny + y;
            F.vo[0][id] = vxo;
            F.vo[1][id] = vyo;
            v2 = v2 > u ? v2 : u;
          }
        }
        ring_x[t % R][lane] = vxo;
        ring_y[t % R][lane] = vyo;
      }
      __syncthreads();
    }
  } else {
#pragma unroll 1
    for (int t = 0; t <= T; t++) {
      const int c = t - 1 - 2 * ORD;                   // stress row x0 + c, from ring rows c .. c + 2 ORD
      if (c >= 0 && own) {
        const VRingL<ORD> V{ring_x, ring_y, c % R, lane};
        stress2d_core<VAR, VISCO, ORD, LC, FAST>(A, x0 + c, y, V, F.so[0], F.so[1], F.so[3]);
      }
      __syncthreads();
    }
  }
  wave_max_to_slot(v2, A.v2slots + blockIdx.x);
}

// Receivers of a velocity channel, after the velocity pass that wrote it
// (one workgroup; step *A.step + i as in k_velocity2d).
__global__ __launch_bounds__(256) void k_record2d(StepArgs A, int i) { record_receivers(A, *A.step + i); }

// The same for every member of a batch that records a velocity channel
// (workgroup m serves member m).
__global__ __launch_bounds__(256) void k_record2d_batch(BatchArgs B, int i) {
  const StepArgs &A = B.args[blockIdx.x];
  if (A.rec_after) record_receivers(A, *A.step + i);
}

__global__ __launch_bounds__(64) void k_advance(int *step, int n) {
  if (threadIdx.x == 0) *step += n;
}

// every member's counter of a batch
__global__ __launch_bounds__(256) void k_advance_batch(BatchArgs B, int n) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < B.n) *B.sparse[m].step += n;
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

// The stress update of batch tile b (of the launch's tiles in dispatch order
// b, mapped to an XCD-contiguous tile order as tile2d does).
template <int VAR, bool VISCO, int ORD, int LC, bool FAST>
__device__ __forceinline__ void batch_stress_tile(const BatchArgs &B, int b) {
  const int i = batch_instance(B, b);
  const StepArgs &A = B.args[i];
  const int t = b - B.tile_beg[i];
  const int nyb = (A.ny + BY - 1) / BY;
  const int y = (t % nyb) * BY + threadIdx.x;
  const int x0 = (t / nyb) * BX * B.rows + threadIdx.y;
  for (int k = 0; k < B.rows; k++) stress2d_cell<VAR, VISCO, ORD, LC, FAST>(A, x0 + k * BX, y);
}

// The velocity update of batch tile b at each member's step *step + i; a
// member's first tile records its receivers (stress channels).
template <int VAR, bool VISCO, int ORD, bool FAST>
__device__ __forceinline__ void batch_velocity_tile(const BatchArgs &B, int b, int i) {
  const int m = batch_instance(B, b);
  const StepArgs &A = B.args[m];
  const int t = b - B.tile_beg[m];
  const int nyb = (A.ny + BY - 1) / BY;
  const int y = (t % nyb) * BY + threadIdx.x;
  const int x0 = (t / nyb) * BX * B.rows + threadIdx.y;
  unsigned v2 = 0;
  const int s = *A.step + i;
  for (int k = 0; k < B.rows; k++) {
    const unsigned u = velocity2d_cell<VAR, VISCO, ORD, FAST>(A, x0 + k * BX, y, s);
    v2 = v2 > u ? v2 : u;
  }
  if (t == 0 && !A.rec_after) record_receivers(A, s);
  wave_max_to_slot(v2, A.v2slots + t);
}

template <int VAR, bool VISCO, int ORD, int LC = 0, bool FAST = false>
__global__ __launch_bounds__(BX * BY) void k_stress2d_batch(BatchArgs B) {
  batch_stress_tile<VAR, VISCO, ORD, LC, FAST>(B, xcd_tile(blockIdx.x, gridDim.x, 1));
}

// One launch over the tiles of every member; each member's step is
// *step + i and its first tile records its receivers.
template <int VAR, bool VISCO, int ORD, bool FAST = false>
__global__ __launch_bounds__(BX * BY) void k_velocity2d_batch(BatchArgs B, int i) {
  batch_velocity_tile<VAR, VISCO, ORD, FAST>(B, xcd_tile(blockIdx.x, gridDim.x, 1), i);
}

// ---- persistent batch: nsteps steps of a group in one cooperative launch ----
// Grid-wide barrier of a persistent launch: a two-level arrival tree over
// device-scope atomics in `bar` (VSE_GRID_BAR_WORDS words, zero at the launch),
// counters VSE_GRID_BAR_STRIDE words apart so no two share a cache line.
// Workgroup b arrives on first-level counter b % ns (ns = min(grid, 64)); the
// last arrival of each counter for barrier k (counter == k * its members)
// arrives on the top counter, and the last of those publishes epoch k.  The
// counters only grow, so a barrier needs no reset.  Every wavefront first
// waits until its own stores are acknowledged; then thread 0 alone fences at
// device scope (the L2s of the XCDs are not coherent with each other): a
// release before its arrival (its XCD's L2 written back), a full fence
// between the levels, an acquire after the wait (stale lines invalidated).
// The wait is bounded: past VSE_GRID_BAR_SPIN polls the workgroup sets the
// fault word and goes on, so a launch always ends, and the host reports the
// fault (the results of such a launch are not valid).
__device__ __forceinline__ void grid_barrier(unsigned *bar, unsigned k) {
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned g = gridDim.x, ns = g < VSE_GRID_BAR_SUB ? g : VSE_GRID_BAR_SUB, sub = blockIdx.x % ns;
    const unsigned members = g / ns + (sub < g % ns ? 1u : 0u);
    unsigned *const top = bar + VSE_GRID_BAR_SUB * VSE_GRID_BAR_STRIDE;
    unsigned *const epoch = top + VSE_GRID_BAR_STRIDE;
    unsigned *const fault = epoch + VSE_GRID_BAR_STRIDE;
    const unsigned a = atomicAdd(bar + sub * VSE_GRID_BAR_STRIDE, 1u) + 1u;
    if (a == k * members) {
      __threadfence();
      const unsigned t = atomicAdd(top, 1u) + 1u;
      if (t == k * ns) {
        __threadfence();
        atomicMax(epoch, k);
      }
    }
    unsigned spins = 0;
    while (__hip_atomic_load(epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < k) {
      __builtin_amdgcn_s_sleep(2);
      ++spins;
      // give up past the spin limit, or at once after another workgroup did
      if (spins == VSE_GRID_BAR_SPIN ||
          ((spins & 255u) == 0 && __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        atomicMax(fault, 1u);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// The graph schedule launches one stress and one velocity kernel per step.
// Here one launch runs nsteps steps of every member: its workgroups (all
// resident at once -- the cooperative launch guarantees that or refuses to
// start) stride over the same `total` tiles in the same order, b = vb mapped
// as xcd_tile(vb, total, 1), and a grid-wide barrier stands where each kernel
// boundary was.  The cell updates are the same code, so the results are the
// same bits.  Every workgroup runs every barrier (the loop bounds are
// uniform over the grid).  Velocity-channel receivers are gathered after the
// barrier that ends each velocity pass; the next stress pass does not write
// them, and the barrier after it orders the gather before the velocity pass
// that does.  The counters move by nsteps after a last barrier, once no
// workgroup reads them any more.
template <int VAR, bool VISCO, int ORD, int LC, bool FAST>
__global__ __launch_bounds__(BX * BY) void k_batch2d_persistent(BatchArgs B, int total, int nsteps, int rec_after,
                                                                unsigned *bar) {
  unsigned k = 0;
  for (int i = 0; i < nsteps; i++) {
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x)
      batch_stress_tile<VAR, VISCO, ORD, LC, FAST>(B, xcd_tile(vb, total, 1));
    grid_barrier(bar, ++k);
    for (int vb = blockIdx.x; vb < total; vb += gridDim.x)
      batch_velocity_tile<VAR, VISCO, ORD, FAST>(B, xcd_tile(vb, total, 1), i);
    grid_barrier(bar, ++k);
    if (rec_after)
      for (int m = blockIdx.x; m < B.n; m += gridDim.x)
        if (B.args[m].rec_after) record_receivers(B.args[m], *B.args[m].step + i);
  }
  grid_barrier(bar, ++k);
  if (blockIdx.x == 0)
    for (int m = threadIdx.x + threadIdx.y * BY; m < B.n; m += BX * BY) *B.sparse[m].step += nsteps;
}

template <int VAR, bool VISCO, bool FAST>
hipError_t launch_ord(bool stress, int ord, const StepArgs &A, const KTune &t, int i, hipStream_t st) {
  const dim3 block(BY, BX);
  const int rows = t.xchunk > 1 ? t.xchunk : 1;
  const int nty = (A.ny + BY - 1) / BY, ntx = ((A.nx + BX - 1) / BX + rows - 1) / rows;
  const dim3 grid((unsigned)(nty * ntx));
  const int rm = t.remap;
  const bool march = t.nt == 1;      // velocity pass: x-row march with register windows
  const bool l3 = VISCO && VAR == VSE_VAR_WGSL && A.L == 3;
#define VSE_CASE(O)                                                                                                \
  case O:                                                                                                          \
    if (stress && l3) hipLaunchKernelGGL((k_stress2d<VAR, VISCO, O, 3, FAST>), grid, block, 0, st, A, nty, rm, rows); \
    else if (stress) hipLaunchKernelGGL((k_stress2d<VAR, VISCO, O, 0, FAST>), grid, block, 0, st, A, nty, rm, rows); \
    else if (march) hipLaunchKernelGGL((k_velocity2d_march<VAR, VISCO, O, FAST>), grid, block, 0, st, A, nty, rm, rows, i); \
    else hipLaunchKernelGGL((k_velocity2d<VAR, VISCO, O, FAST>), grid, block, 0, st, A, nty, rm, rows, i);          \
    break;
  switch (ord) {
    VSE_CASE(2) VSE_CASE(3) VSE_CASE(4) VSE_CASE(5) VSE_CASE(6)
    default: return hipErrorInvalidValue;
  }
#undef VSE_CASE
  return hipGetLastError();
}

template <int VAR, bool VISCO, bool FAST>
hipError_t launch_batch_ord(bool stress, int ord, const BatchArgs &B, int total_tiles, int i, hipStream_t st) {
  dim3 block(BY, BX), grid(total_tiles);
  const bool l3 = VISCO && VAR == VSE_VAR_WGSL && B.L == 3;
#define VSE_CASE(O)                                                                              \
  case O:                                                                                        \
    if (stress && l3) hipLaunchKernelGGL((k_stress2d_batch<VAR, VISCO, O, 3, FAST>), grid, block, 0, st, B); \
    else if (stress) hipLaunchKernelGGL((k_stress2d_batch<VAR, VISCO, O, 0, FAST>), grid, block, 0, st, B); \
    else hipLaunchKernelGGL((k_velocity2d_batch<VAR, VISCO, O, FAST>), grid, block, 0, st, B, i);          \
    break;
  switch (ord) {
    VSE_CASE(2) VSE_CASE(3) VSE_CASE(4) VSE_CASE(5) VSE_CASE(6)
    default: return hipErrorInvalidValue;
  }
#undef VSE_CASE
  return hipGetLastError();
}

// The persistent kernel of one instantiation (for the launch and the
// occupancy query).
template <int VAR, bool VISCO, bool FAST>
const void *persistent_fn(int ord, bool l3) {
#define VSE_CASE(O)                                                               \
  case O:                                                                         \
    return l3 ? (const void *)&k_batch2d_persistent<VAR, VISCO, O, 3, FAST>        \
              : (const void *)&k_batch2d_persistent<VAR, VISCO, O, 0, FAST>;
  switch (ord) {
    VSE_CASE(2) VSE_CASE(3) VSE_CASE(4) VSE_CASE(5) VSE_CASE(6)
    default: return nullptr;
  }
#undef VSE_CASE
}

const void *persistent_kernel(int variant, int visco, int ord, int math, int L) {
  const bool fast = math == VSE_MATH_FAST;
  if (variant == VSE_VAR_TAICHI) return persistent_fn<VSE_VAR_TAICHI, false, false>(ord, false);
  if (visco) return fast ? persistent_fn<VSE_VAR_WGSL, true, true>(ord, L == 3)
                         : persistent_fn<VSE_VAR_WGSL, true, false>(ord, L == 3);
  return fast ? persistent_fn<VSE_VAR_WGSL, false, true>(ord, false)
              : persistent_fn<VSE_VAR_WGSL, false, false>(ord, false);
}

// FAST applies to the WGSL variant (the Taichi variant already multiplies by f32(1/d)).
hipError_t dispatch(bool stress, int variant, int visco, int ord, int math, const StepArgs &A, const KTune &t, int i,
                    hipStream_t st) {
  const bool fast = math == VSE_MATH_FAST;
  if (variant == VSE_VAR_TAICHI) return launch_ord<VSE_VAR_TAICHI, false, false>(stress, ord, A, t, i, st);
  if (visco) return fast ? launch_ord<VSE_VAR_WGSL, true, true>(stress, ord, A, t, i, st)
                         : launch_ord<VSE_VAR_WGSL, true, false>(stress, ord, A, t, i, st);
  return fast ? launch_ord<VSE_VAR_WGSL, false, true>(stress, ord, A, t, i, st)
              : launch_ord<VSE_VAR_WGSL, false, false>(stress, ord, A, t, i, st);
}

}  // namespace

hipError_t vse_launch_batch2d(bool stress, int variant, int visco, int ord, int math, const BatchArgs &B,
                              int total_tiles, int i, hipStream_t st) {
  const bool fast = math == VSE_MATH_FAST;
  if (variant == VSE_VAR_TAICHI)
    return launch_batch_ord<VSE_VAR_TAICHI, false, false>(stress, ord, B, total_tiles, i, st);
  if (visco) return fast ? launch_batch_ord<VSE_VAR_WGSL, true, true>(stress, ord, B, total_tiles, i, st)
                         : launch_batch_ord<VSE_VAR_WGSL, true, false>(stress, ord, B, total_tiles, i, st);
  return fast ? launch_batch_ord<VSE_VAR_WGSL, false, true>(stress, ord, B, total_tiles, i, st)
              : launch_batch_ord<VSE_VAR_WGSL, false, false>(stress, ord, B, total_tiles, i, st);
}

hipError_t vse_batch2d_persistent_grid(int variant, int visco, int ord, int math, int L, int total_tiles, int *grid) {
  *grid = 0;
  const void *fn = persistent_kernel(variant, visco, ord, math, L);
  if (!fn || total_tiles < 1) return hipErrorInvalidValue;
  int dev = 0, ncu = 0, coop = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
  if (e == hipSuccess && !coop) e = hipErrorNotSupported;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BX * BY, 0);
  if (e != hipSuccess) return e;
  long g = std::min<long>((long)per_cu * ncu, total_tiles);
  if (g >= 8) g &= ~7L;
  if (g < 1) return hipErrorInvalidConfiguration;
  *grid = (int)g;
  return hipSuccess;
}

hipError_t vse_launch_batch2d_persistent(int variant, int visco, int ord, int math, const BatchArgs &B,
                                         int total_tiles, int grid, int nsteps, int rec_after, unsigned *bar,
                                         hipStream_t st) {
  const void *fn = persistent_kernel(variant, visco, ord, math, B.L);
  if (!fn || !bar || grid < 1 || grid > total_tiles || nsteps < 1 || nsteps > VSE_GRID_BAR_MAX_STEPS)
    return hipErrorInvalidValue;
  // the counters start at zero for every launch (the fault word is the caller's)
  hipError_t e = hipMemsetAsync(bar, 0, VSE_GRID_BAR_FAULT * sizeof(unsigned), st);
  if (e != hipSuccess) return e;
  BatchArgs b = B;
  int total = total_tiles, n = nsteps, ra = rec_after ? 1 : 0;
  void *args[] = {(void *)&b, (void *)&total, (void *)&n, (void *)&ra, (void *)&bar};
  return hipLaunchCooperativeKernel(fn, dim3((unsigned)grid), dim3(BY, BX), args, 0, st);
}

// Fused 2D step: grid of strips (64 - 2 ORD columns) x x-chunks.
long vse_fused2d_blocks(int nx, int ny, int ord, int xchunk) {
  const int W = 64 - 2 * ord;
  return (long)((ny + W - 1) / W) * ((nx + xchunk - 1) / xchunk);
}

hipError_t vse_launch_fused2d(int variant, int visco, int ord, int math, const FusedArgs &F, int xchunk, int i,
                              hipStream_t st) {
  if (xchunk < 1 || ord < 2 || ord > VSE_MAX_ORD) return hipErrorInvalidValue;
  const StepArgs &A = F.A;
  const int nstrip = (A.ny + (64 - 2 * ord) - 1) / (64 - 2 * ord);
  const dim3 grid((unsigned)vse_fused2d_blocks(A.nx, A.ny, ord, xchunk)), block(128);
  const bool fast = math == VSE_MATH_FAST;
  const bool l3 = visco && variant == VSE_VAR_WGSL && A.L == 3;
#define VSE_F2(VAR, VISCO, FASTV)                                                                              \
  switch (ord) {                                                                                              \
    case 2: if (l3) hipLaunchKernelGGL((k_fused2d<VAR, VISCO, 2, 3, FASTV>), grid, block, 0, st, F, xchunk, nstrip, i); \
            else hipLaunchKernelGGL((k_fused2d<VAR, VISCO, 2, 0, FASTV>), grid, block, 0, st, F, xchunk, nstrip, i); break; \
    case 3: hipLaunchKernelGGL((k_fused2d<VAR, VISCO, 3, 0, FASTV>), grid, block, 0, st, F, xchunk, nstrip, i); break; \
    case 4: if (l3) hipLaunchKernelGGL((k_fused2d<VAR, VISCO, 4, 3, FASTV>), grid, block, 0, st, F, xchunk, nstrip, i); \
            else hipLaunchKernelGGL((k_fused2d<VAR, VISCO, 4, 0, FASTV>), grid, block, 0, st, F, xchunk, nstrip, i); break; \
    case 5: hipLaunchKernelGGL((k_fused2d<VAR, VISCO, 5, 0, FASTV>), grid, block, 0, st, F, xchunk, nstrip, i); break; \
    case 6: hipLaunchKernelGGL((k_fused2d<VAR, VISCO, 6, 0, FASTV>), grid, block, 0, st, F, xchunk, nstrip, i); break; \
    default: return hipErrorInvalidValue;                                                                     \
  }
  if (variant == VSE_VAR_TAICHI) { VSE_F2(VSE_VAR_TAICHI, false, false) }
  else if (visco) { if (fast) { VSE_F2(VSE_VAR_WGSL, true, true) } else { VSE_F2(VSE_VAR_WGSL, true, false) } }
  else { if (fast) { VSE_F2(VSE_VAR_WGSL, false, true) } else { VSE_F2(VSE_VAR_WGSL, false, false) } }
#undef VSE_F2
  return hipGetLastError();
}

hipError_t vse_launch_advance_batch(const BatchArgs &B, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_advance_batch, dim3((unsigned)((B.n + 255) / 256)), dim3(256), 0, st, B, n);
  return hipGetLastError();
}

long vse_tiles2d(int nx, int ny, int rows) {
  return (long)((ny + BY - 1) / BY) * (((nx + BX - 1) / BX + rows - 1) / rows);
}

hipError_t vse_launch_stress2d(int variant, int visco, int ord, int math, const StepArgs &A, const KTune &t,
                               hipStream_t st) {
  return dispatch(true, variant, visco, ord, math, A, t, 0, st);
}
hipError_t vse_launch_velocity2d(int variant, int visco, int ord, int math, const StepArgs &A, const KTune &t, int i,
                                 hipStream_t st) {
  return dispatch(false, variant, visco, ord, math, A, t, i, st);
}
hipError_t vse_launch_advance(int *step, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, step, n);
  return hipGetLastError();
}
hipError_t vse_launch_record2d(const StepArgs &A, int i, hipStream_t st) {
  hipLaunchKernelGGL(k_record2d, dim3(1), dim3(256), 0, st, A, i);
  return hipGetLastError();
}
hipError_t vse_launch_record2d_batch(const BatchArgs &B, int i, hipStream_t st) {
  hipLaunchKernelGGL(k_record2d_batch, dim3((unsigned)B.n), dim3(256), 0, st, B, i);
  return hipGetLastError();
}
